"""The oracle's row checks (oracle/cells_ref.c ref_ledgers; the reference's row_checks,
models/surface_pump_checks.py:112-192, over surface_pump.py:685-696's LEDGERS and the water
operator's, operators/osmotic_water.py:240, with ledger_limits' bounds, surface_pump.py:752-800).

- Accepted trajectories of the plain world (IMEX Euler, water) and of calcium worlds (ROS2) pass
  every ledger.
- The energy ledger is pinned by the physics it books, not by its formula: the step conserves
  energy in the limit dt -> 0, so the ledger's residual over a fixed span falls with the method's
  order (IMEX Euler: halving dt halves it; ROS2: at least three times smaller). Each plausible
  slip (a term dropped, the work's sign flipped, the capacitor energy without its cross term, the
  mixing energy without its logarithm, water's dilution with the wrong sign) leaves a residual
  that does not fall, and the pin misses it.
- Every identity ledger trips on a booking error planted in its own field, and on nothing else
  that is planted elsewhere.
"""

import numpy as np
import pytest

import oracle
from cells_amd.world import F, R, StateLayout, build, config0, config1, initial_state


def trajectory(p, method, dt, steps):
    w = oracle.world_struct(p)
    sym = oracle.world_symbolics(p)[1] if method == "ros2" else None
    y = initial_state(p)
    prev = y
    for _ in range(steps):
        prev = y
        if method == "ros2":
            y, ok = oracle.step_ros2(p, y, dt, sym, world=w)
        else:
            y, ok = oracle.step_euler(p, y, dt, world=w)
        assert ok.all()
    return prev, y


@pytest.fixture(scope="module")
def plain():
    p = build(config0())
    return p, trajectory(p, "euler", 1e-3, 400)


@pytest.fixture(scope="module")
def calcium():
    p = build(config1(worlds=2, cells=(40, 70), seed=3))
    return p, trajectory(p, "ros2", 5e-3, 80)


def test_accepted_trajectories_pass_every_ledger(plain, calcium):
    for (p, (prev, y)), method, steps in ((plain, "euler", 400), (calcium, "ros2", 80)):
        bits, ratios, terms = oracle.ledgers(p, prev, y, steps, method)
        assert bits.tolist() == [0] * p.W, (bits, ratios)
        energy = oracle.LEDGER_RATIOS.index("energy")
        assert np.all(np.delete(ratios, energy, axis=1) <= 0.05), ratios  # the identities: rounding
        assert np.all(ratios[:, energy] <= 0.2), ratios  # the method's energy error, early in a run
        assert np.all(terms[:, 6] > 0)  # heat and work flowed


def _slipped(p, y, terms, slip):
    """The energy residual of world 0 with one plausible slip in its terms."""
    L = StateLayout.of(p)
    c0, c1 = int(p.cell_off[0]), int(p.cell_off[1])
    mixing, capacitor, heat, work, water, residual, _ = terms[0]
    q = F * 1e-15 * sum(p.valence[s] * L.cell(y, L.dn + s)[c0:c1] for s in range(p.S))
    C = p.cap[c0:c1]
    if slip == "no_capacitor":
        return residual - capacitor
    if slip == "work_sign":
        return residual + 2 * work
    if slip == "no_heat":
        return residual - heat
    if slip == "capacitor_without_cross_term":
        return residual - capacitor + 0.5e12 * np.sum(C * (q / C) ** 2)
    if slip == "mixing_without_logarithm":
        RT = R * p.config.temperature
        lin = sum(np.sum(L.cell(y, L.dn + s)[c0:c1] * (np.log(p.n0[s][c0:c1] / p.vol0[c0:c1]) - 1))
                  for s in range(p.S))
        return residual - mixing + 1e-3 * RT * lin
    if slip == "dilution_sign":
        RT = R * p.config.temperature
        dv = L.cell(y, L.dvol)[c0:c1]
        total = sum(p.n0[s][c0:c1] + L.cell(y, L.dn + s)[c0:c1] for s in range(p.S)) + p.imperm[c0:c1]
        dilution = -1e-3 * RT * np.sum(total * np.log1p(dv / p.vol0[c0:c1]))
        return residual - 2 * dilution
    raise ValueError(slip)


@pytest.mark.parametrize("method,cfg,dt,span,order", [
    ("euler", config0(), 1e-3, 0.4, (1.8, 2.2)),
    ("ros2", config1(worlds=1, cells=(40, 40), seed=5), 4e-3, 0.4, (3.0, 8.0)),
])
def test_energy_residual_falls_with_the_method_order_and_each_slip_misses(method, cfg, dt, span, order):
    p = build(cfg)
    out = {}
    for h in (dt, dt / 2):
        n = int(round(span / h))
        prev, y = trajectory(p, method, h, n)
        _, _, terms = oracle.ledgers(p, prev, y, n, method)
        out[h] = (y, terms)
    r1, r2 = out[dt][1][0, 5], out[dt / 2][1][0, 5]
    assert abs(r1) > 0 and order[0] <= r1 / r2 <= order[1], (r1, r2, r1 / r2)
    slips = ["no_capacitor", "work_sign", "no_heat", "capacitor_without_cross_term", "mixing_without_logarithm"]
    if p.water:
        slips.append("dilution_sign")
    for slip in slips:
        s1 = _slipped(p, out[dt][0], out[dt][1], slip)
        s2 = _slipped(p, out[dt / 2][0], out[dt / 2][1], slip)
        # a slip's residual stays put when dt halves (it is no discretisation error)
        assert not (order[0] <= s1 / s2 <= order[1]), (slip, s1, s2)
        assert abs(s2) > 10 * abs(r2), (slip, s2, r2)


# (field kind, field, amount, the checks it must trip)
PLANTS = [
    ("cell", "pwork", 1e-9, {"metabolic_work"}),
    ("cell", "cwork", 1e-9, {"metabolic_work"}),
    ("cell", "mtr+1", 1e-9, {"local", "outside"}),
    ("cell", "dn+0", 1e-9, {"local", "total", "electrode"}),  # the charge moved too
    ("world", "rwork", 1e-9, {"reservoir_work"}),
    ("world", "ework", 1e-9, {"electrode_work"}),
    ("world", "echg", 1e-19, {"global_charge", "electrode"}),
    ("world", "sup+2", 1e-9, {"total", "outside", "global_charge", "reservoir_work"}),  # work is booked on it
    ("cell", "mdiss", 50.0, {"energy"}),
]


@pytest.mark.parametrize("kind,field,amount,want", PLANTS, ids=[f for _, f, _, _ in PLANTS])
def test_a_booking_error_trips_its_own_ledgers(calcium, kind, field, amount, want):
    p, (prev, y) = calcium
    L = StateLayout.of(p)
    bad = y.copy()
    name, _, off = field.partition("+")
    f = getattr(L, name) + (int(off) if off else 0)
    if kind == "cell":
        L.cell(bad, f)[int(p.cell_off[1]) + 3] += amount
    else:
        L.world(bad, f)[1] += amount
    bits, ratios, _ = oracle.ledgers(p, prev, bad, 80, "ros2")
    assert bits[0] == 0
    tripped = {k for k, v in zip(oracle.LEDGER_RATIOS, ratios[1]) if v > 1.0}
    assert tripped == want, (field, tripped)
    assert bits[1] == sum(hip_bit(k) for k in want)


def hip_bit(name):
    from cells_amd import lib

    return lib.CHECKS[name]


def test_water_ledgers_trip_on_their_own_fields(plain):
    p, (prev, y) = plain
    L = StateLayout.of(p)
    for field, where, want in (("wtr", "cell", {"water"}), ("wwork", "world", {"bath_water_work"}),
                               ("wsup", "world", {"water", "bath_water_work"})):  # work is booked on it
        bad = y.copy()
        if where == "cell":
            L.cell(bad, getattr(L, field))[5] += 1e-9
        else:
            L.world(bad, getattr(L, field))[0] += 1e-9
        bits, ratios, _ = oracle.ledgers(p, prev, bad, 400, "euler")
        tripped = {k for k, v in zip(oracle.LEDGER_RATIOS, ratios[0]) if v > 1.0}
        assert tripped == want, (field, tripped)
    # heat that fell since the previous row
    bad = y.copy()
    L.cell(bad, L.wdiss)[2] = L.cell(prev, L.wdiss)[2] - 1.0
    bits, _, _ = oracle.ledgers(p, prev, bad, 400, "euler")
    assert bits[0] & hip_bit("dissipation")


def test_the_readouts_energy_balance_vanishes_and_each_slip_breaks_it(calcium):
    """The readouts' energy rates (models/surface_pump_v3.py:2221-2262) at any state: the free
    energy's rate (chemical + capacitor) plus the dissipation less the work is zero in exact
    arithmetic, since the chemical rate sum ln(c) dN/dt is the ledger's mixing term's derivative
    and the capacitor rate its capacitor term's. The oracle's balance is below 1e-13 of the
    dissipation; a capacitor term with the wrong sign, the mixing rate without its logarithm (the
    concentration itself) or the pump heat left out each leave an imbalance."""
    p, (prev, y) = calcium
    L = StateLayout.of(p)
    rng = np.random.default_rng(5)
    states = [y, prev]
    for _ in range(3):  # states off the trajectory: neutral salt moved, a little charge, gates
        z = y.copy()
        salt = rng.normal(scale=0.02, size=p.ncell) * p.n0[0]
        L.cell(z, L.dn + 0)[:] += salt
        L.cell(z, L.dn + 2)[:] += salt
        L.cell(z, L.dn + 1)[:] += rng.normal(scale=1e-7, size=p.ncell) * p.n0[1]
        L.edge(z, L.gate)[:] = rng.uniform(0.2, 1.0, size=p.nedge)
        states.append(z)
    names = oracle.ENERGY_RATES
    for z in states:
        er = oracle.energy_rates(p, z)
        d = er[:, names.index("dissipation")]
        assert np.all(d > 0)
        assert np.all(np.abs(er[:, names.index("energy_balance")]) <= 1e-13 * d), er
        cap = er[:, names.index("capacitor")]
        assert np.all(np.abs(er[:, 6] - 2 * cap) > 1e3 * 1e-13 * d)  # the capacitor's sign
        r, _ = oracle.rates(p, z)
        for k in range(p.W):
            c0, c1 = int(p.cell_off[k]), int(p.cell_off[k + 1])
            vol = p.vol0[c0:c1]
            lin = sum(np.sum((p.n0[s][c0:c1] + L.cell(z, L.dn + s)[c0:c1]) / vol * L.cell(r, L.dn + s)[c0:c1])
                      for s in range(p.S))
            slipped = er[k, 6] - er[k, 0] + 1e-3 * R * p.config.temperature * lin
            assert abs(slipped) > 1e-6 * d[k]
            pump_heat = np.sum(L.cell(r, L.pdiss)[c0:c1])
            assert abs(er[k, 6] - pump_heat) > 1e-6 * d[k]
