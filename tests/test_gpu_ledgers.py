"""The device row checks' full ledger set on the GPU, through the C-ABI, against the oracle's
(oracle/cells_ref.c ref_ledgers; the reference's row_checks, models/surface_pump_checks.py:112-192,
over surface_pump.py:685-696's LEDGERS and the water operator's, operators/osmotic_water.py:240).

- On accepted oracle trajectories (IMEX Euler with water; ROS2 with calcium) the library's bits
  equal the oracle's (none set), its energy ratio equals the oracle's within 1e-9 relative (the
  residual is the method's energy error, not rounding), and every identity ledger sits at
  rounding level in both.
- A booking error planted in one field trips the same ledgers on the GPU as in the oracle, one
  plant per ledger bit.
- In the step's fused epilogue, an input state that books heat that never flowed refuses its
  world with the energy bit alone; the other worlds step.
"""

import sys

import numpy as np
import pytest

import oracle
from cells_amd import lib as hip
from cells_amd.world import StateLayout, build, config0, config1, initial_state
from tests.test_oracle_ledgers import PLANTS, trajectory

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def torch():
    import torch as t

    if not t.cuda.is_available():
        pytest.skip("no GPU")
    return t


def ensemble(p, method):
    from cells_amd.ensemble import Ensemble

    return Ensemble(p, method)


CASES = [("euler", config0(worlds=2, seed=4), 1e-3, 300), ("ros2", config1(worlds=3, cells=(40, 260), seed=9), 5e-3, 60)]


@pytest.mark.parametrize("method,cfg,dt,steps", CASES, ids=["plain", "calcium"])
def test_ledgers_on_accepted_rows_match_the_oracle(torch, method, cfg, dt, steps):
    p = build(cfg)
    prev, y = trajectory(p, method, dt, steps)
    bits_o, ratios_o, _ = oracle.ledgers(p, prev, y, steps, method)
    e = ensemble(p, method)
    say(f"{method}: device checks on an oracle trajectory of {steps} steps")
    failed, ratios = e.device_checks(prev, y, steps)
    failed, ratios = failed.cpu().numpy(), ratios.cpu().numpy()
    assert failed.tolist() == bits_o.tolist() == [0] * p.W, (failed, ratios)
    k = hip.LEDGER_RATIOS.index("energy")
    assert np.all(ratios_o[:, k] > 0)
    assert np.allclose(ratios[:, k], ratios_o[:, k], rtol=1e-9, atol=0), (ratios[:, k], ratios_o[:, k])
    rest = np.delete(ratios, k, axis=1)
    assert np.all(rest <= 0.05) and np.all(np.delete(ratios_o, k, axis=1) <= 0.05), (ratios, ratios_o)
    e.close()


def test_each_planted_booking_error_trips_the_oracles_ledgers(torch):
    p = build(config1(worlds=2, cells=(40, 70), seed=3))
    prev, y = trajectory(p, "ros2", 5e-3, 80)
    L = StateLayout.of(p)
    e = ensemble(p, "ros2")
    seen = 0
    for kind, field, amount, want in PLANTS:
        bad = y.copy()
        name, _, off = field.partition("+")
        f = getattr(L, name) + (int(off) if off else 0)
        if kind == "cell":
            L.cell(bad, f)[int(p.cell_off[1]) + 3] += amount
        else:
            L.world(bad, f)[1] += amount
        bits_o, _, _ = oracle.ledgers(p, prev, bad, 80, "ros2")
        failed, _ = e.device_checks(prev, bad, 80)
        got = failed.cpu().numpy().tolist()
        say(f"plant {field}: bits {got} (oracle {bits_o.tolist()})")
        assert got == bits_o.tolist() and got[0] == 0
        assert got[1] == sum(hip.CHECKS[k] for k in want)
        seen |= got[1]
    # capacitor charge: no stored field carries it (every charge derives from the amounts); a
    # non-finite amount is the domain's
    bad = y.copy()
    L.cell(bad, L.dn)[int(p.cell_off[0])] = np.nan
    failed, _ = e.device_checks(prev, bad, 80)
    assert failed.cpu().numpy()[0] & hip.CHECKS["domain"]
    e.close()
    want_seen = {"local", "total", "outside", "global_charge", "electrode", "energy", "reservoir_work",
                 "electrode_work", "metabolic_work"}
    assert seen == sum(hip.CHECKS[k] for k in want_seen)


def test_water_plants_trip_the_oracles_ledgers(torch):
    p = build(config0(worlds=2, seed=4))
    prev, y = trajectory(p, "euler", 1e-3, 300)
    L = StateLayout.of(p)
    e = ensemble(p, "euler")
    for field, where in (("wtr", "cell"), ("wwork", "world"), ("wsup", "world"), ("wdiss", "fall")):
        bad = y.copy()
        if where == "cell":
            L.cell(bad, getattr(L, field))[int(p.cell_off[1]) + 5] += 1e-9
        elif where == "world":
            L.world(bad, getattr(L, field))[1] += 1e-9
        else:
            i = int(p.cell_off[1]) + 2
            L.cell(bad, L.wdiss)[i] = L.cell(prev, L.wdiss)[i] - 1.0
        bits_o, _, _ = oracle.ledgers(p, prev, bad, 300, "euler")
        failed, _ = e.device_checks(prev, bad, 300)
        got = failed.cpu().numpy().tolist()
        say(f"water plant {field}: bits {got} (oracle {bits_o.tolist()})")
        assert got == bits_o.tolist() and got[0] == 0 and got[1] != 0
    e.close()


def test_the_fused_epilogue_refuses_heat_that_never_flowed(torch):
    p = build(config1(worlds=3, cells=(50, 120), seed=13))
    L = StateLayout.of(p)
    e = ensemble(p, "ros2")
    y0 = initial_state(p)
    y1, ok = e.step(y0, 5e-3, nsteps=40)
    assert bool(ok.all())
    bad = y1.cpu().numpy().copy()
    i = int(p.cell_off[2]) + 7
    L.cell(bad, L.mdiss)[i] += 1e3  # pJ of membrane heat that no transfer produced
    e.step_checks(True, steps_before=40)
    out, ok = e.step(bad, 5e-3, nsteps=5)
    bits = e.check_bits().cpu().numpy()
    say(f"fused epilogue: ok {ok.cpu().numpy().tolist()}, bits {bits.tolist()}")
    assert ok.cpu().numpy().tolist() == [1, 1, 0]
    assert bits[2] == hip.CHECKS["energy"] and not bits[:2].any()
    assert np.array_equal(out.cpu().numpy()[L.world_base:][2::p.W], bad[L.world_base:][2::p.W])
    e.step_checks(False)
    e.close()


def test_energy_readouts_match_the_oracle_and_balance(torch):
    """cells_energy_rates (the readouts' energy rates, models/surface_pump_v3.py:2221-2262) against
    the oracle's at trajectory states: every term within 1e-12 of the dissipation, the balance
    within 1e-12 of it (zero in exact arithmetic)."""
    p = build(config1(worlds=3, cells=(40, 260), seed=9))
    _, y = trajectory(p, "ros2", 5e-3, 30)
    e = ensemble(p, "ros2")
    got = e.energy_rates(y).cpu().numpy()
    want = oracle.energy_rates(p, y)
    d = want[:, hip.ENERGY_RATES.index("dissipation")]
    say(f"energy readouts: balance / dissipation {(got[:, 6] / d).tolist()}")
    for k in range(len(hip.ENERGY_RATES) - 1):
        assert np.all(np.abs(got[:, k] - want[:, k]) <= 1e-12 * d), (hip.ENERGY_RATES[k], got[:, k], want[:, k])
    assert np.all(np.abs(got[:, 6]) <= 1e-12 * d)
    cur = hip.ENERGY_RATES.index("electrode_current")
    assert np.allclose(got[:, cur], want[:, cur], rtol=1e-10, atol=1e-25)
    e.close()
