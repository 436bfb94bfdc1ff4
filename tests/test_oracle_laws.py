"""The oracle's point laws against independent high-precision values and closed forms (CPU).

- Constant-field (GHK) route flux against the 70-digit Decimal formula the reference's junction
  check uses (validation/betse/junction_transport/check.py:10-23 of the reference): x (ca e^x - cb)
  / (e^x - 1) times A P / 1e-15, and its identities (zero flux at the Nernst potential, the Fick
  limit at V = 0, B(x) - B(-x) = -x, dissipation >= 0).
- The route slope against a Decimal central difference of the same formula.
- The junction gate's rates against BETSE's decimal rate equations as the reference's independent
  gate reference states them (validation/betse/junction_gate/reference.py:9-20): alpha = 1.3
  exp(-77 s) per s, beta = 1000 b / (1 + 50 b) with b = 0.0013 exp(140 s) (s in volts), to 1e-14.
- The Na/K- and Ca-ATPases: zero cycles at zero affinity, the second law (cycles x cycle free
  energy <= 0), exact zero at zero capacity, and the magnitude against a Decimal evaluation of the
  mass-action expression (laws/pumps/nak_atp_mass_action_positive.py:63-98).
"""

import ctypes
import math
from decimal import Decimal, localcontext

import numpy as np
import pytest

import oracle
from cells_amd.world import F, R, build, config0, config1

T = 310.0


def L():
    return oracle.lib()


def decimal_flux(ca, cb, dv, z, pa):
    with localcontext() as c:
        c.prec = 70
        ca, cb, dv, z, pa = (Decimal(repr(float(v))) for v in (ca, cb, dv, z, pa))
        Fd = Decimal(6.02214076e23 * 1.602176634e-19)
        Rd = Decimal(6.02214076e23 * 1.380649e-23)
        x = z * Fd * dv / (Rd * Decimal(T))
        if x == 0:
            drive = ca - cb
        else:
            e = x.exp()
            drive = x * (ca * e - cb) / (e - 1)
        return float(pa * drive * Decimal("1e15"))


VOLTAGES = [-0.12, -0.08, -0.03, -1e-3, -2e-6, 0.0, 3e-6, 2e-3, 0.04, 0.09, 0.12]


@pytest.mark.parametrize("z", [1.0, -1.0, 2.0])
def test_ghk_flux_meets_the_decimal_formula(z):
    for dv in VOLTAGES:
        for ca, cb in ((12.0, 145.0), (140.0, 5.0), (1e-4, 2.0), (10.0, 10.0)):
            got = L().ref_selective_flux(ctypes.c_double(ca), ctypes.c_double(cb),
                                         ctypes.c_double(dv), ctypes.c_double(z),
                                         ctypes.c_double(T), ctypes.c_double(4.7e-18),
                                         ctypes.c_double(1.0))
            want = decimal_flux(ca, cb, dv, z, 4.7e-18)
            scale = 4.7e-18 / 1e-15 * (abs(ca) + abs(cb)) * (1 + abs(z * F * dv / (R * T)))
            assert abs(got - want) <= 2e-14 * scale, (z, dv, ca, cb, got, want)


def test_ghk_identities():
    lib = L()
    for x in (-30.0, -3.0, -1e-3, -5e-5, 0.0, 5e-5, 1e-3, 2.0, 25.0):
        bx = lib.ref_bernoulli(ctypes.c_double(x))
        bm = lib.ref_bernoulli(ctypes.c_double(-x))
        assert abs((bx - bm) - (-x)) <= 1e-14 * max(1.0, abs(x))
    # zero flux at the Nernst potential, for each valence
    for z, ci, co in ((1.0, 12.0, 145.0), (-1.0, 10.0, 110.0), (2.0, 1e-4, 2.0)):
        vn = R * T / (z * F) * (math.log(co) - math.log(ci))
        j = lib.ref_selective_flux(*(ctypes.c_double(v) for v in (ci, co, vn, z, T, 1e-18, 1.0)))
        one_way = 1e-18 / 1e-15 * ci * lib.ref_bernoulli(ctypes.c_double(-z * F * vn / (R * T)))
        assert abs(j) <= 1e-13 * one_way
        # dissipation J mu >= 0 on either side of it
        for dv in (vn - 0.01, vn + 0.01):
            j = lib.ref_selective_flux(*(ctypes.c_double(v) for v in (ci, co, dv, z, T, 1e-18, 1.0)))
            mu = R * T * (math.log(ci) - math.log(co)) + z * F * dv
            assert j * mu > 0
    # Fick limit at V = 0
    j = lib.ref_selective_flux(*(ctypes.c_double(v) for v in (12.0, 145.0, 0.0, 1.0, T, 1e-18, 1.0)))
    assert math.isclose(j, 1e-18 / 1e-15 * (12.0 - 145.0), rel_tol=1e-15)


def test_route_slope_is_the_derivative():
    lib = L()
    for z in (1.0, -1.0, 2.0):
        for dv in (-0.1, -0.02, -5e-4, 0.0, 4e-4, 0.03, 0.1):
            ca, cb, pa = 12.0, 145.0, 3e-18
            got = lib.ref_route_slope(*(ctypes.c_double(v) for v in (ca, cb, dv, z, T, pa, 1.0)))
            h = 1e-7
            with localcontext() as c:
                c.prec = 50
                num = (Decimal(decimal_flux(ca, cb, dv + h, z, pa)) -
                       Decimal(decimal_flux(ca, cb, dv - h, z, pa))) / Decimal(2 * h)
            assert math.isclose(got, float(num), rel_tol=1e-7), (z, dv, got, float(num))
            # the slope of the charge current is positive (z dJ/dv >= 0)
            assert z * got > 0


def gate_decimal(v, threshold):
    with localcontext() as c:
        c.prec = 65
        offset = (abs(Decimal(repr(float(v)))) - Decimal(repr(float(threshold)))) * 1000
        a = Decimal("0.0013") * (-Decimal("0.077") * offset).exp()
        b = Decimal("0.0013") * (Decimal("0.14") * offset).exp()
        return float(1000 * a), float(1000 * b / (1 + 50 * b))


def test_gate_rhs_meets_the_decimal_rates():
    lib = L()
    worst = 0.0
    for threshold in (0.0, 0.015, 0.06, 0.12):
        for minimum in (0.0, 0.1, 0.4, 1.0):
            gate = minimum + 0.3 * (1 - minimum)
            for v in np.linspace(-0.12, 0.12, 25):
                a, b = gate_decimal(v, threshold)
                want = a * (1 - gate) - b * (gate - minimum)
                got = lib.ref_gate_rhs(*(ctypes.c_double(x) for x in (gate, v, threshold, minimum)))
                worst = max(worst, abs(got - want) / (a + b))
                # at the target the rate is zero
                target = (a + b * minimum) / (a + b)
                at = lib.ref_gate_rhs(*(ctypes.c_double(x) for x in (target, v, threshold, minimum)))
                assert abs(at) <= 1e-13 * (a + b)
    assert worst <= 1e-14


def _world():
    p = build(config1(worlds=1, cells=(4, 4)))
    return p, oracle.world_struct(p)


def _nak(w, na, ko, mu_na, mu_k, cap=0.05, avail=1.0):
    out = [ctypes.c_double() for _ in range(3)]
    L().ref_nak_kinetics(ctypes.byref(w), *(ctypes.c_double(v) for v in (na, ko, mu_na, mu_k, cap, avail)),
                         *(ctypes.byref(o) for o in out))
    return [o.value for o in out]


def test_nak_pump_stalls_at_zero_affinity_and_obeys_the_second_law():
    p, w = _world()
    c = p.config
    RT = R * T
    drive = c.dg0 + RT * (math.log(c.adp) + math.log(c.pi) - math.log(c.atp) - math.log(1000.0))
    rng = np.random.default_rng(0)
    for _ in range(200):
        mu_na, mu_k = rng.uniform(-30000, 30000, 2)
        cyc, energy, d = _nak(w, 12.0, 5.0, mu_na, mu_k)
        assert math.isclose(d, drive, rel_tol=1e-15)
        assert cyc * energy <= 0
    # zero affinity: pick mu_k so that drive - 3 mu_na + 2 mu_k = 0
    mu_na = -10000.0
    mu_k = (3 * mu_na - drive) / 2
    cyc, energy, _ = _nak(w, 12.0, 5.0, mu_na, mu_k)
    assert abs(energy) <= 1e-9 and abs(cyc) <= 1e-12 * 0.05
    # zero capacity or availability: an exact zero
    assert _nak(w, 12.0, 5.0, -9000.0, 1000.0, cap=0.0)[0] == 0.0
    assert _nak(w, 12.0, 5.0, -9000.0, 1000.0, avail=0.0)[0] == 0.0


def test_nak_pump_magnitude_meets_a_decimal_evaluation():
    p, w = _world()
    c = p.config
    for na, ko, mu_na, mu_k in ((12.0, 5.0, -12000.0, 2800.0), (30.0, 0.3, -5000.0, 9000.0),
                                (3.0, 20.0, -20000.0, -1000.0)):
        cyc, _, _ = _nak(w, na, ko, mu_na, mu_k)
        with localcontext() as ctx:
            ctx.prec = 60
            D = lambda v: Decimal(repr(float(v)))  # noqa: E731
            RT = D(R) * D(T)
            drive = D(c.dg0) + RT * (D(c.adp).ln() + D(c.pi).ln() - D(c.atp).ln() - D(1000).ln())
            a = (drive - 3 * D(mu_na) + 2 * D(mu_k)) / RT
            sat = (1 / (1 + (D(c.km_na) / D(na)) ** 3)) * (1 / (1 + (D(c.km_k) / D(ko)) ** 2)) * (
                1 / (1 + D(c.km_atp) / D(c.atp)))
            want = D(0.05) * sat * (1 - a.exp())  # mass action: capacity x saturation x (1 - e^a)
        assert math.isclose(cyc, float(want), rel_tol=1e-12), (cyc, float(want))


def test_ca_pump_second_law_and_stall():
    p, w = _world()
    out = [ctypes.c_double() for _ in range(3)]
    rng = np.random.default_rng(1)
    for _ in range(200):
        ci = 10 ** rng.uniform(-7, -3)
        v = rng.uniform(-0.12, 0.12)
        L().ref_ca_kinetics(ctypes.byref(w), *(ctypes.c_double(x) for x in (ci, 2.0, v, 0.002, 1.0)),
                            *(ctypes.byref(o) for o in out))
        cyc, energy = out[0].value, out[1].value
        assert cyc * energy <= 0


def test_logaddexp_is_numpys():
    for a, b in ((0.0, 0.0), (0.0, 3.0), (0.0, -40.0), (0.0, 800.0), (2.5, -1.0)):
        got = L().ref_logaddexp(ctypes.c_double(a), ctypes.c_double(b))
        assert math.isclose(got, float(np.logaddexp(a, b)), rel_tol=1e-15)


def test_the_oracle_meets_the_committed_decimal_vectors():
    """tests/golden/decimal_laws.json (written by tests/golden/make_decimal_vectors.py)."""
    import json
    import os

    data = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "decimal_laws.json")))
    lib = L()
    for g in data["ghk"]:
        got = lib.ref_selective_flux(*(ctypes.c_double(g[k]) for k in ("c_from", "c_to", "dv", "z")),
                                     ctypes.c_double(data["temperature"]), ctypes.c_double(g["pa"]),
                                     ctypes.c_double(1.0))
        scale = g["pa"] / 1e-15 * (g["c_from"] + g["c_to"]) * (1 + abs(g["z"] * F * g["dv"] / (R * T)))
        assert abs(got - g["flux"]) <= 2e-14 * scale, g
    for g in data["gate"]:
        minimum, gate = 0.1, 0.4
        want = g["alpha"] * (1 - gate) - g["beta"] * (gate - minimum)
        got = lib.ref_gate_rhs(*(ctypes.c_double(x) for x in (gate, g["v"], g["threshold"], minimum)))
        assert abs(got - want) <= 1e-14 * (g["alpha"] + g["beta"]), g
