"""The oracle's N2 step (oracle/cells_ref.c ref_step_n2: the reference's ARS(2,2,2) with implicit
slope conductances, numerics/surface_imex.py:38-79, as models/surface_pump_v3.py:3547-3955
imex_step runs it for cells on one held bath) and its exponential gate update
(numerics/exponential_gates.py:37-51), pinned by mathematics the code does not restate:

- the gate update is the exact solution of dg/dt = (g_inf(t) - g) / tau with tau constant and
  g_inf linear in time (checked against that ODE's closed form in 50-digit Decimal), it equals the
  constant-kinetics relaxation when both ends agree, and it stays between the start gate and the
  two targets for any step; its lag weight's series meets the closed form at the switch;
- N2 converges at second order to the solution of the same ODE computed by another method (the
  oracle's ROS2 at a step 50 times smaller): halving dt cuts the error at least threefold;
- N2 runs close every ledger with N2's own energy bound (2e-3 of the flows,
  models/surface_pump.py:64) from the first step, keep each gate in [minimum, 1], and agree with
  ROS2 and IMEX Euler runs to within those methods' errors.
"""

import math
from decimal import Decimal, localcontext

import numpy as np
import pytest

import oracle
from cells_amd.world import StateLayout, build, config0, config1, initial_state
from tests.helpers import conc, relerr, vmem


def advance(g, ta, tau_a, tb, tau_b, h):
    return oracle.lib().ref_gate_advance(g, ta, tau_a, tb, tau_b, h)


def test_gate_update_solves_the_linear_target_ode_exactly():
    rng = np.random.default_rng(3)
    for _ in range(200):
        g0, a, b = rng.uniform(0.1, 1.0, 3)
        tau = 10 ** rng.uniform(-3, 0)
        h = 10 ** rng.uniform(-4, 0)
        got = advance(g0, a, tau, b, tau, h)
        with localcontext() as ctx:
            ctx.prec = 50
            D = lambda v: Decimal(repr(float(v)))  # noqa: E731
            z = D(h) / D(tau)
            s_tau = (D(b) - D(a)) / z  # the target's slope times tau
            want = D(b) - s_tau + (D(g0) - D(a) + s_tau) * (-z).exp()
        assert math.isclose(got, float(want), rel_tol=1e-12, abs_tol=1e-15), (g0, a, b, tau, h)


def test_gate_update_is_the_relaxation_at_constant_kinetics_and_stays_between_its_ends():
    rng = np.random.default_rng(4)
    for _ in range(500):
        g0, t0 = rng.uniform(0.1, 1.0, 2)
        tau = 10 ** rng.uniform(-4, 1)
        h = 10 ** rng.uniform(-5, 1)
        got = advance(g0, t0, tau, t0, tau, h)
        want = t0 + (g0 - t0) * math.exp(-h / tau)
        assert math.isclose(got, want, rel_tol=1e-13, abs_tol=1e-15)
        ta, tb = rng.uniform(0.1, 1.0, 2)
        tau_b = 10 ** rng.uniform(-4, 1)
        g = advance(g0, ta, tau, tb, tau_b, h)
        lo, hi = min(g0, ta, tb), max(g0, ta, tb)
        assert lo - 1e-15 <= g <= hi + 1e-15, (g, lo, hi)
    # the lag weight's series and closed form meet at the switch (z = 1e-3)
    z = 1e-3
    closed = (math.expm1(-z) + z) / z
    assert math.isclose(oracle.lib().ref_lag_weight(z * (1 - 1e-12)), closed, rel_tol=1e-12)
    assert oracle.lib().ref_lag_weight(0.0) == 0.0


@pytest.fixture(scope="module")
def calcium():
    p = build(config1(worlds=2, cells=(40, 70), seed=3))
    return p, oracle.world_struct(p), oracle.world_symbolics(p)[1]


def run(step, p, w, sym, dt, span, y=None):
    y = initial_state(p) if y is None else y
    for k in range(int(round(span / dt))):
        y, ok = step(p, y, dt, sym, world=w)
        assert ok.all(), k
    return y


def test_n2_converges_at_second_order_to_another_method(calcium):
    p, w, sym = calcium
    span = 0.2
    fine = run(oracle.step_ros2, p, w, sym, 1e-4, span)
    errs = []
    for dt in (1e-2, 5e-3, 2.5e-3):
        y = run(oracle.step_n2, p, w, sym, dt, span)
        errs.append(max(relerr(vmem(p, y), vmem(p, fine)), relerr(conc(p, y), conc(p, fine))))
    assert errs[0] / errs[1] >= 3.0 and errs[1] / errs[2] >= 3.0, errs
    assert errs[-1] < 1e-7


def test_n2_closes_every_ledger_from_the_first_step_and_keeps_its_gates(calcium):
    p, w, sym = calcium
    L = StateLayout.of(p)
    y = initial_state(p)
    for k in range(1, 61):
        prev = y
        y, ok = oracle.step_n2(p, y, 5e-3, sym, world=w)
        assert ok.all()
        bits, ratios, _ = oracle.ledgers(p, prev, y, k, "n2", world=w)
        assert bits.tolist() == [0] * p.W, (k, ratios)
        g = L.edge(y, L.gate)
        assert np.all(g >= p.config.gate_minimum) and np.all(g <= 1.0)
    # the same span by ROS2 and by IMEX Euler (at a tenth of the step) lands within their errors
    ros2 = run(oracle.step_ros2, p, w, sym, 5e-3, 0.3)
    assert relerr(vmem(p, y), vmem(p, ros2)) < 1e-5 and relerr(conc(p, y), conc(p, ros2)) < 1e-5


def test_n2_with_water_runs_the_plain_world():
    p = build(config0())
    w = oracle.world_struct(p)
    _, sym = oracle.world_symbolics(p)
    y = initial_state(p)
    for k in range(1, 201):
        prev = y
        y, ok = oracle.step_n2(p, y, 1e-3, sym, world=w)
        assert ok.all(), k
    bits, ratios, _ = oracle.ledgers(p, prev, y, 200, "n2", world=w)
    assert bits.tolist() == [0], ratios
    yl = y
    ye = initial_state(p)
    for _ in range(2000):
        ye, ok = oracle.step_euler(p, ye, 1e-4, world=w)
        assert ok.all()
    assert relerr(vmem(p, yl), vmem(p, ye)) < 1e-3 and relerr(conc(p, yl), conc(p, ye)) < 1e-3
