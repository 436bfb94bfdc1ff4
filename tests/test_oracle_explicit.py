"""The oracle's explicit tableaux on cells over a held bath (oracle/cells_ref.c): ARS(2,2,2)
(numerics/surface_imex.py:53-57, its implicit part empty without faces) and the classical RK4 of an
empty tableau (numerics/fixed_step.py:25-71), pinned by their orders against a fine RK4 solution of
the same ODE (ARS: halving dt quarters the error; RK4: cuts it at least twelvefold; IMEX Euler: halves
it), and by their ledgers (every check passes, RK4's energy within the absolute floor alone,
ENERGY_RTOL["rk4"] = 0, models/surface_pump.py:63)."""

import numpy as np
import pytest

import oracle
from cells_amd.world import build, config0, config1, initial_state
from tests.helpers import conc, relerr, vmem


@pytest.fixture(scope="module")
def worlds():
    p = build(config1(worlds=2, cells=(40, 70), seed=3))
    return p, oracle.world_struct(p)


def run(step, p, w, dt, span):
    y = initial_state(p)
    for k in range(int(round(span / dt))):
        y, ok = step(p, y, dt, world=w)
        assert ok.all(), (k, dt)
    return y


def error(p, y, ref):
    return max(relerr(vmem(p, y), vmem(p, ref)), relerr(conc(p, y), conc(p, ref)))


@pytest.mark.parametrize("name,dts,low,high", [
    ("ars222", (4e-5, 2e-5, 1e-5), 3.8, 4.2),
    ("euler", (4e-5, 2e-5, 1e-5), 1.9, 2.1),
    ("rk4", (8e-4, 4e-4, 2e-4), 12.0, 20.0),
])
def test_orders_against_a_fine_rk4_solution(worlds, name, dts, low, high):
    p, w = worlds
    step = {"ars222": oracle.step_ars222, "euler": oracle.step_euler, "rk4": oracle.step_rk4}[name]
    span = 0.02 if name != "rk4" else 0.04
    fine = run(oracle.step_rk4, p, w, dts[-1] / 8, span)
    errs = [error(p, run(step, p, w, dt, span), fine) for dt in dts]
    ratios = [errs[i] / errs[i + 1] for i in range(2)]
    assert all(low <= r <= high for r in ratios), (errs, ratios)


def test_ars222_books_a_falling_junction_heat_at_the_larger_step(worlds):
    """ARS(2,2,2)'s explicit weight delta < 0 books a negative junction heat increment where a
    junction's power falls by more than 59 % within a step (models/surface_pump_v3.py:3784-3790,
    the reason for the reference's trapezoid heat option); the dissipation check, and it alone,
    catches it at 0.5 ms on these worlds."""
    p, w = worlds
    y = initial_state(p)
    seen = 0
    for k in range(1, 31):
        prev = y
        y, ok = oracle.step_ars222(p, y, 5e-4, world=w)
        bits, _, _ = oracle.ledgers(p, prev, y, k, "ars222", world=w)
        seen |= int(np.bitwise_or.reduce(bits))
    assert seen == 16, seen


@pytest.mark.parametrize("name,dt", [("ars222", 2e-4), ("rk4", 4e-4)])
def test_ledgers_close(worlds, name, dt):
    p, w = worlds
    step = {"ars222": oracle.step_ars222, "rk4": oracle.step_rk4}[name]
    y = initial_state(p)
    for k in range(1, 51):
        prev = y
        y, ok = step(p, y, dt, world=w)
        assert ok.all()
        bits, ratios, _ = oracle.ledgers(p, prev, y, k, name, world=w)
        assert bits.tolist() == [0] * p.W, (k, ratios)


def test_ars222_runs_the_plain_world_with_water_and_rk4_refuses_it():
    p = build(config0())
    w = oracle.world_struct(p)
    y = initial_state(p)
    for k in range(1, 101):
        prev = y
        y, ok = oracle.step_ars222(p, y, 2e-4, world=w)
        assert ok.all(), k
    bits, ratios, _ = oracle.ledgers(p, prev, y, 100, "ars222", world=w)
    assert bits.tolist() == [0], ratios
    _, ok = oracle.step_rk4(p, initial_state(p), 1e-4, world=w)
    assert not ok.any()
