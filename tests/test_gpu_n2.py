"""N2 on the GPU (CELLS_METHOD_N2: the reference's ARS(2,2,2) with implicit slope conductances,
numerics/surface_imex.py:38-79 and models/surface_pump_v3.py:3547-3955), through the C-ABI, against
the oracle's ref_step_n2:

- calcium worlds (one above 2048 cells) over 100 steps of 5 ms, and the plain world with water
  over 300 steps of 1 ms, at 1e-9 relative in the potentials and concentrations, the ledgers and
  gates included;
- the reference-order factor gives the same trajectory; the batched launch equals the plain one
  bit for bit;
- the fused ledger checks pass every N2 step with N2's energy bound; a zero stage-residual
  bound refuses every world with the residual bit.
"""

import sys

import numpy as np
import pytest

import oracle
from cells_amd import lib as hip
from cells_amd.world import StateLayout, build, config0, config1, initial_state
from tests.helpers import conc, relerr, vmem

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


def ensemble(p, factor="supernodal"):
    from cells_amd.ensemble import Ensemble

    return Ensemble(p, "n2", factor=factor)


def oracle_run(p, dt, steps):
    w = oracle.world_struct(p)
    _, sym = oracle.world_symbolics(p)
    y = initial_state(p)
    for k in range(steps):
        y, ok = oracle.step_n2(p, y, dt, sym, world=w)
        assert ok.all(), k
    return y


def fields_close(p, got, want, tol):
    L = StateLayout.of(p)
    err = max(relerr(vmem(p, got), vmem(p, want)), relerr(conc(p, got), conc(p, want)))
    assert err <= tol, err
    for f in range(L.cell_fields):
        a, b = L.cell(got, f), L.cell(want, f)
        assert relerr(a, b, floor=1e-9 * max(np.abs(b).max(), 1e-300)) <= tol, ("cell field", f)
    for f in range(L.edge_fields):
        a, b = L.edge(got, f), L.edge(want, f)
        assert relerr(a, b, floor=1e-9 * max(np.abs(b).max(), 1e-300)) <= tol, ("edge field", f)
    for f in range(L.world_fields):
        a, b = L.world(got, f), L.world(want, f)
        assert relerr(a, b, floor=1e-9 * max(np.abs(b).max(), 1e-300)) <= tol, ("world field", f)
    return err


@pytest.mark.parametrize("cfg,dt,steps", [
    (config1(worlds=3, cells=(300, 2300), seed=12), 5e-3, 100),
    (config0(), 1e-3, 300),
], ids=["calcium", "plain-water"])
def test_n2_trajectory_matches_the_oracle(torch, cfg, dt, steps):
    p = build(cfg)
    e = ensemble(p)
    y0 = initial_state(p)
    say(f"N2 {steps} steps of {dt}: cells {np.diff(p.cell_off).tolist()}")
    got, ok = e.step(y0, dt, nsteps=steps)
    assert bool(ok.all())
    want = oracle_run(p, dt, steps)
    err = fields_close(p, got.cpu().numpy(), want, 1e-9)
    say(f"N2: worst relative error {err:.3e}")
    e.close()


def test_n2_pairwise_factor_and_batched_equals_plain(torch):
    p = build(config1(worlds=4, cells=(60, 200), seed=31))
    y0 = initial_state(p)
    e = ensemble(p, "pairwise")
    got, ok = e.step(y0, 5e-3, nsteps=40)
    assert bool(ok.all())
    fields_close(p, got.cpu().numpy(), oracle_run(p, 5e-3, 40), 1e-9)
    e.close()
    e = ensemble(p)
    batched, ok = e.step(y0, 5e-3, nsteps=20, batched=True)
    plain, ok2 = e.step(y0, 5e-3, nsteps=20, batched=False)
    assert bool(ok.all()) and bool(ok2.all()) and torch.equal(batched, plain)
    e.close()


def test_n2_ledgers_pass_and_the_residual_bound_refuses(torch):
    p = build(config1(worlds=3, cells=(50, 150), seed=5))
    e = ensemble(p)
    y0 = initial_state(p)
    e.step_checks(True, steps_before=0)
    y, ok = e.step(y0, 5e-3, nsteps=30)
    assert bool(ok.all()) and int(e.check_bits().abs().sum()) == 0
    failed, ratios = e.device_checks(y0, y, 30)
    k = hip.LEDGER_RATIOS.index("energy")
    say(f"N2 ledger ratios (worst): {ratios.cpu().numpy().max(axis=0).round(4).tolist()}")
    assert failed.cpu().numpy().tolist() == [0, 0, 0] and float(ratios[:, k].max()) < 1.0
    e.step_checks(False)
    e.stage_tolerance(0.0, 0.0)
    out, ok = e.step(y0, 5e-3, nsteps=2)
    bits = e.check_bits().cpu().numpy()
    assert not bool(ok.any()) and all(b & hip.REFUSE_RESIDUAL for b in bits)
    assert np.array_equal(out.cpu().numpy(), y0)
    e.close()
