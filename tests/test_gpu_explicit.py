"""The explicit tableaux on the GPU (CELLS_METHOD_ARS222, CELLS_METHOD_RK4), through the C-ABI,
against the oracle's ref_step_ars222 / ref_step_rk4:

- ARS(2,2,2) on calcium worlds and on the plain world with water, RK4 on calcium worlds: every
  field at 1e-9 relative after 60 steps;
- batched equals plain bit for bit; the fused ledger checks pass with each method's energy bound
  (RK4's the absolute floor alone);
- RK4 refuses worlds with water at create (the reference's water RK4 couples the water rates).
"""

import sys

import numpy as np
import pytest

import oracle
from cells_amd import lib as hip
from cells_amd.world import build, config0, config1, initial_state
from tests.test_gpu_n2 import fields_close

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


STEP = {"ars222": lambda *a, **k: oracle.step_ars222(*a, **k), "rk4": lambda *a, **k: oracle.step_rk4(*a, **k)}


@pytest.mark.parametrize("method,cfg,dt", [
    ("ars222", config1(worlds=3, cells=(200, 900), seed=8), 2e-4),
    ("ars222", config0(), 2e-4),
    ("rk4", config1(worlds=3, cells=(200, 900), seed=8), 4e-4),
], ids=["ars222-calcium", "ars222-water", "rk4-calcium"])
def test_trajectory_matches_the_oracle(torch, method, cfg, dt):
    p = build(cfg)
    e = ensemble(p, method)
    y0 = initial_state(p)
    say(f"{method}: 60 steps of {dt}")
    got, ok = e.step(y0, dt, nsteps=60)
    assert bool(ok.all())
    w = oracle.world_struct(p)
    want = y0
    for k in range(60):
        want, okr = STEP[method](p, want, dt, world=w)
        assert okr.all(), k
    err = fields_close(p, got.cpu().numpy(), want, 1e-9)
    say(f"{method}: worst relative error {err:.3e}")
    e.close()


@pytest.mark.parametrize("method,dt", [("ars222", 2e-4), ("rk4", 4e-4)])
def test_batched_equals_plain_and_the_ledgers_pass(torch, method, dt):
    p = build(config1(worlds=4, cells=(60, 200), seed=31))
    e = ensemble(p, method)
    y0 = initial_state(p)
    batched, ok = e.step(y0, dt, nsteps=10, batched=True)
    plain, ok2 = e.step(y0, dt, nsteps=10, batched=False)
    assert bool(ok.all()) and bool(ok2.all()) and torch.equal(batched, plain)
    e.step_checks(True, steps_before=0)
    y, ok = e.step(y0, dt, nsteps=30)
    bits = e.check_bits().cpu().numpy()
    say(f"{method}: fused checks bits {bits.tolist()}")
    assert bool(ok.all()) and not bits.any()
    e.step_checks(False)
    e.close()


def test_rk4_refuses_water_worlds(torch):
    from cells_amd.ensemble import Ensemble

    with pytest.raises(hip.CellsError, match="RK4 covers worlds without water") as err:
        Ensemble(build(config0()), "rk4")
    assert err.value.code == hip.CELLS_EINVAL
