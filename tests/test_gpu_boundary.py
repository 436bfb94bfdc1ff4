"""The C-ABI's buffer contract on the GPU: cells_step on the caller's output buffer.

In place (y_out == y_in) and out of place give the same bits, batched and plain; output buffers at
more addresses than the captured-graph cache holds give the same bits; a recorded step in place
records its input (the reverse from the record equals the recomputing reverse)."""

import sys

import numpy as np
import pytest

from cells_amd.world import build, config0, config1, initial_state

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


@pytest.mark.parametrize("method,cfg,dt", [("ros2", config1(worlds=5, cells=(20, 120), seed=12), 5e-3),
                                           ("euler", config0(worlds=3, cells=(9, 30), seed=13), 1e-3)])
def test_in_place_steps_equal_out_of_place(torch, method, cfg, dt):
    from cells_amd.ensemble import Ensemble

    p = build(cfg)
    e = Ensemble(p, method)
    y0 = e.to_device(initial_state(p))
    for batched in (True, False):
        ref, ok_ref = e.step(y0, dt, nsteps=7, batched=batched)
        y = y0.clone()
        say(f"{method} in place, batched {batched}")
        got, ok = e.step(y, dt, nsteps=7, out=y, batched=batched)
        assert got.data_ptr() == y.data_ptr()
        assert torch.equal(got, ref) and torch.equal(ok, ok_ref) and bool(ok.all())
        # odd and even counts in place, one after another
        y = y0.clone()
        e.step(y, dt, nsteps=3, out=y, batched=batched)
        e.step(y, dt, nsteps=4, out=y, batched=batched)
        assert torch.equal(y, ref)
    e.close()


def test_many_output_addresses_give_the_same_steps(torch):
    from cells_amd.ensemble import Ensemble

    p = build(config1(worlds=4, cells=(30, 90), seed=14))
    e = Ensemble(p, "ros2")
    y0 = e.to_device(initial_state(p))
    ref, _ = e.step(y0, 5e-3, nsteps=5)
    outs = [torch.empty_like(y0) for _ in range(6)]  # more than the four cached graphs
    for rep in range(2):
        for o in outs:
            got, ok = e.step(y0, 5e-3, nsteps=5, out=o)
            assert bool(ok.all()) and torch.equal(got, ref)
    e.close()


def test_a_recorded_step_in_place_records_its_input(torch):
    from cells_amd.ensemble import Ensemble
    from cells_amd import lib as hip
    import ctypes

    p = build(config1(worlds=3, cells=(20, 49), seed=8, species=("Na", "K", "Cl"), inside_mM=(12.0, 140.0, 10.0),
                      bath_mM=(145.0, 5.0, 110.0), permeability=(1e-10, 5e-9, 1e-9)))
    e = Ensemble(p, "ros2")
    e.reserve_tape(1)
    y0, _ = e.step(e.to_device(initial_state(p)), 5e-3, nsteps=3)
    ref, _ = e.step(y0, 5e-3, nsteps=1)
    y = y0.clone()
    ok = torch.empty(p.W, dtype=torch.int32, device=y.device)
    ctl = hip.ControlStruct(1, 1, 1, 1, 1)
    v = lambda x: ctypes.c_void_p(x.data_ptr())  # noqa: E731
    hip.check(hip.load().cells_step_record(e._h, e._stream(), v(y), v(y), 5e-3, ctypes.byref(ctl), 0, v(ok)))
    assert torch.equal(y, ref) and bool(ok.all())
    ybar = torch.as_tensor(np.random.default_rng(1).normal(size=y.numel()), device=y.device)
    a = e.step_vjp(y0, ybar, 5e-3)
    b = e.step_vjp_recorded(0, ybar, 5e-3)
    for u, w in zip(a, b):
        assert torch.equal(u, w)
    e.close()
