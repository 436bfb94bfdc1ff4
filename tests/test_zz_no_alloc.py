"""cells_step_vjp_noalloc through the C-ABI: the adjoint of one ROS2 step on the library's fixed
device workspace equals cells_step_vjp's (which allocates its own) bit for bit, on two calls in a
row, the second on a smaller ensemble (the fixed array keeps the first call's contents); an
ensemble whose workspace exceeds the array is refused with CELLS_EINVAL before any launch."""

import sys

import numpy as np
import pytest

from cells_amd import lib as hip
from cells_amd.world import StateLayout, build, config1, initial_state

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


def three_ion(worlds, cells, seed):
    return build(config1(worlds=worlds, cells=cells, species=("Na", "K", "Cl"),
                         inside_mM=(12.0, 140.0, 10.0), bath_mM=(145.0, 5.0, 110.0),
                         permeability=(1e-10, 5e-9, 1e-9), seed=seed))


def test_noalloc_adjoint_equals_the_allocating_one_on_two_calls(torch):
    from cells_amd.ensemble import Ensemble

    for k, p in enumerate((three_ion(3, (200, 400), 8), three_ion(2, (12, 30), 5))):
        e = Ensemble(p, "ros2")
        L = StateLayout.of(p)
        y, ok = e.step(initial_state(p), 5e-3, nsteps=4)
        assert bool(ok.all())
        rng = np.random.default_rng(3 + k)
        ybar = np.zeros(L.total)
        for s in range(p.S):
            L.cell(ybar, L.dn + s)[:] = rng.normal(size=p.ncell)
        L.edge(ybar, L.gate)[:] = rng.normal(size=p.nedge)
        say(f"no-alloc adjoint, call {k + 1}: {p.W} worlds, {p.ncell} cells")
        got = e.step_vjp(y, ybar, 5e-3, noalloc=True)
        want = e.step_vjp(y, ybar, 5e-3)
        assert bool(got[4].all()) and bool(want[4].all())
        for a, b in zip(got, want):
            assert torch.equal(a, b)
        e.close()


def test_noalloc_adjoint_refuses_an_ensemble_above_its_workspace(torch):
    from cells_amd.ensemble import Ensemble

    p = three_ion(300, (1024, 1024), 3)  # 13 doubles per cell and 12 per junction: about 11.1 M doubles
    e = Ensemble(p, "ros2")
    y = initial_state(p)
    with pytest.raises(hip.CellsError, match="64 MiB") as err:
        e.step_vjp(y, np.zeros_like(y), 5e-3, noalloc=True)
    assert err.value.code == hip.CELLS_EINVAL
    e.close()
