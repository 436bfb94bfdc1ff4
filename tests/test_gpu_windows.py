"""GPU checks of the windowed driver (cells_amd.windows, the execution layer's vector windows over
cells_step with the ledger checks as each step's epilogue).

- Windows (7 steps, observations every 5) end bit for bit where one 30-step call ends, and each
  observation is the cell potentials of the state a plain run reaches at that step.
- A run resumed from a mid-run checkpoint ends bit for bit where the uninterrupted run ends.
- refusals="world": a world whose start state leaves the domain refuses step 0 and stays put, the
  others run on; refusals="batch" raises, naming the world and step.
- The run summary carries the fields of the ionic ensemble's (ensemble/ionic.py:182-205).
"""

import numpy as np
import pytest

from cells_amd import lib as hip
from cells_amd.windows import NumericalError, WindowPolicy, cell_potentials, run_windowed
from cells_amd.world import StateLayout, build, config1, initial_state

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as t

    if not t.cuda.is_available():
        pytest.skip("no GPU")
    return t


@pytest.fixture(scope="module")
def ens(torch):
    from cells_amd.ensemble import Ensemble

    p = build(config1(worlds=6, cells=(64, 300), seed=21))
    e = Ensemble(p, "ros2")
    yield e
    e.close()


def test_windows_equal_one_call_and_observe_the_plain_run(torch, ens):
    p = ens.packed
    y0 = initial_state(p)
    whole, ok = ens.step(y0, 5e-3, nsteps=30)
    assert bool(ok.all())
    res = run_windowed(ens, y0, 5e-3, 30, WindowPolicy(7, observe_every=5))
    assert torch.equal(res.checkpoint.state, whole)
    assert res.checkpoint.step_index.tolist() == [30] * p.W
    y = ens.to_device(y0)
    for k in range(5, 31, 5):
        y, _ = ens.step(y, 5e-3, nsteps=5)
        assert np.array_equal(res.observations[k], cell_potentials(p, y))
    assert not res.refused
    s = res.summary(p)
    assert s["integrated_world_steps"] == 30 * p.W and abs(s["through_time_s"] - 0.15) < 1e-12
    assert len(s["final_voltage_V"]) == p.W
    for k in ("model", "physical_model", "world_id_start", "deterministic_replicas", "compile_seconds",
              "execute_seconds", "resources", "elapsed_seconds", "backend", "precision", "world_batching"):
        assert k in s, k  # ensemble/ionic.py:182-205's fields
    assert s["resources"]["executed_world_steps"] == 30 * p.W
    acc = s["ionic_accounting"]
    assert set(acc) == set(hip.LEDGER_RATIOS) and all(v <= 1.0 for v in acc.values()), acc


def test_resume_from_a_window_checkpoint(torch, ens):
    p = ens.packed
    y0 = initial_state(p)
    saved = {}
    full = run_windowed(ens, y0, 5e-3, 20, WindowPolicy(7), sink=lambda i, cp: saved.setdefault(i, cp))
    cp = saved[2]  # after 14 steps
    assert cp.step_index.tolist() == [14] * p.W
    rest = run_windowed(ens, None, 5e-3, 6, WindowPolicy(7), checkpoint=cp)
    assert torch.equal(rest.checkpoint.state, full.checkpoint.state)
    assert rest.checkpoint.step_index.tolist() == [20] * p.W


def test_a_refusing_world_stops_alone_or_ends_a_batch_run(torch, ens):
    p = ens.packed
    L = StateLayout.of(p)
    y0 = initial_state(p)
    bad = 3
    e0, e1 = int(p.edge_off[bad]), int(p.edge_off[bad + 1])
    L.edge(y0, L.gate)[e0:e1] = 1.5  # a gate above 1: outside the domain
    res = run_windowed(ens, y0, 5e-3, 12, WindowPolicy(5, refusals="world"))
    assert list(res.refused) == [bad]
    assert res.refused[bad].step == 0
    assert res.checkpoint.step_index.tolist() == [12 if w != bad else 0 for w in range(p.W)]
    got = res.checkpoint.state.cpu().numpy()
    assert np.array_equal(L.edge(got, L.gate)[e0:e1], L.edge(y0, L.gate)[e0:e1])
    with pytest.raises(NumericalError, match=f"world {bad} refused step 0"):
        run_windowed(ens, y0, 5e-3, 12, WindowPolicy(5, refusals="batch"))


def test_resume_from_checkpoint_files_is_bitwise(torch, ens, tmp_path):
    """A run written to checkpoint directories (the reference's container) and read back into a
    new start state continues bit for bit."""
    from cells_amd.snapshot import read_cohort, write_cohort

    p = ens.packed
    y0 = initial_state(p)
    whole, _ = ens.step(y0, 5e-3, nsteps=20)
    half, _ = ens.step(y0, 5e-3, nsteps=10)
    write_cohort(tmp_path / "cohort", p, half, 10 * 5e-3, 10)
    back, times, steps = read_cohort(tmp_path / "cohort", p)
    assert steps.tolist() == [10] * p.W
    rest, _ = ens.step(back, 5e-3, nsteps=10)
    assert torch.equal(rest, whole)
