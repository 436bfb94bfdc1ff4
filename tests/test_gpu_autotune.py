"""GPU checks of the plan autotuner (cells_amd.autotune): it times each candidate on the caller's
worlds and leaves the fastest in force; plans that differ only in how far supernodes are joined
(explicit zeros) or where the grid tree is cut give the same steps to rounding (1e-12 of each
field's scale)."""

import numpy as np
import pytest

from cells_amd.autotune import autotune_ensemble, autotune_field, plan_options, set_plan_options
from cells_amd.world import build, config1, initial_state

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as t

    if not t.cuda.is_available():
        pytest.skip("no GPU")
    return t


@pytest.fixture()
def restore():
    before = plan_options()
    yield before
    set_plan_options(**before)


def test_autotune_picks_and_applies_the_fastest_lattice_plan(torch, restore):
    from cells_amd.ensemble import Ensemble

    p = build(config1(worlds=12, cells=(256, 1500), seed=31))
    cands = [{"relax_zeros": 0}, {"relax_zeros": 64}]
    res = autotune_ensemble(p, "ros2", 5e-3, steps=2, candidates=cands)
    assert [c for c, _ in res.table] == cands
    assert res.best in cands and res.ms_per_step == min(m for _, m in res.table)
    assert plan_options()["relax_zeros"] == res.best["relax_zeros"]
    outs = []
    for c in cands:
        set_plan_options(**c)
        e = Ensemble(p, "ros2")
        y, ok = e.step(initial_state(p), 5e-3, nsteps=4)
        assert bool(ok.all())
        outs.append(y.cpu().numpy())
        e.close()
    scale = np.maximum(np.abs(outs[0]).max(), 1e-300)
    assert np.max(np.abs(outs[0] - outs[1])) <= 1e-12 * scale


def test_autotune_field_plan(torch, restore):
    from cells_amd.field import FieldConfig, FieldEnsemble, FieldLayout, build_field, initial_field_state

    fp = build_field(FieldConfig(worlds=2, cells=36, row_width=6, side=40, spacing=5, seed=32))
    cands = [{"grid_relax_zeros": 0, "grid_subtrees": 8}, {"grid_relax_zeros": 3, "grid_subtrees": 32}]
    res = autotune_field(fp, steps=3, refresh_every=5, candidates=cands)
    assert res.best in cands
    outs = []
    for c in cands:
        set_plan_options(**c)
        e = FieldEnsemble(fp, refresh_every=5)
        y, ok = e.step(initial_field_state(fp), 1e-3, nsteps=6)
        assert bool(ok.all())
        outs.append(y.cpu().numpy())
        e.close()
    L = FieldLayout.of(fp)
    for name, view, count in (("cell", L.cell, L.cell_fields), ("node", L.node, L.node_fields)):
        for f in range(count):
            a, b = view(outs[0], f), view(outs[1], f)
            scale = max(float(np.max(np.abs(b))), 1e-300)
            assert float(np.max(np.abs(a - b))) <= 1e-9 * scale, (name, f)
