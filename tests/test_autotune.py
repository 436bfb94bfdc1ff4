"""Plan options (cells_set_plan_options / cells_get_plan_options) on the host: defaults, partial
updates, range refusals, restoring."""

import pytest

from cells_amd.autotune import plan_options, set_plan_options
from cells_amd.lib import CellsError


def test_plan_options_set_get_and_refuse():
    before = plan_options()
    assert before == {"relax_rows": 128, "relax_zeros": 64, "grid_relax_zeros": 3, "grid_subtrees": 32,
                      "widen_groups": 1, "largest_first": 1}
    try:
        now = set_plan_options(relax_zeros=8, grid_subtrees=48)
        assert now == dict(before, relax_zeros=8, grid_subtrees=48) == plan_options()
        for bad in ({"relax_rows": 0}, {"relax_rows": 129}, {"relax_zeros": -1}, {"grid_subtrees": 0},
                    {"widen_groups": 2}, {"largest_first": -1}):
            with pytest.raises(CellsError, match="plan options out of range"):
                set_plan_options(**bad)
        assert plan_options() == now
        with pytest.raises(ValueError, match="unknown plan options"):
            set_plan_options(worlds_per_wave=2)
    finally:
        set_plan_options(**before)
    assert plan_options() == before
