"""The windowed driver's host logic (no GPU): window boundaries and policy validation."""

import pytest

from cells_amd.windows import WindowPolicy, _boundaries


def test_windows_end_at_every_window_and_observation_step():
    assert _boundaries(0, 30, 7, 5) == [5, 7, 10, 14, 15, 20, 21, 25, 28, 30]
    assert _boundaries(0, 10, 4, 0) == [4, 8, 10]
    assert _boundaries(14, 6, 7, 0) == [20]  # resumed runs count windows from their start
    assert _boundaries(14, 6, 7, 5) == [15, 20]


@pytest.mark.parametrize("kw", [dict(window_steps=0), dict(window_steps=5000), dict(window_steps=4, observe_every=-1),
                                dict(window_steps=4, refusals="never")])
def test_bad_policies_are_refused(kw):
    with pytest.raises(ValueError):
        WindowPolicy(**kw)
