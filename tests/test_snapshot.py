"""Checkpoints in the reference's container (records/io.py:86-230): a world's leaves by the
reference's names and axes, the manifest's checksums, and the refusals of its loader."""

import json

import numpy as np
import pytest

from cells_amd.snapshot import leaves, read_cohort, read_snapshot, write_cohort, write_snapshot
from cells_amd.world import StateLayout, build, config0, config1


@pytest.mark.parametrize("cfg", [config1(worlds=3, cells=(20, 40), seed=2), config0(worlds=2)],
                         ids=["calcium", "water"])
def test_cohort_round_trip_is_exact(tmp_path, cfg):
    p = build(cfg)
    rng = np.random.default_rng(1)
    y = rng.normal(size=StateLayout.of(p).total)
    write_cohort(tmp_path / "c", p, y, time=[0.1 * k for k in range(p.W)], step_index=[7] * p.W)
    got, times, steps = read_cohort(tmp_path / "c", p)
    assert np.array_equal(got, y)
    assert times.tolist() == [0.1 * k for k in range(p.W)] and steps.tolist() == [7] * p.W


def test_leaves_carry_the_reference_names_and_axes():
    p = build(config1(worlds=2, cells=(30, 30), seed=3))
    lv = leaves(p, np.zeros(StateLayout.of(p).total), 1)
    n, E, S = 30, int(p.edge_off[2] - p.edge_off[1]), p.S
    assert lv["amount_increment"].shape == (S, n)
    assert lv["junction_transferred"].shape == (S, E)
    assert lv["junction_gate"].shape == (E,)
    assert lv["reservoir_supplied"].shape == (S, 1)
    assert lv["calcium_pump_cycles"].shape == (n,)
    assert lv["face_transferred"].shape == (S, 0)


def test_the_loader_refuses_changed_files(tmp_path):
    p = build(config1(worlds=1, cells=(25, 25), seed=4))
    y = np.arange(StateLayout.of(p).total, dtype=float)
    d = write_snapshot(tmp_path / "w", p, y, 0, 0.5, 100)
    out = np.zeros_like(y)
    assert read_snapshot(d, p, out, 0) == {"time": 0.5, "step_index": 100, "event_index": 0}
    assert np.array_equal(out, y)
    m = json.loads((d / "manifest.json").read_text())
    m["clock"]["step_index"] = 101
    (d / "manifest.json").write_text(json.dumps(m))
    with pytest.raises(ValueError, match="manifest checksum"):
        read_snapshot(d, p, out, 0)
    d2 = write_snapshot(tmp_path / "w2", p, y, 0, 0.5, 100)
    raw = bytearray((d2 / "state.npz").read_bytes())
    raw[-20] ^= 1
    (d2 / "state.npz").write_bytes(bytes(raw))
    with pytest.raises(ValueError, match="state checksum"):
        read_snapshot(d2, p, out, 0)
    other = build(config1(worlds=1, cells=(26, 26), seed=4))
    d3 = write_snapshot(tmp_path / "w3", p, y, 0, 0.5, 100)
    with pytest.raises(ValueError, match="size"):
        read_snapshot(d3, other, np.zeros(StateLayout.of(other).total), 0)
