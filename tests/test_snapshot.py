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


def test_the_loader_refuses_another_definition_source_or_runtime(tmp_path, monkeypatch):
    from cells_amd import snapshot

    p = build(config1(worlds=1, cells=(25, 25), seed=4))
    y = np.arange(StateLayout.of(p).total, dtype=float)
    out = np.zeros_like(y)
    # the same lattice under another configuration (another seed: other parameters)
    other = build(config1(worlds=1, cells=(25, 25), seed=5))
    d = write_snapshot(tmp_path / "a", p, y, 0, 0.5, 100)
    with pytest.raises(ValueError, match="definition differs"):
        read_snapshot(d, other, out, 0)
    with pytest.raises(ValueError, match="definition differs"):
        read_snapshot(d, p, out, 0, solver="euler")
    # another build of the library's sources
    real = snapshot.source_hashes()
    monkeypatch.setattr(snapshot, "source_hashes", lambda: {**real, "cells_hip.hip": "0" * 64})
    with pytest.raises(ValueError, match="source differs"):
        read_snapshot(d, p, out, 0)
    assert read_snapshot(d, p, out, 0, continuity=True)["step_index"] == 100  # admitted explicitly
    monkeypatch.setattr(snapshot, "source_hashes", lambda: real)
    monkeypatch.setattr(snapshot, "runtime_identity", lambda: {**snapshot.RUNTIME, "precision": "float32"})
    with pytest.raises(ValueError, match="runtime differs"):
        read_snapshot(d, p, out, 0)


def test_the_checkpoint_carries_the_world_layout_and_its_plan(tmp_path):
    from cells_amd import symbolic
    from cells_amd.snapshot import layout_leaves

    p = build(config1(worlds=2, cells=(30, 40), seed=6))
    y = np.arange(StateLayout.of(p).total, dtype=float)
    d = write_snapshot(tmp_path / "w", p, y, 1, 0.0, 0)
    c0, c1 = int(p.cell_off[1]), int(p.cell_off[2])
    e0, e1 = int(p.edge_off[1]), int(p.edge_off[2])
    sym = symbolic.bind(c1 - c0, p.snd[e0:e1] - c0, p.rcv[e0:e1] - c0)
    with np.load(d / "state.npz", allow_pickle=False) as a:
        for k in ("order", "rank", "column_start", "row", "column", "parent", "level"):
            assert np.array_equal(a["layout_operator_" + k], getattr(sym, k))
        assert np.array_equal(a["layout_junction_area_permeability"], p.paj[:, e0:e1])
    # a layout leaf changed inside a consistent directory (manifest re-signed) is refused
    import json

    from cells_amd.snapshot import _sha256, digest

    with np.load(d / "state.npz", allow_pickle=False) as a:
        arrays = {k: a[k] for k in a.files}
    arrays["layout_operator_order"] = arrays["layout_operator_order"][::-1].copy()
    np.savez(d / "state.npz", **arrays)
    m = json.loads((d / "manifest.json").read_text())
    m.pop("manifest_sha256")
    m["state_sha256"] = _sha256(d / "state.npz")
    m["manifest_sha256"] = digest(m)
    (d / "manifest.json").write_text(json.dumps(m))
    with pytest.raises(ValueError, match="layout differs"):
        read_snapshot(d, p, np.zeros_like(y), 1)
    assert set(layout_leaves(p, 1)) <= set(arrays)
