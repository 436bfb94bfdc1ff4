"""Complete numerical checkpoints in the reference's container format.

The reference writes a checkpoint as a directory (records/io.py:86-135, `write_snapshot`): a
`state.npz` holding the state's leaves by name (models/surface_tissue.py:31-44 STATE_AXES and the
kinetic and calcium additions, models/surface_kinetic.py:36-44, models/surface_pump_v2.py:44-50;
water's operators/osmotic_water.py:76-83) and a `manifest.json` (schema_version 1, kind
"cells_complete_numerical_checkpoint", the definition and its digest, the clock, the runtime and
source identities, the state file's SHA-256, and the manifest's own digest over its canonical
JSON, definitions/specs.py:25-31), written last so a partial directory cannot be restored. A
batch writes one such directory per world (execution/batch_io's cohort).

Here a world's leaves are cut out of the packed field-major state and named as the reference
names them, with its axes ([species, slot] or [slot]); the held bath is the one extracellular
slot (no faces, no channels: those leaves are empty). Beside them, as the reference stores its
layout's leaves in the same file (records/io.py:107), go the world's bound layout: its per-cell
and per-junction parameters and its stage matrix's symbolic analysis (fill-reducing order, rank,
column starts, entry rows and columns, elimination tree, levels), under `layout_*` names. Reading
refuses what the reference's loader refuses (records/io.py:138-230): a changed manifest, state
file, clock fields or leaf set, a definition other than the target world's, another source or
runtime identity (unless the caller admits it with `continuity=True`, the counterpart of the
reference's continuity request, which relaxes those two comparisons only), and a layout that
differs from the target world's bound one (validate_checkpoint's comparison).
"""

from __future__ import annotations

import hashlib
import json
import os
from pathlib import Path

import numpy as np

from cells_amd.world import Packed, StateLayout

SCHEMA_VERSION = 1
KIND = "cells_complete_numerical_checkpoint"


def canonical_json(value):
    return json.dumps(value, sort_keys=True, separators=(",", ":"), allow_nan=False)


def digest(value):
    return hashlib.sha256(canonical_json(value).encode()).hexdigest()


def _sha256(path):
    h = hashlib.sha256()
    with open(path, "rb") as f:
        for block in iter(lambda: f.read(1 << 20), b""):
            h.update(block)
    return h.hexdigest()


def source_hashes():
    """SHA-256 of the library's sources (the reference's source identity, records/io.py)."""
    here = Path(__file__).resolve().parent / "csrc"
    return {p.name: _sha256(p) for p in sorted(here.iterdir()) if p.suffix in (".hip", ".inc", ".cpp", ".h")}


RUNTIME = {"backend": "rocm", "library": "libcells_hip", "precision": "float64"}


def runtime_identity():
    """The runtime a checkpoint records (the reference's runtime_identity, records/io.py)."""
    return dict(RUNTIME)


def layout_leaves(packed: Packed, world: int) -> dict:
    """World `world`'s bound layout: its parameters (world-local junction ends) and its stage
    matrix's symbolic analysis (numerics/sparse_cholesky.py:69-85's Symbolic arrays, from the
    library's host bind, bit for bit the reference's)."""
    from cells_amd import symbolic

    c0, c1 = int(packed.cell_off[world]), int(packed.cell_off[world + 1])
    e0, e1 = int(packed.edge_off[world]), int(packed.edge_off[world + 1])
    snd = packed.snd[e0:e1] - c0
    rcv = packed.rcv[e0:e1] - c0
    out = {
        "layout_volume_pL": packed.vol0[c0:c1].copy(),
        "layout_membrane_capacitance_F": packed.cap[c0:c1].copy(),
        "layout_membrane_area_permeability": packed.pam[:, c0:c1].copy(),
        "layout_pump_capacity": packed.nak_cap[c0:c1].copy(),
        "layout_calcium_pump_capacity": packed.ca_cap[c0:c1].copy(),
        "layout_amount_reference": packed.n0[:, c0:c1].copy(),
        "layout_initial_cell_charge_C": packed.q0[c0:c1].copy(),
        "layout_water_conductance": packed.lpa[c0:c1].copy(),
        "layout_impermeant_fmol": packed.imperm[c0:c1].copy(),
        "layout_junction_sender": snd.astype(np.int64),
        "layout_junction_receiver": rcv.astype(np.int64),
        "layout_junction_area_permeability": packed.paj[:, e0:e1].copy(),
    }
    key = (c1 - c0, snd.astype(np.int64).tobytes(), rcv.astype(np.int64).tobytes())
    sym = _SYMBOLICS.get(key)
    if sym is None:
        sym = _SYMBOLICS[key] = symbolic.bind(c1 - c0, snd, rcv)  # worlds of one lattice share it
    for k in ("order", "rank", "column_start", "row", "column", "parent", "level"):
        out["layout_operator_" + k] = np.asarray(getattr(sym, k), dtype=np.int64)
    return out


_SYMBOLICS: dict = {}


def leaves(packed: Packed, y, world: int) -> dict:
    """World `world`'s state leaves by the reference's names and axes."""
    L = StateLayout.of(packed)
    y = np.asarray(y)
    S = packed.S
    c0, c1 = int(packed.cell_off[world]), int(packed.cell_off[world + 1])
    e0, e1 = int(packed.edge_off[world]), int(packed.edge_off[world + 1])

    def cell(f):
        return L.cell(y, f)[c0:c1].copy()

    def edge(f):
        return L.edge(y, f)[e0:e1].copy()

    def wfield(f):
        return np.array([L.world(y, f)[world]])

    out = {
        "amount_increment": np.stack([cell(L.dn + s) for s in range(S)]),
        "extracellular_amount_increment": np.zeros((S, 1)),  # the held bath
        "membrane_transferred": np.stack([cell(L.mtr + s) for s in range(S)]),
        "junction_transferred": np.stack([edge(L.jtr + s) for s in range(S)]),
        "face_transferred": np.zeros((S, 0)),
        "reservoir_supplied": np.stack([wfield(L.sup + s) for s in range(S)]),
        "membrane_dissipated_energy_pJ": cell(L.mdiss),
        "junction_dissipated_energy_pJ": edge(L.jdiss),
        "face_dissipated_energy_pJ": np.zeros(0),
        "electrode_charge_transferred_C": wfield(L.echg),
        "reservoir_work_pJ": wfield(L.rwork),
        "electrode_work_pJ": wfield(L.ework),
        "channel_gate": np.zeros((0, 0)),
        "channel_transferred": np.zeros((S, 0)),
        "channel_dissipated_energy_pJ": np.zeros((S, 0)),
        "junction_gate": edge(L.gate),
        "pump_cycles": cell(L.pcyc),
        "pump_dissipated_energy_pJ": cell(L.pdiss),
        "metabolic_work_pJ": cell(L.pwork),
    }
    if packed.calcium:
        out["calcium_pump_cycles"] = cell(L.ccyc)
        out["calcium_pump_dissipated_energy_pJ"] = cell(L.cdiss)
        out["calcium_metabolic_work_pJ"] = cell(L.cwork)
    if packed.water:
        out["volume_increment_pL"] = cell(L.dvol)
        out["extracellular_volume_increment_pL"] = np.zeros(0)  # finite outside nodes only
        out["water_transferred_pL"] = cell(L.wtr)
        out["water_dissipated_energy_pJ"] = cell(L.wdiss)
        out["bath_water_supplied_pL"] = wfield(L.wsup)
        out["bath_water_work_pJ"] = wfield(L.wwork)
    return out


def place(packed: Packed, y, world: int, arrays: dict):
    """Write a world's leaves back into the packed state y (in place)."""
    L = StateLayout.of(packed)
    S = packed.S
    c0, c1 = int(packed.cell_off[world]), int(packed.cell_off[world + 1])
    e0, e1 = int(packed.edge_off[world]), int(packed.edge_off[world + 1])
    for s in range(S):
        L.cell(y, L.dn + s)[c0:c1] = arrays["amount_increment"][s]
        L.cell(y, L.mtr + s)[c0:c1] = arrays["membrane_transferred"][s]
        L.edge(y, L.jtr + s)[e0:e1] = arrays["junction_transferred"][s]
        L.world(y, L.sup + s)[world] = arrays["reservoir_supplied"][s][0]
    for f, name in ((L.mdiss, "membrane_dissipated_energy_pJ"), (L.pcyc, "pump_cycles"),
                    (L.pdiss, "pump_dissipated_energy_pJ"), (L.pwork, "metabolic_work_pJ")):
        L.cell(y, f)[c0:c1] = arrays[name]
    L.edge(y, L.jdiss)[e0:e1] = arrays["junction_dissipated_energy_pJ"]
    L.edge(y, L.gate)[e0:e1] = arrays["junction_gate"]
    for f, name in ((L.echg, "electrode_charge_transferred_C"), (L.rwork, "reservoir_work_pJ"),
                    (L.ework, "electrode_work_pJ")):
        L.world(y, f)[world] = arrays[name][0]
    if packed.calcium:
        for f, name in ((L.ccyc, "calcium_pump_cycles"), (L.cdiss, "calcium_pump_dissipated_energy_pJ"),
                        (L.cwork, "calcium_metabolic_work_pJ")):
            L.cell(y, f)[c0:c1] = arrays[name]
    if packed.water:
        for f, name in ((L.dvol, "volume_increment_pL"), (L.wtr, "water_transferred_pL"),
                        (L.wdiss, "water_dissipated_energy_pJ")):
            L.cell(y, f)[c0:c1] = arrays[name]
        L.world(y, L.wsup)[world] = arrays["bath_water_supplied_pL"][0]
        L.world(y, L.wwork)[world] = arrays["bath_water_work_pJ"][0]


def world_spec(packed: Packed, world: int, solver: str) -> dict:
    """The definition a world's checkpoint carries: its configuration and its own size."""
    import dataclasses

    cfg = {k: (list(v) if isinstance(v, tuple) else v)
           for k, v in dataclasses.asdict(packed.config).items()}
    return {"model": "cells.model.resolved_pump_tissue@3", "solver": solver, "config": cfg,
            "world": int(world), "cells": int(packed.cell_off[world + 1] - packed.cell_off[world]),
            "edges": int(packed.edge_off[world + 1] - packed.edge_off[world])}


def write_snapshot(directory, packed: Packed, y, world: int, time: float, step_index: int,
                   solver: str = "ros2") -> Path:
    """One world's complete checkpoint directory (records/io.py:86-135's layout)."""
    path = Path(directory)
    path.mkdir(parents=True, exist_ok=False)
    np.savez(path / "state.npz", **leaves(packed, y, world), **layout_leaves(packed, world))
    spec = world_spec(packed, world, solver)
    manifest = {
        "schema_version": SCHEMA_VERSION,
        "kind": KIND,
        "spec": spec,
        "spec_sha256": digest(spec),
        "clock": {"time": float(time), "step_index": int(step_index), "event_index": 0},
        "random_state": None,
        "runtime": runtime_identity(),
        "source_hashes": source_hashes(),
        "state_sha256": _sha256(path / "state.npz"),
    }
    manifest["manifest_sha256"] = digest(manifest)
    with open(path / "manifest.json", "x") as f:  # written last
        json.dump(manifest, f, indent=2, sort_keys=True, allow_nan=False)
        f.write("\n")
    return path


def read_snapshot(directory, packed: Packed, y, world: int, solver: str = "ros2", continuity: bool = False):
    """Restore one world's checkpoint into y (in place); returns its clock. Refuses (ValueError,
    the reference's ConfigError messages) a changed manifest or state file, unknown clock fields,
    a definition other than world_spec(packed, world, solver), another source or runtime identity
    (continuity=True admits those two), another leaf set, and a layout other than the world's."""
    path = Path(directory)
    with open(path / "manifest.json") as f:
        manifest = json.load(f)
    fields = {"schema_version", "kind", "spec", "spec_sha256", "clock", "random_state", "runtime",
              "source_hashes", "state_sha256", "manifest_sha256"}
    if not isinstance(manifest, dict) or set(manifest) != fields:
        raise ValueError("Invalid checkpoint manifest fields")
    fingerprint = manifest.pop("manifest_sha256")
    if digest(manifest) != fingerprint:
        raise ValueError("Checkpoint manifest checksum mismatch")
    if type(manifest["schema_version"]) is not int or manifest["schema_version"] != SCHEMA_VERSION:
        raise ValueError("Unsupported checkpoint schema")
    if manifest["kind"] != KIND:
        raise ValueError("Not a complete Cells numerical checkpoint")
    if manifest["random_state"] is not None:
        raise ValueError("These deterministic models have no random state")
    if not continuity:
        if manifest["source_hashes"] != source_hashes():
            raise ValueError("Checkpoint model/engine source differs from this installation")
        if manifest["runtime"] != runtime_identity():
            raise ValueError("Checkpoint runtime differs; cross-runtime migration is not supported")
    spec = manifest["spec"]
    if digest(spec) != manifest["spec_sha256"]:
        raise ValueError("Checkpoint specification checksum mismatch")
    n = int(packed.cell_off[world + 1] - packed.cell_off[world])
    E = int(packed.edge_off[world + 1] - packed.edge_off[world])
    if not isinstance(spec, dict) or spec.get("cells") != n or spec.get("edges") != E:
        raise ValueError("Checkpoint world size differs from the target world")
    if spec != world_spec(packed, world, solver):
        raise ValueError("Checkpoint definition differs from the target world")
    if _sha256(path / "state.npz") != manifest["state_sha256"]:
        raise ValueError("Checkpoint state checksum mismatch")
    clock = manifest["clock"]
    if not isinstance(clock, dict) or set(clock) != {"time", "step_index", "event_index"}:
        raise ValueError("Invalid checkpoint clock fields")
    if isinstance(clock["time"], bool) or not isinstance(clock["time"], (int, float)):
        raise ValueError("Invalid checkpoint time")
    state_names = set(leaves(packed, np.zeros(StateLayout.of(packed).total), world))
    layout = layout_leaves(packed, world)
    with np.load(path / "state.npz", allow_pickle=False) as archive:
        if set(archive.files) != state_names | set(layout):
            raise ValueError("Invalid checkpoint state fields")
        for k, v in layout.items():
            got = archive[k]
            if got.shape != v.shape or got.dtype != v.dtype or not np.array_equal(got, v):
                raise ValueError(f"Checkpoint layout differs from the target world ({k})")
        place(packed, y, world, {k: archive[k] for k in state_names})
    return clock


def write_cohort(directory, packed: Packed, y, time, step_index, solver: str = "ros2") -> Path:
    """Every world's checkpoint under one directory (world_00000, ...); `time` and `step_index`
    per world (arrays) or shared (scalars)."""
    root = Path(directory)
    root.mkdir(parents=True, exist_ok=False)
    yh = y.cpu().numpy() if hasattr(y, "cpu") else np.asarray(y)
    times = np.broadcast_to(np.asarray(time, dtype=float), (packed.W,))
    steps = np.broadcast_to(np.asarray(step_index, dtype=np.int64), (packed.W,))
    for k in range(packed.W):
        write_snapshot(root / f"world_{k:05d}", packed, yh, k, float(times[k]), int(steps[k]), solver)
    return root


def read_cohort(directory, packed: Packed, solver: str = "ros2", continuity: bool = False):
    """(state, times [W], step counts [W]) restored from write_cohort's directory."""
    root = Path(directory)
    y = np.zeros(StateLayout.of(packed).total)
    times = np.zeros(packed.W)
    steps = np.zeros(packed.W, dtype=np.int64)
    names = sorted(os.listdir(root))
    if names != [f"world_{k:05d}" for k in range(packed.W)]:
        raise ValueError("Cohort directory does not hold one checkpoint per world")
    for k in range(packed.W):
        clock = read_snapshot(root / names[k], packed, y, k, solver, continuity)
        times[k], steps[k] = clock["time"], clock["step_index"]
    return y, times, steps
