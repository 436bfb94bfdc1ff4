"""Windowed runs of a device ensemble: the execution layer's vector windows over the HIP step.

The reference runs a batch of worlds in windows of steps (execution/vector_windows.py:872-1088,
`_vector_window_rounds`): each window is one device launch sequence; between windows the host
reads each world's first failing step and failure bits, records observations and checkpoints,
and, under refusals="world", drops the worlds that refused (the others go on) or, under
refusals="batch", stops with an error naming the first refusal. The ionic ensemble summarises a
run (ensemble/ionic.py:163-205): the time reached, the integrated world steps, each world's final
membrane voltages, the ledger accounting (the largest residual-to-bound ratios), phase timings.

Here a window is one `cells_step` call (a replayed step graph) with the ledger checks as each
step's epilogue (`cells_step_checks`). A world's refusal is located at step granularity by
replaying that world alone from the window's start, one step per call, as the reference
re-derives a refusal from its window's start checkpoint (`_check_refusal`). Observations (each
world's cell potentials) are taken at every `observe_every` steps; windows end at observation
steps so no observation falls inside one.
"""

from __future__ import annotations

import dataclasses
import time

import numpy as np

from cells_amd import lib as _lib
from cells_amd.world import F, StateLayout


class NumericalError(RuntimeError):
    """A refused step under refusals="batch" (execution/errors.py of the reference)."""


@dataclasses.dataclass
class WindowPolicy:
    """Steps per window (execution/vector_windows.py's WindowPolicy), observation stride and the
    refusal policy ("batch": a refusal ends the run; "world": the refusing world stops)."""

    window_steps: int
    observe_every: int = 0
    refusals: str = "batch"

    def __post_init__(self):
        if type(self.window_steps) is not int or not 1 <= self.window_steps <= 4096:
            raise ValueError("window_steps must be an integer in [1, 4096]")
        if type(self.observe_every) is not int or self.observe_every < 0:
            raise ValueError("observe_every must be a nonnegative integer")
        if self.refusals not in ("batch", "world"):
            raise ValueError('refusals must be "batch" or "world"')


@dataclasses.dataclass
class WorldRefusal:
    world: int
    step: int  # the refused step's index (0-based, from the run's start state)
    bits: int  # ledger check bits (cells_device_checks' meanings; 0: a domain refusal)


@dataclasses.dataclass
class Checkpoint:
    """Every world's state at a window boundary, with each world's accepted step count (a world
    that refused stays at its last accepted state and count)."""

    state: object  # device tensor
    step_index: np.ndarray  # [W] int64
    dt: float

    @property
    def time(self):
        return self.step_index * self.dt


@dataclasses.dataclass
class RunResult:
    checkpoint: Checkpoint
    refused: dict  # world -> WorldRefusal
    observations: dict  # absolute step -> [ncell] cell potentials (V)
    accounting: dict  # the largest ratio of each ledger's residual to its bound over the worlds
    timings: dict
    resources: dict = dataclasses.field(default_factory=dict)
    world_id_start: int = 0

    def summary(self, packed):
        """ensemble/ionic.py:182-205's summary fields that this run defines."""
        ncell = np.diff(packed.cell_off)
        final = self.observations.get("final")
        return {
            "model": "cells.model.ionic_worlds@1",
            "physical_model": "cells.model.resolved_pump_tissue@3",
            "worlds": int(packed.W),
            "world_id_start": int(self.world_id_start),
            "deterministic_replicas": True,
            "through_time_s": float(self.checkpoint.time.min()),
            "integrated_world_steps": int(self.checkpoint.step_index.sum()),
            "final_voltage_V": [final[int(a):int(a) + int(n)].tolist()
                                for a, n in zip(packed.cell_off[:-1], ncell)],
            "final_voltage_axis": "membrane_patch",
            "ionic_accounting": self.accounting,
            "refused_worlds": sorted(self.refused),
            "compile_seconds": float(self.timings.get("compile_seconds", 0.0)),
            "execute_seconds": float(self.timings.get("execute_seconds", 0.0)),
            "phase_timings": self.timings,
            "resources": self.resources,
            "elapsed_seconds": float(self.timings.get("elapsed_seconds", 0.0)),
            "backend": "gpu",
            "precision": "float64",
            # every world in one captured launch sequence per step: the vectorised windows' form
            "world_batching": "vmap",
        }


def cell_potentials(packed, y):
    """Each cell's membrane potential (V): the bath plus its charge over its capacitance."""
    L = StateLayout.of(packed)
    yh = y.cpu().numpy() if hasattr(y, "cpu") else np.asarray(y)
    q = packed.q0 + 1e-15 * F * sum(packed.valence[s] * L.cell(yh, L.dn + s) for s in range(packed.S))
    return packed.config.bath_V + q / packed.cap


def _boundaries(start, steps, window, observe):
    """Window ends (absolute steps) from `start` to `start + steps`: every `window` steps and at
    every multiple of `observe`."""
    end = start + steps
    cuts = set(range(start + window, end, window)) | {end}
    if observe:
        first = (start // observe + 1) * observe
        cuts |= set(range(first, end, observe))
    return sorted(c for c in cuts if c > start)


def run_windowed(ensemble, y0, dt, steps, policy: WindowPolicy, ctl=None, checkpoint=None, sink=None,
                 material_atol=1e-12, world_id_start=0):
    """`steps` steps of every world in windows; from `checkpoint` when given (its state and step
    counts), else from y0 at step 0. sink(window_index, checkpoint) runs after each window.
    world_id_start: the first world's id in a larger ensemble (reported in the summary)."""
    import torch

    packed = ensemble.packed
    W = packed.W
    began = time.perf_counter()
    timings = {"compile_seconds": 0.0, "execute_seconds": 0.0, "locate_seconds": 0.0, "observe_seconds": 0.0,
               "windows": 0}
    if checkpoint is not None:
        y = ensemble.to_device(checkpoint.state)
        counts = np.array(checkpoint.step_index, dtype=np.int64)
        if checkpoint.dt != dt:
            raise ValueError("a checkpoint resumes at its own dt")
    else:
        y = ensemble.to_device(y0)
        counts = np.zeros(W, dtype=np.int64)
    start = int(counts.max()) if checkpoint is not None else 0
    if checkpoint is not None and not np.all((counts == start) | (counts < start)):
        raise ValueError("inconsistent checkpoint step counts")
    first_state = y.clone()
    live = np.ones(W, dtype=bool)
    if checkpoint is not None:
        live = counts == start  # worlds behind the others refused before the checkpoint
    refused = {}
    observations = {}
    at = start
    window = 0
    for cut in _boundaries(start, steps, policy.window_steps, policy.observe_every):
        n = cut - at
        ensemble.step_checks(True, material_atol, at)
        t0 = time.perf_counter()
        out, ok = ensemble.step(y, dt, n, ctl=ctl, active=live.astype(np.int32))
        okh = ok.cpu().numpy().astype(bool)
        bits = ensemble.check_bits().cpu().numpy()
        # the first window's call captures the step graph (the compile of a jitted window)
        timings["compile_seconds" if window == 0 else "execute_seconds"] += time.perf_counter() - t0
        failed = np.flatnonzero(live & ~okh)
        if failed.size:
            t1 = time.perf_counter()
            located = [_locate(ensemble, y, dt, ctl, int(w), at, n, material_atol) for w in failed]
            timings["locate_seconds"] += time.perf_counter() - t1
            if policy.refusals == "batch":
                ensemble.step_checks(False)
                r = min(located, key=lambda r: (r.step, r.world))
                raise NumericalError(f"world {r.world} refused step {r.step} (check bits {r.bits})")
            for r in located:
                refused[r.world] = r
                counts[r.world] = r.step
            live &= okh
        counts[live] = cut
        y = out
        at = cut
        window += 1
        if policy.observe_every and cut % policy.observe_every == 0:
            t1 = time.perf_counter()
            observations[cut] = cell_potentials(packed, y)
            timings["observe_seconds"] += time.perf_counter() - t1
        cp = Checkpoint(y, counts.copy(), dt)
        if sink:
            sink(window, cp)
    ensemble.step_checks(False)
    torch.cuda.synchronize()
    observations["final"] = cell_potentials(packed, y)
    _, ratios = ensemble.device_checks(first_state, y, float(max(at - start, 1)), material_atol)
    r = ratios.cpu().numpy()
    accounting = {k: float(r[:, i].max()) for i, k in enumerate(_lib.LEDGER_RATIOS)}
    timings["windows"] = window
    timings["elapsed_seconds"] = time.perf_counter() - began
    free, total = torch.cuda.mem_get_info(ensemble.device)
    resources = {
        "executed_world_steps": int(counts.sum() - (0 if checkpoint is None else np.sum(checkpoint.step_index))),
        "state_bytes": int(8 * ensemble.state_size),
        "device_memory_in_use_bytes": int(total - free),
    }
    return RunResult(Checkpoint(y, counts, dt), refused, observations, accounting, timings, resources,
                     world_id_start)


def _locate(ensemble, y, dt, ctl, world, at, n, material_atol):
    """The first refused step of `world` in the window that starts at step `at` (its state in y),
    replaying that world alone one step at a time."""
    active = np.zeros(ensemble.packed.W, dtype=np.int32)
    active[world] = 1
    cur = y
    for k in range(n):
        ensemble.step_checks(True, material_atol, at + k)
        nxt, ok = ensemble.step(cur, dt, 1, ctl=ctl, active=active)
        if not bool(ok[world]):
            bits = int(ensemble.check_bits()[world])
            return WorldRefusal(world, at + k, bits)
        cur = nxt
    raise RuntimeError(f"world {world} refused in its window but not when replayed alone")
