"""Plan autotuner: picks the multifrontal plans' schedule knobs by timing them on the GPU.

The batched configuration asks for packed launches with an autotuned schedule (uneven worlds,
one launch per stage). In this library the launch geometry is fixed (one workgroup per world,
largest worlds first; inside a world each front goes to a thread, an 8-lane group, a wavefront or
the whole workgroup by its order), and what shapes the schedule is the plan: how far chains of
supernodes are joined into larger fronts (relaxed supernodes: fewer, taller tree levels against
explicit zeros) and, for the spatial grids, at which depth the tree is cut into per-workgroup
subtrees and per-depth launches above. These are plan options (cells_set_plan_options), read when
an ensemble or field is created; the autotuner creates one per candidate on the caller's own
worlds, times a few captured steps, and leaves the fastest set in place. The library defaults are
the winners of such sweeps on the batched (4096 worlds) and spatial (64 worlds, 256^2) workloads.
"""

from __future__ import annotations

import ctypes
import dataclasses
import itertools
import time

from cells_amd import lib as _lib

_FIELDS = ("relax_rows", "relax_zeros", "grid_relax_zeros", "grid_subtrees", "widen_groups", "largest_first")


def plan_options() -> dict:
    """The plan options the next create reads."""
    vals = [ctypes.c_int32() for _ in _FIELDS]
    _lib.check(_lib.load().cells_get_plan_options(*(ctypes.byref(v) for v in vals)))
    return {k: v.value for k, v in zip(_FIELDS, vals)}


def set_plan_options(**kw) -> dict:
    """Set some plan options (the others kept); returns the options now in force."""
    unknown = set(kw) - set(_FIELDS)
    if unknown:
        raise ValueError(f"unknown plan options {sorted(unknown)}")
    cur = plan_options()
    cur.update({k: int(v) for k, v in kw.items()})
    _lib.check(_lib.load().cells_set_plan_options(*(cur[k] for k in _FIELDS)))
    return cur


@dataclasses.dataclass
class TuneResult:
    best: dict  # the fastest candidate's options (now in force when applied)
    ms_per_step: float
    table: list  # (options, ms per step) per candidate, in sweep order

    def summary(self):
        return {"best": self.best, "ms_per_step": self.ms_per_step,
                "candidates": [{"options": o, "ms_per_step": m} for o, m in self.table]}


def _sweep(candidates, run, apply):
    before = plan_options()
    table = []
    try:
        for cand in candidates:
            set_plan_options(**cand)
            table.append((dict(cand), run()))
    finally:
        set_plan_options(**before)
    best, ms = min(table, key=lambda t: t[1])
    if apply:
        set_plan_options(**best)
    return TuneResult(best, ms, table)


def autotune_ensemble(packed, method="ros2", dt=5e-3, steps=3, candidates=None, apply=True) -> TuneResult:
    """Times `steps` captured steps of `packed` (after one warm step) per candidate set of lattice
    plan options (default: explicit zeros per joined column 16, 64 or unbounded with fronts up to
    128; fronts up to 96; the level widening off; the worlds in their own order)."""
    import torch

    from cells_amd.ensemble import Ensemble
    from cells_amd.world import initial_state

    if candidates is None:
        base = {"relax_rows": 128, "relax_zeros": 64, "widen_groups": 1, "largest_first": 1}
        candidates = [dict(base, relax_zeros=z) for z in (16, 64, 4096)]
        candidates += [dict(base, relax_rows=96), dict(base, widen_groups=0), dict(base, largest_first=0)]
    y0 = initial_state(packed)

    def run():
        e = Ensemble(packed, method)
        try:
            y = e.to_device(y0)
            e.step(y, dt, nsteps=1)
            torch.cuda.synchronize()
            t = time.perf_counter()
            e.step(y, dt, nsteps=steps)
            torch.cuda.synchronize()
            return (time.perf_counter() - t) / steps * 1e3
        finally:
            e.close()

    return _sweep(candidates, run, apply)


def autotune_field(fp, dt=1e-3, steps=10, refresh_every=50, candidates=None, apply=True) -> TuneResult:
    """Times `steps` steps of the spatial worlds fp (after one warm step that builds the operator)
    per candidate set of grid plan options (default: 2, 3 or 4 zeros per joined column and a cut at
    24, 32 or 48 fronts)."""
    import torch

    from cells_amd.field import FieldEnsemble, initial_field_state

    if candidates is None:
        candidates = [{"grid_relax_zeros": z, "grid_subtrees": c}
                      for z, c in itertools.product((2, 3, 4), (24, 32, 48))]
    y0 = initial_field_state(fp)

    def run():
        e = FieldEnsemble(fp, refresh_every=refresh_every)
        try:
            y = e.to_device(y0)
            y, _ = e.step(y, dt, nsteps=1)
            torch.cuda.synchronize()
            t = time.perf_counter()
            e.step(y, dt, nsteps=steps)
            torch.cuda.synchronize()
            return (time.perf_counter() - t) / steps * 1e3
        finally:
            e.close()

    return _sweep(candidates, run, apply)
