"""The reference's model-module contract over the HIP step.

The reference's execution layer drives a model module through a few functions
(registry/models.py:1-6 of the reference; models/surface_pump_v3.py):
`initial_values(spec) -> (state, layout)`, `checked_step(state, dt, control, p, active) ->
(state, ok)`, plus `readouts`/`domain_valid`. This module gives those names over packed worlds and
a device ensemble, with the reference's meanings:

- a state is one FP64 vector (cells_amd.world.StateLayout) instead of a pytree of leaves;
- `checked_step` returns the new state where a world's step is valid and the old one where it is
  not (models/surface_pump_v3.py:3934-3935), with ok per world, never raising inside the step;
- the runner's error on a refused step (execution/runner.py:505-510) is `run(..., refusals=
  "raise")`, and the vector windows' "world" policy (execution/vector_windows.py:20-26), where a
  failing world stops and the others continue, is `refusals="world"`;
- `pack_states` / `unpack_states` map the reference's per-world state pytrees (the leaves of
  STATE_TYPE by name: models/surface_tissue.py:31-44, surface_kinetic.py:36-44,
  surface_pump_v2.py:44-50, water's operators/osmotic_water.py:76-83) to and from the packed
  vector, world by world, with the names and axes the checkpoints use.
"""

from __future__ import annotations

import numpy as np

from cells_amd.world import F, Config, Packed, StateLayout, build, initial_state

MODEL_ID = "cells.model.resolved_pump_tissue@3"
# the reference's solver ids (numerics/surface_imex.py:36-45) and this build's ROS2
SOLVERS = {"imex_euler": "euler", "ros2": "ros2", "imex_ars222_implicit_conductances": "n2",
           "imex_ars222": "ars222", "rk4": "rk4",
           "cells.solver.imex_euler_face_electromigration@1": "euler",
           "cells.solver.imex_ars222_face_electromigration@1": "ars222",
           "cells.solver.imex_ars222_implicit_conductances@1": "n2"}


class NumericalError(RuntimeError):
    """A refused step under refusals="raise" (execution/errors.py of the reference)."""


class Layout:
    """The bound layout: packed worlds plus their device ensemble (built once per definition)."""

    def __init__(self, packed: Packed, solver: str = "imex_euler", device: int = 0):
        from cells_amd.ensemble import Ensemble

        if solver not in SOLVERS:
            raise ValueError(f"unknown solver {solver!r}; expected one of {sorted(SOLVERS)}")
        self.packed = packed
        self.state_layout = StateLayout.of(packed)
        self.solver = solver
        self.ensemble = Ensemble(packed, SOLVERS[solver], device)


def initial_values(config: Config, solver: str = "imex_euler", device: int = 0):
    """(state, layout) of a configuration's worlds (models/surface_tissue.py:106)."""
    packed = build(config)
    layout = Layout(packed, solver, device)
    return layout.ensemble.to_device(initial_state(packed)), layout


def checked_step(state, dt, control, p: Layout, active=None):
    """One checked step of every world (models/surface_pump_v3.py:3989)."""
    return p.ensemble.step(state, dt, 1, ctl=control, active=active)


def run(state, dt, steps, control, p: Layout, *, refusals="raise", batched=True):
    """`steps` checked steps; refusals="raise" raises at a refused world, "world" lets the others
    continue (the refused world keeps its last accepted state)."""
    out, ok = p.ensemble.step(state, dt, steps, ctl=control, batched=batched)
    okh = ok.cpu().numpy().astype(bool)
    if refusals == "raise" and not okh.all():
        bad = np.flatnonzero(~okh).tolist()
        raise NumericalError(f"worlds {bad[:8]} refused a step")
    return out, ok


def potentials(state, p: Layout):
    """Cell potentials (V) of a state: the held bath plus each cell's charge over its capacitor."""
    pk, L = p.packed, p.state_layout
    y = state.cpu().numpy() if hasattr(state, "cpu") else np.asarray(state)
    q = pk.q0 + 1e-15 * F * sum(pk.valence[s] * L.cell(y, L.dn + s) for s in range(pk.S))
    return pk.config.bath_V + q / pk.cap


def unpack_states(packed: Packed, y) -> list:
    """Each world's state pytree (a dict of the reference's leaf names to arrays, with its axes)
    cut out of the packed state y."""
    from cells_amd.snapshot import leaves

    yh = y.cpu().numpy() if hasattr(y, "cpu") else np.asarray(y)
    return [leaves(packed, yh, k) for k in range(packed.W)]


def pack_states(packed: Packed, states) -> np.ndarray:
    """The packed state of per-world pytrees (one dict of leaf names per world, in world order, as
    unpack_states gives them). Every leaf the packed layout holds must be present with its shape;
    the leaves the lattice worlds have no slots for (faces, channels, finite outside nodes) must
    be empty or zero."""
    from cells_amd.snapshot import leaves, place

    if len(states) != packed.W:
        raise ValueError(f"need one state per world ({packed.W}), got {len(states)}")
    L = StateLayout.of(packed)
    y = np.zeros(L.total)
    for k, st in enumerate(states):
        want = leaves(packed, y, k)
        missing = sorted(set(want) - set(st))
        if missing:
            raise ValueError(f"world {k}: missing state leaves {missing}")
        arrays = {}
        for name, template in want.items():
            a = np.asarray(st[name], dtype=np.float64)
            if a.shape != template.shape:
                raise ValueError(f"world {k}: leaf {name} has shape {a.shape}, expected {template.shape}")
            arrays[name] = a
        for name in ("extracellular_amount_increment", "face_transferred", "face_dissipated_energy_pJ",
                     "channel_gate", "channel_transferred", "channel_dissipated_energy_pJ",
                     "extracellular_volume_increment_pL"):
            if name in arrays and np.any(arrays[name] != 0):
                raise ValueError(f"world {k}: leaf {name} has no slots in a lattice on a held bath")
        place(packed, y, k, arrays)
    return y
