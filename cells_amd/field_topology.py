"""bind's capacitance topology on the host through the C-ABI (cells_capacitor_topology).

Reference: src/cells/numerics/surface_field.py `bind` (:253-340, the union-find capacitor
components, the representatives of `representative` :238-250, the floating basis, the offsets' gauge
freedom and the return-path check) over the layout checks of numerics/surface_transport.py:194-216.
A layout is the reference's field-layout fields (cell_active, outside_active, reservoir,
membrane_active, membrane_cell, membrane_outside, face_active, face_sender, face_receiver,
face_factor [species, face], valence [species, 1]); refusals raise ValueError with the reference's
messages.
"""

import ctypes
from typing import NamedTuple

import numpy as np

from cells_amd import lib as _lib

_i32p = ctypes.POINTER(ctypes.c_int32)
_u8p = ctypes.POINTER(ctypes.c_uint8)


class Topology(NamedTuple):
    component: np.ndarray  # [n + o] capacitor component (by its lowest active node), -1 inactive
    fixed: np.ndarray  # [n + o] bool: inactive, reservoir or a floating component's representative
    floating_basis: np.ndarray  # [n + o, k] 1.0 on each floating component's members
    offset_free: np.ndarray  # [k] bool: the component's representative is not the gauge


def bind_topology(layout, gauge_mask, return_paths=()):
    """The capacitance topology bind derives from a field layout (dict or object of its fields)."""
    get = layout.get if isinstance(layout, dict) else (lambda k: getattr(layout, k))
    cell_active = np.ascontiguousarray(get("cell_active"), dtype=bool)
    outside_active = np.ascontiguousarray(get("outside_active"), dtype=bool)
    n, o = cell_active.size, outside_active.size
    gauge = np.asarray(gauge_mask)
    if gauge.shape != (o,) or gauge.dtype != np.bool_:
        raise ValueError("Resolved field gauge must be an outside boolean mask")
    valence = np.asarray(get("valence"), dtype=float)
    face_factor = np.asarray(get("face_factor"), dtype=float)
    face_active = np.asarray(get("face_active"), dtype=bool)
    conductive = face_active & np.any((valence != 0) & (face_factor > 0), axis=0)
    u8 = lambda a: np.ascontiguousarray(a, dtype=np.uint8)  # noqa: E731
    i32 = lambda a: np.ascontiguousarray(a, dtype=np.int32)  # noqa: E731
    mc, mo = i32(get("membrane_cell")), i32(get("membrane_outside"))
    ma = u8(get("membrane_active"))
    fs, fr, fc = i32(get("face_sender")), i32(get("face_receiver")), u8(conductive)
    rp = i32(np.asarray(return_paths, dtype=np.int64).reshape(-1))
    ca, oa, res, g = u8(cell_active), u8(outside_active), u8(get("reservoir")), u8(gauge)
    size = n + o
    component, floating = np.empty(size, np.int32), np.empty(size, np.int32)
    fixed, free = np.empty(size, np.uint8), np.empty(size, np.uint8)
    counts = np.zeros(2, np.int32)
    L = _lib.load()
    rc = L.cells_capacitor_topology(n, o, mc.size, mc.ctypes.data_as(_i32p), mo.ctypes.data_as(_i32p),
                                    ma.ctypes.data_as(_u8p), ca.ctypes.data_as(_u8p), oa.ctypes.data_as(_u8p),
                                    res.ctypes.data_as(_u8p), g.ctypes.data_as(_u8p), fs.size, fs.ctypes.data_as(_i32p),
                                    fr.ctypes.data_as(_i32p), fc.ctypes.data_as(_u8p), rp.size // 2,
                                    rp.ctypes.data_as(_i32p), component.ctypes.data_as(_i32p),
                                    floating.ctypes.data_as(_i32p), fixed.ctypes.data_as(_u8p),
                                    free.ctypes.data_as(_u8p), counts.ctypes.data_as(_i32p))
    if rc == _lib.CELLS_EINVAL:
        raise ValueError(L.cells_last_error().decode())
    _lib.check(rc)
    k = int(counts[1])
    basis = np.zeros((size, k))
    members = floating >= 0
    basis[np.flatnonzero(members), floating[members]] = 1.0
    return Topology(component.astype(np.int64), fixed.astype(bool), basis, free[:k].astype(bool))
