"""Host symbolic analysis through the C-ABI (cells_minimum_degree, cells_bind).

The reference's counterpart is src/cells/numerics/sparse_cholesky.py: `minimum_degree` (:137-162)
and `bind` (:165-215), whose integer outputs these reproduce bit for bit. Runs on the host only;
no GPU is needed.
"""

import ctypes
from typing import NamedTuple

import numpy as np

from cells_amd import lib as _lib

_i32p = ctypes.POINTER(ctypes.c_int32)
_i64p = ctypes.POINTER(ctypes.c_int64)


class Symbolic(NamedTuple):
    size: int
    order: np.ndarray
    rank: np.ndarray
    column_start: np.ndarray
    row: np.ndarray  # entry rows, as nodes
    column: np.ndarray  # entry columns, as nodes
    parent: np.ndarray
    level: np.ndarray


def _coords(rows, columns):
    rows = np.ascontiguousarray(rows, dtype=np.int32)
    columns = np.ascontiguousarray(columns, dtype=np.int32)
    if rows.shape != columns.shape or rows.ndim != 1:
        raise ValueError("A sparse pattern's coordinates must be two equal one-dimensional arrays")
    return rows, columns


def minimum_degree(size, rows, columns):
    rows, columns = _coords(rows, columns)
    if type(size) is not int or size < 1:
        raise ValueError("A sparse pattern needs a positive integer size")
    out = np.empty(size, dtype=np.int32)
    rc = _lib.load().cells_minimum_degree(size, rows.size, rows.ctypes.data_as(_i32p),
                                          columns.ctypes.data_as(_i32p), out.ctypes.data_as(_i32p))
    if rc == _lib.CELLS_EINVAL:
        raise ValueError(_lib.load().cells_last_error().decode())
    _lib.check(rc)
    return out.astype(np.int64)


def relaxed_padding(size, rows, columns, max_rows=128, max_zeros=64):
    """(the bound pattern's order, the relaxed supernodes' explicit zeros as (rows, columns) node
    arrays) through cells_relaxed_padding: binding the pattern plus them with that order gives the
    multifrontal plan's pattern."""
    rows, columns = _coords(rows, columns)
    L = _lib.load()
    handle = ctypes.c_void_p()
    rc = L.cells_bind(size, rows.size, rows.ctypes.data_as(_i32p), columns.ctypes.data_as(_i32p), None,
                      ctypes.byref(handle))
    if rc == _lib.CELLS_EINVAL:
        raise ValueError(L.cells_last_error().decode())
    _lib.check(rc)
    try:
        count = ctypes.c_int64()
        _lib.check(L.cells_relaxed_padding(handle, max_rows, max_zeros, None, 0, ctypes.byref(count)))
        pairs = np.empty(2 * max(count.value, 1), np.int32)
        _lib.check(L.cells_relaxed_padding(handle, max_rows, max_zeros, pairs.ctypes.data_as(_i32p), count.value,
                                           ctypes.byref(count)))
        n = size
        order = np.empty(n, np.int32)
        nnz = ctypes.c_int64()
        levels = ctypes.c_int32()
        _lib.check(L.cells_symbolic_sizes(handle, ctypes.byref(nnz), ctypes.byref(levels)))
        f = nnz.value
        scratch = dict(rank=np.empty(n, np.int32), column_start=np.empty(n + 1, np.int64), row=np.empty(f, np.int32),
                       column=np.empty(f, np.int32), parent=np.empty(n, np.int32), level=np.empty(n, np.int32))
        _lib.check(L.cells_symbolic_export(
            handle, order.ctypes.data_as(_i32p), scratch["rank"].ctypes.data_as(_i32p),
            scratch["column_start"].ctypes.data_as(_i64p), scratch["row"].ctypes.data_as(_i32p),
            scratch["column"].ctypes.data_as(_i32p), scratch["parent"].ctypes.data_as(_i32p),
            scratch["level"].ctypes.data_as(_i32p)))
    finally:
        L.cells_symbolic_free(handle)
    p = pairs[:2 * count.value].astype(np.int64)
    return order.astype(np.int64), p[0::2], p[1::2]


def bind(size, rows, columns, order=None):
    if type(size) is not int or size < 1:
        raise ValueError("A sparse pattern needs a positive integer size")
    rows, columns = _coords(rows, columns)
    L = _lib.load()
    handle = ctypes.c_void_p()
    if order is not None:
        order = np.ascontiguousarray(order, dtype=np.int32)
        if order.shape != (size,):
            raise ValueError("The elimination order must be a permutation")
    rc = L.cells_bind(size, rows.size, rows.ctypes.data_as(_i32p), columns.ctypes.data_as(_i32p),
                      None if order is None else order.ctypes.data_as(_i32p), ctypes.byref(handle))
    if rc == _lib.CELLS_EINVAL:
        raise ValueError(L.cells_last_error().decode())
    _lib.check(rc)
    try:
        nnz = ctypes.c_int64()
        levels = ctypes.c_int32()
        _lib.check(L.cells_symbolic_sizes(handle, ctypes.byref(nnz), ctypes.byref(levels)))
        n, f = size, nnz.value
        arrays = dict(order=np.empty(n, np.int32), rank=np.empty(n, np.int32),
                      column_start=np.empty(n + 1, np.int64), row=np.empty(f, np.int32),
                      column=np.empty(f, np.int32), parent=np.empty(n, np.int32),
                      level=np.empty(n, np.int32))
        _lib.check(L.cells_symbolic_export(
            handle, *(arrays[k].ctypes.data_as(_i64p if k == "column_start" else _i32p)
                      for k in ("order", "rank", "column_start", "row", "column", "parent",
                                "level"))))
    finally:
        L.cells_symbolic_free(handle)
    return Symbolic(size, *(arrays[k].astype(np.int64) for k in
                            ("order", "rank", "column_start", "row", "column", "parent", "level")))
