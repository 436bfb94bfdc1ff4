"""The sparse Cholesky module's matrix-vector product on the device (cells_product_plan, cells_product).

Reference: src/cells/numerics/sparse_cholesky.py, `product_plan` (:444-464) and `product` (:467-471).
A plan is built once on the host from a symmetric pattern (either or both triangles of its
off-diagonal coordinates; the whole diagonal is implied): each row's entries are its neighbours and
itself in ascending column order, rows in order, so a value array on the pattern is row-major. The
product sums each row's products (each formed and rounded on its own, as the reference's
optimization barrier keeps them) in adjacent pairs, step by step: the reference's fold, bit for bit.
The same errors as the reference's `_neighbours` refuse a malformed pattern.
"""

import ctypes
from typing import NamedTuple

import numpy as np

from cells_amd import lib as _lib
from cells_amd.symbolic import _coords

_i32p = ctypes.POINTER(ctypes.c_int32)
_i64p = ctypes.POINTER(ctypes.c_int64)


class ProductPlan(NamedTuple):
    """The product's plan (the reference's ProductPlan: row, column, diagonal; the row sums'
    pairwise plan is each row's run of entries, `row_start`)."""

    row: np.ndarray  # [nnz] int64: each entry's row
    column: np.ndarray  # [nnz] int64: each entry's column
    diagonal: np.ndarray  # [n] int64: each row's diagonal entry
    row_start: np.ndarray  # [n + 1] int64: row i's entries are row_start[i]:row_start[i + 1]


def product_plan(size, rows, columns):
    """The ProductPlan of the symmetric pattern of `size` nodes with off-diagonal entries at (rows,
    columns) (either or both triangles) and its whole diagonal."""
    if type(size) is not int or size < 1:
        raise ValueError("A sparse pattern needs a positive integer size")
    rows, columns = _coords(rows, columns)
    L = _lib.load()
    nnz = ctypes.c_int64()
    args = (size, rows.size, rows.ctypes.data_as(_i32p), columns.ctypes.data_as(_i32p))
    rc = L.cells_product_plan_size(*args, ctypes.byref(nnz))
    if rc == _lib.CELLS_EINVAL:
        raise ValueError(L.cells_last_error().decode())
    _lib.check(rc)
    row_start = np.empty(size + 1, np.int64)
    column = np.empty(nnz.value, np.int32)
    diagonal = np.empty(size, np.int64)
    _lib.check(L.cells_product_plan(*args, row_start.ctypes.data_as(_i64p), column.ctypes.data_as(_i32p),
                                    diagonal.ctypes.data_as(_i64p)))
    row = np.repeat(np.arange(size, dtype=np.int64), np.diff(row_start))
    return ProductPlan(row, column.astype(np.int64), diagonal, row_start)


class DevicePlan:
    """A ProductPlan's row starts and columns on a device (built once, reused by every product)."""

    def __init__(self, plan: ProductPlan, device=0):
        import torch

        self.plan = plan
        self.device = torch.device("cuda", device) if isinstance(device, int) else torch.device(device)
        self.n = int(plan.row_start.size - 1)
        self.nnz = int(plan.column.size)
        self.row_start = torch.from_numpy(np.ascontiguousarray(plan.row_start, np.int64)).to(self.device)
        self.column = torch.from_numpy(np.ascontiguousarray(plan.column, np.int32)).to(self.device)


def _launch(dp, values, x, y, count, values_stride, vector_stride):
    import torch

    stream = ctypes.c_void_p(torch.cuda.current_stream(dp.device).cuda_stream)
    _lib.check(_lib.load().cells_product(
        stream, dp.n, ctypes.c_void_p(dp.row_start.data_ptr()), ctypes.c_void_p(dp.column.data_ptr()),
        ctypes.c_void_p(values.data_ptr()), ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(y.data_ptr()), count,
        values_stride, vector_stride))


def product(values, x, plan):
    """A x for the matrix A whose entries on the plan's pattern are `values` [nnz] (x [n] or [n, ...]
    on the plan's device, float64; trailing axes are independent)."""
    import torch

    dp = plan if isinstance(plan, DevicePlan) else DevicePlan(plan, x.device)
    values = values.to(dp.device, torch.float64).contiguous()
    if values.shape != (dp.nnz,):
        raise ValueError("values must hold one entry per pattern position")
    if x.shape[0] != dp.n:
        raise ValueError("x must have one row per node")
    cols = x.reshape(dp.n, -1).to(dp.device, torch.float64).t().contiguous()  # [k, n]
    y = torch.empty_like(cols)
    _launch(dp, values, cols, y, cols.shape[0], 0, dp.n)
    return y.t().reshape(x.shape).contiguous()


def product_batched(values, x, plan):
    """A_b x_b for a batch of matrices on one pattern: values [B, nnz], x [B, n] -> [B, n]."""
    import torch

    dp = plan if isinstance(plan, DevicePlan) else DevicePlan(plan, x.device)
    values = values.to(dp.device, torch.float64).contiguous()
    x = x.to(dp.device, torch.float64).contiguous()
    if values.ndim != 2 or values.shape[1] != dp.nnz or x.shape != (values.shape[0], dp.n):
        raise ValueError("values [B, nnz] and x [B, n] on the plan's pattern")
    y = torch.empty_like(x)
    _launch(dp, values, x, y, values.shape[0], dp.nnz, dp.n)
    return y
