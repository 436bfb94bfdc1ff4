"""Oracle (test infrastructure only): the sparse Cholesky's host symbolic analysis, restated.

Only tests, the smoke check and the benchmark's checker legs (its CPU baseline and its parity checks,
outside every timed region) may import this module; the product's own symbolic analysis is the C++ in cells_amd/csrc/symbolic.cpp, which these functions
check bit for bit.

Restates src/cells/numerics/sparse_cholesky.py of the reference:
- `neighbours`    follows `_neighbours`      (sparse_cholesky.py:123-135)
- `minimum_degree` follows `minimum_degree`  (sparse_cholesky.py:137-162): a heap of
  (current degree, original degree, index); stale entries skipped; the eliminated node's
  neighbours become a clique.
- `bind`          follows `bind`             (sparse_cholesky.py:165-215): rank, union-rule
  column structures, parent = min(below), column-major entry positions (diagonal first, then
  rows ascending by rank), levels (leaves 0, a parent above its children).
- `row_structures` follows `_row_structures` (sparse_cholesky.py:229-235).
- `pairwise_tree_sum` is the fold's order (`_pairwise_steps`/`fold`, sparse_cholesky.py:244-322):
  adjacent pairs, step by step, an odd tail passing through.
- `product_plan`  follows `product_plan`     (sparse_cholesky.py:444-464): each row's entries, its
  neighbours and itself in ascending column order, rows in order (so the value array on the
  pattern is row-major), each row's diagonal position; the row sums' pairwise plan is the fold
  over each row's entries in that order.
- `product`       follows `product`          (sparse_cholesky.py:467-471): each row's products
  value x x[column], summed by `pairwise_tree_sum` (the products are formed and rounded before the
  sum, as the reference's optimization barrier keeps them).

Pinned by: the O(n^2) definition of minimum degree (tests/unit/test_sparse_cholesky.py:198-211 of
the reference, restated in tests/test_oracle_symbolic.py), the arrowhead special cases
(:221-228), level order (:231-238) and the malformed-pattern refusals (:252-268).
"""

import heapq

import numpy as np


def neighbours(size, rows, columns):
    rows = np.asarray(rows, dtype=np.int64)
    columns = np.asarray(columns, dtype=np.int64)
    if rows.shape != columns.shape or rows.ndim != 1:
        raise ValueError("A sparse pattern's coordinates must be two equal one-dimensional arrays")
    if rows.size and (min(rows.min(), columns.min()) < 0 or max(rows.max(), columns.max()) >= size):
        raise ValueError("A sparse pattern's coordinates must lie within its size")
    adjacency = [set() for _ in range(size)]
    for i, j in zip(rows.tolist(), columns.tolist()):
        if i != j:
            adjacency[i].add(j)
            adjacency[j].add(i)
    return adjacency


def minimum_degree(size, rows, columns):
    adjacency = neighbours(size, rows, columns)
    original = [len(a) for a in adjacency]
    heap = [(original[i], original[i], i) for i in range(size)]
    heapq.heapify(heap)
    alive = [True] * size
    order = []
    while heap:
        degree, _, node = heapq.heappop(heap)
        if not alive[node] or degree != len(adjacency[node]):
            continue
        alive[node] = False
        order.append(node)
        clique = adjacency[node]
        for a in clique:
            adjacency[a].discard(node)
            adjacency[a] |= clique - {a}
            heapq.heappush(heap, (len(adjacency[a]), original[a], a))
        adjacency[node] = set()
    return np.asarray(order, dtype=np.int64)


class Symbolic:
    """order, rank, column_start, row, column, parent, level, structure (per rank, ascending)."""

    def __init__(self, **kw):
        self.__dict__.update(kw)


def bind(size, rows, columns, order=None):
    if type(size) is not int or size < 1:
        raise ValueError("A sparse pattern needs a positive integer size")
    adjacency = neighbours(size, rows, columns)
    if order is None:
        order = minimum_degree(size, rows, columns)
    order = np.asarray(order, dtype=np.int64)
    if order.shape != (size,) or sorted(order.tolist()) != list(range(size)):
        raise ValueError("The elimination order must be a permutation")
    rank = np.empty(size, dtype=np.int64)
    rank[order] = np.arange(size)
    below = [set() for _ in range(size)]
    for i in range(size):
        r = int(rank[i])
        below[r].update(int(rank[a]) for a in adjacency[i] if rank[a] > r)
    parent = np.full(size, -1, dtype=np.int64)
    for k in range(size):
        if below[k]:
            p = min(below[k])
            parent[k] = p
            below[p] |= below[k] - {p}
    structure = [sorted(b) for b in below]
    column_start = np.zeros(size + 1, dtype=np.int64)
    column_start[1:] = np.cumsum([1 + len(b) for b in structure])
    entry_row, entry_column = [], []
    for k, rows_k in enumerate(structure):
        for r in (k, *rows_k):
            entry_row.append(int(order[r]))
            entry_column.append(int(order[k]))
    level = np.zeros(size, dtype=np.int64)
    for k in range(size):
        if parent[k] >= 0:
            level[parent[k]] = max(level[parent[k]], level[k] + 1)
    return Symbolic(
        size=size,
        order=order,
        rank=rank,
        column_start=column_start,
        row=np.asarray(entry_row, dtype=np.int64),
        column=np.asarray(entry_column, dtype=np.int64),
        parent=parent,
        level=level,
        structure=structure,
    )


def row_structures(symbolic):
    rows = [[] for _ in range(symbolic.size)]
    for m, rows_m in enumerate(symbolic.structure):
        for r in rows_m:
            rows[r].append(m)
    return rows


def pairwise_tree_sum(values):
    """The fold's order for one segment: adjacent pairs step by step; +0.0 when empty."""
    values = list(values)
    if not values:
        return 0.0
    while len(values) > 1:
        values = [values[i] + values[i + 1] if i + 1 < len(values) else values[i]
                  for i in range(0, len(values), 2)]
    return values[0]


class ProductPlan:
    """A product's plan on a symmetric pattern with its whole diagonal, as row starts (CSR)."""

    def __init__(self, row_start, column, diagonal):
        self.row_start = row_start  # [n + 1] int64
        self.column = column  # [nnz] int64, ascending within each row
        self.diagonal = diagonal  # [n] int64: each row's diagonal entry
        self.row = np.repeat(np.arange(row_start.size - 1, dtype=np.int64), np.diff(row_start))


def product_plan(size, rows, columns):
    """The ProductPlan of the symmetric pattern of `size` nodes with off-diagonal entries at (rows,
    columns), either or both triangles, and its whole diagonal."""
    adjacency = neighbours(size, rows, columns)
    row_start, column, diagonal = [0], [], []
    for i in range(size):
        entries = sorted(adjacency[i] | {i})
        diagonal.append(row_start[-1] + entries.index(i))
        column.extend(entries)
        row_start.append(len(column))
    return ProductPlan(np.asarray(row_start, dtype=np.int64), np.asarray(column, dtype=np.int64),
                       np.asarray(diagonal, dtype=np.int64))


def product(values, x, plan):
    """A x for the matrix A whose entries on the plan's pattern are `values` (x [n] or [n, k]; its
    trailing axes are independent): each row's products summed in adjacent pairs, step by step."""
    values = np.asarray(values, dtype=np.float64)
    x = np.asarray(x, dtype=np.float64)
    flat = x.reshape(x.shape[0], -1)
    y = np.empty_like(flat)
    for i in range(plan.row_start.size - 1):
        a, b = int(plan.row_start[i]), int(plan.row_start[i + 1])
        terms = values[a:b, None] * flat[plan.column[a:b]]
        for k in range(flat.shape[1]):
            y[i, k] = pairwise_tree_sum(terms[:, k])
    return y.reshape(x.shape)
