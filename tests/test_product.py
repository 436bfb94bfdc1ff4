"""The sparse Cholesky module's product (reference numerics/sparse_cholesky.py:444-471) on the CPU.

Restates the reference's tests/unit/test_sparse_cholesky.py:155-166 (products against dense on its
patterns, a zero entry on the pattern kept) for the oracle's product (oracle/symbolic.py), pins the
fold's tree with known answers that other summation trees miss, checks the signed zero of a row of
-0.0 terms, and pins the product's host plan (cells_amd.sparse via the C-ABI) to the oracle's.
"""

import numpy as np
import pytest

from cells_amd import sparse
from oracle import symbolic as ref
from tests.test_symbolic import PATTERNS, coordinates, spd


@pytest.mark.parametrize("name", sorted(PATTERNS))
def test_oracle_products_against_dense(name):
    pattern = PATTERNS[name]
    n, rows, columns = coordinates(pattern)
    plan = ref.product_plan(n, rows, columns)
    a = spd(pattern, 9)
    a[0, 0] = 0.0  # a zero entry on the pattern stays an entry
    x = np.random.default_rng(3).uniform(-1.0, 1.0, (n, 2))
    values = a[plan.row, plan.column]
    assert values.size == int(pattern.sum())
    y = ref.product(values, x, plan)
    assert np.allclose(y, a @ x, rtol=1e-13, atol=1e-14)
    # the plan's pattern is the matrix's: every nonzero of a has a position
    assert np.array_equal(np.sort(plan.row * n + plan.column), np.flatnonzero(pattern.reshape(-1)))
    assert np.array_equal(plan.column[plan.diagonal], np.arange(n))


def _one_row(terms):
    """A pattern whose row 0 holds len(terms) entries (row 0 joined to nodes 1..), its values on
    row 0 equal to `terms` and x = 1, so y[0] is the fold of `terms` in order."""
    n = len(terms)
    plan = ref.product_plan(n, np.zeros(n - 1, np.int64), np.arange(1, n))
    values = np.ones(plan.column.size)
    values[: n] = terms
    return values, np.ones(n), plan


@pytest.mark.parametrize("terms, expected", [
    # (1 + 1e16) + -1e16: the odd tail waits a step; 1 + (1e16 + -1e16) would be 1.0
    ([1.0, 1e16, -1e16], 0.0),
    # ((1 + 1e16) + (-1e16 + 1)) + ((1 + 1) + -0): a serial fold and halves of three both give 3.0
    ([1.0, 1e16, -1e16, 1.0, 1.0, 1.0], 2.0),
    # (1e16 + 1) + (-1e16 + 1): a serial fold gives 1.0
    ([1e16, 1.0, -1e16, 1.0], 0.0),
])
def test_the_rows_are_summed_in_adjacent_pairs(terms, expected):
    values, x, plan = _one_row(terms)
    assert ref.product(values, x, plan)[0] == expected


def test_a_row_of_negative_zero_terms_keeps_its_sign():
    values, x, plan = _one_row([-0.0] * 5)
    y = ref.product(values, x, plan)
    assert y[0] == 0.0 and np.signbit(y[0])


@pytest.mark.parametrize("name", sorted(PATTERNS))
def test_host_plan_equals_the_oracle_plan(name):
    n, rows, columns = coordinates(PATTERNS[name])
    for r, c in ((rows, columns), (columns, rows), (np.r_[rows, columns], np.r_[columns, rows])):
        got, want = sparse.product_plan(n, r, c), ref.product_plan(n, r, c)
        for k in ("row_start", "column", "diagonal"):
            assert np.array_equal(getattr(got, k), getattr(want, k)), k
        assert np.array_equal(got.row, want.row)


def test_host_plan_refuses_what_the_reference_refuses():
    with pytest.raises(ValueError, match="must lie within its size"):
        sparse.product_plan(3, [0, 3], [1, 2])
    with pytest.raises(ValueError, match="must lie within its size"):
        ref.product_plan(3, [0, 3], [1, 2])
    with pytest.raises(ValueError, match="two equal one-dimensional arrays"):
        sparse.product_plan(3, [0, 1], [1])
    with pytest.raises(ValueError, match="positive integer size"):
        sparse.product_plan(0, [], [])
