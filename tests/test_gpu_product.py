"""The device product (cells_product) against the oracle's, bit for bit (reference
numerics/sparse_cholesky.py:467-471; its own test, tests/unit/test_sparse_cholesky.py:155-166,
asks the device product to equal the host's bit for bit).

The reference's patterns with x [n, 2]; a batch of matrices on one lattice stage pattern; a hub row
of 12 000 entries (above the 8191 terms of a 13-level counter stack); rows of -0.0 terms."""

import sys

import numpy as np
import pytest

from cells_amd import sparse
from cells_amd.world import lattice
from oracle import symbolic as ref
from tests.test_symbolic import PATTERNS, coordinates, spd

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def torch():
    import torch as t

    if not t.cuda.is_available():
        pytest.skip("no GPU")
    return t


def _bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).tobytes()


@pytest.mark.parametrize("name", sorted(PATTERNS))
def test_product_equals_the_oracle_bit_for_bit(torch, name):
    say(f"product on {name}")
    pattern = PATTERNS[name]
    n, rows, columns = coordinates(pattern)
    plan = sparse.product_plan(n, rows, columns)
    a = spd(pattern, 9)
    a[0, 0] = 0.0
    x = np.random.default_rng(3).uniform(-1.0, 1.0, (n, 2))
    values = a[plan.row, plan.column]
    want = ref.product(values, x, ref.product_plan(n, rows, columns))
    got = sparse.product(torch.from_numpy(values).cuda(), torch.from_numpy(x).cuda(), plan).cpu().numpy()
    assert _bits(got) == _bits(want)
    assert np.allclose(got, a @ x, rtol=1e-13, atol=1e-14)


def test_batched_products_on_a_lattice_stage_pattern(torch):
    say("batched products on a 37 x 37 + 11 lattice")
    n = 37 * 37 + 11
    snd, rcv = lattice(n)
    plan = sparse.product_plan(n, snd, rcv)
    rng = np.random.default_rng(5)
    B = 3
    values = rng.uniform(-1.0, 1.0, (B, plan.column.size)) * 10.0 ** rng.uniform(-6, 6, (B, plan.column.size))
    x = rng.uniform(-1.0, 1.0, (B, n))
    rp = ref.product_plan(n, snd, rcv)
    want = np.stack([ref.product(values[b], x[b], rp) for b in range(B)])
    dp = sparse.DevicePlan(plan)
    got = sparse.product_batched(torch.from_numpy(values).cuda(), torch.from_numpy(x).cuda(), dp).cpu().numpy()
    assert _bits(got) == _bits(want)


def test_a_hub_row_above_the_short_stack(torch):
    say("hub row of 12000 entries")
    n = 12000
    plan = sparse.product_plan(n, np.zeros(n - 1, np.int64), np.arange(1, n))
    rng = np.random.default_rng(7)
    values = rng.uniform(-1.0, 1.0, plan.column.size) * 10.0 ** rng.uniform(-8, 8, plan.column.size)
    x = rng.uniform(-1.0, 1.0, n)
    want = ref.product(values, x, ref.product_plan(n, np.zeros(n - 1, np.int64), np.arange(1, n)))
    got = sparse.product(torch.from_numpy(values).cuda(), torch.from_numpy(x).cuda(), plan).cpu().numpy()
    assert _bits(got) == _bits(want)


def test_rows_of_negative_zero_terms_keep_their_sign(torch):
    say("negative zeros")
    n, rows, columns = coordinates(PATTERNS["grid_5x5"])
    plan = sparse.product_plan(n, rows, columns)
    values = np.full(plan.column.size, -0.0)
    got = sparse.product(torch.from_numpy(values).cuda(), torch.ones(n, dtype=torch.float64).cuda(), plan)
    got = got.cpu().numpy()
    assert np.all(got == 0.0) and np.all(np.signbit(got))
