"""The sparse Cholesky's symbolic analysis and numeric passes on the CPU.

Restates the reference's tests/unit/test_sparse_cholesky.py checks (patterns, the O(n^2)
definition of minimum degree, arrowhead, level order, refusals, the pairwise fold with signed zeros,
factor/solve against dense) for the oracle (oracle/symbolic.py, oracle/cells_ref.c), and pins the
product's C++ symbolic analysis (cells_amd.symbolic via the C-ABI) bit for bit to the oracle's.
"""

import numpy as np
import pytest

import oracle
from cells_amd import symbolic as product
from cells_amd.world import lattice
from oracle import symbolic as ref


def spd(pattern, seed):
    rng = np.random.default_rng(seed)
    a = np.where(np.tril(pattern, -1), rng.uniform(-1.0, 1.0, pattern.shape), 0.0)
    a = a + a.T
    a[np.diag_indices_from(a)] = np.abs(a).sum(axis=1) + rng.uniform(0.5, 2.0, a.shape[0])
    return a


def arrowhead(n, hub=0):
    p = np.eye(n, dtype=bool)
    p[hub, :] = p[:, hub] = True
    return p


def grid(m):
    n = m * m
    p = np.eye(n, dtype=bool)
    for r in range(m):
        for c in range(m):
            i = r * m + c
            for j in (i + 1 if c + 1 < m else None, i + m if r + 1 < m else None):
                if j is not None:
                    p[i, j] = p[j, i] = True
    return p


def random_pattern(n, density, seed):
    rng = np.random.default_rng(seed)
    p = np.tril(rng.uniform(size=(n, n)) < density, -1)
    return p | p.T | np.eye(n, dtype=bool)


PATTERNS = {
    "diagonal": np.eye(9, dtype=bool),
    "arrowhead_hub_first": arrowhead(12, hub=0),
    "arrowhead_hub_last": arrowhead(12, hub=11),
    "grid_5x5": grid(5),
    "random_40": random_pattern(40, 0.08, 3),
}


def coordinates(pattern):
    rows, columns = np.nonzero(np.tril(pattern, -1))
    return pattern.shape[0], rows, columns


def reference_minimum_degree(pattern):
    """The definition, scanning every live node at each step (O(n^2))."""
    n = pattern.shape[0]
    adjacency = [set(np.flatnonzero(pattern[i]).tolist()) - {i} for i in range(n)]
    original = [len(a) for a in adjacency]
    alive, order = set(range(n)), []
    for _ in range(n):
        node = min(alive, key=lambda i: (len(adjacency[i]), original[i], i))
        for a in adjacency[node]:
            adjacency[a].discard(node)
            adjacency[a] |= adjacency[node] - {a}
        alive.discard(node)
        order.append(node)
    return order


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_heap_minimum_degree_is_the_definition(seed):
    pattern = random_pattern(60, 0.06, seed)
    want = reference_minimum_degree(pattern)
    assert ref.minimum_degree(*coordinates(pattern)).tolist() == want
    assert product.minimum_degree(*coordinates(pattern)).tolist() == want


def test_a_plausible_tie_break_mistake_misses_the_definition():
    """Ties by index alone (no original degree) give another order on these patterns, so the
    definition test above would catch that mistake."""
    differs = 0
    for seed in (1, 2, 3):
        pattern = random_pattern(60, 0.06, seed)
        n = pattern.shape[0]
        adjacency = [set(np.flatnonzero(pattern[i]).tolist()) - {i} for i in range(n)]
        alive, order = set(range(n)), []
        for _ in range(n):
            node = min(alive, key=lambda i: (len(adjacency[i]), i))
            for a in adjacency[node]:
                adjacency[a].discard(node)
                adjacency[a] |= adjacency[node] - {a}
            alive.discard(node)
            order.append(node)
        differs += order != reference_minimum_degree(pattern)
    assert differs > 0


def test_hub_goes_last_with_no_fill():
    pattern = arrowhead(30, hub=0)
    for bind in (ref.bind, product.bind):
        s = bind(*coordinates(pattern))
        assert int(s.order[-1]) == 0
        assert int(s.column_start[-1]) == np.tril(pattern).sum()
        assert int(s.level.max()) + 1 == 2
        natural = bind(*coordinates(pattern), order=np.arange(30))
        assert int(natural.column_start[-1]) == 30 * 31 // 2


def test_every_level_follows_its_children():
    s = product.bind(*coordinates(PATTERNS["random_40"]))
    for k, p in enumerate(s.parent.tolist()):
        if p >= 0:
            assert s.level[p] > s.level[k]
            assert p > k


def _same(a, b):
    assert a.order.tolist() == b.order.tolist()
    assert a.rank.tolist() == b.rank.tolist()
    assert a.column_start.tolist() == b.column_start.tolist()
    assert a.row.tolist() == b.row.tolist()
    assert a.column.tolist() == b.column.tolist()
    assert a.parent.tolist() == b.parent.tolist()
    assert a.level.tolist() == b.level.tolist()


@pytest.mark.parametrize("name", sorted(PATTERNS))
def test_product_symbolic_is_the_oracle_bit_for_bit(name):
    args = coordinates(PATTERNS[name])
    _same(product.bind(*args), ref.bind(*args))


@pytest.mark.parametrize("n", [1, 2, 5, 64, 256, 1000, 2176, 4096])
def test_product_symbolic_on_the_lattices(n):
    snd, rcv = lattice(n)
    _same(product.bind(n, snd, rcv), ref.bind(n, snd, rcv))


def test_lattice_edge_counts():
    for n in (64, 256, 300, 1000, 2176, 4096):
        side = int(np.floor(np.sqrt(n)))
        rem = n - side * side
        snd, rcv = lattice(n)
        assert len(snd) == 2 * side * (side - 1) + rem
        assert np.all(snd < rcv)
        assert len(set(zip(snd.tolist(), rcv.tolist()))) == len(snd)
    assert len(lattice(64)[0]) == 112


@pytest.mark.parametrize("args", [(0, [], []), (3, [0, 1], [1]), (3, [0], [3]), (3, [-1], [0])],
                         ids=["empty", "unequal coordinates", "outside", "negative"])
def test_bind_refuses_a_malformed_pattern(args):
    for bind in (ref.bind, product.bind):
        with pytest.raises(ValueError):
            bind(*args)


def test_bind_refuses_an_order_that_is_not_a_permutation():
    for bind in (ref.bind, product.bind):
        with pytest.raises(ValueError):
            bind(3, [0], [1], order=[0, 0, 1])


def _adjacent_tree(values):
    values = list(values)
    while len(values) > 1:
        values = [values[i] + values[i + 1] if i + 1 < len(values) else values[i]
                  for i in range(0, len(values), 2)]
    return values[0] if values else 0.0


def _stack_tree(values):
    """The kernels' evaluation of the fold: a binary-counter stack of partial sums."""
    st = []
    for t in values:
        st.append([t, 1])
        while len(st) >= 2 and st[-1][1] == st[-2][1]:
            b = st.pop()
            st[-1][0] = st[-1][0] + b[0]
            st[-1][1] *= 2
    if not st:
        return 0.0
    s = st[-1][0]
    for i in range(len(st) - 2, -1, -1):
        s = st[i][0] + s
    return s


def test_the_pairwise_fold_tree_and_its_stack_form_agree_with_signed_zeros():
    rng = np.random.default_rng(11)
    for n in list(range(0, 70)) + [127, 128, 129, 255, 300]:
        v = (rng.uniform(-1, 1, n) * 10.0 ** rng.uniform(-8, 8, n)).tolist()
        a, b, c = _adjacent_tree(v), _stack_tree(v), ref.pairwise_tree_sum(v)
        assert np.float64(a).tobytes() == np.float64(b).tobytes() == np.float64(c).tobytes()
    for n in range(0, 9):
        zeros = [-0.0] * n
        assert np.signbit(_stack_tree(zeros)) == (n > 0)


def _oracle_symbolic(pattern):
    n, rows, cols = coordinates(pattern)
    s = ref.bind(n, rows, cols)

    class P:  # a one-world stand-in for symbolic_structs
        W = 1
        cell_off = np.array([0, n])
        edge_off = np.array([0, len(rows)])
        snd = np.asarray(rows)
        rcv = np.asarray(cols)

    return s, oracle.symbolic_structs(P, [s])


def dense_factor(values, s):
    lower = np.zeros((s.size, s.size))
    lower[s.rank[s.row], s.rank[s.column]] = values
    return lower


@pytest.mark.parametrize("name", sorted(PATTERNS))
def test_oracle_factor_and_solve_against_dense(name):
    pattern = PATTERNS[name]
    a = spd(pattern, seed=len(name))
    b = np.random.default_rng(7).uniform(-1.0, 1.0, a.shape[0])
    s, structs = _oracle_symbolic(pattern)
    values = oracle.factor(structs[0], a[s.row, s.column])
    x = oracle.solve(structs[0], values, b)
    lower = dense_factor(values, s)
    permuted = a[np.ix_(s.order, s.order)]
    assert np.allclose(lower @ lower.T, permuted, rtol=1e-13, atol=1e-13)
    assert np.allclose(x, np.linalg.solve(a, b), rtol=1e-12, atol=1e-13)


@pytest.mark.parametrize("n", [1, 2, 7, 64, 65, 300])
def test_a_diagonal_system_is_the_dense_path_bit_for_bit(n):
    rng = np.random.default_rng(n)
    d = rng.uniform(0.3, 3.0, n) * 10.0 ** rng.uniform(-3, 3, n)
    b = rng.uniform(-1.0, 1.0, n) * 10.0 ** rng.uniform(-20, 0, n)
    b[: n // 3] = 0.0
    b[n // 3: n // 2] = -0.0
    s, structs = _oracle_symbolic(np.eye(n, dtype=bool))
    values = oracle.factor(structs[0], np.diag(d)[s.row, s.column])
    root = np.sqrt(d)
    assert values[s.rank].tobytes() == root.tobytes()
    x = oracle.solve(structs[0], values, b)
    expect = (b * (1.0 / root)) * (1.0 / root)  # the device's reciprocal pivots
    assert x.tobytes() == expect.tobytes()
    assert np.array_equal(np.signbit(x), np.signbit(b))


def test_a_matrix_that_is_not_positive_definite_gives_a_non_finite_factor():
    pattern = arrowhead(6, hub=2)
    a = spd(pattern, 5)
    a[2, 2] = -1.0
    s, structs = _oracle_symbolic(pattern)
    with np.errstate(invalid="ignore"):
        values = oracle.factor(structs[0], a[s.row, s.column])
    assert not np.all(np.isfinite(values))


# ---- relaxed supernodes (the multifrontal plan's pattern) ---------------------------------------
#
# The plan joins chains of supernodes with explicit zeros (cells_relaxed_padding). Pinned by what
# the elimination fixes: the padded pattern keeps the order's elimination tree; every padded entry
# lies outside the filled pattern, so the Cholesky factor of any SPD matrix on the pattern is zero
# there (checked against numpy's dense Cholesky); the oracle's factor on the padded pattern equals
# its factor on the bound one at every true entry; the joined fronts are the padded pattern's
# fundamental supernodes and fewer levels deep.


def _supernode_levels(s):
    n = s.size
    col_len = np.diff(s.column_start)
    nchild = np.bincount(s.parent[s.parent >= 0], minlength=n)
    first, of = [], np.zeros(n, int)
    k = 0
    while k < n:
        j = k
        while j + 1 < n and s.parent[j] == j + 1 and nchild[j + 1] == 1 and col_len[j] == col_len[j + 1] + 1:
            j += 1
        of[k:j + 1] = len(first)
        first.append((k, j + 1))
        k = j + 1
    lev = np.zeros(len(first), int)
    for t, (a, b) in enumerate(first):
        p = s.parent[b - 1]
        if p >= 0:
            lev[of[p]] = max(lev[of[p]], lev[t] + 1)
    return len(first), int(lev.max()) + 1


@pytest.mark.parametrize("n", [1, 2, 3, 7, 64, 300, 1216, 2176])
def test_relaxed_supernodes_keep_the_tree_and_pad_only_structural_zeros(n):
    snd, rcv = lattice(n)
    exact = product.bind(n, snd, rcv)
    order, pr, pc = product.relaxed_padding(n, snd, rcv)
    assert np.array_equal(order, exact.order)
    relaxed = product.bind(n, np.concatenate((snd, pr)), np.concatenate((rcv, pc)), order=order)
    assert np.array_equal(relaxed.parent, exact.parent)
    assert int(relaxed.column_start[-1]) == int(exact.column_start[-1]) + pr.size
    true = set(zip(exact.row.tolist(), exact.column.tolist()))
    assert not any((int(a), int(b)) in true or (int(b), int(a)) in true for a, b in zip(pr, pc))
    assert true <= set(zip(relaxed.row.tolist(), relaxed.column.tolist()))
    fronts0, levels0 = _supernode_levels(exact)
    fronts1, levels1 = _supernode_levels(relaxed)
    assert fronts1 <= fronts0 and levels1 <= levels0
    if n >= 1216:
        assert levels1 < levels0 / 2  # the chains were what made these trees deep
    # an SPD matrix on the pattern: its dense factor vanishes at the padded entries
    rng = np.random.default_rng(n)
    A = np.zeros((n, n))
    w = rng.uniform(0.1, 1.0, snd.size)
    A[snd, rcv] = A[rcv, snd] = -w
    A[np.arange(n), np.arange(n)] = -A.sum(axis=1) + rng.uniform(0.5, 1.5, n)
    P = A[np.ix_(order, order)]
    Ld = np.linalg.cholesky(P)
    rank = np.empty(n, int)
    rank[order] = np.arange(n)
    hi, lo = np.maximum(rank[pr], rank[pc]), np.minimum(rank[pr], rank[pc])
    assert np.all(np.abs(Ld[hi, lo]) <= 1e-14 * np.abs(Ld).max())
    # the oracle's factor on either pattern agrees at the true entries and is zero at the padding
    syms = [ref.bind(n, snd, rcv, order=order),
            ref.bind(n, np.concatenate((snd, pr)), np.concatenate((rcv, pc)), order=order)]
    factors = []
    for s in syms:
        st = oracle.Symbolic()
        keep = [np.ascontiguousarray(x, dtype=np.int64) for x in
                (s.column_start, s.rank[s.row], s.order, s.column_start[s.rank[np.arange(n)]], np.zeros(1))]
        st.n, st.nnz = n, int(s.column_start[-1])
        st.column_start, st.row_rank, st.order, st.diag_pos, st.edge_pos = [
            x.ctypes.data_as(oracle._i64p) for x in keep]
        v = P[s.rank[s.row], s.rank[s.column]]
        factors.append((oracle.factor(st, v), s, keep))
    (f0, s0, _), (f1, s1, _) = factors
    pos1 = {(int(r), int(c)): q for q, (r, c) in enumerate(zip(s1.row, s1.column))}
    at = np.array([pos1[(int(r), int(c))] for r, c in zip(s0.row, s0.column)])
    assert np.max(np.abs(f1[at] - f0)) <= 1e-13 * np.max(np.abs(f0))
    rest = np.ones(f1.size, bool)
    rest[at] = False
    assert np.all(f1[rest] == 0.0)


def lattice_pattern(n):
    snd, rcv = lattice(n)
    p = np.eye(n, dtype=bool)
    p[snd, rcv] = p[rcv, snd] = True
    return p


@pytest.mark.parametrize("pattern", [
    pytest.param(lambda: random_pattern(60, 0.06, 11), id="random_60"),
    pytest.param(lambda: random_pattern(90, 0.03, 12), id="random_90"),
    pytest.param(lambda: grid(9), id="grid_9x9"),
    pytest.param(lambda: lattice_pattern(150), id="lattice_150"),
    pytest.param(lambda: lattice_pattern(300), id="lattice_300"),
])
def test_bound_pattern_is_exactly_the_dense_factor_nonzeros(pattern):
    """bind's pattern is the fill, no more and no less: the nonzeros of the dense Cholesky factor
    of a generic SPD matrix on the pattern (random values, so no entry cancels by accident) in
    bind's order are exactly its entries (L L^T = P A P^T alone would admit extra entries; fill
    entries decay with distance on these diagonally dominant matrices, so the comparison is with
    the exact zeros, not a threshold). The product's bind gives the same pattern."""
    pat = pattern()
    n, rows, cols = coordinates(pat)
    s = ref.bind(n, rows, cols)
    a = spd(pat, seed=n)
    lower = np.linalg.cholesky(a[np.ix_(s.order, s.order)])
    # a structural zero of the dense factor is an exact zero (every product in its sum holds one)
    nonzero = set(zip(*np.nonzero(lower)))
    pattern_ranks = set(zip(s.rank[s.row].tolist(), s.rank[s.column].tolist()))
    assert pattern_ranks == {(int(i), int(j)) for i, j in nonzero}
    assert len(pattern_ranks) == len(s.row)  # no entry twice
    got = product.bind(n, rows, cols)
    assert set(zip(got.rank[got.row].tolist(), got.rank[got.column].tolist())) == pattern_ranks


def test_a_pattern_with_a_dropped_fill_entry_misses_the_dense_nonzeros():
    """The planted mistake the check above must catch: a bind that loses one fill entry (the
    union rule skipping a child) no longer covers the dense factor's nonzeros."""
    pat = grid(6)
    n, rows, cols = coordinates(pat)
    s = ref.bind(n, rows, cols)
    lower = np.linalg.cholesky(spd(pat, seed=5)[np.ix_(s.order, s.order)])
    nonzero = {(int(i), int(j)) for i, j in zip(*np.nonzero(lower))}
    original = {(int(min(a, b)), int(max(a, b))) for a, b in zip(s.rank[rows], s.rank[cols])}
    fill = [(int(r), int(c)) for r, c in zip(s.rank[s.row], s.rank[s.column])
            if r != c and (int(c), int(r)) not in original]
    assert fill, "the grid's elimination creates fill"
    dropped = set(zip(s.rank[s.row].tolist(), s.rank[s.column].tolist())) - {fill[0]}
    assert dropped != nonzero
