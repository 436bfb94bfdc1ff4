"""The multifrontal schedule the device factor follows, restated in numpy on the CPU.

Fundamental supernodes (column j joins j + 1 when j + 1 is j's parent, has no other child and
holds j's rows but j + 1 itself), fronts assembled from the supernode's columns plus the
children's update matrices through the children's row positions in the parent's front, a
right-looking partial factor over the supernode's columns (pivot sqrt, column times the pivot's
reciprocal), the trailing triangle passed up. The result must be the factor on the bound pattern:
checked against the oracle's factor (the reference's adjacent-pairs order) and a dense Cholesky
of the permuted matrix on lattices of the batched configuration's sizes, and every front's rows
must sit in its parent's front (the extend-add is a gather with no misses).
"""

import numpy as np
import pytest

import oracle
from cells_amd.world import lattice
from oracle import symbolic as sy


def supernodes(s):
    n = s.size
    st, par = s.structure, s.parent
    kids = np.zeros(n, int)
    for k in range(n):
        if par[k] >= 0:
            kids[par[k]] += 1
    out, k = [], 0
    while k < n:
        j = k
        while j + 1 < n and par[j] == j + 1 and kids[j + 1] == 1 and len(st[j]) == len(st[j + 1]) + 1:
            j += 1
        out.append((k, j + 1))
        k = j + 1
    return out


def multifrontal(s, A):
    """L (dense, by rank) of the permuted SPD matrix A (by rank) through the fronts."""
    n = s.size
    L = np.zeros((n, n))
    sns = supernodes(s)
    of = {}
    for t, (a, b) in enumerate(sns):
        for j in range(a, b):
            of[j] = t
    rows = [list(range(a, b)) + list(s.structure[b - 1]) for a, b in sns]
    update = {}
    children = {t: [] for t in range(len(sns))}
    for t, (a, b) in enumerate(sns):
        p = s.parent[b - 1]
        if p >= 0:
            children[of[p]].append(t)
    for t, (a, b) in enumerate(sns):  # children precede parents in column order
        R = rows[t]
        m, J = len(R), b - a
        F = np.zeros((m, m))
        F[:, :J] = np.tril(A[np.ix_(R, R[:J])])
        pos = {r: i for i, r in enumerate(R)}
        for c in children[t]:
            Sc = rows[c][sns[c][1] - sns[c][0]:]
            rel = [pos[r] for r in Sc]  # a KeyError here is a front row missing from the parent
            U = update.pop(c)
            for bb in range(len(Sc)):
                for aa in range(bb, len(Sc)):
                    F[rel[aa], rel[bb]] += U[aa, bb]
        for jj in range(J):
            d = np.sqrt(F[jj, jj])
            F[jj + 1:, jj] *= 1.0 / d
            F[jj, jj] = d
            for k in range(jj + 1, m):
                F[k:, k] -= F[k:, jj] * F[k, jj]
        for jj in range(J):
            L[R[jj:], a + jj] = F[jj:, jj]
        if m > J:
            update[t] = F[J:, J:]
    assert not update
    return L


@pytest.mark.parametrize("n", [7, 64, 300, 1024, 2176])
def test_fronts_give_the_factor(n):
    snd, rcv = lattice(n)
    s = sy.bind(n, list(snd), list(rcv))
    rng = np.random.default_rng(n)
    off = rng.uniform(-1, 0, len(snd))
    A = np.zeros((n, n))
    A[snd, rcv] = off
    A[rcv, snd] = off
    A[np.arange(n), np.arange(n)] = 0.5 + rng.uniform(0.5, 2.0, n) - A.sum(axis=1)
    P = A[np.ix_(s.order, s.order)]  # by rank
    L = multifrontal(s, P)
    assert np.allclose(L @ L.T, P, rtol=0, atol=1e-12 * np.abs(P).max())
    # the pattern: nothing outside the symbolic structure, and the dense Cholesky agrees
    mask = np.zeros((n, n), bool)
    for k in range(n):
        mask[k, k] = True
        mask[list(s.structure[k]), k] = True
    assert not np.any(L[~mask])
    ref = np.linalg.cholesky(P)
    assert np.max(np.abs(L - ref)) <= 1e-12 * np.max(np.abs(ref))
    # and the oracle's factor (adjacent-pairs order) on the same positions
    if n >= 64:
        from cells_amd.world import build, config1

        p = build(config1(worlds=1, cells=(n, n), seed=1))
        syms, sym = oracle.world_symbolics(p)
        cs = syms[0].column_start
        vals = np.zeros(int(cs[-1]))
        # positions by rank: column k's entries are its diagonal then its structure, ascending
        for k in range(n):
            rr = [k] + list(s.structure[k])
            vals[cs[k]:cs[k + 1]] = P[rr, k]
        f = oracle.factor(sym[0], vals)
        got = np.concatenate([L[[k] + list(s.structure[k]), k] for k in range(n)])
        assert np.max(np.abs(got - f)) <= 1e-12 * np.max(np.abs(f))


def test_every_supernode_fits_a_workgroup_front_at_the_batched_sizes():
    """Fronts of the lattices U{256..4096} stay within order 127 (the workgroup's LDS)."""
    for n in (256, 1000, 2176, 3000, 4096):
        snd, rcv = lattice(n)
        s = sy.bind(n, list(snd), list(rcv))
        worst = max(b - a + len(s.structure[b - 1]) for a, b in supernodes(s))
        assert worst <= 127, (n, worst)


def tri_locate(T, e):
    """The device's decode of entry e of a packed lower column-major triangle of order T (float32
    root of the column-start quadratic, then one correction either way)."""
    h = np.float32(2 * T + 1)
    b = int((h - np.sqrt(h * h - np.float32(8) * np.float32(e))) * np.float32(0.5))
    b = min(max(b, 0), T - 1)
    if b + 1 < T and (b + 1) * T - (b + 1) * b // 2 <= e:
        b += 1
    if b * T - b * (b - 1) // 2 > e:
        b -= 1
    return b, e - (b * T - b * (b - 1) // 2) + b


def test_packed_triangle_decode_is_exact_up_to_order_200():
    for T in range(1, 201):
        e = 0
        for b in range(T):
            for a in range(b, T):
                assert tri_locate(T, e) == (b, a), (T, e)
                e += 1


def supernodal_solve(s, L, b):
    """x = L^-T L^-1 b (by rank) through the fronts: forward with update vectors passed to the
    parent, backward with the rows below gathered from the ancestors' solution."""
    n = s.size
    sns = supernodes(s)
    of = {}
    for t, (a, bb) in enumerate(sns):
        for j in range(a, bb):
            of[j] = t
    rows = [list(range(a, bb)) + list(s.structure[bb - 1]) for a, bb in sns]
    children = {t: [] for t in range(len(sns))}
    for t, (a, bb) in enumerate(sns):
        p = s.parent[bb - 1]
        if p >= 0:
            children[of[p]].append(t)
    y = b.copy()
    vec = {}
    for t, (a, bb) in enumerate(sns):  # forward, children first
        R = rows[t]
        J = bb - a
        f = np.zeros(len(R))
        f[:J] = y[a:bb]
        pos = {r: i for i, r in enumerate(R)}
        for c in children[t]:
            Sc = rows[c][sns[c][1] - sns[c][0]:]
            for aa, r in enumerate(Sc):
                f[pos[r]] += vec[c][aa]
        for jj in range(J):
            x = f[jj] * (1.0 / L[a + jj, a + jj])
            f[jj + 1:] -= L[R[jj + 1:], a + jj] * x
            y[a + jj] = x
        vec[t] = f[J:]
    for t in range(len(sns) - 1, -1, -1):  # backward, parents first
        a, bb = sns[t]
        R = rows[t]
        J = bb - a
        g = y[R[J:]]
        xr = np.array([y[a + k] - np.dot(L[R[J:], a + k], g) for k in range(J)])
        for jj in range(J - 1, -1, -1):
            z = xr[jj] * (1.0 / L[a + jj, a + jj])
            y[a + jj] = z
            xr[:jj] -= L[a + jj, a:a + jj] * z
    return y


@pytest.mark.parametrize("n", [7, 64, 300, 1024, 2176])
def test_supernodal_solve_meets_a_dense_solve(n):
    snd, rcv = lattice(n)
    s = sy.bind(n, list(snd), list(rcv))
    rng = np.random.default_rng(n + 1)
    off = rng.uniform(-1, 0, len(snd))
    A = np.zeros((n, n))
    A[snd, rcv] = off
    A[rcv, snd] = off
    A[np.arange(n), np.arange(n)] = 0.5 + rng.uniform(0.5, 2.0, n) - A.sum(axis=1)
    P = A[np.ix_(s.order, s.order)]
    L = np.linalg.cholesky(P)
    b = rng.uniform(-1, 1, n)
    x = supernodal_solve(s, L, b)
    ref = np.linalg.solve(P, b)
    assert np.max(np.abs(x - ref)) <= 1e-12 * np.max(np.abs(ref))
