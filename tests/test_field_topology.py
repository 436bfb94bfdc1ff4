"""bind's capacitance topology (reference numerics/surface_field.py:238-340) through the C-ABI.

Pinned three ways: (1) the reference's own field fixtures (validation/tissue/surface_field_fixtures.py
and surface_fixtures.py: the "open" layout of 2 cells and 5 outside nodes and its "closed",
"one_owner", "anchored", "all_clamped" and "split" variants, restated here as data) with their
components, representatives and offsets worked out by hand; (2) the reference's refusal cases
(tests/unit/test_surface_field.py:275-291: return path, missing gauge, extra gauge, inactive
reservoir, no charged path) and its representative test (:326-331: the closed layout fixes outside
nodes 0 and 4); (3) random layouts against an independent breadth-first search of the same rules.
"""

import copy

import numpy as np
import pytest

from cells_amd.field_topology import bind_topology


def floating_index(t):
    """Each node's floating component (-1 where none) from the floating basis."""
    b = t.floating_basis
    if b.shape[1] == 0:
        return -np.ones(b.shape[0], int)
    return np.where(b.any(axis=1), b.argmax(axis=1), -1)


def open_layout():
    return dict(
        valence=[[1.0], [1.0], [-1.0]],
        cell_active=[True, True],
        outside_active=[True] * 5,
        reservoir=[False] * 4 + [True],
        membrane_active=[True] * 4,
        membrane_cell=[0, 0, 1, 1],
        membrane_outside=[0, 1, 2, 3],
        face_active=[True] * 5,
        face_sender=[0, 2, 1, 4, 0],
        face_receiver=[1, 3, 4, 2, 3],
        face_factor=[[1e-19, 2e-19, 1e-19, 2e-19, 1e-19], [2e-19, 3e-19, 2e-19, 1e-19, 3e-19],
                     [1.5e-19, 1e-19, 2e-19, 1e-19, 2e-19]],
    )


def case(name):
    p, gauge = open_layout(), [False] * 5
    if name in ("closed", "one_owner"):
        p["reservoir"] = [False] * 5
        gauge[4] = True
        if name == "one_owner":
            p["membrane_outside"] = [0, 0, 2, 2]
    elif name == "anchored":
        p["reservoir"][1] = True
    elif name == "all_clamped":
        p["reservoir"] = [True] * 5
    elif name == "split":
        for key in ("membrane_active", "membrane_cell", "membrane_outside"):
            p[key].append(p[key][0])
    return p, np.asarray(gauge)


# nodes: cells 0, 1, then outside nodes 0..4 at 2..6
EXPECTED = {
    # components {c0, o0, o1}, {c1, o2, o3} float (representatives o0, o2), {o4} is the held bath
    "open": ([0, 1, 0, 0, 1, 1, 2], [0, 0, 1, 0, 1, 0, 1], [0, 1, 0, 0, 1, 1, -1], [1, 1]),
    "split": ([0, 1, 0, 0, 1, 1, 2], [0, 0, 1, 0, 1, 0, 1], [0, 1, 0, 0, 1, 1, -1], [1, 1]),
    # no bath: {o4} floats with the gauge as its representative, whose offset is not free
    "closed": ([0, 1, 0, 0, 1, 1, 2], [0, 0, 1, 0, 1, 0, 1], [0, 1, 0, 0, 1, 1, 2], [1, 1, 0]),
    # each cell owns one outside node; o1 and o3 are components of their own
    "one_owner": ([0, 1, 0, 2, 1, 3, 4], [0, 0, 1, 1, 1, 1, 1], [0, 1, 0, 2, 1, 3, 4], [1, 1, 1, 1, 0]),
    # o1 clamps the first cell's component
    "anchored": ([0, 1, 0, 0, 1, 1, 2], [0, 0, 0, 1, 1, 0, 1], [-1, 0, -1, -1, 0, 0, -1], [1]),
    "all_clamped": ([0, 1, 0, 0, 1, 1, 2], [0, 0, 1, 1, 1, 1, 1], [-1] * 7, []),
}


@pytest.mark.parametrize("name", sorted(EXPECTED))
def test_the_reference_fixtures_topologies(name):
    p, gauge = case(name)
    t = bind_topology(p, gauge)
    component, fixed, floating, free = EXPECTED[name]
    assert t.component.tolist() == component
    assert t.fixed.astype(int).tolist() == fixed
    got_floating = floating_index(t)
    assert got_floating.tolist() == floating
    assert t.offset_free.astype(int).tolist() == free


def test_the_closed_layouts_representatives_are_outside_nodes_0_and_4():
    p, gauge = case("closed")
    fixed = bind_topology(p, gauge).fixed
    assert fixed[2 + 0] and fixed[2 + 4]


@pytest.mark.parametrize("fault, message", [
    ("return_path", "return path to a clamp/gauge"),
    ("missing_gauge", "exactly one active outside gauge"),
    ("extra_gauge", "exactly one active outside gauge"),
    ("inactive_reservoir", "Reservoir must be an active outside owner"),
    ("no_charged_path", "return path to a clamp/gauge"),
])
def test_the_reference_refusals(fault, message):
    p, gauge = case("open")
    if fault == "return_path":
        p["face_factor"] = [[0.0] * 5 for _ in range(3)]
    elif fault == "missing_gauge":
        p["reservoir"] = [False] * 5
    elif fault == "extra_gauge":
        gauge[0] = True
    elif fault == "inactive_reservoir":
        p["outside_active"][4] = False
    else:
        p["valence"] = [[0.0]] * 3
    with pytest.raises(ValueError, match=message):
        bind_topology(p, gauge)


def test_a_return_path_admits_a_component_without_a_face():
    """A floating component whose only path to the clamp is a forms face's return path is admitted
    (surface_field.py:224-236)."""
    p, gauge = case("open")
    p["face_factor"] = [[0.0] * 5 for _ in range(3)]
    with pytest.raises(ValueError):
        bind_topology(p, gauge)
    t = bind_topology(p, gauge, return_paths=[(1, 4), (4, 2)])
    assert t.offset_free.tolist() == [True, True]


def _bfs_topology(n, o, mc, mo, ma, cact, oact, res, gauge, fs, fr, fc, rp):
    """The same rules by breadth-first search (an independent check of the union-find)."""
    size = n + o
    active = np.r_[cact, oact]
    adj = [[] for _ in range(size)]
    for c, x, a in zip(mc, mo, ma):
        if a:
            adj[c].append(n + x)
            adj[n + x].append(c)
    comp = -np.ones(size, int)
    members = []
    for s in range(size):
        if not active[s] or comp[s] >= 0:
            continue
        comp[s] = len(members)
        todo, mem = [s], [s]
        while todo:
            u = todo.pop()
            for v in adj[u]:
                if comp[v] < 0:
                    comp[v] = comp[s]
                    todo.append(v)
                    mem.append(v)
        members.append(sorted(mem))
    fixed = ~active.copy()
    fixed[n + np.flatnonzero(res)] = True
    floating, free = -np.ones(size, int), []
    for k, mem in enumerate(members):
        if any(i >= n and res[i - n] for i in mem):
            continue
        g = [i for i in mem if i >= n and gauge[i - n]]
        outs = [i for i in mem if i >= n]
        if not g and not outs:
            return None
        rep = g[0] if g else outs[0]
        fixed[rep] = True
        for i in mem:
            floating[i] = len(free)
        free.append(not g)
    reached = {comp[n + k] for k in range(o) if (res[k] or gauge[k])}
    links = [(comp[n + a], comp[n + b]) for a, b, c in zip(fs, fr, fc) if c] + [(comp[n + a], comp[n + b]) for a, b in rp]
    graph = [[] for _ in members]
    for a, b in links:
        graph[a].append(b)
        graph[b].append(a)
    todo = list(reached)
    while todo:
        u = todo.pop()
        for v in graph[u]:
            if v not in reached:
                reached.add(v)
                todo.append(v)
    if len(reached) != len(members):
        return None
    return comp, fixed, floating, free


def test_random_layouts_against_breadth_first_search():
    rng = np.random.default_rng(17)
    agreed = refused = 0
    for trial in range(300):
        n, o = int(rng.integers(1, 7)), int(rng.integers(1, 9))
        m, f = int(rng.integers(0, 10)), int(rng.integers(0, 12))
        mc, mo = rng.integers(0, n, m), rng.integers(0, o, m)
        oact = rng.uniform(size=o) < 0.9
        cact = np.ones(n, bool)
        res = (rng.uniform(size=o) < 0.25) & oact
        gauge = np.zeros(o, bool)
        if not res.any() and oact.any():
            gauge[int(rng.choice(np.flatnonzero(oact)))] = True
        ma = oact[mo].copy()
        fs, fr = rng.integers(0, o, f), rng.integers(0, o, f)
        fa = oact[fs] & oact[fr] & (fs != fr)
        fac = rng.uniform(size=(1, f)) * (rng.uniform(size=(1, f)) < 0.7)
        layout = dict(valence=[[1.0]], cell_active=cact, outside_active=oact, reservoir=res, membrane_active=ma,
                      membrane_cell=mc, membrane_outside=mo, face_active=fa, face_sender=fs, face_receiver=fr,
                      face_factor=fac)
        want = _bfs_topology(n, o, mc, mo, ma, cact, oact, res, gauge, fs, fr, fa & (fac[0] > 0), [])
        if not oact.any():
            continue
        if want is None:
            with pytest.raises(ValueError):
                bind_topology(layout, gauge)
            refused += 1
            continue
        t = bind_topology(layout, gauge)
        comp, fixed, floating, free = want
        assert t.component.tolist() == comp.tolist(), trial
        assert t.fixed.tolist() == fixed.tolist(), trial
        got = floating_index(t)
        assert got.tolist() == floating.tolist(), trial
        assert t.offset_free.tolist() == free, trial
        agreed += 1
    assert agreed > 50 and refused > 20, (agreed, refused)


def test_grid_worlds_without_a_charged_face_path_are_refused_at_create():
    """cells_field_create derives the grid world's capacitor components with the same topology:
    with every face factor zero no interior component reaches the held ring, and bind's refusal
    comes back (before any device is touched)."""
    import ctypes

    from cells_amd import lib
    from cells_amd.field import build_field, small_field

    fp = build_field(small_field())
    fp.face_factor[:] = 0.0
    w, g = lib.world_struct(fp.cells), lib.field_grid_struct(fp)
    h = ctypes.c_void_p()
    L = lib.load()
    assert L.cells_field_create(ctypes.byref(w), ctypes.byref(g), 50, 0, ctypes.byref(h)) == lib.CELLS_EINVAL
    assert b"return path to a clamp/gauge" in L.cells_last_error()
