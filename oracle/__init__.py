"""Oracle (test infrastructure only): ctypes access to oracle/_build/libcells_ref.so.

Only tests, the smoke check and the benchmark's checker legs import this package: its CPU
baseline and its parity checks, each outside every timed region. It is the checker, never the
thing measured or shipped: the product (cells_amd) must not import it.
"""

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "libcells_ref.so")

_i64p = ctypes.POINTER(ctypes.c_int64)
_i32p = ctypes.POINTER(ctypes.c_int32)
_f64p = ctypes.POINTER(ctypes.c_double)


class World(ctypes.Structure):
    _fields_ = [
        ("S", ctypes.c_int32), ("W", ctypes.c_int32), ("water", ctypes.c_int32),
        ("calcium", ctypes.c_int32), ("ncell", ctypes.c_int64), ("nedge", ctypes.c_int64),
        ("cell_off", _i64p), ("edge_off", _i64p),
        ("valence", ctypes.c_double * 4), ("bath_c", ctypes.c_double * 4),
        ("temperature", ctypes.c_double), ("bath_V", ctypes.c_double),
        ("atp", ctypes.c_double), ("adp", ctypes.c_double), ("pi", ctypes.c_double),
        ("km_na", ctypes.c_double), ("km_k", ctypes.c_double), ("km_atp", ctypes.c_double),
        ("dg0", ctypes.c_double), ("ca_km_ca", ctypes.c_double), ("ca_km_atp", ctypes.c_double),
        ("ca_dg0", ctypes.c_double), ("gate_threshold", ctypes.c_double),
        ("gate_minimum", ctypes.c_double), ("water_K", ctypes.c_double),
        ("water_P0", ctypes.c_double), ("water_vmin", ctypes.c_double),
        ("water_vmax", ctypes.c_double), ("bath_pressure", ctypes.c_double),
        ("vol0", _f64p), ("cap", _f64p), ("pam", _f64p), ("nak_cap", _f64p), ("ca_cap", _f64p),
        ("n0", _f64p), ("q0", _f64p), ("lpa", _f64p), ("imperm", _f64p),
        ("snd", _i64p), ("rcv", _i64p), ("paj", _f64p),
        ("inc_ptr", _i64p), ("inc_edge", _i64p), ("inc_out", _i32p),
    ]


class Control(ctypes.Structure):
    _fields_ = [(k, ctypes.c_double) for k in ("membrane", "junction", "pump", "calcium_pump", "water")]


class Symbolic(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64), ("nnz", ctypes.c_int64), ("column_start", _i64p),
                ("row_rank", _i64p), ("order", _i64p), ("diag_pos", _i64p), ("edge_pos", _i64p)]


def build():
    """Compile the oracle with its Makefile (gcc)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = ctypes.CDLL(LIB)
        for name in ("ref_bernoulli", "ref_bernoulli_slope", "ref_bernoulli_derivative", "ref_logaddexp"):
            getattr(_lib, name).restype = ctypes.c_double
        for name in ("ref_selective_flux", "ref_route_slope", "ref_gate_rhs"):
            getattr(_lib, name).restype = ctypes.c_double
        for name in ("ref_lag_weight", "ref_gate_advance"):
            getattr(_lib, name).restype = ctypes.c_double
        _lib.ref_lag_weight.argtypes = [ctypes.c_double]
        _lib.ref_gate_advance.argtypes = [ctypes.c_double] * 6
        _lib.ref_membrane_rate.restype = ctypes.c_double
        _lib.ref_membrane_rate.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.c_int64,
                                           ctypes.c_void_p]
    return _lib


def _p(a, t):
    return a.ctypes.data_as(t)


def world_struct(p):
    """A World over the arrays of a cells_amd.world.Packed (kept alive on the struct)."""
    c = p.config
    keep = {}

    def arr(name, dtype):
        a = np.ascontiguousarray(getattr(p, name), dtype=dtype)
        keep[name] = a
        return a

    w = World()
    w.S, w.W, w.water, w.calcium = p.S, p.W, int(p.water), int(p.calcium)
    w.ncell, w.nedge = p.ncell, p.nedge
    w.cell_off = _p(arr("cell_off", np.int64), _i64p)
    w.edge_off = _p(arr("edge_off", np.int64), _i64p)
    for s in range(p.S):
        w.valence[s] = p.valence[s]
        w.bath_c[s] = p.bath_c[s]
    w.temperature, w.bath_V = c.temperature, c.bath_V
    for k in ("atp", "adp", "pi", "km_na", "km_k", "km_atp", "dg0", "ca_km_ca", "ca_km_atp",
              "ca_dg0", "gate_threshold", "gate_minimum", "water_K", "water_P0", "water_vmin",
              "water_vmax", "bath_pressure"):
        setattr(w, k, getattr(c, k))
    for k in ("vol0", "cap", "pam", "nak_cap", "ca_cap", "n0", "q0", "lpa", "imperm", "paj"):
        setattr(w, k, _p(arr(k, np.float64), _f64p))
    for k in ("snd", "rcv", "inc_ptr", "inc_edge"):
        setattr(w, k, _p(arr(k, np.int64), _i64p))
    w.inc_out = _p(arr("inc_out", np.int32), _i32p)
    w._keep = keep
    return w


def control(membrane=1.0, junction=1.0, pump=1.0, calcium_pump=1.0, water=1.0):
    return Control(membrane, junction, pump, calcium_pump, water)


def step_euler(p, y, dt, ctl=None, world=None):
    w = world if world is not None else world_struct(p)
    ctl = ctl or control()
    y = np.ascontiguousarray(y, dtype=np.float64)
    out = np.empty_like(y)
    valid = np.zeros(p.W, dtype=np.int32)
    lib().ref_step_euler(ctypes.byref(w), _p(y, _f64p), _p(out, _f64p), ctypes.c_double(dt),
                         ctypes.byref(ctl), _p(valid, _i32p))
    return out, valid.astype(bool)


def step_ars222(p, y, dt, ctl=None, world=None):
    """One ARS(2,2,2) step of every world (its implicit part empty on a held bath)."""
    return _explicit_step("ref_step_ars222", p, y, dt, ctl, world)


def step_rk4(p, y, dt, ctl=None, world=None):
    """One classical RK4 step of every world (the reference's empty tableau)."""
    return _explicit_step("ref_step_rk4", p, y, dt, ctl, world)


def _explicit_step(name, p, y, dt, ctl, world):
    w = world if world is not None else world_struct(p)
    ctl = ctl or control()
    y = np.ascontiguousarray(y, dtype=np.float64)
    out = np.empty_like(y)
    valid = np.zeros(p.W, dtype=np.int32)
    getattr(lib(), name)(ctypes.byref(w), _p(y, _f64p), _p(out, _f64p), ctypes.c_double(dt), ctypes.byref(ctl),
                         _p(valid, _i32p))
    return out, valid.astype(bool)


def symbolic_structs(p, symbolics):
    """Symbolic structs per world from oracle.symbolic.bind results (one per world, in order),
    with each lattice entry's factor position."""
    out = (Symbolic * p.W)()
    keep = []
    for k, s in enumerate(symbolics):
        c0, c1 = int(p.cell_off[k]), int(p.cell_off[k + 1])
        e0, e1 = int(p.edge_off[k]), int(p.edge_off[k + 1])
        n = c1 - c0
        row_rank = s.rank[s.row]
        diag_pos = s.column_start[s.rank[np.arange(n)]]
        a = s.rank[p.snd[e0:e1] - c0]
        b = s.rank[p.rcv[e0:e1] - c0]
        lo, hi = np.minimum(a, b), np.maximum(a, b)
        pos = {}
        for kk in range(n):
            for q in range(int(s.column_start[kk]), int(s.column_start[kk + 1])):
                pos[(int(row_rank[q]), kk)] = q
        edge_pos = np.array([pos[(int(h), int(l))] for h, l in zip(hi, lo)], dtype=np.int64)
        arrays = [np.ascontiguousarray(x, dtype=np.int64) for x in
                  (s.column_start, row_rank, s.order, diag_pos, edge_pos)]
        keep.append(arrays)
        out[k].n, out[k].nnz = n, int(s.column_start[-1])
        (out[k].column_start, out[k].row_rank, out[k].order, out[k].diag_pos,
         out[k].edge_pos) = [_p(x, _i64p) for x in arrays]
    out._keep = keep
    return out


def step_ros2(p, y, dt, sym, ctl=None, world=None):
    w = world if world is not None else world_struct(p)
    ctl = ctl or control()
    y = np.ascontiguousarray(y, dtype=np.float64)
    out = np.empty_like(y)
    valid = np.zeros(p.W, dtype=np.int32)
    lib().ref_step_ros2(ctypes.byref(w), sym, _p(y, _f64p), _p(out, _f64p), ctypes.c_double(dt),
                        ctypes.byref(ctl), _p(valid, _i32p))
    return out, valid.astype(bool)


def step_n2(p, y, dt, sym, ctl=None, world=None):
    """One N2 step (the reference's ARS(2,2,2) with implicit slope conductances) of every world."""
    w = world if world is not None else world_struct(p)
    ctl = ctl or control()
    y = np.ascontiguousarray(y, dtype=np.float64)
    out = np.empty_like(y)
    valid = np.zeros(p.W, dtype=np.int32)
    lib().ref_step_n2(ctypes.byref(w), sym, _p(y, _f64p), _p(out, _f64p), ctypes.c_double(dt),
                      ctypes.byref(ctl), _p(valid, _i32p))
    return out, valid.astype(bool)


def factor(sym_one, values):
    v = np.ascontiguousarray(values, dtype=np.float64).copy()
    lib().ref_factor(ctypes.byref(sym_one), _p(v, _f64p))
    return v


def solve(sym_one, factor_values, rhs):
    f = np.ascontiguousarray(factor_values, dtype=np.float64)
    b = np.ascontiguousarray(rhs, dtype=np.float64)
    x = np.empty_like(b)
    lib().ref_solve(ctypes.byref(sym_one), _p(f, _f64p), _p(b, _f64p), _p(x, _f64p))
    return x


def rates(p, y, ctl=None, world=None):
    """Explicit rates of every field at y (the state's own potentials), world supplies included."""
    w = world if world is not None else world_struct(p)
    ctl = ctl or control()
    y = np.ascontiguousarray(y, dtype=np.float64)
    r = np.zeros_like(y)
    phi = np.zeros(p.ncell)
    valid = np.zeros(p.W, dtype=bool)
    for k in range(p.W):
        lib().ref_potentials(ctypes.byref(w), k, _p(y, _f64p), _p(phi, _f64p))
        valid[k] = bool(lib().ref_rates(ctypes.byref(w), k, _p(y, _f64p), _p(phi, _f64p),
                                        ctypes.byref(ctl), _p(r, _f64p), None, None))
    return r, valid


# IMEX Euler's energy bound (models/surface_pump.py:63); ROS2, which the reference lacks, takes it
# too (its stiff start books energy errors above ARS(2,2,2)'s 2e-3 in the first steps)
ENERGY_RTOL = {"euler": 2e-2, "ros2": 2e-2, "n2": 2e-3, "ars222": 2e-3, "rk4": 0.0}  # surface_pump.py:63-64
LEDGER_RATIOS = ("local", "total", "global_charge", "water", "outside", "charge", "electrode",
                 "energy", "reservoir_work", "electrode_work", "metabolic_work", "bath_water_work")
ENERGY_TERMS = ("mixing", "capacitor", "heat", "work", "water", "residual", "flows")


def ledgers(p, prev, y, steps, method, material_atol=1e-12, world=None):
    """The row checks of every world's state y against prev: (bits [W], ratios [W, 12] in
    LEDGER_RATIOS order, energy terms [W, 7] in ENERGY_TERMS order)."""
    w = world if world is not None else world_struct(p)
    prev = np.ascontiguousarray(prev, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    bits = np.zeros(p.W, dtype=np.int64)
    ratios = np.zeros((p.W, len(LEDGER_RATIOS)))
    terms = np.zeros((p.W, len(ENERGY_TERMS)))
    for k in range(p.W):
        r = np.zeros(len(LEDGER_RATIOS))
        t = np.zeros(len(ENERGY_TERMS))
        bits[k] = lib().ref_ledgers(ctypes.byref(w), k, _p(prev, _f64p), _p(y, _f64p),
                                    ctypes.c_double(steps), ctypes.c_double(material_atol),
                                    ctypes.c_double(ENERGY_RTOL[method]), _p(r, _f64p), _p(t, _f64p))
        ratios[k], terms[k] = r, t
    return bits, ratios, terms


ENERGY_RATES = ("chemical", "capacitor", "dissipation", "reservoir_power", "electrode_power", "metabolic",
                "energy_balance", "electrode_current")


def energy_rates(p, y, ctl=None, world=None):
    """The readouts' energy rates [W, 8] (ENERGY_RATES order) of every world at y."""
    w = world if world is not None else world_struct(p)
    ctl = ctl or control()
    y = np.ascontiguousarray(y, dtype=np.float64)
    out = np.zeros((p.W, len(ENERGY_RATES)))
    for k in range(p.W):
        o = np.zeros(len(ENERGY_RATES))
        lib().ref_energy_rates(ctypes.byref(w), k, _p(y, _f64p), ctypes.byref(ctl), _p(o, _f64p))
        out[k] = o
    return out


def world_symbolics(p):
    """oracle.symbolic.bind of every world's lattice pattern, and their structs."""
    from oracle.symbolic import bind

    syms = []
    cache = {}
    for k in range(p.W):
        c0, c1 = int(p.cell_off[k]), int(p.cell_off[k + 1])
        e0, e1 = int(p.edge_off[k]), int(p.edge_off[k + 1])
        n = c1 - c0
        if n not in cache:
            cache[n] = bind(n, p.snd[e0:e1] - c0, p.rcv[e0:e1] - c0)
        syms.append(cache[n])
    return syms, symbolic_structs(p, syms)


def ros2_vjp(p, y, dt, sym, ybar, ctl=None, world=None):
    """Reverse of one ROS2 step at y: (cotangent on y's amounts and gates, pam_bar [S, n],
    nak_cap_bar [n], paj_bar [S, E]) for a cotangent ybar on the step's output."""
    w = world if world is not None else world_struct(p)
    ctl = ctl or control()
    y = np.ascontiguousarray(y, dtype=np.float64)
    yb = np.ascontiguousarray(ybar, dtype=np.float64)
    out = np.zeros_like(y)
    pam_bar = np.zeros((p.S, p.ncell))
    cap_bar = np.zeros(p.ncell)
    paj_bar = np.zeros((p.S, p.nedge))
    ok = lib().ref_ros2_vjp(ctypes.byref(w), sym, _p(y, _f64p), ctypes.c_double(dt), ctypes.byref(ctl),
                            _p(yb, _f64p), _p(out, _f64p), _p(pam_bar, _f64p), _p(cap_bar, _f64p),
                            _p(paj_bar, _f64p))
    if not ok:
        raise ValueError("the step at y is refused, or the world has water or calcium")
    return out, pam_bar, cap_bar, paj_bar


# ---------------------------------------------------------------- spatial worlds

class FieldGrid(ctypes.Structure):
    _fields_ = [("side", ctypes.c_int32), ("pad", ctypes.c_int32), ("cell_node", _i32p),
                ("node_volume", ctypes.c_double), ("face_factor", ctypes.c_double * 4),
                ("node_n0", _f64p), ("node_q0", _f64p)]


def field_grid_struct(fp):
    g = FieldGrid()
    keep = [np.ascontiguousarray(fp.cell_node, dtype=np.int32),
            np.ascontiguousarray(fp.node_n0, dtype=np.float64),
            np.ascontiguousarray(fp.node_q0, dtype=np.float64)]
    g.side = fp.side
    g.cell_node = _p(keep[0], _i32p)
    g.node_volume = fp.node_volume
    for s in range(fp.S):
        g.face_factor[s] = fp.face_factor[s]
    g.node_n0 = _p(keep[1], _f64p)
    g.node_q0 = _p(keep[2], _f64p)
    g._keep = keep
    return g


class FieldOperator:
    """The oracle's operator system of a spatial world (cells, then interior nodes row-major):
    its minimum-degree symbolic (oracle.symbolic.bind, the reference's sparse solver), the
    entry positions, and the frozen operator (weights, scales, factors) the steps carry."""

    def __init__(self, fp):
        from oracle.symbolic import bind

        G, n = fp.side, fp.config.cells
        ni = (G - 2) * (G - 2)
        self.nsys = n + ni
        local = fp.cell_node[:n].astype(np.int64)

        def interior(node):
            r, c = np.divmod(node, G)
            return (r - 1) * (G - 2) + (c - 1)

        a, b = fp.face_nodes()
        ring = fp.ring()
        inner = ~ring[a] & ~ring[b]
        rows = np.concatenate((np.arange(n), n + interior(a[inner])))
        cols = np.concatenate((n + interior(local), n + interior(b[inner])))
        s = bind(self.nsys, rows, cols)
        row_rank = s.rank[s.row]
        pos = {}
        for kk in range(self.nsys):
            for q in range(int(s.column_start[kk]), int(s.column_start[kk + 1])):
                pos[(int(row_rank[q]), kk)] = q

        def at(u, v):
            ru, rv = int(s.rank[u]), int(s.rank[v])
            return pos[(max(ru, rv), min(ru, rv))]

        diag_pos = s.column_start[s.rank[np.arange(self.nsys)]]
        cap_pos = [at(i, n + interior(local[i])) for i in range(n)]
        face_pos = np.full(len(a), -1, dtype=np.int64)
        for q in np.nonzero(inner)[0]:
            face_pos[q] = at(n + interior(a[q]), n + interior(b[q]))
        edge_pos = np.concatenate((np.asarray(cap_pos, dtype=np.int64), face_pos))
        self.symbolic = s
        self.nnz = int(s.column_start[-1])
        arrays = [np.ascontiguousarray(x, dtype=np.int64) for x in
                  (s.column_start, row_rank, s.order, diag_pos, edge_pos)]
        self._keep = arrays
        self.structs = (Symbolic * fp.W)()
        for k in range(fp.W):
            st = self.structs[k]
            st.n, st.nnz = self.nsys, self.nnz
            st.column_start, st.row_rank, st.order, st.diag_pos, st.edge_pos = [_p(x, _i64p) for x in arrays]
        self.weight = np.zeros(fp.W * fp.S * fp.faces)
        self.scale = np.zeros(fp.W * self.nsys)
        self.factor = np.zeros(fp.W * self.nnz)
        self.kirchhoff = self._offset_pattern(fp, interior, a, b, inner)

    @staticmethod
    def _offset_pattern(fp, interior, a, b, inner):
        """The field solve's offset system (one offset per interior node, row-major): the face
        Laplacian's pattern plus each junction's pair of cell nodes; edge_pos lists each face's
        entry (-1 where it touches the ring), then each junction's (world-local order). Every
        world shares the grid, the placement and the lattice, so one pattern serves them all."""
        from oracle.symbolic import bind

        p, G, n = fp.cells, fp.side, fp.config.cells
        ni = (G - 2) * (G - 2)
        e0, e1 = int(p.edge_off[0]), int(p.edge_off[1])
        for k in range(1, fp.W):
            f0, f1 = int(p.edge_off[k]), int(p.edge_off[k + 1])
            c0 = int(p.cell_off[k])
            if not (np.array_equal(p.snd[f0:f1] - c0, p.snd[e0:e1]) and
                    np.array_equal(p.rcv[f0:f1] - c0, p.rcv[e0:e1])):
                raise ValueError("the worlds' junction graphs differ")
        local = fp.cell_node[:n].astype(np.int64)
        ja, jb = interior(local[p.snd[e0:e1]]), interior(local[p.rcv[e0:e1]])
        rows = np.concatenate((interior(a[inner]), ja))
        cols = np.concatenate((interior(b[inner]), jb))
        s = bind(ni, rows, cols)
        row_rank = s.rank[s.row]
        pos = {}
        for kk in range(ni):
            for q in range(int(s.column_start[kk]), int(s.column_start[kk + 1])):
                pos[(int(row_rank[q]), kk)] = q

        def at(u, v):
            ru, rv = int(s.rank[u]), int(s.rank[v])
            return pos[(max(ru, rv), min(ru, rv))]

        diag_pos = s.column_start[s.rank[np.arange(ni)]]
        face_pos = np.full(len(a), -1, dtype=np.int64)
        for q in np.nonzero(inner)[0]:
            face_pos[q] = at(interior(a[q]), interior(b[q]))
        junction_pos = np.array([at(u, v) for u, v in zip(ja, jb)], dtype=np.int64)
        arrays = [np.ascontiguousarray(x, dtype=np.int64) for x in
                  (s.column_start, row_rank, s.order, diag_pos, np.concatenate((face_pos, junction_pos)))]
        structs = (Symbolic * fp.W)()
        for k in range(fp.W):
            st = structs[k]
            st.n, st.nnz = ni, int(s.column_start[-1])
            st.column_start, st.row_rank, st.order, st.diag_pos, st.edge_pos = [_p(x, _i64p) for x in arrays]
        structs._keep = arrays
        return structs


def field_step(fp, y, dt, op, refresh, ctl=None, world=None, grid=None, start="solve"):
    """One oracle step of every spatial world: (state, valid [W]). refresh rebuilds op's frozen
    operator from y first. start="solve": the step-start potentials are the field solve at y (the
    reference's); "carried": the state's node potentials."""
    if start not in ("solve", "carried"):
        raise ValueError('start must be "solve" or "carried"')
    w = world if world is not None else world_struct(fp.cells)
    g = grid if grid is not None else field_grid_struct(fp)
    ctl = ctl or control()
    y = np.ascontiguousarray(y, dtype=np.float64)
    out = np.empty_like(y)
    valid = np.zeros(fp.W, dtype=np.int32)
    lib().ref_field_step(ctypes.byref(w), ctypes.byref(g), op.structs,
                         op.kirchhoff if start == "solve" else None, _p(y, _f64p), _p(out, _f64p),
                         ctypes.c_double(dt), ctypes.byref(ctl), int(bool(refresh)),
                         _p(op.weight, _f64p), _p(op.scale, _f64p), _p(op.factor, _f64p),
                         _p(valid, _i32p))
    return out, valid.astype(bool)


def field_run(fp, y, dt, steps, refresh_every, op=None, ctl=None, start="solve"):
    """steps oracle steps, the operator refreshed every refresh_every steps (from step 0)."""
    op = op or FieldOperator(fp)
    w, g = world_struct(fp.cells), field_grid_struct(fp)
    ok = np.ones(fp.W, dtype=bool)
    for t in range(steps):
        y, v = field_step(fp, y, dt, op, t % refresh_every == 0, ctl, w, g, start)
        ok &= v
    return y, ok


def field_start(fp, y, op, world, ctl=None):
    """The field solve of one spatial world at y: (cell potentials [n], node potentials
    [side^2], Newton iterations, converged)."""
    w, g = world_struct(fp.cells), field_grid_struct(fp)
    ctl = ctl or control()
    y = np.ascontiguousarray(y, dtype=np.float64)
    n = fp.config.cells
    phi = np.zeros(n + fp.side * fp.side)
    it = ctypes.c_int32()
    ok = lib().ref_field_start(ctypes.byref(w), ctypes.byref(g), ctypes.byref(op.kirchhoff[world]), int(world),
                               _p(y, _f64p), ctypes.byref(ctl), _p(phi, _f64p), ctypes.byref(it))
    return phi[:n], phi[n:], int(it.value), bool(ok)


def logmean(left, right):
    m, lg = ctypes.c_double(), ctypes.c_double()
    lib().ref_logmean(ctypes.c_double(left), ctypes.c_double(right), ctypes.byref(m), ctypes.byref(lg))
    return m.value, lg.value


FIELD_ENERGY_RATES = ("chemical", "capacitor", "dissipation", "reservoir_power", "electrode_power",
                      "energy_balance", "electrode_current", "capacitor_energy", "membrane_heat", "junction_heat",
                      "face_heat")


def field_energy_rates(fp, y, op, world, ctl=None, phi=None):
    """The field solve's energy rates [11] (FIELD_ENERGY_RATES order) of one spatial world at y, with
    the oracle's field solve for the potentials (or the given (cell, node) potentials); returns
    (rates, converged)."""
    w, g = world_struct(fp.cells), field_grid_struct(fp)
    ctl = ctl or control()
    y = np.ascontiguousarray(y, dtype=np.float64)
    converged = True
    if phi is None:
        pc, pn, _, converged = field_start(fp, y, op, world, ctl)
    else:
        pc, pn = phi
    both = np.ascontiguousarray(np.concatenate((pc, pn)), dtype=np.float64)
    out = np.zeros(len(FIELD_ENERGY_RATES))
    ok = lib().ref_field_energy_rates(ctypes.byref(w), ctypes.byref(g), int(world), _p(y, _f64p), ctypes.byref(ctl),
                                      _p(both, _f64p), _p(out, _f64p))
    return out, bool(converged and ok)
