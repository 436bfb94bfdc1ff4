"""Spatial worlds: a lattice of cells embedded in an extracellular finite-volume grid.

Every world is the same square grid of side x side extracellular nodes (row-major) whose outer ring
is held at the bath (the reference's reservoirs as clamps: held concentrations, potential 0), with
a face between each pair of 4-neighbours (5-point finite volumes). A lattice of cells sits inside
it, each cell's one membrane facing the grid node at its site instead of a bath. The step is the
reference's IMEX Euler with implicit face electromigration (models/surface_pump.py:292-362,
numerics/surface_imex.py): membranes, junctions, pumps, gates and face diffusion explicit, the
face drift implicit through the node-space operator (C + dt L) phi = q.

Faces, world-local: horizontal faces r (side - 1) + c between (r, c) and (r, c + 1), then vertical
faces side (side - 1) + r side + c between (r, c) and (r + 1, c); the first node is the sender.
Face factor D_s A / L with A = h x depth x ecs_fraction and L = h (numerics/surface_imex.py:123-132
takes it per species and face), node volume h^2 x depth x ecs_fraction.

State (field-major, FieldLayout): cell fields (amount increments [S], membrane transfers [S],
membrane heat, Na/K cycles, pump heat, metabolic work) over every world's cells, edge fields
(junction transfers [S], junction heat, gate), node fields (amount increments [S], potential) over
W x side^2, face fields (transfers [S], heat) over W x F, world fields (ring supplies [S], reservoir
work). The node potential is the last accepted step's operator potential: the next step's start
potential at the nodes (a cell's is its node's plus its charge over its capacitance).
"""

from __future__ import annotations

import dataclasses

import numpy as np

from cells_amd.world import Config, Packed, build, row_widths


@dataclasses.dataclass
class FieldConfig:
    """One spatial configuration. Defaults: 64 worlds of 2048 cells in rows of 46 (a 45 x 46
    lattice) in a 256 x 256 grid, Na/K/Cl with D = 1.33e-9, 1.96e-9, 2.03e-9 m^2/s."""

    worlds: int = 64
    cells: int = 2048
    row_width: int = 46
    side: int = 256
    spacing: int = 5  # grid nodes between neighbouring cells' membrane nodes
    # Grid geometry the configuration leaves open, declared here: 5 um nodes (so dt D / h^2 =
    # 0.08 at dt = 1 ms: explicit face diffusion well inside its bound), a slab of the cells'
    # height (the cube root of 2 pL) and an extracellular volume fraction of 0.2.
    node_spacing_m: float = 5e-6
    depth_m: float = 2e-15 ** (1.0 / 3.0)
    ecs_fraction: float = 0.2
    diffusivity: tuple = (1.33e-9, 1.96e-9, 2.03e-9)
    seed: int = 2
    cell_kw: dict = dataclasses.field(default_factory=dict)  # other Config fields of the cells

    def cells_config(self) -> Config:
        return Config(worlds=self.worlds, cells=(self.cells, self.cells), row_width=self.row_width,
                      seed=self.seed, **self.cell_kw)


def small_field(worlds=2, seed=5, **kw):
    """A spatial world the oracle steps in seconds: 6 cells (2 rows of 3) in a 14 x 14 grid."""
    base = dict(worlds=worlds, cells=6, row_width=3, side=14, spacing=3, seed=seed)
    base.update(kw)
    return FieldConfig(**base)


@dataclasses.dataclass
class FieldPacked:
    """The packed spatial worlds: the cells' Packed plus the grid (cells_field_grid's arrays)."""

    config: FieldConfig
    cells: Packed
    side: int
    cell_node: np.ndarray   # [ncell] int32, world-local grid node of each cell's membrane
    node_volume: float      # pL
    face_factor: np.ndarray  # [S] m^3/s
    node_n0: np.ndarray     # [S, side^2] reference amounts, fmol
    node_q0: np.ndarray     # [W * side^2] fixed node charges, C

    @property
    def W(self):
        return self.cells.W

    @property
    def S(self):
        return self.cells.S

    @property
    def nodes(self):
        return self.side * self.side

    @property
    def faces(self):
        return 2 * self.side * (self.side - 1)

    def ring(self):
        """[side^2] bool: the held outer ring."""
        r, c = np.divmod(np.arange(self.nodes), self.side)
        return (r == 0) | (c == 0) | (r == self.side - 1) | (c == self.side - 1)

    def face_nodes(self):
        """(sender, receiver) world-local nodes of every face."""
        G = self.side
        idx = np.arange(G * G).reshape(G, G)
        a = np.concatenate((idx[:, :-1].ravel(), idx[:-1, :].ravel()))
        b = np.concatenate((idx[:, 1:].ravel(), idx[1:, :].ravel()))
        return a, b


def placement(n, row_width, side, spacing):
    """World-local grid node of each of the n cells: lattice row r, column c at grid
    (r0 + spacing r, c0 + spacing c), the lattice centred in the grid's interior."""
    widths = row_widths(n, row_width)
    rows, cols = len(widths), int(widths.max())
    r0 = (side - 1 - spacing * (rows - 1)) // 2
    c0 = (side - 1 - spacing * (cols - 1)) // 2
    node = []
    for r, w in enumerate(widths.tolist()):
        for c in range(w):
            node.append((r0 + spacing * r) * side + c0 + spacing * c)
    node = np.asarray(node, dtype=np.int64)
    rr, cc = np.divmod(node, side)
    if rr.min() < 1 or cc.min() < 1 or rr.max() > side - 2 or cc.max() > side - 2:
        raise ValueError("The cell lattice does not fit inside the grid's held ring")
    if len(np.unique(node)) != n:
        raise ValueError("Two cells share a grid node")
    return node


def build_field(cfg: FieldConfig) -> FieldPacked:
    p = build(cfg.cells_config())
    if p.water or p.calcium:
        raise ValueError("Spatial worlds carry no water and no Ca-ATPase")
    local = placement(cfg.cells, cfg.row_width, cfg.side, cfg.spacing)
    cell_node = np.tile(local, p.W).astype(np.int32)
    h, depth, frac = cfg.node_spacing_m, cfg.depth_m, cfg.ecs_fraction
    node_volume = h * h * depth * frac * 1e15  # pL
    S = p.S
    if len(cfg.diffusivity) != S:
        raise ValueError("One diffusivity per species")
    face_factor = np.asarray(cfg.diffusivity, dtype=float) * (h * depth * frac) / h
    nodes = cfg.side * cfg.side
    node_n0 = np.ascontiguousarray(np.asarray(p.bath_c, dtype=float)[:, None] * np.full(nodes, node_volume))
    # each cell's node holds the opposite of the cell's initial charge (the pair starts neutral)
    node_q0 = np.zeros(p.W * nodes)
    node_q0[np.arange(p.W).repeat(cfg.cells) * nodes + cell_node] = -p.q0
    return FieldPacked(config=cfg, cells=p, side=cfg.side, cell_node=cell_node,
                       node_volume=node_volume, face_factor=face_factor, node_n0=node_n0,
                       node_q0=node_q0)


@dataclasses.dataclass
class FieldLayout:
    """Field-major blocks of a spatial state (the oracle's ref_field_layout, the library's)."""

    S: int
    ncell: int
    nedge: int
    W: int
    side: int

    def __post_init__(self):
        S = self.S
        self.dn, self.mtr = 0, S
        self.mdiss, self.pcyc, self.pdiss, self.pwork = 2 * S, 2 * S + 1, 2 * S + 2, 2 * S + 3
        self.cell_fields = 2 * S + 4
        self.jtr, self.jdiss, self.gate = 0, S, S + 1
        self.edge_fields = S + 2
        self.dno, self.phi = 0, S
        self.node_fields = S + 1
        self.ftr, self.fheat = 0, S
        self.face_fields = S + 1
        self.sup, self.rwork = 0, S
        self.world_fields = S + 1
        self.nodes = self.side * self.side
        self.faces = 2 * self.side * (self.side - 1)
        self.edge_base = self.cell_fields * self.ncell
        self.node_base = self.edge_base + self.edge_fields * self.nedge
        self.face_base = self.node_base + self.node_fields * self.nodes * self.W
        self.world_base = self.face_base + self.face_fields * self.faces * self.W
        self.total = self.world_base + self.world_fields * self.W

    @classmethod
    def of(cls, fp: FieldPacked):
        return cls(fp.S, fp.cells.ncell, fp.cells.nedge, fp.W, fp.side)

    def cell(self, y, f):
        return y[f * self.ncell:(f + 1) * self.ncell]

    def edge(self, y, f):
        b = self.edge_base + f * self.nedge
        return y[b:b + self.nedge]

    def node(self, y, f):
        """[W, side^2] view of node field f."""
        b = self.node_base + f * self.W * self.nodes
        return y[b:b + self.W * self.nodes].reshape(self.W, self.nodes)

    def face(self, y, f):
        """[W, F] view of face field f."""
        b = self.face_base + f * self.W * self.faces
        return y[b:b + self.W * self.faces].reshape(self.W, self.faces)

    def world(self, y, f):
        b = self.world_base + f * self.W
        return y[b:b + self.W]


def initial_field_state(fp: FieldPacked) -> np.ndarray:
    L = FieldLayout.of(fp)
    y = np.zeros(L.total)
    L.edge(y, L.gate)[:] = fp.cells.config.initial_gate
    return y


def amounts(fp: FieldPacked, y):
    """(cell amounts [S, ncell], node amounts [S, W, side^2]) of a state, fmol."""
    L = FieldLayout.of(fp)
    cells = np.stack([fp.cells.n0[s] + L.cell(y, L.dn + s) for s in range(fp.S)])
    nodes = np.stack([fp.node_n0[s][None, :] + L.node(y, L.dno + s) for s in range(fp.S)])
    return cells, nodes


def charges(fp: FieldPacked, y):
    """(cell charges [ncell], node charges [W, side^2]) of a state, C (fixed charges included)."""
    from cells_amd.world import F

    L = FieldLayout.of(fp)
    z = fp.cells.valence
    cq = fp.cells.q0 + F * 1e-15 * sum(z[s] * L.cell(y, L.dn + s) for s in range(fp.S))
    nq = fp.node_q0.reshape(fp.W, -1) + F * 1e-15 * sum(z[s] * L.node(y, L.dno + s) for s in range(fp.S))
    return cq, nq


class FieldEnsemble:
    """Spatial worlds bound on one GPU (cells_field_create): the cells, the grid, the operator's
    nested-dissection plan and the stepping workspace. The operator is rebuilt at the first step
    and then every refresh_every steps. start="solve" (the reference's): each step starts from the
    field solve at its state (numerics/surface_field.py:614-855), by at most max_iterations chord
    steps; start="carried": from the state's node potentials (the last operator solve's)."""

    def __init__(self, fp: FieldPacked, refresh_every: int = 50, device: int = 0, start: str = "solve",
                 max_iterations: int = 8):
        import ctypes

        from cells_amd import lib as _lib
        from cells_amd.ensemble import _torch

        torch = _torch()
        self.torch = torch
        self.packed = fp
        self.layout = FieldLayout.of(fp)
        self.device = torch.device("cuda", device)
        self._world = _lib.world_struct(fp.cells)
        self._grid = _lib.field_grid_struct(fp)
        handle = ctypes.c_void_p()
        torch.cuda.init()
        _lib.check(_lib.load().cells_field_create(ctypes.byref(self._world), ctypes.byref(self._grid),
                                                  int(refresh_every), device, ctypes.byref(handle)))
        self._h = handle
        self.state_size = int(_lib.load().cells_field_state_size(handle))
        if self.state_size != self.layout.total:
            raise RuntimeError("spatial state layout disagrees with the library's")
        if start not in ("solve", "carried"):
            raise ValueError('start must be "solve" or "carried"')
        self.start = start
        _lib.check(_lib.load().cells_field_set_start(handle, int(start == "solve"), int(max_iterations)))

    def close(self):
        from cells_amd import lib as _lib

        if getattr(self, "_h", None):
            _lib.load().cells_field_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _stream(self):
        import ctypes

        return ctypes.c_void_p(self.torch.cuda.current_stream(self.device).cuda_stream)

    def to_device(self, y):
        t = self.torch
        if isinstance(y, np.ndarray):
            y = t.from_numpy(np.ascontiguousarray(y, dtype=np.float64))
        return y.to(self.device, dtype=t.float64).contiguous()

    def step(self, y, dt, nsteps=1, ctl=None, active=None, out=None, ok=None):
        """nsteps steps of every world; (new state, ok per world) on the device."""
        import ctypes

        from cells_amd import lib as _lib
        from cells_amd.ensemble import control

        t = self.torch
        y = self.to_device(y)
        if y.numel() != self.state_size:
            raise ValueError("state size disagrees with the worlds")
        out = t.empty_like(y) if out is None else out
        ok = t.empty(self.packed.W, dtype=t.int32, device=self.device) if ok is None else ok
        act = None
        if active is not None:
            act = t.as_tensor(np.asarray(active, dtype=np.int32)).to(self.device)
        ctl = ctl or control()
        _lib.check(_lib.load().cells_field_step(
            self._h, self._stream(), ctypes.c_void_p(y.data_ptr()), ctypes.c_void_p(out.data_ptr()),
            float(dt), int(nsteps), ctypes.byref(ctl),
            None if act is None else ctypes.c_void_p(act.data_ptr()), ctypes.c_void_p(ok.data_ptr())))
        return out, ok

    def refresh(self):
        from cells_amd import lib as _lib

        _lib.check(_lib.load().cells_field_refresh(self._h))

    def potentials(self, y, dt, ctl=None):
        """The field solve at y (cells_field_potentials): (node potentials [W, side^2], chord steps
        [W], ok [W]) on the device; a cell's potential is its node's plus its charge over its
        capacitance."""
        import ctypes

        from cells_amd import lib as _lib
        from cells_amd.ensemble import control

        t = self.torch
        y = self.to_device(y)
        if y.numel() != self.state_size:
            raise ValueError("state size disagrees with the worlds")
        W, nodes = self.packed.W, self.packed.side * self.packed.side
        phi = t.empty((W, nodes), dtype=t.float64, device=self.device)
        it = t.empty(W, dtype=t.int32, device=self.device)
        ok = t.empty(W, dtype=t.int32, device=self.device)
        ctl = ctl or control()
        _lib.check(_lib.load().cells_field_potentials(
            self._h, self._stream(), ctypes.c_void_p(y.data_ptr()), float(dt), ctypes.byref(ctl),
            ctypes.c_void_p(phi.data_ptr()), ctypes.c_void_p(it.data_ptr()), ctypes.c_void_p(ok.data_ptr())))
        return phi, it, ok

    def energy_rates(self, y, dt, ctl=None):
        """The field solve's energy terms at y (cells_field_energy_rates): ([W, 11] on the device,
        columns as lib.FIELD_ENERGY_RATES (numerics/surface_field.py:792-835), ok [W]: the field
        solve met its target)."""
        import ctypes

        from cells_amd import lib as _lib
        from cells_amd.ensemble import control

        t = self.torch
        y = self.to_device(y)
        if y.numel() != self.state_size:
            raise ValueError("state size disagrees with the worlds")
        out = t.empty((self.packed.W, len(_lib.FIELD_ENERGY_RATES)), dtype=t.float64, device=self.device)
        ok = t.empty(self.packed.W, dtype=t.int32, device=self.device)
        ctl = ctl or control()
        _lib.check(_lib.load().cells_field_energy_rates(
            self._h, self._stream(), ctypes.c_void_p(y.data_ptr()), float(dt), ctypes.byref(ctl),
            ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(ok.data_ptr())))
        return out, ok

    def plan_stats(self):
        """dict(nnz, flops, fronts, max_front, levels) of the operator's factor plan (per world)."""
        import ctypes

        from cells_amd import lib as _lib

        nnz, flops = ctypes.c_int64(), ctypes.c_double()
        fronts, mf, lev = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        _lib.check(_lib.load().cells_field_plan_stats(self._h, ctypes.byref(nnz), ctypes.byref(flops),
                                                      ctypes.byref(fronts), ctypes.byref(mf),
                                                      ctypes.byref(lev)))
        return dict(nnz=nnz.value, flops=flops.value, fronts=fronts.value, max_front=mf.value,
                    levels=lev.value)

    def kernel_times(self, y, dt, reps=3, ctl=None):
        """[(kernel name, mean ms per launch)] over reps steps from y (a refresh at the first)."""
        import ctypes

        from cells_amd import lib as _lib
        from cells_amd.ensemble import control

        y = self.to_device(y)
        cap = 32
        ms = np.zeros(cap)
        names = ctypes.create_string_buffer(32 * cap)
        count = ctypes.c_int32()
        ctl = ctl or control()
        _lib.check(_lib.load().cells_field_kernel_times(
            self._h, self._stream(), ctypes.c_void_p(y.data_ptr()), float(dt), ctypes.byref(ctl),
            int(reps), cap, ms.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), names,
            ctypes.byref(count)))
        raw = names.raw
        return [(raw[32 * k:32 * k + 32].split(b"\0")[0].decode(), float(ms[k]))
                for k in range(count.value)]
