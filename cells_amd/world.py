"""Synthetic tissue worlds on held baths, packed for the batched step.

A world is a square lattice of cells (side L = floor(sqrt(n)), the remaining cells appended to the
last row) with 4-neighbour gap junctions, every cell's one membrane patch facing one held bath. The
packed form is the data layout every kernel reads: worlds back to back, cells and edges in world
order, per-cell and per-edge parameters species-major ([S][ncell], [S][nedge]) so a wavefront reads
one species of 64 consecutive cells in one coalesced load.

The definitions these worlds stand for are the reference's resolved pump tissues
(configs/resolved_pump_tissue.json of the reference): membranes with constant-field permeabilities
(laws/junctions/selective.py:28-50), a Na/K-ATPase per membrane
(laws/pumps/nak_atp_mass_action_positive.py), optionally the Ca-ATPase
(laws/pumps/ca_atp_mass_action_positive.py), junctions with the voltage-relaxation gate
(laws/junctions/junction_voltage_relaxation.py) and optionally osmotic water
(operators/osmotic_water.py). Values the configurations leave open are fixed here and named in
`Config` (see the docstring of each field).
"""

from __future__ import annotations

import dataclasses
import math

import numpy as np

AVOGADRO = 6.02214076e23
F = AVOGADRO * 1.602176634e-19
R = AVOGADRO * 1.380649e-23
MOLAR_VOLUME_WATER = 18.07e-6  # m^3/mol (validation/tissue/pump_water_fixtures.py:30 of the reference)


@dataclasses.dataclass
class Config:
    """One tissue configuration. Defaults are the plain 8x8 Na/K/Cl world."""

    worlds: int = 1
    cells: tuple = (64, 64)  # uniform integer range [lo, hi] of cells per world
    species: tuple = ("Na", "K", "Cl")
    temperature: float = 310.0
    inside_mM: tuple = (12.0, 140.0, 10.0)
    bath_mM: tuple = (145.0, 5.0, 110.0)
    permeability: tuple = (1e-10, 5e-9, 1e-9)  # m/s, each scaled per cell by U(0.9, 1.1)
    # Ca permeability (4-ion worlds): the configurations give none; 1e-12 m/s is declared here.
    calcium_permeability: float = 1e-12
    cell_volume_m3: float = 2e-15
    capacitance_F_m2: float = 1e-2
    # Membrane area: a cube of the cell's volume (6 V^(2/3)). A sphere's area puts the 1e-9 S
    # junctions past the explicit-rate bound at dt = 1 ms; the cube admits them (dt x rate 0.84).
    junction_S: float = 1e-9  # V = 0 GHK slope conductance, lognormal sigma ln 1.1
    junction_sigma: float = math.log(1.1)
    # Junction relative permeabilities Na:K:Cl(:Ca), the reference fixture's 1:2:1 (Ca 0.1 declared)
    junction_weights: tuple = (1.0, 2.0, 1.0, 0.1)
    pump_capacity: float = 0.05  # fmol cycles/s, scaled per cell by U(0.9, 1.1)
    calcium_pump_capacity: float = 0.002
    atp: float = 1.5
    adp: float = 0.1
    pi: float = 0.1
    km_na: float = 12.0
    km_k: float = 0.2
    km_atp: float = 0.5
    dg0: float = -37000.0
    ca_km_ca: float = 0.001
    ca_km_atp: float = 0.5
    ca_dg0: float = -30500.0
    gate_threshold: float = 0.015
    gate_minimum: float = 0.1
    initial_gate: float = 1.0
    initial_vmem: float = -0.060  # V; sets each cell's initial charge C_i V0
    water: bool = False
    pf_m_s: float = 3e-5
    water_K: float = 60.0
    water_P0: float = 40.0
    water_vmin: float = 0.5
    water_vmax: float = 2.0
    bath_pressure: float = 0.0
    bath_V: float = 0.0
    seed: int = 0
    # Rows of this many cells (the last row holds the rest); 0: the square rule of `lattice`.
    row_width: int = 0


def config0(**kw):
    """Plain method: one 8x8 world, Na/K/Cl, water on, float64."""
    return Config(**{"water": True, **kw})


def config1(worlds=4096, **kw):
    """Batched ensemble: worlds of U{256..4096} cells, Na/K/Cl/Ca, no water."""
    base = dict(
        worlds=worlds,
        cells=(256, 4096),
        species=("Na", "K", "Cl", "Ca"),
        inside_mM=(12.0, 140.0, 10.0, 1e-4),
        bath_mM=(145.0, 5.0, 110.0, 2.0),
        permeability=(1e-10, 5e-9, 1e-9, None),
        seed=1,
    )
    base.update(kw)
    return Config(**base)


def config4(worlds=512, **kw):
    """Stationary analysis: worlds of 64x64 = 4096 cells, Na/K/Cl/Ca, junction conductances
    spread by a lognormal of sigma 0.3, no water."""
    base = dict(cells=(4096, 4096), junction_sigma=0.3, seed=4)
    base.update(kw)
    return config1(worlds=worlds, **base)


def row_widths(n, row_width=0):
    """Cells per lattice row: L = floor(sqrt n) rows of L with the remaining n - L^2 appended to
    the last row, or (row_width > 0) full rows of row_width and the rest in a last, shorter row."""
    if n < 1:
        raise ValueError("A lattice needs at least one cell")
    if row_width > 0:
        full, rest = divmod(n, row_width)
        return np.array([row_width] * full + ([rest] if rest else []), dtype=np.int64)
    side = math.isqrt(n)
    widths = np.full(side, side, dtype=np.int64)
    widths[-1] += n - side * side
    return widths


def lattice(n, row_width=0):
    """(senders, receivers) of the lattice of n cells in rows of `row_widths(n, row_width)`; each
    cell's right edge, then its down edge."""
    widths = row_widths(n, row_width)
    side = len(widths)
    start = np.concatenate(([0], np.cumsum(widths)))
    snd, rcv = [], []
    for r in range(side):
        w = int(widths[r])
        cols = np.arange(w)
        own = start[r] + cols
        right = np.stack((own[:-1], own[1:]), axis=1)
        if r + 1 < side:
            below = min(w, int(widths[r + 1]))
            down = np.stack((own[:below], start[r + 1] + cols[:below]), axis=1)
        else:
            down = np.zeros((0, 2), dtype=np.int64)
        # interleave: per cell its right edge then its down edge
        pairs = []
        ri, di = 0, 0
        for c in range(w):
            if c + 1 < w:
                pairs.append(right[ri]); ri += 1
            if c < len(down):
                pairs.append(down[di]); di += 1
        if pairs:
            p = np.asarray(pairs, dtype=np.int64)
            snd.append(p[:, 0]); rcv.append(p[:, 1])
    if not snd:
        return np.zeros(0, np.int64), np.zeros(0, np.int64)
    return np.concatenate(snd), np.concatenate(rcv)


@dataclasses.dataclass
class Packed:
    """The packed worlds: the arrays the C-ABI's cells_world struct points at."""

    config: Config
    S: int
    W: int
    calcium: bool
    water: bool
    cell_off: np.ndarray
    edge_off: np.ndarray
    valence: np.ndarray
    bath_c: np.ndarray
    vol0: np.ndarray
    cap: np.ndarray
    pam: np.ndarray  # [S, ncell]
    nak_cap: np.ndarray
    ca_cap: np.ndarray
    n0: np.ndarray  # [S, ncell]
    q0: np.ndarray
    lpa: np.ndarray
    imperm: np.ndarray
    snd: np.ndarray
    rcv: np.ndarray
    paj: np.ndarray  # [S, nedge]
    inc_ptr: np.ndarray
    inc_edge: np.ndarray
    inc_out: np.ndarray

    @property
    def ncell(self):
        return int(self.cell_off[-1])

    @property
    def nedge(self):
        return int(self.edge_off[-1])

    def sizes(self):
        return np.diff(self.cell_off)


VALENCE = {"Na": 1.0, "K": 1.0, "Cl": -1.0, "Ca": 2.0}


def build(cfg: Config) -> Packed:
    rng = np.random.default_rng(cfg.seed)
    S = len(cfg.species)
    lo, hi = cfg.cells
    sizes = rng.integers(lo, hi + 1, size=cfg.worlds) if hi > lo else np.full(cfg.worlds, lo)
    sizes = sizes.astype(np.int64)
    cell_off = np.concatenate(([0], np.cumsum(sizes)))
    ncell = int(cell_off[-1])
    lattices = {}
    snd_l, rcv_l, esz = [], [], []
    for w, n in enumerate(sizes.tolist()):
        if n not in lattices:
            lattices[n] = lattice(n, cfg.row_width)
        s, r = lattices[n]
        snd_l.append(s + cell_off[w])
        rcv_l.append(r + cell_off[w])
        esz.append(len(s))
    edge_off = np.concatenate(([0], np.cumsum(esz))).astype(np.int64)
    snd = np.concatenate(snd_l).astype(np.int64) if snd_l else np.zeros(0, np.int64)
    rcv = np.concatenate(rcv_l).astype(np.int64) if rcv_l else np.zeros(0, np.int64)
    nedge = len(snd)
    valence = np.array([VALENCE[s] for s in cfg.species])
    inside = np.asarray(cfg.inside_mM, dtype=float)
    vol0 = np.full(ncell, cfg.cell_volume_m3 * 1e15)  # pL
    area = 6.0 * cfg.cell_volume_m3 ** (2.0 / 3.0)
    cap = np.full(ncell, cfg.capacitance_F_m2 * area)
    perms = [cfg.calcium_permeability if p is None else p for p in cfg.permeability]
    pam = np.empty((S, ncell))
    for s in range(S):
        pam[s] = area * perms[s] * rng.uniform(0.9, 1.1, ncell)
    nak_cap = cfg.pump_capacity * rng.uniform(0.9, 1.1, ncell)
    calcium = S > 3
    ca_cap = cfg.calcium_pump_capacity * rng.uniform(0.9, 1.1, ncell) if calcium else np.zeros(ncell)
    n0 = inside[:, None] * vol0[None, :]
    q0 = cap * cfg.initial_vmem
    lp = cfg.pf_m_s * MOLAR_VOLUME_WATER / (R * cfg.temperature)  # m/(Pa s)
    lpa = np.full(ncell, lp * area * 1e15) if cfg.water else np.zeros(ncell)  # pL/(Pa s)
    bath = np.asarray(cfg.bath_mM, dtype=float)
    if cfg.water:
        # impermeant osmolytes that put each cell at the bath's water potential initially
        total = float(bath.sum()) + (cfg.water_P0 - cfg.bath_pressure) / (R * cfg.temperature)
        imperm = vol0 * total - n0.sum(axis=0)
    else:
        imperm = np.zeros(ncell)
    # junctions: V = 0 GHK slope G = F^2/RT sum_s z_s^2 (A P)_s c_s at the initial concentrations
    g = cfg.junction_S * rng.lognormal(0.0, cfg.junction_sigma, nedge)
    weights = np.asarray(cfg.junction_weights[:S], dtype=float)
    per_unit = F * F / (R * cfg.temperature) * float(np.sum(valence**2 * weights * inside))
    paj = (weights[:, None] * (g / per_unit)[None, :]).copy()
    # incidence: each cell's out-edges (it is the sender) ascending, then its in-edges ascending
    e_idx = np.arange(nedge, dtype=np.int64)
    owner = np.concatenate((snd, rcv))
    kind = np.concatenate((np.zeros(nedge, np.int64), np.ones(nedge, np.int64)))
    edges2 = np.concatenate((e_idx, e_idx))
    order = np.lexsort((edges2, kind, owner))
    inc_edge = edges2[order]
    counts = np.bincount(owner, minlength=ncell)
    inc_ptr = np.concatenate(([0], np.cumsum(counts))).astype(np.int64)
    inc_out = np.bincount(snd, minlength=ncell).astype(np.int32)
    return Packed(
        config=cfg, S=S, W=cfg.worlds, calcium=calcium, water=cfg.water,
        cell_off=cell_off.astype(np.int64), edge_off=edge_off, valence=valence, bath_c=bath,
        vol0=vol0, cap=cap, pam=np.ascontiguousarray(pam), nak_cap=nak_cap, ca_cap=ca_cap,
        n0=np.ascontiguousarray(n0), q0=q0, lpa=lpa, imperm=imperm, snd=snd, rcv=rcv,
        paj=np.ascontiguousarray(paj), inc_ptr=inc_ptr, inc_edge=inc_edge.astype(np.int64),
        inc_out=inc_out,
    )


@dataclasses.dataclass
class StateLayout:
    """Field-major SoA blocks: cell fields x ncell, then edge fields x nedge, then world fields x W.

    Cell fields: amount increments [S], membrane transfers [S], membrane heat, Na/K cycles, pump
    heat, metabolic work, (Ca cycles, Ca heat, Ca work), (volume increment, water transfer, water
    heat). Edge fields: junction transfers [S], junction heat, gate. World fields: bath supplies
    [S], electrode charge, reservoir work, electrode work, (bath water supply, bath water work)."""

    S: int
    calcium: bool
    water: bool
    ncell: int
    nedge: int
    W: int

    def __post_init__(self):
        S = self.S
        f = 0
        self.dn = f; f += S
        self.mtr = f; f += S
        self.mdiss = f; f += 1
        self.pcyc = f; f += 1
        self.pdiss = f; f += 1
        self.pwork = f; f += 1
        self.ccyc = self.cdiss = self.cwork = -1
        if self.calcium:
            self.ccyc, self.cdiss, self.cwork = f, f + 1, f + 2; f += 3
        self.dvol = self.wtr = self.wdiss = -1
        if self.water:
            self.dvol, self.wtr, self.wdiss = f, f + 1, f + 2; f += 3
        self.cell_fields = f
        self.jtr, self.jdiss, self.gate = 0, S, S + 1
        self.edge_fields = S + 2
        f = 0
        self.sup = f; f += S
        self.echg, self.rwork, self.ework = f, f + 1, f + 2; f += 3
        self.wsup = self.wwork = -1
        if self.water:
            self.wsup, self.wwork = f, f + 1; f += 2
        self.world_fields = f
        self.edge_base = self.cell_fields * self.ncell
        self.world_base = self.edge_base + self.edge_fields * self.nedge
        self.total = self.world_base + self.world_fields * self.W

    @classmethod
    def of(cls, p: Packed):
        return cls(p.S, p.calcium, p.water, p.ncell, p.nedge, p.W)

    def cell(self, y, f):
        return y[f * self.ncell:(f + 1) * self.ncell]

    def edge(self, y, f):
        b = self.edge_base + f * self.nedge
        return y[b:b + self.nedge]

    def world(self, y, f):
        b = self.world_base + f * self.W
        return y[b:b + self.W]


def initial_state(p: Packed) -> np.ndarray:
    L = StateLayout.of(p)
    y = np.zeros(L.total)
    L.edge(y, L.gate)[:] = p.config.initial_gate
    return y


def head(p: Packed, k: int) -> Packed:
    """The first k worlds of a packed ensemble, with the same parameters (a sub-ensemble)."""
    if not 1 <= k <= p.W:
        raise ValueError("need 1 <= k <= W")
    nc, ne = int(p.cell_off[k]), int(p.edge_off[k])
    return dataclasses.replace(
        p, W=k, cell_off=p.cell_off[:k + 1].copy(), edge_off=p.edge_off[:k + 1].copy(),
        vol0=p.vol0[:nc], cap=p.cap[:nc], pam=np.ascontiguousarray(p.pam[:, :nc]),
        nak_cap=p.nak_cap[:nc], ca_cap=p.ca_cap[:nc], n0=np.ascontiguousarray(p.n0[:, :nc]),
        q0=p.q0[:nc], lpa=p.lpa[:nc], imperm=p.imperm[:nc], snd=p.snd[:ne], rcv=p.rcv[:ne],
        paj=np.ascontiguousarray(p.paj[:, :ne]), inc_ptr=p.inc_ptr[:nc + 1].copy(),
        inc_edge=p.inc_edge[:2 * ne].copy(), inc_out=p.inc_out[:nc].copy(),
    )


def worlds(p: Packed, a: int, b: int) -> Packed:
    """Worlds [a, b) of a packed ensemble as an ensemble of their own (the same parameters)."""
    if not 0 <= a < b <= p.W:
        raise ValueError("need 0 <= a < b <= W")
    c0, c1 = int(p.cell_off[a]), int(p.cell_off[b])
    e0, e1 = int(p.edge_off[a]), int(p.edge_off[b])
    return dataclasses.replace(
        p, W=b - a, cell_off=p.cell_off[a:b + 1] - c0, edge_off=p.edge_off[a:b + 1] - e0,
        vol0=p.vol0[c0:c1], cap=p.cap[c0:c1], pam=np.ascontiguousarray(p.pam[:, c0:c1]),
        nak_cap=p.nak_cap[c0:c1], ca_cap=p.ca_cap[c0:c1], n0=np.ascontiguousarray(p.n0[:, c0:c1]),
        q0=p.q0[c0:c1], lpa=p.lpa[c0:c1], imperm=p.imperm[c0:c1], snd=p.snd[e0:e1] - c0,
        rcv=p.rcv[e0:e1] - c0, paj=np.ascontiguousarray(p.paj[:, e0:e1]),
        inc_ptr=p.inc_ptr[c0:c1 + 1] - p.inc_ptr[c0], inc_edge=p.inc_edge[2 * e0:2 * e1] - e0,
        inc_out=p.inc_out[c0:c1].copy(),
    )


def take_state(p: Packed, y: np.ndarray, k: int) -> np.ndarray:
    """The state of the first k worlds (head(p, k)'s layout) out of a state of p."""
    L = StateLayout.of(p)
    sub = head(p, k)
    Ls = StateLayout.of(sub)
    out = np.zeros(Ls.total)
    for f in range(L.cell_fields):
        Ls.cell(out, f)[:] = L.cell(y, f)[:sub.ncell]
    for f in range(L.edge_fields):
        Ls.edge(out, f)[:] = L.edge(y, f)[:sub.nedge]
    for f in range(L.world_fields):
        Ls.world(out, f)[:] = L.world(y, f)[:k]
    return out
