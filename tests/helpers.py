"""Shared helpers: physical readouts of a state and error measures."""

import numpy as np

from cells_amd.world import F, StateLayout


def vmem(p, y):
    L = StateLayout.of(p)
    q = p.q0 + 1e-15 * F * sum(p.valence[s] * L.cell(y, L.dn + s) for s in range(p.S))
    return p.config.bath_V + q / p.cap


def volume(p, y):
    L = StateLayout.of(p)
    return p.vol0 + (L.cell(y, L.dvol) if p.water else 0.0)


def conc(p, y):
    L = StateLayout.of(p)
    v = volume(p, y)
    return np.stack([(p.n0[s] + L.cell(y, L.dn + s)) / v for s in range(p.S)])


def relerr(a, b, floor=0.0):
    a = np.asarray(a, dtype=float)
    b = np.asarray(b, dtype=float)
    scale = np.maximum(np.abs(b), floor)
    with np.errstate(divide="ignore", invalid="ignore"):
        e = np.where(scale > 0, np.abs(a - b) / scale, np.abs(a - b))
    return float(np.max(e)) if e.size else 0.0


def material_residual(p, y):
    """Per species and cell: dN + membrane transfer - pump stoichiometry x cycles - junction net
    inflow (zero for a closed ledger), relative to the summed magnitudes."""
    L = StateLayout.of(p)
    worst = 0.0
    for s in range(p.S):
        dn = L.cell(y, L.dn + s)
        mtr = L.cell(y, L.mtr + s)
        jt = L.edge(y, L.jtr + s)
        net = np.zeros(p.ncell)
        np.add.at(net, p.snd, -jt)
        np.add.at(net, p.rcv, jt)
        pump = np.zeros(p.ncell)
        cyc = L.cell(y, L.pcyc)
        if s == 0:
            pump = 3.0 * cyc
        elif s == 1:
            pump = -2.0 * cyc
        if s == 3 and p.calcium:
            pump = pump + L.cell(y, L.ccyc)
        res = dn + mtr + pump - net
        mag = np.abs(dn) + np.abs(mtr) + np.abs(pump) + np.abs(net) + 1e-300
        worst = max(worst, float(np.max(np.abs(res) / mag.max())))
        # the bath: sum of cell amounts equals the bath's supply per world
        for k in range(p.W):
            c0, c1 = p.cell_off[k], p.cell_off[k + 1]
            tot = dn[c0:c1].sum()
            sup = L.world(y, L.sup + s)[k]
            m = np.abs(dn[c0:c1]).sum() + abs(sup) + 1e-300
            worst = max(worst, abs(tot - sup) / m)
    return worst
