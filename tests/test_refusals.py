"""Step refusals and junction multiplicity on the CPU: the oracle's per-stage explicit-rate
refusal (models/surface_pump_v3.py:3640-3642 with membrane_rate :4588), the oracle's sum of
parallel junctions into one stage-matrix entry (numerics/surface_field.py:413-418), and the
product's refusal of a repeated pair of cells (checked before any device is touched).
"""

import ctypes
import dataclasses

import numpy as np
import pytest

import oracle
from cells_amd import lib as hip
from cells_amd.world import F, R, StateLayout, build, config0, config1, initial_state


def with_edges(p, snd, rcv, paj):
    """p with its junctions replaced (one world; incidence rebuilt as the packer builds it)."""
    assert p.W == 1
    snd = np.asarray(snd, np.int64)
    rcv = np.asarray(rcv, np.int64)
    ne = len(snd)
    e_idx = np.arange(ne, dtype=np.int64)
    owner = np.concatenate((snd, rcv))
    kind = np.concatenate((np.zeros(ne, np.int64), np.ones(ne, np.int64)))
    edges2 = np.concatenate((e_idx, e_idx))
    order = np.lexsort((edges2, kind, owner))
    counts = np.bincount(owner, minlength=p.ncell)
    return dataclasses.replace(
        p, snd=snd, rcv=rcv, paj=np.ascontiguousarray(paj), edge_off=np.array([0, ne], np.int64),
        inc_edge=edges2[order].astype(np.int64),
        inc_ptr=np.concatenate(([0], np.cumsum(counts))).astype(np.int64),
        inc_out=np.bincount(snd, minlength=p.ncell).astype(np.int32))


def doubled(p, edge=5, reverse=False):
    """p with junction `edge` split into two junctions of half its permeabilities between the same
    pair of cells (the second appended, optionally pointing the other way)."""
    a, b = int(p.snd[edge]), int(p.rcv[edge])
    paj = p.paj.copy()
    paj[:, edge] *= 0.5
    snd = np.append(p.snd, b if reverse else a)
    rcv = np.append(p.rcv, a if reverse else b)
    return with_edges(p, snd, rcv, np.concatenate((paj, paj[:, edge:edge + 1]), axis=1))


def cell_fields(p, y):
    L = StateLayout.of(p)
    return np.concatenate([L.cell(y, f) for f in range(L.cell_fields)])


@pytest.mark.parametrize("reverse", [False, True])
def test_parallel_junctions_sum_into_one_stage_entry(reverse):
    """Two junctions of half the permeabilities between one pair carry what one junction carries:
    the GHK flux is linear in the permeability and both gates start and stay equal, so the ROS2
    trajectories agree to rounding. An assembly that stored one junction's conductance instead of
    their sum would halve that stage-matrix entry and miss by far more."""
    p = build(config1(worlds=1, cells=(36, 36), seed=9))
    q = doubled(p, reverse=reverse)
    ya, yb = initial_state(p), initial_state(q)
    _, sa = oracle.world_symbolics(p)
    _, sb = oracle.world_symbolics(q)
    wa, wb = oracle.world_struct(p), oracle.world_struct(q)
    for _ in range(30):
        ya, oka = oracle.step_ros2(p, ya, 5e-3, sa, world=wa)
        yb, okb = oracle.step_ros2(q, yb, 5e-3, sb, world=wb)
        assert oka.all() and okb.all()
    a, b = cell_fields(p, ya), cell_fields(q, yb)
    scale = np.abs(a).max()
    assert np.max(np.abs(a - b)) <= 1e-11 * scale
    La, Lb = StateLayout.of(p), StateLayout.of(q)
    # the two halves' transfers add up to the single junction's
    for s in range(p.S):
        ja, jb = La.edge(ya, La.jtr + s), Lb.edge(yb, Lb.jtr + s)
        sign = -1.0 if reverse else 1.0
        assert abs(jb[5] + sign * jb[-1] - ja[5]) <= 1e-11 * np.abs(ja).max()


def test_the_product_refuses_a_repeated_pair_of_cells():
    q = doubled(build(config1(worlds=1, cells=(36, 36), seed=9)))
    L = hip.load()
    h = ctypes.c_void_p()
    w = hip.world_struct(q)
    rc = L.cells_ensemble_create(ctypes.byref(w), hip.METHOD_ROS2, 0, ctypes.byref(h))
    assert rc == hip.CELLS_EINVAL
    assert b"same pair of cells" in L.cells_last_error()


def membrane_rate(p, y, membrane=1.0):
    """The explicit stage's rate per cell from the definitions (F^2/RT sum z^2 A P max(c) / C)."""
    L = StateLayout.of(p)
    vol = p.vol0 + (L.cell(y, L.dvol) if p.water else 0.0)
    t = np.zeros(p.ncell)
    for s in range(p.S):
        c = (p.n0[s] + L.cell(y, L.dn + s)) / vol
        t += p.valence[s] ** 2 * p.pam[s] * np.maximum(c, p.bath_c[s])
    return F * F / (R * p.config.temperature) * t * membrane / p.cap


def test_oracle_euler_refuses_a_stage_above_the_membrane_rate_bound():
    p = build(config0(water=False))
    y = initial_state(p)
    rate = membrane_rate(p, y).max()
    # the oracle's own rate is the definition's, cell by cell
    w = oracle.world_struct(p)
    own = [oracle.lib().ref_membrane_rate(ctypes.byref(w), y.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                          i, ctypes.byref(oracle.control())) for i in range(p.ncell)]
    assert np.allclose(own, membrane_rate(p, y), rtol=1e-14, atol=0)
    _, ok = oracle.step_euler(p, y, 0.98 / rate)
    assert ok.all()
    out, ok = oracle.step_euler(p, y, 1.02 / rate)
    assert not ok.any() and out.tobytes() == y.tobytes()
    # half the membrane availability halves the rate: the same step is admitted
    _, ok = oracle.step_euler(p, y, 1.02 / rate, ctl=oracle.control(membrane=0.5))
    assert ok.all()
    # a cell loaded with K relaxes faster: refused at a step its declared state admits
    L = StateLayout.of(p)
    yk = y.copy()
    L.cell(yk, L.dn + 1)[7] += p.n0[1][7]
    assert membrane_rate(p, yk)[7] > 1.5 * rate
    _, ok = oracle.step_euler(p, yk, 0.98 / rate)
    assert not ok.any()
