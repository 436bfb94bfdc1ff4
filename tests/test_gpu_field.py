"""GPU checks of the spatial worlds (cells_field_*) through the C-ABI.

- Parity: the library's steps against the oracle's (tests restate the reference's IMEX Euler with
  implicit face electromigration on the cells-plus-free-nodes operator, minimum-degree ordered and
  factored in the reference's order, each step from the field solve at its state) from the same
  seeded worlds, with the operator rebuilt every step and every 5 steps: every field within 1e-9
  of its scale; once more from the carried potentials. Grids up to 130 x 130, whose largest fronts
  (above 100 rows) run in HBM scratch and through the workgroup solve.
- The field solve alone (cells_field_potentials) against the oracle's Newton at states some steps
  into a run: node and cell potentials within 1e-9 of their scale.
- The field solve's energy terms (cells_field_energy_rates) against the oracle's at the device's
  potentials, and their balance, which vanishes at the field solve's potentials.
- At the full 256 x 256 grid with 2048 cells: material closes per species against the ring
  supplies (1e-12 of the amounts), every interior node with its cell is neutral after each step
  (1e-12 of the cells' charge), and the refusal path keeps a world's state.
"""

import sys

import numpy as np
import pytest

import oracle
from cells_amd.field import (FieldConfig, FieldLayout, amounts, build_field, charges,
                             initial_field_state, small_field)

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


def compare(fp, got, want, tol=1e-9):
    """Largest field-scaled difference between two spatial states, per block. A world's ring
    supply is the net of the ring faces' transfers, which cancel: it is scaled by their gross
    transfer (the sum of |transfer| over the faces that touch the ring), the reservoir work by
    1e-3 sum_s |mu_s| times that."""
    from cells_amd.world import R

    L = FieldLayout.of(fp)
    worst = {}
    blocks = [("cell", L.cell, L.cell_fields), ("edge", L.edge, L.edge_fields),
              ("node", L.node, L.node_fields), ("face", L.face, L.face_fields)]
    for name, view, count in blocks:
        for f in range(count):
            a, b = view(got, f), view(want, f)
            scale = max(float(np.max(np.abs(b))), 1e-300)
            worst[(name, f)] = float(np.max(np.abs(a - b))) / scale
    ring = fp.ring()
    fa, fb = fp.face_nodes()
    touch = ring[fa] | ring[fb]
    gross = np.stack([np.abs(L.face(want, L.ftr + s)[:, touch]).sum(axis=1) for s in range(fp.S)])
    mu = R * fp.cells.config.temperature * np.log(fp.cells.bath_c)
    for s in range(fp.S):
        d = np.abs(L.world(got, L.sup + s) - L.world(want, L.sup + s))
        worst[("world", s)] = float(np.max(d / np.maximum(gross[s], 1e-300)))
    d = np.abs(L.world(got, L.rwork) - L.world(want, L.rwork))
    worst[("world", fp.S)] = float(np.max(d / np.maximum(1e-3 * (np.abs(mu)[:, None] * gross).sum(axis=0), 1e-300)))
    bad = {k: v for k, v in worst.items() if v > tol}
    return worst, bad


@pytest.mark.parametrize("cfg,steps,refresh,start", [
    (small_field(worlds=2, seed=5), 60, 1, "solve"),
    (small_field(worlds=2, seed=6), 60, 5, "solve"),
    (small_field(worlds=2, seed=6), 60, 5, "carried"),
    (FieldConfig(worlds=2, cells=36, row_width=6, side=40, spacing=5, seed=7), 20, 5, "solve"),
    (FieldConfig(worlds=2, cells=100, row_width=10, side=130, spacing=11, seed=8), 8, 4, "solve"),
], ids=["14-grid-refresh1", "14-grid-refresh5", "14-grid-carried", "40-grid", "130-grid"])
def test_field_steps_match_the_oracle(torch, cfg, steps, refresh, start):
    from cells_amd.field import FieldEnsemble

    fp = build_field(cfg)
    say(f"field parity: side {cfg.side}, {cfg.cells} cells, {cfg.worlds} worlds, {steps} steps, "
        f"refresh every {refresh}, start {start}")
    y0 = initial_field_state(fp)
    want, ok_ref = oracle.field_run(fp, y0, 1e-3, steps, refresh, start=start)
    assert ok_ref.all()
    e = FieldEnsemble(fp, refresh_every=refresh, start=start)
    say(f"plan {e.plan_stats()}")
    got, ok = e.step(y0, 1e-3, nsteps=steps)
    assert bool(ok.all())
    worst, bad = compare(fp, got.cpu().numpy(), want)
    say(f"worst {max(worst.values()):.3e}")
    assert not bad, bad
    e.close()


@pytest.mark.parametrize("cfg,steps", [
    (small_field(worlds=2, seed=18), 12),
    (FieldConfig(worlds=2, cells=36, row_width=6, side=40, spacing=5, seed=19), 10),
    (FieldConfig(worlds=2, cells=100, row_width=10, side=130, spacing=11, seed=20), 4),
], ids=["14-grid", "40-grid", "130-grid"])
def test_field_solve_matches_the_oracle(torch, cfg, steps):
    """The step-start potentials (cells_field_potentials, the chord iteration on the operator's
    factor) against the oracle's damped Newton (numerics/surface_field.py:614-855) at a state some
    steps into a run: the same root, met to the same target."""
    from cells_amd.field import FieldEnsemble

    fp = build_field(cfg)
    say(f"field solve parity: side {cfg.side}, {cfg.cells} cells, {steps} steps")
    y, ok_ref = oracle.field_run(fp, initial_field_state(fp), 1e-3, steps, 3)
    assert ok_ref.all()
    op = oracle.FieldOperator(fp)
    e = FieldEnsemble(fp, refresh_every=3)
    e.step(initial_field_state(fp), 1e-3, nsteps=1)  # an operator one step old: the chord's matrix
    phi, iterations, ok = e.potentials(y, 1e-3)
    assert bool(ok.all())
    phi = phi.cpu().numpy()
    say(f"chord steps {iterations.cpu().tolist()}")
    n = fp.config.cells
    cq, _ = charges(fp, y)
    for k in range(fp.W):
        cells, nodes, newton, conv = oracle.field_start(fp, y, op, k)
        assert conv
        scale = np.max(np.abs(nodes))
        assert scale > 0
        assert np.max(np.abs(phi[k] - nodes)) <= 1e-9 * scale, (k, np.max(np.abs(phi[k] - nodes)) / scale)
        c0 = int(fp.cells.cell_off[k])
        got_cells = phi[k][fp.cell_node[c0:c0 + n]] + cq[c0:c0 + n] / fp.cells.cap[c0:c0 + n]
        assert np.max(np.abs(got_cells - cells)) <= 1e-9 * np.max(np.abs(cells))
    e.close()


@pytest.mark.parametrize("cfg,steps", [
    (small_field(worlds=2, seed=21), 8),
    (FieldConfig(worlds=2, cells=36, row_width=6, side=40, spacing=5, seed=22), 6),
    (FieldConfig(worlds=2, cells=100, row_width=10, side=130, spacing=11, seed=23), 3),
], ids=["14-grid", "40-grid", "130-grid"])
def test_field_energy_rates_match_the_oracle(torch, cfg, steps):
    """The field solve's energy terms (cells_field_energy_rates; numerics/surface_field.py:792-835)
    against the oracle's at the device's own potentials: every rate within 1e-10 of the world's
    dissipation, the capacitor energy within 1e-12 of itself; the balance vanishes on the device
    (below 1e-8 of the dissipation: the chord iteration meets the field solve's continuity target)
    and the dissipation is the three routes' heat. Against the oracle's own field solve (a Newton
    root met to the same target) the terms agree within 1e-6 of the dissipation."""
    from cells_amd.field import FieldEnsemble

    fp = build_field(cfg)
    say(f"field energy rates: side {cfg.side}, {cfg.cells} cells, {steps} steps")
    y, ok_ref = oracle.field_run(fp, initial_field_state(fp), 1e-3, steps, 3)
    assert ok_ref.all()
    op = oracle.FieldOperator(fp)
    e = FieldEnsemble(fp, refresh_every=3)
    got, ok_e = e.energy_rates(y, 1e-3)
    assert bool(ok_e.all())
    got = got.cpu().numpy()
    phi, _, ok = e.potentials(y, 1e-3)
    assert bool(ok.all())
    phi = phi.cpu().numpy()
    names = oracle.FIELD_ENERGY_RATES
    ix = {n: names.index(n) for n in names}
    rates = [ix[n] for n in names if n != "capacitor_energy"]
    n = fp.config.cells
    cq, _ = charges(fp, y)
    for k in range(fp.W):
        c0 = int(fp.cells.cell_off[k])
        cells = phi[k][fp.cell_node[c0:c0 + n]] + cq[c0:c0 + n] / fp.cells.cap[c0:c0 + n]
        want, valid = oracle.field_energy_rates(fp, y, op, k, phi=(cells, phi[k]))
        assert valid
        d = want[ix["dissipation"]]
        assert d > 0
        err = np.abs(got[k] - want)
        assert np.max(err[rates]) <= 1e-10 * d, (k, dict(zip(names, err / d)))
        ce = ix["capacitor_energy"]
        assert err[ce] <= 1e-12 * abs(want[ce])
        assert abs(got[k][ix["energy_balance"]]) <= 1e-8 * d, (k, got[k][ix["energy_balance"]] / d)
        heat = got[k][ix["membrane_heat"]] + got[k][ix["junction_heat"]] + got[k][ix["face_heat"]]
        assert abs(got[k][ix["dissipation"]] - heat) <= 1e-12 * d
        solved, conv = oracle.field_energy_rates(fp, y, op, k)
        assert conv
        assert np.max(np.abs(got[k] - solved)[rates]) <= 1e-6 * d
    e.close()


def test_field_steps_in_pieces_equal_one_call(torch):
    """Steps split over calls (the operator's refresh count carried across) give the same bits."""
    from cells_amd.field import FieldEnsemble

    fp = build_field(small_field(worlds=2, seed=9))
    y0 = initial_field_state(fp)
    a = FieldEnsemble(fp, refresh_every=4)
    whole, _ = a.step(y0, 1e-3, nsteps=10)
    b = FieldEnsemble(fp, refresh_every=4)
    y = y0
    for n in (3, 4, 3):
        y, ok = b.step(y, 1e-3, nsteps=n)
        assert bool(ok.all())
    assert torch.equal(whole, y)
    a.close()
    b.close()


def test_full_grid_closes_material_and_charge(torch):
    from cells_amd.field import FieldEnsemble

    fp = build_field(FieldConfig(worlds=2))
    L = FieldLayout.of(fp)
    e = FieldEnsemble(fp, refresh_every=50)
    stats = e.plan_stats()
    say(f"full grid plan {stats}")
    assert stats["max_front"] <= 512
    y0 = initial_field_state(fp)
    y = e.to_device(y0)
    ring = fp.ring()
    c0, n0 = amounts(fp, y0)
    for chunk in range(3):
        y, ok = e.step(y, 1e-3, nsteps=20)
        assert bool(ok.all())
        yh = y.cpu().numpy()
        c1, n1 = amounts(fp, yh)
        for s in range(fp.S):
            for k in range(fp.W):
                cs = slice(int(fp.cells.cell_off[k]), int(fp.cells.cell_off[k + 1]))
                before = c0[s, cs].sum() + n0[s, k, ~ring].sum()
                after = c1[s, cs].sum() + n1[s, k, ~ring].sum()
                sup = L.world(yh, L.sup + s)[k]
                assert abs(after - before - sup) <= 1e-12 * before, (s, k, after - before - sup)
        cq, nq = charges(fp, yh)
        for k in range(fp.W):
            cs = slice(int(fp.cells.cell_off[k]), int(fp.cells.cell_off[k + 1]))
            q = nq[k].copy()
            q[fp.cell_node[cs]] += cq[cs]
            assert np.max(np.abs(q[~ring])) <= 1e-12 * np.max(np.abs(cq[cs]))
        phi = L.node(yh, L.phi)
        say(f"after {20 * (chunk + 1)} steps: node potentials {phi.min():.3e} .. {phi.max():.3e} V")
    start, iterations, ok = e.potentials(y, 1e-3)
    assert bool(ok.all())
    say(f"field solve at the end: chord steps {iterations.cpu().tolist()}")
    e.close()


def test_a_refused_world_keeps_its_state(torch):
    """A world whose step leaves its domain (a membrane availability above 1 is refused by the
    pump's domain, as the reference's control check refuses it) keeps its input state; ok = 0."""
    from cells_amd.ensemble import control
    from cells_amd.field import FieldEnsemble

    fp = build_field(small_field(worlds=2, seed=10))
    e = FieldEnsemble(fp, refresh_every=1)
    y0 = e.to_device(initial_field_state(fp))
    y1, ok = e.step(y0, 1e-3, nsteps=1, ctl=control(pump=1.5))
    assert not bool(ok.any())
    assert torch.equal(y1, y0)
    y2, ok = e.step(y0, 1e-3, nsteps=2, active=[1, 0])
    assert ok.cpu().tolist() == [1, 0]
    L = FieldLayout.of(fp)
    got = y2.cpu().numpy()
    want, _ = oracle.field_run(fp, initial_field_state(fp), 1e-3, 2, 1)
    n = fp.config.cells
    assert np.array_equal(L.cell(got, L.dn)[n:], np.zeros(n))  # world 1 untouched
    assert np.max(np.abs(L.cell(got, L.dn)[:n] - L.cell(want, L.dn)[:n])) <= 1e-9 * np.max(
        np.abs(L.cell(want, L.dn)[:n]))
    e.close()


def test_an_unconverged_field_solve_refuses_the_step(torch):
    """A world whose field solve misses the Newton target within its chord-step bound refuses the
    step and keeps its state, as the reference refuses an unconverged field (models/
    surface_kinetic.py readouts: valid = converged & ...). The bound is forced from the host: one
    chord step, where these states need more."""
    from cells_amd.field import FieldEnsemble

    fp = build_field(small_field(worlds=2, seed=21))
    y, ok_ref = oracle.field_run(fp, initial_field_state(fp), 1e-3, 6, 1)
    assert ok_ref.all()
    full = FieldEnsemble(fp, refresh_every=1)
    _, chord, ok = full.potentials(y, 1e-3)
    assert bool(ok.all()) and int(chord.min()) >= 2
    full.close()
    short = FieldEnsemble(fp, refresh_every=1, max_iterations=1)
    y0 = short.to_device(y)
    y1, ok = short.step(y0, 1e-3, nsteps=1)
    assert not bool(ok.any())
    assert torch.equal(y1, y0)
    _, chord, ok = short.potentials(y0, 1e-3)
    assert not bool(ok.any()) and chord.cpu().tolist() == [1, 1]
    short.close()
