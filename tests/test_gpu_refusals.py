"""The step's refusals on the GPU, through the C-ABI, against the oracle and the definitions.

- Definition-time rate bounds (models/surface_pump.py:402-555 StepChecks, explicit_rate_bound,
  junction_gate_rate_bound; surface_pump_v2.py:583-617; surface_pump_v3.py:4981-5017): the
  bounds the library reports equal the definitions' (computed here from the packed parameters),
  the configurations' own steps are admitted and a dt beyond a bound is CELLS_EINVAL with the
  reference's message, before anything is launched.
- The explicit stage's membrane rate (surface_pump_v3.py:3640-3642): IMEX Euler refuses a world
  whose stage rate times dt exceeds 1, as the oracle does, with the rate bit.
- The stage solves' charge residual (numerics/surface_imex.py:269-278 as surface_pump.py:325-330
  checks it): the default bound admits every step; a zero bound refuses every world with the
  residual bit, and restoring the bound restores the step exactly (the captured graph follows
  the tolerance).
- The captured step graph follows the ledger checks' tolerance alone (a call that changes only
  the tolerance is checked against the new one).
"""

import sys

import numpy as np
import pytest

import oracle
from cells_amd import lib as hip
from cells_amd.world import F, R, StateLayout, build, config0, config1, initial_state
from tests.helpers import conc, relerr, vmem

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


def ensemble(p, method):
    from cells_amd.ensemble import Ensemble

    return Ensemble(p, method)


def definition_bounds(p, method):
    """The reference's definition-time bounds from the packed parameters (declared state)."""
    T = p.config.temperature
    slope = F * F / (R * T)
    c = p.n0 / p.vol0[None, :]
    z2 = p.valence[:, None] ** 2
    if method == "euler":
        gj = slope * np.sum(z2 * p.paj * np.maximum(c[:, p.snd], c[:, p.rcv]), axis=0)
        rows = np.zeros(p.ncell)
        np.add.at(rows, p.snd, gj)
        np.add.at(rows, p.rcv, gj)
        gm = slope * np.sum(z2 * p.pam * np.maximum(c, p.bath_c[:, None]), axis=0)
        explicit = max(2 * np.max(rows / p.cap), np.max(gm / p.cap))
        gate = 1.3 * np.exp(77 * p.config.gate_threshold) + 20 if p.config.gate_minimum < 1 else 0.0
    else:
        x = p.valence[:, None] * F * (p.q0 / p.cap)[None, :] / (R * T)
        b = np.where(np.abs(x) < 1e-4, 1 - x / 2, -x / np.expm1(-x))  # B(-x)
        mat = p.pam * b
        np.add.at(mat, (slice(None), p.snd), p.paj)
        np.add.at(mat, (slice(None), p.rcv), p.paj)
        explicit = 2 * np.max(mat / (p.vol0 * 1e-15)[None, :])
        gate = 0.0
    ca = 2 * np.max(p.ca_cap / p.config.ca_km_ca / p.vol0) if p.calcium else 0.0
    return explicit, gate, ca


def test_definition_bounds_admit_the_configurations_and_refuse_beyond(torch):
    for cfg, method, dt in ((config0(), "euler", 1e-3), (config1(worlds=3, cells=(40, 90), seed=2), "ros2", 5e-3)):
        p = build(cfg)
        e = ensemble(p, method)
        b = e.step_bounds()
        want = definition_bounds(p, method)
        say(f"{method}: bounds {b} (definitions {want}), dt x explicit {dt * b['explicit']:.3f}")
        assert np.allclose([b["explicit"], b["gate"], b["calcium"]], want, rtol=1e-12, atol=0)
        assert dt * b["explicit"] <= 1 and dt * b["gate"] <= 0.125 and dt * b["calcium"] <= 1
        y0 = initial_state(p)
        _, ok = e.step(y0, dt, nsteps=2)
        assert bool(ok.all())
        with pytest.raises(hip.CellsError, match="explicit-rate stability bound") as err:
            e.step(y0, 1.01 / b["explicit"], nsteps=1)
        assert err.value.code == hip.CELLS_EINVAL
        e.close()
    # the plain configuration's junctions set its bound: dt x rate 0.84 for a cell with the mean
    # junctions, 0.885 for the worst cell of the seeded lognormal spread, at 1 ms
    p = build(config0())
    e = ensemble(p, "euler")
    assert 0.84 <= 1e-3 * e.step_bounds()["explicit"] <= 0.95
    e.close()


def loaded(p, world, cell, factor=4.0):
    """The initial state with cell `cell` of `world` holding (1 + factor) times its K and as much
    extra Cl (charge-neutral): its membranes conduct, and relax, several times faster."""
    L = StateLayout.of(p)
    y = initial_state(p)
    i = int(p.cell_off[world]) + cell
    extra = factor * p.n0[1][i]
    L.cell(y, L.dn + 1)[i] += extra
    L.cell(y, L.dn + 2)[i] += extra
    return y


def test_euler_refuses_a_stage_above_the_membrane_rate_as_the_oracle_does(torch):
    p = build(config0(worlds=3, cells=(16, 16), seed=2, water=False))
    e = ensemble(p, "euler")
    dt = 0.98 / e.step_bounds()["explicit"]  # admitted by the definition
    y = loaded(p, 1, 5)
    out, ok = e.step(y, dt, nsteps=1)
    bits = e.check_bits().cpu().numpy()
    ref, okr = oracle.step_euler(p, y, dt)
    say(f"loaded cell: ok {ok.cpu().numpy().tolist()}, oracle {okr.tolist()}, bits {bits.tolist()}")
    assert ok.cpu().numpy().tolist() == [1, 0, 1] and okr.tolist() == [True, False, True]
    assert bits[1] & hip.REFUSE_EXPLICIT_RATE and not bits[[0, 2]].any()
    got = out.cpu().numpy()
    L = StateLayout.of(p)
    c0, c1 = p.cell_off[1], p.cell_off[2]
    assert np.array_equal(L.cell(got, L.dn + 1)[c0:c1], L.cell(y, L.dn + 1)[c0:c1])  # kept
    assert relerr(conc(p, got), conc(p, ref)) <= 1e-12
    # at a third of the membrane availability the same step is admitted by both
    from cells_amd.ensemble import control

    out, ok = e.step(y, dt, nsteps=1, ctl=control(membrane=0.3))
    ref, okr = oracle.step_euler(p, y, dt, ctl=oracle.control(membrane=0.3))
    assert bool(ok.all()) and okr.all()
    got = out.cpu().numpy()
    assert max(relerr(vmem(p, got), vmem(p, ref)), relerr(conc(p, got), conc(p, ref))) <= 1e-12
    e.close()


def test_stage_residual_bound_refuses_and_the_graph_follows_it(torch):
    p = build(config1(worlds=3, cells=(60, 200), seed=7))
    e = ensemble(p, "ros2")
    y0 = initial_state(p)
    ref, ok = e.step(y0, 5e-3, nsteps=3)
    assert bool(ok.all()) and int(e.check_bits().abs().sum()) == 0
    e.stage_tolerance(0.0, 0.0)  # no rounding admitted: every world's residual trips it
    say("stage residual bound 0")
    out, ok = e.step(y0, 5e-3, nsteps=3)
    bits = e.check_bits().cpu().numpy()
    assert not bool(ok.any()) and all(b & hip.REFUSE_RESIDUAL for b in bits)
    assert np.array_equal(out.cpu().numpy(), y0)
    e.stage_tolerance()  # the defaults again: the same steps as before
    again, ok = e.step(y0, 5e-3, nsteps=3)
    assert bool(ok.all()) and torch.equal(again, ref)
    # the oracle's residual check admits the same steps
    _, sym = oracle.world_symbolics(p)
    yr = y0
    for _ in range(3):
        yr, okr = oracle.step_ros2(p, yr, 5e-3, sym)
        assert okr.all()
    e.close()


def test_the_step_graph_follows_the_check_tolerance(torch):
    p = build(config1(worlds=3, cells=(30, 80), seed=17))
    L = StateLayout.of(p)
    e = ensemble(p, "ros2")
    bad = initial_state(p)
    L.cell(bad, L.mtr + 1)[int(p.cell_off[1])] += 1e-3  # world 1's material ledger is off by 1e-3
    e.step_checks(True, material_atol=1.0, steps_before=0)
    _, ok = e.step(bad, 5e-3, nsteps=2)
    assert bool(ok.all()) and int(e.check_bits().abs().sum()) == 0
    e.step_checks(True, material_atol=1e-12, steps_before=0)  # only the tolerance changes
    say("check tolerance 1e-12 after 1.0")
    _, ok = e.step(bad, 5e-3, nsteps=2)
    bits = e.check_bits().cpu().numpy()
    assert ok.cpu().numpy().tolist() == [1, 0, 1] and bits[1] & 2
    e.step_checks(False)
    e.close()
