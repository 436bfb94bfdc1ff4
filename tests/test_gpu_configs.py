"""GPU checks at the configurations' own world shapes, through the C-ABI.

- Batched ensemble: ROS2 trajectories of a world of more than 2048 cells (beside a small one)
  over 200 steps against the oracle, 1e-9 relative.
- Adjoint calibration: worlds of 32x32 = 1024 cells with the six log-parameters; one reverse
  step against the oracle's adjoint at that size (1e-9); the calibration gradient over 200 steps
  (a Vmem target every 2 steps, 100 rows) with the binomial schedule over 45 snapshots, against
  central differences of the HIP loss along 16 random directions (h = 1e-6, 1e-7 relative), and
  the taped schedule bit for bit the recomputing one.
- Newton-Krylov: worlds of 64x64 = 4096 cells, 4 ions, junction conductances lognormal sigma
  0.3: every world converges, the oracle's rates vanish at the result (below 1e-6 of those at
  the start), and the FP32-Krylov result equals the all-FP64 one within 1e-9 mM.
"""

import sys

import numpy as np
import pytest

import oracle
from cells_amd.world import StateLayout, build, config1, config4, initial_state
from tests.helpers import conc, material_residual, relerr, vmem

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


def three_ion(worlds, cells, seed):
    return build(config1(worlds=worlds, cells=cells, species=("Na", "K", "Cl"),
                         inside_mM=(12.0, 140.0, 10.0), bath_mM=(145.0, 5.0, 110.0),
                         permeability=(1e-10, 5e-9, 1e-9), seed=seed))


def test_ros2_trajectory_of_a_world_above_2048_cells(torch):
    from cells_amd.ensemble import Ensemble

    p = build(config1(worlds=2, cells=(2900, 3300), seed=44))
    sizes = np.diff(p.cell_off)
    assert sizes.max() > 2048, sizes
    e = Ensemble(p, "ros2")
    _, sym = oracle.world_symbolics(p)
    w = oracle.world_struct(p)
    y0 = initial_state(p)
    say(f"large-world ROS2: cells {sizes.tolist()}")
    got, ok = e.step(y0, 5e-3, nsteps=200)
    assert bool(ok.all())
    got = got.cpu().numpy()
    ref = y0
    for k in range(200):
        ref, okr = oracle.step_ros2(p, ref, 5e-3, sym, world=w)
        assert okr.all(), k
    err = max(relerr(vmem(p, got), vmem(p, ref)), relerr(conc(p, got), conc(p, ref)))
    say(f"large-world ROS2: worst relative error {err:.3e}")
    assert err <= 1e-9
    assert material_residual(p, got) <= 1e-12
    e.close()


def test_adjoint_step_at_the_calibration_shape_matches_the_oracle(torch):
    from cells_amd.ensemble import Ensemble

    p = three_ion(2, (1024, 1024), seed=3)
    L = StateLayout.of(p)
    _, sym = oracle.world_symbolics(p)
    w = oracle.world_struct(p)
    y = initial_state(p)
    for _ in range(5):
        y, _ = oracle.step_ros2(p, y, 5e-3, sym, world=w)
    rng = np.random.default_rng(11)
    ybar = np.zeros_like(y)
    for s in range(p.S):
        L.cell(ybar, L.dn + s)[:] = rng.normal(size=p.ncell)
    L.edge(ybar, L.gate)[:] = rng.normal(size=p.nedge)
    yin_o, pam_o, cap_o, paj_o = oracle.ros2_vjp(p, y, 5e-3, sym, ybar, world=w)
    e = Ensemble(p, "ros2")
    say("adjoint step at 32x32")
    yin, pam, cap, paj, valid = e.step_vjp(y, ybar, 5e-3)
    assert bool(valid.all())
    yin = yin.cpu().numpy()
    for s in range(p.S):
        a, b = L.cell(yin, L.dn + s), L.cell(yin_o, L.dn + s)
        assert relerr(a, b, floor=1e-9 * np.abs(b).max()) <= 1e-9, s
    a, b = L.edge(yin, L.gate), L.edge(yin_o, L.gate)
    assert relerr(a, b, floor=1e-9 * np.abs(b).max()) <= 1e-9
    for got, want in ((pam, pam_o), (cap, cap_o), (paj[:, :p.nedge], paj_o)):
        g = got.cpu().numpy()
        assert relerr(g, want, floor=1e-9 * np.abs(want).max()) <= 1e-9
    e.close()


def test_calibration_gradient_at_the_calibration_shape(torch):
    from cells_amd.adjoint import Calibration
    from cells_amd.ensemble import Ensemble

    p = three_ion(2, (1024, 1024), seed=3)
    e = Ensemble(p, "ros2")
    y0 = initial_state(p)
    dt, steps, every = 5e-3, 200, 2
    rng = np.random.default_rng(7)
    hidden = rng.normal(scale=0.1, size=(6, p.ncell))
    samples = steps // every
    cal = Calibration(e, y0, dt, steps, every, [np.zeros(p.ncell)] * samples)
    cal.set_theta(hidden)
    y = e.to_device(y0)
    targets = []
    for _ in range(samples):
        y, ok = e.step(y, dt, nsteps=every)
        assert bool(ok.all())
        targets.append(vmem(p, y.cpu().numpy()))
    cal = Calibration(e, y0, dt, steps, every, targets)
    theta = np.zeros((6, p.ncell))
    say("calibration at 32x32: 45-snapshot recomputing schedule")
    loss, grad = cal.loss_and_gradient(theta, snapshots=45, tape=0)
    assert loss > 0 and cal.reverse_steps == steps
    say(f"calibration: loss {loss:.6e}, {cal.forward_steps} forward steps")
    for trial in range(16):
        d = rng.normal(size=theta.shape)
        h = 1e-6
        fd = (cal.loss(theta + h * d) - cal.loss(theta - h * d)) / (2 * h)
        an = float(np.sum(grad * d))
        assert abs(fd - an) <= 1e-7 * abs(fd), (trial, fd, an, abs(fd - an) / abs(fd))
    e.reserve_tape(5)
    say("calibration: taped schedule (5-step units)")
    loss_t, grad_t = cal.loss_and_gradient(theta, snapshots=45, tape=5)
    assert loss_t == loss and np.array_equal(grad_t, grad)
    e.close()


def test_newton_krylov_at_the_precision_shape(torch):
    from cells_amd.ensemble import Ensemble

    p = build(config4(worlds=2))
    assert np.all(np.diff(p.cell_off) == 4096)
    e = Ensemble(p, "ros2")
    y0 = initial_state(p)
    say("stationary 64x64 (fp32 inner)")
    y32, conv, step, iters, rep = e.stationary(y0, fp32_inner=True)
    say(f"report {rep}")
    assert bool(conv.all()) and float(step.max()) <= 1e-10
    ys = y32.cpu().numpy()
    L = StateLayout.of(p)

    def coordinates(r):
        return np.concatenate([L.cell(r, L.dn + s) / p.vol0 for s in range(p.S)] + [L.edge(r, L.gate)])

    r0 = coordinates(oracle.rates(p, y0)[0])
    r1, ok = oracle.rates(p, ys)
    assert ok.all()
    r1 = coordinates(r1)
    na = p.S * p.ncell
    for sl in (slice(0, na), slice(na, None)):
        assert np.max(np.abs(r1[sl])) <= 1e-6 * np.max(np.abs(r0[sl])), np.max(np.abs(r1[sl]))
    say("stationary 64x64 (all fp64)")
    y64, conv64, _, _, rep64 = e.stationary(y0, fp32_inner=False)
    assert bool(conv64.all()), rep64
    y64 = y64.cpu().numpy()
    for s in range(p.S):
        assert np.max(np.abs((L.cell(ys, L.dn + s) - L.cell(y64, L.dn + s)) / p.vol0)) <= 1e-9
    assert np.max(np.abs(L.edge(ys, L.gate) - L.edge(y64, L.gate))) <= 1e-9
    e.close()
