"""GPU parity of the HIP step against the oracle (FP64 CPU restatement), through the C-ABI.

Bars: rates within 1e-12 relative (ulp-level differences of the FP64 transcendentals); trajectories
of Vmem, concentrations and volumes within 1e-9 relative over the plain configuration's 1e4 steps;
the numeric factor and the solve bit for bit in the reference-order (pairwise) mode (same sums in
the same order, no FMA contraction), and within 1e-13 (factor) and 1e-12 (solve) in the default
supernodal mode;
batched (one graph for all worlds) equal to plain (one world per launch) bit for bit; material
ledgers closed to 1e-12 of their magnitudes.
"""

import sys

import numpy as np
import pytest

import oracle
from cells_amd.world import Config, StateLayout, build, config0, config1, initial_state
from tests.helpers import conc, material_residual, relerr, vmem, volume

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


def ensemble(p, method="euler"):
    from cells_amd.ensemble import Ensemble

    return Ensemble(p, method)


def test_native_library_loads_on_gpu(torch):
    from cells_amd import lib

    L = lib.load()
    for name in lib.EXPORTS:
        assert hasattr(L, name)


def test_rates_match_oracle(torch):
    for name, cfg in (("config0", config0()), ("config1-small", config1(worlds=3, cells=(40, 120)))):
        say(f"rates parity {name}")
        p = build(cfg)
        y = initial_state(p)
        # perturb the amounts so every route carries flux both ways: Na and Cl by the same
        # amount (no charge), K by a few mV's worth of charge, Ca by 5 %
        L = StateLayout.of(p)
        rng = np.random.default_rng(3)
        a = p.n0[0] * rng.uniform(-0.05, 0.05, p.ncell)
        L.cell(y, L.dn + 0)[:] = a
        L.cell(y, L.dn + 2)[:] = a
        L.cell(y, L.dn + 1)[:] = rng.uniform(-5e-4, 5e-4, p.ncell)
        if p.S > 3:
            L.cell(y, L.dn + 3)[:] = p.n0[3] * rng.uniform(-0.05, 0.05, p.ncell)
        L.edge(y, L.gate)[:] = rng.uniform(0.5, 1.0, p.nedge)
        e = ensemble(p)
        r, valid = e.rates(y)
        r = r.cpu().numpy()
        ro, vo = oracle.rates(p, y)
        assert valid.cpu().numpy().all() and vo.all()
        for f in range(L.cell_fields):
            a, b = L.cell(r, f), L.cell(ro, f)
            assert relerr(a, b, floor=1e-12 * np.abs(b).max() + 1e-300) <= 1e-12, (name, "cell", f)
        for f in range(L.edge_fields):
            a, b = L.edge(r, f), L.edge(ro, f)
            assert relerr(a, b, floor=1e-12 * np.abs(b).max() + 1e-300) <= 1e-12, (name, "edge", f)
        for s in range(p.S):
            a, b = L.world(r, L.sup + s), L.world(ro, L.sup + s)
            assert relerr(a, b, floor=1e-12 * np.abs(b).max() + 1e-300) <= 1e-11, (name, "supply", s)
        e.close()


def test_plain_config_trajectory_matches_oracle(torch):
    """Config 0: one 8x8 world, IMEX Euler dt = 1 ms, water half-steps, 1e4 steps."""
    p = build(config0())
    e = ensemble(p)
    y_gpu = initial_state(p)
    y_ref = initial_state(p)
    w = oracle.world_struct(p)
    worst = 0.0
    for block in range(10):
        say(f"config0 block {block}")
        out, ok = e.step(y_gpu, 1e-3, nsteps=1000)
        assert ok.cpu().numpy().all()
        y_gpu = out.cpu().numpy()
        for _ in range(1000):
            y_ref, okr = oracle.step_euler(p, y_ref, 1e-3, world=w)
            assert okr.all()
        worst = max(worst, relerr(vmem(p, y_gpu), vmem(p, y_ref)),
                    relerr(conc(p, y_gpu), conc(p, y_ref)),
                    relerr(volume(p, y_gpu), volume(p, y_ref)))
    say(f"config0 worst relative error {worst:.3e}")
    assert worst <= 1e-9
    assert material_residual(p, y_gpu) <= 1e-12
    e.close()


@pytest.mark.parametrize("method,cfg,dt,steps", [
    ("ros2", config1(worlds=12, cells=(1, 7), seed=3), 5e-3, 40),
    ("euler", config0(worlds=5, cells=(1, 4), seed=4), 1e-3, 200),
], ids=["ros2-1-to-7-cells", "euler-water-1-to-4-cells"])
def test_tiny_and_ragged_worlds_match_the_oracle(torch, method, cfg, dt, steps):
    """Worlds of one cell (no junctions: a 1 x 1 stage matrix) up to a few cells, mixed in one
    batch: the trajectory against the oracle (1e-9), batched against plain bit for bit, material
    closed."""
    p = build(cfg)
    say(f"tiny worlds {method}: cells {np.diff(p.cell_off).tolist()}, junctions {np.diff(p.edge_off).tolist()}")
    e = ensemble(p, method)
    y0 = initial_state(p)
    got, ok = e.step(y0, dt, nsteps=steps)
    assert ok.cpu().numpy().all()
    got = got.cpu().numpy()
    want = y0
    w = oracle.world_struct(p)
    if method == "ros2":
        _, sym = oracle.world_symbolics(p)
        for _ in range(steps):
            want, okr = oracle.step_ros2(p, want, dt, sym, world=w)
            assert okr.all()
    else:
        for _ in range(steps):
            want, okr = oracle.step_euler(p, want, dt, world=w)
            assert okr.all()
    err = max(relerr(vmem(p, got), vmem(p, want)), relerr(conc(p, got), conc(p, want)))
    if p.water:
        err = max(err, relerr(volume(p, got), volume(p, want)))
    say(f"tiny worlds worst relative error {err:.3e}")
    assert err <= 1e-9
    assert material_residual(p, got) <= 1e-12
    a, oka = e.step(y0, dt, nsteps=5, batched=True)
    b, okb = e.step(y0, dt, nsteps=5, batched=False)
    assert oka.cpu().numpy().all() and okb.cpu().numpy().all()
    assert torch.equal(a, b)
    e.close()


def test_batched_equals_plain_bitwise(torch):
    for method, cfg, dt in (("euler", config0(worlds=4, cells=(20, 70), seed=5), 1e-3),
                            ("ros2", config1(worlds=4, cells=(30, 90), seed=6), 5e-3)):
        say(f"batched vs plain {method}")
        p = build(cfg)
        e = ensemble(p, method)
        y = initial_state(p)
        a, oka = e.step(y, dt, nsteps=7, batched=True)
        b, okb = e.step(y, dt, nsteps=7, batched=False)
        assert oka.cpu().numpy().all() and okb.cpu().numpy().all()
        assert torch.equal(a, b)
        e.close()


def test_ros2_trajectory_matches_oracle(torch):
    p = build(config1(worlds=3, cells=(256, 420), seed=7))
    syms, sym = oracle.world_symbolics(p)
    e = ensemble(p, "ros2")
    y_gpu = initial_state(p)
    y_ref = initial_state(p)
    w = oracle.world_struct(p)
    for block in range(4):
        say(f"ros2 block {block}")
        out, ok = e.step(y_gpu, 5e-3, nsteps=25)
        assert ok.cpu().numpy().all()
        y_gpu = out.cpu().numpy()
        for _ in range(25):
            y_ref, okr = oracle.step_ros2(p, y_ref, 5e-3, sym, world=w)
            assert okr.all()
    err = max(relerr(vmem(p, y_gpu), vmem(p, y_ref)), relerr(conc(p, y_gpu), conc(p, y_ref)))
    say(f"ros2 worst relative error {err:.3e}")
    assert err <= 1e-9
    assert material_residual(p, y_gpu) <= 1e-12
    e.close()


def front_orders(n):
    """Front orders of the default multifrontal plan of the n-cell lattice: the relaxed pattern's
    fundamental supernodes (column j joins j + 1 when j + 1 is its parent with no other child and
    holds its rows but itself), each front the order of its first column."""
    from cells_amd import lib as hip
    from cells_amd import symbolic
    from cells_amd.world import lattice

    relax_rows, relax_zeros = (np.zeros(1, np.int32) for _ in range(2))
    hip.check(hip.load().cells_get_plan_options(relax_rows.ctypes.data_as(hip._i32p),
                                                relax_zeros.ctypes.data_as(hip._i32p), None, None, None, None))
    snd, rcv = lattice(n)
    order, pr, pc = symbolic.relaxed_padding(n, snd, rcv, int(relax_rows[0]), int(relax_zeros[0]))
    s = symbolic.bind(n, np.concatenate((snd, pr)), np.concatenate((rcv, pc)), order)
    count = np.diff(s.column_start)
    kids = np.bincount(s.parent[s.parent >= 0], minlength=n)
    orders, k = [], 0
    while k < n:
        j = k
        while j + 1 < n and s.parent[j] == j + 1 and kids[j + 1] == 1 and count[j] == count[j + 1] + 1:
            j += 1
        orders.append(int(count[k]))
        k = j + 1
    return orders


@pytest.mark.parametrize("mode,cells,seed", [("pairwise", (64, 300), 8), ("supernodal", (64, 300), 8),
                                             ("supernodal", (1500, 4096), 3), ("supernodal", (4096, 4096), 5)])
def test_factor_and_solve_against_oracle(torch, mode, cells, seed):
    """Pairwise mode: the factor bit for bit, and the solve on it bit for bit the oracle's solve.
    Supernodal mode: the factor within 1e-13 of the largest entry (Jacobi-scaled SPD stage
    matrices) and the supernodal solve on it within 1e-12 of the oracle's solve on it; worlds of
    4096 cells hold workgroup fronts of order 91-100 (in LDS) and 101-128 (in HBM scratch). The
    pairwise plan's layout is the bound pattern's; the supernodal plan's is the relaxed pattern's,
    whose extra entries (the joined chains' explicit zeros) stay zero."""
    from cells_amd.ensemble import Ensemble

    p = build(config1(worlds=3, cells=cells, seed=seed))
    syms, sym = oracle.world_symbolics(p)
    if cells == (4096, 4096):
        # the plan of these worlds holds fronts of both workgroup kinds (checked on the host)
        orders = front_orders(4096)
        assert any(91 <= o <= 100 for o in orders) and any(101 <= o <= 128 for o in orders), max(orders)
    e = Ensemble(p, "ros2", factor=mode)
    rng = np.random.default_rng(9)
    values = np.zeros(e.factor_size)
    rhs = rng.uniform(-1, 1, p.ncell)
    expect_f, where, extra = [], [], []
    for k in range(p.W):
        d, ep, off = e.positions(k)
        c0, c1 = p.cell_off[k], p.cell_off[k + 1]
        e0, e1 = p.edge_off[k], p.edge_off[k + 1]
        n = c1 - c0
        off_diag = rng.uniform(-1, 0, e1 - e0)
        diag = np.full(n, 0.5) + rng.uniform(0.5, 2.0, n)
        np.add.at(diag, p.snd[e0:e1] - c0, -off_diag)
        np.add.at(diag, p.rcv[e0:e1] - c0, -off_diag)
        # Jacobi scaling, as the stage assembly does
        sc = 1.0 / np.sqrt(diag)
        off_diag = off_diag * sc[p.snd[e0:e1] - c0] * sc[p.rcv[e0:e1] - c0]
        diag = np.ones(n)
        nnz = int(syms[k].column_start[-1])
        size = (e.positions(k + 1)[2] if k + 1 < p.W else e.factor_size) - off
        # the plan's position of every entry of the bound pattern
        pos = e.factor_positions(k, syms[k].row, syms[k].column)
        assert np.all(pos >= 0) and np.unique(pos).size == nnz
        if mode == "pairwise":
            assert size == nnz and np.array_equal(pos, np.arange(nnz))
        v = np.zeros(size)
        v[d] = diag
        v[ep] = off_diag
        values[off:off + size] = v
        # the oracle's positions are the same pattern: check, then factor there
        vo = np.zeros(nnz)
        vo[np.asarray([sym[k].diag_pos[i] for i in range(n)])] = diag
        vo[np.asarray([sym[k].edge_pos[q] for q in range(e1 - e0)])] = off_diag
        assert np.array_equal(v[pos], vo) and np.count_nonzero(v) == np.count_nonzero(vo)
        expect_f.append(oracle.factor(sym[k], vo))
        where.append(off + pos)
        rest = np.ones(size, bool)
        rest[pos] = False
        extra.append(off + np.nonzero(rest)[0])
    say(f"factor ({mode}, cells {cells}, {sum(x.size for x in extra)} explicit zeros)")
    fg = e.factor(values).cpu().numpy()
    xg = e.solve(fg, rhs).cpu().numpy()
    for k in range(p.W):
        f = expect_f[k]
        got = fg[where[k]]
        assert np.all(fg[extra[k]] == 0.0)
        if mode == "pairwise":
            assert got.tobytes() == f.tobytes()
        else:
            assert np.max(np.abs(got - f)) <= 1e-13 * np.max(np.abs(f))
        c0, c1 = p.cell_off[k], p.cell_off[k + 1]
        want = oracle.solve(sym[k], got, rhs[c0:c1])
        if mode == "pairwise":
            assert xg[c0:c1].tobytes() == want.tobytes()
        else:
            err = np.max(np.abs(xg[c0:c1] - want)) / np.max(np.abs(want))
            assert err <= 1e-12, err
    e.close()


def test_a_refused_world_keeps_its_state_and_the_others_step(torch):
    p = build(config0(worlds=3, cells=(16, 16), seed=2))
    L = StateLayout.of(p)
    y = initial_state(p)
    # world 1: an amount below zero is outside the domain, so its every step refuses
    c0 = p.cell_off[1]
    L.cell(y, L.dn)[c0] = -2.0 * p.n0[0][c0]
    e = ensemble(p)
    out, ok = e.step(y, 1e-3, nsteps=3)
    out = out.cpu().numpy()
    ok = ok.cpu().numpy()
    assert ok.tolist() == [1, 0, 1]
    yo = y.copy()
    ref = y.copy()
    w = oracle.world_struct(p)
    for _ in range(3):
        ref, okr = oracle.step_euler(p, ref, 1e-3, world=w)
    assert okr.tolist() == [True, False, True]
    c1 = p.cell_off[2]
    for f in range(L.cell_fields):
        assert np.array_equal(L.cell(out, f)[c0:c1], L.cell(yo, f)[c0:c1])
    assert relerr(vmem(p, out)[:c0], vmem(p, ref)[:c0]) <= 1e-12
    e.close()


def test_device_checks_pass_on_accepted_rows_and_catch_a_planted_ledger_error(torch):
    p = build(config1(worlds=3, cells=(40, 90), seed=12))
    e = ensemble(p, "ros2")
    y0 = initial_state(p)
    y1, ok = e.step(y0, 5e-3, nsteps=20)
    assert ok.cpu().numpy().all()
    y2, ok = e.step(y1, 5e-3, nsteps=20)
    failed, ratios = e.device_checks(y1, y2, steps=40)
    assert failed.cpu().numpy().tolist() == [0, 0, 0], ratios.cpu().numpy()
    r = ratios.cpu().numpy()
    assert np.all(r[:, :3] <= 1.0)
    # a planted booking error: one cell's amount moved without its transfer (material ledgers)
    bad = y2.cpu().numpy().copy()
    L = StateLayout.of(p)
    c = p.cell_off[1] + 3
    L.cell(bad, L.dn + 1)[c] += 1e-9 * abs(p.n0[1][c])
    failed, _ = e.device_checks(y1, bad, steps=40)
    f = failed.cpu().numpy().tolist()
    assert f[0] == 0 and f[2] == 0 and f[1] & 2 and f[1] & 4
    # heat that decreased since the previous row
    bad = y2.cpu().numpy().copy()
    L.cell(bad, L.mdiss)[p.cell_off[2]] = L.cell(y1.cpu().numpy(), L.mdiss)[p.cell_off[2]] - 1.0
    failed, _ = e.device_checks(y1, bad, steps=40)
    assert failed.cpu().numpy().tolist()[2] & 16
    e.close()


def _three_ion(worlds=2, cells=(12, 30), seed=5):
    return build(config1(worlds=worlds, cells=cells, species=("Na", "K", "Cl"),
                         inside_mM=(12.0, 140.0, 10.0), bath_mM=(145.0, 5.0, 110.0),
                         permeability=(1e-10, 5e-9, 1e-9), seed=seed))


def test_step_adjoint_matches_the_oracle_adjoint(torch):
    p = _three_ion()
    L = StateLayout.of(p)
    _, sym = oracle.world_symbolics(p)
    w = oracle.world_struct(p)
    y = initial_state(p)
    for _ in range(10):
        y, _ = oracle.step_ros2(p, y, 5e-3, sym, world=w)
    rng = np.random.default_rng(1)
    ybar = np.zeros_like(y)
    for s in range(p.S):
        L.cell(ybar, L.dn + s)[:] = rng.normal(size=p.ncell)
    L.edge(ybar, L.gate)[:] = rng.normal(size=p.nedge)
    yin_o, pam_o, cap_o, paj_o = oracle.ros2_vjp(p, y, 5e-3, sym, ybar, world=w)
    e = ensemble(p, "ros2")
    say("adjoint parity")
    yin, pam, cap, paj, valid = e.step_vjp(y, ybar, 5e-3)
    assert valid.cpu().numpy().all()
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


def test_step_adjoint_meets_central_differences_of_the_hip_step(torch):
    p = _three_ion(seed=9)
    L = StateLayout.of(p)
    e = ensemble(p, "ros2")
    y, _ = e.step(initial_state(p), 5e-3, nsteps=10)
    y = y.cpu().numpy()
    rng = np.random.default_rng(2)
    ybar = np.zeros_like(y)
    for s in range(p.S):
        L.cell(ybar, L.dn + s)[:] = rng.normal(size=p.ncell)
    L.edge(ybar, L.gate)[:] = rng.normal(size=p.nedge)
    yin, *_ = e.step_vjp(y, ybar, 5e-3)
    yin = yin.cpu().numpy()

    def dot(a, b):
        t = sum(np.dot(L.cell(a, L.dn + s), L.cell(b, L.dn + s)) for s in range(p.S))
        return t + np.dot(L.edge(a, L.gate), L.edge(b, L.gate))

    for trial in range(3):
        v = np.zeros_like(y)
        for s in range(p.S):
            L.cell(v, L.dn + s)[:] = rng.normal(size=p.ncell) * 1e-6
        L.edge(v, L.gate)[:] = rng.normal(size=p.nedge) * 1e-3
        h = 1e-3
        plus = e.step(y + h * v, 5e-3)[0].cpu().numpy()
        minus = e.step(y - h * v, 5e-3)[0].cpu().numpy()
        fd = (dot(ybar, plus) - dot(ybar, minus)) / (2 * h)
        assert abs(fd - dot(yin, v)) <= 1e-6 * abs(fd), (trial, fd, dot(yin, v))
    e.close()


def test_calibration_gradient_meets_central_differences(torch):
    """Loss over 40 ROS2 steps (4 sample rows), binomial checkpointing with 3 slots: the
    gradient in theta against central differences of the HIP loss (h = 1e-6)."""
    from cells_amd.adjoint import Calibration
    from cells_amd.model import potentials  # noqa: F401
    from tests.helpers import vmem as vm

    p = _three_ion(worlds=2, cells=(16, 25), seed=21)
    e = ensemble(p, "ros2")
    y0 = initial_state(p)
    rng = np.random.default_rng(4)
    hidden = rng.normal(scale=0.2, size=(6, p.ncell))
    cal = Calibration(e, y0, 5e-3, 40, 10, [np.zeros(p.ncell)] * 4)
    cal.set_theta(hidden)
    y = e.to_device(y0)
    targets = []
    for _ in range(4):
        y, _ = e.step(y, 5e-3, nsteps=10)
        targets.append(vm(p, y.cpu().numpy()))
    cal = Calibration(e, y0, 5e-3, 40, 10, targets)
    theta = np.zeros((6, p.ncell))
    say("calibration gradient")
    loss, grad = cal.loss_and_gradient(theta, snapshots=3)
    assert loss > 0
    for trial in range(4):
        d = rng.normal(size=theta.shape)
        h = 1e-6
        fd = (cal.loss(theta + h * d) - cal.loss(theta - h * d)) / (2 * h)
        an = float(np.sum(grad * d))
        assert abs(fd - an) <= 1e-6 * abs(fd), (trial, fd, an)
    e.close()


def test_recorded_steps_reproduce_the_step_and_its_adjoint(torch):
    """cells_step_record gives cells_step's state bit for bit, and the reverse of each recorded
    step equals cells_step_vjp's (which recomputes the forward internals) bit for bit."""
    p = _three_ion(worlds=3, cells=(20, 49), seed=8)
    e = ensemble(p, "ros2")
    e.reserve_tape(3)
    L = StateLayout.of(p)
    dt = 5e-3
    y = e.to_device(initial_state(p))
    rng = np.random.default_rng(2)
    ys, bars = [], []
    say("recorded steps")
    for j in range(3):
        ref, ok_ref = e.step(y, dt, nsteps=1)
        got, ok = e.step_record(y, dt, j)
        assert torch.equal(ref, got) and bool(ok.all()) and bool(ok_ref.all())
        ys.append(y)
        y = got
    for j in range(3):
        ybar = torch.as_tensor(rng.normal(size=L.total), device=y.device)
        a = e.step_vjp(ys[j], ybar, dt)
        b = e.step_vjp_recorded(j, ybar, dt)
        for u, v in zip(a, b):
            assert torch.equal(u, v)
    e.close()


def test_taped_calibration_gradient_equals_the_recomputing_one(torch):
    """Binomial schedule over 6-step recorded units vs the one-step recomputing schedule: the same
    loss (and the plain loss sweep's) and the same gradient, bit for bit."""
    from cells_amd.adjoint import Calibration
    from tests.helpers import vmem as vm

    p = _three_ion(worlds=2, cells=(16, 25), seed=21)
    e = ensemble(p, "ros2")
    y0 = initial_state(p)
    rng = np.random.default_rng(9)
    y = e.to_device(y0)
    targets = []
    for _ in range(5):
        y, _ = e.step(y, 5e-3, nsteps=8)
        targets.append(vm(p, y.cpu().numpy()) + rng.normal(scale=1e-4, size=p.ncell))
    cal = Calibration(e, y0, 5e-3, 40, 8, targets)
    theta = rng.normal(scale=0.05, size=(6, p.ncell))
    say("calibration: recomputing schedule")
    l0, g0 = cal.loss_and_gradient(theta, snapshots=3, tape=0)
    e.reserve_tape(6)
    say("calibration: taped schedule")
    l1, g1 = cal.loss_and_gradient(theta, snapshots=3, tape=6)
    assert cal.recorded_steps == 40
    l2 = cal.loss(theta)
    assert l0 == l1 == l2
    assert np.array_equal(g0, g1)
    e.close()


def test_autograd_function_gradient_meets_central_differences(torch):
    """torch.autograd over three steps of the Step function (cells_step forward, cells_step_vjp
    backward): d loss / d (pam, nak_cap, paj, y0) against central differences of the HIP loss,
    the loss a fixed weighting of the final amounts."""
    from cells_amd.adjoint import step_function

    p = _three_ion(worlds=2, cells=(16, 25), seed=31)
    e = ensemble(p, "ros2")
    L = StateLayout.of(p)
    dev, f64 = torch.device("cuda", 0), torch.float64
    rng = np.random.default_rng(12)
    wts = np.zeros(L.total)
    wts[:p.S * p.ncell] = rng.normal(size=p.S * p.ncell)  # the amounts' block
    wt = torch.as_tensor(wts, device=dev)
    step = step_function(e, 5e-3)
    y0 = torch.as_tensor(initial_state(p), device=dev)
    base = [torch.as_tensor(np.ascontiguousarray(a), device=dev) for a in (p.pam, p.nak_cap, p.paj)]

    def loss(y, pam, cap, paj):
        for _ in range(3):
            y = step(y, pam, cap, paj)
        return (wt * y).sum()

    args = [y0.clone().requires_grad_(True)] + [b.clone().requires_grad_(True) for b in base]
    say("autograd function")
    value = loss(*args)
    value.backward()
    grads = [a.grad.clone() for a in args]
    for trial in range(3):
        dirs = [torch.as_tensor(rng.normal(size=tuple(a.shape)), device=dev) for a in args]
        dirs[0][p.S * p.ncell:] = 0.0  # perturb only the amounts of the state
        scale = [torch.full_like(a, 1e-4) for a in args]  # fmol for amounts
        scale[1:] = [b.abs() for b in base]  # parameters: relative
        hs = 1e-6
        with torch.no_grad():
            plus = [a.detach() + hs * s * d for a, s, d in zip(args, scale, dirs)]
            minus = [a.detach() - hs * s * d for a, s, d in zip(args, scale, dirs)]
            fd = (loss(*plus) - loss(*minus)).item() / (2 * hs)
        an = sum(float((g * s * d).sum()) for g, s, d in zip(grads, scale, dirs))
        assert abs(fd - an) <= 1e-6 * abs(fd) + 1e-18, (trial, fd, an)
    e.close()


def test_fused_ledger_checks_refuse_only_the_corrupted_world(torch):
    """With the checks as the step's epilogue: a clean run is bit for bit the unchecked one with no
    bits; a world whose material ledger is corrupted refuses its step (state kept, ok = 0, the
    material bit set) while the others advance exactly as without checks."""
    p = build(config1(worlds=4, cells=(30, 120), seed=17))
    L = StateLayout.of(p)
    e = ensemble(p, "ros2")
    y0 = initial_state(p)
    ref, ok_ref = e.step(y0, 5e-3, nsteps=4)
    e.step_checks(True, material_atol=1e-12, steps_before=0)
    say("fused checks, clean")
    got, ok = e.step(y0, 5e-3, nsteps=4)
    assert torch.equal(ref, got) and bool(ok.all()) and bool(ok_ref.all())
    assert int(e.check_bits().abs().sum()) == 0
    # corrupt world 2's first cell's membrane transfer of K
    bad = y0.copy()
    i = int(p.cell_off[2])
    L.cell(bad, L.mtr + 1)[i] += 1e-3
    e.step_checks(True, material_atol=1e-12, steps_before=0)
    say("fused checks, planted ledger error")
    got, ok = e.step(bad, 5e-3, nsteps=4)
    bits = e.check_bits().cpu().numpy()
    okn = ok.cpu().numpy()
    assert okn.tolist() == [1, 1, 0, 1] and bits[2] & 2 and bits[[0, 1, 3]].sum() == 0
    g = got.cpu().numpy()
    r = ref.cpu().numpy()
    for w in (0, 1, 3):
        c0, c1 = p.cell_off[w], p.cell_off[w + 1]
        assert np.array_equal(L.cell(g, L.dn)[c0:c1], L.cell(r, L.dn)[c0:c1])
    c0, c1 = p.cell_off[2], p.cell_off[3]
    assert np.array_equal(L.cell(g, L.dn)[c0:c1], L.cell(bad, L.dn)[c0:c1])
    e.step_checks(False)
    e.close()


def test_resume_from_a_saved_state_is_bitwise(torch, tmp_path):
    """Ten batched steps equal four steps, the state written to disk (npy) and read back into a
    newly created ensemble of the same worlds, then six steps — bit for bit (the state vector is
    the whole checkpoint: plans are rebuilt identically from the worlds)."""
    p = build(config1(worlds=3, cells=(40, 150), seed=23))
    e = ensemble(p, "ros2")
    y0 = initial_state(p)
    full, ok = e.step(y0, 5e-3, nsteps=10)
    part, _ = e.step(y0, 5e-3, nsteps=4)
    path = tmp_path / "state.npy"
    np.save(path, part.cpu().numpy())
    e.close()
    say("resume from a saved state")
    e2 = ensemble(p, "ros2")
    back = np.load(path)
    rest, ok2 = e2.step(back, 5e-3, nsteps=6)
    assert bool(ok.all()) and bool(ok2.all())
    assert torch.equal(full.cpu(), rest.cpu())
    e2.close()
