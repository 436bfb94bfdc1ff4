"""GPU checks of the stationary analysis (Newton–Krylov) through the C-ABI.

- Its residual is the step's rates in the reference's coordinates (amount rates over the cell
  volume in mM/s, gate rates), against the oracle's rates at the same state (1e-12 relative).
- Its forward-tangent Jacobian products against central differences of the residual (1e-6).
- Newton converges for every world (correction below 1e-10 mM), the oracle's rates at the result
  are zero to a millionth of those at the start, and the FP32-Krylov path (FP64 refinement)
  lands within 1e-9 mM of the all-FP64 path.
"""

import sys

import numpy as np
import pytest

import oracle
from cells_amd.world import StateLayout, build, config4, initial_state

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


def coordinates(p, r):
    """The oracle's rates r (state layout) in the stationary coordinates."""
    L = StateLayout.of(p)
    amounts = np.concatenate([L.cell(r, L.dn + s) / p.vol0 for s in range(p.S)])
    return np.concatenate([amounts, L.edge(r, L.gate)])


def moved(p, y, v, h):
    """y with amounts moved by h v (mM, times the volume) and gates by h v."""
    L = StateLayout.of(p)
    out = y.copy()
    for s in range(p.S):
        L.cell(out, L.dn + s)[:] += h * v[s * p.ncell:(s + 1) * p.ncell] * p.vol0
    L.edge(out, L.gate)[:] += h * v[p.S * p.ncell:]
    return out


def small(seed=6, worlds=3, cells=(30, 200)):
    return build(config4(worlds=worlds, cells=cells, seed=seed))


def stepped(e, p, steps=20):
    y = e.to_device(initial_state(p))
    y, ok = e.step(y, 5e-3, nsteps=steps)
    assert bool(ok.all())
    return y.cpu().numpy()


def test_residual_is_the_oracle_rates_in_coordinates(torch):
    from cells_amd.ensemble import Ensemble

    p = small()
    e = Ensemble(p, "ros2")
    y = stepped(e, p)
    say("stationary residual")
    r, valid = e.stationary_residual(y)
    assert bool(valid.all())
    ref, ok = oracle.rates(p, y)
    assert ok.all()
    want = coordinates(p, ref)
    got = r.cpu().numpy()
    na = p.S * p.ncell
    for sl in (slice(0, na), slice(na, None)):
        assert np.max(np.abs(got[sl] - want[sl])) <= 1e-12 * np.max(np.abs(want[sl]))
    e.close()


def test_tangent_products_meet_central_differences(torch):
    """Amount rows and gate rows separately. The gate law depends on |V_a - V_b|, which has a
    kink at zero: the central difference (its potential step is about 0.2 uV) is compared only on
    the junctions with |V_a - V_b| >= 0.1 mV, where it is a derivative."""
    from cells_amd.ensemble import Ensemble

    p = small(seed=7)
    e = Ensemble(p, "ros2")
    y = stepped(e, p, steps=200)
    L = StateLayout.of(p)
    from tests.helpers import vmem

    v_m = vmem(p, y)
    dv = np.abs(v_m[p.snd] - v_m[p.rcv])
    rng = np.random.default_rng(3)
    na = p.S * p.ncell
    say(f"stationary tangent (min |dV| over junctions {dv.min():.3e} V)")
    h = 1e-6
    for trial in range(3):
        v = rng.normal(size=na + p.nedge)
        v[:na] *= 1e-2
        v[na:] *= 1e-3
        jv = e.stationary_jvp(y, v).cpu().numpy()
        fp, _ = e.stationary_residual(moved(p, y, v, h))
        fm, _ = e.stationary_residual(moved(p, y, v, -h))
        fd = (fp.cpu().numpy() - fm.cpu().numpy()) / (2 * h)
        smooth = np.concatenate([np.ones(na, bool), dv >= 1e-4])
        assert smooth[na:].sum() >= p.nedge // 4
        for name, sl in (("amounts", slice(0, na)), ("gates", slice(na, None))):
            keep = smooth[sl]
            err = np.abs(jv[sl] - fd[sl]) * keep
            worst = int(np.argmax(err))
            msg = (trial, name, float(np.linalg.norm(err) / np.linalg.norm(fd[sl])), worst,
                   float(jv[sl][worst]), float(fd[sl][worst]))
            assert np.linalg.norm(err) <= 1e-5 * np.linalg.norm(fd[sl] * keep), msg
    e.close()


def test_newton_converges_to_a_zero_of_the_oracle_rates(torch):
    from cells_amd.ensemble import Ensemble

    p = small(seed=8, worlds=4, cells=(64, 256))
    e = Ensemble(p, "ros2")
    y0 = initial_state(p)
    say("stationary newton (fp32 inner)")
    y32, conv, step, iters, rep = e.stationary(y0, fp32_inner=True)
    say(f"report {rep}, iterations {iters.cpu().numpy().tolist()}")
    assert bool(conv.all()), (rep, step.cpu().numpy())
    assert float(step.max()) <= 1e-10
    ys = y32.cpu().numpy()
    r0 = coordinates(p, oracle.rates(p, y0)[0])
    r1, ok = oracle.rates(p, ys)
    assert ok.all()
    r1 = coordinates(p, r1)
    na = p.S * p.ncell
    for sl in (slice(0, na), slice(na, None)):
        assert np.max(np.abs(r1[sl])) <= 1e-6 * np.max(np.abs(r0[sl])), np.max(np.abs(r1[sl]))
    say("stationary newton (all fp64)")
    y64, conv64, _, _, rep64 = e.stationary(y0, fp32_inner=False)
    assert bool(conv64.all()), rep64
    L = StateLayout.of(p)
    y64 = y64.cpu().numpy()
    for s in range(p.S):
        d = (L.cell(ys, L.dn + s) - L.cell(y64, L.dn + s)) / p.vol0
        assert np.max(np.abs(d)) <= 1e-9
    assert np.max(np.abs(L.edge(ys, L.gate) - L.edge(y64, L.gate))) <= 1e-9
    e.close()
