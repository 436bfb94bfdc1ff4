"""The oracle's adjoint of the ROS2 step against central finite differences (CPU).

The reference has no adjoint; the check the configuration names is central differences of the
forward step (h relative to each direction's scale). Directions are small enough that the step
stays in its linear regime: amounts by ~1e-6 fmol (a few uV of membrane potential), gates by 1e-3,
parameters by their own magnitude times h.
"""

import dataclasses

import numpy as np
import pytest

import oracle
from cells_amd.world import StateLayout, build, config1, initial_state


def world(cells, seed=5):
    return build(config1(worlds=2, cells=cells, species=("Na", "K", "Cl"),
                         inside_mM=(12.0, 140.0, 10.0), bath_mM=(145.0, 5.0, 110.0),
                         permeability=(1e-10, 5e-9, 1e-9), seed=seed))


@pytest.fixture(scope="module")
def setup():
    p = world((12, 30))
    L = StateLayout.of(p)
    _, sym = oracle.world_symbolics(p)
    w = oracle.world_struct(p)
    y = initial_state(p)
    for _ in range(10):
        y, ok = oracle.step_ros2(p, y, 5e-3, sym, world=w)
        assert ok.all()
    rng = np.random.default_rng(0)
    ybar = np.zeros_like(y)
    for s in range(p.S):
        L.cell(ybar, L.dn + s)[:] = rng.normal(size=p.ncell)
    L.edge(ybar, L.gate)[:] = rng.normal(size=p.nedge)
    return p, L, sym, w, y, ybar, rng


def dot(L, p, a, b):
    t = sum(np.dot(L.cell(a, L.dn + s), L.cell(b, L.dn + s)) for s in range(p.S))
    return t + np.dot(L.edge(a, L.gate), L.edge(b, L.gate))


def test_state_cotangent_meets_central_differences(setup):
    p, L, sym, w, y, ybar, rng = setup
    yin, _, _, _ = oracle.ros2_vjp(p, y, 5e-3, sym, ybar, world=w)
    for trial in range(4):
        v = np.zeros_like(y)
        for s in range(p.S):
            L.cell(v, L.dn + s)[:] = rng.normal(size=p.ncell) * 1e-6
        L.edge(v, L.gate)[:] = rng.normal(size=p.nedge) * 1e-3
        h = 1e-3
        plus, _ = oracle.step_ros2(p, y + h * v, 5e-3, sym, world=w)
        minus, _ = oracle.step_ros2(p, y - h * v, 5e-3, sym, world=w)
        fd = (dot(L, p, ybar, plus) - dot(L, p, ybar, minus)) / (2 * h)
        an = dot(L, p, yin, v)
        assert abs(fd - an) <= 1e-6 * abs(fd), (trial, fd, an)


@pytest.mark.parametrize("name", ["pam", "nak_cap", "paj"])
def test_parameter_cotangents_meet_central_differences(setup, name):
    p, L, sym, w, y, ybar, rng = setup
    _, pam_bar, cap_bar, paj_bar = oracle.ros2_vjp(p, y, 5e-3, sym, ybar, world=w)
    bar = {"pam": pam_bar, "nak_cap": cap_bar, "paj": paj_bar}[name]
    base = getattr(p, name)
    d = rng.normal(size=base.shape) * base
    h = 1e-4
    plus = dataclasses.replace(p, **{name: base + h * d})
    minus = dataclasses.replace(p, **{name: base - h * d})
    a, _ = oracle.step_ros2(plus, y, 5e-3, sym)
    b, _ = oracle.step_ros2(minus, y, 5e-3, sym)
    fd = (dot(L, p, ybar, a) - dot(L, p, ybar, b)) / (2 * h)
    an = float(np.sum(bar * d))
    assert abs(fd - an) <= 1e-6 * abs(fd), (fd, an)


def test_a_planted_sign_error_in_the_cotangent_misses_the_differences(setup):
    """Flipping the adjoint's correction sign (a plausible slip) is visible to this check."""
    p, L, sym, w, y, ybar, rng = setup
    yin, _, _, _ = oracle.ros2_vjp(p, y, 5e-3, sym, ybar, world=w)
    v = np.zeros_like(y)
    for s in range(p.S):
        L.cell(v, L.dn + s)[:] = rng.normal(size=p.ncell) * 1e-6
    h = 1e-3
    plus, _ = oracle.step_ros2(p, y + h * v, 5e-3, sym, world=w)
    minus, _ = oracle.step_ros2(p, y - h * v, 5e-3, sym, world=w)
    fd = (dot(L, p, ybar, plus) - dot(L, p, ybar, minus)) / (2 * h)
    # the identity part alone (dropping everything the step adds) is far off
    ident = dot(L, p, ybar, v)
    assert abs(fd - ident) > 1e-3 * abs(fd)
    assert abs(fd - dot(L, p, yin, v)) <= 1e-6 * abs(fd)
