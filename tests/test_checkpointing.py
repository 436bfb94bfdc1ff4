"""The binomial checkpointing schedule of cells_amd.adjoint on the CPU (no device calls).

A stand-in ensemble whose "state" is the step index checks that the reverse sweep visits every
step exactly once, last to first, each from the right state, within the snapshot budget, and
that the forward recomputation stays within the binomial bound (r sweeps, r the least repetition
count with C(c + r, r) >= steps; Griewank and Walther 2000).
"""

import math

import numpy as np
import pytest

from cells_amd import adjoint


class Probe(adjoint.Calibration):
    """State = step index; targets at every `every`-th step; tape of `unit` recorded steps
    (unit 0: every reverse step recomputes its forward internals)."""

    def __init__(self, steps, unit=0, every=7):  # no device
        self.forward_steps = 0
        self.recorded_steps = 0
        self.reverse_steps = 0
        self.visited = []
        self.max_live = 0
        self.steps = steps
        self.unit = max(unit, 1)
        self.taped = unit > 0
        self.samples = list(range(every, steps + 1, every))
        self.targets = {k: None for k in self.samples}
        self.frontier = 0
        self.loss_rows = []
        self.cotangent_rows = []
        self.slots = {}
        self.e = self
        self.weight, self.acc, self.scratch, self.ybar = 1.0, None, None, 0

    # the ensemble's calls, on integer states
    def vmem_loss(self, y, target, weight, loss, ybar=None):
        (self.loss_rows if ybar is None else self.cotangent_rows).append(y)

    def _advance(self, y, n):
        self.forward_steps += n
        return y + n

    def _adjoint_step(self, y_a, a):
        assert y_a == a
        b = a + 1
        if b in self.targets:
            if b > self.frontier:
                self.loss_rows.append(b)
            self.cotangent_rows.append(b)
        self.frontier = max(self.frontier, b)
        self.visited.append(a)

    def _record_step(self, y, slot):
        assert 0 <= slot < self.unit
        self.slots[slot] = y
        self.recorded_steps += 1
        return y + 1

    def _reverse_recorded(self, slot, a):
        assert self.slots[slot] == a
        self.visited.append(a)


def run(steps, snaps, unit=0):
    pr = Probe(steps, unit)
    store = {0: 0}

    class Watch(dict):
        def __setitem__(s, k, v):
            super().__setitem__(k, v)
            pr.max_live = max(pr.max_live, len(s) - 1)

    store = Watch(store)
    units = -(-steps // pr.unit)
    pr._reverse(store, 0, units, snaps)
    return pr


@pytest.mark.parametrize("steps,snaps", [(1, 3), (7, 2), (100, 5), (2000, 45), (500, 0), (64, 64)])
def test_every_step_is_reversed_once_in_order(steps, snaps):
    pr = run(steps, snaps)
    assert pr.visited == list(range(steps - 1, -1, -1))
    assert pr.max_live <= snaps


@pytest.mark.parametrize("steps,snaps,unit", [(1, 3, 4), (7, 2, 3), (100, 5, 8), (2000, 45, 45),
                                              (2000, 45, 20), (500, 0, 16), (64, 64, 64)])
def test_taped_units_reverse_every_step_once_in_order(steps, snaps, unit):
    pr = run(steps, snaps, unit)
    assert pr.visited == list(range(steps - 1, -1, -1))
    assert pr.max_live <= snaps
    # every step is recorded exactly once, just before its reversal
    assert pr.recorded_steps == steps


@pytest.mark.parametrize("steps,snaps,unit", [(200, 45, 0), (200, 45, 45), (2000, 45, 45),
                                              (1000, 3, 10), (30, 0, 0)])
def test_loss_rows_are_added_once_and_every_cotangent_once(steps, snaps, unit):
    """The loss of each sample row is accumulated exactly once (on the sweep that first reaches
    it) and its cotangent is seeded exactly once, in reverse order."""
    pr = run(steps, snaps, unit)
    assert sorted(pr.loss_rows) == pr.samples
    assert pr.cotangent_rows == pr.samples[::-1]


def test_a_tape_of_45_steps_meets_the_one_sweep_bound():
    """2000 steps, 45 snapshots, 45-step tape: one plain forward sweep over all but the last unit
    (no recomputation beyond it), one recorded sweep, one reverse sweep."""
    pr = run(2000, 45, 45)
    assert pr.forward_steps <= 2000 - 1
    assert pr.recorded_steps == 2000


@pytest.mark.parametrize("steps,snaps", [(100, 5), (2000, 45), (1000, 10)])
def test_recomputation_within_the_binomial_bound(steps, snaps):
    r = 0
    while math.comb(snaps + r, r) < steps:
        r += 1
    pr = run(steps, snaps)
    assert pr.forward_steps <= r * steps


def test_enough_snapshots_means_no_recomputation_beyond_one_sweep():
    pr = run(300, 300)
    assert pr.forward_steps <= 300


def test_theta_chain_rule_matches_finite_differences():
    """theta_cotangent is the adjoint of theta_parameters (checked on a linear functional)."""
    from cells_amd.world import build, config1

    p = build(config1(worlds=2, cells=(6, 9), species=("Na", "K", "Cl"),
                      inside_mM=(12.0, 140.0, 10.0), bath_mM=(145.0, 5.0, 110.0),
                      permeability=(1e-10, 5e-9, 1e-9)))
    base = (p.pam.copy(), p.nak_cap.copy(), p.paj.copy())
    rng = np.random.default_rng(0)
    theta = rng.normal(scale=0.1, size=(6, p.ncell))
    w = [rng.normal(size=a.shape) for a in base]

    def f(th):
        return sum(float(np.sum(wi * x)) for wi, x in zip(w, adjoint.theta_parameters(p, base, th)))

    params = adjoint.theta_parameters(p, base, theta)
    tb = adjoint.theta_cotangent(p, params, (w[0], w[1], w[2]))
    d = rng.normal(size=theta.shape)
    h = 1e-6
    fd = (f(theta + h * d) - f(theta - h * d)) / (2 * h)
    assert abs(fd - float(np.sum(tb * d))) <= 1e-7 * abs(fd)
