"""Calibration gradients through the ROS2 time loop: the adjoint step under binomial checkpointing.

The reference has no adjoint and no calibration loop (it differentiates only forward,
analysis/stationary.py:505-517); this is the build's own, for the adjoint-calibration
configuration: per-cell log-parameters theta [6, ncell] =
  (log P_Na, log P_K, log P_Cl, log Na/K pump capacity, log membrane leak scale,
   log gap-junction scale),
acting as
  pam[s, i]  = pam0[s, i] exp(theta[s, i] + theta[4, i])          (s = Na, K, Cl)
  nak[i]     = nak0[i] exp(theta[3, i])
  paj[s, e]  = paj0[s, e] exp((theta[5, a(e)] + theta[5, b(e)]) / 2)
The "leak scale" multiplies every membrane permeability of a cell (a redundant direction of the
three log permeabilities, declared so). The loss is the mean squared Vmem misfit over sample
rows: L = 1/(K n) sum_k sum_i (V_i(t_k) - target_k,i)^2.

Checkpointing: Griewank's binomial schedule (Griewank and Walther 2000, ACM TOMS 26:19, "revolve")
over units of `tape` steps: with c free snapshot slots and l units to reverse, the first split keeps
the right part within beta(c-1, r) = C(c-1+r, r) units, r the least repetition count with
beta(c, r) >= l. Snapshots are device copies of the state. A unit is reversed by stepping it
forward once more while recording every step (cells_step_record: the factor, slopes, stage
increments and stage state of each step in a device tape slot), then running the reverse steps
from the records, last to first, with no forward recomputation inside them. With tape = 0 a unit
is one step and each reverse step recomputes its forward internals (cells_step_vjp). The loss is
accumulated on the sweep that first reaches each sample row, so the gradient needs no separate
loss pass.

"""

from __future__ import annotations

import math

import numpy as np

from cells_amd.ensemble import control
from cells_amd.world import Packed


def beta(c: int, r: int) -> int:
    return math.comb(c + r, r)


def split(length: int, free: int) -> int:
    """Length of the right part of [a, a + length) kept after the first advance (binomial)."""
    r = 0
    while beta(free, r) < length:
        r += 1
    return max(1, min(length - 1, beta(free - 1, r)))


def theta_parameters(p: Packed, base, theta):
    """(pam [S, n], nak_cap [n], paj [S, E]) of theta [6, n] on base = (pam0, nak0, paj0)."""
    pam0, nak0, paj0 = base
    pam = pam0 * np.exp(theta[:3] + theta[4][None, :])
    nak = nak0 * np.exp(theta[3])
    paj = paj0 * np.exp(0.5 * (theta[5][p.snd] + theta[5][p.rcv]))[None, :]
    return pam, nak, paj


def theta_cotangent(p: Packed, params, bars):
    """theta_bar [6, n] from the parameters' cotangents (the chain rule of theta_parameters)."""
    pam, nak, paj = params
    pam_bar, cap_bar, paj_bar = bars
    tb = np.zeros((6, p.ncell))
    g = pam_bar * pam
    tb[:3] = g
    tb[4] = g.sum(axis=0)
    tb[3] = cap_bar * nak
    h = 0.5 * (paj_bar[:, :p.nedge] * paj).sum(axis=0)
    np.add.at(tb[5], p.snd, h)
    np.add.at(tb[5], p.rcv, h)
    return tb


class Calibration:
    """Loss and gradient of the Vmem misfit over `steps` ROS2 steps of an ensemble."""

    def __init__(self, ensemble, y0, dt, steps, sample_every, targets, ctl=None):
        self.e = ensemble
        self.t = ensemble.torch
        self.p = ensemble.packed
        if self.p.S != 3 or self.p.water or self.p.calcium:
            raise ValueError("calibration covers 3-ion worlds without water")
        self.y0 = ensemble.to_device(y0)
        self.dt = float(dt)
        self.steps = int(steps)
        self.every = int(sample_every)
        self.samples = list(range(self.every, self.steps + 1, self.every))
        self.targets = {k: self.t.as_tensor(np.asarray(targets[i]), dtype=self.t.float64,
                                            device=ensemble.device)
                        for i, k in enumerate(self.samples)}
        self.weight = 1.0 / (len(self.samples) * self.p.ncell)
        self.ctl = ctl or control()
        self.base = (self.p.pam.copy(), self.p.nak_cap.copy(), self.p.paj.copy())
        self.forward_steps = 0
        self.recorded_steps = 0
        self.reverse_steps = 0

    def set_theta(self, theta):
        params = theta_parameters(self.p, self.base, theta)
        self.e.set_parameters(*params)
        return params

    def _advance(self, y, n):
        if n == 0:
            return y.clone()
        out, ok = self.e.step(y, self.dt, nsteps=n, ctl=self.ctl)
        if not bool(ok.cpu().numpy().all()):
            raise FloatingPointError("a world refused a step of the calibration run")
        self.forward_steps += n
        return out

    def loss(self, theta):
        """The loss alone (one forward sweep)."""
        self.set_theta(theta)
        t = self.t
        acc = t.zeros(self.p.W, dtype=t.float64, device=self.e.device)
        y, last = self.y0, 0
        for k in self.samples:
            y = self._advance(y, k - last)
            last = k
            self.e.vmem_loss(y, self.targets[k], self.weight, acc)
        return float(acc.sum().cpu())

    def loss_and_gradient(self, theta, snapshots=45, tape=None):
        """(loss, theta_bar [6, n]) with the binomial schedule over `snapshots` free slots, in
        units of `tape` recorded steps (None: the ensemble's reserved tape; 0: no tape)."""
        t = self.t
        params = self.set_theta(theta)
        T = self.e.tape_slots if tape is None else int(tape)
        if T > self.e.tape_slots:
            raise ValueError(f"tape of {T} steps exceeds the {self.e.tape_slots} reserved slots")
        W = self.p.W
        S, n, E = self.p.S, self.p.ncell, max(self.p.nedge, 1)
        dev, f64 = self.e.device, t.float64
        self.pam_bar = t.zeros((S, n), dtype=f64, device=dev)
        self.cap_bar = t.zeros(n, dtype=f64, device=dev)
        self.paj_bar = t.zeros((S, E), dtype=f64, device=dev)
        self.ybar = t.zeros_like(self.y0)
        self.scratch = t.zeros(W, dtype=f64, device=dev)
        self.acc = t.zeros(W, dtype=f64, device=dev)
        self.frontier = 0
        self.unit = max(T, 1)
        self.taped = T > 0
        units = -(-self.steps // self.unit)
        store = {0: self.y0}
        self._reverse(store, 0, units, snapshots)
        bars = (self.pam_bar.cpu().numpy(), self.cap_bar.cpu().numpy(), self.paj_bar.cpu().numpy())
        return float(self.acc.sum().cpu()), theta_cotangent(self.p, params, bars)

    def _pos(self, u):
        return min(u * self.unit, self.steps)

    def _advance_to(self, y, a, b):
        """The state at step b from the state y at step a; adds the loss of every sample row in
        (frontier, b] on the way."""
        pos = a
        for k in self.samples:
            if max(a, self.frontier) < k <= b:
                y = self._advance(y, k - pos)
                pos = k
                self.e.vmem_loss(y, self.targets[k], self.weight, self.acc)
        y = self._advance(y, b - pos)
        self.frontier = max(self.frontier, b)
        return y

    def _reverse_unit(self, y_a, u):
        """ybar <- reverse of the steps of unit u, from the state y_a at its start."""
        a, b = self._pos(u), self._pos(u + 1)
        if not self.taped:
            self._adjoint_step(y_a, a)
            return
        at = {}
        y = y_a
        for j in range(b - a):
            y = self._record_step(y, j)
            k = a + j + 1
            if k in self.targets:
                at[k] = y
                if k > self.frontier:
                    self.e.vmem_loss(y, self.targets[k], self.weight, self.acc)
        self.frontier = max(self.frontier, b)
        for j in range(b - a - 1, -1, -1):
            k = a + j + 1
            if k in at:
                self.e.vmem_loss(at[k], self.targets[k], self.weight, self.scratch, self.ybar)
            self._reverse_recorded(j, a + j)

    def _record_step(self, y, slot):
        y, ok = self.e.step_record(y, self.dt, slot, ctl=self.ctl)
        if not bool(ok.cpu().numpy().all()):
            raise FloatingPointError("a world refused a step of the calibration run")
        self.recorded_steps += 1
        return y

    def _reverse_recorded(self, slot, a):
        """ybar <- reverse of step a -> a+1 from its record in tape `slot`."""
        self.ybar = self.e.step_vjp_recorded(slot, self.ybar, self.dt, ctl=self.ctl,
                                             pam_bar=self.pam_bar, cap_bar=self.cap_bar,
                                             paj_bar=self.paj_bar)[0]
        self.reverse_steps += 1

    def _adjoint_step(self, y_a, a):
        """ybar <- reverse of step a -> a+1 at y_a (forward internals recomputed), after adding
        the loss cotangent at a+1."""
        b = a + 1
        if b in self.targets:
            y_b = self._advance(y_a, 1)
            if b > self.frontier:
                self.e.vmem_loss(y_b, self.targets[b], self.weight, self.acc)
            self.e.vmem_loss(y_b, self.targets[b], self.weight, self.scratch, self.ybar)
        self.frontier = max(self.frontier, b)
        yin, _, _, _, valid = self.e.step_vjp(y_a, self.ybar, self.dt, ctl=self.ctl,
                                             pam_bar=self.pam_bar, cap_bar=self.cap_bar,
                                             paj_bar=self.paj_bar)
        if not bool(valid.cpu().numpy().all()):
            raise FloatingPointError("a world refused a step of the reverse sweep")
        self.ybar = yin
        self.reverse_steps += 1

    def _reverse(self, store, a, b, free):
        """Reverse units [a, b) from store[pos(a)] with `free` snapshot slots."""
        length = b - a
        if length == 1:
            self._reverse_unit(store[self._pos(a)], a)
            return
        if free <= 0:
            for j in range(b - 1, a - 1, -1):
                self._reverse_unit(self._advance_to(store[self._pos(a)], self._pos(a), self._pos(j)), j)
            return
        right = split(length, free)
        m = b - right
        store[self._pos(m)] = self._advance_to(store[self._pos(a)], self._pos(a), self._pos(m))
        self._reverse(store, m, b, free - 1)
        del store[self._pos(m)]
        self._reverse(store, a, m, free)


def step_function(ensemble, dt, ctl=None):
    """One batched ROS2 step as a torch.autograd.Function of (y, pam [S, n], nak_cap [n],
    paj [S, E]): forward replaces the ensemble's learnable parameters and steps through
    cells_step; backward calls cells_step_vjp at the saved input. Cotangents flow through the
    amounts and gates (the fields the step's rates read); ledger fields pass theirs through."""
    torch = ensemble.torch
    ctl = ctl or control()

    class Step(torch.autograd.Function):
        @staticmethod
        def forward(fctx, y, pam, nak_cap, paj):
            ensemble.set_parameters(pam.detach(), nak_cap.detach(), paj.detach())
            out, ok = ensemble.step(y.detach(), dt, nsteps=1, ctl=ctl)
            if not bool(ok.all()):
                raise FloatingPointError("a world refused the step")
            fctx.save_for_backward(y.detach(), pam.detach(), nak_cap.detach(), paj.detach())
            return out

        @staticmethod
        def backward(fctx, ybar):
            y, pam, nak_cap, paj = fctx.saved_tensors
            ensemble.set_parameters(pam, nak_cap, paj)  # this step's parameters
            yin, pam_bar, cap_bar, paj_bar, valid = ensemble.step_vjp(y, ybar.contiguous(), dt, ctl=ctl)
            if not bool(valid.all()):
                raise FloatingPointError("a world refused the step in the reverse sweep")
            return yin, pam_bar, cap_bar, paj_bar[:, :paj.shape[1]]

    return Step.apply
