"""Device ensembles of worlds: the host side of the batched step, over the C-ABI.

PyTorch provides device memory and streams only; every step runs in libcells_hip.so.

Counterparts in the reference (src/cells/ of Eidosoma/Cells):
- `Ensemble.step(..., batched=True)`  execution/vector_windows.py:348-413, `vmap(checked_step)`
  inside one scan, one launch sequence for every world (here a captured HIP graph per step);
- `Ensemble.step(..., batched=False)` execution/batch.py:321-357, `lax.map` over worlds, one
  world per launch;
- `Ensemble.rates`                    models/surface_pump_v2.py:370-450, explicit_rates at the
  state's own potentials;
- `Ensemble.factor` / `.solve`        numerics/sparse_cholesky.py:474-514 on the stage pattern.
"""

from __future__ import annotations

import ctypes

import numpy as np

from cells_amd import lib as _lib
from cells_amd.world import Packed, StateLayout

_i64p = ctypes.POINTER(ctypes.c_int64)
_i32p = ctypes.POINTER(ctypes.c_int32)


def _torch():
    import torch

    if not torch.cuda.is_available():
        raise RuntimeError("no GPU: the HIP step has no CPU fallback")
    return torch


def control(membrane=1.0, junction=1.0, pump=1.0, calcium_pump=1.0, water=1.0):
    """Held availabilities in [0, 1] (definitions/surface_pump_v2_spec.py:118-123)."""
    return _lib.ControlStruct(membrane, junction, pump, calcium_pump, water)


class Ensemble:
    """Packed worlds bound on one GPU: parameters, per-shape plans and the stepping workspace."""

    def __init__(self, packed: Packed, method: str = "euler", device: int = 0,
                 factor: str = "supernodal"):
        torch = _torch()
        self.torch = torch
        self.packed = packed
        self.layout = StateLayout.of(packed)
        self.method = {"euler": _lib.METHOD_EULER, "ros2": _lib.METHOD_ROS2, "n2": _lib.METHOD_N2,
                       "ars222": _lib.METHOD_ARS222, "rk4": _lib.METHOD_RK4}[method]
        flag = {"supernodal": 0, "pairwise": _lib.FACTOR_PAIRWISE}[factor]
        self.device = torch.device("cuda", device)
        L = _lib.load()
        self._world = _lib.world_struct(packed)
        handle = ctypes.c_void_p()
        torch.cuda.init()
        _lib.check(L.cells_ensemble_create(ctypes.byref(self._world), self.method | flag, device,
                                           ctypes.byref(handle)))
        self._h = handle
        self.tape_slots = 0
        self.state_size = int(L.cells_state_size(handle))
        if self.state_size != self.layout.total:
            raise RuntimeError("state layout disagrees with the library's")

    def close(self):
        if getattr(self, "_h", None):
            _lib.load().cells_ensemble_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _stream(self):
        return ctypes.c_void_p(self.torch.cuda.current_stream(self.device).cuda_stream)

    def to_device(self, y):
        t = self.torch
        if isinstance(y, np.ndarray):
            y = t.from_numpy(np.ascontiguousarray(y, dtype=np.float64))
        return y.to(self.device, dtype=t.float64).contiguous()

    def step(self, y, dt, nsteps=1, ctl=None, active=None, batched=True, out=None, ok=None):
        """nsteps checked steps of every world; returns (new state, ok per world) on the device."""
        t = self.torch
        y = self.to_device(y)
        if y.numel() != self.state_size:
            raise ValueError("state size disagrees with the worlds")
        out = t.empty_like(y) if out is None else out
        ok = t.empty(self.packed.W, dtype=t.int32, device=self.device) if ok is None else ok
        act = None
        if active is not None:
            act = t.as_tensor(np.asarray(active, dtype=np.int32)).to(self.device)
        ctl = ctl or control()
        _lib.check(_lib.load().cells_step(
            self._h, self._stream(), ctypes.c_void_p(y.data_ptr()), ctypes.c_void_p(out.data_ptr()),
            float(dt), int(nsteps), ctypes.byref(ctl),
            None if act is None else ctypes.c_void_p(act.data_ptr()),
            ctypes.c_void_p(ok.data_ptr()), 1 if batched else 0))
        return out, ok

    def rates(self, y, ctl=None):
        t = self.torch
        y = self.to_device(y)
        r = t.zeros_like(y)
        valid = t.empty(self.packed.W, dtype=t.int32, device=self.device)
        ctl = ctl or control()
        _lib.check(_lib.load().cells_rates(self._h, self._stream(), ctypes.c_void_p(y.data_ptr()),
                                           ctypes.byref(ctl), ctypes.c_void_p(r.data_ptr()),
                                           ctypes.c_void_p(valid.data_ptr())))
        return r, valid

    @property
    def factor_size(self):
        return int(_lib.load().cells_factor_size(self._h))

    def positions(self, w):
        """(diag_pos [n] by node, edge_pos [E] by world edge, factor offset) of world w."""
        p = self.packed
        n = int(p.cell_off[w + 1] - p.cell_off[w])
        E = int(p.edge_off[w + 1] - p.edge_off[w])
        d = np.empty(n, np.int64)
        e = np.empty(max(E, 1), np.int64)
        off = ctypes.c_int64()
        _lib.check(_lib.load().cells_world_positions(self._h, w, d.ctypes.data_as(_i64p),
                                                     e.ctypes.data_as(_i64p), ctypes.byref(off)))
        return d, e[:E], off.value

    def factor_positions(self, w, rows, cols):
        """Positions in world w's factor slice of the entries (rows, cols) (world-local nodes), -1
        where the plan's pattern has none."""
        r = np.ascontiguousarray(rows, dtype=np.int32)
        c = np.ascontiguousarray(cols, dtype=np.int32)
        out = np.empty(max(r.size, 1), np.int64)
        _lib.check(_lib.load().cells_world_factor_positions(
            self._h, w, r.ctypes.data_as(_i32p), c.ctypes.data_as(_i32p), r.size, out.ctypes.data_as(_i64p)))
        return out[:r.size]

    def factor(self, values):
        v = self.to_device(values).clone()
        _lib.check(_lib.load().cells_factor(self._h, self._stream(), ctypes.c_void_p(v.data_ptr())))
        return v

    def solve(self, factor_values, rhs):
        f = self.to_device(factor_values)
        b = self.to_device(rhs)
        x = self.torch.empty_like(b)
        _lib.check(_lib.load().cells_solve(self._h, self._stream(), ctypes.c_void_p(f.data_ptr()),
                                           ctypes.c_void_p(b.data_ptr()),
                                           ctypes.c_void_p(x.data_ptr())))
        return x

    def energy_rates(self, y, ctl=None):
        """The readouts' energy rates of every world at y: [W, len(lib.ENERGY_RATES)] on the device
        (pJ/s; the electrode current in A)."""
        t = self.torch
        y = self.to_device(y)
        out = t.empty((self.packed.W, len(_lib.ENERGY_RATES)), dtype=t.float64, device=self.device)
        c = ctl or control()
        if self.packed.water:
            raise ValueError("the energy readouts cover worlds without water")
        _lib.check(_lib.load().cells_energy_rates(self._h, self._stream(), ctypes.c_void_p(y.data_ptr()),
                                                  ctypes.byref(c), ctypes.c_void_p(out.data_ptr())))
        return out

    def device_checks(self, previous, y, steps, material_atol=1e-12):
        """(failed bits [W], ratios [W, len(lib.LEDGER_RATIOS)]) of a state against its predecessor:
        the bits of lib.CHECKS, and per ledger of lib.LEDGER_RATIOS its largest |residual| / bound."""
        t = self.torch
        a, b = self.to_device(previous), self.to_device(y)
        failed = t.empty(self.packed.W, dtype=t.int32, device=self.device)
        ratios = t.empty((self.packed.W, len(_lib.LEDGER_RATIOS)), dtype=t.float64, device=self.device)
        _lib.check(_lib.load().cells_device_checks(
            self._h, self._stream(), ctypes.c_void_p(a.data_ptr()), ctypes.c_void_p(b.data_ptr()),
            float(steps), float(material_atol), ctypes.c_void_p(failed.data_ptr()),
            ctypes.c_void_p(ratios.data_ptr())))
        return failed, ratios

    def step_vjp(self, y, ybar, dt, ctl=None, pam_bar=None, cap_bar=None, paj_bar=None, noalloc=False):
        """Reverse of one ROS2 step at y: (cotangent on y, pam_bar, cap_bar, paj_bar, valid).
        Parameter cotangents are added to the given tensors (zeros when None). noalloc: through
        cells_step_vjp_noalloc (the library's fixed workspace instead of one it allocates)."""
        t = self.torch
        p = self.packed
        y, ybar = self.to_device(y), self.to_device(ybar)
        yin = t.zeros_like(y)
        dev, f64 = self.device, t.float64
        pam_bar = t.zeros((p.S, p.ncell), dtype=f64, device=dev) if pam_bar is None else pam_bar
        cap_bar = t.zeros(p.ncell, dtype=f64, device=dev) if cap_bar is None else cap_bar
        paj_bar = t.zeros((p.S, max(p.nedge, 1)), dtype=f64, device=dev) if paj_bar is None else paj_bar
        valid = t.empty(p.W, dtype=t.int32, device=dev)
        ctl = ctl or control()
        v = lambda x: ctypes.c_void_p(x.data_ptr())  # noqa: E731
        entry = _lib.load().cells_step_vjp_noalloc if noalloc else _lib.load().cells_step_vjp
        _lib.check(entry(self._h, self._stream(), v(y), v(ybar), float(dt), ctypes.byref(ctl), v(yin),
                         v(pam_bar), v(cap_bar), v(paj_bar), v(valid)))
        return yin, pam_bar, cap_bar, paj_bar, valid

    def step_checks(self, enabled=True, material_atol=1e-12, steps_before=0):
        """Run the ledger checks as each step's epilogue (a failing world refuses the step)."""
        _lib.check(_lib.load().cells_step_checks(self._h, 1 if enabled else 0, float(material_atol),
                                                 int(steps_before)))

    def step_bounds(self):
        """dict(explicit, gate, calcium): the definition-time rate bounds (1/s) that every step
        checks its dt against (dt x explicit <= 1, dt x gate <= 0.125, dt x calcium <= 1)."""
        out = np.zeros(3)
        _lib.check(_lib.load().cells_step_bounds(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return dict(explicit=float(out[0]), gate=float(out[1]), calcium=float(out[2]))

    def stage_tolerance(self, atol=3e-25, rtol=2e-12):
        """The stage solves' residual bound (defaults: the reference's charge tolerances)."""
        _lib.check(_lib.load().cells_set_stage_tolerance(self._h, float(atol), float(rtol)))

    def check_bits(self):
        """Failure bits per world of the last step call (device int32 [W]): the ledger checks'
        bits when they are on, and the refusal reasons (lib.REFUSE_*) always."""
        t = self.torch
        bits = t.empty(self.packed.W, dtype=t.int32, device=self.device)
        _lib.check(_lib.load().cells_step_check_bits(self._h, self._stream(), ctypes.c_void_p(bits.data_ptr())))
        return bits

    def reserve_tape(self, slots):
        """Reserve `slots` adjoint step records on the device (0 frees them); returns bytes."""
        nbytes = ctypes.c_int64()
        _lib.check(_lib.load().cells_tape_reserve(self._h, int(slots), ctypes.byref(nbytes)))
        self.tape_slots = int(slots)
        return nbytes.value

    def step_record(self, y, dt, slot, ctl=None):
        """One batched ROS2 step (as step(..., nsteps=1)) that records itself in tape `slot`:
        (y_next, ok [W])."""
        t = self.torch
        y = self.to_device(y)
        out = t.empty_like(y)
        ok = t.empty(self.packed.W, dtype=t.int32, device=self.device)
        ctl = ctl or control()
        _lib.check(_lib.load().cells_step_record(
            self._h, self._stream(), ctypes.c_void_p(y.data_ptr()), ctypes.c_void_p(out.data_ptr()),
            float(dt), ctypes.byref(ctl), int(slot), ctypes.c_void_p(ok.data_ptr())))
        return out, ok

    def step_vjp_recorded(self, slot, ybar, dt, ctl=None, pam_bar=None, cap_bar=None, paj_bar=None):
        """step_vjp of the step recorded in tape `slot`: (cotangent on y, pam_bar, cap_bar, paj_bar,
        valid)."""
        t = self.torch
        p = self.packed
        ybar = self.to_device(ybar)
        yin = t.zeros_like(ybar)
        dev, f64 = self.device, t.float64
        pam_bar = t.zeros((p.S, p.ncell), dtype=f64, device=dev) if pam_bar is None else pam_bar
        cap_bar = t.zeros(p.ncell, dtype=f64, device=dev) if cap_bar is None else cap_bar
        paj_bar = t.zeros((p.S, max(p.nedge, 1)), dtype=f64, device=dev) if paj_bar is None else paj_bar
        valid = t.empty(p.W, dtype=t.int32, device=dev)
        ctl = ctl or control()
        v = lambda x: ctypes.c_void_p(x.data_ptr())  # noqa: E731
        _lib.check(_lib.load().cells_step_vjp_recorded(self._h, self._stream(), int(slot), v(ybar),
                                                       float(dt), ctypes.byref(ctl), v(yin),
                                                       v(pam_bar), v(cap_bar), v(paj_bar), v(valid)))
        return yin, pam_bar, cap_bar, paj_bar, valid

    def stationary(self, y0, ctl=None, tolerance=1e-10, max_newton=60, max_cycles=20, linear_rtol=1e-8,
                   fp32_inner=True):
        """Steady state of every world (damped Newton, GMRES with block-Jacobi, FP32 Krylov under
        FP64 refinement when fp32_inner): (y, converged [W], step [W], iterations [W], report)."""
        t = self.torch
        y0 = self.to_device(y0)
        out = t.empty_like(y0)
        W = self.packed.W
        conv = t.empty(W, dtype=t.int32, device=self.device)
        step = t.empty(W, dtype=t.float64, device=self.device)
        iters = t.empty(W, dtype=t.int32, device=self.device)
        opt = _lib.StationaryOptions(float(tolerance), int(max_newton), int(max_cycles), float(linear_rtol),
                                     1 if fp32_inner else 0)
        rep = _lib.StationaryReport()
        ctl = ctl or control()
        v = lambda x: ctypes.c_void_p(x.data_ptr())  # noqa: E731
        _lib.check(_lib.load().cells_stationary(self._h, self._stream(), v(y0), v(out), ctypes.byref(ctl),
                                                ctypes.byref(opt), v(conv), v(step), v(iters),
                                                ctypes.byref(rep)))
        report = dict(converged=rep.converged, newton_iterations=rep.newton_iterations,
                      krylov_steps=rep.krylov_steps, worst_step=rep.worst_step)
        return out, conv, step, iters, report

    def stationary_residual(self, y, ctl=None):
        """(f(y) in coordinate units [S ncell + nedge], valid [W])."""
        t = self.torch
        y = self.to_device(y)
        p = self.packed
        r = t.empty(p.S * p.ncell + p.nedge, dtype=t.float64, device=self.device)
        valid = t.empty(p.W, dtype=t.int32, device=self.device)
        ctl = ctl or control()
        v = lambda x: ctypes.c_void_p(x.data_ptr())  # noqa: E731
        _lib.check(_lib.load().cells_stationary_residual(self._h, self._stream(), v(y), ctypes.byref(ctl),
                                                         v(r), v(valid)))
        return r, valid

    def stationary_jvp(self, y, direction, ctl=None):
        """J(y) direction (coordinate units) from the forward-tangent kernels."""
        t = self.torch
        y = self.to_device(y)
        d = self.to_device(direction)
        out = t.empty_like(d)
        ctl = ctl or control()
        v = lambda x: ctypes.c_void_p(x.data_ptr())  # noqa: E731
        _lib.check(_lib.load().cells_stationary_jvp(self._h, self._stream(), v(y), ctypes.byref(ctl), v(d),
                                                    v(out)))
        return out

    def vmem_loss(self, y, target, weight, loss, ybar=None):
        """loss[w] += weight sum_i (V_i - target_i)^2; ybar (if given) += its cotangent."""
        v = lambda x: ctypes.c_void_p(x.data_ptr())  # noqa: E731
        _lib.check(_lib.load().cells_vmem_loss(self._h, self._stream(), v(self.to_device(y)),
                                               v(target), float(weight), v(loss),
                                               None if ybar is None else v(ybar)))

    def set_parameters(self, pam=None, nak_cap=None, paj=None):
        """Replace the learnable parameters on the device (tensors on this device)."""
        v = lambda x: None if x is None else ctypes.c_void_p(self.to_device(x).data_ptr())  # noqa: E731
        keep = [self.to_device(x) if x is not None else None for x in (pam, nak_cap, paj)]
        _lib.check(_lib.load().cells_set_parameters(self._h, self._stream(), *(
            None if x is None else ctypes.c_void_p(x.data_ptr()) for x in keep)))
        self.torch.cuda.current_stream(self.device).synchronize()

    def plan_stats(self):
        """dict(nnz, flops, max_levels, shapes) of the bound factor plans."""
        nnz, lev, shapes = ctypes.c_int64(), ctypes.c_int32(), ctypes.c_int32()
        flops = ctypes.c_double()
        _lib.check(_lib.load().cells_plan_stats(self._h, ctypes.byref(nnz), ctypes.byref(flops),
                                                ctypes.byref(lev), ctypes.byref(shapes)))
        return dict(nnz=nnz.value, flops=flops.value, max_levels=lev.value, shapes=shapes.value)

    def kernel_times(self, y, dt, reps=3, ctl=None):
        """[(kernel name, mean ms)] of one step's launches, in launch order (HIP events)."""
        y = self.to_device(y)
        cap = 64
        ms = np.zeros(cap)
        names = ctypes.create_string_buffer(32 * cap)
        count = ctypes.c_int32()
        ctl = ctl or control()
        _lib.check(_lib.load().cells_kernel_times(
            self._h, self._stream(), ctypes.c_void_p(y.data_ptr()), float(dt), ctypes.byref(ctl),
            int(reps), cap, ms.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), names,
            ctypes.byref(count)))
        raw = names.raw
        return [(raw[32 * k:32 * k + 32].split(b"\0")[0].decode(), float(ms[k]))
                for k in range(count.value)]
