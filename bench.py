"""Benchmark: cell-steps/s of the batched HIP tissue step on one MI355X.

Workload: the batched ensemble configuration (4096 worlds of U{256..4096} cells, Na/K/Cl/Ca,
ROS2 with the sparse-Cholesky stage solves, dt = 5 ms), synthetic seeded parameters. A step is one
ROS2 step of every world. The timed region is exactly --steps steps of the packed ensemble (one
captured HIP graph per step), with the state resident in HBM before it starts.

Beside the headline: the same step one world per launch ("plain", timed on the ensemble's first
--plain-worlds worlds), the batched/plain ratio, the dominant kernel's roofline figures from HIP
events (`roofline`; the other sparse kernel, factor or stage solve, on the same terms in
`roofline_others`), a parity check of the first worlds against the FP64 oracle, the oracle's own
throughput on this host (CPU baseline), and the N2, stationary, spatial and adjoint legs.
"""

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
PROFILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "ros2_step_4096_worlds.md")


def measured_traffic(kernel):
    """HBM bytes per dispatch of `kernel` from the committed rocprofv3 PMC summary (FETCH_SIZE and
    WRITE_SIZE scaled by a 1 GiB copy), or None."""
    try:
        for line in open(PROFILE):
            cells = [c.strip() for c in line.strip().strip("|").split("|")]
            if len(cells) == 6 and cells[0].split("<")[0] == kernel and cells[3].endswith("GB"):
                return float(cells[3].split()[0]) * 1e9
    except OSError:
        pass
    return None
BYTES_PER_CELL_STEP = 539.0  # one read + write of the state, one read of the parameters (4 ions)


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def vmem(p, y):
    """Each cell's membrane potential (V) in a state: bath + (q0 + F' sum_s z_s dN_s) / C."""
    from cells_amd.world import F, StateLayout

    L = StateLayout.of(p)
    q = p.q0 + 1e-15 * F * sum(p.valence[s] * L.cell(y, L.dn + s) for s in range(p.S))
    return p.config.bath_V + q / p.cap


def conc(p, y):
    """Concentrations [S, ncell] (mM) in a state."""
    from cells_amd.world import StateLayout

    L = StateLayout.of(p)
    v = p.vol0 + (L.cell(y, L.dvol) if p.water else 0.0)
    return np.stack([(p.n0[s] + L.cell(y, L.dn + s)) / v for s in range(p.S)])


def relerr(a, b):
    """Largest |a - b| / |b| (0 where both vanish)."""
    a, b = np.asarray(a, dtype=float), np.asarray(b, dtype=float)
    with np.errstate(divide="ignore", invalid="ignore"):
        e = np.where(np.abs(b) > 0, np.abs(a - b) / np.abs(b), np.abs(a - b))
    return float(np.max(e)) if e.size else 0.0


def host_cores():
    """(usable core ids, CPU model) of this host."""
    cores = sorted(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else list(range(os.cpu_count() or 1))
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return cores, model


def cpu_baseline(p, seconds, dt, max_threads=16):
    """The oracle (FP64 C restatement) on the host's cores, one world per thread (the oracle's
    calls release the interpreter), on the ensemble's first worlds, for about `seconds`."""
    from concurrent.futures import ThreadPoolExecutor

    import oracle
    from cells_amd.world import initial_state, worlds

    cores, model = host_cores()
    threads = max(1, min(len(cores), max_threads, p.W))
    subs = [worlds(p, k, k + 1) for k in range(threads)]
    prepared = []
    for sub in subs:
        _, sym = oracle.world_symbolics(sub)
        prepared.append((sub, sym, oracle.world_struct(sub)))
    deadline = [0.0]

    def run(i):
        sub, sym, w = prepared[i]
        y = initial_state(sub)
        steps = 0
        while True:
            y, ok = oracle.step_ros2(sub, y, dt, sym, world=w)
            steps += 1
            if time.perf_counter() >= deadline[0]:
                return steps

    t0 = time.perf_counter()
    deadline[0] = t0 + seconds
    with ThreadPoolExecutor(threads) as ex:
        counts = list(ex.map(run, range(threads)))
    el = time.perf_counter() - t0
    cells = sum(sub.ncell * c for (sub, _, _), c in zip(prepared, counts))
    return {
        "value": cells / el, "unit": "cell-steps/s", "cores": threads, "kind": "port",
        "sample": f"oracle ROS2 steps of the ensemble's first {threads} worlds "
                  f"({sum(s.ncell for s in subs)} cells), one world per thread on {threads} of the host's "
                  f"{len(cores)} usable cores, {sum(counts)} world-steps in {el:.1f} s",
        "host": {"nproc": os.cpu_count(), "usable_cores": len(cores), "core_list": cores,
                 "cpu_model": model, "threads": threads},
    }


def stationary_leg(worlds):
    """Newton-Krylov steady state of the precision configuration (64x64-cell worlds, Na/K/Cl/Ca,
    junction conductances lognormal sigma 0.3) from the initial state: FP32 Krylov under FP64
    refinement, then the all-FP64 path, timed; the difference of the two steady states."""
    import torch

    from cells_amd.ensemble import Ensemble
    from cells_amd.world import StateLayout, build, config4, initial_state

    p = build(config4(worlds=worlds))
    e = Ensemble(p, "ros2")
    y0 = e.to_device(initial_state(p))
    log(f"stationary leg: {p.W} worlds x {p.ncell // p.W} cells")
    e.stationary(y0, max_newton=1)  # workspace and first touch
    out = {}
    states = {}
    for name, fp32 in (("fp32_inner", True), ("fp64", False)):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        y, conv, step, iters, rep = e.stationary(y0, fp32_inner=fp32)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        states[name] = y.cpu().numpy()
        out[name] = {"seconds": dt, "solves_per_s": p.W / dt, "converged": int(conv.sum()),
                     "newton_iterations": rep["newton_iterations"], "krylov_steps": rep["krylov_steps"],
                     "worst_step": rep["worst_step"]}
        log(f"stationary leg ({name}): {dt:.3f} s, {out[name]}")
    L = StateLayout.of(p)
    a, b = states["fp32_inner"], states["fp64"]
    diff = max(float(np.max(np.abs((L.cell(a, L.dn + s) - L.cell(b, L.dn + s)) / p.vol0))) for s in range(p.S))
    gdiff = float(np.max(np.abs(L.edge(a, L.gate) - L.edge(b, L.gate))))
    e.close()
    return {"worlds": p.W, "cells_per_world": p.ncell // p.W, **out,
            "fp32_vs_fp64_max_mM": diff, "fp32_vs_fp64_max_gate": gdiff}


def copy_probe(mib):
    """Plain device copies of known size (the calibration of the HBM byte counters): the runtime's
    copy (hipMemcpy, `__amd_rocclr_copyBuffer`) and an ordinary elementwise kernel (torch's
    multiply by one), each reading and writing mib MiB per launch."""
    import torch

    n = mib * (1 << 20) // 8
    src = torch.ones(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    for _ in range(3):
        torch.mul(src, 1.0, out=dst)
    torch.cuda.synchronize()
    log(f"copy probe: {mib} MiB read + {mib} MiB written per launch (runtime copy, then elementwise)")
    del src, dst


def spatial_oracle_operator():
    """The oracle's operator of the spatial configuration's grid (its symbolic analysis takes about
    a minute on the host), built on a thread of its own while the device legs run."""
    import threading

    box = {}

    def build():
        import oracle
        from cells_amd.field import FieldConfig, build_field

        fp = build_field(FieldConfig())
        box["operator"] = oracle.FieldOperator(fp)

    t = threading.Thread(target=build, daemon=True)
    t.start()
    return t, box


def spatial_checks(fp, y0, y, phi, oracle_op):
    """The spatial run's ledger closure and its field potentials against the oracle: material per
    species and world against the ring supplies, relative to the amounts; every interior node's
    charge with its cell's, relative to the cells' charge; world 0's node potentials from the
    device field solve against the oracle's damped Newton at the same state."""
    import oracle
    from cells_amd.field import FieldLayout, amounts, charges

    L = FieldLayout.of(fp)
    ring = fp.ring()
    c0, n0 = amounts(fp, y0)
    c1, n1 = amounts(fp, y)
    material = 0.0
    for s in range(fp.S):
        for k in range(fp.W):
            cs = slice(int(fp.cells.cell_off[k]), int(fp.cells.cell_off[k + 1]))
            before = c0[s, cs].sum() + n0[s, k, ~ring].sum()
            after = c1[s, cs].sum() + n1[s, k, ~ring].sum()
            material = max(material, abs(after - before - L.world(y, L.sup + s)[k]) / before)
    cq, nq = charges(fp, y)
    neutral = 0.0
    for k in range(fp.W):
        cs = slice(int(fp.cells.cell_off[k]), int(fp.cells.cell_off[k + 1]))
        q = nq[k].copy()
        q[fp.cell_node[cs]] += cq[cs]
        neutral = max(neutral, float(np.max(np.abs(q[~ring])) / np.max(np.abs(cq[cs]))))
    out = {"material_closure_rel": material, "charge_neutrality_rel": neutral}
    if oracle_op is not None:
        cells, nodes, _, conv = oracle.field_start(fp, y, oracle_op, 0)
        scale = float(np.max(np.abs(nodes)))
        out["potential_error_vs_oracle"] = float(np.max(np.abs(phi[0] - nodes))) / scale
        out["potential_error_world"] = 0
        out["oracle_field_solve_converged"] = bool(conv)
    return out


def spatial_leg(steps, refresh, starts=("solve", "carried"), tune_plan=True, oracle_job=None):
    """The spatial configuration: 64 worlds of 2048 cells (a 45 x 46 lattice) in a 256 x 256
    extracellular grid with a held ring, Na/K/Cl, IMEX Euler dt = 1 ms with the implicit face drift
    through the nested-dissection operator, rebuilt every `refresh` steps; each step starts from
    the field solve at its state (chord steps on the operator's factor). Times `steps` steps (a
    multiple of refresh: the factor cost is amortised as it is in a run) and one step's kernels,
    then the same steps from the carried potentials; prices the operator solve against its
    algorithmic bytes (both sweeps read the factor once, 2 x 8 x nnz per world) and the factor
    against the FP64 matrix peak (sum of column counts squared per world)."""
    import numpy as np
    import torch

    from cells_amd.field import FieldConfig, FieldEnsemble, build_field, initial_field_state

    fp = build_field(FieldConfig())
    tune = None
    if tune_plan:
        from cells_amd.autotune import autotune_field

        tune = autotune_field(fp, steps=10, refresh_every=refresh)
        log(f"spatial leg: autotuned grid plan options {tune.best} ({tune.ms_per_step:.2f} ms/step over 10 steps)")
    out = {}
    for start in starts:
        e = FieldEnsemble(fp, refresh_every=refresh, start=start)
        plan = e.plan_stats()
        log(f"spatial leg ({start}): {fp.W} worlds x {fp.config.cells} cells in {fp.side}^2 grids, plan {plan}")
        y = e.to_device(initial_field_state(fp))
        kt = e.kernel_times(y, 1e-3, reps=3)
        y, ok = e.step(y, 1e-3, nsteps=refresh)  # warm: one operator refresh and its steps
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        y, ok = e.step(y, 1e-3, nsteps=steps)
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        ms = el / steps * 1e3
        log(f"spatial leg ({start}): {ms:.3f} ms/step")
        out[start] = (kt, ms, el, bool(ok.all()), plan)
        if start == "solve":
            phi, chord, _ = e.potentials(y, 1e-3)
            chord = chord.cpu().numpy()
            op = None
            if oracle_job is not None:
                log("spatial leg: waiting for the oracle's operator")
                oracle_job[0].join()
                op = oracle_job[1].get("operator")
            checks = spatial_checks(fp, initial_field_state(fp), y.cpu().numpy(), phi.cpu().numpy(), op)
            log(f"spatial leg checks: {checks}")
        e.close()
    kt, ms, el, accepted, plan = out[starts[0]]
    if starts[0] != "solve":
        chord = np.zeros(1)
        checks = None
    solve_ms = sum(v for k, v in kt if k.startswith("k_fd_solve") and "field" not in k)
    field_ms = sum(v for k, v in kt if k.startswith("k_fd_k") or k.startswith("k_fd_solve field"))
    factor_ms = sum(v for k, v in kt if k.startswith("k_fd_factor") or k == "k_fd_assemble")
    solve_bytes = 2 * 8 * plan["nnz"] * fp.W
    achieved = solve_bytes / (solve_ms * 1e-3) / 1e9
    tflops = plan["flops"] * fp.W / (factor_ms * 1e-3) / 1e12  # the reference's cost model
    log(f"spatial leg: operator solve {solve_ms:.3f} ms ({achieved:.0f} GB/s), field solve {field_ms:.3f} ms "
        f"({float(np.mean(chord)):.2f} chord steps), factor {factor_ms:.2f} ms per refresh")
    return {"worlds": fp.W, "cells_per_world": fp.config.cells, "grid": f"{fp.side}x{fp.side}",
            "steps": steps, "refresh_every": refresh, "autotune": tune.summary() if tune else None,
            "start_potentials": "field solve" if starts[0] == "solve" else "carried",
            "ms_per_step": ms, "cell_steps_per_s": fp.W * fp.config.cells * steps / el,
            "all_worlds_accepted": all(v[3] for v in out.values()), "plan": plan, "checks": checks,
            "kernel_ms": {k: round(v, 4) for k, v in kt},
            "field_solve": {"ms_per_step": field_ms, "chord_steps_mean": float(np.mean(chord)),
                            "chord_steps_max": int(np.max(chord))},
            "carried_start": ({"ms_per_step": out["carried"][1],
                               "cell_steps_per_s": fp.W * fp.config.cells * steps / out["carried"][2]}
                              if "carried" in out and starts[0] == "solve" else None),
            "solve": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                      "frac": achieved / HBM_PEAK_GBS, "ms_per_step": solve_ms},
            "factor": {"bound": "mfma", "achieved": tflops, "peak": 78.6, "unit": "TFLOP/s",
                       "frac": tflops / 78.6, "ms_per_refresh": factor_ms}}


def adjoint_leg(steps, snapshots, tape, fd_directions=2):
    """Calibration gradient on the adjoint configuration's worlds (256 worlds of 32x32 cells,
    Na/K/Cl, ROS2 dt = 5 ms, 6 log-parameters per cell, Vmem targets every 20 steps): one loss
    (forward sweep) and one loss-and-gradient with binomial checkpointing, timed."""
    import torch

    import threading

    from cells_amd.adjoint import Calibration
    from cells_amd.ensemble import Ensemble
    from cells_amd.world import build, config1, initial_state

    dt, every = 5e-3, 20
    p = build(config1(worlds=256, cells=(1024, 1024), species=("Na", "K", "Cl"),
                      inside_mM=(12.0, 140.0, 10.0), bath_mM=(145.0, 5.0, 110.0),
                      permeability=(1e-10, 5e-9, 1e-9), seed=3))
    e = Ensemble(p, "ros2")
    y0 = initial_state(p)
    rng = np.random.default_rng(7)
    hidden = rng.normal(scale=0.1, size=(6, p.ncell))
    samples = steps // every
    cal = Calibration(e, y0, dt, steps, every, [np.zeros(p.ncell)] * samples)
    cal.set_theta(hidden)
    y = e.to_device(y0)
    targets = []
    for _ in range(samples):
        y, _ = e.step(y, dt, nsteps=every)
        targets.append(vmem(p, y.cpu().numpy()))
    cal = Calibration(e, y0, dt, steps, every, targets)
    theta = np.zeros((6, p.ncell))
    tape_bytes = e.reserve_tape(tape)
    log(f"adjoint leg: {p.W} worlds x {p.ncell // p.W} cells, {steps} steps, {snapshots} snapshots, "
        f"tape of {tape} steps ({tape_bytes / 2**30:.2f} GiB)")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    loss = cal.loss(theta)  # the step graph was captured while making the targets
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t0
    cal.forward_steps = cal.recorded_steps = cal.reverse_steps = 0
    # peak device use while the gradient runs: the least free memory, sampled every 5 ms
    free0, total = torch.cuda.mem_get_info()
    least = [free0]
    stop = threading.Event()

    def sample():
        while not stop.is_set():
            least[0] = min(least[0], torch.cuda.mem_get_info()[0])
            stop.wait(0.005)

    sampler = threading.Thread(target=sample, daemon=True)
    sampler.start()
    t0 = time.perf_counter()
    loss2, grad = cal.loss_and_gradient(theta, snapshots=snapshots, tape=tape)
    torch.cuda.synchronize()
    t_grad = time.perf_counter() - t0
    stop.set()
    sampler.join()
    least[0] = min(least[0], torch.cuda.mem_get_info()[0])
    log(f"adjoint leg: forward {t_fwd:.3f} s, loss+gradient {t_grad:.3f} s "
        f"({cal.forward_steps} forward + {cal.recorded_steps} recorded + {cal.reverse_steps} reverse "
        f"steps), peak device use {(total - least[0]) / 2**30:.2f} GiB")
    # the gradient against central differences of the HIP loss along random directions
    fd = []
    for trial in range(fd_directions):
        d = rng.normal(size=theta.shape)
        h = 1e-6
        diff = (cal.loss(theta + h * d) - cal.loss(theta - h * d)) / (2 * h)
        an = float(np.sum(grad * d))
        fd.append(abs(diff - an) / abs(diff))
    if fd:
        log(f"adjoint leg: central differences along {len(fd)} directions, worst relative error {max(fd):.2e}")
    e.close()
    return {"worlds": p.W, "cells_per_world": p.ncell // p.W, "steps": steps,
            "snapshots": snapshots, "tape_steps": tape, "tape_bytes": tape_bytes,
            "grads_per_s": 1.0 / t_grad, "forward_s": t_fwd,
            "gradient_s": t_grad, "gradient_over_forward": t_grad / t_fwd,
            "forward_steps": cal.forward_steps, "recorded_steps": cal.recorded_steps,
            "reverse_steps": cal.reverse_steps, "loss": loss, "loss_from_gradient_sweep": loss2,
            "snapshot_bytes": snapshots * 8 * int(y.numel()), "peak_device_bytes": int(total - least[0]),
            "grad_norm": float(np.linalg.norm(grad)),
            "fd_check": {"directions": len(fd), "h": 1e-6, "max_rel_err": max(fd) if fd else None}}


def n2_leg(p, dt, warmup, steps, parity_worlds, ros2_ms, parity):
    """The reference's N2 (ARS(2,2,2) with implicit slope conductances) on the same worlds: its
    batched step time beside ROS2's, its kernels, and its trajectory against the oracle's N2 on the
    first worlds."""
    import torch

    from cells_amd.ensemble import Ensemble
    from cells_amd.world import head, initial_state, take_state

    log("N2 on the same worlds")
    e = Ensemble(p, "n2")
    y0 = e.to_device(initial_state(p))
    out = torch.empty_like(y0)
    ok = torch.empty(p.W, dtype=torch.int32, device=y0.device)
    y, _ = e.step(y0, dt, nsteps=warmup, out=out, ok=ok)
    yw = y.clone()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    y, ok = e.step(yw, dt, nsteps=steps, out=out, ok=ok)
    torch.cuda.synchronize()
    ms = 1e3 * (time.perf_counter() - t0) / steps
    final = y.cpu().numpy()
    accepted = bool(ok.cpu().numpy().all())
    kernels = {}
    for name, t in e.kernel_times(yw, dt, reps=2):
        kernels[name] = round(kernels.get(name, 0.0) + t, 4)
    e.close()
    log(f"N2: {ms:.3f} ms/step ({p.ncell * 1e3 / ms:.4g} cell-steps/s), all accepted {accepted}; kernels {kernels}")
    err = None
    if parity:
        import oracle  # the checker of the parity leg

        sub = head(p, parity_worlds)
        _, sym = oracle.world_symbolics(sub)
        ws = oracle.world_struct(sub)
        ref = initial_state(sub)
        for _ in range(warmup + steps):
            ref, okr = oracle.step_n2(sub, ref, dt, sym, world=ws)
        got = take_state(p, final, parity_worlds)
        err = max(relerr(vmem(sub, got), vmem(sub, ref)), relerr(conc(sub, got), conc(sub, ref)))
        log(f"N2 parity max relative error {err:.3e}")
    return {"ms_per_step": ms, "cell_steps_per_s": p.ncell * 1e3 / ms, "over_ros2_ms": ms / ros2_ms,
            "all_worlds_accepted": accepted, "kernel_ms": kernels, "parity_max_rel_err": err,
            "parity_worlds": parity_worlds if parity else None}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--worlds", type=int, default=4096)
    ap.add_argument("--plain-worlds", type=int, default=64)
    ap.add_argument("--plain-steps", type=int, default=2)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--parity-worlds", type=int, default=3)
    ap.add_argument("--skip", default="",
                    help="comma list of legs to skip: tune,plain,parity,cpu,adjoint,stationary,spatial,n2")
    ap.add_argument("--spatial-steps", type=int, default=100)
    ap.add_argument("--spatial-start", choices=("both", "solve", "carried"), default="both",
                    help="the spatial leg's step-start potentials: the field solve, the carried ones, or both")
    ap.add_argument("--spatial-only", action="store_true",
                    help="run only the spatial leg and print its JSON (for profiling it alone)")
    ap.add_argument("--adjoint-steps", type=int, default=2000, help="horizon of the adjoint leg")
    ap.add_argument("--adjoint-snapshots", type=int, default=45)
    ap.add_argument("--stationary-worlds", type=int, default=512)
    ap.add_argument("--adjoint-tape", type=int, default=45,
                    help="recorded steps per unit of the adjoint's binomial schedule (0: none)")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="also run a plain device copy of this many MiB (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "the benchmark runs on one GPU: independent worlds shard across "
                          "GPUs with no exchange, and no multi-GPU figure is measured"}))
        sys.exit(2)
    skip = set(filter(None, args.skip.split(",")))
    if args.spatial_only:
        if args.copy_probe:
            copy_probe(args.copy_probe)
        starts = ("solve", "carried") if args.spatial_start == "both" else (args.spatial_start,)
        skip = set(filter(None, args.skip.split(",")))
        print(json.dumps({"spatial": spatial_leg(args.spatial_steps, 50, starts, "tune" not in skip)}), flush=True)
        return
    oracle_job = spatial_oracle_operator() if "spatial" not in skip and "parity" not in skip else None

    import torch

    import oracle
    from cells_amd.ensemble import Ensemble
    from cells_amd.world import build, config1, head, initial_state, take_state

    dt = 5e-3
    log("building worlds")
    p = build(config1(worlds=args.worlds))
    tune = None
    if "tune" not in skip:
        from cells_amd.autotune import autotune_ensemble

        tune = autotune_ensemble(p, "ros2", dt, steps=3)
        log(f"autotuned plan options {tune.best} ({tune.ms_per_step:.2f} ms/step); candidates "
            + "; ".join(f"{c}: {m:.2f}" for c, m in tune.table))
    log(f"{p.W} worlds, {p.ncell} cells, {p.nedge} edges; binding")
    t0 = time.perf_counter()
    e = Ensemble(p, "ros2")
    stats = e.plan_stats()
    log(f"bound in {time.perf_counter() - t0:.1f} s; plans {stats}")
    y0 = e.to_device(initial_state(p))
    out = torch.empty_like(y0)
    ok = torch.empty(p.W, dtype=torch.int32, device=y0.device)
    y, _ = e.step(y0, dt, nsteps=args.warmup, out=out, ok=ok)  # also captures the step graph
    yw = y.clone()
    torch.cuda.synchronize()
    log("timing batched")
    t0 = time.perf_counter()
    y, ok = e.step(yw, dt, nsteps=args.steps, out=out, ok=ok)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_step = 1e3 * elapsed / args.steps
    cell_steps = p.ncell * args.steps / elapsed
    all_ok = bool(ok.cpu().numpy().all())
    final = y.cpu().numpy()
    log(f"batched: {ms_per_step:.3f} ms/step, {cell_steps:.4g} cell-steps/s, all accepted {all_ok}")
    # the same steps with the ledger checks as each step's epilogue
    e.step_checks(True, material_atol=1e-12, steps_before=args.warmup)
    e.step(yw, dt, nsteps=1, out=out, ok=ok)  # captures the checked graph
    torch.cuda.synchronize()
    e.step_checks(True, material_atol=1e-12, steps_before=args.warmup)
    t0 = time.perf_counter()
    yc, okc = e.step(yw, dt, nsteps=args.steps, out=out, ok=ok)
    torch.cuda.synchronize()
    checked_ms = 1e3 * (time.perf_counter() - t0) / args.steps
    checked_bits = int(e.check_bits().abs().sum())
    e.step_checks(False)
    log(f"with the ledger checks as epilogue: {checked_ms:.3f} ms/step, failure bits {checked_bits}")

    if args.copy_probe:
        copy_probe(args.copy_probe)

    log("kernel times")
    times = e.kernel_times(yw, dt, reps=3)
    per_kernel = {}
    for name, ms in times:
        per_kernel[name] = per_kernel.get(name, 0.0) + ms
    log(f"kernels (ms per step): {json.dumps({k: round(v, 4) for k, v in per_kernel.items()})}")
    dominant = max(per_kernel, key=per_kernel.get)

    def launches_of(name):
        return [ms for kname, ms in times if kname == name]

    def factor_roofline():
        # the factor's algorithmic bytes: L written once (8 nnz of the bound pattern) and the stage
        # matrix it assembles read once (diagonal, junction conductances, Jacobi scale: 8 (2 n + E));
        # its flops, sum_j c_j^2, per byte put it below the FP64 matrix ridge (78.6 TFLOP/s over
        # 8 TB/s), so the roofline that bounds it is HBM's
        fact_ms = float(np.mean(launches_of("k_factor")))
        algo = 8 * stats["nnz"] + 8 * (2 * p.ncell + p.nedge)
        intensity = stats["flops"] / algo
        ridge = FP64_PEAK_TFLOPS * 1e12 / (HBM_PEAK_GBS * 1e9)
        if intensity < ridge:
            achieved = algo / (fact_ms * 1e-3) / 1e9
            r = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                 "frac": achieved / HBM_PEAK_GBS}
        else:
            achieved = stats["flops"] / (fact_ms * 1e-3) / 1e12
            r = {"bound": "mfma", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                 "frac": achieved / FP64_PEAK_TFLOPS}
        r.update({"traffic": measured_traffic("k_factor_mf"),
                  "traffic_source": "profiles/ros2_step_4096_worlds.md (rocprofv3 PMC, copy-calibrated)",
                  "kernel": "k_factor_mf", "launch_ms": fact_ms,
                  "algorithmic": f"8 nnz(L) + 8 (2 ncell + nedge) = {algo:.4g} bytes and sum over worlds "
                                 f"of sum_j c_j^2 = {stats['flops']:.4g} flops per launch",
                  "intensity_flop_per_byte": intensity, "ridge_flop_per_byte": ridge,
                  "fp64_tflops": stats["flops"] / (fact_ms * 1e-3) / 1e12})
        return r

    def solve_roofline():
        # one stage solve: the factor read by both sweeps (2 x 8 x nnz(L)) plus the right-hand side's
        # amounts (S per cell), the Jacobi scale, the result and the rank-ordered vector (read and
        # written by each sweep): 8 (S + 6) bytes per cell
        avg_ms = float(np.mean(launches_of("k_solve_stage")))
        algo = 2 * 8 * stats["nnz"] + 8 * (p.S + 6) * p.ncell
        achieved = algo / (avg_ms * 1e-3) / 1e9
        return {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic("k_solve_stage"),
                "traffic_source": "profiles/ros2_step_4096_worlds.md (rocprofv3 PMC, copy-calibrated)",
                "kernel": "k_solve_stage", "launch_ms": avg_ms,
                "algorithmic": f"2 x 8 x nnz(L) + 8 (S + 6) ncell = {algo:.4g} bytes per launch"}

    if dominant == "k_factor":
        roofline = factor_roofline()
    elif dominant == "k_solve_stage":
        roofline = solve_roofline()
    else:
        avg_ms = float(np.mean(launches_of(dominant)))
        achieved = BYTES_PER_CELL_STEP * p.ncell / (avg_ms * 1e-3) / 1e9
        roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(dominant), "kernel": dominant,
                    "launch_ms": avg_ms}
    # the sparse kernels that are not the dominant one, on the same terms
    roofline_others = [f() for name, f in (("k_factor", factor_roofline), ("k_solve_stage", solve_roofline))
                       if name != dominant and name in per_kernel]
    step_gbs = BYTES_PER_CELL_STEP * p.ncell / (ms_per_step * 1e-3) / 1e9

    n2 = None
    if "n2" not in skip:
        n2 = n2_leg(p, dt, args.warmup, args.steps, args.parity_worlds, ms_per_step, "parity" not in skip)

    plain = sub_batched = None
    if "plain" not in skip:
        k = min(args.plain_worlds, p.W)
        sub = head(p, k)
        log(f"timing plain (one world per launch) on the first {k} worlds ({sub.ncell} cells)")
        es = Ensemble(sub, "ros2")
        ys = es.to_device(initial_state(sub))
        ysw, _ = es.step(ys, dt, nsteps=1, batched=False)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        es.step(ysw, dt, nsteps=args.plain_steps, batched=False)
        torch.cuda.synchronize()
        plain = sub.ncell * args.plain_steps / (time.perf_counter() - t0)
        es.step(ysw, dt, nsteps=1, batched=True)  # captures the sub-ensemble's step graph
        torch.cuda.synchronize()
        nb = max(args.plain_steps, 10)
        t0 = time.perf_counter()
        es.step(ysw, dt, nsteps=nb, batched=True)
        torch.cuda.synchronize()
        sub_batched = sub.ncell * nb / (time.perf_counter() - t0)
        log(f"plain: {plain:.4g} cell-steps/s; batched on the same worlds {sub_batched:.4g}")
        es.close()

    parity = None
    if "parity" not in skip:
        log("parity against the oracle")
        k = args.parity_worlds
        sub = head(p, k)
        _, sym = oracle.world_symbolics(sub)
        ws = oracle.world_struct(sub)
        ref = initial_state(sub)
        for _ in range(args.warmup + args.steps):
            ref, okr = oracle.step_ros2(sub, ref, dt, sym, world=ws)
        got = take_state(p, final, k)
        parity = max(relerr(vmem(sub, got), vmem(sub, ref)), relerr(conc(sub, got), conc(sub, ref)))
        log(f"parity max relative error {parity:.3e}")

    sta = None
    if "stationary" not in skip:
        sta = stationary_leg(args.stationary_worlds)

    spatial = None
    if "spatial" not in skip:
        spatial = spatial_leg(args.spatial_steps, 50, tune_plan="tune" not in skip, oracle_job=oracle_job)

    adj = None
    if "adjoint" not in skip:
        adj = adjoint_leg(args.adjoint_steps, args.adjoint_snapshots, args.adjoint_tape)

    cpu = None
    if "cpu" not in skip:
        log("cpu baseline")
        cpu = cpu_baseline(p, args.cpu_seconds, dt)

    line = {
        "metric": "cell-steps/s batched vs plain HIP at <=1e-9 rel err (FP64 oracle); adjoint grads/s",
        "value": cell_steps,
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded lattices and parameters)",
        "config": {"workload": "batched ensemble: 4096 worlds, cells U{256..4096}, Na/K/Cl/Ca, "
                               "ROS2 dt=5e-3 with sparse-Cholesky stage solves",
                   "worlds": p.W, "cells": p.ncell, "edges": p.nedge, "dt": dt,
                   "method": "ros2"},
        "plain_value": plain,
        "plain_worlds": min(args.plain_worlds, p.W) if plain else None,
        "ms_per_step_with_checks": checked_ms,
        "check_failures": checked_bits,
        "batched_over_plain": (sub_batched / plain) if plain else None,
        "batched_over_plain_note": ("like for like: the same first plain_worlds worlds, captured graph "
                                    "replays against one world per launch; batched_full_over_plain is the "
                                    "whole ensemble's value over the same plain figure") if plain else None,
        "batched_full_over_plain": (cell_steps / plain) if plain else None,
        "all_worlds_accepted": all_ok,
        "parity_max_rel_err": parity,
        "parity_worlds": args.parity_worlds if parity is not None else None,
        "step_effective_GBs": step_gbs,
        "kernel_ms": {kk: round(v, 4) for kk, v in per_kernel.items()},
        "plans": stats,
        "roofline": roofline,
        "roofline_others": roofline_others,
        "adjoint": adj,
        "n2": n2,
        "stationary": sta,
        "spatial": spatial,
        "autotune": tune.summary() if tune is not None else None,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    e.close()


if __name__ == "__main__":
    main()
