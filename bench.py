#!/usr/bin/env python3
"""Chorin projection steps/s on one MI355X: plain FP64 path vs the mixed
FP32/FP64 path, at 500k P1 pressure DOFs (about 2M P2 velocity DOFs per
component), every linear solve to a 1e-10 relative residual.

Workload: seeded channel 5.4 x 1 with a hole (r = 0.12 at (0.5, 0.49)),
Re = 100 with a parabolic inflow of peak 1.5, convection reassembled every
step into the fixed CSR pattern, a seeded 1e-3 perturbation of the
inflow-initialised field.  The requested tau = 5e-4 violates the explicit
scheme's diffusion bound (tau nu lambda_max(M^-1 K) <= 2) at this mesh size:
the library's time-step guard (clamp policy) runs it at 1 / lambda_max(M^-1
nu K); both values are reported.

Prints one JSON line.  Progress goes to stderr."""

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_workload(nodes, seed, gpu_assembly=False):
    import numpy as np
    from nsfem_amd import fem, mesher
    t0 = time.perf_counter()
    m = mesher.channel_mesh(nodes, seed=seed)
    t1 = time.perf_counter()
    peak, re = 1.5, 100.0
    diameter = 2 * m["radius"]
    nu = (2.0 / 3.0) * peak * diameter / re
    cond = fem.DfgConditions(viscosity=nu, peakVelocity=peak, channelHeight=m["height"], profileScale=4.0)
    ctx = fem.ChorinContext(mesher.to_triangle_mesh(m), cond, gpu_assembly=gpu_assembly)
    t2 = time.perf_counter()
    log(f"mesh {t1 - t0:.1f}s (V={len(m['xy'])}, E={len(m['triangles'])}), context {t2 - t1:.1f}s "
        f"({'GPU' if gpu_assembly else 'host'} assembly, N={ctx.numVelocityNodes})")
    return m, cond, ctx


def initial_velocity(ctx, cond, seed):
    """Dirichlet values plus a seeded uniform(-1, 1) * 1e-3 * peak
    perturbation on the interior velocity nodes (both components)."""
    import numpy as np
    N = ctx.numVelocityNodes
    rng = np.random.default_rng(seed)
    u = rng.uniform(-1, 1, 2 * N) * 1e-3 * cond.peakVelocity
    ids, vals = ctx.dirichlet("vx")
    idy, valy = ctx.dirichlet("vy")
    u[ids] = vals
    u[idy + N] = valy
    return u


def run_path(name, ctx, cfg, tau, u0, warmup, steps, autotune, make=None):
    import numpy as np
    from nsfem_amd import _lib, fem
    ch = make(cfg) if make else fem.Chorin(ctx, cfg)
    ch.set_time_step(tau)
    tau_info = ch.time_step_info()
    t0 = time.perf_counter()
    shape = ch.autotune(30) if autotune else ch.shape()
    tune_s = time.perf_counter() - t0 if autotune else 0.0
    ch.upload_velocity(u0)
    log(f"  {name}: autotuned in {tune_s:.1f} s; first step")
    first = ch.step()
    u1, p1 = ch.download()
    warm = [ch.step() for _ in range(warmup)]
    log(f"  {name}: warm-up done; pressure iterations of the first and warm-up steps: " +
        " ".join(str(int(s["itersPressure"])) for s in [first] + warm))
    _lib.call("nsfem_synchronize")
    stats, walls = [], []
    t0 = time.perf_counter()
    for k in range(steps):
        t = time.perf_counter()
        stats.append(ch.step())
        walls.append(1e3 * (time.perf_counter() - t))
        if k % 5 == 4:
            log(f"  {name}: step {k + 1}/{steps}, {walls[-1]:.1f} ms")
    _lib.call("nsfem_synchronize")
    dt = time.perf_counter() - t0
    med = float(np.median(walls))
    for k, (w, s) in enumerate(zip(walls, stats)):
        if w > 4 * med:
            log(f"  {name}: step {k + 1} took {w:.1f} ms (median {med:.1f}): " +
                ", ".join(f"{key} {s[key]:.4g}" for key in sorted(s)))
    u, p = ch.download()
    if not (np.isfinite(u).all() and np.isfinite(p).all()):
        raise RuntimeError(f"{name}: non-finite state")
    agg = {k: float(np.mean([s[k] for s in stats])) for k in stats[0]}
    res = {"steps_per_s": steps / dt, "ms_per_step": 1e3 * dt / steps,
           "iters": {"tentative": agg["itersTent"], "pressure": agg["itersPressure"], "final": agg["itersFinal"]},
           "outer": {"tentative": agg["outerTent"], "pressure": agg["outerPressure"], "final": agg["outerFinal"]},
           "max_relres": max(max(s["relTent"], s["relPressure"], s["relFinal"]) for s in stats),
           "time_step": tau_info,
           "captured_steps": int(sum(s["capturedStep"] for s in stats)),
           "step_wall_ms": {"min": min(walls), "median": float(np.median(walls)), "max": max(walls)},
           "step_graph": ch.step_graph_info(),
           "ms": {"convection": agg["msConvection"], "tentative": agg["msTentative"], "pressure": agg["msPressure"],
                  "final": agg["msFinal"]},
           "shape": {"velocity_lanes": shape[0], "velocity_rows_per_block": shape[1], "pressure_lanes": shape[2],
                     "pressure_rows_per_block": shape[3], "velocity_sweep_lanes": shape[4],
                     "velocity_sweep_rows_per_block": shape[5], "pressure_sweep_lanes": shape[6],
                     "pressure_sweep_rows_per_block": shape[7],
                     "velocity_direction_update_rho_rows": list(shape[8:11]),
                     "pressure_direction_update_rho_rows": list(shape[11:14]),
                     "velocity_pressure_direction_lanes": list(shape[14:16]),
                     "velocity_pressure_fused_reductions": list(shape[16:18]),
                     "autotuned": bool(autotune), "autotune_s": tune_s}}
    log(f"{name}: {res['steps_per_s']:.2f} steps/s ({res['ms_per_step']:.2f} ms/step), iters {res['iters']}, "
        f"outer {res['outer']}, max relres {res['max_relres']:.2e}, phase ms {res['ms']}")
    return ch, res, (u1, p1), (u, p)


def ensemble_chorin(args, ctx, cfg, B, scales):
    from nsfem_amd import fem
    ch = fem.Chorin(ctx, cfg, instances=B, inflow_scale=scales)
    apply_amg_smoother(args, ch, cfg)
    if args.velocity_precond == "fsai":
        ch.set_solver_option("preconditioner", 1, "velocity")
    return ch


def run_ensemble(args, ctx, cond, u0, target, precision="mixed", tau=None):
    """B flows on the one mesh (instance b: inflow scale 1 - 0.4 b / (B - 1),
    start u0 for b = 0, other seeds otherwise) stepped together: every
    matrix is shared, the velocity solves are one pcg_mixed solve over 2B
    columns and the pressure one solve over B columns (--ensemble-pressure:
    pcg_mixed, independent columns sharing each matrix read, or bpcg_mixed,
    one block Krylov space)."""
    import numpy as np
    from nsfem_amd import F64, fem
    B = args.ensemble
    scales = [1.0 - 0.4 * b / max(1, B - 1) for b in range(B)]
    starts = [u0]
    N = ctx.numVelocityNodes
    for b in range(1, B):
        u = initial_velocity(ctx, cond, 1 + b)
        ids, _ = ctx.dirichlet("vx")
        idy, _ = ctx.dirichlet("vy")
        u[ids] *= scales[b]
        u[idy + N] *= scales[b]
        starts.append(u)
    pressure = args.ensemble_pressure or ("amg_pcg_mixed" if args.pressure_solver == "amg" else "pcg_mixed")
    velocity = "pcg_mixed"
    if precision == "fp64":
        pressure, velocity = pressure.replace("_mixed", ""), "pcg"
    # tau: the single flow's step on the same mesh (already through the guard;
    # the ensemble's own estimate of lambda_max differs by the power
    # iteration's tolerance), so instance 0 can be compared with the single flow
    cfg = fem.ChorinConfig(fem.SolverConfig(velocity, 5000, target, 1), fem.SolverConfig(pressure, 50000, target, 1),
                           dtype=F64, timeStepPolicy=fem.TAU_CLAMP if tau is None else fem.TAU_UNCHECKED,
                           pressureStart=fem.PRESSURE_START_EXTRAPOLATE if args.pressure_start == "extrapolate" else 0,
                           velocityStart=velocity_start_code(args.velocity_start))
    ch, r, _, last = run_path(f"ensemble{B}_{precision}", ctx, cfg, args.tau if tau is None else tau, np.concatenate(starts),
                              args.warmup, args.steps,
                              not args.no_autotune, make=lambda c: ensemble_chorin(args, ctx, c, B, scales))
    del ch
    r["instances"] = B
    r["inflow_scales"] = scales
    r["solvers"] = {"velocity": velocity, "pressure": pressure}
    r["instance_steps_per_s"] = r["steps_per_s"] * B
    return r, last


def kernel_table(ch, res, reps=30, with_iteration=True, amg_pressure=False, velocity_fsai=False):
    """Per-kernel average launch time (HIP events on the library stream,
    replaying the launches one solver iteration makes), algorithmic bytes per
    launch, GB/s, and each kernel's estimated share of one step: launches per
    iteration times the measured iterations per step."""
    it = res["iters"] if res else None
    vel_iters = it["tentative"] + it["final"] if it else None
    out = {}
    for solver, iters in (("velocity", vel_iters), ("pressure", it["pressure"] if it else None)):
        kinds = ("direction", "sweep", "rho", "update", "finalize") + (("iteration",) if with_iteration else ())
        if solver == "pressure" and amg_pressure:
            kinds = ("direction", "update", "rho", "vcycle", "iteration")
        if solver == "velocity" and velocity_fsai:
            kinds = ("direction", "update", "fsai") + (("iteration",) if with_iteration else ())
        for k in kinds:
            ms, b, per_it = ch.time_kernel(f"{solver}_{k}", reps=reps)
            out[f"{solver}_{k}"] = {"avg_ms": ms, "bytes": b, "GBps": b / (ms * 1e-3) / 1e9 if b else None,
                                    "launches_per_iteration": per_it,
                                    "per_step": per_it * iters if it else None,
                                    "ms_per_step": ms * per_it * iters if it else None}
    for k in ("convection", "spmm"):
        ms, b, _ = ch.time_kernel(k, reps=reps)
        out[k] = {"avg_ms": ms, "bytes": b, "GBps": b / (ms * 1e-3) / 1e9, "per_step": 1, "ms_per_step": ms}
    return out


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic_chorin_500k_mixed.md")
LLC_BYTES = 256e6  # MI355X Infinity Cache


def apply_amg_smoother(args, ch, cfg):
    """--amg-chebyshev FROM[,DEGREE[,FRACTION]] on an AMG pressure solver."""
    if not args.amg_chebyshev or not cfg.pressureSolver.method.startswith("amg"):
        return
    vals = [float(v) for v in args.amg_chebyshev.split(",")]
    for key, v in zip(("amgChebyshevFrom", "amgChebyshevDegree", "amgChebyshevFraction"), vals):
        ch.set_solver_option(key, v, "pressure")


def velocity_start_code(name):
    """The ChorinConfig.velocityStart value of a --velocity-start choice."""
    from nsfem_amd import fem
    return {"zero": fem.VELOCITY_START_ZERO, "last": fem.VELOCITY_START_LAST,
            "extrapolate": fem.VELOCITY_START_EXTRAPOLATE}[name]


def copy_peak_gbps(mib=2048, reps=10):
    """Achievable streaming bandwidth on this device (GB/s, read + write): a
    2 GiB device-to-device copy, far beyond the 256 MB Infinity Cache,
    repeated back to back and timed with HIP events (nsfem_copy_bandwidth)."""
    import ctypes
    from nsfem_amd import _lib
    v = ctypes.c_double()
    _lib.call("nsfem_copy_bandwidth", ctypes.c_size_t(mib << 20), int(reps), ctypes.byref(v))
    return v.value / 1e9


def measured_traffic(kind):
    """HBM bytes per launch of a mixed-path kernel kind, from the committed
    rocprofv3 counter summary (FETCH_SIZE + WRITE_SIZE, calibrated); None if
    the summary does not list it."""
    try:
        with open(TRAFFIC_FILE) as f:
            for line in f:
                parts = line.split()
                if len(parts) == 3 and parts[0] == "traffic_bytes_per_launch" and parts[1] == kind:
                    return float(parts[2])
    except OSError:
        pass
    return None


def host_cpu_info():
    """Host cores usable by this process and the CPU model (/proc/cpuinfo's
    model name, what lscpu reports)."""
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": usable, "os_cpu_count": os.cpu_count(), "model": model}


def cpu_baseline(ctx, cond, tau, u0, budget_s=45.0, max_steps=10):
    """The CPU baseline: the FP64 C++ restatement of the step (oracle/
    cpu_step.cpp; OpenMP over the usable host cores, or OMP_NUM_THREADS),
    every solve SGS-PCG to 1e-10 like the GPU paths, from the same initial
    state.  One warm-up step (its state is the reference for the GPU paths'
    first step), then up to max_steps timed steps, stopping early once
    budget_s has passed (at least 2)."""
    import numpy as np
    from oracle import nsfem_oracle as O
    t0 = time.perf_counter()
    xyv, ev = ctx.velocity_mesh()
    octx = {"N": ctx.numVelocityNodes, "V": ctx.numPressureNodes, "M": ctx.scipy("velocityMass"),
            "K": ctx.scipy("velocityStiffness"), "Kp_int": ctx.scipy("pressureStiffnessInternal"),
            "VPD": ctx.scipy("velocityPressureDiv"), "PVD": ctx.scipy("pressureVelocityDiv"),
            "vm": {"elements": ev, "xy": xyv}, "dvx": ctx.dirichlet("vx"), "dvy": ctx.dirichlet("vy"),
            "int_p": ctx.internal_pressure(), "viscosity": cond.viscosity}
    info = host_cpu_info()
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or info["nproc"]
    cpu = O.CpuStep(octx, tau, rtol=1e-10, threads=threads)
    cpu.set_velocity(u0)
    t1 = time.perf_counter()
    warm_iters = cpu.run(1)
    first = cpu.state()
    t2 = time.perf_counter()
    iters, times = [], []
    while len(times) < max_steps and (len(times) < 2 or sum(times) < budget_s):
        t = time.perf_counter()
        iters.append(cpu.run(1)[0].tolist())
        times.append(time.perf_counter() - t)
    steps = len(times)
    rate = steps / sum(times)
    log(f"cpu baseline ({cpu.threads} threads, {info['model']}): set-up {t1 - t0:.1f}s, warm-up step {t2 - t1:.1f}s "
        f"(iterations {warm_iters[0].tolist()}), {steps} steps at {1.0 / rate:.2f} s/step")
    res = {"value": rate, "unit": "steps/s", "cores": cpu.threads, "kind": "port",
           "sample": f"{steps} Chorin steps of the same 500k-DOF workload after 1 warm-up step, from the same initial state: "
                     f"FP64 C++ restatement (oracle/cpu_step.cpp), OpenMP on {cpu.threads} threads, every solve SGS-PCG "
                     f"to 1e-10 relative",
           "seconds_per_step": 1.0 / rate, "iterations_per_step": iters, "host": info}
    return res, first


def build_batch(count, seed):
    """Independent channel-with-hole instances of uneven sizes (the seeded
    spec of mesher.uneven_instances), one block-diagonal context."""
    from nsfem_amd import fem, mesher
    t0 = time.perf_counter()
    inst = mesher.uneven_instances(count, seed)
    meshes, conds, info = [], [], []
    for d in inst:
        m = d["mesh"]
        meshes.append(mesher.to_triangle_mesh(m))
        conds.append(fem.DfgConditions(viscosity=d["viscosity"], peakVelocity=d["peak"], channelHeight=m["height"], profileScale=4.0))
        info.append({"pressure_dofs": len(m["xy"]), "radius": d["radius"], "peak": d["peak"], "re": d["re"]})
    t1 = time.perf_counter()
    ctx = fem.ChorinContext.batch(meshes, conds)
    t2 = time.perf_counter()
    log(f"batch of {count}: meshes {t1 - t0:.1f}s, context {t2 - t1:.1f}s, V = {ctx.numPressureNodes}, "
        f"N = {ctx.numVelocityNodes}")
    return ctx, info


def batch_leg(args, count, step_graph, with_kernels):
    """count uneven instances (C4) in one block-diagonal batch, FP64 and
    mixed: steps/s of the whole batch (every instance advances one step)."""
    import numpy as np
    from nsfem_amd import F64, fem
    t_setup = time.perf_counter()
    ctx, info = build_batch(count, args.seed)
    rng = np.random.default_rng(1)
    N = ctx.numVelocityNodes
    u0 = rng.uniform(-1, 1, 2 * N) * 1e-3
    for comp, off in (("vx", 0), ("vy", N)):
        ids, vals = ctx.dirichlet(comp)
        u0[ids + off] = vals
    setup_s = time.perf_counter() - t_setup
    target = 1e-10
    res = {}
    for name, method in (("fp64", "pcg"), ("mixed", "pcg_mixed")):
        cfg = fem.ChorinConfig(fem.SolverConfig(method, 5000, target, 1), fem.SolverConfig(method, 50000, target, 1), dtype=F64,
                               timeStepPolicy=fem.TAU_CLAMP, stepGraph=1 if step_graph else 0)
        def make(c):
            ch = fem.Chorin(ctx, c)
            if args.batch_precond == "fsai":
                ch.set_solver_option("preconditioner", 1, "both")
            return ch
        # the guard clamps tau to the batch's limit (the tightest instance's)
        ch, r, _, _ = run_path(f"batch{count}_{name}", ctx, cfg, args.tau, u0, args.warmup, args.steps, not args.no_autotune,
                               make)
        r["kernels"] = kernel_table(ch, r, args.kernel_reps, velocity_fsai=args.batch_precond == "fsai") if with_kernels else {}
        del ch
        r["instance_steps_per_s"] = r["steps_per_s"] * count
        r["pressure_dof_steps_per_s"] = r["steps_per_s"] * ctx.numPressureNodes
        res[name] = r
    return {"workload": f"chorin_step_batch_{count}_uneven", "instances": count, "captured_steps": bool(step_graph),
            "pressure_dofs_total": ctx.numPressureNodes, "velocity_dofs_per_component_total": N,
            "tau_requested": args.tau, "tau_used": res["mixed"]["time_step"]["tauUsed"],
            "lambda_max_nu": res["mixed"]["time_step"]["lambdaNu"], "solver_rtol": target, "seed": args.seed,
            "solvers": "pcg / pcg_mixed (per-instance convergence in one block-diagonal solve)",
            "preconditioner": "factorized sparse approximate inverse (block-diagonal)" if args.batch_precond == "fsai"
            else "one symmetric multicolour GS sweep",
            "instances_detail": info, "fp64": res["fp64"], "mixed": res["mixed"],
            "mixed_over_fp64": res["mixed"]["steps_per_s"] / res["fp64"]["steps_per_s"], "setup_s": setup_s,
            "_ctx": ctx}


def run_batch(args):
    """Batched uneven instances in one launch sequence (C4) as the bench line."""
    b = batch_leg(args, args.batch, args.step_graph, not args.no_kernels)
    ctx = b.pop("_ctx")
    res = {"fp64": b["fp64"], "mixed": b["mixed"]}
    N = ctx.numVelocityNodes
    target = b["solver_rtol"]
    info = b["instances_detail"]
    setup_s = b["setup_s"]
    out = {
        "metric": "batched Chorin instance-steps/s, mixed vs FP64 GPU, same residual",
        "value": res["mixed"]["instance_steps_per_s"], "unit": "instance-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": res["mixed"]["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64",
        "data": "synthetic: seeded channel meshes and seeded initial perturbation",
        "config": {"workload": f"chorin_step_batch_{args.batch}_uneven", "instances": args.batch,
                   "pressure_dofs_total": ctx.numPressureNodes, "velocity_dofs_per_component_total": N,
                   "tau_requested": args.tau, "tau_used": res["mixed"]["time_step"]["tauUsed"],
                   "lambda_max_nu": res["mixed"]["time_step"]["lambdaNu"], "solver_rtol": target, "seed": args.seed,
                   "instances_detail": info},
        "fp64": res["fp64"], "mixed": res["mixed"],
        "mixed_over_fp64": res["mixed"]["steps_per_s"] / res["fp64"]["steps_per_s"], "setup_s": setup_s,
    }
    print(json.dumps(out))
    return 0


def run_poisson(args):
    """Pressure-Poisson kernel alone: the pure-Neumann P1 Laplacian of a
    seeded channel mesh, uniform(-1, 1) right-hand side projected to zero
    mean, PCG to 1e-10 relative residual: FP64, mixed (FP64 Krylov, FP32
    preconditioner) and mixed_restart (FP32 PCG inner solves under FP64
    iterative refinement)."""
    import numpy as np
    from nsfem_amd import cu, fem, mesher
    t0 = time.perf_counter()
    m = mesher.channel_mesh(args.poisson, seed=args.seed)
    ctx = fem.ChorinContext(mesher.to_triangle_mesh(m), fem.DfgConditions(0.01, 1.0, m["height"], 4.0), gpu_assembly=True)
    K = ctx.scipy("pressureStiffness")
    setup_s = time.perf_counter() - t0
    n = K.shape[0]
    rng = np.random.default_rng(1)
    b = rng.uniform(-1, 1, n)
    b -= b.mean()
    out = {}
    for name, method, restart in (("fp64", "pcg", 0), ("mixed", "pcg_mixed", 0), ("mixed_restart", "pcg_mixed", 1),
                                  ("fp64_amg", "amg_pcg", 0), ("mixed_amg", "amg_pcg_mixed", 0)):
        s = cu.createSolver(method, 1, K, 50000, 1e-10, 1)
        s.setOption("projectMean", 1)
        if restart:
            s.setOption("mixedRestart", 1)
        amg = method.startswith("amg")
        shape = s.autotune(10) if not (args.no_autotune or amg) else s.shape()
        shape = tuple(shape) + s.sweepShape()
        times, iters = [], []
        for rep in range(args.warmup + args.steps):
            s.getRhs().upload(b)
            s.getSol().memsetZero()
            t = time.perf_counter()
            s.solve()
            dt = time.perf_counter() - t
            if rep >= args.warmup:
                times.append(dt)
                iters.append(s.getLastIterations())
        x = s.getSol().download()
        rel = float(np.linalg.norm(b - K @ x) / np.linalg.norm(b))
        out[name] = {"solves_per_s": 1.0 / float(np.mean(times)), "ms_per_solve": 1e3 * float(np.mean(times)),
                     "iterations": float(np.mean(iters)), "true_relres": rel, "shape": shape,
                     "ms_per_iteration": 1e3 * float(np.mean(times)) / float(np.mean(iters)),
                     "outer_iterations": s.getLastOuterIterations()}
        log(f"{name}: {out[name]['ms_per_solve']:.1f} ms per solve, {out[name]['iterations']:.0f} iterations, "
            f"true relres {rel:.2e}")
        del s
    res = {
        "metric": "pressure-Poisson PCG solves/s, mixed vs FP64 GPU, 1e-10 relative residual",
        "value": out["mixed"]["solves_per_s"], "unit": "solves/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": out["mixed"]["ms_per_solve"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic: seeded mesh and right-hand side",
        "config": {"workload": f"poisson_neumann_{n}", "rows": n, "nnz": int(K.nnz), "seed": args.seed},
        "fp64": out["fp64"], "mixed": out["mixed"], "mixed_restart": out["mixed_restart"],
        "mixed_over_fp64": out["mixed"]["solves_per_s"] / out["fp64"]["solves_per_s"], "setup_s": setup_s,
        "preconditioners": {"fp64, mixed, mixed_restart": "one symmetric multicolour GS sweep",
                            "fp64_amg, mixed_amg": "smoothed-aggregation V-cycle, multicolour GS smoothing on the finest level, Chebyshev on the coarser"},
        "fp64_amg": out["fp64_amg"], "mixed_amg": out["mixed_amg"],
        "mixed_amg_over_fp64": out["mixed_amg"]["solves_per_s"] / out["fp64"]["solves_per_s"],
    }
    print(json.dumps(res))
    return 0


def p1_laplacian(xy, tri):
    """Pure-Neumann P1 stiffness matrix of a triangulation, CSR: per triangle
    K_ij += |T| grad(phi_i).grad(phi_j) = (E_i . E_j) / (4 |T|), with E_i the
    edge opposite vertex i."""
    import numpy as np
    import scipy.sparse as sp
    p = xy[tri]
    e = np.stack([p[:, 2] - p[:, 1], p[:, 0] - p[:, 2], p[:, 1] - p[:, 0]], 1)
    det = np.abs(e[:, 2, 0] * e[:, 0, 1] - e[:, 2, 1] * e[:, 0, 0])  # 2 |T|
    loc = np.einsum("eik,ejk->eij", e, e) / (2.0 * det)[:, None, None]
    rows = np.repeat(tri, 3, axis=1).ravel()
    cols = np.tile(tri, (1, 3)).ravel()
    n = len(xy)
    k = sp.coo_matrix((loc.ravel(), (rows, cols)), shape=(n, n)).tocsr()
    k.sum_duplicates()
    k.sort_indices()
    return k


def run_multi_rhs(args):
    """Multi-right-hand-side stress configuration: the pure-Neumann P1
    Laplacian of a seeded channel mesh with about args.multi_rhs rows, 32
    zero-mean uniform(-1, 1) right-hand sides, FP32 preconditioner with FP64
    Krylov vectors and reductions, mean projection, to 1e-8 relative residual
    per column.  Block PCG (bpcg_mixed: one Krylov space per tile, k x k step
    matrices, verified true residuals) in RHS tiles of 4, 8, 16 or 32
    columns, the tile width picked by an autotuner: each width solves its
    first tile, whose time x (32 / width) predicts the time for all 32
    columns; then all 32 columns are solved at the chosen width from
    device-resident right-hand sides.  The independent-column
    solver (pcg_mixed tiles, one CG per column sharing the matrix) runs the
    same columns for comparison."""
    import numpy as np
    from nsfem_amd import cu, mesher
    nrhs = 32
    tol = 1e-8
    t0 = time.perf_counter()
    m = mesher.channel_mesh(args.multi_rhs, seed=args.seed)
    K = p1_laplacian(m["xy"].astype(np.float64), m["triangles"])
    del m
    n = K.shape[0]
    rng = np.random.default_rng(1)
    B = rng.uniform(-1, 1, (nrhs, n))
    B -= B.mean(axis=1, keepdims=True)
    setup_s = time.perf_counter() - t0
    log(f"multi-rhs: {n} rows, {K.nnz} nnz, {nrhs} right-hand sides, set-up {setup_s:.1f} s")
    Bd = cu.vec(B.ravel())
    Xd = cu.vec(nrhs * n, np.float64)

    def column_view(v, first, count):
        return cu.vec.view(v.ptr.value + first * n * 8, count * n, np.float64, owner=v)

    def solve_all(s, width):
        its = []
        t = time.perf_counter()
        for first in range(0, nrhs, width):
            column_view(Bd, first, width).copyTo(s.getRhs())
            s.getSol().memsetZero()
            s.solve()
            its.append(s.getLastIterations())
            s.getSol().copyTo(column_view(Xd, first, width))
        return time.perf_counter() - t, its

    def check():
        X = Xd.download().reshape(nrhs, n)
        rel = [float(np.linalg.norm(B[c] - K @ X[c]) / np.linalg.norm(B[c])) for c in range(nrhs)]
        mean = max(float(abs(X[c].mean()) / np.max(np.abs(X[c]))) for c in range(nrhs))
        return max(rel), mean

    # block PCG: tile-width autotuner.  A block's convergence rate depends on
    # its width only after the first iterations, so each width solves its
    # first tile completely; time x (32 / width) predicts the whole solve.
    tiles = {}
    for width in (4, 8, 16, 32):
        s = cu.createSolver("bpcg_mixed", width, K, 20000, tol, 1)
        s.setOption("projectMean", 1)
        column_view(Bd, 0, width).copyTo(s.getRhs())
        s.getSol().memsetZero()
        t = time.perf_counter()
        s.solve()
        dt = time.perf_counter() - t
        done = s.getLastIterations()
        tiles[width] = {"first_tile_s": dt, "first_tile_iterations": done, "ms_per_iteration": 1e3 * dt / max(1, done),
                        "predicted_s_all_columns": dt * (nrhs // width)}
        log(f"  block tile {width:2d}: first tile {done} iterations in {dt:.2f} s ({1e3 * dt / max(1, done):.2f} ms per "
            f"iteration), predicted {tiles[width]['predicted_s_all_columns']:.2f} s for {nrhs} columns")
        del s
    best = min(tiles, key=lambda w: tiles[w]["predicted_s_all_columns"])
    s = cu.createSolver("bpcg_mixed", best, K, 20000, tol, 1)
    s.setOption("projectMean", 1)
    times, iters = [], []
    for rep in range(args.warmup + args.steps):
        dt, its = solve_all(s, best)
        if rep >= args.warmup:
            times.append(dt)
            iters.append(its)
        log(f"  block PCG, {nrhs} columns in tiles of {best}: {dt:.2f} s, iterations {its}")
    block_rel, block_mean = check()
    t_block = float(np.mean(times))
    del s

    # independent columns (one CG per column, tiles of 8 sharing the matrix)
    s = cu.createSolver("pcg_mixed", 8, K, 50000, tol, 1)
    s.setOption("projectMean", 1)
    if not args.no_autotune:
        s.autotune(5)
    dt_ind, its_ind = solve_all(s, 8)
    ind_rel, ind_mean = check()
    log(f"  independent columns (pcg_mixed tiles of 8): {dt_ind:.2f} s, iterations {its_ind}")
    del s

    # the AMG-preconditioned CG over tiles of 16 columns (each column its own
    # CG scalars, one V-cycle for the tile)
    s = cu.createSolver("amg_pcg_mixed", 16, K, 2000, tol, 1)
    s.setOption("projectMean", 1)
    amg_times = []
    for rep in range(1 + args.steps):
        dt_amg, its_amg = solve_all(s, 16)
        if rep >= 1:
            amg_times.append(dt_amg)
    amg_rel, amg_mean = check()
    dt_amg = float(np.mean(amg_times))
    log(f"  AMG-preconditioned columns (amg_pcg_mixed tiles of 16): {dt_amg:.2f} s, iterations {its_amg}")
    del s
    res = {
        "metric": "multi-RHS PCG-MCGS: right-hand sides solved per s to 1e-8 (32 columns, block PCG, autotuned RHS tile)",
        "value": nrhs / t_block, "unit": "columns/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * t_block, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+f64", "data": "synthetic: seeded mesh and right-hand sides",
        "config": {"workload": f"multi_rhs_neumann_{n}x{nrhs}", "rows": n, "nnz": int(K.nnz), "rhs": nrhs, "tol": tol,
                   "method": "bpcg_mixed", "tile": best, "seed": args.seed},
        "block": {"seconds": t_block, "iterations_per_tile": iters[-1] if iters else [], "max_true_relres": block_rel,
                  "max_rel_mean": block_mean, "tiles": {str(w): v for w, v in tiles.items()}},
        "independent": {"seconds": dt_ind, "tile": 8, "iterations_per_tile": its_ind, "max_true_relres": ind_rel,
                        "max_rel_mean": ind_mean},
        "block_over_independent": dt_ind / t_block, "setup_s": setup_s,
        "amg": {"seconds": dt_amg, "tile": 16, "iterations_per_tile": its_amg, "max_true_relres": amg_rel,
                "max_rel_mean": amg_mean, "method": "amg_pcg_mixed (smoothed-aggregation V-cycle, multicolour GS smoothing on the finest level, Chebyshev on the coarser)"},
        "amg_over_block": t_block / dt_amg,
    }
    log(f"multi-rhs: block tile {best}, {t_block:.2f} s per {nrhs}-column solve (independent {dt_ind:.2f} s), "
        f"max true relres {block_rel:.2e}")
    print(json.dumps(res))
    return 0


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--nodes", type=int, default=500_000, help="P1 pressure DOFs")
    ap.add_argument("--tau", type=float, default=5e-4)
    ap.add_argument("--seed", type=int, default=20241103)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--paths", default="fp64,mixed,fp64_fast,fp64_mcgs,mixed_mcgs",
                    help="comma list of fp64 (the plain path: the reference's velocity solves and starts), mixed (the "
                         "headline path: FP32 preconditioners, --velocity-precond, --velocity-start, --pressure-start), "
                         "fp64_fast (the mixed path's algorithm in FP64), fp64_mcgs, mixed_mcgs (plain, with the "
                         "one-level multicolour GS pressure preconditioner); fp64 and mixed use the pressure "
                         "preconditioner of --pressure-solver")
    ap.add_argument("--shape", default="0,0,0,0",
                    help="launch shape velocityLanes,velocityRowsPerBlock,pressureLanes,pressureRowsPerBlock (0 = default), "
                         "optionally followed by the same four for the colour sweeps and then by the velocity and "
                         "pressure direction,update,rho rows, the velocity and pressure direction lanes and "
                         "whether each fuses its reductions (18 numbers, as the JSON records them)")
    ap.add_argument("--fast-shape", default="2,256,2,512,512,4096,1024,2,1",
                    help="velocity launch shape of the fast paths (mixed, fp64_fast): lanes, rows per block, sweep "
                         "lanes, sweep rows, direction / update / rho rows, direction lanes, fused reductions, and "
                         "optionally the approximate inverse's rows (0: the direction kernel's map); the "
                         "measured best on MI355X (the autotuner's direction-row candidates are multiples of the base "
                         "rows it picks first, and from a 1,024-row base miss these 512-row workgroups, 10 %% per "
                         "iteration); 'auto' autotunes them like the other paths; an explicit "
                         "--shape takes precedence")
    ap.add_argument("--no-kernels", action="store_true", help="skip the per-kernel timing table")
    ap.add_argument("--step-graph", action="store_true",
                    help="each step as one captured graph with device-side convergence (after a first regular step)")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default (or --shape) launch shapes")
    ap.add_argument("--kernels-only", action="store_true",
                    help="no time steps: build the workload and time each kernel kind (--kernel-reps launches); "
                         "for hardware-counter passes, which need few dispatches")
    ap.add_argument("--kernel-reps", type=int, default=30)
    ap.add_argument("--host-assembly", action="store_true", help="assemble the matrices on the host instead of the GPU")
    ap.add_argument("--poisson", type=int, default=0,
                    help="instead of the Chorin step: the pure-Neumann P1 Laplacian with about this many rows alone")
    ap.add_argument("--multi-rhs", type=int, default=0,
                    help="instead of the Chorin step: 32 right-hand sides on the pure-Neumann P1 Laplacian with about "
                         "this many rows, RHS tile width autotuned (4, 8, 16, 32)")
    ap.add_argument("--batch", type=int, default=0,
                    help="instead of the 500k workload: this many uneven instances (10k-250k pressure DOFs) in one batch")
    ap.add_argument("--cpu-budget", type=float, default=45.0,
                    help="seconds of timed CPU-baseline steps (at most 10 steps, at least 2)")
    ap.add_argument("--eager", action="store_true",
                    help="launch solver iterations from the host (no graph replay; for kernel tracing tools)")
    ap.add_argument("--ensemble", type=int, default=8,
                    help="also run an ensemble of this many flows on the same mesh (inflow scales 1.0 down to 0.6; "
                         "mixed velocity solves over 2B columns, pressure over B); 0 = skip")
    ap.add_argument("--pressure-start", default="extrapolate", choices=["last", "extrapolate"],
                    help="pressure solves' initial guess: the last pressure (as the reference), or 2 p_n - p_n-1 "
                         "(default; the same solution to the solver tolerance in fewer iterations)")
    ap.add_argument("--velocity-start", default="extrapolate", choices=["zero", "last", "extrapolate"],
                    help="velocity solves' initial guess: zero (as the reference), the last solution, or 2 a_n - a_n-1 "
                         "(default)")
    ap.add_argument("--velocity-precond", default="fsai", choices=["sgs", "fsai"],
                    help="velocity solves' preconditioner (single flows): one symmetric multicolour GS sweep (as the "
                         "reference) or the factorized sparse approximate inverse (default)")
    ap.add_argument("--batch-precond", default="fsai", choices=["sgs", "fsai"],
                    help="the batches' (C4) velocity and pressure preconditioner: one symmetric multicolour GS sweep or "
                         "the factorized sparse approximate inverse (default)")
    ap.add_argument("--amg-chebyshev", default="",
                    help="AMG pressure solves: FROM[,DEGREE[,FRACTION]] = the options amgChebyshevFrom, "
                         "amgChebyshevDegree, amgChebyshevFraction (default: the library's)")
    ap.add_argument("--pressure-solver", default="amg", choices=["mcgs", "amg"],
                    help="pressure preconditioner: one symmetric multicolour GS sweep (pcg, pcg_mixed) or the "
                         "smoothed-aggregation V-cycle (amg_pcg, amg_pcg_mixed)")
    ap.add_argument("--batch-leg", type=int, default=48,
                    help="also run this many uneven instances (10k-250k pressure DOFs) as one captured batch; 0 = skip")
    ap.add_argument("--ensemble-pressure", default=None, choices=["pcg_mixed", "bpcg_mixed", "amg_pcg_mixed"],
                    help="the ensemble's pressure solver (default: amg_pcg_mixed, or pcg_mixed with --pressure-solver mcgs)")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="also run 3 in-place scalings of an array this size (known traffic, to scale HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1"}))
        log("bench.py: only --gpus 1 is supported (single-GPU build)")
        return 2

    import numpy as np
    from nsfem_amd import F64, _lib, cu, fem

    if args.batch > 0:
        return run_batch(args)
    if args.poisson > 0:
        return run_poisson(args)
    if args.multi_rhs > 0:
        return run_multi_rhs(args)

    if args.copy_calibration_mib > 0:
        # a plain streaming kernel of known traffic: each launch reads and
        # writes the whole array once (scaleKernel<double>)
        n = args.copy_calibration_mib * (1 << 20) // 8
        a = cu.vec(n, np.float64)
        a.memsetZero()
        for _ in range(3):
            cu.scale(n, a, 1.0)
        _lib.call("nsfem_synchronize")
        log(f"calibration: 3 in-place scalings of {args.copy_calibration_mib} MiB (read + write)")
        del a

    t_setup = time.perf_counter()
    m, cond, ctx = build_workload(args.nodes, args.seed, gpu_assembly=not args.host_assembly)
    u0 = initial_velocity(ctx, cond, 1)
    target = 1e-10
    shape = [int(v) for v in args.shape.split(",")]
    if len(shape) not in (4, 8, 14, 16, 18):
        raise SystemExit("--shape takes 4, 8, 14, 16 or 18 numbers")
    vl, vr, pl, pr, vsl, vsr, psl, psr = (shape + [0] * 4)[:8]
    kind_rows = shape[8:14] if len(shape) >= 14 else None
    dir_lanes = shape[14:16] if len(shape) >= 16 else None
    fused_modes = shape[16:18] if len(shape) == 18 else None

    fast_shape = None
    if args.fast_shape != "auto" and args.shape == "0,0,0,0":
        fast_shape = [int(v) for v in args.fast_shape.split(",")]
        if len(fast_shape) not in (9, 10):
            raise SystemExit("--fast-shape takes 9 or 10 numbers or 'auto'")

    def make_chorin(cfg, fast=True):
        ch = fem.Chorin(ctx, cfg)
        if fused_modes is not None:
            for which, v in zip(("velocity", "pressure"), fused_modes):
                ch.set_solver_option("fusedReductions", v, which)
        if kind_rows is not None:
            for which, rows in (("velocity", kind_rows[0:3]), ("pressure", kind_rows[3:6])):
                for key, v in zip(("rowsDirection", "rowsUpdate", "rowsRho"), rows):
                    ch.set_solver_option(key, v, which)
        if dir_lanes is not None:
            for which, v in zip(("velocity", "pressure"), dir_lanes):
                ch.set_solver_option("lanesDirection", v, which)
        if fast and fast_shape is not None:
            for key, v in zip(("rowsDirection", "rowsUpdate", "rowsRho", "lanesDirection", "fusedReductions",
                               "rowsPreconditioner"), fast_shape[4:]):
                ch.set_solver_option(key, v, "velocity")
        apply_amg_smoother(args, ch, cfg)
        if fast and args.velocity_precond == "fsai":
            ch.set_solver_option("preconditioner", 1, "velocity")
        return ch

    def solvers(method, one_level=False, fast=False):
        pressure = method if (one_level or args.pressure_solver == "mcgs") else method.replace("pcg", "amg_pcg")
        v = fast_shape[:4] if fast and fast_shape is not None else (vl, vr, vsl, vsr)
        return (fem.SolverConfig(method, 5000, target, 1, *v),
                fem.SolverConfig(pressure, 50000, target, 1, pl, pr, psl, psr))

    eager = 1 if args.eager else 0
    sg = 1 if args.step_graph else 0
    ps = fem.PRESSURE_START_EXTRAPOLATE if args.pressure_start == "extrapolate" else fem.PRESSURE_START_LAST
    vs = velocity_start_code(args.velocity_start)
    def config(method, one_level=False, fast=False):
        # the plain paths keep the reference's velocity solves (zero start, the
        # GS sweep) and pressure start, so that the FP64 one matches the
        # reference to 1e-10; the fast ones take the options
        return fem.ChorinConfig(*solvers(method, one_level, fast), dtype=F64, eagerLaunches=eager, timeStepPolicy=fem.TAU_CLAMP,
                                stepGraph=sg, pressureStart=ps if fast else fem.PRESSURE_START_LAST,
                                velocityStart=vs if fast else fem.VELOCITY_START_ZERO)

    # fp64 / mixed / fp64_fast: the pressure preconditioner of
    # --pressure-solver; the *_mcgs paths: one symmetric multicolour GS sweep
    # (the one-level pcg)
    configs = {"fp64": config("pcg"), "mixed": config("pcg_mixed", fast=True), "fp64_fast": config("pcg", fast=True),
               "fp64_mcgs": config("pcg", True), "mixed_mcgs": config("pcg_mixed", True)}
    fast_paths = {"mixed", "fp64_fast"}
    amg_paths = {"fp64", "mixed", "fp64_fast"} if args.pressure_solver == "amg" else set()
    paths = [p.strip() for p in args.paths.split(",") if p.strip()]
    for p in paths:
        if p not in configs:
            raise SystemExit(f"unknown path {p}; choose from {', '.join(configs)}")
    if args.kernels_only:
        out = {"mode": "kernels-only", "config": {"workload": "chorin_step_channel_500k", "nodes": args.nodes},
               "kernels": {}}
        for name in paths:
            ch = make_chorin(configs[name], name in fast_paths)
            ch.upload_velocity(u0)
            out["kernels"][name] = kernel_table(ch, None, args.kernel_reps, amg_pressure=name in amg_paths,
                                    velocity_fsai=name in fast_paths and args.velocity_precond == "fsai")
            out[f"{name}_shape"] = ch.shape()
            del ch
        print(json.dumps(out))
        return 0
    setup_s = time.perf_counter() - t_setup

    results, kernels, firsts, lasts = {}, {}, {}, {}
    for name in paths:
        cfg = configs[name]
        # the library's guard clamps the requested tau to 1 / lambda_max(M^-1 nu K)
        # the fast paths' velocity shape is fixed (--fast-shape) and their
        # pressure solver (AMG) has no launch shape to tune
        tune = not args.no_autotune and not (name in fast_paths and fast_shape is not None and name in amg_paths)
        ch, r, first, last = run_path(name, ctx, cfg, args.tau, u0, args.warmup, args.steps, tune,
                                      lambda c, f=name in fast_paths: make_chorin(c, f))
        ti = r["time_step"]
        log(f"{name}: tau requested {ti['tauRequested']:g}, used {ti['tauUsed']:.4g} "
            f"(lambda_max(M^-1 nu K) = {ti['lambdaNu']:.4g})")
        r["pressure_solver"] = cfg.pressureSolver.method
        r["velocity_solver"] = cfg.velocitySolver.method
        # the per-kernel table for the two headline paths (the comparators are timed whole)
        kernels[name] = {} if (args.no_kernels or name not in ("fp64", "mixed")) else \
            kernel_table(ch, r, amg_pressure=name in amg_paths,
                         velocity_fsai=name in fast_paths and args.velocity_precond == "fsai")
        results[name], firsts[name], lasts[name] = r, first, last
        del ch

    def rel(a, b):
        return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))

    batch = None
    if args.batch_leg > 0:
        batch = batch_leg(args, args.batch_leg, True, False)
        del batch["_ctx"]
        log(f"batch of {args.batch_leg}: FP64 {batch['fp64']['instance_steps_per_s']:.1f}, mixed "
            f"{batch['mixed']['instance_steps_per_s']:.1f} instance-steps/s (captured steps)")

    ensemble = None
    if args.ensemble > 0:
        tau_single = results[paths[0]]["time_step"]["tauUsed"] if paths else None
        ensemble, ens_last = run_ensemble(args, ctx, cond, u0, target, tau=tau_single)
        # the same ensemble in FP64 throughout: what batching alone gives
        ens64, _ = run_ensemble(args, ctx, cond, u0, target, precision="fp64", tau=tau_single)
        ensemble["fp64_ensemble"] = ens64
        ensemble["mixed_over_fp64_ensemble"] = ensemble["steps_per_s"] / ens64["steps_per_s"]
        if "fp64" in results:
            # instance 0 has the single flow's inflow and start
            N = ctx.numVelocityNodes
            ensemble["instance0_vs_fp64"] = {"velocity": rel(ens_last[0][:2 * N], lasts["fp64"][0]),
                                            "pressure": rel(ens_last[1][:ctx.numPressureNodes], lasts["fp64"][1])}
            ensemble["instance_steps_over_fp64_steps"] = ensemble["instance_steps_per_s"] / results["fp64"]["steps_per_s"]
        if "mixed" in results:
            ensemble["instance_steps_over_mixed_steps"] = ensemble["instance_steps_per_s"] / results["mixed"]["steps_per_s"]
        log(f"ensemble: {ensemble['instance_steps_per_s']:.2f} instance-steps/s, "
            f"{ensemble.get('instance_steps_over_fp64_steps', 0):.2f}x the FP64 single flow")

    errors = {}
    if "fp64" in results and "mixed" in results:
        errors["mixed_vs_fp64_after_steps"] = {"velocity": rel(lasts["mixed"][0], lasts["fp64"][0]),
                                               "pressure": rel(lasts["mixed"][1], lasts["fp64"][1])}
    if "fp64_fast" in results and "mixed" in results:
        errors["mixed_vs_fp64_fast_after_steps"] = {"velocity": rel(lasts["mixed"][0], lasts["fp64_fast"][0]),
                                                    "pressure": rel(lasts["mixed"][1], lasts["fp64_fast"][1])}
    if "fp64" in results and "fp64_mcgs" in results:
        errors["fp64_vs_fp64_mcgs_after_steps"] = {"velocity": rel(lasts["fp64"][0], lasts["fp64_mcgs"][0]),
                                                   "pressure": rel(lasts["fp64"][1], lasts["fp64_mcgs"][1])}
    head = "mixed" if "mixed" in results else paths[0]
    tau = results[head]["time_step"]["tauUsed"]
    lam = results[head]["time_step"]["lambdaNu"] / cond.viscosity
    cpu = None
    if not args.no_cpu:
        cpu, (uo, po) = cpu_baseline(ctx, cond, tau, u0, budget_s=args.cpu_budget)
        for name in results:
            errors[f"{name}_vs_oracle_step1"] = {"velocity": rel(firsts[name][0], uo), "pressure": rel(firsts[name][1], po)}
    log(f"errors: {errors}")
    rmx, kernmx = results[head], kernels[head]
    r64 = results.get("fp64")

    # Dominant kernel of the headline path by estimated time per step.
    roofline = roofline_llc = None
    if kernmx:
        # single kernels only (an iteration or a V-cycle is a sequence of launches)
        dom_name = max((k for k in kernmx if not k.endswith(("_iteration", "_vcycle"))), key=lambda k: kernmx[k]["ms_per_step"])
        dom = kernmx[dom_name]
        peak = 8000.0
        traffic = measured_traffic(dom_name) if head == "mixed" and args.nodes == 500_000 else None
        copy = copy_peak_gbps()
        log(f"achievable copy bandwidth: {copy:.0f} GB/s")
        def line(name):
            k = kernmx[name]
            t = measured_traffic(name) if head == "mixed" and args.nodes == 500_000 else None
            return {"bound": "hbm", "kernel": name, "achieved": k["GBps"], "peak": peak, "unit": "GB/s",
                    "frac": k["GBps"] / peak, "traffic": t, "bytes_per_launch": k["bytes"], "avg_launch_ms": k["avg_ms"],
                    # a launch's algorithmic bytes within the 256 MB Infinity Cache: its operands can stay
                    # cache-resident between launches, so the HBM fraction understates what bounds it
                    "working_set": "exceeds the 256 MB Infinity Cache" if k["bytes"] > LLC_BYTES else "Infinity-Cache resident",
                    "traffic_source": (os.path.relpath(TRAFFIC_FILE, ROOT) + " (counter passes of the same kernel and "
                                       "workload at the launch shape that file names, committed; not measured in "
                                       "this run)") if t else None,
                    "achievable_copy": copy, "frac_of_achievable": k["GBps"] / copy}

        roofline = line(dom_name)
        big = [k for k in kernmx if not k.endswith(("_iteration", "_vcycle")) and kernmx[k]["bytes"] > LLC_BYTES]
        roofline_llc = line(max(big, key=lambda k: kernmx[k]["ms_per_step"])) if big else None
    out = {
        "metric": "Chorin steps/s @500k p-DOFs, mixed vs FP64 GPU, same residual; SpMV % HBM peak",
        "value": rmx["steps_per_s"],
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": rmx["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if head == "mixed" else "f64",
        "data": "synthetic: seeded channel mesh and seeded initial perturbation",
        "config": {"workload": "chorin_step_channel_500k", "pressure_dofs": ctx.numPressureNodes,
                   "velocity_dofs_per_component": ctx.numVelocityNodes, "elements": ctx.numElements,
                   "reynolds": 100, "peak_inflow": cond.peakVelocity, "viscosity": cond.viscosity,
                   "tau_requested": args.tau, "tau_used": tau, "lambda_max": lam, "solver_rtol": target,
                   "mixed_scheme": "FP64 Krylov vectors, SpMV and reductions; FP32 preconditioners",
                   "preconditioners": {"velocity": {"fast paths (mixed, fp64_fast)": "factorized sparse approximate inverse"
                                                    if args.velocity_precond == "fsai" else "one symmetric multicolour GS sweep",
                                                    "plain paths (fp64, fp64_mcgs, mixed_mcgs)": "one symmetric multicolour GS sweep"},
                                       "pressure": ("smoothed-aggregation V-cycle, multicolour GS smoothing on the finest level, Chebyshev on the coarser"
                                                    if args.pressure_solver == "amg" else "one symmetric multicolour GS sweep")},
                   "starts": {"fast paths": {"velocity": args.velocity_start, "pressure": args.pressure_start},
                              "plain paths": {"velocity": "zero", "pressure": "last"}},
                   "launch_shapes": {"fast paths velocity": ("fixed " + args.fast_shape) if fast_shape is not None
                                     else ("--shape" if args.shape != "0,0,0,0" else "autotuned"),
                                     "other paths": "--shape" if args.shape != "0,0,0,0" else
                                     ("defaults" if args.no_autotune else "autotuned")},
                   "pressure_solver": args.pressure_solver,
                   "amg_chebyshev_from_degree_fraction": args.amg_chebyshev or "1,2,0.2 (library default)",
                   "seed": args.seed},
        "path": head,
        "fp64": r64,
        "mixed": results.get("mixed"),
        "mixed_over_fp64": (results["mixed"]["steps_per_s"] / r64["steps_per_s"]) if (r64 and "mixed" in results) else None,
        "fp64_fast": results.get("fp64_fast"),
        "mixed_over_fp64_fast": (results["mixed"]["steps_per_s"] / results["fp64_fast"]["steps_per_s"])
        if ("mixed" in results and "fp64_fast" in results) else None,
        "errors": errors,
        "roofline": roofline,
        "roofline_llc_exceeding": roofline_llc,
        "fp64_mcgs": results.get("fp64_mcgs"),
        "mixed_mcgs": results.get("mixed_mcgs"),
        "mixed_mcgs_over_fp64_mcgs": (results["mixed_mcgs"]["steps_per_s"] / results["fp64_mcgs"]["steps_per_s"])
        if ("mixed_mcgs" in results and "fp64_mcgs" in results) else None,
        "mixed_over_fp64_mcgs": (rmx["steps_per_s"] / results["fp64_mcgs"]["steps_per_s"]) if "fp64_mcgs" in results else None,
        "kernels": kernels,
        "cpu_baseline": cpu,
        "ensemble": ensemble,
        "batch": batch,
        "setup_s": setup_s,
    }
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
