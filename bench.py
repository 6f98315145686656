#!/usr/bin/env python
"""Benchmark: linearised SSA solve to a 1e-6 relative residual on one MI355X.

Workload (default): the marine ice-sheet configuration at ~1M vertices
(seeded Delaunay mesh on a 1600 km square, ~4M combined Aa+Ac SOR points).
Set-up, untimed: mesh generation, secondary mesh data, one general update,
SIA, and `--picard0` Picard iterations of solve_SSA whose linear systems are
solved to 1e-6 (so the viscosity is that of a flowing state); the next
Picard step's linear system is then frozen (viscosity, sliding, right-hand
side).

One step = reset the velocities to zero, then solve the linear system to
||b - A w||_2 / ||b||_2 <= 1e-6 (FP64 norm).  Solvers timed:
  mixed_gmres        FP64 iterative refinement around restarted GMRES in FP32,
                     preconditioned by an FP32 aggregation-multigrid K-cycle
                     (headline path)
  fp64_krylov_mg     the same method with FP64 Krylov vectors around the same
                     FP32 cycle
  fp64_gmres         FP64 GMRES preconditioned by FP64 SOR sweeps
  fp64_sor           the reference iteration: FP64 SOR sweeps with the config
                     omega (the plain FP64 path)

Metric (`value`): Mvertex-sweeps/s = nVAaAc * (fine-level sweeps this path
performs) / (seconds this path needs to the tolerance) / 1e6.  The time to the
tolerance is `ms_per_step`; `effective_Mvertex_sweeps_per_s` restates it in
reference-SOR sweeps.  The roofline object is for the kernel that takes the
most time in the headline solve, timed live inside the timed region with HIP
events on the library's stream; the CPU baseline is the CPU oracle timed on a
bounded number of sweeps of the same system.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "SSA solve Mvertex-sweeps/s to 1e-6 residual, FP64 vs mixed; adjoint/forward ratio"
HBM_PEAK = 8.0e12


def log(*a):
    print(*a, file=sys.stderr, flush=True)


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_sor_colour_marine1M_hbm_traffic.txt")


def measured_traffic(kernel, nA):
    """HBM bytes per launch of `kernel` from the rocprofv3 FETCH_SIZE/WRITE_SIZE
    summary (copy-calibrated), when it was taken on a mesh of this size."""
    try:
        with open(TRAFFIC_FILE) as f:
            for line in f:
                p = line.split()
                if len(p) == 4 and p[0] == "traffic" and p[1] == kernel and int(p[2]) == nA:
                    return int(p[3])
    except (OSError, ValueError):
        pass
    return None


def cpu_threads():
    """Threads for the CPU baseline: every CPU this process may run on, within
    the OMP_NUM_THREADS share when that is set; and a description of the host
    (logical CPUs, physical cores and sockets from lscpu)."""
    allowed = len(os.sched_getaffinity(0))
    share = int(os.environ.get("OMP_NUM_THREADS", allowed) or allowed)
    threads = max(1, min(allowed, share))
    host = {"cpus_in_affinity": allowed, "omp_num_threads": os.environ.get("OMP_NUM_THREADS"),
            "threads_used": threads}
    try:
        import subprocess
        r = subprocess.run(["lscpu", "-p=CORE,SOCKET"], capture_output=True, text=True, timeout=20)
        rows = {tuple(l.split(",")) for l in r.stdout.splitlines() if l and not l.startswith("#")}
        host["physical_cores"] = len(rows)
        host["sockets"] = len({s for _, s in rows})
        m = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        for line in m.splitlines():
            if line.startswith("Model name:"):
                host["model"] = line.split(":", 1)[1].strip()
    except Exception as e:  # the description is informative only
        host["lscpu_error"] = repr(e)
    return threads, host


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nv", type=int, default=1_000_000)
    ap.add_argument("--seed", type=int, default=11)
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--picard0", type=int, default=3)
    ap.add_argument("--omega", type=float, default=0.0, help="SOR omega (0: the config's 1.3)")
    ap.add_argument("--omega-mixed", type=float, default=0.0, help="omega of the GMRES preconditioner")
    ap.add_argument("--check", type=int, default=256, help="residual checks of the plain SOR path")
    ap.add_argument("--max-sweeps", type=int, default=400000)
    ap.add_argument("--gmres", type=int, default=0,
                    help="GMRES restart length (default 20 with the multigrid preconditioner, 40 with SOR)")
    ap.add_argument("--precond", type=int, default=0,
                    help="SOR sweeps per preconditioner application (sor: default 5) or smoothing sweeps "
                         "before and after the coarse correction (multigrid: default 2)")
    ap.add_argument("--pc", default="multigrid", choices=["sor", "multigrid"],
                    help="preconditioner of the mixed_gmres path")
    ap.add_argument("--orth-passes", type=int, default=1, help="Gram-Schmidt passes (1 or 2)")
    ap.add_argument("--full-sweeps", action="store_true",
                    help="preconditioner sweeps repeat the colour at each turn of direction")
    ap.add_argument("--krylov", default="gmres", choices=["gmres", "bicgstab"],
                    help="Krylov method of the mixed_gmres / fp64_gmres paths")
    ap.add_argument("--mode", default="both",
                    choices=["both", "mixed_gmres", "fp64_gmres", "fp64_sor"])
    ap.add_argument("--cpu-seconds", type=float, default=25.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-ladder", action="store_true", help="skip the mixed path at tolerances 1e-7..1e-9")
    ap.add_argument("--mark", action="store_true",
                    help="launch a one-element torch kernel before and after the headline's timed solves "
                         "(brackets them in a kernel trace)")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-step", action="store_true", help="skip the mixed-vs-FP64 dynamics steps")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched ensemble")
    ap.add_argument("--no-tune", action="store_true", help="keep the default sweep launch shape")
    ap.add_argument("--no-ref-steps", action="store_true", help="skip the FP64 reference-step configs")
    ap.add_argument("--no-rows", action="store_true",
                    help="skip the paths around the solve (geometry, physical properties, SIA-3D, time step, "
                         "upwind derivatives, remapping)")
    ap.add_argument("--remap-nv", type=int, default=200_000, help="vertices of the remapping meshes")
    ap.add_argument("--batch-n", type=int, default=64, help="members of the batched ensemble")
    ap.add_argument("--batch-min", type=int, default=5000)
    ap.add_argument("--batch-max", type=int, default=100_000)
    ap.add_argument("--batch-picard", type=int, default=3)
    ap.add_argument("--step-picard", type=int, default=3, help="Picard iterations of those steps")
    ap.add_argument("--step-tol", type=float, default=1e-9, help="tighter tolerance of those steps")
    ap.add_argument("--adj-nv", type=int, default=200_000, help="vertices of the adjoint mesh")
    ap.add_argument("--adj-picard", type=int, default=5, help="Picard steps of the adjoint forward")
    ap.add_argument("--adj-tol", type=float, default=1e-11, help="linear-solve tolerance of the gradient's solves")
    ap.add_argument("--adj-fd", type=int, default=32, help="finite-difference directions to check")
    ap.add_argument("--adj-every", type=int, default=1, help="checkpoint every c-th Picard iterate")
    args = ap.parse_args()
    args.pc_type = 1 if args.pc == "multigrid" else 0     # UFE_PRECOND_MULTIGRID / UFE_PRECOND_SOR
    args.sor_sweeps = args.precond if (args.precond > 0 and args.pc == "sor") else 5
    args.pc_sweeps = args.precond if args.precond > 0 else (2 if args.pc == "multigrid" else 5)
    args.gmres_sor = args.gmres if args.gmres > 0 else 40
    args.gmres_pc = args.gmres if args.gmres > 0 else (20 if args.pc == "multigrid" else 40)

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(json.dumps({"error": "this build runs on one GPU only (the sharded multi-GPU path is "
                                   "not built); run with --gpus 1"}))
        sys.exit(2)

    import numpy as np
    from ufemism_amd import abi, geometry
    from ufemism_amd.dynamics import frozen_ssa_system

    ens = None
    if not args.no_batch:       # CPU-only mesh generation in forked workers, before any GPU context
        from ufemism_amd import batch
        tb = time.time()
        ens = batch.ensemble(n=args.batch_n, nv_min=args.batch_min, nv_max=args.batch_max, seed=args.seed,
                             picard=args.batch_picard, workers=cpu_threads()[0])
        log(f"[batch] {len(ens)} meshes, {sum(w.mesh.nV for w in ens)} vertices, generated in {time.time() - tb:.1f}s")
    remap_wls = None
    if not args.no_rows:        # two meshes to remap between, generated before any GPU context
        tb = time.time()
        remap_wls = [geometry.marine(n_target=args.remap_nv, seed=args.seed + 100, picard=1),
                     geometry.marine(n_target=int(args.remap_nv * 1.25), seed=args.seed + 200, picard=1)]
        log(f"[rows] remapping meshes nV={remap_wls[0].mesh.nV} -> {remap_wls[1].mesh.nV} "
            f"generated in {time.time() - tb:.1f}s")
    t0 = time.time()
    wl = geometry.marine(n_target=args.nv, seed=args.seed, picard=args.picard0)
    if args.omega > 0:
        wl.params.SSA_SOR_omega = args.omega
    log(f"[bench] mesh nV={wl.mesh.nV} generated in {time.time() - t0:.1f}s")
    ctx = abi.Context(0)
    t1 = time.time()
    step, st0 = frozen_ssa_system(ctx, wl, picard=args.picard0, warm_tol=1e-6)
    sizes = ctx.sizes
    warm_s = time.time() - t1
    log(f"[bench] mesh build+upload and {st0.picard_iterations} warm-up Picard steps {warm_s:.1f}s sizes={sizes}")
    tuning = None
    if not args.no_tune:
        t = ctx.sweep_autotune(10)
        tuning = {"threads_per_block": t.threads_per_block, "schedule": t.schedule, "loads": t.loads,
                  "candidates": t.candidates()}
        log(f"[bench] sweep autotune: {t.threads_per_block} threads, schedule {t.schedule}, loads {t.loads}")
    nA = sizes["nVAaAc"]
    sb = ctx.sweep_bytes()
    omega = wl.params.SSA_SOR_omega
    omega_mixed = args.omega_mixed if args.omega_mixed > 0 else omega

    def solve(path, tol=args.tol):
        ctx.fill("U_SSA_AaAc", 0.0)
        ctx.fill("V_SSA_AaAc", 0.0)
        if path == "fp64_sor":      # the reference iteration: FP64 SOR sweeps, config omega
            return ctx.ssa_linear_solve(tol, args.max_sweeps, precision=abi.UFE_FP64,
                                        omega=omega, check_interval=args.check)
        if path == "fp64_gmres":    # FP64 GMRES around FP64 SOR sweeps
            return ctx.ssa_linear_solve(tol, args.max_sweeps, precision=abi.UFE_FP64, omega=omega,
                                        gmres=args.gmres_sor, precond_sweeps=args.sor_sweeps,
                                        orth_passes=args.orth_passes, precond_type=abi.UFE_PRECOND_SOR)
        # mixed_gmres: FP32 Krylov; fp64_krylov_mg: FP64 Krylov; both around the same preconditioner
        mixed = path == "mixed_gmres"
        return ctx.ssa_linear_solve(tol, args.max_sweeps, precision=abi.UFE_MIXED if mixed else abi.UFE_FP64,
                                    omega=omega_mixed, gmres=args.gmres_pc, precond_sweeps=args.pc_sweeps,
                                    orth_passes=args.orth_passes,
                                    method=1 if args.krylov == "bicgstab" else 0,
                                    full_sweeps=args.full_sweeps, precond_type=args.pc_type)

    def marker():
        """A one-element torch kernel: brackets the headline's timed solves in a
        kernel trace (--mark), outside the timed region."""
        import torch
        torch.ones(1, device="cuda").add_(1.0)
        torch.cuda.synchronize()

    def run(path, steps, warmup, tol=args.tol):
        for _ in range(warmup):
            solve(path, tol)
        mark = args.mark and path == order[0] and tol == args.tol
        if mark:
            marker()
        ctx.profile_read()
        ctx.profile_enable(True)
        ctx.synchronize()
        ta = time.perf_counter()
        sweeps, last = 0, None
        for _ in range(steps):
            last = solve(path, tol)
            sweeps += last.sweeps
        ctx.synchronize()
        el = time.perf_counter() - ta
        if mark:
            marker()
        ctx.profile_enable(False)
        prof = ctx.profile_read()
        sol = (ctx.get("U_SSA_AaAc"), ctx.get("V_SSA_AaAc"))
        res = dict(seconds_per_solve=el / steps, sweeps_per_solve=last.sweeps,
                   matvecs_per_solve=last.matvecs, refinements=last.refinements,
                   rel_residual=last.rel_residual, converged=last.status == 0,
                   raw_Mvertex_sweeps_per_s=nA * sweeps / el / 1e6,
                   kernel_launches=prof.total_launches, kernel_launches_timed=prof.launches,
                   kernel_avg_us=1e6 * prof.seconds / max(prof.launches, 1),
                   kernel_bytes_per_launch=prof.algorithmic_bytes / max(prof.launches, 1),
                   kernel_GBps=prof.algorithmic_bytes / max(prof.seconds, 1e-30) / 1e9,
                   kernel_time_share=prof.seconds * prof.total_launches / max(prof.launches, 1) / el)
        log(f"[bench] {path} tol {tol:g}: {res}")
        return res, sol

    # the converged FP64 solution every path's velocities are measured against
    st_ref = solve("fp64_krylov_mg", 1e-11)     # near the FP64 floor of this path at 1M
    U_ref, V_ref = ctx.get("U_SSA_AaAc"), ctx.get("V_SSA_AaAc")
    den = float(np.sqrt(np.sum(U_ref ** 2 + V_ref ** 2)))
    log(f"[bench] FP64 reference solution: residual {st_ref.rel_residual:.2e}, {st_ref.matvecs} matvecs")

    def verr(sol):
        u, v = sol
        return float(np.sqrt(np.sum((u - U_ref) ** 2 + (v - V_ref) ** 2)) / den)

    paths = {}
    order = ["mixed_gmres", "fp64_krylov_mg", "fp64_gmres", "fp64_sor"] if args.mode == "both" else [args.mode]
    for path in order:
        steps, warm = (args.steps, args.warmup) if path == "mixed_gmres" else (1, 0 if path == "fp64_sor" else 1)
        log(f"[bench] running {path} ({steps} timed, {warm} warm-up)")
        paths[path], sol = run(path, steps, warm)
        paths[path]["velocity_error_vs_fp64_converged"] = verr(sol)

    head = order[0]
    R0 = paths[head]
    sor_ok = "fp64_sor" in paths and paths["fp64_sor"]["converged"]
    ref_sweeps = paths["fp64_sor"]["sweeps_per_solve"] if sor_ok else None
    for path, r in paths.items():
        # reference-equivalent rate: the reference iteration's sweep count to the
        # same residual, divided by this path's time to that residual
        r["effective_Mvertex_sweeps_per_s"] = (nA * ref_sweeps / r["seconds_per_solve"] / 1e6
                                               if (ref_sweeps and r["converged"]) else None)
    # the mixed path at tighter tolerances: where its velocities come within 1e-6
    # of the converged FP64 solution
    ladder = {}
    if head == "mixed_gmres":
        ladder[f"{args.tol:g}"] = {"ms": round(1e3 * R0["seconds_per_solve"], 3), "matvecs": R0["matvecs_per_solve"],
                                   "rel_residual": R0["rel_residual"],
                                   "velocity_error": R0["velocity_error_vs_fp64_converged"]}
    if head == "mixed_gmres" and not args.no_ladder:
        for tol in (1e-7, 1e-8, 1e-9):
            r, sol = run("mixed_gmres", max(1, args.steps // 2), 1, tol)
            ladder[f"{tol:g}"] = {"ms": round(1e3 * r["seconds_per_solve"], 3), "matvecs": r["matvecs_per_solve"],
                                  "rel_residual": r["rel_residual"], "velocity_error": verr(sol)}
        if "fp64_krylov_mg" in paths:
            for tol in (1e-8, 1e-9):
                r, sol = run("fp64_krylov_mg", 1, 1, tol)
                ladder[f"{tol:g}"]["fp64_krylov_mg"] = {"ms": round(1e3 * r["seconds_per_solve"], 3),
                                                        "velocity_error": verr(sol)}
    R = paths[head]
    kname = {"mixed_gmres": "k_sor_colour_f32"}.get(head, "k_sor_colour")
    if head == "mixed_gmres" and tuning:
        kname = f"k_sor_colour_f32_tuned<{tuning['threads_per_block']},{tuning['schedule']},{tuning['loads']}>"
    meets = [k for k, v in ladder.items() if v["velocity_error"] <= 1e-6]
    out = {
        "metric": METRIC,
        "value": round(R["raw_Mvertex_sweeps_per_s"], 3),
        "unit": "Mvertex-sweeps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(1e3 * R["seconds_per_solve"], 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+f64" if head == "mixed_gmres" else "f64", "data": "synthetic",
        "config": {
            "workload": f"marine ice sheet, seeded Delaunay mesh nV={sizes['nV']} "
                        f"(nVAaAc={nA}) on a 1600 km square; linearised SSA after "
                        f"{st0.picard_iterations} Picard steps solved to 1e-6, solved from zero to "
                        f"FP64 relative residual {args.tol:g}",
            "nV": sizes["nV"], "nVAaAc": nA, "mean_degree": round(sb["mean_degree"], 4),
            "path": head, "omega": omega, "omega_inner": omega_mixed,
            "krylov": args.krylov, "gmres_restart": args.gmres_pc, "preconditioner": args.pc,
            "precond_sweeps": args.pc_sweeps,
            "orth_passes": args.orth_passes,
            "value_definition": "nVAaAc * (fine-level SOR sweeps this path performs) / "
                                "(seconds to the tolerance) / 1e6; ms_per_step is the time to the tolerance",
        },
        "roofline": {
            "bound": "hbm", "achieved": round(R["kernel_GBps"], 2), "peak": HBM_PEAK / 1e9,
            "unit": "GB/s", "frac": round(R["kernel_GBps"] * 1e9 / HBM_PEAK, 4),
            "traffic": measured_traffic(kname, nA),
            "kernel": kname, "bytes_per_launch": round(R["kernel_bytes_per_launch"]),
            "avg_launch_us": round(R["kernel_avg_us"], 3),
            "kernel_time_share": round(R["kernel_time_share"], 4),
            "traffic_source": "rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE per launch, scaled by a 1 GiB "
                              "copy (read x2.0, write x1.0); null when not measured on this mesh size",
        },
        "time_to_tolerance_ms": round(1e3 * R["seconds_per_solve"], 3),
        "effective_Mvertex_sweeps_per_s": (round(R["effective_Mvertex_sweeps_per_s"], 1)
                                           if R["effective_Mvertex_sweeps_per_s"] else None),
        "paths": paths,
        "mixed_tolerance_ladder": ladder,
        "loosest_tolerance_with_velocity_error_le_1e-6": (max(meets, key=float) if meets else None),
        "sweep_tuning": tuning,
        "multigrid": ctx.multigrid_info() if args.pc == "multigrid" else None,
        "warmup_picard": {"iterations": st0.picard_iterations, "seconds": round(warm_s, 3)},
    }
    if sor_ok:
        out["speedup_vs_fp64_sor"] = round(paths["fp64_sor"]["seconds_per_solve"] / R["seconds_per_solve"], 3)
    for p in ("fp64_gmres", "fp64_krylov_mg"):
        if p in paths and head != p:
            out[f"speedup_vs_{p}"] = round(paths[p]["seconds_per_solve"] / R["seconds_per_solve"], 3)

    if not args.no_cpu:
        try:
            from oracle import pyoracle
            threads, host = cpu_threads()
            om = pyoracle.mesh_from_arrays(wl.mesh, step.ma.arr, step.ma.colour_nV)
            fields = {k: ctx.get(k) for k in ("Hi_AaAc", "Hb_AaAc", "SL_AaAc", "eta_AaAc", "S_AaAc",
                                              "dHs_dx_shelf_AaAc", "dHs_dy_shelf_AaAc")}
            fields["U_SSA_AaAc"] = np.zeros(nA)
            fields["V_SSA_AaAc"] = np.zeros(nA)
            pyoracle.time_sor_sweeps(om, fields, wl.params, 2, threads)       # first touch
            dt1, n1 = pyoracle.time_sor_sweeps(om, fields, wl.params, 8, threads)
            nsw = max(2, int(args.cpu_seconds / max(dt1 / max(n1, 1), 1e-6)))
            dt, n = pyoracle.time_sor_sweeps(om, fields, wl.params, nsw, threads)
            out["cpu_baseline"] = {
                "value": round(nA * n / dt / 1e6, 3), "unit": "Mvertex-sweeps/s", "cores": threads,
                "kind": "port",
                "sample": f"{n} FP64 SOR sweeps of the CPU oracle (OpenMP, {threads} threads) "
                          f"on the same {nA}-point linear system, {dt:.1f}s",
                "host": host}
            try:   # the same on one thread, about a third of the time
                d1, m1 = pyoracle.time_sor_sweeps(om, fields, wl.params, 2, 1)
                ns1 = max(2, int(args.cpu_seconds / 3 / max(d1 / max(m1, 1), 1e-6)))
                d1, m1 = pyoracle.time_sor_sweeps(om, fields, wl.params, ns1, 1)
                out["cpu_baseline"]["single_thread"] = {
                    "value": round(nA * m1 / d1 / 1e6, 3), "unit": "Mvertex-sweeps/s", "cores": 1,
                    "sample": f"{m1} FP64 SOR sweeps of the CPU oracle on one thread, {d1:.1f}s"}
            except Exception as e:
                out["cpu_baseline"]["single_thread"] = {"value": None, "error": repr(e)}
            finally:
                pyoracle.set_threads(threads)
        except Exception as e:  # the CPU baseline is reported, never required
            out["cpu_baseline"] = {"value": None, "error": repr(e)}
    if not args.no_step:
        out["precision_error"] = bench_precision(args, ctx, step, wl)
    if not args.no_rows:
        out["rows"] = bench_rows(args, ctx, step, wl, remap_wls)
    if not args.no_ref_steps:
        out["reference_steps"] = bench_reference_steps(args, ctx)
    if not args.no_adjoint:
        out["adjoint"] = bench_adjoint(args, ctx)
    if not args.no_batch:
        out["batch"] = bench_batch(args, ens)
    print(json.dumps(out), flush=True)
    ctx.close()


def bench_rows(args, ctx, step, wl, remap_wls, reps=10):
    """The paths around the solve on the 1M mesh, each timed over `reps`
    calls (every call ends with a stream synchronisation): the mesh geometry
    builder, ice_physical_properties, solve_SIA_3D, the critical-time-step
    reduction and the upwind derivatives; HBM rates use the algorithmic bytes
    below (nZ = 15, mean degree n).  Remapping: build and field rates between
    two marine meshes."""
    import numpy as np
    import torch
    from ufemism_amd import abi
    from ufemism_amd.dynamics import Step

    nV, nAc, nZ = ctx.sizes["nV"], ctx.sizes["nAc"], wl.params.nZ
    n = 6.0
    res = {}

    def timed(fn, k=reps):
        fn()
        ctx.synchronize()
        t = time.perf_counter()
        for _ in range(k):
            fn()
        ctx.synchronize()
        return (time.perf_counter() - t) / k

    def row(name, sec, nbytes, extra=None):
        r = {"ms": round(sec * 1e3, 3)}
        if nbytes:
            r["algorithmic_bytes"] = int(nbytes)
            r["GBps"] = round(nbytes / sec / 1e9, 1)
            r["hbm_frac"] = round(nbytes / sec / 8e12, 3)
        if extra:
            r.update(extra)
        res[name] = r
        log(f"[rows] {name}: {r}")

    tg = time.perf_counter()
    geo = ctx.mesh_geometry(wl.mesh)
    row("mesh_geometry", time.perf_counter() - tg, None,
        {"note": "Tricc, A, Cw, TriA, R of the 1M mesh, host-to-device copies and checks included",
         "area_error": geo["stats"].area_error})
    step.reset()
    ctx.update_general_ice_model_data()
    ctx.set("Ti", wl.temperature())
    # Aa: Ti, Hi, mask in; Ti_pmp, A, Cpi, Ki, A_mean out; Ac: Aci, two Ti gathers, mask in; Ti_Ac, A, mean out
    row("ice_physical_properties", timed(lambda: ctx.ice_physical_properties(1)),
        nV * (40 * nZ + 20) + nAc * (32 * nZ + 20))
    ctx.solve_SIA()
    # U_3D/V_3D: 5 Aa fields + A_flow in, U_3D/V_3D out; W_3D: 6 Aa fields, Nx/Ny, own and
    # neighbour U_3D/V_3D per layer in, W_3D out
    sia3d = nV * ((24 * nZ + 44) + (56 + 20 * (n + 1) + 16 * (n + 1) * nZ + 8 * nZ))
    row("solve_SIA_3D", timed(ctx.solve_SIA_3D), sia3d)
    # per Ac point: D_SIA_Ac, Aci, both vertices' V, U_SSA, V_SSA; per vertex: A, U_SSA, V_SSA, U_3D, V_3D
    row("critical_timesteps", timed(ctx.critical_timesteps), nAc * 80 + nV * (24 + 16 * nZ),
        {"dt_crit": [round(x, 6) for x in ctx.critical_timesteps()]})
    U = torch.randn(nV * nZ, dtype=torch.float64, device="cuda") * 100
    V = torch.randn_like(U) * 100
    d = torch.randn_like(U)
    dx, dy = torch.empty_like(U), torch.empty_like(U)
    torch.cuda.synchronize()
    up = lambda: ctx.upwind_derivatives_3D(U.data_ptr(), V.data_ptr(), d.data_ptr(), dx.data_ptr(), dy.data_ptr(), nZ)
    t_up = timed(up)
    row("upwind_derivatives_3D", t_up, None, {"Mvertex_layers_per_s": round(nV * nZ / t_up / 1e6, 1)})
    del U, V, d, dx, dy
    if remap_wls:
        ca, cb = abi.Context(0), abi.Context(0)
        try:
            Step(ca, remap_wls[0])
            Step(cb, remap_wls[1])
            ma, mb = ca.mesh_build(remap_wls[0].mesh), cb.mesh_build(remap_wls[1].mesh)
            tb = time.perf_counter()
            st = cb.remap_build(ca, ma, mb)
            tbuild = time.perf_counter() - tb
            nVa, nVb = remap_wls[0].mesh.nV, remap_wls[1].mesh.nV
            src = torch.rand(nVa * nZ, dtype=torch.float64, device="cuda")
            dst = torch.empty(nVb * nZ, dtype=torch.float64, device="cuda")
            torch.cuda.synchronize()
            rates = {}
            for m in ("trilin", "nearest_neighbour", "cons_1st_order", "cons_2nd_order"):
                sec = timed(lambda m=m: cb.remap_field(ca, m, src.data_ptr(), dst.data_ptr(), nZ))
                rates[m] = {"ms": round(sec * 1e3, 3), "Mvertex_layers_per_s": round(nVb * nZ / sec / 1e6, 1)}
            tr = time.perf_counter()
            cb.remap_ice_model(ca)
            res["remap"] = {"nV_src": nVa, "nV_dst": nVb, "build_seconds": round(tbuild, 3),
                            "build_kernel_seconds": round(st.seconds, 3), "entries": st.n_entries,
                            "fallback_vertices": st.n_fallback, "apply_nZ_15": rates,
                            "remap_ice_model_ms": round((time.perf_counter() - tr) * 1e3, 3)}
            log(f"[rows] remap: {res['remap']}")
        finally:
            ca.close()
            cb.close()
    step.reset()
    return res


def bench_precision(args, ctx, step, wl):
    """One full dynamics step (general update, SIA, SSA with K Picard
    iterations, thickness update) with every linearised system solved by
    GMRES-IR in mixed precision and by FP64 GMRES around the same
    preconditioner, to the same tolerance: the relative L2 error of the mixed
    step's u, v and H against the FP64 step."""
    import copy
    import numpy as np
    from ufemism_amd import abi

    def rel(a, b):
        return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))

    p = copy.deepcopy(wl.params)
    p.SSA_max_outer_loops = args.step_picard
    res = {"picard_iterations": args.step_picard}
    for tol in (args.tol, args.step_tol):
        outs = {}
        for name, mode in (("fp64", abi.UFE_FP64_KRYLOV), ("mixed", abi.UFE_MIXED)):
            step.reset()
            for f in ("U_SSA", "V_SSA", "Ux_SSA_Ac", "Uy_SSA_Ac", "N_AaAc"):
                ctx.fill(f, 0.0)
            # both around the same preconditioner: only the Krylov precision differs
            ctx.set_params(p, mode, linear_rel_tol=tol, gmres_restart=args.gmres_pc,
                           precond_sweeps=args.pc_sweeps, precond_type=args.pc_type)
            ctx.synchronize()
            ta = time.perf_counter()
            r = step.run()
            el = time.perf_counter() - ta
            outs[name] = (r, el)
            log(f"[step] tol {tol:g} {name}: {el:.2f}s picard {r['picard']} sweeps {r['sweeps']}")
        (a, ta_), (b, tb_) = outs["fp64"], outs["mixed"]
        res[f"tol_{tol:g}"] = {
            "seconds": {"fp64": round(ta_, 3), "mixed": round(tb_, 3)},
            "sweeps": {"fp64": a["sweeps"], "mixed": b["sweeps"]},
            "relL2_mixed_vs_fp64": {k: rel(b[k], a[k]) for k in
                                   ("U_SSA", "V_SSA", "Ux_SSA_Ac", "Uy_SSA_Ac", "Hi")},
        }
    ctx.set_params(wl.params, abi.UFE_FP64)
    return res


def bench_reference_steps(args, ctx):
    """The FP64 configurations that run the reference arithmetic and stopping
    rules on the GPU: the 8k dome (one full dynamics step, 10 Picard) against
    the CPU oracle's step, and the 200k marine SSA (30 Picard to RN 1e-6)
    against the oracle's sweep rate on a bounded sample of the same system."""
    import numpy as np
    from ufemism_amd import geometry
    from ufemism_amd.dynamics import Step

    out = {}
    threads = cpu_threads()[0]

    def rel(a, b):
        return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))

    # C1: dome, one step, GPU vs oracle
    wl = geometry.dome(8000, seed=1, picard=10)
    st = Step(ctx, wl)
    st.run()                                   # warm-up
    st.reset()
    ctx.synchronize()
    ta = time.perf_counter()
    g = st.run()
    tg = time.perf_counter() - ta
    nA = ctx.sizes["nVAaAc"]
    c1 = {"workload": f"dome nV={wl.mesh.nV} (nVAaAc={nA}), general update + SIA + SSA (10 Picard, "
                      f"reference SOR and stopping rules, omega 1.3) + thickness update, FP64",
          "gpu_seconds": round(tg, 4), "picard": g["picard"], "sweeps": g["sweeps"]}
    sw = int(sum(g["sweeps"]))
    c1["gpu_Mvertex_sweeps_per_s"] = round(nA * sw / tg / 1e6, 1)
    if not args.no_cpu:
        from oracle import pyoracle
        ta = time.perf_counter()
        ref = pyoracle.run_step(wl, threads=threads)
        tc = time.perf_counter() - ta
        c1.update({"cpu_seconds": round(tc, 3), "cpu_threads": threads, "cpu_sweeps": ref["sweeps"],
                   "cpu_Mvertex_sweeps_per_s": round(nA * int(sum(ref["sweeps"])) / tc / 1e6, 1),
                   "speedup": round(tc / tg, 2),
                   "relL2_vs_oracle": {k: rel(g[k], ref[k]) for k in ("U_SSA", "V_SSA", "Hi")},
                   "sweeps_equal": list(g["sweeps"]) == list(ref["sweeps"])})
    log(f"[ref] dome 8k: {c1}")
    out["dome8k_step"] = c1
    # C2: marine 200k, FP64 reference SSA
    wl = geometry.marine(200_000, seed=2, picard=30, RN_tol=1e-6)
    st = Step(ctx, wl)
    ctx.update_general_ice_model_data()
    ctx.solve_SIA()
    ctx.synchronize()
    ta = time.perf_counter()
    s2 = ctx.solve_SSA()
    tg = time.perf_counter() - ta
    nA = ctx.sizes["nVAaAc"]
    sweeps = list(s2.sweeps)[: s2.picard_iterations]
    c2 = {"workload": f"marine nV={wl.mesh.nV} (nVAaAc={nA}), solve_SSA: Picard to RN 1e-6 (at most 30), "
                      f"reference SOR and stopping rules, FP64",
          "gpu_seconds": round(tg, 3), "picard": s2.picard_iterations, "picard_converged": bool(s2.picard_converged),
          "sweeps_total": int(sum(sweeps)),
          "gpu_Mvertex_sweeps_per_s": round(nA * sum(sweeps) / tg / 1e6, 1)}
    log(f"[ref] marine 200k SSA: {c2}")
    if not args.no_cpu:
        from oracle import pyoracle
        ctx.ssa_prepare()
        om = pyoracle.mesh_from_arrays(wl.mesh, st.ma.arr, st.ma.colour_nV)
        fields = {k: ctx.get(k) for k in ("Hi_AaAc", "Hb_AaAc", "SL_AaAc", "eta_AaAc", "S_AaAc",
                                          "dHs_dx_shelf_AaAc", "dHs_dy_shelf_AaAc", "U_SSA_AaAc", "V_SSA_AaAc")}
        pyoracle.time_sor_sweeps(om, fields, wl.params, 2, threads)
        dt, n = pyoracle.time_sor_sweeps(om, fields, wl.params, 200, threads)
        rate = nA * n / dt / 1e6
        c2.update({"cpu_Mvertex_sweeps_per_s": round(rate, 1), "cpu_threads": threads,
                   "cpu_sample": f"{n} oracle SOR sweeps of this system, {dt:.1f}s",
                   "cpu_seconds_estimated": round(nA * sum(sweeps) / rate / 1e6, 1),
                   "speedup_estimated": round(nA * sum(sweeps) / rate / 1e6 / tg, 1)})
    log(f"[ref] marine 200k SSA cpu: {c2}")
    out["marine200k_ssa"] = c2
    return out


def bench_batch(args, wls):
    """Batched ensemble: independent SSA solves (K Picard steps each) on seeded
    meshes with sizes log-uniform in [batch-min, batch-max], dome or marine,
    A uniform in [1e-17, 1e-16], log-normal friction; every linearised system to
    the relative residual --tol in mixed precision (with the --pc preconditioner,
    the members' union aggregated as one mesh for multigrid), one hipGraph launch
    per GMRES cycle of all members, the sweep's launch shape autotuned; error per
    member against the same batch in FP64 with SOR sweeps as preconditioner."""
    import numpy as np
    from ufemism_amd import abi, batch

    for wl in wls:
        wl.params.SSA_RN_tol = 0.0
    ctx = abi.Context(0)
    t1 = time.time()
    B = batch.Batch(ctx, wls)
    log(f"[batch] union nVAaAc={B.nA} built and uploaded in {time.time() - t1:.1f}s")
    B.reset()
    ctx.update_general_ice_model_data()
    ctx.ssa_prepare()
    tune = ctx.sweep_autotune(10)
    cands = tune.candidates()
    log(f"[batch] autotune: {tune.threads_per_block} threads, schedule {tune.schedule}, loads {tune.loads}")

    def run(prec, graph=True):
        B.reset()
        ctx.update_general_ice_model_data()
        ctx.synchronize()
        ta = time.perf_counter()
        mixed = prec == abi.UFE_MIXED       # the multigrid preconditioner needs the mixed path
        st, rel = ctx.batch_solve_SSA(args.batch_picard, rel_tol=args.tol, precision=prec,
                                      gmres_restart=args.gmres_pc if mixed else args.gmres_sor,
                                      precond_sweeps=args.pc_sweeps if mixed else args.sor_sweeps,
                                      use_graph=graph,
                                      precond_type=args.pc_type if mixed else abi.UFE_PRECOND_SOR)
        ctx.synchronize()
        el = time.perf_counter() - ta
        return st, rel, el, B.member_values("U_SSA_AaAc"), B.member_values("V_SSA_AaAc")

    run(abi.UFE_MIXED)                                   # warm-up: capture and allocations
    st, rel, el, U1, V1 = run(abi.UFE_MIXED)
    log(f"[batch] mixed: {el:.3f}s cycles {st.cycles} sweeps/member {st.sweeps} converged {st.converged}")
    st0, rel0, el0, U0, V0 = run(abi.UFE_FP64)
    log(f"[batch] fp64: {el0:.3f}s cycles {st0.cycles} sweeps/member {st0.sweeps}")
    errs = [float(np.sqrt((np.sum((U1[b] - U0[b]) ** 2) + np.sum((V1[b] - V0[b]) ** 2)) /
                          max(np.sum(U0[b] ** 2 + V0[b] ** 2), 1e-300))) for b in range(len(wls))]
    rows = B.nA
    ctx.close()
    return {
        "workload": f"{len(wls)} independent SSA solves ({args.batch_picard} Picard steps each) on seeded "
                    f"meshes of {args.batch_min}-{args.batch_max} vertices (log-uniform), dome or marine, "
                    f"A ~ U[1e-17, 1e-16], log-normal friction; linear tolerance {args.tol:g}",
        "members": len(wls), "vertices": int(sum(w.mesh.nV for w in wls)), "nVAaAc": int(rows),
        "member_nV_min": int(min(w.mesh.nV for w in wls)), "member_nV_max": int(max(w.mesh.nV for w in wls)),
        "preconditioner": args.pc,
        "seconds": round(el, 4), "fp64_seconds": round(el0, 4), "speedup_vs_fp64": round(el0 / el, 3),
        "gmres_cycles": st.cycles, "sweeps_per_member": st.sweeps, "converged": st.converged,
        "aggregate_Mvertex_sweeps_per_s": round(rows * st.sweeps / el / 1e6, 1),
        "solves_per_s": round(len(wls) / el, 2),
        "max_member_rel_residual": float(np.max(rel)),
        "member_velocity_error_vs_fp64": {"max": max(errs), "median": float(np.median(errs))},
        "tuning": {"threads_per_block": tune.threads_per_block, "schedule": tune.schedule,
                   "loads": tune.loads, "candidates": cands},
        "graph": "one hipGraph launch per GMRES cycle of all members",
    }


def bench_adjoint(args, ctx):
    """Adjoint of K Picard steps on the marine mesh: J = 1/2 |u - u_obs|^2 with
    u_obs from a friction field perturbed log-normally (sigma 0.3); times the
    forward and the backward and checks the gradient by central differences."""
    import numpy as np
    import torch
    from ufemism_amd import geometry
    from ufemism_amd.adjoint import AdjointOptions, SSAProblem, ssa_velocities
    from ufemism_amd.dynamics import Step

    t0 = time.time()
    wl = geometry.marine(n_target=args.adj_nv, seed=args.seed + 1, picard=args.adj_picard)
    step = Step(ctx, wl)
    ctx.update_general_ice_model_data()
    log(f"[adjoint] mesh nV={wl.mesh.nV} nVAaAc={ctx.sizes['nVAaAc']} set up in {time.time() - t0:.1f}s")
    mg = args.pc == "multigrid"       # FP64 Krylov around the FP32 cycle, both directions
    prob = SSAProblem(ctx, AdjointOptions(picard_iterations=args.adj_picard, lin_rel_tol=args.adj_tol,
                                          checkpoint_every=args.adj_every, precond_type=args.pc_type,
                                          gmres_restart=args.gmres_pc if mg else 40,
                                          precond_sweeps=args.pc_sweeps if mg else 5))
    dev = prob.dev
    VA = step.ma["VAaAc"]
    log_tau0 = torch.log(torch.tensor(wl.tau_c_override(VA), dtype=torch.float64, device=dev))
    A0 = torch.full((prob.nA,), wl.A_flow, dtype=torch.float64, device=dev)
    g = torch.tensor(geometry.gaussian_field(VA, 50e3, args.seed + 4242), dtype=torch.float64, device=dev)
    U_obs, V_obs = prob.forward(log_tau0 + 0.3 * g, A0)
    log(f"[adjoint] observations: {prob.forward_stats.forward_seconds:.2f}s")

    def grad():
        lt = log_tau0.clone().requires_grad_(True)
        a = A0.clone().requires_grad_(True)
        torch.cuda.synchronize()
        ta = time.perf_counter()
        U, V = ssa_velocities(lt, a, prob)
        torch.cuda.synchronize()
        tb = time.perf_counter()
        J = 0.5 * (((U - U_obs) ** 2).sum() + ((V - V_obs) ** 2).sum())
        J.backward()
        torch.cuda.synchronize()
        tc = time.perf_counter()
        return lt.grad, a.grad, tb - ta, tc - tb, float(J)

    grad()                                          # warm-up (allocations)
    g_tau, g_A, tf, tb, J = grad()
    fs, bs = prob.forward_stats, prob.backward_stats
    log(f"[adjoint] forward {tf:.3f}s backward {tb:.3f}s ratio {tb / tf:.2f}")
    h = 1e-6
    # The central differences need J(+) - J(-), about 2e-6 dJ, free of solver
    # noise: their forwards use FP64 GMRES around FP64 SOR sweeps to 1e-12, a
    # map smooth in the friction field (an FP32 preconditioner rounds each
    # correction, which leaves noise of the size of the tolerance in J).
    import dataclasses
    fd_tol = 1e-12
    prob.opts = dataclasses.replace(prob.opts, lin_rel_tol=fd_tol, precond_type=0, gmres_restart=40,
                                    precond_sweeps=5)
    gnorm = float(torch.linalg.norm(g_tau))
    gen = torch.Generator(device="cpu").manual_seed(args.seed)
    errs = []
    for k in range(args.adj_fd):
        d = torch.randn(prob.nA, generator=gen, dtype=torch.float64).to(dev)
        ad = float((g_tau * d).sum())
        Up, Vp = prob.forward(log_tau0 + h * d, A0)
        Um, Vm = prob.forward(log_tau0 - h * d, A0)
        fd = 0.5 * float(((Up - Um) * (Up + Um - 2 * U_obs)).sum() +
                         ((Vp - Vm) * (Vp + Vm - 2 * V_obs)).sum()) / (2 * h)
        errs.append(abs(fd - ad) / gnorm)
        log(f"[adjoint] FD direction {k}: adjoint {ad:+.9e} fd {fd:+.9e} |fd - ad| / |g| {errs[-1]:.1e}")
    return {
        "workload": f"marine ice sheet nV={wl.mesh.nV} (nVAaAc={prob.nA}); J = 1/2 |u - u_obs|^2 over "
                    f"AaAc (u, v) after {args.adj_picard} Picard steps, u_obs from friction perturbed "
                    f"log-normally (sigma 0.3); gradient w.r.t. log tau_c and A",
        "picard_iterations": args.adj_picard, "lin_rel_tol": args.adj_tol,
        "checkpoint_every": args.adj_every, "checkpoints": fs.checkpoints,
        "forward_seconds": round(tf, 4), "backward_seconds": round(tb, 4),
        "adjoint_forward_ratio": round(tb / tf, 3),
        "forward_sweeps": fs.forward_sweeps, "adjoint_sweeps": bs.adjoint_sweeps,
        "recompute_sweeps": bs.recompute_sweeps,
        "max_forward_residual": fs.max_forward_residual, "max_adjoint_residual": bs.max_adjoint_residual,
        "preconditioner": args.pc,
        "J": J, "fd_directions": args.adj_fd, "fd_step": h, "fd_lin_rel_tol": fd_tol,
        "fd_error_definition": "|fd - g.d| / |g| (d has unit normal entries, so g.d is typically |g|)",
        "fd_forward_preconditioner": "sor",
        "fd_max_error": max(errs) if errs else None,
        "fd_errors": [float(f"{e:.3e}") for e in errs],
    }


if __name__ == "__main__":
    main()
