#!/usr/bin/env python
"""Benchmark: linearised SSA solve to a 1e-6 relative residual on one MI355X.

Workload (default): the marine ice-sheet configuration at ~1M vertices
(seeded Delaunay mesh on a 1600 km square, ~4M combined Aa+Ac SOR points).
Set-up, untimed: mesh generation, secondary mesh data, one general update,
SIA, and `--picard0` Picard iterations of the reference SSA (FP64, reference
stopping rule) to reach a representative viscosity.  The linear system is
then frozen (viscosity, sliding, right-hand side).

One step = reset the velocities to zero, then solve the linear system to
||b - A w||_2 / ||b||_2 <= 1e-6 (FP64 norm), in FP64 sweeps or in mixed
precision (FP32 sweeps, FP64 iterative refinement).

Metric: Mvertex-sweeps/s = nVAaAc * sweeps / time / 1e6 (whole job, one GPU).
The roofline object is for the dominant kernel (the SOR colour pass), timed
live with HIP events on the library's stream; the CPU baseline is the CPU
oracle timed on a bounded number of sweeps of the same system.
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
    ap.add_argument("--omega-mixed", type=float, default=0.0, help="inner omega of the mixed path")
    ap.add_argument("--check", type=int, default=32)
    ap.add_argument("--max-sweeps", type=int, default=200000)
    ap.add_argument("--mode", default="both", choices=["both", "fp64", "mixed"])
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(json.dumps({"error": "this build runs on one GPU only; multi-GPU sharding is "
                                   "design-only (see DESIGN.md); run with --gpus 1"}))
        sys.exit(2)

    import numpy as np
    from ufemism_amd import abi, geometry
    from ufemism_amd.dynamics import Step

    t0 = time.time()
    wl = geometry.marine(n_target=args.nv, seed=args.seed, picard=args.picard0)
    if args.omega > 0:
        wl.params.SSA_SOR_omega = args.omega
    log(f"[bench] mesh nV={wl.mesh.nV} generated in {time.time() - t0:.1f}s")
    ctx = abi.Context(0)
    t1 = time.time()
    step = Step(ctx, wl)
    sizes = ctx.sizes
    log(f"[bench] mesh build+upload {time.time() - t1:.1f}s sizes={sizes}")
    ctx.update_general_ice_model_data()
    ctx.solve_SIA()
    t2 = time.time()
    st0 = ctx.solve_SSA()
    ref_ssa_s = time.time() - t2
    log(f"[bench] reference SSA ({st0.picard_iterations} Picard, sweeps "
        f"{list(st0.sweeps)[:st0.picard_iterations]}) {ref_ssa_s:.2f}s")
    ctx.ssa_prepare()
    nA = sizes["nVAaAc"]
    sb = ctx.sweep_bytes()
    omega = wl.params.SSA_SOR_omega
    omega_mixed = args.omega_mixed if args.omega_mixed > 0 else omega

    def solve(mode):
        ctx.fill("U_SSA_AaAc", 0.0)
        ctx.fill("V_SSA_AaAc", 0.0)
        return ctx.ssa_linear_solve(args.tol, args.max_sweeps,
                                    precision=abi.UFE_MIXED if mode == "mixed" else abi.UFE_FP64,
                                    omega=omega_mixed if mode == "mixed" else omega,
                                    check_interval=args.check)

    modes = ["fp64", "mixed"] if args.mode == "both" else [args.mode]
    results = {}
    W = {}
    for mode in modes:
        for _ in range(args.warmup):
            solve(mode)
        ctx.profile_read()            # discard
        ctx.profile_enable(True)
        ctx.synchronize()
        ta = time.perf_counter()
        sweeps, last = 0, None
        for _ in range(args.steps):
            last = solve(mode)
            sweeps += last.sweeps
        ctx.synchronize()
        tb = time.perf_counter()
        ctx.profile_enable(False)
        prof = ctx.profile_read()
        el = tb - ta
        W[mode] = (ctx.get("U_SSA_AaAc"), ctx.get("V_SSA_AaAc"))
        bytes_per_sweep = sb["fp64"] if mode == "fp64" else sb["fp32"]
        results[mode] = dict(
            value=nA * sweeps / el / 1e6, ms_per_step=1e3 * el / args.steps,
            sweeps_per_solve=last.sweeps, rel_residual=last.rel_residual,
            refinements=last.refinements, converged=last.status == 0,
            kernel_launches=prof.launches,
            kernel_avg_us=1e6 * prof.seconds / max(prof.launches, 1),
            kernel_bytes_per_launch=prof.algorithmic_bytes / max(prof.launches, 1),
            kernel_GBps=prof.algorithmic_bytes / max(prof.seconds, 1e-30) / 1e9,
            sweep_time_share=prof.seconds / el,
            algorithmic_bytes_per_sweep=bytes_per_sweep)
        log(f"[bench] {mode}: {results[mode]}")

    head = "mixed" if "mixed" in results else "fp64"
    R = results[head]
    out = {
        "metric": METRIC, "value": round(R["value"], 3), "unit": "Mvertex-sweeps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(R["ms_per_step"], 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32+f64" if head == "mixed" else "f64", "data": "synthetic",
        "config": {
            "workload": f"marine ice sheet, seeded Delaunay mesh nV={sizes['nV']} "
                        f"(nVAaAc={nA}) on a 1600 km square; linearised SSA after "
                        f"{st0.picard_iterations} FP64 Picard iterations, solved from zero to "
                        f"relative residual {args.tol:g}",
            "nV": sizes["nV"], "nVAaAc": nA, "mean_degree": round(sb["mean_degree"], 4),
            "omega": omega, "omega_mixed": omega_mixed, "check_interval": args.check,
            "path": head,
        },
        "roofline": {
            "bound": "hbm", "achieved": round(R["kernel_GBps"], 2), "peak": HBM_PEAK / 1e9,
            "unit": "GB/s", "frac": round(R["kernel_GBps"] * 1e9 / HBM_PEAK, 4), "traffic": None,
            "kernel": "k_sor_colour_f32" if head == "mixed" else "k_sor_colour",
            "bytes_per_launch": round(R["kernel_bytes_per_launch"]),
            "avg_launch_us": round(R["kernel_avg_us"], 3),
        },
        "paths": results,
        "reference_ssa_seconds": round(ref_ssa_s, 3),
    }
    if "fp64" in results and "mixed" in results:
        out["speedup_mixed_vs_fp64_time"] = round(results["fp64"]["ms_per_step"] /
                                                  results["mixed"]["ms_per_step"], 3)
        u64, v64 = W["fp64"]
        u32, v32 = W["mixed"]
        ref = np.sqrt(np.sum(u64 ** 2 + v64 ** 2))
        out["mixed_error_vs_fp64_relL2"] = float(np.sqrt(np.sum((u32 - u64) ** 2 + (v32 - v64) ** 2)) / ref)

    if not args.no_cpu:
        try:
            from oracle import pyoracle
            threads = min(16, os.cpu_count() or 1)
            om = pyoracle.mesh_from_arrays(wl.mesh, step.ma.arr, step.ma.colour_nV)
            fields = {k: ctx.get(k) for k in ("Hi_AaAc", "Hb_AaAc", "SL_AaAc", "eta_AaAc", "S_AaAc",
                                              "dHs_dx_shelf_AaAc", "dHs_dy_shelf_AaAc")}
            fields["U_SSA_AaAc"] = np.zeros(nA)
            fields["V_SSA_AaAc"] = np.zeros(nA)
            dt1, n1 = pyoracle.time_sor_sweeps(om, fields, wl.params, 2, threads)
            nsw = max(2, int(args.cpu_seconds / max(dt1 / max(n1, 1), 1e-6)))
            dt, n = pyoracle.time_sor_sweeps(om, fields, wl.params, nsw, threads)
            out["cpu_baseline"] = {
                "value": round(nA * n / dt / 1e6, 3), "unit": "Mvertex-sweeps/s", "cores": threads,
                "kind": "port",
                "sample": f"{n} FP64 SOR sweeps of the CPU oracle (OpenMP, {threads} threads) "
                          f"on the same {nA}-point linear system, {dt:.1f}s"}
        except Exception as e:  # the CPU baseline is reported, never required
            out["cpu_baseline"] = {"value": None, "error": repr(e)}
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
