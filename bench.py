"""Benchmark: RK4 cell-stage updates per second of the FP64 viscous
second-order flat plate (256x128x64 cells, 2.1M) on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload flat_plate_c2]

A step is one RK4 time step (4 stages) over the whole mesh, inputs resident
in HBM.  value = cells * 4 * K / (wall time of K steps), timed between two
stream synchronisations.  The JSON line also carries:
  roofline      the dominant kernel's algorithmic bytes per launch divided by
                its mean launch time (HIP events on the solver's stream),
                against the 8 TB/s HBM peak; traffic = FETCH+WRITE bytes per
                launch from the committed rocprofv3 counter summary of this
                workload, when present;
  cpu_baseline  the CPU oracle (a single-threaded port of the reference's
                serial arithmetic) timed on a bounded sample of the same
                physics on this host.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def algorithmic_bytes(kernel: str, nc: int, nif: int, nbf: int, second_order: bool) -> float:
    """Compulsory HBM bytes of one launch: every array the kernel needs read
    once, every output written once (FP64 values, int32 indices)."""
    F = nif + nbf
    if kernel == "k_flux":
        # V 40, grad 120, phi 40, x_c 24, vol 8, U_acc 40 read; U_acc 40, R 40 written per
        # cell; left/right 8, n/t/b 72, x_f 24 per face.
        if second_order:
            return 352.0 * nc + 104.0 * F
        return 168.0 * nc + 80.0 * F       # V, vol, U_acc in; U_acc, R out; conn, n/t/b
    if kernel in ("k_grad", "k_grad_limiter"):
        # V 40, vol 8, x_c 24 read; grad 120 (+ phi 40) written; conn 8, n 24, x_f 24 per face
        return (232.0 if kernel == "k_grad_limiter" else 192.0) * nc + 56.0 * F
    if kernel == "k_prim":
        return 120.0 * nc                  # U_n, R read; V written
    if kernel == "k_dt_partial":
        return 48.0 * nc + 24.0 * F
    return 0.0


def cpu_baseline(opts) -> dict:
    """Times the CPU oracle on a bounded sample: the same physics and options
    on a 128x64x32 mesh (1/8 of the cells), 2 RK4 steps."""
    import dataclasses

    from oracle import oracle as O

    kw = {f.name: getattr(opts, f.name) for f in dataclasses.fields(opts)
          if f.name not in ("output_results", "output_frequency")}
    kw.update(nx=max(1, opts.nx // 2), ny=max(1, opts.ny // 2), nz=max(1, opts.nz // 2))
    steps = 2
    m = O.OracleMesh(O.make_options(**kw))
    s = O.OracleSolver(m)
    s.init()
    t0 = time.perf_counter()
    s.step(steps)
    dt = time.perf_counter() - t0
    return {"value": m.ncells * 4 * steps / dt, "unit": "cell-stage updates/s", "cores": 1,
            "kind": "port",
            "sample": f"oracle (single-thread C++ restatement of the reference's serial path), "
                      f"{kw['nx']}x{kw['ny']}x{kw['nz']} = {m.ncells} cells, {steps} RK4 steps, "
                      f"same physics/options, {dt:.1f} s"}


def load_traffic(workload: str, kernel: str):
    p = os.path.join(ROOT, "profiles", f"{kernel}_{workload}_traffic.json")
    if not os.path.exists(p):
        return None
    try:
        with open(p) as f:
            return float(json.load(f)["hbm_bytes_per_launch"])
    except Exception:
        return None


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="flat_plate_c2")
    ap.add_argument("--profile-steps", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--precision", type=int, default=None,
                    help="override the workload's arithmetic (0 FP64, 1 FP32 flux/gradient/limiter)")
    ap.add_argument("--compare-fp64", action="store_true",
                    help="also run the workload in FP64 and report the per-variable relative L2 error")
    ap.add_argument("--copy-ceiling", action="store_true",
                    help="only measure the plain copy kernel (1 GiB, read + write) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU runs are not part of it "
                                   "(cases and mesh blocks shard across GPUs as described in DESIGN.md)"}))
        return 2

    import miniaero_amd as M
    from miniaero_amd import _lib
    from miniaero_amd.configs import WORKLOADS

    copy_bytes = 1 << 30
    copy_ms = _lib.measure_copy(copy_bytes, 10)
    copy_gbs = 2 * copy_bytes / (copy_ms * 1e-3) / 1e9
    if args.copy_ceiling:
        print(json.dumps({"copy_bytes": copy_bytes, "ms_per_copy": copy_ms, "GBps_read_plus_write": copy_gbs}))
        return 0

    opts = WORKLOADS[args.workload]()
    if args.precision is not None:
        opts = opts.replace(precision=args.precision)
    t0 = time.perf_counter()
    md = M.Parallel3DMesh(opts.nx, opts.ny, opts.nz, opts.lx, opts.ly, opts.lz, opts.problem_type,
                          opts.angle, options=opts).fillMeshData()
    solver = M.TimeSolverExplicitRK4(md, opts)
    setup_s = time.perf_counter() - t0
    sizes = solver.sizes()
    nc, nif, nbf = sizes["cells"], sizes["internal_faces"], sizes["boundary_faces"]
    print(f"[bench] {args.workload}: {nc} cells, {nif} internal + {nbf} boundary faces, "
          f"setup {setup_s:.1f} s, device {sizes['device_bytes'] / 2**30:.2f} GiB", file=sys.stderr,
          flush=True)

    solver.initialize()
    solver.step(args.warmup)
    solver.synchronize()
    t0 = time.perf_counter()
    solver.step(args.steps)
    solver.synchronize()
    elapsed = time.perf_counter() - t0
    if solver.nonfinite():
        print("[bench] non-finite state after the timed steps", file=sys.stderr)
        return 1
    value = nc * 4 * args.steps / elapsed
    print(f"[bench] {args.steps} steps in {elapsed:.3f} s", file=sys.stderr, flush=True)

    precision_error = None
    if args.compare_fp64 and opts.precision != 0:
        # same steps from the same start in FP64; error of the timed run's state
        import numpy as np
        U = solver.state()
        ref = M.TimeSolverExplicitRK4(md, opts.replace(precision=0))
        ref.initialize()
        ref.step(args.warmup + args.steps)
        R = ref.state()
        ref.close()
        mom = np.linalg.norm(R[:, 1:4])
        scale = [np.linalg.norm(R[:, 0]), mom, mom, mom, np.linalg.norm(R[:, 4])]
        names = ("rho", "rho_u", "rho_v", "rho_w", "rho_E")
        precision_error = {
            "metric": "relative L2 vs the FP64 run after warmup+steps; momentum components "
                      "scaled by the momentum-vector norm",
            "steps": args.warmup + args.steps,
            "whole_state": float(np.linalg.norm(U - R) / np.linalg.norm(R)),
            **{n: float(np.linalg.norm(U[:, k] - R[:, k]) / scale[k]) for k, n in enumerate(names)},
            "raw_per_component": {n: float(np.linalg.norm(U[:, k] - R[:, k]) / max(np.linalg.norm(R[:, k]), 1e-300))
                                  for k, n in enumerate(names)}}
        print(f"[bench] precision error: {json.dumps(precision_error)}", file=sys.stderr, flush=True)

    prof = solver.profile(args.profile_steps)
    so = bool(opts.second_order_space)
    dom = max(prof, key=lambda k: prof[k][0] * prof[k][1])
    launches, mean_ms = prof[dom]
    alg = algorithmic_bytes(dom, nc, nif, nbf, so)
    achieved = alg / (mean_ms * 1e-3) / 1e9
    kernels = {k: {"launches": v[0], "mean_ms": round(v[1], 5),
                   "GBps": round(algorithmic_bytes(k, nc, nif, nbf, so) / (v[1] * 1e-3) / 1e9, 1)}
               for k, v in prof.items()}
    print(f"[bench] kernels: {json.dumps(kernels)}", file=sys.stderr, flush=True)
    traffic = load_traffic(args.workload, dom)

    cpu = None
    if not args.no_cpu_baseline:
        try:
            cpu = cpu_baseline(opts)
        except Exception as e:  # the oracle is optional on a machine without a compiler
            print(f"[bench] cpu baseline skipped: {e}", file=sys.stderr)

    t_final, dt_last = solver.time()
    solver.close()
    line = {
        "metric": "RK4 cell-stage updates/s",
        "value": value,
        "unit": "cell-stage updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if opts.precision == 0 else "f32 arithmetic, f64 state",
        "data": "synthetic",
        "config": {"workload": args.workload, "mesh": f"{opts.nx}x{opts.ny}x{opts.nz}", "cells": nc,
                   "internal_faces": nif, "boundary_faces": nbf,
                   "physics": "viscous Roe, Green-Gauss, Venkat, 2nd order" if so else "first order",
                   "cfl": opts.cfl, "last_dt": dt_last,
                   "precision": "fp64" if opts.precision == 0 else "fp32 flux/gradient/limiter, fp64 state"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": dom, "algorithmic_bytes_per_launch": alg,
                     "mean_launch_ms": round(mean_ms, 5), "copy_ceiling_GBps": round(copy_gbs, 1),
                     "kernels": kernels},
        "cpu_baseline": cpu,
    }
    if precision_error is not None:
        line["precision_error"] = precision_error
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
