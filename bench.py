"""Benchmark: RK4 cell-stage updates per second of the FP64 viscous
second-order flat plate (256x128x64 cells, 2.1M) on one MI355X, with the
other headline figures of the path in the same JSON line.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload suite]

A step is one RK4 time step (4 stages) over the whole mesh, inputs resident
in HBM.  value = cells * 4 * K / (wall time of K steps), timed between two
stream synchronisations.  The JSON line also carries:
  roofline      the dominant kernel's algorithmic bytes per launch divided by
                its mean launch time (HIP events on the solver's stream),
                against the 8 TB/s HBM peak; traffic = FETCH+WRITE bytes per
                launch from the committed rocprofv3 counter summary of this
                workload, when present; path = the whole step's algorithmic
                bytes over its wall time;
  cpu_baseline  the CPU oracle (a single-threaded port of the reference's
                serial arithmetic) timed on a bounded sample of the same
                physics on this host's cores: one oracle process per core
                (at most 16), all running the sample at once; the value is
                their aggregate rate;
  workloads     (default suite) the FP64 path on the 512x256x64 plate (8.4M
                cells) with its roofline, the precision form on that mesh
                (500 steps, per-variable error against FP64), the 512-case
                batched form with its per-case latency, the first 200
                requests of the served stream with p50/p99 latency, and the
                3D Sod case run in full on one CPU core and on the GPU.
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


def algorithmic_bytes(kernel: str, nc: int, nif: int, nbf: int, second_order: bool,
                      t: int = 8, g: int = 8, face_values: bool = False) -> float:
    """Compulsory HBM bytes of one launch: every array the kernel needs read
    once, every output written once.  t = bytes of a primitive, gradient or
    limiter value (8 FP64, 4 in the FP32-arithmetic form), g = bytes of a
    geometry value (cell centroid and volume, face area/tangent/binormal,
    face centroid; also 4 in the FP32-arithmetic form); the state U, the
    RK accumulator and the residual are always FP64 (40 B per cell each);
    connectivity is 8 B per face.  face_values: the inviscid second-order
    scheme, where the gradient kernel writes each cell's extrapolated value
    at each of its 6 faces (30 values per cell; gradients and limiters are not
    stored while stepping) and the flux kernel reads 5 of them per face side
    instead of the cells' primitives, gradients and limiters."""
    F = nif + nbf
    if kernel == "k_flux":
        # V, grad, phi, x_c, vol, U_acc read and U_acc written per cell;
        # stages 0-2 also read U_n (the same array as U_acc's input on stage
        # 0), every stage writes the next stage's V (stage 3 the next
        # step's), stage 3 writes R: averaged over the four launches of a
        # step, 110 + 5 t bytes per cell; per face left/right,
        # area/tangent/binormal, x_f
        per_cell = 110.0 + 5.0 * t
        if face_values:
            return (g + per_cell) * nc + (8 + 9 * g + 10 * t) * F
        if second_order:
            return (25 * t + 4 * g + per_cell) * nc + (8 + 12 * g) * F
        return (5 * t + g + per_cell) * nc + (8 + 9 * g) * F
    if kernel in ("k_grad", "k_grad_limiter"):
        # V, vol, x_c read; grad (+ phi) written; left/right, area, x_f per face
        vals = 25 if kernel == "k_grad_limiter" else 20
        if face_values:
            vals = 5 + 30  # V read; face values written
        return (vals * t + 4 * g) * nc + (8 + 6 * g) * F
    if kernel == "k_prim":
        return (80.0 + 5 * t) * nc          # U_n, R read; V written
    if kernel == "k_dt_partial":
        # U_n, the cell's 5 summed face-geometry values and vol read per cell
        return 88.0 * nc
    return 0.0


def path_bytes_per_step(nc: int, nif: int, nbf: int, second_order: bool, viscous: bool,
                        fp32: bool = False) -> float:
    """Compulsory HBM bytes of one RK4 step of the whole path (4 stages):
    the per-stage figures of the fused target design -- 824 N_c + 168 F
    (second order, inviscid or viscous), 536 N_c + 112 F (first-order
    viscous), 208 N_c + 80 F (first-order inviscid); FP64 sizes.  fp32: the
    precision form's second-order figure with gradients, limiters and
    geometry at 4 bytes and the state, accumulator and residual at 8 --
    572 N_c + 96 F (pass A 184 N_c + 20 F, B 132 N_c + 20 F, C 256 N_c + 56 F)."""
    F = nif + nbf
    if second_order and fp32:
        per = 572.0 * nc + 96.0 * F
    elif second_order:
        per = 824.0 * nc + 168.0 * F
    elif viscous:
        per = 536.0 * nc + 112.0 * F
    else:
        per = 208.0 * nc + 80.0 * F
    return 4.0 * per


def box_faces(nx: int, ny: int, nz: int):
    """Cells, internal faces and boundary faces of an nx x ny x nz box mesh."""
    nc = nx * ny * nz
    nif = (nx - 1) * ny * nz + nx * (ny - 1) * nz + nx * ny * (nz - 1)
    nbf = 2 * (ny * nz + nx * nz + nx * ny)
    return nc, nif, nbf


def _oracle_run(cases, steps: int):
    """The CPU oracle over the given option dicts, one after another for
    `steps` RK4 steps each: (cells, seconds of stepping)."""
    from oracle import oracle as O

    cells, secs = 0, 0.0
    for kw in cases:
        m = O.OracleMesh(O.make_options(**kw))
        s = O.OracleSolver(m)
        s.init()
        t0 = time.perf_counter()
        s.step(steps)
        secs += time.perf_counter() - t0
        cells += m.ncells
    return cells, secs


def _cpu_cores() -> int:
    """Host cores this process may run on, at most 16 (one GPU's share of the host)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def _oracle_rate(cases, steps: int):
    """Aggregate cell-stage updates/s of the CPU oracle on all usable host
    cores: one single-threaded oracle process per core (started as child
    processes, which touch no GPU), each running every case of `cases` for
    `steps` RK4 steps; total updates over the slowest process's stepping
    time.  Returns (rate, cells per process, seconds, processes)."""
    import dataclasses
    import subprocess

    kws = [{f.name: getattr(o, f.name) for f in dataclasses.fields(o)
            if f.name not in ("output_results", "output_frequency")} for o in cases]
    spec = json.dumps({"cases": kws, "steps": steps})
    n = _cpu_cores()
    print(f"[bench] cpu baseline: {n} oracle processes, {len(cases)} case(s) x {steps} steps each",
          file=sys.stderr, flush=True)
    procs = [subprocess.Popen([sys.executable, os.path.abspath(__file__), "--cpu-worker"], stdin=subprocess.PIPE,
                              stdout=subprocess.PIPE, text=True) for _ in range(n)]
    for p in procs:  # every worker gets its input before any output is awaited, so all run at once
        p.stdin.write(spec)
        p.stdin.close()
    outs = [p.stdout.read() for p in procs]
    for p in procs:
        p.wait()
    if any(p.returncode != 0 for p in procs):
        raise RuntimeError("an oracle worker failed")
    res = [json.loads(o.strip().splitlines()[-1]) for o in outs]
    cells = res[0]["cells"]
    secs = max(r["secs"] for r in res)
    return n * cells * 4 * steps / secs, cells, secs, n


def cpu_worker() -> int:
    """--cpu-worker: run the oracle on the cases given as JSON on stdin and
    print {"cells", "secs"}."""
    spec = json.loads(sys.stdin.read())
    cells, secs = _oracle_run(spec["cases"], spec["steps"])
    print(json.dumps({"cells": cells, "secs": secs}), flush=True)
    return 0


def cpu_baseline(opts) -> dict:
    """Times the CPU oracle on a bounded sample: the same physics and options
    on a mesh with half the cells in each direction (1/8 of the cells), 5
    RK4 steps."""
    steps = 5
    sample = opts.replace(nx=max(1, opts.nx // 2), ny=max(1, opts.ny // 2), nz=max(1, opts.nz // 2))
    rate, cells, secs, n = _oracle_rate([sample], steps)
    return {"value": rate, "unit": "cell-stage updates/s", "cores": n, "kind": "port",
            "sample": f"oracle (single-thread C++ restatement of the reference's serial path), one process "
                      f"per core on {n} cores, each {sample.nx}x{sample.ny}x{sample.nz} = {cells} cells, "
                      f"{steps} RK4 steps, same physics/options, {secs:.1f} s"}


def cpu_baseline_batch(cases, min_cells: int = 600_000, steps: int = 3) -> dict:
    """The CPU oracle over the batch's first cases, one after another, until
    they hold min_cells cells, `steps` RK4 steps each."""
    chosen, n = [], 0
    for o in cases:
        chosen.append(o)
        n += o.nx * o.ny * o.nz
        if n >= min_cells:
            break
    rate, cells, secs, n = _oracle_rate(chosen, steps)
    return {"value": rate, "unit": "cell-stage updates/s", "cores": n, "kind": "port",
            "sample": f"oracle (single-thread C++ restatement of the reference's serial path), one process "
                      f"per core on {n} cores, each running the first {len(chosen)} cases of the sample "
                      f"({cells} cells), {steps} RK4 steps each, {secs:.1f} s"}


def load_traffic(workload: str, kernel: str):
    p = os.path.join(ROOT, "profiles", f"{kernel}_{workload}_traffic.json")
    if not os.path.exists(p):
        return None
    try:
        with open(p) as f:
            return float(json.load(f)["hbm_bytes_per_launch"])
    except Exception:
        return None


def precision_errors(U, R) -> dict:
    """Relative L2 of U against the FP64 state R: whole state and per
    variable, momentum components scaled by the momentum-vector norm."""
    import numpy as np
    mom = np.linalg.norm(R[:, 1:4])
    scale = [np.linalg.norm(R[:, 0]), mom, mom, mom, np.linalg.norm(R[:, 4])]
    names = ("rho", "rho_u", "rho_v", "rho_w", "rho_E")
    return {"whole_state": float(np.linalg.norm(U - R) / np.linalg.norm(R)),
            **{n: float(np.linalg.norm(U[:, k] - R[:, k]) / scale[k]) for k, n in enumerate(names)}}


def timed_run(solver, steps: int, warmup: int, graph: int = 0, persistent: bool = False):
    solver.initialize()
    if persistent:
        solver.set_persistent(True)
        graph = 0
    if graph:
        solver.set_graph(graph)
    solver.step(warmup)
    if graph:
        # capture the graphs of both ping-pong parities before timing
        solver.step(graph)
        solver.step(1)
        solver.step(graph)
    solver.synchronize()
    t0 = time.perf_counter()
    solver.step(steps)
    solver.synchronize()
    return time.perf_counter() - t0


def roofline(solver, profile_steps, nc, nif, nbf, so, fp32, traffic, copy_gbs, face_values=False):
    prof = solver.profile(profile_steps)
    t = g = 4 if fp32 else 8
    dom = max(prof, key=lambda k: prof[k][0] * prof[k][1])
    launches, mean_ms = prof[dom]
    alg = algorithmic_bytes(dom, nc, nif, nbf, so, t, g, face_values)
    achieved = alg / (mean_ms * 1e-3) / 1e9
    kernels = {k: {"launches": v[0], "mean_ms": round(v[1], 5),
                   "GBps": round(algorithmic_bytes(k, nc, nif, nbf, so, t, g, face_values) / (v[1] * 1e-3) / 1e9,
                                 1)}
               for k, v in prof.items()}
    print(f"[bench] kernels: {json.dumps(kernels)}", file=sys.stderr, flush=True)
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic(dom),
            "kernel": dom, "algorithmic_bytes_per_launch": alg,
            "mean_launch_ms": round(mean_ms, 5), "copy_ceiling_GBps": round(copy_gbs, 1),
            "kernels": kernels}


def bench_single(args, M, opts, copy_gbs) -> dict:
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
    shape, wg, tuned_ms = solver.tuning()
    print(f"[bench] tile shape {shape}, workgroups per CU {wg or 'all tiles'}"
          + (f" (autotuned, {tuned_ms:.3f} ms/step/Mcell)" if tuned_ms else ""), file=sys.stderr, flush=True)
    elapsed = timed_run(solver, args.steps, args.warmup, args.graph, args.persistent)
    if solver.nonfinite():
        raise RuntimeError("non-finite state after the timed steps")
    print(f"[bench] {args.steps} steps in {elapsed:.3f} s", file=sys.stderr, flush=True)
    warm_total = args.warmup + (2 * args.graph + 1 if args.graph else 0)

    extra = {}
    if args.compare_fp64 and opts.precision != 0:
        # same steps from the same start in FP64; error of the timed run's state
        U = solver.state()
        ref = M.TimeSolverExplicitRK4(md, opts.replace(precision=0))
        ref.initialize()
        ref.step(warm_total + args.steps)
        err = precision_errors(U, ref.state())
        ref.close()
        extra["precision_error"] = {
            "metric": "relative L2 vs the FP64 run after warmup+steps; momentum components "
                      "scaled by the momentum-vector norm", "steps": warm_total + args.steps, **err}
        print(f"[bench] precision error: {json.dumps(extra['precision_error'])}", file=sys.stderr, flush=True)

    so = bool(opts.second_order_space)
    roof = roofline(solver, args.profile_steps, nc, nif, nbf, so, opts.precision != 0,
                    lambda k: load_traffic(args.workload, k), copy_gbs,
                    face_values=so and not opts.viscous)
    if opts.precision == 0 or so:
        pb = path_bytes_per_step(nc, nif, nbf, so, bool(opts.viscous), fp32=opts.precision != 0)
        path_gbs = pb / (elapsed / args.steps) / 1e9
        roof["path"] = {"algorithmic_bytes_per_step": pb, "achieved_GBps": round(path_gbs, 1),
                        "frac": round(path_gbs / HBM_PEAK_GBS, 4)}
    cpu = None
    if not args.no_cpu_baseline:
        try:
            cpu = cpu_baseline(opts)
        except Exception as e:  # the oracle is optional on a machine without a compiler
            print(f"[bench] cpu baseline skipped: {e}", file=sys.stderr)
    _, dt_last = solver.time()
    solver.close()
    physics = ("viscous " if opts.viscous else "inviscid ") + "Roe, " + (
        ("Green-Gauss, " + ("Van Albada" if opts.limiter else "Venkat") + ", 2nd order") if so else "1st order")
    config = {"workload": args.workload, "mesh": f"{opts.nx}x{opts.ny}x{opts.nz}", "cells": nc,
              "internal_faces": nif, "boundary_faces": nbf, "physics": physics,
              "cfl": opts.cfl, "last_dt": dt_last,
              "precision": "fp64" if opts.precision == 0 else "fp32 flux/gradient/limiter, fp64 state",
              "tile_shape": list(shape), "workgroups_per_cu": wg,
              "autotuned_ms_per_step_per_Mcell": tuned_ms or None,
              "graph_steps": 0 if args.persistent else args.graph,
              "step_launch": "persistent (one cooperative launch per step() call)" if args.persistent
              else "per kernel"}
    return dict(value=nc * 4 * args.steps / elapsed, elapsed=elapsed, config=config, roofline=roof,
                cpu_baseline=cpu, fp32=opts.precision != 0, extra=extra)


def bench_batch(args, M, cases, copy_gbs) -> dict:
    import numpy as np
    t0 = time.perf_counter()
    from miniaero_amd.mesh import generate_meshes
    meshes = generate_meshes(cases)
    gen_s = time.perf_counter() - t0
    solver = M.BatchSolver(list(zip(meshes, cases)))
    setup_s = time.perf_counter() - t0
    sizes = solver.sizes()
    nc, nif, nbf = sizes["cells"], sizes["internal_faces"], sizes["boundary_faces"]
    print(f"[bench] {args.workload}: {len(cases)} cases, {nc} cells, {nif} internal + {nbf} boundary "
          f"faces, mesh generation {gen_s:.1f} s, setup {setup_s:.1f} s, device "
          f"{sizes['device_bytes'] / 2**30:.2f} GiB", file=sys.stderr, flush=True)
    shape, wg, tuned_ms = solver.tuning()
    print(f"[bench] tile shape {shape}, workgroups per CU {wg or 'all tiles'}"
          + (f" (autotuned, {tuned_ms:.3f} ms/step/Mcell)" if tuned_ms else ""), file=sys.stderr, flush=True)
    elapsed = timed_run(solver, args.steps, args.warmup, args.graph, args.persistent)
    if solver.nonfinite():
        raise RuntimeError("non-finite state after the timed steps")
    print(f"[bench] {args.steps} steps in {elapsed:.3f} s", file=sys.stderr, flush=True)
    warm_total = args.warmup + (2 * args.graph + 1 if args.graph else 0)
    extra = {}
    fp32 = cases[0].precision != 0
    if args.compare_fp64 and fp32:
        # compared after at least the cases' own step count
        n_cmp = max(warm_total + args.steps, cases[0].ntimesteps)
        if n_cmp > warm_total + args.steps:
            solver.step(n_cmp - warm_total - args.steps)
        U = [solver.case_state(k) for k in range(len(cases))]
        ref = M.BatchSolver(list(zip(meshes, [o.replace(precision=0, autotune=0) for o in cases])))
        ref.initialize()
        ref.step(n_cmp)
        errs = [precision_errors(U[k], ref.case_state(k)) for k in range(len(cases))]
        ref.close()
        worst = {k: max(e[k] for e in errs) for k in errs[0]}
        extra["precision_error"] = {
            "metric": "relative L2 vs the FP64 run after warmup+steps, worst case of the batch; "
                      "momentum components scaled by the momentum-vector norm",
            "steps": n_cmp, **worst}
        print(f"[bench] precision error: {json.dumps(extra['precision_error'])}", file=sys.stderr, flush=True)
    del meshes
    roof = roofline(solver, args.profile_steps, nc, nif, nbf, True, fp32,
                    lambda k: load_traffic(args.workload, k), copy_gbs,
                    face_values=bool(cases[0].second_order_space) and not cases[0].viscous)
    pb = path_bytes_per_step(nc, nif, nbf, True, bool(cases[0].viscous), fp32=fp32)
    path_gbs = pb / (elapsed / args.steps) / 1e9
    roof["path"] = {"algorithmic_bytes_per_step": pb, "achieved_GBps": round(path_gbs, 1),
                    "frac": round(path_gbs / HBM_PEAK_GBS, 4)}
    cpu = None
    if not args.no_cpu_baseline:
        try:
            cpu = cpu_baseline_batch(cases)
        except Exception as e:
            print(f"[bench] cpu baseline skipped: {e}", file=sys.stderr)
    solver.close()
    cells = np.array([o.nx * o.ny * o.nz for o in cases])
    # every case of the batch completes with the batch: its latency is the
    # time of the workload's step count at the measured step time
    per_case_ms = elapsed / args.steps * cases[0].ntimesteps * 1e3
    config = {"workload": args.workload, "cases": len(cases), "cells": nc,
              "cells_per_case": {"min": int(cells.min()), "mean": float(cells.mean()), "max": int(cells.max())},
              "internal_faces": nif, "boundary_faces": nbf,
              "physics": "inviscid Roe, Green-Gauss, Van Albada, 2nd order, Sod-type states",
              "dt": cases[0].dt, "steps_per_case": cases[0].ntimesteps,
              "precision": "fp32 flux/gradient/limiter, fp64 state" if fp32 else "fp64",
              "tile_shape": list(shape), "workgroups_per_cu": wg,
              "autotuned_ms_per_step_per_Mcell": tuned_ms or None,
              "graph_steps": 0 if args.persistent else args.graph,
              "step_launch": "persistent (one cooperative launch per step() call)" if args.persistent
              else "per kernel",
              "case_latency_ms": {"p50": round(per_case_ms, 2), "p99": round(per_case_ms, 2),
                                  "note": "all cases run in the same launches and finish together"}}
    return dict(value=nc * 4 * args.steps / elapsed, elapsed=elapsed, config=config, roofline=roof,
                cpu_baseline=cpu, fp32=fp32, extra=extra)


def bench_served(args, M, stream, copy_gbs) -> dict:
    """Replays the arrival stream in real time against a Server: each request
    is submitted at its arrival time, finished cases are collected as they
    come.  value = sum over cases of cells * 4 * steps / (last completion -
    first submission); latency = completion - submission."""
    import numpy as np
    srv = M.Server(hbm_budget_bytes=args.hbm_budget, max_active=args.max_active,
                   setup_threads=args.setup_threads, steps_in_flight=args.steps_in_flight)
    # one small case first, so that module loading is not in the first request's latency
    warm = stream[0][1].replace(nx=8, ny=4, nz=2, ntimesteps=2)
    wid = srv.submit(warm)
    while True:
        got = srv.poll(timeout=5.0)
        if got:
            if got[0]["status"] != 0:
                raise RuntimeError(f"warm-up case failed: {got[0]}")
            srv.release(wid)
            break
    st0 = srv.stats()
    print(f"[bench] served: {len(stream)} requests, region {st0['region_size'] / 1e9:.1f} GB, "
          f"max_active {args.max_active or 16}, setup threads {args.setup_threads}", file=sys.stderr, flush=True)
    done, opts_of = {}, {}
    t0 = srv.now()
    last_note = time.time()
    peak_used = 0

    def collect(timeout):
        nonlocal peak_used
        for r in srv.poll(timeout=timeout):
            done[r["id"]] = r
            if args.compare_fp64 and r["status"] == 0 and len([1 for x in done.values() if "U" in x]) < 6:
                r["U"] = srv.state(r["id"], r["cells"])
            srv.release(r["id"])
        peak_used = max(peak_used, srv.stats()["region_used"])

    for arrival, o in stream:
        while True:
            wait = t0 + arrival - srv.now()
            if wait <= 0:
                break
            collect(min(wait, 0.005))
        opts_of[srv.submit(o)] = o
        if time.time() - last_note > 10:
            print(f"[bench] {len(opts_of)} submitted, {len(done)} finished", file=sys.stderr, flush=True)
            last_note = time.time()
    while len(done) < len(opts_of):
        collect(1.0)
        if time.time() - last_note > 10:
            print(f"[bench] {len(done)} of {len(opts_of)} finished", file=sys.stderr, flush=True)
            last_note = time.time()
    plan_cache = srv.plan_cache()
    srv.close()
    bad = [r for r in done.values() if r["status"] != 0]
    if bad:
        raise RuntimeError(f"{len(bad)} served cases failed, e.g. {bad[0]}")
    rs = list(done.values())
    lat = np.array([r["t_finish"] - r["t_submit"] for r in rs])
    wait = np.array([r["t_admit"] - r["t_submit"] for r in rs])
    work = sum(r["cells"] * 4 * r["steps"] for r in rs)
    first = min(r["t_submit"] for r in rs)
    span = max(r["t_finish"] for r in rs) - first
    extra = {}
    if args.compare_fp64:
        errs = []
        for r in rs:
            if "U" not in r:
                continue
            o = opts_of[r["id"]].replace(precision=0)
            md = M.Parallel3DMesh(o.nx, o.ny, o.nz, o.lx, o.ly, o.lz, o.problem_type, o.angle,
                                  options=o).fillMeshData()
            ref = M.TimeSolverExplicitRK4(md, o)
            ref.initialize()
            ref.step(o.ntimesteps)
            errs.append(precision_errors(r["U"], ref.state()))
            ref.close()
        if errs:
            extra["precision_error"] = {
                "metric": f"relative L2 vs FP64 runs of the first {len(errs)} finished cases, worst case; "
                          "momentum components scaled by the momentum-vector norm",
                **{k: max(e[k] for e in errs) for k in errs[0]}}
    cells = np.array([r["cells"] for r in rs])
    path_bytes = 0.0
    for r in rs:
        o = opts_of[r["id"]]
        nc, nif, nbf = box_faces(o.nx, o.ny, o.nz)
        path_bytes += path_bytes_per_step(nc, nif, nbf, True, bool(o.viscous), fp32=o.precision != 0) * r["steps"]
    path_gbs = path_bytes / span / 1e9
    roof = {"bound": "hbm", "achieved": round(path_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(path_gbs / HBM_PEAK_GBS, 4), "traffic": None,
            "kernel": "path: every kernel of every case, algorithmic bytes of the precision form over "
                      "the span from the first submission to the last completion"}
    config = {"workload": args.workload, "requests": len(rs),
              "arrivals": "Poisson, 200 per second (seeded)", "cells_per_case":
                  {"min": int(cells.min()), "mean": float(cells.mean()), "max": int(cells.max())},
              "physics": "Roe, Green-Gauss, Venkat, 2nd order; half viscous flat plate, half slip wall",
              "precision": "fp32 flux/gradient/limiter, fp64 state",
              "max_active": args.max_active or 16, "setup_threads": args.setup_threads,
              "steps_in_flight": args.steps_in_flight or 2,
              "layout_plans": {"reused": plan_cache["hits"], "planned": plan_cache["misses"]},
              "region_bytes": st0["region_size"],
              "peak_region_bytes_used": int(peak_used),
              "case_latency_s": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                                 "max": float(lat.max())},
              "admission_wait_s": {"p50": float(np.percentile(wait, 50)), "p99": float(np.percentile(wait, 99))},
              "span_s": span, "offered_cell_stages_per_s": work / max(stream[-1][0] - stream[0][0], 1e-9)}
    return dict(value=work / span, elapsed=span, config=config, roofline=roof, cpu_baseline=None, fp32=True,
                extra=extra, steps=None)


def _sub(args, **over):
    a = argparse.Namespace(**vars(args))
    for k, v in over.items():
        setattr(a, k, v)
    return a


def _compact(r, steps) -> dict:
    """The fields of one workload's result that go into the suite line."""
    out = {"value": r["value"], "unit": "cell-stage updates/s"}
    if r.get("steps", 1) is not None:
        out["ms_per_step"] = r["elapsed"] / steps * 1e3
        out["steps"] = steps
    roof = r.get("roofline")
    if roof:
        out["roofline"] = {k: roof[k] for k in ("bound", "achieved", "peak", "unit", "frac", "traffic", "kernel",
                                                "mean_launch_ms", "path") if k in roof}
    out["cpu_baseline"] = r.get("cpu_baseline")
    out["dtype"] = "f32 arithmetic, f64 state" if r["fp32"] else "f64"
    out["config"] = r["config"]
    out.update(r.get("extra", {}))
    return out


def sod_c1_full(M) -> dict:
    """The 3D Sod case (64x16x16, first order, 100 steps) run in full on the
    GPU -- with a launch per kernel, as one captured graph of all 100 steps
    and as one persistent launch -- and in full on one host core by the CPU
    oracle; every final state is compared with the oracle's bit for bit."""
    import dataclasses

    import numpy as np

    from miniaero_amd.configs import sod_c1
    from oracle import oracle as O

    o = sod_c1()
    md = M.Parallel3DMesh(o.nx, o.ny, o.nz, o.lx, o.ly, o.lz, o.problem_type, o.angle, options=o).fillMeshData()

    def run(mode):
        # the whole case, after an untimed run of it (which captures the graph)
        s = M.TimeSolverExplicitRK4(md, o)
        try:
            if mode == "graph":
                s.set_graph(o.ntimesteps)
            elif mode == "persistent":
                s.set_persistent(True)
            s.initialize()
            s.step(o.ntimesteps)
            s.synchronize()
            s.initialize()
            t0 = time.perf_counter()
            s.step(o.ntimesteps)
            s.synchronize()
            return time.perf_counter() - t0, s.state()
        finally:
            s.close()

    kw = {f.name: getattr(o, f.name) for f in dataclasses.fields(o)
          if f.name not in ("output_results", "output_frequency")}
    om = O.OracleMesh(O.make_options(**kw))
    os_ = O.OracleSolver(om)
    os_.init()
    t0 = time.perf_counter()
    os_.step(o.ntimesteps)
    cpu_s = time.perf_counter() - t0
    Uref = os_.state()
    work = om.ncells * 4 * o.ntimesteps
    modes = {}
    for mode in ("per_kernel", "graph", "persistent"):
        sec, U = run(mode)
        modes[mode] = {"value": work / sec, "ms_per_step": sec / o.ntimesteps * 1e3,
                       "matches_oracle_bitwise": bool(np.array_equal(U, Uref))}
    best = max(modes, key=lambda m: modes[m]["value"])
    gpu_s = work / modes[best]["value"]
    return {"value": work / gpu_s, "unit": "cell-stage updates/s", "ms_per_step": gpu_s / o.ntimesteps * 1e3,
            "steps": o.ntimesteps, "dtype": "f64",
            "note": "16k cells: the GPU time is launch latency, not bandwidth; value = the fastest step mode",
            "step_mode": best, "step_modes": modes,
            "matches_oracle_bitwise": all(m["matches_oracle_bitwise"] for m in modes.values()),
            "cpu_baseline": {"value": work / cpu_s, "unit": "cell-stage updates/s", "cores": 1, "kind": "port",
                             "sample": f"the whole case ({om.ncells} cells, {o.ntimesteps} RK4 steps) on one "
                                       f"core by the oracle (single-thread C++ restatement of the reference's "
                                       f"serial path), {cpu_s:.1f} s"},
            "config": {"workload": "sod_c1", "mesh": f"{o.nx}x{o.ny}x{o.nz}", "cells": om.ncells,
                       "physics": "inviscid Roe, 1st order, Extrapolate on every side", "dt": o.dt}}


def plain_fp64_on_cases(M, cases, timed_steps: int = 20, graph: int = 10) -> dict:
    """The plain FP64 path on a sample of a workload's own cases, one case at
    a time, as the reference runs them (one case per process).  Each case is
    built alone in FP64 with the default tiles and stepped in graphs of
    `graph` steps; its step time is measured over `timed_steps` steps after
    one untimed graph.  The aggregate is the sample's work at each case's own
    step count over the summed stepping time.  Mesh generation and setup are
    left out, which favours the plain path."""
    from miniaero_amd.mesh import generate_meshes
    plain = [o.replace(precision=0, autotune=0) for o in cases]
    meshes = generate_meshes(plain)
    work = secs = 0.0
    rates = []
    for md, o in zip(meshes, plain):
        s = M.TimeSolverExplicitRK4(md, o)
        try:
            s.set_graph(graph)
            s.initialize()
            s.step(graph)
            s.synchronize()
            t0 = time.perf_counter()
            s.step(timed_steps)
            s.synchronize()
            t = (time.perf_counter() - t0) / timed_steps
            if s.nonfinite():
                raise RuntimeError("non-finite state in the plain-path sample")
        finally:
            s.close()
        cells = o.nx * o.ny * o.nz
        work += cells * 4.0 * o.ntimesteps
        secs += t * o.ntimesteps
        rates.append(cells * 4.0 / t)
    cells = [o.nx * o.ny * o.nz for o in plain]
    return {"value": work / secs, "unit": "cell-stage updates/s", "cases": len(plain),
            "cells_per_case": {"min": min(cells), "max": max(cells)},
            "per_case_rate": {"min": min(rates), "max": max(rates)},
            "method": f"each sampled case alone in FP64 (default tiles, graphs of {graph} steps), step time "
                      f"over {timed_steps} steps after one untimed graph; aggregate = sample work at each "
                      f"case's own step count / summed stepping time; setup excluded"}


def vs_plain(M, r, cases, n: int = 20) -> None:
    """Adds the plain FP64 path on about n of the cases (an even stride over
    the list) and the ratio to r's extra fields."""
    stride = max(1, len(cases) // n)
    p = plain_fp64_on_cases(M, cases[::stride])
    p["sample"] = f"every {stride}th case"
    r.setdefault("extra", {}).update(plain_fp64_same_cases=p,
                                     vs_plain_fp64_same_cases=round(r["value"] / p["value"], 3))
    print(f"[bench] plain FP64 path on {p['cases']} of the same cases: {p['value']:.3e} cell-stage updates/s",
          file=sys.stderr, flush=True)


def bench_suite(args, M, copy_gbs) -> dict:
    """The primary line (FP64 viscous plate, 2.1M cells) plus every other
    headline figure of the path, each under workloads[name]."""
    from miniaero_amd.configs import BATCH_WORKLOADS, SERVED_WORKLOADS, WORKLOADS

    def note(msg):
        print(f"[bench] suite: {msg}", file=sys.stderr, flush=True)

    note("flat_plate_c2 (primary)")
    r = bench_single(_sub(args, workload="flat_plate_c2", graph=0), M, WORKLOADS["flat_plate_c2"](), copy_gbs)
    plain = r["value"]
    wl = {}

    def attempt(name, fn):
        note(name)
        t0 = time.perf_counter()
        try:
            wl[name] = fn()
        except Exception as e:  # a failed leg is reported, the others still run
            wl[name] = {"error": f"{type(e).__name__}: {e}"}
        wl[name]["bench_seconds"] = round(time.perf_counter() - t0, 1)
        if "value" in wl[name]:
            wl[name]["vs_plain_fp64_c2"] = round(wl[name]["value"] / plain, 3)
        note(f"{name} done in {wl[name]['bench_seconds']} s")

    same_physics = {"see": "cpu_baseline of the primary line: the same physics, timed on the same host"}

    def fp64_large():
        a = _sub(args, workload="flat_plate_c3_fp64", graph=0, no_cpu_baseline=True, compare_fp64=False)
        x = _compact(bench_single(a, M, WORKLOADS["flat_plate_c3"]().replace(precision=0), copy_gbs), a.steps)
        x["cpu_baseline"] = same_physics
        return x

    def precision_c3():
        o = WORKLOADS["flat_plate_c3"]()
        a = _sub(args, workload="flat_plate_c3", graph=0, no_cpu_baseline=True, compare_fp64=True,
                 warmup=min(args.warmup, 5), steps=o.ntimesteps - min(args.warmup, 5))
        x = _compact(bench_single(a, M, o, copy_gbs), a.steps)
        x["cpu_baseline"] = same_physics
        return x

    def against_plain(x, cases, sample):
        # the target's comparison: the plain FP64 path on the same cases (a
        # failure here is reported without losing the workload's own figures)
        try:
            p = plain_fp64_on_cases(M, cases[::sample])
        except Exception as e:
            x["plain_fp64_same_cases"] = {"error": f"{type(e).__name__}: {e}"}
            return
        p["sample"] = f"every {sample}th case"
        x["plain_fp64_same_cases"] = p
        x["vs_plain_fp64_same_cases"] = round(x["value"] / p["value"], 3)
        note(f"plain FP64 path on {p['cases']} of the same cases: {p['value']:.3e} cell-stage updates/s")

    def batched_c4():
        # the mixed-precision form, with every case's error against its FP64 run after its 100 steps
        a = _sub(args, workload="batched_c4", graph=2, compare_fp64=True)
        cases = [o.replace(precision=1) for o in BATCH_WORKLOADS["batched_c4"]()]
        x = _compact(bench_batch(a, M, cases, copy_gbs), a.steps)
        against_plain(x, cases, 16)
        return x

    def served_200():
        a = _sub(args, workload="served_c5", cases=200, compare_fp64=False)
        stream = SERVED_WORKLOADS["served_c5"](200)
        x = _compact(bench_served(a, M, stream, copy_gbs), a.steps)
        if not args.no_cpu_baseline:
            small = [o for _, o in stream if o.nx * o.ny * o.nz <= 300_000]
            x["cpu_baseline"] = cpu_baseline_batch(small, min_cells=300_000, steps=2)
        against_plain(x, [o for _, o in stream], 10)
        return x

    attempt("flat_plate_fp64_8M", fp64_large)
    attempt("precision_c3", precision_c3)
    attempt("batched_c4", batched_c4)
    attempt("served_c5_first200", served_200)
    attempt("sod_c1", lambda: sod_c1_full(M))
    r["extra"] = dict(r.get("extra", {}), workloads=wl)
    return r


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="suite",
                    help="suite (default: the primary flat-plate line plus every other headline workload), "
                         "or one of flat_plate_c2, flat_plate_c3, sod_c1, batched_c4, served_c5")
    ap.add_argument("--profile-steps", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--precision", type=int, default=None,
                    help="override the workload's arithmetic (0 FP64, 1 FP32 flux/gradient/limiter)")
    ap.add_argument("--compare-fp64", action="store_true",
                    help="also run the workload in FP64 and report the per-variable relative L2 error")
    ap.add_argument("--graph", type=int, default=None,
                    help="steps per captured HIP graph (even; 0 = plain launches); default 2 for "
                         "batched workloads, 0 otherwise")
    ap.add_argument("--persistent", action="store_true",
                    help="run the timed steps as one cooperative launch with grid-wide barriers "
                         "between phases instead of a launch per kernel (graphs off)")
    ap.add_argument("--autotune", type=int, default=None,
                    help="1: pick the flux tile shape by timing candidates at setup (default: the "
                         "workload's setting)")
    ap.add_argument("--cases", type=int, default=200, help="served workload: requests in the stream")
    ap.add_argument("--max-active", type=int, default=32, help="served workload: cases on the GPU at once")
    ap.add_argument("--hbm-budget", type=float, default=200e9, help="served workload: device bytes for cases")
    ap.add_argument("--setup-threads", type=int, default=14,
                    help="served workload: host threads generating and planning requests")
    ap.add_argument("--steps-in-flight", type=int, default=0,
                    help="served workload: RK4 steps queued ahead per scheme group (0 = the server's default, 2)")
    ap.add_argument("--wg-per-cu", type=int, default=None,
                    help="resident workgroups per CU for the flux kernel (0: one per tile)")
    ap.add_argument("--vs-plain", type=int, nargs="?", const=20, default=0, metavar="N",
                    help="batched and served workloads: also time the plain FP64 path on about N (default "
                         "20) of the same cases, one at a time, and report the ratio")
    ap.add_argument("--copy-ceiling", action="store_true",
                    help="only measure the plain copy kernel (1 GiB, read + write) and exit")
    ap.add_argument("--cpu-worker", action="store_true", help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.cpu_worker:
        return cpu_worker()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU runs are not part of it "
                                   "(cases and mesh blocks shard across GPUs as described in DESIGN.md)"}))
        return 2

    import miniaero_amd as M
    from miniaero_amd import _lib
    from miniaero_amd.configs import BATCH_WORKLOADS, SERVED_WORKLOADS, WORKLOADS

    copy_bytes = 1 << 30
    copy_ms = _lib.measure_copy(copy_bytes, 10)
    copy_gbs = 2 * copy_bytes / (copy_ms * 1e-3) / 1e9
    if args.copy_ceiling:
        print(json.dumps({"copy_bytes": copy_bytes, "ms_per_copy": copy_ms, "GBps_read_plus_write": copy_gbs}))
        return 0

    if args.graph is None:
        args.graph = 2 if args.workload in BATCH_WORKLOADS else 0
    try:
        if args.workload == "suite":
            r = bench_suite(args, M, copy_gbs)
        elif args.workload in SERVED_WORKLOADS:
            stream = SERVED_WORKLOADS[args.workload](args.cases)
            r = bench_served(args, M, stream, copy_gbs)
            if args.vs_plain:
                vs_plain(M, r, [o for _, o in stream], args.vs_plain)
        elif args.workload in BATCH_WORKLOADS:
            cases = BATCH_WORKLOADS[args.workload]()
            if args.precision is not None:
                cases = [o.replace(precision=args.precision) for o in cases]
            if args.autotune is not None:
                cases = [o.replace(autotune=args.autotune) for o in cases]
            if args.wg_per_cu is not None:
                cases = [o.replace(workgroups_per_cu=args.wg_per_cu) for o in cases]
            r = bench_batch(args, M, cases, copy_gbs)
            if args.vs_plain:
                vs_plain(M, r, cases, args.vs_plain)
        else:
            opts = WORKLOADS[args.workload]()
            if args.precision is not None:
                opts = opts.replace(precision=args.precision)
            if args.autotune is not None:
                opts = opts.replace(autotune=args.autotune)
            if args.wg_per_cu is not None:
                opts = opts.replace(workgroups_per_cu=args.wg_per_cu)
            r = bench_single(args, M, opts, copy_gbs)
    except RuntimeError as e:
        print(f"[bench] {e}", file=sys.stderr)
        return 1
    line = {
        "metric": "RK4 cell-stage updates/s",
        "value": r["value"],
        "unit": "cell-stage updates/s",
        "n_gpus": 1,
        "steps": args.steps if r.get("steps", 1) is not None else None,
        "warmup": args.warmup,
        "ms_per_step": r["elapsed"] / args.steps * 1e3 if r.get("steps", 1) is not None else None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 arithmetic, f64 state" if r["fp32"] else "f64",
        "data": "synthetic",
        "config": r["config"],
        "roofline": r["roofline"],
        "cpu_baseline": r["cpu_baseline"],
    }
    line.update(r["extra"])
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
