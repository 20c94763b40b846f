"""The benchmark configurations themselves, against the oracle.

C1 (64x16x16 first-order Sod, Extrapolate on every side, 100 steps) in full
and C2 (the 2.1M-cell stretched viscous plate at CFL 0.5) for its first two
steps equal the oracle bit for bit.  C3's FP32 arithmetic form (8.4M cells)
stays within relative L2 1e-5 of its FP64 run, whole state and per
variable.  Cases of the C4 batch, stepped together, equal the same cases run
alone bit for bit (FP64), and their FP32 form stays within 1e-5 of FP64 for
the whole state (per variable the worst of all 512 cases is 2.5e-5, the
momentum of a case whose two pressures nearly match; 1e-4 is asserted)."""
import dataclasses

import numpy as np
import pytest

import miniaero_amd as M
from miniaero_amd.configs import batched_c4, flat_plate_c2, flat_plate_c3, sod_c1
from oracle import oracle as O
from test_gpu_parity import _log, precision_errors

pytestmark = pytest.mark.gpu


def _kw(opts):
    return {f.name: getattr(opts, f.name) for f in dataclasses.fields(opts)
            if f.name not in ("output_results", "output_frequency")}


def _mesh(opts):
    return M.Parallel3DMesh(opts.nx, opts.ny, opts.nz, opts.lx, opts.ly, opts.lz, opts.problem_type,
                            opts.angle, options=opts).fillMeshData()


def _against_oracle(opts, chunks, per_chunk, mode="per_kernel"):
    md = _mesh(opts)
    s = M.TimeSolverExplicitRK4(md, opts)
    cpu = O.OracleSolver(O.OracleMesh(O.make_options(**_kw(opts))))
    try:
        if mode == "graph":
            s.set_graph(per_chunk)
        elif mode == "persistent":
            s.set_persistent(True)
        s.initialize()
        cpu.init()
        assert np.array_equal(s.state(), cpu.state()), "initial state"
        for i in range(chunks):
            s.step(per_chunk)
            cpu.step(per_chunk)
            same = np.array_equal(s.state(), cpu.state())
            _log(f"[gpu-configs] {opts.nx}x{opts.ny}x{opts.nz}: step {(i + 1) * per_chunk}, bitwise={same}")
            assert same, f"differs from the oracle after {(i + 1) * per_chunk} steps"
            t, dt = s.time()
            assert t == cpu.time and dt == cpu.last_dt
    finally:
        s.close()


@pytest.mark.parametrize("mode", ["per_kernel", "graph", "persistent"])
def test_c1_sod_in_full_matches_oracle(mode):
    """C1 in full, checked every 10 steps, with a launch per kernel, as graph
    replays of 10 steps and as one persistent launch per 10 steps."""
    o = sod_c1()
    _against_oracle(o, chunks=10, per_chunk=o.ntimesteps // 10, mode=mode)


@pytest.mark.slow
def test_c2_plate_first_steps_match_oracle():
    _against_oracle(flat_plate_c2(), chunks=2, per_chunk=1)


@pytest.mark.slow
def test_c3_precision_form_within_tolerance():
    o = flat_plate_c3()
    md = _mesh(o)
    runs = {}
    for prec in (0, 1):
        s = M.TimeSolverExplicitRK4(md, o.replace(precision=prec))
        try:
            s.initialize()
            s.step(20)
            runs[prec] = s.state()
        finally:
            s.close()
    err = precision_errors(runs[1], runs[0])
    _log(f"[gpu-configs] C3 FP32 vs FP64 after 20 steps: {err[0]:.2e}, per variable {['%.2e' % x for x in err[1]]}")
    assert err[0] <= 1e-5 and max(err[1]) <= 1e-5
    assert not np.array_equal(runs[0], runs[1])


def test_c4_cases_batched_equal_alone_and_fp32_within_tolerance():
    cases = [o.replace(autotune=0) for o in batched_c4()[:8]]
    meshes = [_mesh(o) for o in cases]
    out = {}
    for prec in (0, 1):
        b = M.BatchSolver([(md, o.replace(precision=prec)) for md, o in zip(meshes, cases)])
        try:
            b.initialize()
            b.step(20)
            out[prec] = [b.case_state(k) for k in range(len(cases))]
        finally:
            b.close()
    for k, (md, o) in enumerate(zip(meshes, cases)):
        s = M.TimeSolverExplicitRK4(md, o)
        try:
            s.initialize()
            s.step(20)
            assert np.array_equal(out[0][k], s.state()), f"case {k} differs from its run alone"
        finally:
            s.close()
        whole, per, _ = precision_errors(out[1][k], out[0][k])
        _log(f"[gpu-configs] C4 case {k} ({o.nx}x{o.ny}x{o.nz}): FP32 vs FP64 {whole:.2e}, "
             f"per variable max {max(per):.2e}")
        assert whole <= 1e-5 and max(per) <= 1e-4
