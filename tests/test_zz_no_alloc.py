"""ma_solver_residual_norms_noalloc: the residual norms without any device
allocation, written to a caller's device array on the solver's stream.

Bar: bit for bit the norms of ma_solver_residual_norms (itself checked against
numpy norms of the oracle's residual), on two calls in a row through the C ABI
-- a larger run, then a smaller one, so the second call reads nothing the
first left in the shared static scratch -- and for a batch of cases, where each
case's row comes from its own tiles.  The device output array is the caller's
(allocated here with the HIP runtime, as a caller would); the entry point
itself allocates nothing.
"""
import ctypes
import ctypes.util

import numpy as np
import pytest

import miniaero_amd as M
from miniaero_amd import _lib

pytestmark = pytest.mark.gpu


def _hip():
    for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
        try:
            h = ctypes.CDLL(name)
            break
        except OSError:
            h = None
    assert h is not None, "HIP runtime library not found"
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipFree.argtypes = [ctypes.c_void_p]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    return h


HIP_D2H = 2


def _noalloc_norms(solver, hip, dev, ncases):
    L = _lib.load()
    _lib.check(L.ma_solver_residual_norms_noalloc(solver._h, dev))
    _lib.check(L.ma_solver_synchronize(solver._h))
    out = np.full((ncases, 5), np.nan)
    assert hip.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), dev, out.nbytes, HIP_D2H) == 0
    return out


def _run(nx, ny, nz, steps, **kw):
    o = M.Options(problem_type=1, lx=2.0, ly=0.002, lz=1.0, nx=nx, ny=ny, nz=nz, ntimesteps=steps, dt=3e-8,
                  second_order_space=1, viscous=1, **kw)
    md = M.Parallel3DMesh(nx, ny, nz, o.lx, o.ly, o.lz, o.problem_type).fillMeshData()
    s = M.TimeSolverExplicitRK4(md, o)
    s.initialize()
    s.step(steps)
    return s


def test_noalloc_norms_equal_the_allocating_entry_point():
    hip = _hip()
    dev = ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(dev), 8 * 5 * 8) == 0
    try:
        for nx, ny, nz, steps in ((64, 32, 16, 3), (16, 8, 4, 2)):   # second call smaller
            print(f"[no-alloc] {nx}x{ny}x{nz}, {steps} steps", flush=True)
            s = _run(nx, ny, nz, steps)
            try:
                ref = s.residual_norms()
                got = _noalloc_norms(s, hip, dev, 1)[0]
                assert np.all(ref > 0)
                assert np.array_equal(got, ref), (got, ref)
            finally:
                s.close()
        # a batch: one row per case, each from its own tiles
        cases = []
        for nx, ny, nz in ((32, 16, 8), (8, 4, 4), (16, 8, 2)):
            o = M.Options(problem_type=0, lx=1.0, ly=0.25, lz=0.25, nx=nx, ny=ny, nz=nz, ntimesteps=2, dt=1e-4,
                          second_order_space=1, viscous=0, limiter=1, ic_kind=1, ic=(1.0, 1.0, 0.2, 0.3, 0.45),
                          bc_kind=(1, 1, 1, 1, 1, 1))
            cases.append((M.Parallel3DMesh(nx, ny, nz, o.lx, o.ly, o.lz, 0).fillMeshData(), o))
        print("[no-alloc] batch of 3 cases", flush=True)
        b = M.BatchSolver(cases)
        try:
            b.initialize()
            b.step(2)
            ref = b.residual_norms()
            got = _noalloc_norms(b, hip, dev, 3)
            assert np.array_equal(got, ref), (got, ref)
            assert len({tuple(r) for r in ref}) == 3
        finally:
            b.close()
    finally:
        hip.hipFree(dev)
