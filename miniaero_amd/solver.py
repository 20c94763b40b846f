"""TimeSolverExplicitRK4: host mirror of the reference's solver class
(TimeSolverExplicitRK4.h:160-539) over the library's C ABI.

    solver = TimeSolverExplicitRK4(mesh_data, options)
    solver.Solve()            # ntimesteps RK4 steps; writes results.0 if asked

All arithmetic runs in the HIP kernels of libminiaero_amd.so on the current
GPU; this class only sequences calls and moves results to the host.
"""
from __future__ import annotations

import ctypes
import sys
import time as _time

import numpy as np

from . import _lib
from .mesh import MeshData
from .options import Options
from .results import write_results


class TimeSolverExplicitRK4:
    def __init__(self, mesh_data: MeshData, options: Options):
        self.mesh_data = mesh_data
        self.options = options
        self._L = _lib.load()
        self._c_opts = options.to_c()
        self._c_mesh = mesh_data.to_c()
        h = ctypes.c_void_p()
        _lib.check(self._L.ma_solver_create(ctypes.byref(self._c_mesh), ctypes.byref(self._c_opts),
                                            ctypes.byref(h)))
        self._h = h
        self.ncells = int(mesh_data.num_owned_cells)

    # -- the reference's entry point ------------------------------------------
    def Solve(self, out=sys.stdout, results_path: str = "results.0") -> None:
        """Initial condition, ntimesteps RK4 steps, then results.<rank> when
        output_results is set (TimeSolverExplicitRK4.h:207-539)."""
        o = self.options
        t0 = _time.time()
        self.initialize()
        freq = max(1, int(o.output_frequency))
        done = 0
        while done < o.ntimesteps:
            n = min(freq - (done % freq), o.ntimesteps - done)
            self.step(n)
            done += n
            if done % freq == 0:
                t, dt = self.time()
                out.write("\nTime Step #%i:  Time = %16.9e; dt = %16.9e\n" % (done, t, dt))
        self.synchronize()
        out.write("\n ... Device Run time: %8.2f seconds ...\n" % (_time.time() - t0))
        if o.output_results:
            write_results(results_path, self.mesh_data.cell_coordinates, self.state())

    # -- finer-grained access ---------------------------------------------------
    def initialize(self) -> None:
        _lib.check(self._L.ma_solver_initialize(self._h))

    def set_state(self, U: np.ndarray) -> None:
        U = np.ascontiguousarray(U, dtype=np.float64).reshape(self.ncells, 5)
        _lib.check(self._L.ma_solver_set_state(self._h, U.ctypes.data_as(_lib._D)))

    def step(self, n: int = 1) -> None:
        _lib.check(self._L.ma_solver_step(self._h, int(n)))

    def tuning(self):
        """(tile brick (i, j, k) in use, resident workgroups per CU, autotuned ms per
        step per million cells or 0)."""
        sh = (ctypes.c_int * 3)()
        wg = ctypes.c_int()
        ms = ctypes.c_double()
        _lib.check(self._L.ma_solver_tuning(self._h, sh, ctypes.byref(wg), ctypes.byref(ms)))
        return tuple(sh), wg.value, ms.value

    def tile_shape(self):
        """(tile brick (i, j, k) in use, autotuned ms per step per million cells or 0)."""
        sh, _, ms = self.tuning()
        return sh, ms

    def set_graph(self, steps_per_graph: int) -> None:
        """Run groups of steps_per_graph steps (even) as one captured HIP graph; 0 = off."""
        _lib.check(self._L.ma_solver_set_graph(self._h, int(steps_per_graph)))

    def set_persistent(self, enabled: bool) -> None:
        """Run each step() call's steps as one cooperative launch with grid-wide
        barriers between phases (waits for it); turns graphs off."""
        _lib.check(self._L.ma_solver_set_persistent(self._h, int(bool(enabled))))

    def synchronize(self) -> None:
        _lib.check(self._L.ma_solver_synchronize(self._h))

    def state(self) -> np.ndarray:
        U = np.empty((self.ncells, 5))
        _lib.check(self._L.ma_solver_get_state(self._h, U.ctypes.data_as(_lib._D)))
        return U

    def residual(self) -> np.ndarray:
        R = np.empty((self.ncells, 5))
        _lib.check(self._L.ma_solver_get_residual(self._h, R.ctypes.data_as(_lib._D)))
        return R

    def gradients(self):
        g = np.empty((self.ncells, 5, 3))
        lim = np.empty((self.ncells, 5))
        _lib.check(self._L.ma_solver_get_gradients(self._h, g.ctypes.data_as(_lib._D),
                                                   lim.ctypes.data_as(_lib._D)))
        return g, lim

    def time(self):
        t, d = ctypes.c_double(), ctypes.c_double()
        _lib.check(self._L.ma_solver_get_time(self._h, ctypes.byref(t), ctypes.byref(d)))
        return t.value, d.value

    def residual_norms(self) -> np.ndarray:
        """L2 norm of the last stage's residual per conserved variable (one row
        per case for a batch)."""
        n = max(1, getattr(self, "ncases", 1))
        out = np.empty((n, 5))
        _lib.check(self._L.ma_solver_residual_norms(self._h, out.ctypes.data_as(_lib._D)))
        return out[0] if not hasattr(self, "ncases") else out

    def nonfinite(self) -> int:
        n = ctypes.c_longlong(0)
        st = self._L.ma_solver_check_finite(self._h, ctypes.byref(n))
        if st not in (_lib.MA_OK, _lib.MA_ERR_NONFINITE):
            _lib.check(st)
        return int(n.value)

    def sizes(self) -> dict:
        c = (ctypes.c_longlong * 4)()
        _lib.check(self._L.ma_solver_sizes(self._h, c))
        return dict(cells=c[0], internal_faces=c[1], boundary_faces=c[2], device_bytes=c[3])

    def stream(self) -> int:
        p = ctypes.c_void_p()
        _lib.check(self._L.ma_solver_stream(self._h, ctypes.byref(p)))
        return int(p.value or 0)

    def profile(self, nsteps: int) -> dict:
        """Runs nsteps with HIP events around every launch; returns
        {kernel: (launches, mean ms)}."""
        mx = 16
        names = (ctypes.c_char_p * mx)()
        launches = (ctypes.c_int * mx)()
        ms = (ctypes.c_double * mx)()
        n = self._L.ma_solver_profile(self._h, int(nsteps), names, launches, ms, mx)
        if n < 0:
            _lib.check(_lib.MA_ERR_DEVICE)
        return {names[i].decode(): (launches[i], ms[i]) for i in range(n)}

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._L.ma_solver_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class BatchSolver(TimeSolverExplicitRK4):
    """Independent cases stepped together in one set of launches
    (ma_batch_create).  The reference runs one mesh per process (Main.C:96-141);
    here each case keeps its own mesh, initial state, inflow and dt or CFL,
    and its result equals a TimeSolverExplicitRK4 run on it alone bit for bit.
    state()/residual()/gradients() cover all cases' cells in order."""

    def __init__(self, cases):
        cases = list(cases)
        if not cases:
            raise ValueError("BatchSolver needs at least one case")
        self.cases = cases
        self.mesh_data = cases[0][0]
        self.options = cases[0][1]
        self._L = _lib.load()
        self._c_meshes = [md.to_c() for md, _ in cases]
        n = len(cases)
        ptrs = (ctypes.POINTER(_lib.ma_mesh_data) * n)(*[ctypes.pointer(m) for m in self._c_meshes])
        opts = (_lib.ma_options * n)(*[o.to_c() for _, o in cases])
        self._keep = (ptrs, opts)
        h = ctypes.c_void_p()
        _lib.check(self._L.ma_batch_create(n, ptrs, opts, ctypes.byref(h)))
        self._h = h
        self.ncases = n
        off = (ctypes.c_int * (n + 1))()
        nn = ctypes.c_int()
        _lib.check(self._L.ma_batch_sizes(self._h, ctypes.byref(nn), off))
        self.offsets = np.array(off[:], dtype=np.int64)
        self.ncells = int(self.offsets[-1])

    def Solve(self, *a, **kw):
        raise NotImplementedError("BatchSolver: step the cases with initialize()/step()")

    def case_state(self, k: int) -> np.ndarray:
        U = np.empty((int(self.offsets[k + 1] - self.offsets[k]), 5))
        _lib.check(self._L.ma_batch_get_case_state(self._h, int(k), U.ctypes.data_as(_lib._D)))
        return U

    def case_time(self, k: int):
        t, d = ctypes.c_double(), ctypes.c_double()
        _lib.check(self._L.ma_batch_get_case_time(self._h, int(k), ctypes.byref(t), ctypes.byref(d)))
        return t.value, d.value
