"""ctypes wrapper of the CPU oracle (oracle/liboracle.so) and of the reference
mesh build (oracle/_ref/libref_mesh.so).

TEST INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and the cpu_baseline
leg of bench.py load this module, as the checker.  The product never does.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_LIB_PATH = os.path.join(HERE, "_ref", "libref_mesh.so")
REF_SRC_DIR = "/root/reference/kokkos"

BC_NAMES = {1: "Extrapolate", 2: "Tangent", 3: "NoSlip", 4: "Inflow"}


class orc_options(ctypes.Structure):
    _fields_ = [
        ("problem_type", ctypes.c_int),
        ("lx", ctypes.c_double), ("ly", ctypes.c_double), ("lz", ctypes.c_double),
        ("angle", ctypes.c_double),
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
        ("ntimesteps", ctypes.c_int),
        ("dt", ctypes.c_double),
        ("second_order_space", ctypes.c_int),
        ("viscous", ctypes.c_int),
        ("bc_kind", ctypes.c_int * 6),
        ("y_stretch", ctypes.c_double),
        ("limiter", ctypes.c_int),
        ("ic_kind", ctypes.c_int),
        ("ic", ctypes.c_double * 5),
        ("use_inflow", ctypes.c_int),
        ("inflow", ctypes.c_double * 5),
        ("prandtl", ctypes.c_double),
        ("cfl", ctypes.c_double),
        ("perturb", ctypes.c_double),
        ("seed", ctypes.c_ulonglong),
    ]


def make_options(**kw) -> orc_options:
    o = orc_options()
    for k, v in kw.items():
        if k in ("bc_kind", "ic", "inflow"):
            arr = getattr(o, k)
            for i, x in enumerate(v):
                arr[i] = x
        else:
            setattr(o, k, v)
    return o


def build(force: bool = False) -> None:
    """Compile liboracle.so, and the reference mesh library when the reference
    sources are present (they are not on the GPU machine)."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)
    if os.path.isdir(REF_SRC_DIR) and (force or not os.path.exists(REF_LIB_PATH)):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


_lib = None
_ref = None

_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)


def _dp(a):
    return a.ctypes.data_as(_D) if a is not None else None


def _ip(a):
    return a.ctypes.data_as(_I) if a is not None else None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.c_void_p
        L.orc_mesh_new.restype = P
        L.orc_mesh_new.argtypes = [ctypes.POINTER(orc_options)]
        L.orc_mesh_free.argtypes = [P]
        L.orc_mesh_counts.argtypes = [P, _I, ctypes.c_int]
        L.orc_mesh_cells.argtypes = [P, _D, _D]
        L.orc_mesh_nodes.argtypes = [P, _D]
        L.orc_mesh_faces.argtypes = [P, ctypes.c_int, _D, _D, _D, _D, _I, _I, _I, _I]
        L.orc_solver_new.restype = P
        L.orc_solver_new.argtypes = [P, ctypes.POINTER(orc_options)]
        L.orc_solver_free.argtypes = [P]
        L.orc_solver_init.argtypes = [P]
        L.orc_solver_set_state.argtypes = [P, _D]
        L.orc_solver_step.argtypes = [P, ctypes.c_int]
        L.orc_solver_get_state.argtypes = [P, _D]
        L.orc_solver_get_residual.argtypes = [P, _D]
        L.orc_solver_get_gradients.argtypes = [P, _D, _D]
        L.orc_solver_time.restype = ctypes.c_double
        L.orc_solver_time.argtypes = [P]
        L.orc_solver_last_dt.restype = ctypes.c_double
        L.orc_solver_last_dt.argtypes = [P]
        L.orc_primitives.argtypes = [_D, _D]
        L.orc_roe_flux.argtypes = [_D, _D, _D, _D, _D, _D]
        L.orc_viscous_flux.argtypes = [_D, _D, _D, ctypes.c_double, _D]
        L.orc_venkat.restype = ctypes.c_double
        L.orc_venkat.argtypes = [ctypes.c_double] * 4
        L.orc_vanalbada.restype = ctypes.c_double
        L.orc_vanalbada.argtypes = [ctypes.c_double] * 3
        L.orc_hex_volume_centroid.argtypes = [_D, _D, _D, _D, _D]
        _lib = L
    return _lib


def ref_available() -> bool:
    return os.path.exists(REF_LIB_PATH)


def ref_lib():
    global _ref
    if _ref is None:
        L = ctypes.CDLL(REF_LIB_PATH)
        P = ctypes.c_void_p
        L.ref_mesh_new.restype = P
        L.ref_mesh_new.argtypes = [ctypes.c_int] * 3 + [ctypes.c_double] * 4
        L.ref_mesh_free.argtypes = [P]
        L.ref_mesh_counts.argtypes = [P, _I]
        L.ref_mesh_cells.argtypes = [P, _D, _D]
        L.ref_mesh_faces.argtypes = [P, ctypes.c_int, _D, _D, _D, _D, _I, _I]
        _ref = L
    return _ref


def _faces_dict(nf):
    return dict(xf=np.zeros((nf, 3)), a=np.zeros((nf, 3)), t=np.zeros((nf, 3)),
                b=np.zeros((nf, 3)), conn=np.zeros((nf, 2), np.int32),
                fidx=np.zeros((nf, 2), np.int32))


class OracleMesh:
    """Mesh built by the oracle's restatement of Parallel3DMesh::fillMeshData."""

    def __init__(self, opts: orc_options):
        L = lib()
        self.opts = opts
        self.h = L.orc_mesh_new(ctypes.byref(opts))
        c = np.zeros(16, np.int32)
        n = L.orc_mesh_counts(self.h, _ip(c), 16)
        self.ncells, self.n_internal, ng = int(c[0]), int(c[1]), int(c[2])
        self.group_sizes = [int(x) for x in c[3:3 + ng]]
        self.cell_xyz = np.zeros((self.ncells, 3))
        self.cell_vol = np.zeros(self.ncells)
        L.orc_mesh_cells(self.h, _dp(self.cell_xyz), _dp(self.cell_vol))
        self.internal = self._faces(-1, self.n_internal)
        self.groups = [self._faces(g, nf) for g, nf in enumerate(self.group_sizes)]

    def nodes(self):
        """Node coordinates as an (nx+1, ny+1, nz+1, 3) array indexed by (i, j, k)."""
        o = self.opts
        X = np.zeros(((o.nx + 1) * (o.ny + 1) * (o.nz + 1), 3))
        lib().orc_mesh_nodes(self.h, _dp(X))
        return X.reshape(o.nx + 1, o.ny + 1, o.nz + 1, 3)

    def _faces(self, g, nf):
        d = _faces_dict(nf)
        d["perm"] = np.zeros(nf, np.int32)
        k = ctypes.c_int(0)
        lib().orc_mesh_faces(self.h, g, _dp(d["xf"]), _dp(d["a"]), _dp(d["t"]), _dp(d["b"]),
                             _ip(d["conn"]), _ip(d["fidx"]), _ip(d["perm"]), ctypes.byref(k))
        d["kind"] = int(k.value)
        return d

    def __del__(self):
        try:
            lib().orc_mesh_free(self.h)
        except Exception:
            pass


class RefMesh:
    """Mesh built by the reference's own MeshProcessor/Face/ElementTopoHexa8 code."""

    def __init__(self, nx, ny, nz, lx, ly, lz, angle=0.0):
        L = ref_lib()
        self.h = L.ref_mesh_new(nx, ny, nz, lx, ly, lz, angle)
        c = np.zeros(9, np.int32)
        L.ref_mesh_counts(self.h, _ip(c))
        self.ncells, self.n_internal = int(c[0]), int(c[1])
        self.group_sizes = [int(x) for x in c[3:9]]
        self.cell_xyz = np.zeros((self.ncells, 3))
        self.cell_vol = np.zeros(self.ncells)
        L.ref_mesh_cells(self.h, _dp(self.cell_xyz), _dp(self.cell_vol))
        self.internal = self._faces(-1, self.n_internal)
        self.groups = [self._faces(g, nf) for g, nf in enumerate(self.group_sizes)]

    def _faces(self, g, nf):
        d = _faces_dict(nf)
        ref_lib().ref_mesh_faces(self.h, g, _dp(d["xf"]), _dp(d["a"]), _dp(d["t"]), _dp(d["b"]),
                                 _ip(d["conn"]), _ip(d["fidx"]))
        return d

    def __del__(self):
        try:
            ref_lib().ref_mesh_free(self.h)
        except Exception:
            pass


class OracleSolver:
    """Fixed-order CPU restatement of TimeSolverExplicitRK4::Solve."""

    def __init__(self, mesh: OracleMesh, opts: orc_options | None = None):
        self.mesh = mesh
        self.opts = opts if opts is not None else mesh.opts
        self.h = lib().orc_solver_new(mesh.h, ctypes.byref(self.opts))
        self.nc = mesh.ncells

    def init(self):
        lib().orc_solver_init(self.h)

    def set_state(self, U):
        U = np.ascontiguousarray(U, dtype=np.float64)
        lib().orc_solver_set_state(self.h, _dp(U))

    def step(self, n=1):
        lib().orc_solver_step(self.h, int(n))

    def state(self):
        U = np.zeros((self.nc, 5))
        lib().orc_solver_get_state(self.h, _dp(U))
        return U

    def residual(self):
        R = np.zeros((self.nc, 5))
        lib().orc_solver_get_residual(self.h, _dp(R))
        return R

    def gradients(self):
        g = np.zeros((self.nc, 5, 3))
        lim = np.zeros((self.nc, 5))
        lib().orc_solver_get_gradients(self.h, _dp(g), _dp(lim))
        return g, lim

    @property
    def time(self):
        return lib().orc_solver_time(self.h)

    @property
    def last_dt(self):
        return lib().orc_solver_last_dt(self.h)

    def __del__(self):
        try:
            lib().orc_solver_free(self.h)
        except Exception:
            pass


def roe_flux(VL, VR, n, t, b):
    F = np.zeros(5)
    args = [np.ascontiguousarray(x, dtype=np.float64) for x in (VL, VR, n, t, b)]
    lib().orc_roe_flux(*[_dp(x) for x in args], _dp(F))
    return F


def viscous_flux(grad, V, a, prandtl=0.0):
    F = np.zeros(5)
    g = np.ascontiguousarray(grad, dtype=np.float64).reshape(15)
    V = np.ascontiguousarray(V, dtype=np.float64)
    a = np.ascontiguousarray(a, dtype=np.float64)
    lib().orc_viscous_flux(_dp(g), _dp(V), _dp(a), float(prandtl), _dp(F))
    return F


def primitives(U):
    V = np.zeros(5)
    U = np.ascontiguousarray(U, dtype=np.float64)
    lib().orc_primitives(_dp(U), _dp(V))
    return V


def venkat(dumax, dumin, du, dx2):
    return lib().orc_venkat(dumax, dumin, du, dx2)


def vanalbada(dumax, dumin, du):
    return lib().orc_vanalbada(dumax, dumin, du)


def hex_volume_centroid(ex, ey, ez):
    vol = np.zeros(1)
    c = np.zeros(3)
    ex, ey, ez = [np.ascontiguousarray(v, dtype=np.float64) for v in (ex, ey, ez)]
    lib().orc_hex_volume_centroid(_dp(ex), _dp(ey), _dp(ez), _dp(vol), _dp(c))
    return float(vol[0]), c
