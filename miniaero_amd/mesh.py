"""Mesh side of the host mirror: Faces, MeshData and Parallel3DMesh.

Mirrors the reference's setup objects (Faces.h:42-71, Cells.h:41-76,
MeshData.h:42-58, Parallel3DMesh.h:56-59).  Parallel3DMesh.fillMeshData()
calls the library's closed-form generator (ma_mesh_generate), which produces
the reference's cells, faces, face order and boundary groups bit for bit.
A MeshData built by hand from numpy arrays is accepted by the solver too.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from .options import Options


@dataclass
class Faces:
    """One Faces<Device> as host arrays (row-major, n x 3 / n x 2 / n)."""
    coordinates: np.ndarray
    face_normal: np.ndarray
    face_tangent: np.ndarray
    face_binormal: np.ndarray
    face_cell_conn: np.ndarray
    cell_flux_index: np.ndarray | None = None
    permute_vector: np.ndarray | None = None

    @property
    def nfaces_(self) -> int:
        return int(self.face_cell_conn.shape[0])

    def _c(self) -> _lib.ma_faces:
        """C view of the arrays; the arrays it points to are kept alive in the
        returned object's `_keep` list."""
        f = _lib.ma_faces()
        f.nfaces = self.nfaces_
        D, I = _lib._D, _lib._I
        keep = [np.ascontiguousarray(self.coordinates, np.float64),
                np.ascontiguousarray(self.face_normal, np.float64),
                np.ascontiguousarray(self.face_tangent, np.float64),
                np.ascontiguousarray(self.face_binormal, np.float64),
                np.ascontiguousarray(self.face_cell_conn, np.int32)]
        f.coordinates, f.face_normal, f.face_tangent, f.face_binormal = [
            a.ctypes.data_as(D) for a in keep[:4]]
        f.face_cell_conn = keep[4].ctypes.data_as(I)
        if self.cell_flux_index is not None:
            a = np.ascontiguousarray(self.cell_flux_index, np.int32)
            keep.append(a)
            f.cell_flux_index = a.ctypes.data_as(I)
        if self.permute_vector is not None:
            a = np.ascontiguousarray(self.permute_vector, np.int32)
            keep.append(a)
            f.permute_vector = a.ctypes.data_as(I)
        f._keep = keep
        return f


@dataclass
class MeshData:
    """MeshData<Device> for one rank (num_ghosts == 0)."""
    num_owned_cells: int
    cell_coordinates: np.ndarray
    cell_volumes: np.ndarray
    internal_faces: Faces
    boundary_faces: list = field(default_factory=list)  # [(name, Faces)]
    num_ghosts: int = 0

    def plan(self) -> dict:
        """Host-only pre-flight check (ma_mesh_plan): validates the arrays and
        reports the device layout the solver would build."""
        c = self.to_c()
        st = (ctypes.c_longlong * 7)()
        _lib.check(_lib.load().ma_mesh_plan(ctypes.byref(c), st))
        return _plan_dict(st)


    def to_c(self) -> _lib.ma_mesh_data:
        """C view of the mesh; every array it points to is kept alive in the
        returned object's `_keep` list, so the view stays valid as long as
        the returned object does."""
        m = _lib.ma_mesh_data()
        m.num_owned_cells = int(self.num_owned_cells)
        m.num_ghosts = int(self.num_ghosts)
        keep = [np.ascontiguousarray(self.cell_coordinates, np.float64),
                np.ascontiguousarray(self.cell_volumes, np.float64)]
        m.cell_coordinates = keep[0].ctypes.data_as(_lib._D)
        m.cell_volumes = keep[1].ctypes.data_as(_lib._D)
        fi = self.internal_faces._c()
        keep.append(fi._keep)
        m.internal_faces = fi
        groups = (_lib.ma_bc_group * max(1, len(self.boundary_faces)))()
        for g, (name, faces) in enumerate(self.boundary_faces):
            groups[g].kind = _lib.BC_KINDS[name]
            fg = faces._c()
            keep.append(fg._keep)
            groups[g].faces = fg
        keep.append(groups)
        m._keep = keep
        m.num_boundary_groups = len(self.boundary_faces)
        m.boundary_faces = ctypes.cast(groups, ctypes.POINTER(_lib.ma_bc_group))
        return m


def _faces_from_view(v: _lib.ma_faces) -> Faces:
    n = v.nfaces

    def arr(p, shape, dt):
        if n == 0:
            return np.zeros(shape, dt)
        return np.ctypeslib.as_array(p, shape=shape)

    return Faces(coordinates=arr(v.coordinates, (n, 3), np.float64),
                 face_normal=arr(v.face_normal, (n, 3), np.float64),
                 face_tangent=arr(v.face_tangent, (n, 3), np.float64),
                 face_binormal=arr(v.face_binormal, (n, 3), np.float64),
                 face_cell_conn=arr(v.face_cell_conn, (n, 2), np.int32),
                 cell_flux_index=arr(v.cell_flux_index, (n, 2), np.int32),
                 permute_vector=arr(v.permute_vector, (n,), np.int32))


class _MeshOwner:
    """Keeps a generated ma_mesh alive while numpy views of it exist."""

    def __init__(self, handle):
        self.handle = handle

    def __del__(self):
        try:
            _lib.load().ma_mesh_destroy(self.handle)
        except Exception:
            pass


class Parallel3DMesh:
    """Parallel3DMesh(nx, ny, nz, lx, ly, lz, problem_type, angle) for one rank
    (Parallel3DMesh.h:56-59).  `options` carries the config-mode extensions
    (y_stretch, bc_kind) when given."""

    def __init__(self, nx, ny, nz, lx, ly, lz, problem_type, angle=0.0, options: Options | None = None):
        base = options if options is not None else Options()
        self.options = base.replace(nx=nx, ny=ny, nz=nz, lx=lx, ly=ly, lz=lz,
                                    problem_type=problem_type, angle=angle)

    def fillMeshData(self) -> MeshData:
        L = _lib.load()
        h = ctypes.c_void_p()
        o = self.options.to_c()
        _lib.check(L.ma_mesh_generate(ctypes.byref(o), ctypes.byref(h)))
        owner = _MeshOwner(h)
        v = _lib.ma_mesh_data()
        _lib.check(L.ma_mesh_view(h, ctypes.byref(v)))
        nc = v.num_owned_cells
        md = MeshData(
            num_owned_cells=nc,
            cell_coordinates=np.ctypeslib.as_array(v.cell_coordinates, shape=(nc, 3)),
            cell_volumes=np.ctypeslib.as_array(v.cell_volumes, shape=(nc,)),
            internal_faces=_faces_from_view(v.internal_faces),
            boundary_faces=[(_lib.BC_NAMES[v.boundary_faces[g].kind],
                             _faces_from_view(v.boundary_faces[g].faces))
                            for g in range(v.num_boundary_groups)])
        md._owner = owner  # the arrays above are views into the generated mesh
        return md


def _plan_dict(st) -> dict:
    return dict(lattice=(st[0], st[1], st[2]), tiles=st[3], max_tile_faces=st[4],
                tile_face_entries=st[5], derived_frames=bool(st[6]))


def plan_batch(meshes) -> dict:
    """ma_batch_plan: the checked layout of several meshes stepped together."""
    cs = [md.to_c() for md in meshes]
    ptrs = (ctypes.POINTER(_lib.ma_mesh_data) * len(cs))(*[ctypes.pointer(c) for c in cs])
    st = (ctypes.c_longlong * 7)()
    _lib.check(_lib.load().ma_batch_plan(len(cs), ptrs, st))
    return _plan_dict(st)


def generate_meshes(options_list, max_workers: int | None = None) -> list:
    """Parallel3DMesh(...).fillMeshData() for every Options of the list, on
    up to max_workers threads (the generator is thread-safe and runs outside
    the interpreter lock); results in list order."""
    from concurrent.futures import ThreadPoolExecutor
    import os

    def one(o):
        return Parallel3DMesh(o.nx, o.ny, o.nz, o.lx, o.ly, o.lz, o.problem_type, o.angle,
                              options=o).fillMeshData()
    workers = max_workers or min(16, os.cpu_count() or 1)
    if workers <= 1 or len(options_list) <= 1:
        return [one(o) for o in options_list]
    with ThreadPoolExecutor(max_workers=workers) as ex:
        return list(ex.map(one, options_list))
