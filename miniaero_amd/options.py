"""Options of a run: the nine values of miniaero.inp (Options.h:47-101) plus
the config-mode extensions of include/miniaero.h."""
from __future__ import annotations

import ctypes
import dataclasses
from dataclasses import dataclass, field

from . import _lib


@dataclass
class Options:
    # Options.h:49-57, with the constructor defaults of Options.h:59-70
    problem_type: int = 0
    lx: float = 1.0
    ly: float = 1.0
    lz: float = 1.0
    angle: float = 0.0
    nx: int = 10
    ny: int = 10
    nz: int = 10
    ntimesteps: int = 1
    dt: float = 5e-8
    output_results: int = 0
    output_frequency: int = 10
    second_order_space: int = 0
    viscous: int = 0
    # extensions; zero values reproduce the reference
    bc_kind: tuple = (0, 0, 0, 0, 0, 0)
    y_stretch: float = 0.0
    limiter: int = 0
    ic_kind: int = 0
    ic: tuple = (0.0, 0.0, 0.0, 0.0, 0.0)
    use_inflow: int = 0
    inflow: tuple = (0.0, 0.0, 0.0, 0.0, 0.0)
    prandtl: float = 0.0
    cfl: float = 0.0
    perturb: float = 0.0
    seed: int = 0
    precision: int = 0        # 0 FP64, 1 FP32 flux/gradient/limiter arithmetic
    tile_shape: tuple = (0, 0, 0)   # flux tile brick (i, j, k); zeros = 4x8x8 (4x4x4 below 256 cells
                                    # per CU), (-1, ..) = consecutive
    autotune: int = 0         # 1: pick the fastest tile shape / resident workgroups at creation
    workgroups_per_cu: int = 0  # flux kernel: resident workgroups per CU (0: one per tile)

    def read_options_file(self, path: str = "miniaero.inp") -> "Options":
        """Options::read_options_file (Options.h:73-101): nine whitespace
        separated values.  A missing or short file raises (the reference
        prints a message and continues with whatever was read)."""
        o = _lib.ma_options()
        _lib.check(_lib.load().ma_options_read(path.encode(), ctypes.byref(o)))
        for f in ("problem_type", "lx", "ly", "lz", "angle", "nx", "ny", "nz", "ntimesteps", "dt",
                  "output_results", "output_frequency", "second_order_space", "viscous"):
            setattr(self, f, getattr(o, f))
        return self

    def to_c(self) -> _lib.ma_options:
        o = _lib.ma_options()
        for f in dataclasses.fields(self):
            v = getattr(self, f.name)
            if f.name in ("bc_kind", "ic", "inflow", "tile_shape"):
                arr = getattr(o, f.name)
                for i, x in enumerate(v):
                    arr[i] = x
            else:
                setattr(o, f.name, v)
        return o

    def replace(self, **kw) -> "Options":
        return dataclasses.replace(self, **kw)

    def as_dict(self) -> dict:
        return dataclasses.asdict(self)
