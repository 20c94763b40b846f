"""Named workloads.

The reference runs only its three hard-coded problems (problem_type 0/1/2,
Parallel3DMesh.h:382-396, TimeSolverExplicitRK4.h:218-223).  The benchmark
workloads below use the config-mode extensions of the options struct to set
the freestream from a Mach and Reynolds number, stretch the mesh toward the
wall, choose boundary kinds per side and pick dt from a CFL number.
"""
from __future__ import annotations

import math

import numpy as np

from .options import Options

RGAS = 287.05    # GasModel.h:39
GAMMA = 1.4      # GasModel.h:51


def sutherland(T: float) -> float:
    """GasModel.h:92-98."""
    return 1.458e-6 * T * math.sqrt(T) / (T + 110.4)


def freestream_state(mach: float, T: float, re_per_m: float, angle_deg: float = 0.0,
                     length: float = 1.0) -> tuple:
    """Conserved state (rho, rho u, rho v, rho w, rho E) of a freestream at
    the given Mach number, temperature and unit Reynolds number; the flow
    direction is turned by angle_deg in the x-y plane."""
    c = math.sqrt(GAMMA * RGAS * T)
    speed = mach * c
    rho = re_per_m * sutherland(T) / (speed * length) if re_per_m > 0 else 1.0
    a = math.radians(angle_deg)
    u, v = speed * math.cos(a), speed * math.sin(a)
    E = rho * (RGAS / (GAMMA - 1.0)) * T + 0.5 * rho * speed * speed
    return (rho, rho * u, rho * v, 0.0, E)


def sod_c1(**kw) -> Options:
    """3D Sod shock tube on [0,1]x[0,0.25]x[0,0.25], 64x16x16, first order,
    Extrapolate on every side, rho/p = 1/1 left, 0.125/0.1 right, dt 1e-4."""
    o = Options(problem_type=0, lx=1.0, ly=0.25, lz=0.25, nx=64, ny=16, nz=16, ntimesteps=100,
                dt=1e-4, second_order_space=0, viscous=0, bc_kind=(1, 1, 1, 1, 1, 1), ic_kind=1,
                ic=(1.0, 1.0, 0.125, 0.1, 0.5))
    return o.replace(**kw)


def flat_plate_c2(**kw) -> Options:
    """Viscous flat plate, 256x128x64 on [0,1]x[0,0.05]x[0,0.025], y-spacing
    growing by 1.05 away from the wall; M 0.3, T 300 K, Re 1e5 per metre,
    Pr 0.72; Inflow -x, Extrapolate +x, NoSlip -y, Tangent +y and +-z; Roe +
    viscous, Green-Gauss, Venkat, second order; CFL 0.5, 200 steps."""
    o = Options(problem_type=1, lx=1.0, ly=0.05, lz=0.025, nx=256, ny=128, nz=64, ntimesteps=200,
                dt=0.0, second_order_space=1, viscous=1, y_stretch=1.05,
                bc_kind=(3, 2, 2, 2, 1, 4), use_inflow=1,
                inflow=freestream_state(0.3, 300.0, 1.0e5), prandtl=0.72, cfl=0.5)
    return o.replace(**kw)


def flat_plate_c3(**kw) -> Options:
    """The 256x128x64 flat plate's physics on 512x256x64, started from the freestream with a seeded
    +-1% perturbation of density and velocity, 500 steps, flux/gradient/limiter
    arithmetic in FP32 with FP64 state and sums."""
    return flat_plate_c2(nx=512, ny=256, nz=64, ntimesteps=500, ic_kind=2, perturb=0.01, seed=3,
                         precision=1).replace(**kw)


def batched_c4(ncases: int = 512, seed: int = 4, **kw) -> list:
    """Independent Sod-type cases stepped together: cell counts N drawn
    log-uniform over [8192, 262144], mesh 4m x m x m with m = round((N/4)^(1/3))
    on [0,1]x[0,0.25]x[0,0.25]; left/right density and pressure U[0.1, 1],
    interface position U[0.3, 0.7]; Extrapolate on every side, second order,
    Van Albada, dt 1e-4 (largest sound speed sqrt(1.4 * 1 / 0.1) = 3.74 on
    dx >= 1/160 gives a Courant number below 0.06), 100 steps; flux tile
    shape picked by the autotuner.  Draws from
    numpy's default_rng(seed) in the order N, rho_L, rho_R, p_L, p_R, x0."""
    rng = np.random.default_rng(seed)
    n = rng.uniform(math.log(8192.0), math.log(262144.0), ncases)
    rho_l, rho_r = rng.uniform(0.1, 1.0, ncases), rng.uniform(0.1, 1.0, ncases)
    p_l, p_r = rng.uniform(0.1, 1.0, ncases), rng.uniform(0.1, 1.0, ncases)
    x0 = rng.uniform(0.3, 0.7, ncases)
    out = []
    for i in range(ncases):
        m = max(1, int(round((math.exp(n[i]) / 4.0) ** (1.0 / 3.0))))
        o = Options(problem_type=0, lx=1.0, ly=0.25, lz=0.25, nx=4 * m, ny=m, nz=m, ntimesteps=100,
                    dt=1e-4, second_order_space=1, viscous=0, limiter=1, ic_kind=1,
                    ic=(float(rho_l[i]), float(p_l[i]), float(rho_r[i]), float(p_r[i]), float(x0[i])),
                    bc_kind=(1, 1, 1, 1, 1, 1), autotune=1)
        out.append(o.replace(**kw))
    return out


def served_case(viscous: bool, mach: float, angle_deg: float, cells: int, steps: int) -> Options:
    """One request of the served stream: the stretched flat-plate box (4:2:1 cells,
    m = round((cells/8)^(1/3)), nx = 4m, ny = 2m, nz = m) with a freestream
    at the given Mach number and inflow angle, T 300 K, Re 1e5 per metre,
    CFL 0.5, second order with Venkat, FP32 flux/gradient/limiter arithmetic
    and FP64 state.  Viscous: the 256x128x64 plate's boundaries and y-stretching.  Inviscid:
    the same box with a slip (Tangent) wall at -y and uniform spacing."""
    m = max(1, int(round((cells / 8.0) ** (1.0 / 3.0))))
    o = flat_plate_c2(nx=4 * m, ny=2 * m, nz=m, ntimesteps=int(steps),
                      inflow=freestream_state(mach, 300.0, 1.0e5, angle_deg), precision=1)
    if not viscous:
        o = o.replace(viscous=0, y_stretch=0.0, bc_kind=(2, 2, 2, 2, 1, 4))
    return o


def served_c5(ncases: int = 2000, seed: int = 5) -> list:
    """The served stream: (arrival time in s, Options) for ncases requests.
    Poisson arrivals at 200 per second (exponential gaps, mean 5 ms), viscous
    or inviscid with probability 1/2, cells log-uniform over [4096, 2097152],
    Mach U[0.2, 0.8], inflow angle U[-5, 5] degrees, steps uniform over
    50..500.  Draws from numpy's default_rng(seed) in the order gaps, kind,
    cells, Mach, angle, steps."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / 200.0, ncases)
    viscous = rng.random(ncases) < 0.5
    cells = np.exp(rng.uniform(math.log(4096.0), math.log(2097152.0), ncases))
    mach = rng.uniform(0.2, 0.8, ncases)
    angle = rng.uniform(-5.0, 5.0, ncases)
    steps = rng.integers(50, 501, ncases)
    arrival = np.cumsum(gaps)
    return [(float(arrival[i]), served_case(bool(viscous[i]), float(mach[i]), float(angle[i]),
                                            float(cells[i]), int(steps[i])))
            for i in range(ncases)]


WORKLOADS = {"sod_c1": sod_c1, "flat_plate_c2": flat_plate_c2, "flat_plate_c3": flat_plate_c3}
BATCH_WORKLOADS = {"batched_c4": batched_c4}
SERVED_WORKLOADS = {"served_c5": served_c5}
