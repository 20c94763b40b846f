"""Pins the CPU oracle before it is trusted as the checker.

1. Mesh: bit-exact against oracle/_ref, the reference's own MeshProcessor.C /
   Face.C / ElementTopoHexa8.C compiled in place (skipped where the reference
   sources are absent, e.g. on the GPU machine).
2. End to end: the reference's six gold result sets, compared with the
   semantics and tolerances of the reference's numeric_text_diff runners.
3. Scalar kernels: closed forms and invariants, chosen so that a dropped term,
   a wrong sign or a transposed operand fails at least one of them.
"""
import math
import os

import numpy as np
import pytest

from conftest import GOLD_CASES, GOLDEN, gold_rows, read_inp, rows_text
from oracle import oracle as O
from miniaero_amd.results import numeric_text_diff

MESH_FIELDS = ("xf", "a", "t", "b", "conn", "fidx")


@pytest.mark.parametrize("kw", [
    dict(nx=128, ny=4, nz=4, lx=0.3048, ly=1.0, lz=1.0, angle=0.0),
    dict(nx=64, ny=32, nz=2, lx=2.0, ly=2.0, lz=1.0, angle=30.0),
    dict(nx=16, ny=32, nz=2, lx=2.0, ly=0.002, lz=1.0, angle=0.0),
    dict(nx=3, ny=2, nz=4, lx=1.0, ly=1.0, lz=1.0, angle=15.0),
    dict(nx=1, ny=1, nz=1, lx=1.0, ly=1.0, lz=1.0, angle=0.0),
])
def test_oracle_mesh_matches_reference_code(kw):
    if not O.ref_available():
        pytest.skip("reference mesh library (oracle/_ref) not built: reference sources absent")
    om = O.OracleMesh(O.make_options(problem_type=0, **kw))
    rm = O.RefMesh(kw["nx"], kw["ny"], kw["nz"], kw["lx"], kw["ly"], kw["lz"], kw["angle"])
    assert om.ncells == rm.ncells and om.n_internal == rm.n_internal
    assert om.group_sizes == rm.group_sizes
    assert np.array_equal(om.cell_xyz, rm.cell_xyz)
    assert np.array_equal(om.cell_vol, rm.cell_vol)
    for f in MESH_FIELDS:
        assert np.array_equal(om.internal[f], rm.internal[f]), f
        for g in range(6):
            assert np.array_equal(om.groups[g][f], rm.groups[g][f]), (g, f)


def test_oracle_mesh_counts_closed_form():
    nx, ny, nz = 5, 3, 4
    om = O.OracleMesh(O.make_options(problem_type=1, nx=nx, ny=ny, nz=nz, lx=1, ly=1, lz=1))
    assert om.ncells == nx * ny * nz
    assert om.n_internal == (nx - 1) * ny * nz + nx * (ny - 1) * nz + nx * ny * (nz - 1)
    assert om.group_sizes == [nx * nz, nx * nz, nx * ny, nx * ny, ny * nz, ny * nz]
    # uniform box: every volume is dx*dy*dz, every internal normal points left -> right
    assert np.allclose(om.cell_vol, 1.0 / (nx * ny * nz), rtol=1e-12)
    L, R = om.internal["conn"][:, 0], om.internal["conn"][:, 1]
    d = om.cell_xyz[R] - om.cell_xyz[L]
    assert np.all(np.einsum("ij,ij->i", d, om.internal["a"]) > 0)
    # processing order is a stable sort by left cell
    perm = om.internal["perm"]
    assert np.all(np.diff(L[perm]) >= 0)
    # boundary kinds for the flat plate (Parallel3DMesh.h:382-396)
    kinds = [g["kind"] for g in om.groups]
    assert kinds == [3, 1, 2, 2, 1, 4]


def _run_oracle_case(case):
    kw = read_inp(os.path.join(GOLDEN, case, "miniaero.inp"))
    o = O.make_options(**{k: v for k, v in kw.items() if k not in ("output_results", "output_frequency")})
    m = O.OracleMesh(o)
    s = O.OracleSolver(m)
    s.init()
    s.step(kw["ntimesteps"])
    return m, s.state()


_cache = {}


@pytest.mark.parametrize("case,ranks,rtol,floor", GOLD_CASES)
def test_oracle_matches_gold(case, ranks, rtol, floor):
    key = open(os.path.join(GOLDEN, case, "miniaero.inp")).read()
    if key not in _cache:
        _cache[key] = _run_oracle_case(case)
    m, U = _cache[key]
    mine = np.hstack([m.cell_xyz, U])
    gold = gold_rows(case, ranks, m.cell_xyz)
    same, msg = numeric_text_diff(rows_text(mine), rows_text(gold), rtol, floor)
    assert same, msg[:2000]


# ---------------------------------------------------------------- pieces --

def _state(rho, u, v, w, T):
    return np.array([rho, u, v, w, T], float)


def _physical_flux(V, n):
    rho, u, v, w, T = V
    p = rho * 287.05 * T
    H = 1004.0 * T + 0.5 * (u * u + v * v + w * w)
    un = n[0] * u + n[1] * v + n[2] * w
    return np.array([rho * un, rho * un * u + n[0] * p, rho * un * v + n[1] * p,
                     rho * un * w + n[2] * p, rho * un * H])


def _frame(n):
    n = np.asarray(n, float)
    i1 = int(np.argmax(np.abs(n)))
    i2, i3 = (i1 + 1) % 3, (i1 + 2) % 3
    t = np.zeros(3)
    den = math.sqrt(n[i1] ** 2 + n[i3] ** 2)
    t[i1], t[i3] = n[i3] / den, -n[i1] / den
    return n, t, np.cross(n, t)


def test_roe_consistency_equals_physical_flux():
    rng = np.random.default_rng(1)
    for _ in range(50):
        V = _state(rng.uniform(0.1, 2), *rng.uniform(-300, 300, 3), rng.uniform(200, 400))
        n, t, b = _frame(rng.normal(size=3) * rng.uniform(0.01, 3))
        F = O.roe_flux(V, V, n, t, b)
        assert np.allclose(F, _physical_flux(V, n), rtol=1e-13, atol=1e-9 * np.abs(_physical_flux(V, n)).max())


def test_roe_supersonic_upwinding():
    # All eigenvalues positive and above the entropy-fix threshold: F = F(V_L).
    # Exact when only the velocity jumps; with a temperature jump the
    # reference's gas model is not self-consistent (Roe uses gamma - 1 = 0.4
    # while Cp = 1004, R = 287.05 imply gamma = 1.40037), which leaves a
    # residue of order 1e-4 of the flux jump in the energy component.
    n, t, b = _frame([0.3, 0.0, 0.0])
    VL, VR = _state(1.0, 900.0, 0, 0, 300.0), _state(1.0, 850.0, 0, 0, 300.0)
    F = O.roe_flux(VL, VR, n, t, b)
    jump = np.abs(_physical_flux(VR, n) - _physical_flux(VL, n)).max()
    assert np.abs(F - _physical_flux(VL, n)).max() <= 1e-12 * jump
    VL = _state(1.0, 900.0, 10.0, -5.0, 300.0)   # Mach ~2.6 in +x
    VR = _state(0.5, 800.0, 0.0, 0.0, 250.0)
    F = O.roe_flux(VL, VR, n, t, b)
    jump = np.abs(_physical_flux(VR, n) - _physical_flux(VL, n)).max()
    assert np.abs(F - _physical_flux(VL, n)).max() <= 1e-4 * jump
    # mirrored: supersonic toward -x takes the right state
    VL, VR = _state(1.0, -850.0, 0, 0, 300.0), _state(1.0, -900.0, 0, 0, 300.0)
    F2 = O.roe_flux(VL, VR, n, t, b)
    jump = np.abs(_physical_flux(VR, n) - _physical_flux(VL, n)).max()
    assert np.abs(F2 - _physical_flux(VR, n)).max() <= 1e-12 * jump


def test_roe_rotation_invariance():
    rng = np.random.default_rng(2)
    for _ in range(20):
        VL = _state(rng.uniform(0.2, 2), *rng.uniform(-200, 200, 3), rng.uniform(200, 400))
        VR = _state(rng.uniform(0.2, 2), *rng.uniform(-200, 200, 3), rng.uniform(200, 400))
        n, t, b = _frame(rng.normal(size=3))
        F = O.roe_flux(VL, VR, n, t, b)
        q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        rot = lambda V: np.concatenate([[V[0]], q @ V[1:4], [V[4]]])
        nq, tq, bq = q @ n, q @ t, q @ b
        Fq = O.roe_flux(rot(VL), rot(VR), nq, tq, bq)
        assert np.allclose(Fq[[0, 4]], F[[0, 4]], rtol=1e-10, atol=1e-8)
        assert np.allclose(Fq[1:4], q @ F[1:4], rtol=1e-10, atol=1e-6)


def test_roe_conservation_antisymmetry():
    # swapping sides and flipping the normal negates the flux
    VL = _state(1.0, 50.0, 20.0, -10.0, 300.0)
    VR = _state(0.4, -30.0, 5.0, 8.0, 260.0)
    n, t, b = _frame([0.2, 0.5, -0.1])
    F = O.roe_flux(VL, VR, n, t, b)
    G = O.roe_flux(VR, VL, -n, -t, b)
    assert np.allclose(F, -G, rtol=1e-11, atol=1e-9)


def test_viscous_couette_stress_and_heat_flux():
    # u = u(y) only, dT/dx only: tau_xy = mu du/dy, heat flux k dT/dx
    T = 300.0
    mu = 1.458e-6 * T * math.sqrt(T) / (T + 110.4)
    k = mu * 1006.0 / 0.71
    g = np.zeros((5, 3))
    g[1, 1] = 2.0e3   # du/dy
    g[4, 0] = 50.0    # dT/dx
    V = _state(1.0, 10.0, 0.0, 0.0, T)
    a = np.array([0.0, 0.7, 0.0])   # face with normal +y
    F = O.viscous_flux(g, V, a)
    assert F[0] == 0.0
    assert math.isclose(F[1], mu * 2.0e3 * 0.7, rel_tol=1e-14)
    assert F[2] == 0.0 and F[3] == 0.0
    assert math.isclose(F[4], mu * 2.0e3 * 0.7 * 10.0, rel_tol=1e-14)
    a = np.array([0.4, 0.0, 0.0])
    F = O.viscous_flux(g, V, a)
    assert math.isclose(F[2], mu * 2.0e3 * 0.4, rel_tol=1e-14)
    assert math.isclose(F[4], k * 50.0 * 0.4, rel_tol=1e-14)
    # dilatation: du/dx = 1 -> tau_xx = 2 mu (1 - 1/3)
    g = np.zeros((5, 3))
    g[1, 0] = 1.0
    F = O.viscous_flux(g, _state(1, 0, 0, 0, T), np.array([1.0, 0, 0]))
    assert math.isclose(F[1], 2 * mu * (2.0 / 3.0), rel_tol=1e-14)
    # Prandtl is honoured
    g = np.zeros((5, 3))
    g[4, 2] = 3.0
    F = O.viscous_flux(g, V, np.array([0, 0, 1.0]), prandtl=0.72)
    assert math.isclose(F[4], mu * 1006.0 / 0.72 * 3.0, rel_tol=1e-14)


def test_venkat_closed_form():
    # du > 0: phi = ((dmax^2 + e)du + 2 du^2 dmax) / (du (dmax^2 + 2du^2 + dmax du + e))
    for dmax, du, e in [(1.0, 0.5, 0.0), (2.0, 3.0, 0.1), (0.1, 1e-3, 1e-6)]:
        exp = ((dmax ** 2 + e) * du + 2 * du * du * dmax) / (du * (dmax ** 2 + 2 * du ** 2 + dmax * du + e))
        assert math.isclose(O.venkat(dmax, -7.0, du, e), exp, rel_tol=1e-15)
    # du < 0 uses dumin; |du| tiny -> 1
    assert math.isclose(O.venkat(9.0, -1.0, -0.5, 0.0), ((1 + 0) * -0.5 + 2 * 0.25 * -1) / (-0.5 * (1 + 0.5 + 0.5)), rel_tol=1e-15)
    assert O.venkat(1.0, -1.0, 1e-41, 0.0) == 1.0
    # dmax = du, e = 0: phi = (du^3 + 2du^3)/(du(4 du^2)) = 3/4
    assert math.isclose(O.venkat(2.0, -1.0, 2.0, 0.0), 0.75, rel_tol=1e-15)


def test_vanalbada_closed_form():
    # y = dmax/du; phi = (4y - y^2)/(y^2 - 4y + 8) for y < 2, clipped to [0, 1]
    assert O.vanalbada(1.0, -1.0, 1.0) == 0.6     # y = 1: 3/5
    assert O.vanalbada(5.0, -1.0, 1.0) == 1.0     # y >= 2
    assert O.vanalbada(1.0, -0.5, -1.0) == (4 * 0.5 - 0.25) / (0.25 - 2 + 8)
    assert O.vanalbada(1.0, -1.0, 1e-20) == 1.0   # |du| < DBL_EPSILON -> y = 2
    assert O.vanalbada(-1.0, -1.0, 1.0) == 0.0    # y = -1 -> negative, clipped


def test_hex_volume_and_centroid():
    # affine hex: 2x2x2 Gauss is exact; unit cube scaled and sheared
    ex = np.array([0, 1, 1, 0, 0, 1, 1, 0], float) * 2.0
    ey = np.array([0, 0, 1, 1, 0, 0, 1, 1], float) * 3.0
    ez = np.array([0, 0, 0, 0, 1, 1, 1, 1], float) * 0.5
    vol, c = O.hex_volume_centroid(ex + 0.3 * ez, ey, ez)
    assert math.isclose(vol, 3.0, rel_tol=1e-12)
    assert np.allclose(c, [1.0 + 0.3 * 0.25, 1.5, 0.25])


def test_uniform_flow_preserved_and_gradient_exact_for_linear_field():
    # uniform state: residual is round-off (the box's closed surfaces sum to ~0)
    kw = dict(problem_type=1, lx=1.0, ly=0.5, lz=0.25, nx=6, ny=5, nz=4, dt=1e-6,
              second_order_space=1, viscous=1, bc_kind=(1, 1, 1, 1, 1, 1))
    m = O.OracleMesh(O.make_options(**kw))
    s = O.OracleSolver(m)
    s.init()
    U0 = s.state()
    s.step(2)
    U = s.state()
    assert np.abs(U - U0).max() <= 1e-9 * np.abs(U0).max()
    # Green-Gauss of a linear density field is exact on interior cells
    x = m.cell_xyz
    Ul = U0.copy()
    Ul[:, 0] = 1.0 + 0.3 * x[:, 0] - 0.2 * x[:, 1] + 0.1 * x[:, 2]
    Ul[:, 4] = U0[:, 4] * Ul[:, 0] / U0[:, 0]
    Ul[:, 1:4] = U0[:, 1:4] * (Ul[:, 0] / U0[:, 0])[:, None]
    # dt = 1e-300 keeps every stage state equal to Ul, so the gradients left
    # by the last stage are those of Ul
    s3 = O.OracleSolver(m, O.make_options(**dict(kw, dt=1e-300)))
    s3.set_state(Ul)
    s3.step(1)
    g, _ = s3.gradients()
    nx, ny, nz = kw["nx"], kw["ny"], kw["nz"]
    idx = np.arange(m.ncells).reshape(nx, ny, nz)[1:-1, 1:-1, 1:-1].ravel()
    assert np.allclose(g[idx, 0, :], [0.3, -0.2, 0.1], rtol=1e-10, atol=1e-12)


@pytest.mark.parametrize("viscous", [0, 1])
def test_cfl_time_step_closed_form(viscous):
    """CFL extension on a uniform state over a uniform box: every cell has
    two faces of area dy*dz, dx*dz, dx*dy on each axis (boundary faces count),
    so dt = cfl * vol / (sum_f(|u.n_f| + c|n_f|) + 4 kappa sum_f |n_f|^2 / vol)
    with c = sqrt(gamma R T) and kappa = max(4/3, gamma/Pr) mu(T) / rho
    (zero when inviscid) is the same for every cell."""
    rho, u, v, w, T = 0.9, 120.0, -35.0, 20.0, 280.0
    Rg, g = 287.05, 1.4
    E = rho * (Rg / (g - 1.0)) * T + 0.5 * rho * (u * u + v * v + w * w)
    nx, ny, nz, lx, ly, lz = 7, 5, 4, 0.7, 0.25, 0.4
    kw = dict(problem_type=1, lx=lx, ly=ly, lz=lz, nx=nx, ny=ny, nz=nz, ntimesteps=1, dt=0.0,
              second_order_space=0, viscous=viscous, bc_kind=(1, 1, 1, 1, 1, 1), use_inflow=1,
              inflow=(rho, rho * u, rho * v, rho * w, E), prandtl=0.72, cfl=0.5)
    s = O.OracleSolver(O.OracleMesh(O.make_options(**kw)))
    s.init()
    s.step(1)
    dx, dy, dz = lx / nx, ly / ny, lz / nz
    ax, ay, az = dy * dz, dx * dz, dx * dy
    c = math.sqrt(g * Rg * T)
    lc = 2 * (abs(u) * ax + abs(v) * ay + abs(w) * az) + 2 * c * (ax + ay + az)
    lv = 2 * (ax * ax + ay * ay + az * az)
    vol = dx * dy * dz
    lam = lc
    if viscous:
        mu = 1.458e-6 * T * math.sqrt(T) / (T + 110.4)
        lam += 4.0 * max(4.0 / 3.0, g / 0.72) * mu / rho * lv / vol
    expect = 0.5 * vol / lam
    assert s.time == pytest.approx(expect, rel=1e-12)


@pytest.mark.parametrize("viscous", [0, 1])
def test_cfl_time_step_bracket_on_a_skewed_mesh(viscous):
    """On a ramp mesh (faces not normal to an axis) the CFL extension's
    convective term sum_d |u_d| S_d, S_d = sum_f |n_f,d|, lies between the
    spectral radius sum_f |u.n_f| and |u| sum_f |n_f| face by face
    (|u.n| <= sum_d |u_d n_d| <= |u||n|), so the step is never larger than the
    exact spectral-radius bound and never smaller than the isotropic one; on a
    box the three coincide (closed form above)."""
    rho, u, v, w, T = 0.9, 150.0, 60.0, -25.0, 280.0
    Rg, g = 287.05, 1.4
    E = rho * (Rg / (g - 1.0)) * T + 0.5 * rho * (u * u + v * v + w * w)
    kw = dict(problem_type=1, lx=1.0, ly=0.5, lz=0.3, nx=9, ny=6, nz=4, angle=20.0, ntimesteps=1, dt=0.0,
              second_order_space=0, viscous=viscous, bc_kind=(1, 1, 1, 1, 1, 1), use_inflow=1,
              inflow=(rho, rho * u, rho * v, rho * w, E), prandtl=0.72, cfl=0.5)
    m = O.OracleMesh(O.make_options(**kw))
    s = O.OracleSolver(m)
    s.init()
    s.step(1)
    nc = m.ncells
    vel = np.array([u, v, w])
    c = math.sqrt(g * Rg * T)
    exact, iso, lv = np.zeros(nc), np.zeros(nc), np.zeros(nc)
    faces = [(m.internal["conn"][:, 0], m.internal["a"]), (m.internal["conn"][:, 1], m.internal["a"])]
    faces += [(G["conn"][:, 0], G["a"]) for G in m.groups]
    skewed = 0
    for cells, a in faces:
        an = np.linalg.norm(a, axis=1)
        skewed += int(np.sum(np.count_nonzero(np.abs(a) > 1e-12 * an[:, None], axis=1) > 1))
        np.add.at(exact, cells, np.abs(a @ vel) + c * an)
        np.add.at(iso, cells, (np.linalg.norm(vel) + c) * an)
        np.add.at(lv, cells, an * an)
    assert skewed > 0
    vol = m.cell_vol
    visc = 0.0
    if viscous:
        mu = 1.458e-6 * T * math.sqrt(T) / (T + 110.4)
        visc = 4.0 * max(4.0 / 3.0, g / 0.72) * mu / rho * lv / vol
    dt_exact = 0.5 * np.min(vol / (exact + visc))
    dt_iso = 0.5 * np.min(vol / (iso + visc))
    assert dt_iso * (1 - 1e-12) <= s.time <= dt_exact * (1 + 1e-12)
    assert s.time < dt_exact * (1 - 1e-6)      # the ramp's faces make the bound strict
