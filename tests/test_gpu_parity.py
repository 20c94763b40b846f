"""GPU parity: the HIP path against the oracle through the C ABI.

Bar (stated per test): the FP64 path is built without FMA contraction and
sums every cell's face contributions in the reference's serial order, so it
is expected to equal the oracle bit for bit.  The tests assert that, and also
the written tolerances: relative L2 <= 1e-12 per RK4 step and <= 1e-10 after
N steps, measured on the whole state and per conserved variable with the
floor max(||U_ref,var||, 1e-8 ||U_ref||) (round-off-level transverse momenta
would otherwise make per-variable ratios meaningless).
"""
import os
import sys

import numpy as np
import pytest

import miniaero_amd as M
from conftest import GOLD_CASES, GOLDEN, gold_rows, read_inp, rows_text
from oracle import oracle as O

pytestmark = pytest.mark.gpu

EXT = ("bc_kind", "y_stretch", "limiter", "ic_kind", "ic", "use_inflow", "inflow", "prandtl",
       "cfl", "perturb", "seed")


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def rel_l2(U, ref):
    """(whole-state relative L2, worst per-variable floored relative L2)."""
    tot = np.linalg.norm(ref)
    whole = np.linalg.norm(U - ref) / tot
    per = [np.linalg.norm(U[:, k] - ref[:, k]) / max(np.linalg.norm(ref[:, k]), 1e-8 * tot)
           for k in range(5)]
    return whole, max(per)


class Pair:
    """The same case on the GPU and in the oracle."""

    def __init__(self, kw):
        self.kw = kw
        _log(f"[gpu-parity] case {kw}")
        ext = {k: v for k, v in kw.items() if k in EXT}
        self.opts = M.Options(**{k: v for k, v in kw.items()
                                 if k in M.Options.__dataclass_fields__})
        self.md = M.Parallel3DMesh(kw["nx"], kw["ny"], kw["nz"], kw["lx"], kw["ly"], kw["lz"],
                                   kw["problem_type"], kw.get("angle", 0.0),
                                   options=M.Options(**ext)).fillMeshData()
        self.gpu = M.TimeSolverExplicitRK4(self.md, self.opts)
        self.gpu.initialize()
        okw = {k: v for k, v in kw.items() if k not in ("output_results", "output_frequency")}
        self.om = O.OracleMesh(O.make_options(**okw))
        self.cpu = O.OracleSolver(self.om)
        self.cpu.init()

    def step(self, n):
        self.gpu.step(n)
        self.cpu.step(n)
        return self.gpu.state(), self.cpu.state()

    def close(self):
        self.gpu.close()


def check_steps(kw, per_step, total, bitwise=True):
    p = Pair(kw)
    try:
        U, R = p.gpu.state(), p.cpu.state()
        assert np.array_equal(U, R), "initial condition differs"
        for i in range(per_step):
            U, R = p.step(1)
            whole, per = rel_l2(U, R)
            _log(f"  step {i + 1}: rel L2 {whole:.3e}, worst variable {per:.3e}, "
                 f"bitwise={np.array_equal(U, R)}")
            assert whole <= 1e-12 and per <= 1e-12
            if bitwise:
                assert np.array_equal(U, R), f"step {i + 1} not bit-identical"
        rest = total - per_step
        if rest > 0:
            U, R = p.step(rest)
        whole, per = rel_l2(U, R)
        _log(f"  after {total} steps: rel L2 {whole:.3e}, worst variable {per:.3e}, "
             f"bitwise={np.array_equal(U, R)}")
        assert whole <= 1e-10 and per <= 1e-10
        if bitwise:
            assert np.array_equal(U, R)
        t_gpu, _ = p.gpu.time()
        assert abs(t_gpu - p.cpu.time) <= 1e-12 * max(abs(p.cpu.time), 1e-300)
        return p, U
    except Exception:
        p.close()
        raise


def _case(name):
    kw = read_inp(os.path.join(GOLDEN, name, "miniaero.inp"))
    return kw


@pytest.mark.parametrize("case,ranks,rtol,floor", GOLD_CASES)
def test_reference_cases_match_oracle_and_gold(case, ranks, rtol, floor):
    kw = _case(case)
    p, U = check_steps(kw, per_step=3, total=kw["ntimesteps"])
    try:
        mine = np.hstack([p.md.cell_coordinates, U])
        gold = gold_rows(case, ranks, p.md.cell_coordinates)
        same, msg = M.numeric_text_diff(rows_text(mine), rows_text(gold), rtol, floor)
        assert same, msg[:2000]
    finally:
        p.close()


CONFIG_CASES = {
    # Sod-type IC, all sides Extrapolate, Van Albada, second order
    "sod_vanalbada": dict(problem_type=0, lx=1.0, ly=0.25, lz=0.25, nx=32, ny=8, nz=8, dt=1e-4,
                          ntimesteps=20, second_order_space=1, viscous=0, limiter=1, ic_kind=1,
                          ic=(1.0, 1.0, 0.125, 0.1, 0.5), bc_kind=(1, 1, 1, 1, 1, 1)),
    # stretched viscous plate with freestream inflow, Tangent top, CFL dt, Pr 0.72
    "plate_cfl_stretch": dict(problem_type=1, lx=1.0, ly=0.05, lz=0.025, nx=24, ny=16, nz=4,
                              dt=0.0, ntimesteps=10, second_order_space=1, viscous=1,
                              y_stretch=1.05, bc_kind=(0, 2, 0, 0, 0, 0), use_inflow=1,
                              inflow=(0.01772, 0.01772 * 104.2, 0.0, 0.0,
                                      0.01772 * 287.05 / 0.4 * 300.0 + 0.5 * 0.01772 * 104.2 ** 2),
                              prandtl=0.72, cfl=0.5),
    # perturbed freestream start, first-order viscous
    "perturbed_first_order_viscous": dict(problem_type=1, lx=1.0, ly=0.05, lz=0.05, nx=12, ny=10,
                                          nz=6, dt=1e-7, ntimesteps=10, second_order_space=0,
                                          viscous=1, ic_kind=2, perturb=0.01, seed=3),
    # the ramp geometry, first order
    "ramp_first_order": dict(problem_type=2, lx=2.0, ly=2.0, lz=1.0, angle=30.0, nx=16, ny=8,
                             nz=3, dt=1e-5, ntimesteps=30, second_order_space=0, viscous=0),
}


@pytest.mark.parametrize("name", sorted(CONFIG_CASES))
def test_config_mode_cases(name):
    p, _ = check_steps(CONFIG_CASES[name], per_step=2, total=CONFIG_CASES[name]["ntimesteps"])
    p.close()


@pytest.mark.parametrize("shape", [(1, 1, 1), (1, 3, 1), (2, 1, 5), (3, 2, 1)])
def test_tiny_and_degenerate_meshes(shape):
    nx, ny, nz = shape
    for so, visc, pt in ((0, 0, 0), (1, 1, 1), (1, 0, 2)):
        kw = dict(problem_type=pt, lx=1.0, ly=0.5, lz=0.5, angle=10.0 if pt == 2 else 0.0,
                  nx=nx, ny=ny, nz=nz, dt=1e-7, ntimesteps=3, second_order_space=so, viscous=visc)
        p, _ = check_steps(kw, per_step=1, total=3)
        p.close()


def test_gradients_limiters_and_residual_match():
    kw = _case("Ramp_Serial")
    kw["ntimesteps"] = 5
    p, _ = check_steps(kw, per_step=0, total=5)
    try:
        g, lim = p.gpu.gradients()
        og, olim = p.cpu.gradients()
        assert np.array_equal(g, og)
        assert np.array_equal(lim, olim)
        assert np.array_equal(p.gpu.residual(), p.cpu.residual())
    finally:
        p.close()


def test_face_value_gradients_on_demand():
    """Inviscid second order keeps no gradients or limiters while stepping;
    the getter computes the last stage's.  They equal the oracle's bit for
    bit after graph replays, on a second call, after set_state (which keeps
    them, as the oracle does) and after a further plain step."""
    kw = dict(CONFIG_CASES["sod_vanalbada"], ntimesteps=5)
    p = Pair(kw)
    try:
        def same_gradients(what):
            g, lim = p.gpu.gradients()
            og, olim = p.cpu.gradients()
            _log(f"  {what}: gradients bitwise={np.array_equal(g, og)}, limiters bitwise={np.array_equal(lim, olim)}")
            assert np.array_equal(g, og) and np.array_equal(lim, olim), what
            return g, lim
        p.gpu.set_graph(2)
        G, R = p.step(4)
        assert np.array_equal(G, R)
        g1, l1 = same_gradients("after 4 steps as 2 graph replays")
        g2, l2 = same_gradients("second call")
        assert np.array_equal(g1, g2) and np.array_equal(l1, l2)
        U = p.gpu.state()
        p.gpu.set_state(U)
        p.cpu.set_state(U)
        same_gradients("after set_state")
        p.gpu.set_graph(0)
        G, R = p.step(1)
        assert np.array_equal(G, R)
        same_gradients("after one more plain step")
    finally:
        p.close()


@pytest.mark.parametrize("name", ["plate_cfl_stretch", "sod_vanalbada", "perturbed_first_order_viscous"])
def test_extreme_exponent_gradients_match(name):
    """Momenta near the bottom of the FP64 range (1e-306, and subnormal) in
    every 7th cell, beside exact zeros: gradient quotients, limiter ratios
    and fluxes then meet subnormal operands and results, which the GPU keeps
    (no flush to zero) as the CPU does.  State, gradients and limiters stay
    bit-identical."""
    kw = dict(CONFIG_CASES[name], ntimesteps=2)
    p = Pair(kw)
    try:
        U = p.gpu.state().copy()
        idx = np.arange(0, U.shape[0], 7)
        U[idx, 2] = U[idx, 0] * 3e-306
        U[idx[::2], 3] = U[idx[::2], 0] * 2.5e-310
        assert np.count_nonzero((U[:, 3] != 0) & (np.abs(U[:, 3]) < 2.2250738585072014e-308)) > 0
        p.gpu.set_state(U)
        p.cpu.set_state(U)
        for i in range(2):
            G, R = p.step(1)
            _log(f"  step {i + 1}: bitwise={np.array_equal(G, R)}")
            assert np.array_equal(G, R), f"step {i + 1} not bit-identical"
        g, lim = p.gpu.gradients()
        og, olim = p.cpu.gradients()
        assert np.array_equal(g, og)
        if kw["second_order_space"]:
            assert np.array_equal(lim, olim)
    finally:
        p.close()


def test_deterministic_and_finite():
    kw = dict(_case("FlatPlate_Parallel"), ntimesteps=20)
    a = Pair(kw)
    b = Pair(kw)
    a.gpu.step(20)
    b.gpu.step(20)
    assert np.array_equal(a.gpu.state(), b.gpu.state())
    assert a.gpu.nonfinite() == 0
    a.close()
    b.close()


def test_solve_writes_results_file(tmp_path):
    kw = _case("3D_Sod_Serial")
    md = M.Parallel3DMesh(kw["nx"], kw["ny"], kw["nz"], kw["lx"], kw["ly"], kw["lz"],
                          kw["problem_type"], kw["angle"]).fillMeshData()
    s = M.TimeSolverExplicitRK4(md, M.Options(**kw))
    out = tmp_path / "results.0"
    import io
    s.Solve(out=io.StringIO(), results_path=str(out))
    s.close()
    same, msg = M.numeric_file_diff(str(out), os.path.join(GOLDEN, "3D_Sod_Serial", "results.gold"))
    assert same, msg[:1000]


@pytest.mark.slow
def test_large_flat_plate_bitwise():
    """Two RK4 steps of a 128x64x32 (262k cell) stretched viscous plate,
    bit-compared with the oracle on every cell."""
    kw = dict(CONFIG_CASES["plate_cfl_stretch"], nx=128, ny=64, nz=32, ntimesteps=2)
    p, _ = check_steps(kw, per_step=1, total=2)
    p.close()


FULL = dict(CONFIG_CASES["plate_cfl_stretch"], nx=256, ny=128, nz=64)


def _gpu_only(kw):
    ext = {k: v for k, v in kw.items() if k in EXT}
    md = M.Parallel3DMesh(kw["nx"], kw["ny"], kw["nz"], kw["lx"], kw["ly"], kw["lz"],
                          kw["problem_type"], kw.get("angle", 0.0), options=M.Options(**ext)).fillMeshData()
    opts = M.Options(**{k: v for k, v in kw.items() if k in M.Options.__dataclass_fields__})
    return md, M.TimeSolverExplicitRK4(md, opts)


@pytest.mark.slow
def test_full_size_deterministic():
    """256x128x64 viscous plate: two independent solvers give identical bits."""
    _log("[gpu-parity] full size determinism 256x128x64")
    md, a = _gpu_only(FULL)
    b = M.TimeSolverExplicitRK4(md, a.options)
    for s in (a, b):
        s.initialize()
        s.step(3)
    Ua, Ub = a.state(), b.state()
    assert np.isfinite(Ua).all()
    assert np.array_equal(Ua, Ub)
    a.close()
    b.close()


@pytest.mark.slow
def test_full_size_freestream_preserved():
    """Uniform flow with Extrapolate on every side stays uniform to round-off
    on the stretched 2.1M-cell mesh."""
    _log("[gpu-parity] full size free-stream preservation")
    kw = dict(FULL, bc_kind=(1, 1, 1, 1, 1, 1), cfl=0.5)
    md, s = _gpu_only(kw)
    s.initialize()
    U0 = s.state()
    s.step(3)
    U = s.state()
    scale = np.abs(U0).max(axis=0)
    assert (np.abs(U - U0).max(axis=0) <= 1e-11 * scale).all()
    s.close()


@pytest.mark.slow
def test_full_size_closed_box_conserves_mass_and_energy():
    """All walls Tangent, perturbed start: sum(vol * rho) and sum(vol * rhoE)
    change only by round-off on the 2.1M-cell mesh."""
    _log("[gpu-parity] full size conservation")
    kw = dict(FULL, bc_kind=(2, 2, 2, 2, 2, 2), viscous=0, ic_kind=2, perturb=0.01, seed=11,
              cfl=0.5)
    md, s = _gpu_only(kw)
    s.initialize()
    vol = md.cell_volumes
    U0 = s.state()
    s.step(3)
    U = s.state()
    for k in (0, 4):
        m0, m1 = np.sum(vol * U0[:, k]), np.sum(vol * U[:, k])
        assert abs(m1 - m0) <= 1e-12 * abs(m0), (k, m0, m1)
    assert np.abs(U - U0).max() > 0   # the flow did evolve
    s.close()


def test_cpp_driver_reproduces_reference_run(tmp_path):
    """The C++ driver (the reference's Main.C sequence over the library) run in
    a case directory: results.0 passes the reference runner's comparison with
    the gold file, and its text equals the oracle's values printed the same way."""
    import shutil
    import subprocess

    from conftest import ROOT
    exe = os.path.join(ROOT, "miniaero_amd", "bin", "miniaero")
    for case in ("3D_Sod_Serial", "FlatPlate_Serial"):
        d = tmp_path / case
        d.mkdir()
        shutil.copy(os.path.join(GOLDEN, case, "miniaero.inp"), d / "miniaero.inp")
        _log(f"[gpu-parity] driver on {case}")
        r = subprocess.run([exe], cwd=d, capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stderr
        assert "Device Run time" in r.stdout
        text = (d / "results.0").read_text()
        gold = open(os.path.join(GOLDEN, case, "results.gold")).read()
        rtol, floor = (1e-3, 1e-6) if case.startswith("3D") else (1e-2, 1e-3)
        same, msg = M.numeric_text_diff(text, gold, rtol, floor)
        assert same, msg[:1000]
        kw = _case(case)
        m = O.OracleMesh(O.make_options(**{k: v for k, v in kw.items()
                                           if k not in ("output_results", "output_frequency")}))
        s = O.OracleSolver(m)
        s.init()
        s.step(kw["ntimesteps"])
        assert text == rows_text(np.hstack([m.cell_xyz, s.state()]))


def precision_errors(U, ref):
    """Error of a state against the FP64 reference: whole-state relative L2,
    and per conserved variable -- rho and rho E against their own norms, each
    momentum component against the norm of the momentum vector (transverse
    momenta are near zero in these flows, so their own norm is no scale).
    Also the raw per-component ratios, for the record."""
    tot = np.linalg.norm(ref)
    mom = np.linalg.norm(ref[:, 1:4])
    scale = [np.linalg.norm(ref[:, 0]), mom, mom, mom, np.linalg.norm(ref[:, 4])]
    per = [np.linalg.norm(U[:, k] - ref[:, k]) / scale[k] for k in range(5)]
    raw = [np.linalg.norm(U[:, k] - ref[:, k]) / max(np.linalg.norm(ref[:, k]), 1e-300) for k in range(5)]
    return np.linalg.norm(U - ref) / tot, per, raw


@pytest.mark.parametrize("name", ["plate_cfl_stretch", "sod_vanalbada"])
def test_fp32_arithmetic_form_error_vs_fp64(name):
    """Precision form: flux/gradient/limiter arithmetic in FP32, state, RK4
    accumulation and residual sums in FP64.  Bar: relative L2 error against
    the FP64 run <= 1e-5, whole state and per conserved variable (momentum
    components scaled by the momentum norm), after the case's steps; the FP32
    run must also differ from the FP64 one (the switch is real)."""
    kw = CONFIG_CASES[name]
    ext = {k: v for k, v in kw.items() if k in EXT}
    md = M.Parallel3DMesh(kw["nx"], kw["ny"], kw["nz"], kw["lx"], kw["ly"], kw["lz"],
                          kw["problem_type"], kw.get("angle", 0.0), options=M.Options(**ext)).fillMeshData()
    base = M.Options(**{k: v for k, v in kw.items() if k in M.Options.__dataclass_fields__})
    runs = {}
    for prec in (0, 1):
        s = M.TimeSolverExplicitRK4(md, base.replace(precision=prec))
        s.initialize()
        s.step(kw["ntimesteps"])
        runs[prec] = s.state()
        s.close()
    whole, per, raw = precision_errors(runs[1], runs[0])
    _log(f"[gpu-parity] fp32 form {name}: whole {whole:.2e}, per variable {['%.2e' % x for x in per]}, "
         f"raw per component {['%.2e' % x for x in raw]}")
    assert whole <= 1e-5 and max(per) <= 1e-5
    assert not np.array_equal(runs[1], runs[0])


def test_stored_face_frames_match_derived_frames():
    """A mesh whose tangents and binormals are scaled by 2 has the same unit
    vectors (scaling by 2 is exact) but no longer follows the reference's
    construction, so the solver reads stored frames instead of rebuilding
    them: both paths must give the same bits."""
    kw = _case("Ramp_Serial")
    md = M.Parallel3DMesh(kw["nx"], kw["ny"], kw["nz"], kw["lx"], kw["ly"], kw["lz"],
                          kw["problem_type"], kw["angle"]).fillMeshData()
    def scaled(F):
        return M.Faces(F.coordinates, F.face_normal, 2.0 * F.face_tangent, 2.0 * F.face_binormal,
                       F.face_cell_conn, F.cell_flux_index, F.permute_vector)
    md2 = M.MeshData(md.num_owned_cells, md.cell_coordinates, md.cell_volumes,
                     scaled(md.internal_faces), [(n, scaled(F)) for n, F in md.boundary_faces])
    assert md.plan()["derived_frames"] and not md2.plan()["derived_frames"]
    out = []
    for mesh in (md, md2):
        s = M.TimeSolverExplicitRK4(mesh, M.Options(**kw))
        s.initialize()
        s.step(20)
        out.append(s.state())
        s.close()
    assert np.array_equal(out[0], out[1])


def test_residual_norms_reduce_the_residual():
    """ma_solver_residual_norms against numpy norms of the oracle's residual
    (bit-identical to the GPU's, so only the summation order differs:
    relative 1e-13), repeatable bit for bit, and zero up to round-off for a
    uniform flow with Extrapolate on every side (the closed form)."""
    kw = dict(_case("FlatPlate_Serial"), ntimesteps=5)
    p, _ = check_steps(kw, per_step=0, total=5)
    try:
        n = p.gpu.residual_norms()
        ref = np.linalg.norm(p.cpu.residual(), axis=0)
        _log(f"[gpu-parity] residual norms {n} vs oracle {ref}")
        assert np.allclose(n, ref, rtol=1e-13, atol=0.0)
        assert np.array_equal(n, p.gpu.residual_norms())
    finally:
        p.close()
    kw = dict(CONFIG_CASES["plate_cfl_stretch"], bc_kind=(1, 1, 1, 1, 1, 1), viscous=0)
    q = Pair(kw)
    try:
        U0 = q.gpu.state()
        q.gpu.step(2)
        n = q.gpu.residual_norms()
        scale = np.linalg.norm(U0, axis=0)
        _log(f"[gpu-parity] uniform-flow residual norms {n}, state norms {scale}")
        assert (n <= 1e-13 * scale).all()
    finally:
        q.close()
