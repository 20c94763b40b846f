"""Host-side checks of the product that need no GPU: the C ABI library loads
and exports every entry point the header declares; the closed-form mesh
generator is bit-identical to the oracle's mesh (itself pinned to the
reference's code); options reading, results formatting and the comparator."""
import os
import re

import numpy as np
import pytest

import miniaero_amd as M
from miniaero_amd import _lib
from conftest import GOLDEN, ROOT
from oracle import oracle as O


def test_library_exports_every_declared_symbol():
    L = _lib.load()
    header = open(os.path.join(ROOT, "include", "miniaero.h")).read()
    declared = set(re.findall(r"^(?:const char\*|ma_status|void|int|double)\s+(ma_[a-z_]+)\s*\(",
                              header, re.M))
    assert declared == set(_lib.EXPORTS)
    for name in declared:
        assert hasattr(L, name), name


def test_no_gpu_is_a_clean_error_here():
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is present")
    md = M.Parallel3DMesh(2, 2, 2, 1, 1, 1, 0).fillMeshData()
    with pytest.raises(M.MiniAeroError) as e:
        M.TimeSolverExplicitRK4(md, M.Options(dt=1e-6))
    assert e.value.status == _lib.MA_ERR_DEVICE


MESH_CASES = [
    dict(problem_type=0, lx=0.3048, ly=1.0, lz=1.0, angle=0.0, nx=128, ny=4, nz=4),
    dict(problem_type=2, lx=2.0, ly=2.0, lz=1.0, angle=30.0, nx=64, ny=32, nz=2),
    dict(problem_type=1, lx=2.0, ly=0.002, lz=1.0, angle=0.0, nx=32, ny=64, nz=2),
    dict(problem_type=1, lx=1.0, ly=0.05, lz=0.025, angle=0.0, nx=8, ny=6, nz=5, y_stretch=1.05,
         bc_kind=(0, 2, 0, 0, 0, 0)),
    dict(problem_type=0, lx=1.0, ly=1.0, lz=1.0, angle=0.0, nx=1, ny=3, nz=1),
    dict(problem_type=2, lx=1.0, ly=1.0, lz=1.0, angle=10.0, nx=1, ny=1, nz=1),
]


def product_mesh(kw):
    ext = {k: v for k, v in kw.items() if k in ("y_stretch", "bc_kind")}
    return M.Parallel3DMesh(kw["nx"], kw["ny"], kw["nz"], kw["lx"], kw["ly"], kw["lz"],
                            kw["problem_type"], kw["angle"], options=M.Options(**ext)).fillMeshData()


@pytest.mark.parametrize("kw", MESH_CASES)
def test_mesh_generator_bit_exact(kw):
    om = O.OracleMesh(O.make_options(**kw))
    md = product_mesh(kw)
    assert md.num_owned_cells == om.ncells
    assert np.array_equal(md.cell_coordinates, om.cell_xyz)
    assert np.array_equal(md.cell_volumes, om.cell_vol)
    pairs = [("xf", "coordinates"), ("a", "face_normal"), ("t", "face_tangent"),
             ("b", "face_binormal"), ("conn", "face_cell_conn"), ("fidx", "cell_flux_index"),
             ("perm", "permute_vector")]
    for a, b in pairs:
        assert np.array_equal(om.internal[a], getattr(md.internal_faces, b)), a
    assert len(md.boundary_faces) == 6
    for g, (name, F) in enumerate(md.boundary_faces):
        assert _lib.BC_KINDS[name] == om.groups[g]["kind"]
        for a, b in pairs:
            assert np.array_equal(om.groups[g][a], getattr(F, b)), (g, a)


def test_options_read(tmp_path):
    o = M.Options().read_options_file(os.path.join(GOLDEN, "FlatPlate_Serial", "miniaero.inp"))
    assert (o.problem_type, o.nx, o.ny, o.nz, o.ntimesteps) == (1, 16, 32, 2, 400)
    assert (o.lx, o.ly, o.lz, o.angle, o.dt) == (2.0, 0.002, 1.0, 0.0, 3e-8)
    assert (o.output_results, o.output_frequency, o.second_order_space, o.viscous) == (1, 100, 1, 1)
    with pytest.raises(M.MiniAeroError):
        M.Options().read_options_file(str(tmp_path / "missing.inp"))
    short = tmp_path / "short.inp"
    short.write_text("0\n1 1 1 0\n4 4\n")
    with pytest.raises(M.MiniAeroError):
        M.Options().read_options_file(str(short))


def test_results_format_matches_gold_text():
    gold = open(os.path.join(GOLDEN, "3D_Sod_Serial", "results.gold")).read()
    rows = np.loadtxt(os.path.join(GOLDEN, "3D_Sod_Serial", "results.gold"))
    text = M.results.format_rows(rows[:, :3], rows[:, 3:])
    assert text == gold


def test_numeric_text_diff_semantics():
    a = "1\t2.5\t1e-07\t\n3\t4\t\n"
    assert M.numeric_text_diff(a, a)[0]
    assert M.numeric_text_diff(a, "1\t2.5001\t1e-07\t\n3\t4\t\n")[0]           # 4e-5 < 1e-3
    assert not M.numeric_text_diff(a, "1\t2.6\t1e-07\t\n3\t4\t\n")[0]
    assert M.numeric_text_diff(a, "1\t2.5\t9e-07\t\n3\t4\t\n")[0]              # below the floor
    assert not M.numeric_text_diff(a, "1\t2.5\t9e-07\t\n3\t4\t\n", floor=1e-8)[0]
    assert not M.numeric_text_diff(a, "1\t2.5\t1e-07\t\n")[0]                 # line count
    assert not M.numeric_text_diff(a, "1\t2.5\t\n3\t4\t\n")[0]                # field count


def test_cpp_driver_reports_missing_input(tmp_path):
    import subprocess
    exe = os.path.join(ROOT, "miniaero_amd", "bin", "miniaero")
    assert os.path.exists(exe), "bin/miniaero not built"
    r = subprocess.run([exe], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "miniaero.inp not found" in r.stderr


@pytest.mark.parametrize("shape", [(256 // 8, 128 // 8, 64 // 8), (7, 9, 10), (1, 1, 1), (1, 3, 1)])
def test_layout_plan_finds_lattice_bricks(shape):
    nx, ny, nz = shape
    md = M.Parallel3DMesh(nx, ny, nz, 1.0, 0.5, 0.25, 1, options=M.Options(y_stretch=1.05)).fillMeshData()
    p = md.plan()
    assert p["lattice"] == (nx, ny, nz)
    ceil = lambda a, b: -(-a // b)
    assert p["tiles"] == ceil(nx, 4) * ceil(ny, 8) * ceil(nz, 8)
    nif = md.internal_faces.nfaces_
    assert nif <= p["tile_face_entries"] <= 2 * nif
    assert p["max_tile_faces"] <= 6 * 256


def test_layout_plan_on_ramp_and_relabelled_cells():
    md = M.Parallel3DMesh(16, 8, 4, 2.0, 2.0, 1.0, 2, 30.0).fillMeshData()
    assert md.plan()["lattice"] == (16, 8, 4)
    # relabel the cells at random: the lattice comes from the connectivity
    rng = np.random.default_rng(5)
    nc = md.num_owned_cells
    new = rng.permutation(nc)           # old -> new
    old = np.argsort(new)
    def relabel(F):
        conn = F.face_cell_conn.copy()
        m = conn >= 0
        conn[m] = new[conn[m]]
        return M.Faces(F.coordinates, F.face_normal, F.face_tangent, F.face_binormal, conn)
    md2 = M.MeshData(nc, md.cell_coordinates[old], md.cell_volumes[old], relabel(md.internal_faces),
                     [(n, relabel(F)) for n, F in md.boundary_faces])
    assert md2.plan()["lattice"] == (16, 8, 4)


def test_layout_plan_rejects_bad_meshes():
    md = M.Parallel3DMesh(3, 3, 3, 1.0, 1.0, 1.0, 0).fillMeshData()
    F = md.internal_faces
    bad = F.face_cell_conn.copy()
    bad[5, 1] = md.num_owned_cells
    with pytest.raises(M.MiniAeroError) as e:
        M.MeshData(md.num_owned_cells, md.cell_coordinates, md.cell_volumes,
                   M.Faces(F.coordinates, F.face_normal, F.face_tangent, F.face_binormal, bad),
                   md.boundary_faces).plan()
    assert e.value.status == _lib.MA_ERR_INVALID
    perm = F.permute_vector.copy()
    perm[0] = perm[1]
    with pytest.raises(M.MiniAeroError) as e:
        M.MeshData(md.num_owned_cells, md.cell_coordinates, md.cell_volumes,
                   M.Faces(F.coordinates, F.face_normal, F.face_tangent, F.face_binormal,
                           F.face_cell_conn, permute_vector=perm), md.boundary_faces).plan()
    assert e.value.status == _lib.MA_ERR_INVALID
    # a seventh face on one cell
    extra = md.boundary_faces + [("Extrapolate", md.boundary_faces[0][1])]
    with pytest.raises(M.MiniAeroError) as e:
        M.MeshData(md.num_owned_cells, md.cell_coordinates, md.cell_volumes, F, extra).plan()
    assert e.value.status == _lib.MA_ERR_UNSUPPORTED
    with pytest.raises(M.MiniAeroError) as e:
        M.MeshData(md.num_owned_cells, md.cell_coordinates, md.cell_volumes, F, md.boundary_faces,
                   num_ghosts=4).plan()
    assert e.value.status == _lib.MA_ERR_UNSUPPORTED


def test_layout_plan_detects_derivable_face_frames():
    md = M.Parallel3DMesh(6, 5, 4, 2.0, 2.0, 1.0, 2, 30.0).fillMeshData()
    assert md.plan()["derived_frames"]
    F = md.internal_faces
    t = F.face_tangent.copy()
    t[3, 0] = np.nextafter(t[3, 0], 9.0)
    bad = M.MeshData(md.num_owned_cells, md.cell_coordinates, md.cell_volumes,
                     M.Faces(F.coordinates, F.face_normal, t, F.face_binormal, F.face_cell_conn),
                     md.boundary_faces)
    assert not bad.plan()["derived_frames"]


def test_batch_plan_concatenates_case_layouts():
    """The batched layout is the cases' layouts side by side: tiles and tile
    face entries add up, the largest tile is the largest of any case, and
    the merged layout passes the library's own consistency check (every cell
    in one tile, every slot pointing at a face of that cell)."""
    from miniaero_amd.mesh import plan_batch
    meshes = [M.Parallel3DMesh(9, 7, 5, 1.0, 0.5, 0.25, 1, options=M.Options(y_stretch=1.05)).fillMeshData(),
              M.Parallel3DMesh(16, 8, 3, 2.0, 2.0, 1.0, 2, 30.0).fillMeshData(),
              M.Parallel3DMesh(1, 1, 1, 1.0, 1.0, 1.0, 0).fillMeshData()]
    # a mesh that is not a lattice: one cell relabelled and a face pair swapped
    md = M.Parallel3DMesh(4, 4, 4, 1.0, 1.0, 1.0, 0).fillMeshData()
    rng = np.random.default_rng(2)
    nc = md.num_owned_cells
    new = rng.permutation(nc)
    old = np.argsort(new)
    def relabel(F):
        conn = F.face_cell_conn.copy()
        m = conn >= 0
        conn[m] = new[conn[m]]
        return M.Faces(F.coordinates, F.face_normal, F.face_tangent, F.face_binormal, conn)
    meshes.append(M.MeshData(nc, md.cell_coordinates[old], md.cell_volumes[old],
                             relabel(md.internal_faces), [(n, relabel(F)) for n, F in md.boundary_faces]))
    parts = [m.plan() for m in meshes]
    p = plan_batch(meshes)
    assert p["lattice"] == (0, 0, 0)
    assert p["tiles"] == sum(q["tiles"] for q in parts)
    assert p["tile_face_entries"] == sum(q["tile_face_entries"] for q in parts)
    assert p["max_tile_faces"] == max(q["max_tile_faces"] for q in parts)
    assert p["derived_frames"] == all(q["derived_frames"] for q in parts)
    assert plan_batch(meshes[:1]) == parts[0]


def test_batch_create_rejects_mixed_schemes():
    md = M.Parallel3DMesh(2, 2, 2, 1, 1, 1, 0).fillMeshData()
    a = M.Options(dt=1e-6, second_order_space=1)
    for b in (a.replace(second_order_space=0), a.replace(viscous=1), a.replace(limiter=1),
              a.replace(precision=1), a.replace(prandtl=0.72)):
        with pytest.raises(M.MiniAeroError) as e:
            M.BatchSolver([(md, a), (md, b)])
        assert e.value.status == _lib.MA_ERR_INVALID
        assert "batched cases must share" in str(e.value)
    with pytest.raises(M.MiniAeroError) as e:
        M.BatchSolver([(md, a), (md, a.replace(dt=0.0))])
    assert e.value.status == _lib.MA_ERR_INVALID


def test_mesh_generation_on_threads_is_bit_exact():
    """Meshes generated concurrently equal meshes generated one at a time
    (the face shuffle draws from a private copy of rand()'s generator)."""
    from miniaero_amd.mesh import generate_meshes
    opts = [M.Options(problem_type=p, nx=nx, ny=ny, nz=nz, lx=1.0, ly=0.5, lz=0.5,
                      angle=10.0 if p == 2 else 0.0)
            for p, nx, ny, nz in ((0, 9, 7, 5), (2, 12, 6, 3), (1, 5, 11, 4), (0, 16, 4, 4))] * 3
    serial = generate_meshes(opts, max_workers=1)
    threaded = generate_meshes(opts, max_workers=8)
    for a, b in zip(serial, threaded):
        for Fa, Fb in [(a.internal_faces, b.internal_faces)] + [(x[1], y[1]) for x, y in
                                                                 zip(a.boundary_faces, b.boundary_faces)]:
            assert np.array_equal(Fa.face_cell_conn, Fb.face_cell_conn)
            assert np.array_equal(Fa.face_normal, Fb.face_normal)
            assert np.array_equal(Fa.permute_vector, Fb.permute_vector)
        assert np.array_equal(a.cell_volumes, b.cell_volumes)


def test_invalid_tile_shape_is_rejected_before_device_use():
    md = M.Parallel3DMesh(2, 2, 2, 1, 1, 1, 0).fillMeshData()
    for shape in ((16, 16, 2), (0, 4, 4), (4, -2, 4)):
        with pytest.raises(M.MiniAeroError) as e:
            M.TimeSolverExplicitRK4(md, M.Options(dt=1e-6, tile_shape=shape))
        assert e.value.status == _lib.MA_ERR_INVALID


def test_server_checks_config_and_device_cleanly():
    with pytest.raises(M.MiniAeroError) as e:
        M.Server(steps_in_flight=-1)
    assert e.value.status == _lib.MA_ERR_INVALID
    with pytest.raises(M.MiniAeroError) as e:
        M.Server(max_active=-1)
    assert e.value.status == _lib.MA_ERR_INVALID
    with pytest.raises(M.MiniAeroError) as e:
        M.Server(max_active=65)                          # above the largest slot limit run
    assert e.value.status == _lib.MA_ERR_INVALID
    if not os.path.exists("/dev/kfd"):
        with pytest.raises(M.MiniAeroError) as e:
            M.Server()
        assert e.value.status == _lib.MA_ERR_DEVICE


def test_served_stream_is_seeded_and_shaped():
    from miniaero_amd.configs import served_c5
    a, b = served_c5(300), served_c5(300)
    assert [(t, o) for t, o in a] == [(t, o) for t, o in b]
    t = np.array([x for x, _ in a])
    assert np.all(np.diff(t) > 0) and 0.5 < t[-1] / 300 * 200 < 2.0   # ~200 arrivals per second
    cells = np.array([o.nx * o.ny * o.nz for _, o in a])
    assert cells.min() >= 4096 and cells.max() <= 2097152
    assert all(o.nx == 4 * o.nz and o.ny == 2 * o.nz and o.precision == 1 for _, o in a)
    visc = np.array([o.viscous for _, o in a])
    assert 0.35 < visc.mean() < 0.65
    steps = np.array([o.ntimesteps for _, o in a])
    assert steps.min() >= 50 and steps.max() <= 500


def test_batch_plan_with_stored_and_derived_frames():
    """One mesh whose tangents do not follow Face.C (scaled by 2) makes the
    whole batch store face frames; the other parts' frames are then rebuilt
    from their area vectors on the host."""
    from miniaero_amd.mesh import plan_batch
    md = M.Parallel3DMesh(6, 5, 4, 2.0, 2.0, 1.0, 2, 30.0).fillMeshData()
    def scaled(F):
        return M.Faces(F.coordinates, F.face_normal, 2.0 * F.face_tangent, 2.0 * F.face_binormal,
                       F.face_cell_conn, F.cell_flux_index, F.permute_vector)
    md2 = M.MeshData(md.num_owned_cells, md.cell_coordinates, md.cell_volumes,
                     scaled(md.internal_faces), [(n, scaled(F)) for n, F in md.boundary_faces])
    assert plan_batch([md, md])["derived_frames"]
    p = plan_batch([md, md2, md])
    assert not p["derived_frames"]
    assert p["tiles"] == 3 * md.plan()["tiles"]


def test_cpp_many_cases_program_reports_missing_device(tmp_path):
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is present")
    import subprocess
    exe = os.path.join(ROOT, "miniaero_amd", "bin", "many_cases")
    assert os.path.exists(exe), "bin/many_cases not built"
    r = subprocess.run([exe], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "no HIP device" in r.stderr


def test_bench_algorithmic_bytes_match_the_stated_figures():
    """The roofline's algorithmic bytes per launch: the per-cell and per-face
    figures the kernel table states, on the 256x128x64 plate (FP64) and in
    the FP32 and face-value forms, and the survey's whole-step figure."""
    import sys
    sys.path.insert(0, ROOT)
    import bench
    nc, nif, nbf = 2097152, 6234112, 114688
    F = nif + nbf
    assert bench.algorithmic_bytes("k_flux", nc, nif, nbf, True, 8, 8) == 382 * nc + 104 * F
    assert bench.algorithmic_bytes("k_flux", nc, nif, nbf, True, 4, 4) == 246 * nc + 56 * F
    assert abs(bench.algorithmic_bytes("k_flux", nc, nif, nbf, True) - 1.4614e9) < 1e5
    assert bench.algorithmic_bytes("k_grad_limiter", nc, nif, nbf, True) == 232 * nc + 56 * F
    # face-value scheme: 5 values per face side instead of the cells' 25
    assert bench.algorithmic_bytes("k_flux", nc, nif, nbf, True, 4, 4, True) == (4 + 130) * nc + (8 + 36 + 40) * F
    assert bench.algorithmic_bytes("k_grad_limiter", nc, nif, nbf, True, 4, 4, True) == (140 + 16) * nc + 32 * F
    assert bench.algorithmic_bytes("k_prim", nc, nif, nbf, True) == 120 * nc
    assert bench.path_bytes_per_step(nc, nif, nbf, True, True) == 4 * (824 * nc + 168 * F)
    assert abs(bench.path_bytes_per_step(nc, nif, nbf, True, True) / (4 * nc) - 1333) < 1


def test_bench_plain_path_aggregate_weights_each_case_by_its_steps(monkeypatch):
    """The plain-path comparison beside the batched and served figures: each
    sampled case runs alone, its step time is taken over the timed steps
    only (setup and the untimed graph excluded), and the aggregate is the
    sample's work at each case's own step count over the summed stepping
    time -- here with a fake solver on a fake clock, two cases whose step
    times and step counts differ."""
    import sys
    sys.path.insert(0, ROOT)
    import bench

    clock = [0.0]
    step_s = {64: 1e-3, 512: 2e-3}            # seconds per step, by nx
    monkeypatch.setattr(bench.time, "perf_counter", lambda: clock[0])
    import miniaero_amd.mesh as mesh_mod
    monkeypatch.setattr(mesh_mod, "generate_meshes", lambda opts: [o.nx for o in opts])

    class FakeSolver:
        def __init__(self, md, o):
            assert o.precision == 0 and o.autotune == 0
            self.dt = step_s[md]
            clock[0] += 100.0                  # setup: must not count
        def set_graph(self, n):
            assert n == 10
        def initialize(self):
            clock[0] += 50.0
        def step(self, n):
            clock[0] += n * self.dt
        def synchronize(self):
            pass
        def nonfinite(self):
            return 0
        def close(self):
            pass

    class FakeM:
        TimeSolverExplicitRK4 = FakeSolver

    a = M.Options(nx=64, ny=4, nz=4, ntimesteps=100, precision=1)
    b = M.Options(nx=512, ny=4, nz=4, ntimesteps=300, precision=1)
    r = bench.plain_fp64_on_cases(FakeM, [a, b], timed_steps=20, graph=10)
    work = 64 * 16 * 4.0 * 100 + 512 * 16 * 4.0 * 300
    secs = 1e-3 * 100 + 2e-3 * 300
    assert r["cases"] == 2
    assert abs(r["value"] - work / secs) < 1e-6 * work / secs
    assert abs(r["per_case_rate"]["min"] - 64 * 16 * 4 / 1e-3) < 1e-3


def test_axis_aligned_face_frame_equals_general_path(tmp_path):
    """The kernels build the Roe face frame of an axis-aligned area vector with
    selects (axis_frame_t, physics.hpp); it must equal the general path's
    divisions and square roots bit for bit, FP64 and FP32 (host IEEE
    arithmetic), and refuse vectors that are not along an axis."""
    import shutil
    import subprocess
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    here = os.path.dirname(os.path.abspath(__file__))
    exe = str(tmp_path / "axis_frame_check")
    subprocess.run([hipcc, "-O2", "-std=c++17", "-ffp-contract=off", "-fno-fast-math", "--offload-arch=gfx950",
                    "-I", os.path.join(here, "..", "miniaero_amd", "csrc"),
                    os.path.join(here, "host_checks", "axis_frame_check.cpp"), "-o", exe],
                   check=True, capture_output=True)
    r = subprocess.run([exe, "2000000"], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    assert "mismatches=0" in r.stdout and "fast=0" not in r.stdout
