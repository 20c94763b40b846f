"""Batched form: independent cases stepped together in one set of launches.

Bar: every case of a batch equals, bit for bit, the same case run alone
(which the parity tests hold bit for bit against the oracle), and its time
and last dt are the same.  Cases differ in mesh size and geometry, initial
state, inflow, fixed dt or CFL number and seed.
"""
import numpy as np
import pytest

import miniaero_amd as M
from oracle import oracle as O

pytestmark = pytest.mark.gpu

EXT = ("bc_kind", "y_stretch", "limiter", "ic_kind", "ic", "use_inflow", "inflow", "prandtl",
       "cfl", "perturb", "seed")


def make_case(kw):
    opts = M.Options(**{k: v for k, v in kw.items() if k in M.Options.__dataclass_fields__})
    ext = {k: v for k, v in kw.items() if k in EXT}
    md = M.Parallel3DMesh(kw["nx"], kw["ny"], kw["nz"], kw["lx"], kw["ly"], kw["lz"],
                          kw["problem_type"], kw.get("angle", 0.0),
                          options=M.Options(**ext)).fillMeshData()
    return md, opts


def plate(**kw):
    base = dict(problem_type=1, lx=1.0, ly=0.05, lz=0.025, nx=16, ny=12, nz=4, dt=1e-7,
                second_order_space=1, viscous=1)
    base.update(kw)
    return base


VISCOUS_CASES = [
    plate(nx=16, ny=32, nz=2, lx=2.0, ly=0.002, lz=1.0, dt=3e-8),
    plate(nx=24, ny=16, nz=4, dt=0.0, y_stretch=1.05, bc_kind=(0, 2, 0, 0, 0, 0), use_inflow=1,
          inflow=(0.01772, 0.01772 * 104.2, 0.0, 0.0,
                  0.01772 * 287.05 / 0.4 * 300.0 + 0.5 * 0.01772 * 104.2 ** 2), cfl=0.5),
    plate(nx=12, ny=10, nz=6, ic_kind=2, perturb=0.01, seed=3),
    dict(problem_type=2, lx=1.0, ly=0.5, lz=0.5, angle=10.0, nx=1, ny=1, nz=1, dt=1e-7,
         second_order_space=1, viscous=1),
    plate(nx=12, ny=10, nz=6, ic_kind=2, perturb=0.02, seed=11, cfl=0.3, dt=0.0),
    dict(problem_type=0, lx=1.0, ly=0.25, lz=0.25, nx=32, ny=8, nz=8, dt=1e-5,
         second_order_space=1, viscous=1, ic_kind=1, ic=(1.0, 1.0, 0.125, 0.1, 0.4),
         bc_kind=(1, 1, 1, 1, 1, 1)),
]

VANALBADA_CASES = [
    dict(problem_type=0, lx=1.0, ly=0.25, lz=0.25, nx=nx, ny=ny, nz=nz, dt=dt,
         second_order_space=1, viscous=0, limiter=1, ic_kind=1, ic=ic, bc_kind=(1, 1, 1, 1, 1, 1))
    for nx, ny, nz, dt, ic in (
        (32, 8, 8, 1e-4, (1.0, 1.0, 0.125, 0.1, 0.5)),
        (17, 5, 3, 2e-4, (2.0, 3.0, 0.5, 0.4, 0.3)),
        (40, 4, 6, 5e-5, (1.0, 1.0, 0.125, 0.1, 0.7)),
    )
] + [dict(problem_type=2, lx=2.0, ly=2.0, lz=1.0, angle=30.0, nx=16, ny=8, nz=3, dt=1e-5,
          second_order_space=1, viscous=0, limiter=1)]


def run_batch_vs_alone(kws, nsteps):
    cases = [make_case(kw) for kw in kws]
    print(f"[gpu-batch] {len(cases)} cases, cells "
          f"{[int(md.num_owned_cells) for md, _ in cases]}, {nsteps} steps", flush=True)
    b = M.BatchSolver(cases)
    try:
        b.initialize()
        b.step(nsteps)
        got = [b.case_state(k) for k in range(len(cases))]
        times = [b.case_time(k) for k in range(len(cases))]
        whole = b.state()
    finally:
        b.close()
    assert np.array_equal(whole, np.vstack(got))
    for k, (md, opts) in enumerate(cases):
        s = M.TimeSolverExplicitRK4(md, opts)
        try:
            s.initialize()
            s.step(nsteps)
            U = s.state()
            t = s.time()
        finally:
            s.close()
        assert np.array_equal(got[k], U), f"case {k} differs from its own run"
        assert times[k] == t, f"case {k}: time {times[k]} vs {t}"
    return got


def test_viscous_batch_matches_cases_alone():
    run_batch_vs_alone(VISCOUS_CASES, 12)


def test_vanalbada_batch_matches_cases_alone():
    run_batch_vs_alone(VANALBADA_CASES, 15)


def test_batch_case_matches_oracle():
    """One batched case against the oracle directly (not only via its own run)."""
    kw = VISCOUS_CASES[2]
    got = run_batch_vs_alone([VISCOUS_CASES[0], kw, VISCOUS_CASES[3]], 5)
    ok = O.make_options(**{k: v for k, v in kw.items()})
    om = O.OracleMesh(ok)
    cpu = O.OracleSolver(om)
    cpu.init()
    cpu.step(5)
    assert np.array_equal(got[1], cpu.state())


def test_many_small_cases():
    rng = np.random.default_rng(5)
    kws = []
    for i in range(40):
        nx, ny, nz = (int(x) for x in rng.integers(1, 9, size=3))
        kws.append(dict(problem_type=0, lx=1.0, ly=0.5, lz=0.5, nx=nx, ny=ny, nz=nz,
                        dt=float(rng.uniform(1e-5, 1e-4)), second_order_space=1, viscous=0, limiter=1,
                        ic_kind=1, ic=(1.0, 1.0, 0.125, 0.1, float(rng.uniform(0.2, 0.8))),
                        bc_kind=(1, 1, 1, 1, 1, 1)))
    run_batch_vs_alone(kws, 6)


@pytest.mark.parametrize("graph_steps,scheme", [(2, "viscous"), (4, "viscous"), (2, "inviscid")])
def test_graph_replays_match_plain_launches(graph_steps, scheme):
    """Steps run as captured HIP graphs (several replays, from both ping-pong
    parities, with eager work and host reads between them) equal plain
    launches bit for bit, for a batch with fixed-dt and CFL cases (viscous)
    and for the inviscid second-order face-value scheme."""
    kws = ([VISCOUS_CASES[1], VISCOUS_CASES[2], VISCOUS_CASES[4]] if scheme == "viscous"
           else VANALBADA_CASES[:3])
    cases = [make_case(kw) for kw in kws]
    a = M.BatchSolver(cases)
    b = M.BatchSolver(cases)
    try:
        a.initialize()
        b.initialize()
        b.set_graph(graph_steps)
        for n in (graph_steps + 1, 3, 2 * graph_steps, 1, graph_steps):
            a.step(n)
            b.step(n)
            Ua, Ub = a.state(), b.state()
            print(f"[gpu-batch] graph {graph_steps}: +{n} steps, bitwise={np.array_equal(Ua, Ub)}",
                  flush=True)
            assert np.array_equal(Ua, Ub)
            assert b.nonfinite() == 0
            for k in range(len(cases)):
                assert a.case_time(k) == b.case_time(k)
        # profiling (plain launches with events) still works after graphs
        b.profile(2)
        a.step(2)
        assert np.array_equal(a.state(), b.state())
    finally:
        a.close()
        b.close()


@pytest.mark.parametrize("scheme,precision", [("viscous", 0), ("viscous", 1), ("inviscid", 0), ("inviscid", 1),
                                              ("first_order", 0), ("first_order_viscous", 1)])
def test_persistent_steps_match_plain_launches(scheme, precision):
    """Steps run as one cooperative launch with grid-wide barriers between the
    phases (ma_solver_set_persistent) equal plain launches bit for bit: state,
    time and last dt of every case, over odd and even step counts, after
    set_state (the launch then computes the first stage's primitives itself)
    and after the residual and gradients are read between calls.  The
    inviscid batch holds a 524k-cell case, more flux tiles than the launch
    has resident workgroups, so each workgroup walks several."""
    big = dict(VANALBADA_CASES[0], nx=128, ny=64, nz=64, dt=2e-5)
    kws = {"viscous": [VISCOUS_CASES[1], VISCOUS_CASES[2], VISCOUS_CASES[4]],
           "inviscid": VANALBADA_CASES[:3] + [big],
           "first_order": [dict(k, second_order_space=0) for k in VANALBADA_CASES[:2] + [big]],
           "first_order_viscous": [dict(k, second_order_space=0) for k in VISCOUS_CASES[1:3]]}[scheme]
    kws = [dict(k, precision=precision) for k in kws]
    cases = [make_case(kw) for kw in kws]
    a = M.BatchSolver(cases)
    b = M.BatchSolver(cases)
    try:
        a.initialize()
        b.initialize()
        b.set_persistent(True)
        for n in (3, 1, 4, 2):
            a.step(n)
            b.step(n)
            Ua, Ub = a.state(), b.state()
            print(f"[gpu-batch] persistent {scheme} fp{64 - 32 * precision}: +{n} steps, "
                  f"bitwise={np.array_equal(Ua, Ub)}", flush=True)
            assert np.array_equal(Ua, Ub)
            for k in range(len(cases)):
                assert a.case_time(k) == b.case_time(k)
        assert np.array_equal(a.residual(), b.residual())
        ga, gb = a.gradients(), b.gradients()
        assert all(np.array_equal(x, y) for x, y in zip(ga, gb))
        U = a.state()
        U[:, 0] *= 1.0 + 1e-3 * np.cos(np.arange(len(U)))
        a.set_state(U)
        b.set_state(U)
        a.step(3)
        b.step(3)
        assert np.array_equal(a.state(), b.state())
        assert b.nonfinite() == 0
        b.set_persistent(False)  # back to plain launches
        a.step(1)
        b.step(1)
        assert np.array_equal(a.state(), b.state())
    finally:
        a.close()
        b.close()


def test_graph_steps_only_on_the_creating_thread():
    """Graph capture and replay are confined to the thread that created the
    solver: from another thread a step that would use a graph is refused
    with MA_ERR_UNSUPPORTED before anything is queued, shorter steps run as
    plain launches, and the run continues on the creating thread with the
    bits of an all-plain run."""
    import threading

    cases = [make_case(VISCOUS_CASES[1])]
    a = M.BatchSolver(cases)
    b = M.BatchSolver(cases)
    out = {}

    def other(n):
        try:
            b.step(n)
            out[n] = "ok"
        except M.MiniAeroError as e:
            out[n] = e.status

    try:
        a.initialize()
        b.initialize()
        b.set_graph(2)
        for n in (2, 1):
            t = threading.Thread(target=other, args=(n,))
            t.start()
            t.join()
        print(f"[gpu-batch] off-thread steps: {out}", flush=True)
        assert out == {2: 2, 1: "ok"}
        b.step(4)          # two graph replays on the creating thread
        a.step(5)
        assert np.array_equal(a.state(), b.state())
    finally:
        a.close()
        b.close()


@pytest.mark.parametrize("shape,wg", [((-1, 0, 0), 0), ((1, 1, 1), 0), ((2, 8, 16), 1), ((8, 8, 4), 0),
                                      ((16, 4, 4), 2), ((3, 5, 7), 0), ((4, 8, 8), 1), ((4, 8, 8), 3)])
def test_tile_shape_does_not_change_results(shape, wg):
    """Every cell sums its faces in the same order whatever the flux tiles
    are and however many workgroups walk them, so any tile shape and any
    resident-workgroup count give the default layout's bits."""
    # 92k cells: more tiles per XCD than 32 CUs x 1 workgroup, so the
    # workgroups of the capped grid walk several tiles each
    kw = plate(nx=64, ny=40, nz=36, y_stretch=1.05, bc_kind=(3, 2, 2, 2, 1, 4), dt=0.0, cfl=0.5)
    md, opts = make_case(kw)
    out = []
    for o in (opts, opts.replace(tile_shape=shape, workgroups_per_cu=wg)):
        s = M.TimeSolverExplicitRK4(md, o)
        try:
            s.initialize()
            s.step(6)
            out.append(s.state())
            got_shape, _ = s.tile_shape()
        finally:
            s.close()
    print(f"[gpu-batch] tile shape {shape} -> {got_shape}, bitwise={np.array_equal(out[0], out[1])}",
          flush=True)
    assert got_shape == ((-1, -1, -1) if shape[0] < 0 else shape)
    assert np.array_equal(out[0], out[1])


def test_small_run_default_tiles():
    """A run with fewer cells than one 4x8x8 tile per compute unit gets 4x4x4
    tiles by default, a larger one 4x8x8; either gives the bits of the other
    shape, so the choice is a matter of speed only."""
    for n, want in ((16, (4, 4, 4)), (64, (4, 8, 8))):
        kw = plate(nx=n, ny=32, nz=32, y_stretch=1.05, bc_kind=(3, 2, 2, 2, 1, 4), dt=0.0, cfl=0.5)
        md, opts = make_case(kw)
        out = []
        for o in (opts, opts.replace(tile_shape=(4, 8, 8) if want == (4, 4, 4) else (4, 4, 4))):
            s = M.TimeSolverExplicitRK4(md, o)
            try:
                s.initialize()
                s.step(4)
                out.append(s.state())
                if len(out) == 1:
                    got, _ = s.tile_shape()
            finally:
                s.close()
        print(f"[gpu-batch] {n * 32 * 32} cells: default tiles {got}, bitwise={np.array_equal(out[0], out[1])}",
              flush=True)
        assert got == want
        assert np.array_equal(out[0], out[1])


@pytest.mark.parametrize("shape,wg", [((-1, 0, 0), 0), ((1, 1, 1), 0), ((3, 5, 7), 0), ((2, 8, 16), 1)])
def test_fp32_viscous_tiles_are_shape_invariant(shape, wg):
    """The FP32 form is held to a tolerance against FP64, but against itself
    it is deterministic: any tile shape and resident-workgroup count gives
    the default layout's bits."""
    kw = plate(nx=64, ny=40, nz=36, y_stretch=1.05, bc_kind=(3, 2, 2, 2, 1, 4), dt=0.0, cfl=0.5)
    md, opts = make_case(kw)
    opts = opts.replace(precision=1)
    out = []
    for o in (opts, opts.replace(tile_shape=shape, workgroups_per_cu=wg)):
        s = M.TimeSolverExplicitRK4(md, o)
        try:
            s.initialize()
            s.step(6)
            out.append(s.state())
        finally:
            s.close()
    print(f"[gpu-batch] fp32 viscous, tile shape {shape}, wg {wg}: bitwise={np.array_equal(out[0], out[1])}",
          flush=True)
    assert np.array_equal(out[0], out[1])


@pytest.mark.parametrize("shape,wg", [((-1, 0, 0), 0), ((1, 1, 1), 0), ((3, 5, 7), 0), ((16, 4, 4), 2),
                                      ((4, 8, 8), 3)])
def test_face_value_scheme_tile_shapes(shape, wg):
    """Inviscid second order (the gradient kernel stores each face's
    extrapolated values; the tile lists carry each face's slot in its two
    cells): any tile shape and resident-workgroup count gives the default
    layout's bits, and those equal the oracle's."""
    kw = dict(problem_type=0, lx=1.0, ly=0.25, lz=0.25, nx=64, ny=40, nz=36, dt=2e-5,
              second_order_space=1, viscous=0, limiter=1, ic_kind=1, ic=(1.0, 1.0, 0.125, 0.1, 0.45),
              bc_kind=(1, 1, 1, 1, 1, 1))
    md, opts = make_case(kw)
    out = []
    for o in (opts, opts.replace(tile_shape=shape, workgroups_per_cu=wg)):
        s = M.TimeSolverExplicitRK4(md, o)
        try:
            s.initialize()
            s.step(4)
            out.append(s.state())
        finally:
            s.close()
    print(f"[gpu-batch] face values, tile shape {shape}, wg {wg}: bitwise={np.array_equal(out[0], out[1])}",
          flush=True)
    assert np.array_equal(out[0], out[1])
    if shape == (3, 5, 7):
        ok = O.make_options(**kw)
        cpu = O.OracleSolver(O.OracleMesh(ok))
        cpu.init()
        cpu.step(4)
        assert np.array_equal(out[1], cpu.state())


def test_autotuned_batch_matches_default():
    kws = VANALBADA_CASES[:3]
    cases = [make_case(kw) for kw in kws]
    tuned = [(md, o.replace(autotune=1)) for md, o in cases]
    out = []
    for cs in (cases, tuned):
        b = M.BatchSolver(cs)
        try:
            b.initialize()
            b.step(8)
            out.append(b.state())
            shape, ms = b.tile_shape()
        finally:
            b.close()
    print(f"[gpu-batch] autotuned shape {shape}, {ms:.4f} ms/step/Mcell", flush=True)
    assert ms > 0.0
    assert np.array_equal(out[0], out[1])


def test_batch_residual_norms_are_per_case():
    cases = [make_case(kw) for kw in VISCOUS_CASES[:3]]
    b = M.BatchSolver(cases)
    try:
        b.initialize()
        b.step(4)
        n = b.residual_norms()
        R = b.residual()
    finally:
        b.close()
    assert n.shape == (3, 5)
    for k in range(3):
        lo, hi = b.offsets[k], b.offsets[k + 1]
        assert np.allclose(n[k], np.linalg.norm(R[lo:hi], axis=0), rtol=1e-13, atol=0.0)


def test_batch_mixing_stored_and_derived_frames():
    """A batch with one mesh whose frames must be stored (tangents scaled by
    2) stores frames for every case; each case still equals its solo run."""
    kw = VANALBADA_CASES[3]
    md, opts = make_case(kw)
    def scaled(F):
        return M.Faces(F.coordinates, F.face_normal, 2.0 * F.face_tangent, 2.0 * F.face_binormal,
                       F.face_cell_conn, F.cell_flux_index, F.permute_vector)
    md2 = M.MeshData(md.num_owned_cells, md.cell_coordinates, md.cell_volumes,
                     scaled(md.internal_faces), [(n, scaled(F)) for n, F in md.boundary_faces])
    cases = [(md, opts), (md2, opts), make_case(VANALBADA_CASES[0])]
    b = M.BatchSolver(cases)
    try:
        b.initialize()
        b.step(6)
        got = [b.case_state(k) for k in range(3)]
    finally:
        b.close()
    for k, (mesh, o) in enumerate(cases):
        s = M.TimeSolverExplicitRK4(mesh, o)
        try:
            s.initialize()
            s.step(6)
            assert np.array_equal(got[k], s.state()), f"case {k}"
        finally:
            s.close()


def test_cpp_facade_batched_and_served_cases(tmp_path):
    """bin/many_cases: the C++ facade's BatchSolver and Server against solo
    TimeSolverExplicitRK4 runs of the same cases, bit for bit."""
    import os
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = os.path.join(root, "miniaero_amd", "bin", "many_cases")
    r = subprocess.run([exe], cwd=tmp_path, capture_output=True, text=True, timeout=240)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0 and r.stdout.strip().endswith("OK")


@pytest.mark.parametrize("graph_steps", [0, 2])
def test_restart_from_a_state_equals_the_uninterrupted_run(graph_steps):
    """set_state after other steps (so the solver holds the primitives of a
    different state), then steps -- plain or as graph replays -- equal the
    uninterrupted run from that state bit for bit (CFL time step: dt comes
    from the state alone)."""
    for kw in (VISCOUS_CASES[4], VANALBADA_CASES[0]):
        md, opts = make_case(kw)
        a = M.TimeSolverExplicitRK4(md, opts)
        b = M.TimeSolverExplicitRK4(md, opts)
        try:
            a.initialize()
            a.step(3)
            U3 = a.state()
            a.step(4)
            U7 = a.state()
            if graph_steps:
                b.set_graph(graph_steps)
            b.initialize()
            b.step(2)
            b.set_state(U3)
            b.step(4)
            same = np.array_equal(b.state(), U7)
            print(f"[gpu-batch] restart, graph {graph_steps}, viscous {opts.viscous}: bitwise={same}", flush=True)
            assert same
        finally:
            a.close()
            b.close()
