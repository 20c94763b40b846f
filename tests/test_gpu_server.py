"""Served form: cases admitted while others run, the cases of one scheme
advanced by the same launches (continuous batching).

Bar: every case that ran to its end equals, bit for bit, the same case run
alone with TimeSolverExplicitRK4 (itself held bit for bit against the oracle
by the parity tests), whatever cases it shared launches with, whenever it
joined and whether its mesh layout was planned for it or reused from an
earlier request of the same mesh; requests are admitted in submission order
with at most max_active on the device; cases that cannot run (bad options,
larger than the server's region) fail alone without disturbing the others;
the region is empty again at the end.
"""
import time

import numpy as np
import pytest

import miniaero_amd as M
from miniaero_amd import _lib
from miniaero_amd.configs import served_case

pytestmark = pytest.mark.gpu


def alone(opts):
    md = M.Parallel3DMesh(opts.nx, opts.ny, opts.nz, opts.lx, opts.ly, opts.lz, opts.problem_type,
                          opts.angle, options=opts).fillMeshData()
    s = M.TimeSolverExplicitRK4(md, opts)
    try:
        s.initialize()
        s.step(opts.ntimesteps)
        return s.state()
    finally:
        s.close()


def test_served_cases_match_their_solo_runs():
    rng = np.random.default_rng(7)
    cases = []
    for i in range(10):
        cases.append(served_case(bool(i % 2), float(rng.uniform(0.2, 0.8)), float(rng.uniform(-5, 5)),
                                 float(np.exp(rng.uniform(np.log(4096), np.log(40000)))),
                                 int(rng.integers(5, 30))))
    cases.append(cases[0].replace(ntimesteps=0))       # initial state only
    cases.append(cases[1].replace(ntimesteps=13))      # odd step count (state in the other buffer)
    bad = cases[2].replace(nx=0)                         # rejected by the mesh generator
    big = served_case(True, 0.5, 0.0, 600000, 5)         # larger than the region below
    srv = M.Server(hbm_budget_bytes=256 << 20, max_active=3, setup_threads=3)
    try:
        ids = {}
        for k, o in enumerate(cases[:6]):
            ids[srv.submit(o)] = o
        ids[srv.submit(bad)] = bad
        time.sleep(0.05)                                 # the rest arrives while cases run
        ids[srv.submit(big)] = big
        for o in cases[6:]:
            ids[srv.submit(o)] = o
            time.sleep(0.01)
        print(f"[gpu-server] submitted {len(ids)} requests", flush=True)
        results = {}
        t_end = time.time() + 120
        while len(results) < len(ids) and time.time() < t_end:
            for r in srv.poll(timeout=1.0):
                results[r["id"]] = r
                U = srv.state(r["id"], r["cells"]) if r["status"] == _lib.MA_OK else None
                results[r["id"]]["U"] = U
                srv.release(r["id"])
        assert len(results) == len(ids), f"only {len(results)} of {len(ids)} finished"
        st = srv.stats()
        pc = srv.plan_cache()
    finally:
        srv.close()
    print(f"[gpu-server] stats {st}, plan cache {pc}", flush=True)
    # the last two requests repeat the meshes of the first two: their layouts come from the cache
    assert pc["hits"] >= 2 and pc["hits"] + pc["misses"] == len(ids)
    assert st["finished"] == len(ids) and st["active"] == 0 and st["waiting"] == 0
    assert st["region_used"] == 0
    by_opts = {i: o for i, o in ids.items()}
    for i, r in results.items():
        o = by_opts[i]
        if o is bad:
            assert r["status"] == _lib.MA_ERR_INVALID
            continue
        if o is big:
            assert r["status"] == _lib.MA_ERR_NOMEM
            continue
        assert r["status"] == _lib.MA_OK, r
        assert r["steps"] == o.ntimesteps
        assert r["t_submit"] <= r["t_admit"] <= r["t_finish"]
        ref = alone(o)
        print(f"[gpu-server] case {i}: {r['cells']} cells, {o.ntimesteps} steps, viscous {o.viscous}, "
              f"bitwise={np.array_equal(r['U'], ref)}", flush=True)
        assert np.array_equal(r["U"], ref), f"case {i} differs from its solo run"
    # admission in submission order, never more than max_active on the device
    ok = sorted((r for r in results.values() if r["status"] == _lib.MA_OK), key=lambda r: r["id"])
    assert all(a["t_admit"] <= b["t_admit"] for a, b in zip(ok, ok[1:]))
    events = sorted([(r["t_admit"], 1) for r in ok] + [(r["t_finish"], -1) for r in ok])
    live, peak = 0, 0
    for _, d in events:
        live += d
        peak = max(peak, live)
    assert 2 <= peak <= 3
