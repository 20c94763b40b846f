"""Served form: Server over the library's ma_server_* entry points.

    srv = Server(hbm_budget_bytes=20e9, max_active=16)
    i = srv.submit(options)          # mesh generated from the options
    for r in srv.poll(timeout=1.0):  # finished cases, completion order
        U = srv.state(r["id"], r["cells"]); srv.release(r["id"])

Cases are planned on host threads, admitted in submission order when a slot
is free and their device bytes fit in the server's HBM region, and the
cases of one scheme are advanced by the same launches (continuous
batching); each result equals the case run alone (TimeSolverExplicitRK4)
bit for bit.  The reference has no counterpart: it
runs one case per process (Main.C:96-141).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .options import Options


class Server:
    def __init__(self, hbm_budget_bytes: float = 0, max_active: int = 0, setup_threads: int = 0,
                 steps_in_flight: int = 0, plan_cache_bytes: float = 0):
        self._L = _lib.load()
        cfg = _lib.ma_server_config(int(hbm_budget_bytes), int(max_active), int(setup_threads),
                                    int(steps_in_flight), int(plan_cache_bytes))
        h = ctypes.c_void_p()
        _lib.check(self._L.ma_server_create(ctypes.byref(cfg), ctypes.byref(h)))
        self._h = h

    def submit(self, options: Options) -> int:
        i = ctypes.c_longlong()
        o = options.to_c()
        _lib.check(self._L.ma_server_submit(self._h, ctypes.byref(o), ctypes.byref(i)))
        return int(i.value)

    def poll(self, max_results: int = 256, timeout: float = 0.0) -> list:
        out = (_lib.ma_case_result * max_results)()
        n = self._L.ma_server_poll(self._h, out, max_results, float(timeout))
        if n < 0:
            _lib.check(_lib.MA_ERR_INVALID)
        return [dict(id=r.id, status=r.status, steps=r.steps, cells=r.cells, device_bytes=r.device_bytes,
                     t_submit=r.t_submit, t_admit=r.t_admit, t_finish=r.t_finish) for r in out[:n]]

    def state(self, case_id: int, cells: int) -> np.ndarray:
        U = np.empty((int(cells), 5))
        _lib.check(self._L.ma_server_get_state(self._h, int(case_id), U.ctypes.data_as(_lib._D)))
        return U

    def release(self, case_id: int) -> None:
        _lib.check(self._L.ma_server_release(self._h, int(case_id)))

    def stats(self) -> dict:
        s = (ctypes.c_longlong * 6)()
        _lib.check(self._L.ma_server_stats(self._h, s))
        return dict(submitted=s[0], waiting=s[1], active=s[2], finished=s[3], region_used=s[4],
                    region_size=s[5])

    def plan_cache(self) -> dict:
        """Requests that reused a cached mesh layout, and those that planned one."""
        h, m = ctypes.c_longlong(), ctypes.c_longlong()
        _lib.check(self._L.ma_server_plan_cache(self._h, ctypes.byref(h), ctypes.byref(m)))
        return dict(hits=h.value, misses=m.value)

    def now(self) -> float:
        return float(self._L.ma_server_now(self._h))

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._L.ma_server_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
