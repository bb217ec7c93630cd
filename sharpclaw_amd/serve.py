"""Request server (sc_server_* of the C ABI): continuous batching of
independent solves on one device.

    with Server(hbm_bytes=8 << 30) as srv:
        rid = srv.submit(params, ctl, q0, tstart, tend, dt)
        ...
        for rid in srv.poll():
            q, stats = srv.result(rid)

Each request runs the sharpclaw() time loop (src/sharpclaw.f90:35-235) on
its own grid; stats carry the loop's record plus the server clock times of
submission, admission and completion (ms since the server was created).
"""
import ctypes as C

import numpy as np

from ._abi import ScParams, ScRequestStats, ScRunCtl, ScServerStats, grid_shape
from ._lib import check, lib


class Server:
    def __init__(self, hbm_bytes=8 << 30, max_active=4096, device=0):
        h = C.c_void_p()
        check(lib().sc_server_create(device, int(hbm_bytes), int(max_active), C.byref(h)))
        self._h = h
        self._shape = {}

    def close(self):
        if getattr(self, "_h", None):
            lib().sc_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def submit(self, params: ScParams, ctl: ScRunCtl, q, tstart, tend, dt):
        shape = grid_shape(params)
        q = np.asfortranarray(q, dtype=np.float64)
        if q.shape != shape:
            raise ValueError(f"q has shape {q.shape}, expected {shape}")
        rid = C.c_longlong()
        check(lib().sc_server_submit(self._h, C.byref(params), C.byref(ctl), float(tstart), float(tend),
                                     float(dt), q, C.byref(rid)))
        self._shape[rid.value] = shape
        return rid.value

    def poll(self, max_ids=4096):
        ids = (C.c_longlong * max_ids)()
        n = C.c_int()
        check(lib().sc_server_poll(self._h, ids, max_ids, C.byref(n)))
        return [ids[k] for k in range(n.value)]

    def wait(self, timeout_ms=100.0):
        n = C.c_int()
        check(lib().sc_server_wait(self._h, float(timeout_ms), C.byref(n)))
        return n.value

    def result(self, rid):
        q = np.zeros(self._shape.pop(rid), order="F")
        st = ScRequestStats()
        check(lib().sc_server_result(self._h, rid, q, C.byref(st)))
        return q, st

    def drain(self, timeout_ms=600e3):
        check(lib().sc_server_drain(self._h, float(timeout_ms)))

    def stats(self):
        st = ScServerStats()
        check(lib().sc_server_get_stats(self._h, C.byref(st)))
        return st.as_dict()
