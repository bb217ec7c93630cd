"""Host-side mirror of the reference's hot-path interface over the C ABI.

  Solver            the device-resident q/qold/RK registers of module Global
                    (src/1d/Global1.f90) with step(), ts(), advance() --
                    step / ts_ssp* / the sharpclaw() loop of src/sharpclaw.f90
  Batch             many independent 1D grids of uneven size advanced by the
                    sharpclaw() loop in one persistent launch
  rp, evec, weno5, weno5_char, flux1
                    single routines with the reference's argument meaning
                    (rp1eu.f:5-6, flux1.f90:105, reconstruct.f90:77,146,
                    flux1.f90:2), on host arrays in the reference layout

Arrays are numpy float64 in Fortran order: q(1-mbc:nx+mbc[,1-mbc:ny+mbc], meqn).
"""
import ctypes as C
import os

import numpy as np

from ._abi import ScBatchResult, ScParams, ScRunCtl, ScRunStats, grid_shape
from ._lib import check, lib


def _f(a):
    return np.asfortranarray(a, dtype=np.float64)


class Solver:
    """One grid on one device (sc_create/sc_destroy)."""

    def __init__(self, params: ScParams):
        self.p = params
        h = C.c_void_p()
        check(lib().sc_create(C.byref(params), C.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().sc_destroy(self._h)
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

    @property
    def shape(self):
        return grid_shape(self.p)

    def set_q(self, q):
        q = _f(q)
        if q.shape != self.shape:
            raise ValueError(f"q has shape {q.shape}, expected {self.shape}")
        check(lib().sc_set_q(self._h, q))

    def get_q(self):
        q = np.zeros(self.shape, order="F")
        check(lib().sc_get_q(self._h, q))
        return q

    def set_aux(self, aux):
        """aux(1-mbc:mx+mbc, maux) of the grid (setaux's array); the device
        keeps the capacity column mcapa."""
        aux = _f(aux)
        want = tuple(self.shape[:-1]) + (self.p.maux,)
        if self.p.maux > 0 and aux.shape != want:
            raise ValueError(f"aux has shape {aux.shape}, expected {want}")
        check(lib().sc_set_aux(self._h, aux))

    def output_begin(self):
        """Start a frame: |q| < 1e-99 -> 0 in q itself (the reference's
        writers), ghosts filled, device snapshot copied to the host
        asynchronously; the solver may be advanced before output_wait()."""
        check(lib().sc_output_begin(self._h))

    def output_wait(self):
        """The grid snapshotted by the last output_begin()."""
        q = np.zeros(self.shape, order="F")
        check(lib().sc_output_wait(self._h, q))
        return q

    def write_restart(self, path, t, dt_next):
        check(lib().sc_write_restart(self._h, os.fsencode(path), t, dt_next))

    def read_restart(self, path):
        """Upload q from a restart file; returns (t, dt_next)."""
        t, d = C.c_double(), C.c_double()
        check(lib().sc_read_restart(self._h, os.fsencode(path), C.byref(t), C.byref(d)))
        return t.value, d.value

    def save_qold(self):
        check(lib().sc_save_qold(self._h))

    def restore_qold(self):
        check(lib().sc_restore_qold(self._h))

    def step(self, dt, t=0.0, want_dq=True):
        """step(): bc on q, dq = dt*L(q); returns (dq or None, cfl)."""
        cfl = C.c_double()
        if want_dq:
            dq = np.zeros(self.shape, order="F")
            check(lib().sc_step(self._h, dt, t, dq.ctypes.data_as(C.c_void_p), C.byref(cfl)))
            return dq, cfl.value
        check(lib().sc_step(self._h, dt, t, None, C.byref(cfl)))
        return None, cfl.value

    def flux(self, q, dq, dt, t=0.0):
        """flux1 / flux2 on the caller's q with its ghost cells as given
        (after the caller's bc): dq(interior) += dt*L(q) in place; returns
        the CFL number (2D: ghost rows and columns included)."""
        q = np.asfortranarray(q, dtype=np.float64)
        if q.shape != self.shape or dq.shape != self.shape or dq.dtype != np.float64 \
                or not dq.flags.f_contiguous:
            raise ValueError(f"q and dq must be float64 arrays of shape {self.shape} (dq Fortran-ordered)")
        cfl = C.c_double()
        check(lib().sc_flux(self._h, q.ctypes.data_as(C.c_void_p), dq.ctypes.data_as(C.c_void_p), dt, t,
                            C.byref(cfl)))
        return cfl.value

    def ts(self, dt, told=0.0):
        """One fused SSP time step from q (call save_qold first, as
        sharpclaw.f90:134 does); returns the last stage's CFL number."""
        cfl = C.c_double()
        check(lib().sc_ts(self._h, dt, told, C.byref(cfl)))
        return cfl.value

    def advance(self, ctl: ScRunCtl, tstart, tend, dt):
        """sharpclaw(tstart, tend, ...): returns (dt_last, stats)."""
        d = C.c_double(dt)
        st = ScRunStats()
        check(lib().sc_advance(self._h, C.byref(ctl), tstart, tend, C.byref(d), C.byref(st)))
        return d.value, st

    def sums(self):
        out = np.zeros(self.p.meqn, order="F")
        check(lib().sc_sums(self._h, out))
        return out

    def set_stage_timing(self, max_launches):
        """Record the device start and end of the next max_launches fused
        stage launches of ts/advance (0: off); advance then takes its
        host-driven loop."""
        check(lib().sc_stage_timing(self._h, int(max_launches)))

    def stage_timing(self):
        """(total ms, launches) of the recorded stage launches."""
        ms, n = C.c_double(), C.c_longlong()
        check(lib().sc_stage_timing_get(self._h, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def set_tile(self, tile):
        """2D tile shape: -1 autotune at first use, 0..3 fixed."""
        check(lib().sc_set_tile(self._h, tile))

    def get_tile(self):
        k, tx, ty = C.c_int(), C.c_int(), C.c_int()
        check(lib().sc_get_tile(self._h, C.byref(k), C.byref(tx), C.byref(ty)))
        return k.value, (tx.value, ty.value)

    def time_stage(self, n, dt):
        ms = C.c_double()
        check(lib().sc_time_stage(self._h, n, dt, C.byref(ms)))
        return ms.value


class Batch:
    """nsolves 1D grids sharing params (sc_batch_create); solve i has mx[i]
    cells on [xlower[i], xupper[i]]."""

    def __init__(self, params: ScParams, mx, xlower, xupper):
        self.p = params
        self.mx = np.ascontiguousarray(mx, dtype=np.int32)
        n = len(self.mx)
        xl = np.ascontiguousarray(np.broadcast_to(xlower, (n,)), dtype=np.float64)
        xu = np.ascontiguousarray(np.broadcast_to(xupper, (n,)), dtype=np.float64)
        h = C.c_void_p()
        check(lib().sc_batch_create(C.byref(params), n, self.mx.ctypes.data_as(C.POINTER(C.c_int)),
                                    xl.ctypes.data_as(C.POINTER(C.c_double)),
                                    xu.ctypes.data_as(C.POINTER(C.c_double)), C.byref(h)))
        self._h = h
        self.n = n

    def close(self):
        if getattr(self, "_h", None):
            lib().sc_batch_destroy(self._h)
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

    def shape(self, i):
        return (int(self.mx[i]) + 2 * self.p.mbc, self.p.meqn)

    def set_q(self, i, q):
        q = _f(q)
        if q.shape != self.shape(i):
            raise ValueError(f"q has shape {q.shape}, expected {self.shape(i)}")
        check(lib().sc_batch_set_q(self._h, i, q))

    def get_q(self, i):
        q = np.zeros(self.shape(i), order="F")
        check(lib().sc_batch_get_q(self._h, i, q))
        return q

    def run(self, ctl: ScRunCtl, tstart, tend, dt):
        """Every solve: sharpclaw(tstart, tend) from dt.  Returns
        (list of ScBatchResult, launch ms)."""
        res = (ScBatchResult * self.n)()
        ms = C.c_double()
        check(lib().sc_batch_run(self._h, C.byref(ctl), tstart, tend, dt, res, C.byref(ms)))
        return list(res), ms.value

    @property
    def block(self):
        b = C.c_int()
        check(lib().sc_batch_block(self._h, C.byref(b)))
        return b.value

    @property
    def parts(self):
        """Workgroups each solve ran on in the last run (1: one workgroup)."""
        out = np.zeros(self.n, dtype=np.int32)
        check(lib().sc_batch_parts(self._h, out.ctypes.data_as(C.POINTER(C.c_int))))
        return out

    def tuning(self):
        """{workgroup size: predicted whole-run ms} of the tuner (0: untimed;
        -1: a size the kernel is not built for with the batch's options)."""
        blocks = (C.c_int * 8)()
        pred = (C.c_double * 8)()
        n = C.c_int()
        check(lib().sc_batch_tuning(self._h, 8, blocks, pred, C.byref(n)))
        return {blocks[k]: pred[k] for k in range(n.value)}


def rp(p, ixy, mx, ql, qr):
    """rp(ixy,...,ql,qr,...) -> (wave, s, amdq, apdq), reference layout."""
    L = mx + 2 * p.mbc
    wave = np.zeros((L, p.meqn, p.mwaves), order="F")
    s = np.zeros((L, p.mwaves), order="F")
    amdq = np.zeros((L, p.meqn), order="F")
    apdq = np.zeros((L, p.meqn), order="F")
    check(lib().sc_rp(C.byref(p), ixy, mx, _f(ql), _f(qr), wave, s, amdq, apdq))
    return wave, s, amdq, apdq


def evec(p, ixy, mx, q):
    L = mx + 2 * p.mbc
    evl = np.zeros((L, p.meqn, p.meqn), order="F")
    evr = np.zeros((L, p.meqn, p.meqn), order="F")
    check(lib().sc_evec(C.byref(p), ixy, mx, _f(q), evl, evr))
    return evl, evr


def weno5(p, q):
    q = _f(q)
    ql = np.zeros_like(q, order="F")
    qr = np.zeros_like(q, order="F")
    check(lib().sc_weno5(C.byref(p), q.shape[0], q, ql, qr))
    return ql, qr


def weno5_char(p, q, evl, evr):
    q = _f(q)
    ql = np.zeros_like(q, order="F")
    qr = np.zeros_like(q, order="F")
    check(lib().sc_weno5_char(C.byref(p), q.shape[0], q, _f(evl), _f(evr), ql, qr))
    return ql, qr


def flux1(p, ixy, mx, q1d, dt):
    q1d = _f(q1d)
    dq = np.zeros_like(q1d, order="F")
    cfl = C.c_double()
    check(lib().sc_flux1(C.byref(p), ixy, mx, q1d, dq, dt, C.byref(cfl)))
    return dq, cfl.value


def copy_bench(nbytes=1 << 30, reps=10, device=0):
    """Achievable HBM bandwidth (GB/s) of a plain device copy kernel."""
    g = C.c_double()
    check(lib().sc_copy_bench(device, nbytes, reps, C.byref(g)))
    return g.value
