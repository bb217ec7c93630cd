"""ORACLE -- test infrastructure only.

ctypes front end of oracle/sc_oracle.c (the C restatement of the reference's
semi-discrete step).  Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg import this module, and only as the checker.

Arrays are numpy float64 in Fortran order, in the reference's layout.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from sharpclaw_amd._abi import ScParams, ScRunCtl, ScRunStats, grid_shape, interior

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", os.environ.get("ORACLE_LIB", "liboracle.so"))

_lib = None


def build(force=False):
    """Compile sc_oracle.c with gcc (no FMA contraction, as the reference build)."""
    if force or not os.path.exists(LIB) or os.path.getmtime(LIB) < os.path.getmtime(
            os.path.join(HERE, "sc_oracle.c")):
        subprocess.check_call(["make", "-s", "-C", HERE, "lib"])
    return LIB


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="F_CONTIGUOUS")
        P = C.POINTER(ScParams)
        L.orc_rp.argtypes = [P, C.c_int, C.c_int, dp, dp, dp, dp, dp, dp]
        L.orc_evec.argtypes = [P, C.c_int, C.c_int, dp, dp, dp]
        L.orc_weno5.argtypes = [C.c_int, C.c_int, C.c_int, dp, dp, dp]
        L.orc_weno5_char.argtypes = [C.c_int, C.c_int, C.c_int, dp, dp, dp, dp, dp]
        L.orc_flux1.argtypes = [P, C.c_int, C.c_int, dp, dp, C.c_double]
        L.orc_flux1.restype = C.c_double
        L.orc_flux2.argtypes = [P, dp, dp, C.c_double, C.c_int]
        L.orc_flux2.restype = C.c_double
        L.orc_bc.argtypes = [P, dp]
        L.orc_step.argtypes = [P, dp, dp, C.c_double, C.c_double]
        L.orc_step.restype = C.c_double
        L.orc_ts.argtypes = [P, dp, dp, dp, dp, C.c_double, C.c_double, C.POINTER(C.c_longlong)]
        L.orc_ts.restype = C.c_double
        L.orc_sharpclaw.argtypes = [P, C.POINTER(ScRunCtl), dp, C.c_double, C.c_double,
                                    C.POINTER(C.c_double), C.POINTER(ScRunStats)]
        L.orc_sums.argtypes = [P, dp, dp]
        L.orc_set_aux.argtypes = [C.c_void_p]
        L.orc_swirl_setaux.argtypes = [P, dp]
        _lib = L
    return _lib


def _f(a):
    return np.asfortranarray(a, dtype=np.float64)


_aux_ref = None


def _set_aux(p, aux):
    """The grid's aux array (1-mbc:mx+mbc, maux) for the next call (1D runs
    with maux > 0: capacity).  Kept referenced while in use."""
    global _aux_ref
    if p.maux > 0:
        if aux is None:
            raise ValueError("maux > 0 needs the aux array")
        _aux_ref = _f(aux)
        lib().orc_set_aux(_aux_ref.ctypes.data_as(C.c_void_p))
    else:
        _aux_ref = None
        lib().orc_set_aux(None)


def rp(p, ixy, mx, ql, qr):
    L = mx + 2 * p.mbc
    ql, qr = _f(ql), _f(qr)
    wave = np.zeros((L, p.meqn, p.mwaves), order="F")
    s = np.zeros((L, p.mwaves), order="F")
    amdq = np.zeros((L, p.meqn), order="F")
    apdq = np.zeros((L, p.meqn), order="F")
    lib().orc_rp(C.byref(p), ixy, mx, ql, qr, wave, s, amdq, apdq)
    return wave, s, amdq, apdq


def evec(p, ixy, mx, q):
    L = mx + 2 * p.mbc
    evl = np.zeros((L, p.meqn, p.meqn), order="F")
    evr = np.zeros((L, p.meqn, p.meqn), order="F")
    lib().orc_evec(C.byref(p), ixy, mx, _f(q), evl, evr)
    return evl, evr


def weno5(q):
    q = _f(q)
    mx2, meqn = q.shape
    ql = np.zeros_like(q, order="F")
    qr = np.zeros_like(q, order="F")
    lib().orc_weno5(mx2, meqn, mx2, q, ql, qr)
    return ql, qr


def weno5_char(q, evl, evr):
    q = _f(q)
    mx2, meqn = q.shape
    ql = np.zeros_like(q, order="F")
    qr = np.zeros_like(q, order="F")
    lib().orc_weno5_char(mx2, meqn, mx2, q, _f(evl), _f(evr), ql, qr)
    return ql, qr


def flux1(p, ixy, mx, q1d, dt, dq1d=None, aux=None):
    _set_aux(p, aux)
    q1d = _f(q1d)
    dq = np.zeros_like(q1d, order="F") if dq1d is None else _f(dq1d).copy(order="F")
    cfl = lib().orc_flux1(C.byref(p), ixy, mx, q1d, dq, dt)
    return dq, cfl


def flux2(p, q, dt, with_ghost_sweeps=True):
    q = _f(q)
    dq = np.zeros_like(q, order="F")
    cfl = lib().orc_flux2(C.byref(p), q, dq, dt, 1 if with_ghost_sweeps else 0)
    return dq, cfl


def bc(p, q):
    q = _f(q).copy(order="F")
    rc = lib().orc_bc(C.byref(p), q)
    if rc:
        raise ValueError("invalid mthbc")
    return q


def step(p, q, dt, t=0.0, aux=None):
    """step(): returns (q with ghosts filled, dq, cfl)."""
    _set_aux(p, aux)
    q = _f(q).copy(order="F")
    dq = np.zeros_like(q, order="F")
    cfl = lib().orc_step(C.byref(p), q, dq, dt, t)
    return q, dq, cfl


def ts(p, q, dt, told=0.0, aux=None):
    """One SSP time step from q (qold = q); returns (q_new, cfl, nevals)."""
    _set_aux(p, aux)
    q = _f(q).copy(order="F")
    qold = q.copy(order="F")
    r1 = np.zeros_like(q, order="F")
    r2 = np.zeros_like(q, order="F")
    n = C.c_longlong(0)
    cfl = lib().orc_ts(C.byref(p), q, qold, r1, r2, dt, told, C.byref(n))
    return q, cfl, n.value


def sharpclaw(p, ctl, q, tstart, tend, dt, aux=None):
    """The time loop of one output interval; returns (q, dt_last, stats)."""
    _set_aux(p, aux)
    q = _f(q).copy(order="F")
    d = C.c_double(dt)
    st = ScRunStats()
    lib().orc_sharpclaw(C.byref(p), C.byref(ctl), q, tstart, tend, C.byref(d), C.byref(st))
    return q, d.value, st


def flush_small(p, q):
    """The frame writers' in-place reset (src/1d/out1.f90:63-68,
    src/2d/out2.f90:60-67): interior values with |q| < 1e-99 become +0."""
    q = _f(q).copy(order="F")
    qi = interior(p, q)
    qi[np.abs(qi) < 1e-99] = 0.0
    return q


def swirl_setaux(p, maux=2):
    """apps/advection/2d/swirl/setaux.f: edge velocities of the swirl
    stream function, aux(1-mbc:nx+mbc, 1-mbc:ny+mbc, maux) (columns past 2
    zero)."""
    aux = np.zeros((p.nx[0] + 2 * p.mbc, p.nx[1] + 2 * p.mbc, maux), order="F")
    lib().orc_swirl_setaux(C.byref(p), aux)
    return aux


def sums(p, q):
    out = np.zeros(p.meqn)
    lib().orc_sums(C.byref(p), _f(q), out)
    return out


def new_grid(p):
    return np.zeros(grid_shape(p), order="F")
