"""sc_step_noalloc (the step of sc_step on caller-owned device arrays, with
no allocation) against sc_step through the C ABI.

The caller's device arrays come from hipMalloc of the HIP runtime the
library itself is linked against (libamdhip64.so.7, already loaded with
it).  Two calls in a row, the second on a smaller grid with slower waves:
the shared CFL scratch must start from zero on each call (a CFL number left
over from the first call would exceed the second's), and q (ghost cells
filled), dq and the CFL number must equal sc_step's bit for bit (the same
kernels on the same data).  Each case runs on a 2D grid (the fused tile
kernel) and on a 1D grid (the 1D stage kernel).
"""
import ctypes as C

import numpy as np
import pytest

import _states
from sharpclaw_amd import Solver, grid_shape, interior, make_params
from sharpclaw_amd._lib import check, lib

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def hip():
    lib()  # loads libamdhip64.so.7 with the library
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipDeviceSynchronize.argtypes = []
    return h


class DevArray:
    def __init__(self, h, nbytes):
        self.h = h
        self.p = C.c_void_p()
        assert h.hipMalloc(C.byref(self.p), nbytes) == 0
        self.nbytes = nbytes

    def put(self, a):
        a = np.asfortranarray(a, dtype=np.float64)
        assert a.nbytes == self.nbytes
        assert self.h.hipMemcpy(self.p, a.ctypes.data_as(C.c_void_p), a.nbytes, H2D) == 0

    def get(self, shape):
        a = np.zeros(shape, order="F")
        assert self.h.hipMemcpy(a.ctypes.data_as(C.c_void_p), self.p, a.nbytes, D2H) == 0
        return a

    def free(self):
        if self.p:
            self.h.hipFree(self.p)
            self.p = None


def noalloc_step(p, q0, dt):
    h = hip()
    L = lib()
    L.sc_step_noalloc.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_double,
                                  C.c_void_p, C.c_void_p]
    L.sc_step_noalloc.restype = C.c_int
    q = DevArray(h, q0.nbytes)
    dq = DevArray(h, q0.nbytes)
    cfl = DevArray(h, 8)
    try:
        q.put(q0)
        dq.put(np.full(q0.shape, np.nan, order="F"))  # every entry must be written
        check(L.sc_step_noalloc(C.byref(p), q.p, dq.p, dt, 0.0, cfl.p, None))
        assert h.hipDeviceSynchronize() == 0
        return q.get(q0.shape), dq.get(q0.shape), float(cfl.get((1,))[0])
    finally:
        q.free()
        dq.free()
        cfl.free()


def reference_step(p, q0, dt):
    with Solver(p) as s:
        s.set_q(q0)
        dq, cfl = s.step(dt)
        return s.get_q(), dq, cfl


def grid_state(p, seed, amp):
    """Smooth physical data in the interior, zero ghost cells."""
    rng = np.random.default_rng(seed)
    q = np.zeros(grid_shape(p), order="F")
    interior(p, q)[:] = interior(p, _states.grid(p, rng, smooth=True, amp=amp))
    return q


CASES = [
    # (physics, big grid, small grid, char_decomp, mthbc)
    ("euler", (512, 384), (100, 70), 2, (1, 1, 3, 3)),
    ("shallow_water", (300, 300), (64, 96), 2, (3, 3, 1, 1)),
    ("euler", (6000,), (700,), 2, (1, 1, 1, 1)),
    ("shallow_water", (4096,), (300,), 0, (3, 3, 3, 3)),
]


@pytest.mark.parametrize("name,big,small,char,bc", CASES)
def test_noalloc_equals_step_two_calls(name, big, small, char, bc):
    out = []
    for nx, amp, seed in ((big, 1.0, 3), (small, 0.2, 4)):
        p = make_params(name, nx, char_decomp=char, time_integrator=3, mthbc=bc)
        q0 = grid_state(p, seed, amp)
        dt = 0.2 * p.dx[0]
        qa, dqa, ca = noalloc_step(p, q0, dt)
        qb, dqb, cb = reference_step(p, q0, dt)
        assert np.array_equal(qa, qb), (name, nx)
        assert np.array_equal(dqa, dqb), (name, nx)
        assert ca == cb, (name, nx, ca, cb)
        out.append(ca)
    # the second grid's waves are slower: a CFL word left from the first call would show
    assert out[1] < out[0]


def test_noalloc_rejects_what_needs_buffers():
    L = lib()
    L.sc_step_noalloc.restype = C.c_int
    for kw in (dict(char_decomp=1), dict(lim_type=1, char_decomp=2, mthlim=(4, 4, 4, 4))):
        p = make_params("euler", (64, 64), time_integrator=3, **kw)
        x = C.c_void_p(1)
        rc = L.sc_step_noalloc(C.byref(p), x, x, C.c_double(1e-3), C.c_double(0.0), x, None)
        assert rc == 2, rc  # SC_ERR_UNSUPPORTED, before any device work
        assert b"grid-sized buffer" in L.sc_last_error()
