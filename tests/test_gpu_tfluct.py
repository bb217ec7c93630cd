"""tfluct_solver = 1 (src/flux1.f90:136-153) on the device, through the C ABI.

The reference takes the in-cell total fluctuation from the application's
tfluct routine; advection (apps/advection/1d/tfluct1.f), acoustics
(apps/acoustics/1d/homogeneous/tfluct.f) and 1D shallow water
(apps/shallowWater/2dRadialDam/1dRadialSymm-qwave/tfluct1sw.f) ship one.
Tolerances are those of test_gpu_parity: float64 steps 1e-12 normwise
relative L-infinity, whole runs max(1e-10, 8x the oracle's 1-ulp
sensitivity), float32 relative L1 1e-5.  The whole runs are also checked
against the reference's own output (tests/golden/run_*_tfluct.npz).
"""
import os

import numpy as np
import pytest

import _states
from conftest import rel_linf
from oracle import oracle as O
from sharpclaw_amd import (SC_FP32, Batch, ScError, ScRunCtl, Server, Solver, interior,
                           make_params)
from sharpclaw_amd import solver as S
from sharpclaw_amd.workloads import dambreak_q

pytestmark = pytest.mark.gpu

TOL = 1e-12
TOL_RUN = 1e-10
NAMES = ["advection", "acoustics", "shallow_water"]
G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _bcs(name):
    bcs = [(1, 1, 1, 1), (2, 2, 1, 1), (3, 3, 1, 1), (1, 3, 1, 1)]
    return [b for b in bcs if name != "advection" or 3 not in b[:2]]


@pytest.mark.parametrize("name", NAMES)
@pytest.mark.parametrize("char", [0, 2])
def test_flux1_tfluct(name, char):
    rng = np.random.default_rng(31)
    p = make_params(name, (40,), char_decomp=char, tfluct_solver=1)
    for mx in (7, 254, 255, 256, 600):
        q1d = _states.random_cells(p, (mx + 6,), rng, smooth=True, amp=0.3 if mx < 32 else 1.0)
        dt = 0.2 * p.dx[0]
        g, gc = S.flux1(p, 1, mx, q1d, dt)
        r, rc = O.flux1(p, 1, mx, q1d, dt)
        assert rel_linf(g[3:3 + mx], r[3:3 + mx]) <= TOL, mx
        assert abs(gc - rc) <= TOL * rc


@pytest.mark.parametrize("name", NAMES)
@pytest.mark.parametrize("char", [0, 2])
def test_step_tfluct(name, char):
    rng = np.random.default_rng(32)
    for bc in _bcs(name):
        p = make_params(name, (300,), char_decomp=char, mthbc=bc, tfluct_solver=1)
        q = _states.grid(p, rng, smooth=True)
        dt = 0.1 * p.dx[0]
        with Solver(p) as s:
            s.set_q(q)
            gdq, gc = s.step(dt)
            gq = s.get_q()
        rq, rdq, rc = O.step(p, q, dt)
        assert np.array_equal(gq, rq), bc
        assert rel_linf(interior(p, gdq), interior(p, rdq)) <= TOL, bc
        assert abs(gc - rc) <= TOL * rc, bc


def _golden_run(fname, p):
    d = np.load(os.path.join(G, fname))
    c = d["ctl"]
    ctl = ScRunCtl(dt_max=float(c[0]), cfl_max=float(c[1]), cfl_desired=float(c[2]),
                   max_steps=int(c[3]), dt_variable=int(c[4]))
    q0 = O.new_grid(p)
    interior(p, q0)[:] = d["q"][0]
    return d, ctl, q0, float(d["t"][0]), float(d["t"][-1]), float(d["dt_initial"])


@pytest.mark.parametrize("fname,kw", [
    ("run_sw1d_dambreak_tfluct.npz",
     dict(name="shallow_water", nx=(200,), char_decomp=2, time_integrator=3, mthbc=(1, 1, 1, 1))),
    ("run_ac1d_pulse_tfluct.npz",
     dict(name="acoustics", nx=(200,), xlower=(-1.0,), xupper=(1.0,), char_decomp=2,
          time_integrator=4, mthbc=(3, 3, 1, 1))),
])
def test_advance_tfluct_matches_reference_run(fname, kw):
    kw = dict(kw)
    p = make_params(kw.pop("name"), kw.pop("nx"), tfluct_solver=1, **kw)
    d, ctl, q0, t0, t1, dt0 = _golden_run(fname, p)
    rq, _, rst = O.sharpclaw(p, ctl, q0, t0, t1, dt0)
    assert np.array_equal(interior(p, rq), d["q"][-1])
    qp = q0.copy(order="F")
    interior(p, qp)[:] = np.nextafter(interior(p, qp), np.inf)
    rq2, _, _ = O.sharpclaw(p, ctl, qp, t0, t1, dt0)
    tol = max(TOL_RUN, 8.0 * rel_linf(interior(p, rq2), interior(p, rq)))
    with Solver(p) as s:
        s.set_q(q0)
        _, gst = s.advance(ctl, t0, t1, dt0)
        gq = s.get_q()
    assert gst.steps_taken - gst.rejected == len(d["dtlog"]) == rst.steps_taken - rst.rejected
    assert gst.t == t1
    err = rel_linf(interior(p, gq), d["q"][-1])
    print(f"{fname}: rel Linf vs reference {err:.3e}, tolerance {tol:.3e}")
    assert err <= tol


def test_fp32_tfluct_l1():
    p = make_params("shallow_water", (400,), char_decomp=2, time_integrator=3, tfluct_solver=1)
    q0 = dambreak_q(400, 2.0, 0.5, 0.45)
    ctl = ScRunCtl(dt_max=1e99, cfl_max=0.5, cfl_desired=0.45, max_steps=50000, dt_variable=1)
    rq, _, _ = O.sharpclaw(p, ctl, q0, 0.0, 0.05, 1e-4)
    p32 = make_params("shallow_water", (400,), char_decomp=2, time_integrator=3, tfluct_solver=1,
                      precision=SC_FP32)
    with Solver(p32) as s:
        s.set_q(q0)
        s.advance(ctl, 0.0, 0.05, 1e-4)
        gq = s.get_q()
    g, o = interior(p, gq), interior(p, rq)
    err = np.sum(np.abs(g - o)) / np.sum(np.abs(o))
    print("tfluct float32 relative L1:", err)
    assert err <= 1e-5


def test_batch_and_server_tfluct():
    """Batched and served dam breaks with tfluct_solver = 1: each solve
    equals its own single-grid run (sc_advance) bitwise and the oracle to
    the whole-run tolerance."""
    mx = np.array([63, 64, 200, 257, 513], dtype=np.int32)
    n = len(mx)
    rng = np.random.default_rng(33)
    hl, hr, x0 = rng.uniform(1, 3, n), rng.uniform(0.1, 1, n), rng.uniform(0.3, 0.7, n)
    base = make_params("shallow_water", (8,), char_decomp=2, time_integrator=3, tfluct_solver=1)
    qs = [dambreak_q(int(mx[i]), hl[i], hr[i], x0[i]) for i in range(n)]
    ctl = ScRunCtl(dt_max=1e99, cfl_max=0.5, cfl_desired=0.45, max_steps=100000, dt_variable=1)
    with Batch(base, mx, np.zeros(n), np.ones(n)) as b:
        for i in range(n):
            b.set_q(i, qs[i])
        res, _ = b.run(ctl, 0.0, 0.05, 1e-4)
        bq = [b.get_q(i) for i in range(n)]
    with Server(hbm_bytes=64 << 20, max_active=8) as srv:
        ids = []
        for i in range(n):
            p = make_params("shallow_water", (int(mx[i]),), char_decomp=2, time_integrator=3,
                            tfluct_solver=1)
            ids.append(srv.submit(p, ctl, qs[i], 0.0, 0.05, 1e-4))
        srv.drain(60e3)
        sq = [srv.result(r)[0] for r in ids]
    for i in range(n):
        p = make_params("shallow_water", (int(mx[i]),), char_decomp=2, time_integrator=3,
                        tfluct_solver=1)
        with Solver(p) as s:
            s.set_q(qs[i])
            _, gst = s.advance(ctl, 0.0, 0.05, 1e-4)
            gq = s.get_q()
        assert np.array_equal(interior(p, bq[i]), interior(p, gq)), i
        assert np.array_equal(interior(p, sq[i]), interior(p, gq)), i
        assert res[i].steps_taken == gst.steps_taken
        rq, _, rst = O.sharpclaw(p, ctl, qs[i], 0.0, 0.05, 1e-4)
        assert rst.steps_taken == gst.steps_taken
        # the dam break's near-still water amplifies rounding (1 ulp in q0
        # moves the oracle's result by up to 6e-8 here): whole-run tolerance
        qp = qs[i].copy(order="F")
        interior(p, qp)[:] = np.nextafter(interior(p, qp), np.inf)
        rq2, _, _ = O.sharpclaw(p, ctl, qp, 0.0, 0.05, 1e-4)
        tol = max(TOL_RUN, 8.0 * rel_linf(interior(p, rq2), interior(p, rq)))
        assert rel_linf(interior(p, gq), interior(p, rq)) <= tol, i


def test_tfluct_rejected_without_a_reference_tfluct():
    for p in (make_params("euler", (16,), tfluct_solver=1),
              make_params("shallow_water", (16, 16), tfluct_solver=1),
              make_params("advection", (16,), tfluct_solver=2)):
        with pytest.raises(ScError):
            Solver(p)
