"""The reconstructions beyond weno5 / weno5_char on the device (1D):
weno5_wave (lim_type 2, char_decomp 1), weno5_trans (2, 3), tvd2_wave
(1, 1) and tvd2_char (1, 2) with every limiter (src/flux1.f90:65-117,
src/reconstruct.f90:244-763), through the C ABI against the oracle, and
whole runs against the reference's own output (tests/golden/run_*.npz).
Tolerances as test_gpu_parity: float64 routines and steps 1e-12 normwise
relative L-infinity; runs max(1e-10, 8x the oracle's 1-ulp sensitivity);
float32 relative L1 1e-5.
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
NAMES = ["advection", "acoustics", "euler", "shallow_water"]
RECON = [(2, 1), (2, 3), (1, 1), (1, 2)]
G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _limiters(lim):
    return (1, 2, 3, 4, 5) if lim == 1 else (5,)


@pytest.mark.parametrize("name", NAMES)
@pytest.mark.parametrize("lim,char", RECON)
def test_flux1_recon(name, lim, char):
    rng = np.random.default_rng(51)
    for mth in _limiters(lim):
        p = make_params(name, (40,), char_decomp=char, lim_type=lim, mthlim=(mth,) * 4)
        for mx in (7, 254, 255, 256, 600):
            q1d = _states.random_cells(p, (mx + 6,), rng, smooth=True,
                                       amp=0.3 if mx < 32 else 1.0)
            dt = 0.2 * p.dx[0]
            g, gc = S.flux1(p, 1, mx, q1d, dt)
            r, rc = O.flux1(p, 1, mx, q1d, dt)
            assert rel_linf(g[3:3 + mx], r[3:3 + mx]) <= TOL, (mth, mx)
            assert abs(gc - rc) <= TOL * rc, (mth, mx)


@pytest.mark.parametrize("name", NAMES)
@pytest.mark.parametrize("lim,char", RECON)
def test_step_recon(name, lim, char):
    rng = np.random.default_rng(52)
    bcs = [(1, 1, 1, 1), (2, 2, 1, 1), (3, 3, 1, 1), (1, 3, 1, 1)]
    if name == "advection":
        bcs = bcs[:2]
    for mth in _limiters(lim):
        for bc in bcs:
            p = make_params(name, (300,), char_decomp=char, lim_type=lim, mthlim=(mth,) * 4,
                            mthbc=bc)
            q = _states.grid(p, rng, smooth=True)
            dt = 0.1 * p.dx[0]
            with Solver(p) as s:
                s.set_q(q)
                gdq, gc = s.step(dt)
                gq = s.get_q()
            rq, rdq, rc = O.step(p, q, dt)
            assert np.array_equal(gq, rq), (mth, bc)
            assert rel_linf(interior(p, gdq), interior(p, rdq)) <= TOL, (mth, bc)
            assert abs(gc - rc) <= TOL * rc, (mth, bc)


def _golden_params(d, name, nx):
    c = [int(v) for v in d["params"]]
    return make_params(name, nx, xlower=tuple(float(v) for v in d["xl"]),
                       xupper=tuple(float(v) for v in d["xu"]), char_decomp=c[0], lim_type=c[1],
                       mthlim=(c[2],) * 4, time_integrator=c[3], mthbc=tuple(c[4:8]))


def _run_tol(p, ctl, q0, t0, t1, dt0, rq):
    qp = q0.copy(order="F")
    interior(p, qp)[:] = np.nextafter(interior(p, qp), np.inf)
    rq2, _, _ = O.sharpclaw(p, ctl, qp, t0, t1, dt0)
    return max(TOL_RUN, 8.0 * rel_linf(interior(p, rq2), interior(p, rq)))


@pytest.mark.parametrize("fname,name,nx", [("run_eu1d_shuosher_trans.npz", "euler", (400,)),
                                           ("run_ac1d_pulse_wave.npz", "acoustics", (200,)),
                                           ("run_adv1d_tvd2wave.npz", "advection", (200,))])
def test_run_matches_reference(fname, name, nx):
    d = np.load(os.path.join(G, fname))
    p = _golden_params(d, name, nx)
    c = d["ctl"]
    ctl = ScRunCtl(dt_max=float(c[0]), cfl_max=float(c[1]), cfl_desired=float(c[2]),
                   max_steps=int(c[3]), dt_variable=int(c[4]))
    q0 = O.new_grid(p)
    interior(p, q0)[:] = d["q"][0]
    t0, t1, dt0 = float(d["t"][0]), float(d["t"][-1]), float(d["dt_initial"])
    rq, _, rst = O.sharpclaw(p, ctl, q0, t0, t1, dt0)
    assert np.array_equal(interior(p, rq), d["q"][-1])
    tol = _run_tol(p, ctl, q0, t0, t1, dt0, rq)
    with Solver(p) as s:
        s.set_q(q0)
        _, gst = s.advance(ctl, t0, t1, dt0)
        gq = s.get_q()
    assert gst.steps_taken - gst.rejected == len(d["dtlog"])
    err = rel_linf(interior(p, gq), d["q"][-1])
    print(f"{fname}: rel Linf vs reference {err:.3e}, tolerance {tol:.3e}")
    assert err <= tol


@pytest.mark.parametrize("mth", [1, 2, 3, 4, 5])
def test_tvd2_char_dambreak_run(mth):
    """tvd2_char whole runs against the oracle (the reference program
    aborts in free() after a tvd2_char run: the routine writes past its
    workspace, reconstruct.f90:598-606)."""
    p = make_params("shallow_water", (300,), char_decomp=2, lim_type=1, mthlim=(mth,) * 4,
                    time_integrator=3)
    q0 = dambreak_q(300, 2.0, 0.5, 0.45)
    ctl = ScRunCtl(dt_max=1e99, cfl_max=0.5, cfl_desired=0.45, max_steps=50000, dt_variable=1)
    rq, _, rst = O.sharpclaw(p, ctl, q0, 0.0, 0.05, 1e-4)
    tol = _run_tol(p, ctl, q0, 0.0, 0.05, 1e-4, rq)
    with Solver(p) as s:
        s.set_q(q0)
        _, gst = s.advance(ctl, 0.0, 0.05, 1e-4)
        gq = s.get_q()
    assert gst.steps_taken == rst.steps_taken
    assert rel_linf(interior(p, gq), interior(p, rq)) <= tol


@pytest.mark.parametrize("lim,char", RECON)
def test_fp32_recon_l1(lim, char):
    p = make_params("euler", (800,), xlower=(-5.0,), xupper=(5.0,), char_decomp=char,
                    lim_type=lim, mthlim=(4, 4, 4, 4), time_integrator=3)
    # the entropy wave of the Shu-Osher problem without its shock: one step
    # of weno5_wave at that shock gives NaN in cells 81-82, in the reference
    # and the oracle alike (same NaN pattern)
    x = p.xlower[0] + (np.arange(1, 801) - 0.5) * p.dx[0]
    q0 = O.new_grid(p)
    qi = interior(p, q0)
    qi[:, 0] = 1.0 + 0.2 * np.sin(5.0 * x)
    qi[:, 1] = 0.3 * qi[:, 0]
    qi[:, 2] = 1.0 / 0.4 + 0.5 * qi[:, 1] ** 2 / qi[:, 0]
    ctl = ScRunCtl(dt_max=1e99, cfl_max=0.5, cfl_desired=0.45, max_steps=50000, dt_variable=1)
    rq, _, _ = O.sharpclaw(p, ctl, q0, 0.0, 0.1, 1e-4)
    p32 = make_params("euler", (800,), xlower=(-5.0,), xupper=(5.0,), char_decomp=char,
                      lim_type=lim, mthlim=(4, 4, 4, 4), time_integrator=3, precision=SC_FP32)
    with Solver(p32) as s:
        s.set_q(q0)
        s.advance(ctl, 0.0, 0.1, 1e-4)
        gq = s.get_q()
    g, o = interior(p, gq), interior(p, rq)
    err = np.sum(np.abs(g - o)) / np.sum(np.abs(o))
    print(f"float32 lim_type {lim} char_decomp {char}: relative L1 {err:.3e}")
    assert err <= 1e-5


def test_recon_options_rejected_where_not_built():
    bad = [make_params("euler", (16,), lim_type=0),
           make_params("euler", (16,), lim_type=1, char_decomp=0),
           make_params("euler", (16,), lim_type=1, char_decomp=3),
           make_params("euler", (16,), lim_type=1, char_decomp=2, mthlim=(6, 1, 1, 1)),
           make_params("advection", (16, 16), char_decomp=3, maux=2),
           make_params("euler", (16,), char_decomp=4)]
    for p in bad:
        with pytest.raises(ScError):
            Solver(p)
    # batches and the server take every 1D reconstruction, not the invalid ones
    p = make_params("shallow_water", (8,), lim_type=1, char_decomp=3)
    with pytest.raises(ScError):
        Batch(p, np.array([64, 65], dtype=np.int32), np.zeros(2), np.ones(2))
    with Server(hbm_bytes=16 << 20, max_active=2) as srv:
        pr = make_params("shallow_water", (64,), lim_type=1, char_decomp=0)
        ctl = ScRunCtl(dt_max=1e99, cfl_max=0.5, cfl_desired=0.45, max_steps=100, dt_variable=1)
        with pytest.raises(ScError):
            srv.submit(pr, ctl, dambreak_q(64, 2.0, 0.5, 0.5), 0.0, 0.01, 1e-4)
