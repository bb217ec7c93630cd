"""The reconstructions beyond weno5 / weno5_char on 2D grids: weno5_wave,
weno5_trans, tvd2_wave and tvd2_char for Euler and shallow water, run as the
reference's flux2 forms a stage (x sweeps over the rows j = 0..ny+1, y sweeps
over the columns i = 0..nx+1, each a 1D flux1 line; flux2.f90:36-92).
Through the C ABI against the oracle (steps on grids of several 127-cell
chunks in either direction, whole runs with every integrator) and against
the reference's own output (tests/golden/run_*2d_*.npz).  Tolerances as
test_gpu_parity: float64 steps 1e-12 normwise relative L-infinity; runs
max(1e-10, 8x the oracle's 1-ulp sensitivity); float32 relative L1 1e-5.
"""
import os

import numpy as np
import pytest

import _states
from conftest import rel_linf
from oracle import oracle as O
from sharpclaw_amd import SC_FP32, ScError, ScRunCtl, Server, Solver, interior, make_params
from sharpclaw_amd import solver as S
from test_gpu_parity import rel_l1, run_tolerance

pytestmark = pytest.mark.gpu

TOL = 1e-12
RECON = [(2, 1), (2, 3), (1, 1), (1, 2)]
G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _limiters(lim):
    return (1, 2, 3, 4, 5) if lim == 1 else (5,)


def bump(p, u0, v0, amp=0.5):
    """Uniform flow (u0, v0) with a smooth compact bump of relative size amp
    around (0.35, 0.6), radius 0.25 of the unit square."""
    x = (np.arange(p.nx[0]) + 0.5) / p.nx[0]
    y = (np.arange(p.nx[1]) + 0.5) / p.nx[1]
    X, Y = np.meshgrid(x, y, indexing="ij")
    r2 = ((X - 0.35) ** 2 + (Y - 0.6) ** 2) / 0.25 ** 2
    b = np.where(r2 < 1.0, (1.0 - r2) ** 3, 0.0)
    q = O.new_grid(p)
    qi = interior(p, q)
    if p.meqn == 4:
        rho, pr = 1.0 + amp * b, 1.0 + 1.6 * amp * b
        qi[..., 0], qi[..., 1], qi[..., 2] = rho, rho * u0, rho * v0
        qi[..., 3] = pr / 0.4 + 0.5 * rho * (u0 * u0 + v0 * v0)
    else:
        h = 1.0 + amp * b
        qi[..., 0], qi[..., 1], qi[..., 2] = h, h * u0, h * v0
    return q


@pytest.mark.parametrize("name", ["euler", "shallow_water"])
@pytest.mark.parametrize("lim,char", RECON)
def test_step_recon2d(name, lim, char):
    """One step (bc + dq = dt L(q) + CFL) on grids of several chunks along
    x (130 x 29) and along y (27 x 260), every boundary kind."""
    rng = np.random.default_rng(61)
    bcs = [(1, 1, 1, 1), (2, 2, 2, 2), (3, 3, 3, 3), (1, 3, 2, 2)]
    for mth in _limiters(lim):
        for shape in ((130, 29), (27, 260)):
            for bc in (bcs if mth == 5 else [bcs[mth % len(bcs)]]):
                p = make_params(name, shape, char_decomp=char, lim_type=lim, mthlim=(mth,) * 4,
                                mthbc=bc)
                q = _states.grid(p, rng, smooth=True)
                dt = 0.1 * min(p.dx[0], p.dx[1])
                print(f"{name} lim_type {lim} char_decomp {char} mthlim {mth} {shape} bc {bc}",
                      flush=True)
                with Solver(p) as s:
                    s.set_q(q)
                    gdq, gc = s.step(dt)
                    gq = s.get_q()
                rq, rdq, rc = O.step(p, q, dt)
                assert np.array_equal(gq, rq), (mth, shape, bc)
                assert rel_linf(interior(p, gdq), interior(p, rdq)) <= TOL, (mth, shape, bc)
                assert abs(gc - rc) <= TOL * rc, (mth, shape, bc)


@pytest.mark.parametrize("ti", [1, 2, 3, 4])
@pytest.mark.parametrize("name,lim,char,mth", [("euler", 2, 1, 5), ("euler", 2, 3, 5),
                                               ("euler", 1, 1, 4), ("euler", 1, 2, 2),
                                               ("shallow_water", 2, 1, 5),
                                               ("shallow_water", 2, 3, 5),
                                               ("shallow_water", 1, 1, 3),
                                               ("shallow_water", 1, 2, 1)])
def test_advance_recon2d(name, lim, char, mth, ti):
    """Whole runs (accept/reject and dt control included) of every
    integrator against the oracle; the steps and rejections must agree
    exactly.  The first dt gives a CFL number above cfl_max, so every run
    starts with a rejected step and its retake."""
    p = make_params(name, (48, 40), char_decomp=char, lim_type=lim, mthlim=(mth,) * 4,
                    time_integrator=ti, mthbc=(1, 3, 2, 2))
    q0 = bump(p, 0.3, -0.2)
    cmax = 2.5 if ti == 4 else 0.5
    dt0 = 0.06 if ti == 4 else 0.02
    ctl = ScRunCtl(dt_max=1e99, cfl_max=cmax, cfl_desired=0.9 * cmax, max_steps=25, dt_variable=1)
    rq, rdt, rst = O.sharpclaw(p, ctl, q0, 0.0, 10.0, dt0)
    assert rst.rejected >= 1
    with Solver(p) as s:
        s.set_q(q0)
        gdt, gst = s.advance(ctl, 0.0, 10.0, dt0)
        gq = s.get_q()
    assert (gst.steps_taken, gst.rejected) == (rst.steps_taken, rst.rejected)
    assert abs(gdt - rdt) <= 1e-10 * rdt
    tol = run_tolerance(p, ctl, q0, 0.0, 10.0, dt0, rq)
    err = rel_linf(interior(p, gq), interior(p, rq))
    print(f"{name} {lim}/{char} mthlim {mth} ti {ti}: {rst.steps_taken} steps, "
          f"{rst.rejected} rejected, rel Linf {err:.3e} (tol {tol:.3e})")
    assert err <= tol


def _golden_params(d, name, nx):
    c = [int(v) for v in d["params"]]
    return make_params(name, nx, xlower=tuple(float(v) for v in d["xl"]),
                       xupper=tuple(float(v) for v in d["xu"]), char_decomp=c[0], lim_type=c[1],
                       mthlim=(c[2],) * 4, time_integrator=c[3], mthbc=tuple(c[4:8]))


@pytest.mark.parametrize("fname,name", [("run_sw2d_pulse_wave.npz", "shallow_water"),
                                        ("run_eu2d_quad_trans.npz", "euler"),
                                        ("run_eu2d_quad_tvd2wave.npz", "euler")])
def test_run_recon2d_matches_reference(fname, name):
    d = np.load(os.path.join(G, fname))
    p = _golden_params(d, name, (24, 24))
    c = d["ctl"]
    ctl = ScRunCtl(dt_max=float(c[0]), cfl_max=float(c[1]), cfl_desired=float(c[2]),
                   max_steps=int(c[3]), dt_variable=int(c[4]))
    q0 = O.new_grid(p)
    interior(p, q0)[:] = d["q"][0]
    t0, t1, dt0 = float(d["t"][0]), float(d["t"][-1]), float(d["dt_initial"])
    rq, _, _ = O.sharpclaw(p, ctl, q0, t0, t1, dt0)
    assert np.array_equal(interior(p, rq), d["q"][-1])
    tol = run_tolerance(p, ctl, q0, t0, t1, dt0, rq)
    with Solver(p) as s:
        s.set_q(q0)
        _, gst = s.advance(ctl, t0, t1, dt0)
        gq = s.get_q()
    assert gst.steps_taken - gst.rejected == len(d["dtlog"])
    err = rel_linf(interior(p, gq), d["q"][-1])
    print(f"{fname}: rel Linf vs reference {err:.3e}, tolerance {tol:.3e}")
    assert err <= tol


@pytest.mark.parametrize("lim,char", RECON)
def test_fp32_recon2d_l1(lim, char):
    p = make_params("euler", (64, 64), char_decomp=char, lim_type=lim, mthlim=(4, 4, 4, 4),
                    time_integrator=3, mthbc=(2, 2, 2, 2))
    q0 = bump(p, 0.3, 0.2, amp=0.3)
    ctl = ScRunCtl(dt_max=1e99, cfl_max=0.5, cfl_desired=0.45, max_steps=40, dt_variable=1)
    rq, _, _ = O.sharpclaw(p, ctl, q0, 0.0, 10.0, 1e-3)
    p.precision = SC_FP32
    with Solver(p) as s:
        s.set_q(q0)
        s.advance(ctl, 0.0, 10.0, 1e-3)
        gq = s.get_q()
    err = rel_l1(interior(p, gq), interior(p, rq))
    print(f"float32 2D lim_type {lim} char_decomp {char}: relative L1 {err:.3e}")
    assert err <= 1e-5


@pytest.mark.parametrize("prec", [0, SC_FP32])
@pytest.mark.parametrize("name,lim,char,mth", [("euler", 2, 1, 5), ("euler", 2, 3, 5), ("euler", 1, 1, 2),
                                               ("euler", 1, 2, 4), ("shallow_water", 2, 1, 5),
                                               ("shallow_water", 2, 3, 5), ("shallow_water", 1, 1, 3),
                                               ("shallow_water", 1, 2, 1)])
@pytest.mark.skip(reason="its first run on the GPU, with a fused-tile variant of these reconstructions "
                         "selected by SC_SWEEPS=0, went silent (killed after 180 s without output) and no "
                         "cause was found in the source or its gfx950 assembly; the variant is not in the "
                         "library, and this case stays off the GPU")
def test_tile_equals_sweeps_bitwise(name, lim, char, mth, prec, monkeypatch):
    """A fused-tile variant of these reconstructions (neighbour data per tile
    row and column in LDS) against the x and y slice sweeps (SC_SWEEPS=1):
    steps and whole runs (CFL, dt, rejections) bit for bit, on grids with
    partial edge tiles and every boundary kind."""
    rng = np.random.default_rng(67 + lim + char)
    for shape, bc in (((50, 47), (1, 3, 2, 2)), ((130, 29), (3, 3, 1, 1)), ((33, 61), (2, 2, 3, 3))):
        p = make_params(name, shape, char_decomp=char, lim_type=lim, mthlim=(mth,) * 4, mthbc=bc,
                        time_integrator=3, precision=prec)
        q = _states.grid(p, rng, smooth=True)
        dt = 0.1 * min(p.dx[0], p.dx[1])
        ctl = ScRunCtl(dt_max=1e99, cfl_max=0.5, cfl_desired=0.45, max_steps=8, dt_variable=1)
        out = []
        for sweeps in ("1", "0"):
            monkeypatch.setenv("SC_SWEEPS", sweeps)
            with Solver(p) as s:
                s.set_q(q)
                dq, cfl = s.step(dt)
                s.set_q(q)
                dtn, st = s.advance(ctl, 0.0, 10.0, dt * 3)
                out.append((dq.tobytes(), cfl, s.get_q().tobytes(), dtn, st.steps_taken, st.rejected,
                            st.cfl_maxused))
        assert out[0] == out[1], (shape, bc)


@pytest.mark.parametrize("env", [{"SC_HOST_LOOP": "0"}, {"SC_HOST_LOOP": "0", "SC_GRAPH": "1"}])
def test_recon2d_ignores_device_loop_request(env, monkeypatch):
    """The slice sweeps have no device-resident loop: a request for one (or
    for its graph replay) runs the host loop, with the same result bit for
    bit."""
    p = make_params("shallow_water", (40, 40), char_decomp=3, time_integrator=3, mthbc=(1, 1, 3, 3))
    q0 = bump(p, 0.2, 0.1)
    ctl = ScRunCtl(dt_max=1e99, cfl_max=0.5, cfl_desired=0.45, max_steps=12, dt_variable=1)
    res = []
    for e in ({}, env):
        for k in ("SC_HOST_LOOP", "SC_GRAPH"):
            monkeypatch.delenv(k, raising=False)
        for k, v in e.items():
            monkeypatch.setenv(k, v)
        with Solver(p) as s:
            s.set_q(q0)
            dt, st = s.advance(ctl, 0.0, 10.0, 1e-3)
            res.append((s.get_q(), dt, st.steps_taken, st.rejected))
    assert np.array_equal(res[0][0], res[1][0])
    assert res[0][1:] == res[1][1:]


def test_recon2d_rejected_where_not_built():
    """2D advection keeps weno5; the server and sc_flux1 on 2D slices keep
    weno5 / weno5_char."""
    for lim, char in RECON:
        with pytest.raises(ScError):
            Solver(make_params("advection", (16, 16), lim_type=lim, char_decomp=char, maux=2))
        p = make_params("euler", (16, 16), lim_type=lim, char_decomp=char)
        q1d = _states.random_cells(p, (22,), np.random.default_rng(3), smooth=True)
        with pytest.raises(ScError):
            S.flux1(p, 2, 16, q1d, 1e-3)
    with Server(hbm_bytes=64 << 20, max_active=2) as srv:
        pr = make_params("shallow_water", (32, 32), lim_type=1, char_decomp=1)
        ctl = ScRunCtl(dt_max=1e99, cfl_max=0.5, cfl_desired=0.45, max_steps=10, dt_variable=1)
        with pytest.raises(ScError):
            srv.submit(pr, ctl, bump(pr, 0.0, 0.0), 0.0, 0.01, 1e-4)
