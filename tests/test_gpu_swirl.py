"""2D advection with edge velocities in aux and a time-dependent b4step
(apps/advection/2d/swirl: rpn2ad1.f, setaux.f, psi.f, b4step.f90) on the
device, through the C ABI: steps against the oracle, the app's own runs
against the reference's output (tests/golden/run_swirl2d*.npz), rejected
steps keeping the velocity factor of their first attempt (sharpclaw.f90:
127-131 call b4step only for a new step).  Tolerances as test_gpu_parity.
"""
import os

import numpy as np
import pytest

import _states
from conftest import rel_linf
from oracle import oracle as O
from sharpclaw_amd import SC_FP32, ScError, ScRunCtl, Solver, interior, make_params

pytestmark = pytest.mark.gpu

TOL = 1e-12
TOL_RUN = 1e-10
G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.mark.parametrize("shape", [(24, 24), (19, 23), (70, 33)])
@pytest.mark.parametrize("tile", [0, 1, 2, 3])
def test_step_swirl(shape, tile):
    rng = np.random.default_rng(71)
    for bc in [(1, 1, 1, 1), (2, 2, 2, 2), (2, 2, 1, 1)]:
        p = make_params("advection", shape, char_decomp=0, mthbc=bc, maux=2)
        q = _states.grid(p, rng, smooth=True)
        aux = O.swirl_setaux(p)
        dt = 0.02 * p.dx[0]
        with Solver(p) as s:
            s.set_tile(tile)
            s.set_aux(aux)
            s.set_q(q)
            gdq, gc = s.step(dt)
        rq, rdq, rc = O.step(p, q, dt, aux=aux)
        assert rel_linf(interior(p, gdq), interior(p, rdq)) <= TOL, bc
        assert abs(gc - rc) <= TOL * rc, bc


def _frames(p, d, aux, solver_p=None):
    c = d["ctl"]
    ctl = ScRunCtl(dt_max=float(c[0]), cfl_max=float(c[1]), cfl_desired=float(c[2]),
                   max_steps=int(c[3]), dt_variable=int(c[4]))
    q0 = O.new_grid(p)
    interior(p, q0)[:] = d["q"][0]
    out = []
    with Solver(solver_p or p) as s:
        s.set_aux(aux)
        s.set_q(q0)
        t, dt = float(d["t"][0]), float(d["dt_initial"])
        for k in range(1, len(d["t"])):
            dt, st = s.advance(ctl, t, float(d["t"][k]), dt)
            t = st.t
            s.output_begin()
            out.append(s.output_wait())
    return ctl, q0, out


@pytest.mark.parametrize("fname", ["run_swirl2d.npz", "run_swirl2d_tperiod.npz"])
def test_run_swirl_matches_reference(fname):
    d = np.load(os.path.join(G, fname))
    p = make_params("advection", (50, 50), char_decomp=0, mthbc=(1, 1, 1, 1), maux=2,
                    time_integrator=4, tperiod=float(d["tperiod"]))
    aux = O.swirl_setaux(p)
    ctl, q0, frames = _frames(p, d, aux)
    # whole-run tolerance from the oracle's own 1-ulp sensitivity
    q, qp = q0.copy(order="F"), q0.copy(order="F")
    interior(p, qp)[:] = np.nextafter(interior(p, qp), np.inf)
    t = float(d["t"][0])
    dt = dt2 = float(d["dt_initial"])
    for k, f in enumerate(frames, start=1):
        q, dt, _ = O.sharpclaw(p, ctl, q, t, float(d["t"][k]), dt, aux=aux)
        qp, dt2, _ = O.sharpclaw(p, ctl, qp, t, float(d["t"][k]), dt2, aux=aux)
        q, qp = O.flush_small(p, q), O.flush_small(p, qp)
        t = float(d["t"][k])
        tol = max(TOL_RUN, 8.0 * rel_linf(interior(p, qp), interior(p, q)))
        err = rel_linf(interior(p, f), d["q"][k])
        print(f"{fname} frame {k}: rel Linf vs reference {err:.3e}, tolerance {tol:.3e}")
        assert err <= tol


def test_retake_keeps_the_velocity_factor():
    """dt_initial far above the CFL limit: the first attempts are rejected
    and retaken with a smaller dt but the b4step factor of the first attempt
    (the reference calls b4step only when not retaking)."""
    p = make_params("advection", (40, 40), char_decomp=0, mthbc=(1, 1, 1, 1), maux=2,
                    time_integrator=3, tperiod=0.7)
    aux = O.swirl_setaux(p)
    q0 = _states.grid(p, np.random.default_rng(72), smooth=True)
    ctl = ScRunCtl(dt_max=1e99, cfl_max=0.5, cfl_desired=0.45, max_steps=5000, dt_variable=1)
    rq, rdt, rst = O.sharpclaw(p, ctl, q0, 0.0, 0.3, 0.2, aux=aux)
    assert rst.rejected > 0
    with Solver(p) as s:
        s.set_aux(aux)
        s.set_q(q0)
        gdt, gst = s.advance(ctl, 0.0, 0.3, 0.2)
        gq = s.get_q()
    assert (gst.steps_taken, gst.rejected) == (rst.steps_taken, rst.rejected)
    assert rel_linf(interior(p, gq), interior(p, rq)) <= TOL_RUN
    # the same through the step-by-step entry points (the Fortran driver's
    # loop): save_qold starts a step, restore_qold + ts retakes it
    with Solver(p) as s:
        s.set_aux(aux)
        s.set_q(q0)
        t, dt, rejected = 0.0, 0.2, 0
        while t < 0.3:
            told = t
            dt = min(dt, 0.3 - told)
            s.save_qold()
            while True:
                cfl = s.ts(dt, told)
                dtn = min(ctl.dt_max, dt * ctl.cfl_desired / cfl)
                if cfl <= ctl.cfl_max:
                    t = told + dt
                    dt = dtn
                    break
                s.restore_qold()
                rejected += 1
                dt = dtn
        hq = s.get_q()
    assert rejected == rst.rejected
    assert rel_linf(interior(p, hq), interior(p, rq)) <= TOL_RUN


def test_fp32_swirl_l1():
    d = np.load(os.path.join(G, "run_swirl2d_tperiod.npz"))
    p = make_params("advection", (50, 50), char_decomp=0, mthbc=(1, 1, 1, 1), maux=2,
                    time_integrator=4, tperiod=1.5)
    p32 = make_params("advection", (50, 50), char_decomp=0, mthbc=(1, 1, 1, 1), maux=2,
                      time_integrator=4, tperiod=1.5, precision=SC_FP32)
    _, _, frames = _frames(p, d, O.swirl_setaux(p), solver_p=p32)
    err = np.sum(np.abs(interior(p, frames[-1]) - d["q"][-1])) / np.sum(np.abs(d["q"][-1]))
    print("float32 swirl relative L1:", err)
    assert err <= 1e-5


def test_swirl_errors():
    p = make_params("advection", (16, 16), maux=2, char_decomp=0)
    with Solver(p) as s:
        s.set_q(O.new_grid(p))
        with pytest.raises(ScError):
            s.step(1e-3)          # velocities not uploaded
    for bad in (make_params("advection", (16, 16), maux=1, char_decomp=0),
                make_params("advection", (16, 16), maux=2, char_decomp=2),
                make_params("advection", (16, 16), maux=2, mcapa=1, char_decomp=0),
                make_params("advection", (16, 16), maux=2, char_decomp=0, mthbc=(3, 3, 1, 1))):
        with pytest.raises(ScError):
            Solver(bad)
