"""Capacity (mcapa > 0: dtdx(i) = dt/(dx*aux(i,mcapa)), src/flux1.f90:56-61,
in the update and the CFL) and the radial shallow-water source of
apps/shallowWater/2dRadialDam/1dRadialSymm-qwave/src1.f90 on the device,
through the C ABI, against the oracle and the reference's own runs
(tests/golden/run_sw1d_radial_src.npz, run_eu1d_capacity.npz).
Tolerances as test_gpu_parity: steps 1e-12, runs max(1e-10, 8x the
oracle's 1-ulp sensitivity), float32 relative L1 1e-5.
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
G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _bcs(name):
    bcs = [(1, 1, 1, 1), (2, 2, 1, 1), (3, 3, 1, 1), (1, 3, 1, 1)]
    return [b for b in bcs if name != "advection" or 3 not in b[:2]]


@pytest.mark.parametrize("name", ["advection", "acoustics", "euler", "shallow_water"])
@pytest.mark.parametrize("char", [0, 2])
def test_step_capacity(name, char):
    rng = np.random.default_rng(61)
    for bc in _bcs(name):
        for maux, mcapa in ((1, 1), (3, 2)):
            p = make_params(name, (300,), char_decomp=char, mthbc=bc, maux=maux, mcapa=mcapa)
            q = _states.grid(p, rng, smooth=True)
            aux = np.asfortranarray(rng.uniform(0.5, 2.0, (306, maux)))
            dt = 0.05 * p.dx[0]
            with Solver(p) as s:
                s.set_aux(aux)
                s.set_q(q)
                gdq, gc = s.step(dt)
            rq, rdq, rc = O.step(p, q, dt, aux=aux)
            assert rel_linf(interior(p, gdq), interior(p, rdq)) <= TOL, (bc, maux)
            assert abs(gc - rc) <= TOL * rc, (bc, maux)


@pytest.mark.parametrize("char", [0, 2])
@pytest.mark.parametrize("tf", [0, 1])
def test_step_radial_source(char, tf):
    rng = np.random.default_rng(62)
    for bc in [(3, 1, 1, 1), (1, 1, 1, 1)]:
        p = make_params("shallow_water", (300,), xlower=(0.0,), xupper=(2.5,), char_decomp=char,
                        mthbc=bc, src_term=1, tfluct_solver=tf)
        q = _states.grid(p, rng, smooth=True)
        dt = 0.05 * p.dx[0]
        with Solver(p) as s:
            s.set_q(q)
            gdq, gc = s.step(dt)
        rq, rdq, rc = O.step(p, q, dt)
        assert rel_linf(interior(p, gdq), interior(p, rdq)) <= TOL, bc
        assert abs(gc - rc) <= TOL * rc, bc


def _replay_frames(p, fname, aux=None, precision_p=None):
    """Each frame interval through Solver.advance, frames through
    output_begin/wait (the writers' reset of |q| < 1e-99)."""
    d = np.load(os.path.join(G, fname))
    c = d["ctl"]
    ctl = ScRunCtl(dt_max=float(c[0]), cfl_max=float(c[1]), cfl_desired=float(c[2]),
                   max_steps=int(c[3]), dt_variable=int(c[4]))
    q0 = O.new_grid(p)
    interior(p, q0)[:] = d["q"][0]
    frames = []
    with Solver(precision_p or p) as s:
        if aux is not None:
            s.set_aux(aux)
        s.set_q(q0)
        dt, t = float(d["dt_initial"]), float(d["t"][0])
        steps = 0
        for k in range(1, len(d["t"])):
            dt, st = s.advance(ctl, t, float(d["t"][k]), dt)
            assert st.t == float(d["t"][k])
            t = st.t
            steps += st.steps_taken - st.rejected
            s.output_begin()
            frames.append(s.output_wait())
    return d, ctl, q0, frames, steps


def _tol(p, ctl, q0, d, aux=None):
    q, qp = q0.copy(order="F"), q0.copy(order="F")
    interior(p, qp)[:] = np.nextafter(interior(p, qp), np.inf)
    dt = dt2 = float(d["dt_initial"])
    t = float(d["t"][0])
    tols = []
    for k in range(1, len(d["t"])):
        q, dt, _ = O.sharpclaw(p, ctl, q, t, float(d["t"][k]), dt, aux=aux)
        qp, dt2, _ = O.sharpclaw(p, ctl, qp, t, float(d["t"][k]), dt2, aux=aux)
        q, qp = O.flush_small(p, q), O.flush_small(p, qp)
        t = float(d["t"][k])
        assert np.array_equal(interior(p, q), d["q"][k])
        tols.append(max(TOL_RUN, 8.0 * rel_linf(interior(p, qp), interior(p, q))))
    return tols


def test_run_radial_source_matches_reference():
    p = make_params("shallow_water", (400,), xlower=(0.0,), xupper=(2.5,), char_decomp=2,
                    time_integrator=4, mthbc=(3, 1, 1, 1), src_term=1)
    d, ctl, q0, frames, steps = _replay_frames(p, "run_sw1d_radial_src.npz")
    assert steps == len(d["dtlog"])
    for k, (f, tol) in enumerate(zip(frames, _tol(p, ctl, q0, d)), start=1):
        err = rel_linf(interior(p, f), d["q"][k])
        print(f"radial source frame {k}: rel Linf vs reference {err:.3e}, tolerance {tol:.3e}")
        assert err <= tol


def test_run_capacity_matches_reference():
    p = make_params("euler", (400,), xlower=(-5.0,), xupper=(5.0,), char_decomp=2,
                    time_integrator=3, mthbc=(1, 1, 1, 1), maux=2, mcapa=2)
    aux = np.load(os.path.join(G, "run_eu1d_capacity.npz"))["aux"]
    d, ctl, q0, frames, steps = _replay_frames(p, "run_eu1d_capacity.npz", aux=aux)
    assert steps == len(d["dtlog"])
    for k, (f, tol) in enumerate(zip(frames, _tol(p, ctl, q0, d, aux=aux)), start=1):
        err = rel_linf(interior(p, f), d["q"][k])
        print(f"capacity frame {k}: rel Linf vs reference {err:.3e}, tolerance {tol:.3e}")
        assert err <= tol


def test_fp32_capacity_and_source_l1():
    p = make_params("euler", (400,), xlower=(-5.0,), xupper=(5.0,), char_decomp=2,
                    time_integrator=3, mthbc=(1, 1, 1, 1), maux=2, mcapa=2)
    aux = np.load(os.path.join(G, "run_eu1d_capacity.npz"))["aux"]
    p32 = make_params("euler", (400,), xlower=(-5.0,), xupper=(5.0,), char_decomp=2,
                      time_integrator=3, mthbc=(1, 1, 1, 1), maux=2, mcapa=2, precision=SC_FP32)
    d, _, _, frames, _ = _replay_frames(p, "run_eu1d_capacity.npz", aux=aux, precision_p=p32)
    err = np.sum(np.abs(interior(p, frames[-1]) - d["q"][-1])) / np.sum(np.abs(d["q"][-1]))
    print("float32 capacity relative L1:", err)
    assert err <= 1e-5
    p = make_params("shallow_water", (400,), xlower=(0.0,), xupper=(2.5,), char_decomp=2,
                    time_integrator=4, mthbc=(3, 1, 1, 1), src_term=1)
    p32 = make_params("shallow_water", (400,), xlower=(0.0,), xupper=(2.5,), char_decomp=2,
                      time_integrator=4, mthbc=(3, 1, 1, 1), src_term=1, precision=SC_FP32)
    d, _, _, frames, _ = _replay_frames(p, "run_sw1d_radial_src.npz", precision_p=p32)
    err = np.sum(np.abs(interior(p, frames[-1]) - d["q"][-1])) / np.sum(np.abs(d["q"][-1]))
    print("float32 radial source relative L1:", err)
    assert err <= 1e-5


def test_aux_and_source_errors():
    p = make_params("euler", (32,), maux=1, mcapa=1)
    with Solver(p) as s:
        s.set_q(O.new_grid(p))
        with pytest.raises(ScError):
            s.step(1e-3)          # capacity without sc_set_aux
    with Solver(make_params("euler", (32,))) as s:
        with pytest.raises(ScError):
            s.set_aux(np.ones((38, 1), order="F"))   # maux = 0
    for bad in (make_params("euler", (16, 16), maux=1, mcapa=1),
                make_params("euler", (32,), maux=1, mcapa=2),
                make_params("euler", (32,), src_term=1),
                make_params("shallow_water", (32,), src_term=2)):
        with pytest.raises(ScError):
            Solver(bad)
    with pytest.raises(ScError):
        S.flux1(make_params("euler", (32,), maux=1, mcapa=1), 1, 32, np.ones((38, 3)), 1e-3)
    base = make_params("shallow_water", (8,), src_term=1)
    with pytest.raises(ScError):
        Batch(base, np.array([64, 65], dtype=np.int32), np.full(2, 0.1), np.ones(2))
    with Server(hbm_bytes=16 << 20, max_active=2) as srv:
        ctl = ScRunCtl(dt_max=1e99, cfl_max=0.5, cfl_desired=0.45, max_steps=100, dt_variable=1)
        with pytest.raises(ScError):
            srv.submit(make_params("shallow_water", (64,), maux=1, mcapa=1), ctl,
                       dambreak_q(64, 2.0, 0.5, 0.5), 0.0, 0.01, 1e-4)
