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


def _kappa2(rng, nx, ny, maux, mcapa, side=None):
    """Random capacities; with side = 0..3 (left, right, bottom, top) that
    side's ghost column / row gets the smallest ones, so the CFL comes from
    its sweep (flux2.f90:36,68 sweep rows 0..ny+1 and columns 0..nx+1)."""
    a = np.asfortranarray(rng.uniform(0.5, 2.0, (nx + 6, ny + 6, maux)))
    if side is not None:
        k = a[:, :, mcapa - 1]
        k[:] = rng.uniform(1.0, 2.0, k.shape)
        sl = {0: (2, slice(None)), 1: (nx + 3, slice(None)), 2: (slice(None), 2), 3: (slice(None), ny + 3)}[side]
        k[sl] = rng.uniform(0.2, 0.3, k[sl].shape)
    return a


@pytest.mark.parametrize("name", ["euler", "shallow_water"])
@pytest.mark.parametrize("char", [0, 2])
@pytest.mark.parametrize("prec", [0, SC_FP32])
def test_step_capacity_2d(name, char, prec):
    """Capacity on 2D grids (flux2.f90 passes each sweep's aux slice to flux1:
    dtdx(i) = dt/(dx(ixy)*aux(i,mcapa)) in the update and the CFL, the ghost
    rows and columns with their own capacities): one step against the
    oracle, every boundary option, grids with partial edge tiles."""
    rng = np.random.default_rng(71 + char)
    for shape in ((20, 20), (50, 47), (33, 61)):
        for bc in [(1, 1, 1, 1), (2, 2, 2, 2), (3, 3, 3, 3), (1, 3, 2, 2), (2, 2, 3, 1)]:
            for maux, mcapa, side in ((1, 1, None), (3, 2, None), (3, 2, 0), (1, 1, 1), (2, 1, 2), (3, 3, 3)):
                p = make_params(name, shape, char_decomp=char, mthbc=bc, maux=maux, mcapa=mcapa,
                                precision=prec)
                q = _states.grid(p, rng, smooth=True)
                aux = _kappa2(rng, *shape, maux, mcapa, side)
                dt = 0.05 * p.dx[0]
                with Solver(p) as s:
                    s.set_aux(aux)
                    s.set_q(q)
                    gdq, gc = s.step(dt)
                p64 = make_params(name, shape, char_decomp=char, mthbc=bc, maux=maux, mcapa=mcapa)
                rq, rdq, rc = O.step(p64, q, dt, aux=aux)
                if prec == SC_FP32:
                    g, o = interior(p, gdq), interior(p, rdq)
                    assert np.sum(np.abs(g - o)) / np.sum(np.abs(o)) <= 1e-5, (shape, bc)
                    assert abs(gc - rc) <= 1e-5 * rc, (shape, bc)
                else:
                    assert rel_linf(interior(p, gdq), interior(p, rdq)) <= TOL, (shape, bc, maux)
                    assert abs(gc - rc) <= TOL * rc, (shape, bc, maux, gc, rc)


def test_step_capacity_2d_matches_reference():
    """The reference's own steps (tests/golden/step_aux2d.npz: the flang-built
    src/ through its step2d, random capacities)."""
    d = np.load(os.path.join(G, "step_aux2d.npz"))
    for k in range(int(d["n"])):
        p = make_params(str(d[f"c{k}_name"]), (20, 20), char_decomp=int(d[f"c{k}_char"]),
                        mthbc=tuple(int(b) for b in d[f"c{k}_bc"]), maux=int(d[f"c{k}_maux"]),
                        mcapa=int(d[f"c{k}_mcapa"]))
        with Solver(p) as s:
            s.set_aux(d[f"c{k}_aux"])
            s.set_q(d[f"c{k}_q"])
            gdq, gc = s.step(float(d[f"c{k}_dt"]))
            assert np.array_equal(s.get_q(), d[f"c{k}_qbc"]), k
        assert rel_linf(interior(p, gdq), interior(p, d[f"c{k}_dq"])) <= TOL, k
        assert abs(gc - float(d[f"c{k}_cfl"])) <= TOL * gc, k


def test_run_capacity_2d_matches_reference():
    """Four Euler quadrants on a capacity-weighted 32x32 grid, 16 SSPRK3 steps
    (tests/golden/run_eu2d_capacity.npz, the reference program), through
    sc_advance, and a 1024^2 grid of the same kind against the oracle's
    first steps."""
    d = np.load(os.path.join(G, "run_eu2d_capacity.npz"))
    p = make_params("euler", (32, 32), char_decomp=2, time_integrator=3, mthbc=(1, 1, 1, 1), maux=2, mcapa=2)
    c = d["ctl"]
    ctl = ScRunCtl(dt_max=float(c[0]), cfl_max=float(c[1]), cfl_desired=float(c[2]),
                   max_steps=int(c[3]), dt_variable=int(c[4]))
    q0 = O.new_grid(p)
    interior(p, q0)[:] = d["q"][0]
    rq, _, rst = O.sharpclaw(p, ctl, q0, 0.0, 10.0, float(d["dt_initial"]), aux=d["aux"])
    assert np.array_equal(interior(p, rq), d["q"][-1])
    qp = q0.copy(order="F")
    interior(p, qp)[:] = np.nextafter(interior(p, qp), np.inf)
    rq2, _, _ = O.sharpclaw(p, ctl, qp, 0.0, 10.0, float(d["dt_initial"]), aux=d["aux"])
    tol = max(TOL_RUN, 8.0 * rel_linf(interior(p, rq2), interior(p, rq)))
    with Solver(p) as s:
        s.set_aux(d["aux"])
        s.set_q(q0)
        _, gst = s.advance(ctl, 0.0, 10.0, float(d["dt_initial"]))
        gq = s.get_q()
    assert (gst.steps_taken, gst.rejected) == (rst.steps_taken, rst.rejected)
    err = rel_linf(interior(p, gq), d["q"][-1])
    print(f"2D capacity run: rel Linf vs reference {err:.3e}, tolerance {tol:.3e}")
    assert err <= tol
    # 1024^2 (4692 tiles, the last tile row partial): two steps against the oracle
    n = 1024
    rng = np.random.default_rng(73)
    p = make_params("shallow_water", (n, n), char_decomp=2, time_integrator=3, mthbc=(3, 1, 2, 2), maux=1, mcapa=1)
    q0 = _states.grid(p, rng, smooth=True)
    x = (np.arange(-2, n + 4) - 0.5) / n
    X, Y = np.meshgrid(x, x, indexing="ij")
    aux = np.asfortranarray((1.0 + 0.5 * np.sin(3.0 * X + 1.0) ** 2 * np.cos(2.0 * Y) ** 2)[:, :, None])
    ctl = ScRunCtl(dt_max=1e99, cfl_max=0.5, cfl_desired=0.45, max_steps=2, dt_variable=1)
    rq, _, rst = O.sharpclaw(p, ctl, q0, 0.0, 10.0, 1e-5, aux=aux)
    with Solver(p) as s:
        s.set_aux(aux)
        s.set_q(q0)
        _, gst = s.advance(ctl, 0.0, 10.0, 1e-5)
        gq = s.get_q()
    assert gst.steps_taken == rst.steps_taken and gst.cfl_last == pytest.approx(rst.cfl_last, rel=1e-12)
    assert rel_linf(interior(p, gq), interior(p, rq)) <= TOL_RUN


def test_capacity_2d_refusals():
    """2D capacity runs take weno5 / weno5_char and the 16x15 tile, and not
    sc_flux (whose ghost slices would need their capacities)."""
    for bad in (make_params("euler", (16, 16), maux=1, mcapa=1, char_decomp=1),
                make_params("euler", (16, 16), maux=1, mcapa=1, lim_type=1, char_decomp=2),
                make_params("euler", (16, 16), maux=1, mcapa=0)):
        with pytest.raises(ScError):
            Solver(bad)
    p = make_params("euler", (40, 40), maux=1, mcapa=1)
    with Solver(p) as s:
        s.set_aux(np.ones((46, 46, 1), order="F"))
        q = O.new_grid(p)
        interior(p, q)[..., 0] = 1.0
        interior(p, q)[..., 3] = 2.5
        s.set_q(q)
        with pytest.raises(ScError):
            s.set_tile(2)
        s.set_tile(0)
        with pytest.raises(ScError):
            s.flux(q, np.zeros_like(q), 1e-3)


def test_aux_and_source_errors():
    p = make_params("euler", (32,), maux=1, mcapa=1)
    with Solver(p) as s:
        s.set_q(O.new_grid(p))
        with pytest.raises(ScError):
            s.step(1e-3)          # capacity without sc_set_aux
    with Solver(make_params("euler", (32,))) as s:
        with pytest.raises(ScError):
            s.set_aux(np.ones((38, 1), order="F"))   # maux = 0
    for bad in (make_params("euler", (16, 16), maux=1, mcapa=1, char_decomp=3),
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
