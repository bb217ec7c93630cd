"""Writes the golden fixtures of tests/golden/ from the reference itself.

The reference SharpClaw (src/ compiled unmodified with flang -O3 by
oracle/ref/Makefile, plus the project-written app hooks in
oracle/ref/hooks/) is run through oracle/_ref/harness_* (single routines)
and oracle/_ref/xsclaw_* (the reference program src/main.f90) on seeded
inputs.  Each fixture stores the inputs and the reference's outputs.

    make -C oracle ref && python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

import _states  # noqa: E402
from oracle import oracle as O  # noqa: E402  (input data only: the swirl aux arrays)
from oracle import refrun as R  # noqa: E402
from sharpclaw_amd import interior, make_params  # noqa: E402

PHYS = [("advection", 1), ("acoustics", 1), ("euler", 1), ("shallow_water", 1), ("euler", 2),
        ("shallow_water", 2)]


def tag(name, ndim):
    return f"{name}{ndim}d"


def save(fname, **arrays):
    np.savez_compressed(os.path.join(HERE, fname), **arrays)
    print("wrote", fname, len(arrays), "arrays")


def routines():
    rng = np.random.default_rng(20260101)
    for name, ndim in PHYS:
        p = make_params(name, (24,) if ndim == 1 else (24, 24), char_decomp=2)
        mx, L = 24, 30
        out = {}
        for ixy in ([1] if ndim == 1 else [1, 2]):
            ql = _states.random_cells(p, (L,), rng)
            qr = _states.random_cells(p, (L,), rng)
            w, s, am, ap = R.rp(p, ixy, mx, ql, qr)
            out.update({f"rp{ixy}_ql": ql, f"rp{ixy}_qr": qr, f"rp{ixy}_wave": w, f"rp{ixy}_s": s,
                        f"rp{ixy}_amdq": am, f"rp{ixy}_apdq": ap})
            q = _states.random_cells(p, (L,), rng)
            evl, evr = R.evec(p, ixy, mx, q)
            out.update({f"evec{ixy}_q": q, f"evec{ixy}_evl": evl, f"evec{ixy}_evr": evr})
        q = _states.random_cells(p, (L,), rng, smooth=True)
        evl, evr = R.evec(p, 1, mx, q)
        ql, qr = R.weno(p, q)
        qlc, qrc = R.weno(p, q, evl, evr)
        out.update({"weno_q": q, "weno_ql": ql, "weno_qr": qr, "wenoc_evl": evl, "wenoc_evr": evr,
                    "wenoc_ql": qlc, "wenoc_qr": qrc})
        save(f"routines_{tag(name, ndim)}.npz", **out)


def steps():
    rng = np.random.default_rng(20260102)
    for name, ndim in PHYS:
        out = {}
        bcs = [(1, 1, 1, 1), (2, 2, 2, 2), (3, 3, 3, 3), (1, 3, 2, 2)]
        if name == "advection":
            bcs = [b for b in bcs if 3 not in b[:2]]
        k = 0
        for bc in bcs:
            for char in (0, 2):
                p = make_params(name, (40,) if ndim == 1 else (12, 12), char_decomp=char, mthbc=bc)
                q = _states.grid(p, rng, smooth=True)
                dt = 0.1 * p.dx[0]
                qb, dq, cfl = R.step(p, q, dt)
                out.update({f"c{k}_q": q, f"c{k}_qbc": qb, f"c{k}_dq": dq,
                            f"c{k}_cfl": np.array(cfl), f"c{k}_dt": np.array(dt),
                            f"c{k}_bc": np.array(bc), f"c{k}_char": np.array(char)})
                k += 1
        out["n"] = np.array(k)
        save(f"step_{tag(name, ndim)}.npz", **out)


def runs():
    # C1: 1D advection, component-wise WENO5, SSPRK3, CFL 0.5, to t = 1
    p = make_params("advection", (500,), char_decomp=0, time_integrator=3, mthbc=(2, 2, 2, 2))
    cd = R.clawdata_for(p, nout=2, outstyle=1, tfinal=1.0, dt_initial=0.001, cfl_max=0.55,
                        cfl_desired=0.5)
    fr, dtl, info = R.run(p, cd, ["1", "200.0 0.3 0.6 0.8"])
    save("run_c1_advection.npz", t=np.array([f[0] for f in fr]), q=np.stack([f[1] for f in fr]),
         dtlog=dtl, ctl=np.array([1e99, 0.55, 0.5, 50000, 1]), dt_initial=np.array(0.001))
    # C2: 1D Euler Shu-Osher, char WENO5 + Roe, SSP(10,4), to t = 0.36 (two frames)
    p = make_params("euler", (1600,), xlower=(-5.0,), xupper=(5.0,), char_decomp=2,
                    time_integrator=4, mthbc=(1, 1, 1, 1))
    cd = R.clawdata_for(p, nout=2, outstyle=1, tfinal=0.36, dt_initial=0.001, cfl_max=2.5,
                        cfl_desired=2.45)
    fr, dtl, info = R.run(p, cd, ["2", "3.857143 2.629369 10.333333 -4.0"])
    save("run_c2_euler.npz", t=np.array([f[0] for f in fr]), q=np.stack([f[1] for f in fr]),
         dtlog=dtl, ctl=np.array([1e99, 2.5, 2.45, 50000, 1]), dt_initial=np.array(0.001))
    # C3 reduced: 2D Euler four quadrants, char WENO5 + Roe rpn2, SSPRK3, 32x32, 12 steps
    p = make_params("euler", (32, 32), char_decomp=2, time_integrator=3, mthbc=(1, 1, 1, 1))
    rng = np.random.default_rng(7)
    st = np.stack([rng.uniform(0.1, 1.5, 4), rng.uniform(-1, 1, 4), rng.uniform(-1, 1, 4),
                   rng.uniform(0.03, 1.5, 4)], axis=1)
    lines = ["1", "0.5 0.5"] + [" ".join(repr(float(v)) for v in row) for row in st]
    cd = R.clawdata_for(p, nout=12, outstyle=3, nstepout=12, nstop=12, dt_initial=1e-3,
                        cfl_max=0.5, cfl_desired=0.45)
    fr, dtl, info = R.run(p, cd, lines)
    save("run_c3_euler2d.npz", t=np.array([f[0] for f in fr]), q=np.stack([f[1] for f in fr]),
         dtlog=dtl, states=st, ctl=np.array([1e99, 0.5, 0.45, 12, 1]), dt_initial=np.array(1e-3))
    # 1D shallow-water dam break, char WENO5 + Roe (efix), SSPRK3, to t = 0.05
    p = make_params("shallow_water", (200,), char_decomp=2, time_integrator=3, mthbc=(1, 1, 1, 1))
    cd = R.clawdata_for(p, nout=1, outstyle=1, tfinal=0.05, dt_initial=1e-4, cfl_max=0.5,
                        cfl_desired=0.45)
    fr, dtl, info = R.run(p, cd, ["3", "2.0 0.5 0.45"])
    save("run_sw1d_dambreak.npz", t=np.array([f[0] for f in fr]), q=np.stack([f[1] for f in fr]),
         dtlog=dtl, ctl=np.array([1e99, 0.5, 0.45, 50000, 1]), dt_initial=np.array(1e-4))
    # 1D acoustics pulse, walls, char WENO5, SSP(10,4), to t = 0.25
    p = make_params("acoustics", (200,), xlower=(-1.0,), xupper=(1.0,), char_decomp=2,
                    time_integrator=4, mthbc=(3, 3, 1, 1))
    cd = R.clawdata_for(p, nout=1, outstyle=1, tfinal=0.25, dt_initial=1e-3, cfl_max=2.5,
                        cfl_desired=2.4)
    fr, dtl, info = R.run(p, cd, ["4", "1.0 100.0 0.0"])
    save("run_ac1d_pulse.npz", t=np.array([f[0] for f in fr]), q=np.stack([f[1] for f in fr]),
         dtlog=dtl, ctl=np.array([1e99, 2.5, 2.4, 50000, 1]), dt_initial=np.array(1e-3))
    # 2D shallow water pulse, periodic, char WENO5, SSPRK3, 24x24, to t = 0.02
    p = make_params("shallow_water", (24, 24), char_decomp=2, time_integrator=3,
                    mthbc=(2, 2, 2, 2))
    cd = R.clawdata_for(p, nout=1, outstyle=1, tfinal=0.02, dt_initial=1e-4, cfl_max=0.5,
                        cfl_desired=0.45)
    fr, dtl, info = R.run(p, cd, ["2", "0.3 60.0 0.4 0.55"])
    save("run_sw2d_pulse.npz", t=np.array([f[0] for f in fr]), q=np.stack([f[1] for f in fr]),
         dtlog=dtl, ctl=np.array([1e99, 0.5, 0.45, 50000, 1]), dt_initial=np.array(1e-4))


def tfluct():
    """tfluct_solver = 1 (flux1.f90:136-153) with the applications' own
    tfluct routines linked in place of the library stub: one step per case,
    and two full runs."""
    rng = np.random.default_rng(20261203)
    out = {}
    k = 0
    for name in ("advection", "acoustics", "shallow_water"):
        for bc in [(1, 1, 1, 1), (2, 2, 2, 2), (3, 3, 3, 3), (1, 3, 1, 1)]:
            if name == "advection" and 3 in bc[:2]:
                continue
            for char in (0, 2):
                p = make_params(name, (40,), char_decomp=char, mthbc=bc, tfluct_solver=1)
                q = _states.grid(p, rng, smooth=True)
                dt = 0.1 * p.dx[0]
                qb, dq, cfl = R.step(p, q, dt)
                out.update({f"c{k}_name": np.array(name), f"c{k}_q": q, f"c{k}_qbc": qb,
                            f"c{k}_dq": dq, f"c{k}_cfl": np.array(cfl), f"c{k}_dt": np.array(dt),
                            f"c{k}_bc": np.array(bc), f"c{k}_char": np.array(char)})
                k += 1
    out["n"] = np.array(k)
    save("step_tfluct1d.npz", **out)
    # 1D shallow-water dam break as run_sw1d_dambreak, total-fluctuation solver
    p = make_params("shallow_water", (200,), char_decomp=2, time_integrator=3, mthbc=(1, 1, 1, 1),
                    tfluct_solver=1)
    cd = R.clawdata_for(p, nout=1, outstyle=1, tfinal=0.05, dt_initial=1e-4, cfl_max=0.5,
                        cfl_desired=0.45)
    fr, dtl, info = R.run(p, cd, ["3", "2.0 0.5 0.45"])
    save("run_sw1d_dambreak_tfluct.npz", t=np.array([f[0] for f in fr]),
         q=np.stack([f[1] for f in fr]), dtlog=dtl, ctl=np.array([1e99, 0.5, 0.45, 50000, 1]),
         dt_initial=np.array(1e-4))
    # 1D acoustics pulse as run_ac1d_pulse, total-fluctuation solver
    p = make_params("acoustics", (200,), xlower=(-1.0,), xupper=(1.0,), char_decomp=2,
                    time_integrator=4, mthbc=(3, 3, 1, 1), tfluct_solver=1)
    cd = R.clawdata_for(p, nout=1, outstyle=1, tfinal=0.25, dt_initial=1e-3, cfl_max=2.5,
                        cfl_desired=2.4)
    fr, dtl, info = R.run(p, cd, ["4", "1.0 100.0 0.0"])
    save("run_ac1d_pulse_tfluct.npz", t=np.array([f[0] for f in fr]),
         q=np.stack([f[1] for f in fr]), dtlog=dtl, ctl=np.array([1e99, 2.5, 2.4, 50000, 1]),
         dt_initial=np.array(1e-3))


def flush():
    """A run whose frames reset values below 1e-99 (out1.f90:63-68): a narrow
    Gaussian (tails down to 1e-218) advected through four frames."""
    p = make_params("advection", (200,), char_decomp=0, time_integrator=3, mthbc=(2, 2, 2, 2))
    cd = R.clawdata_for(p, nout=4, outstyle=1, tfinal=0.2, dt_initial=1e-3, cfl_max=0.55,
                        cfl_desired=0.5)
    fr, dtl, info = R.run(p, cd, ["1", "2000.0 0.5 2.0 2.0"])
    save("run_adv1d_flush.npz", t=np.array([f[0] for f in fr]), q=np.stack([f[1] for f in fr]),
         dtlog=dtl, ctl=np.array([1e99, 0.55, 0.5, 50000, 1]), dt_initial=np.array(1e-3))


RECON = [(2, 1), (2, 3), (1, 1), (1, 2)]   # (lim_type, char_decomp) beyond weno5/weno5_char


def recon():
    """The other reconstructions of flux1.f90:65-117: weno5_wave, weno5_trans,
    tvd2_wave, tvd2_char (every limiter), one step per case in 1D, and a
    run of each but tvd2_char."""
    rng = np.random.default_rng(20261204)
    out = {}
    k = 0
    for name in ("advection", "acoustics", "euler", "shallow_water"):
        for lim, char in RECON:
            for mth in ((1, 2, 3, 4, 5) if lim == 1 else (5,)):
                bcs = [(1, 1, 1, 1), (2, 2, 2, 2), (3, 3, 3, 3), (1, 3, 1, 1)]
                if name == "advection":
                    bcs = bcs[:2]
                # every boundary kind with one limiter, one kind each with the others
                for bc in (bcs if mth == 5 else [bcs[mth % len(bcs)]]):
                    p = make_params(name, (40,), char_decomp=char, mthbc=bc, lim_type=lim,
                                    mthlim=(mth,) * 4)
                    q = _states.grid(p, rng, smooth=True)
                    dt = 0.1 * p.dx[0]
                    qb, dq, cfl = R.step(p, q, dt)
                    out.update({f"c{k}_name": np.array(name), f"c{k}_q": q, f"c{k}_qbc": qb,
                                f"c{k}_dq": dq, f"c{k}_cfl": np.array(cfl), f"c{k}_dt": np.array(dt),
                                f"c{k}_bc": np.array(bc), f"c{k}_char": np.array(char),
                                f"c{k}_lim": np.array(lim), f"c{k}_mthlim": np.array(mth)})
                    k += 1
    out["n"] = np.array(k)
    save("step_recon1d.npz", **out)
    runs_ = [
        # (no tvd2_char run: the reference's tvd2_char writes up to three
        # elements past its workspace (reconstruct.f90:598-606 index dq, hh,
        # u to mx2-1 = maxnx+5 while alloc_recon_workspace sizes them to
        # maxnx+mbc), and the program aborts in free() at its end)
        # Shu-Osher, weno5_trans, SSP(10,4)
        ("run_eu1d_shuosher_trans.npz", "euler", (400,), (-5.0,), (5.0,), 3, 2, 5, 4,
         (1, 1, 1, 1), 0.36, 1e-3, 2.5, 2.45, ["2", "3.857143 2.629369 10.333333 -4.0"]),
        # acoustics pulse between walls, weno5_wave
        ("run_ac1d_pulse_wave.npz", "acoustics", (200,), (-1.0,), (1.0,), 1, 2, 5, 4,
         (3, 3, 1, 1), 0.25, 1e-3, 2.5, 2.4, ["4", "1.0 100.0 0.0"]),
        # advection, tvd2_wave with superbee, periodic, SSPRK3
        ("run_adv1d_tvd2wave.npz", "advection", (200,), (0.0,), (1.0,), 1, 1, 2, 3,
         (2, 2, 2, 2), 0.5, 1e-3, 0.55, 0.5, ["1", "200.0 0.3 0.6 0.8"]),
    ]
    for (fname, name, nx, xl, xu, char, lim, mth, ti, bc, tf, dt0, cmax, cdes, lines) in runs_:
        p = make_params(name, nx, xlower=xl, xupper=xu, char_decomp=char, lim_type=lim,
                        mthlim=(mth,) * 4, time_integrator=ti, mthbc=bc)
        cd = R.clawdata_for(p, nout=1, outstyle=1, tfinal=tf, dt_initial=dt0, cfl_max=cmax,
                            cfl_desired=cdes)
        fr, dtl, info = R.run(p, cd, lines)
        save(fname, t=np.array([f[0] for f in fr]), q=np.stack([f[1] for f in fr]), dtlog=dtl,
             ctl=np.array([1e99, cmax, cdes, 50000, 1]), dt_initial=np.array(dt0),
             params=np.array([char, lim, mth, ti] + list(bc)), xl=np.array(xl), xu=np.array(xu))


def recon2d():
    """The reconstructions of recon() on 2D grids (flux2.f90 calls flux1 per
    row and column): Euler and shallow water, one step per case, and a run
    of each physics.  Square grids only: on a grid with nx != ny the
    reference's rp and evec index the slice with the leading dimension of
    the larger direction (maxnx) and mis-index.  No 2D Euler tvd2_char: with
    meqn 4 > mwaves 3 it reads mthlim(4) past the end of mthlim.  Runs with
    the shallow-water pulse and four Euler quadrants."""
    rng = np.random.default_rng(20261207)
    out = {}
    k = 0
    for name in ("euler", "shallow_water"):
        for lim, char in RECON:
            if name == "euler" and (lim, char) == (1, 2):
                continue
            for mth in ((1, 2, 3, 4, 5) if lim == 1 else (5,)):
                bcs = [(1, 1, 1, 1), (2, 2, 2, 2), (3, 3, 3, 3), (1, 3, 2, 2)]
                for bc in (bcs if mth == 5 else [bcs[mth % len(bcs)]]):
                    p = make_params(name, (12, 12), char_decomp=char, mthbc=bc, lim_type=lim,
                                    mthlim=(mth,) * 4)
                    q = _states.grid(p, rng, smooth=True)
                    dt = 0.1 * p.dx[0]
                    qb, dq, cfl = R.step(p, q, dt)
                    out.update({f"c{k}_name": np.array(name), f"c{k}_q": q, f"c{k}_qbc": qb,
                                f"c{k}_dq": dq, f"c{k}_cfl": np.array(cfl), f"c{k}_dt": np.array(dt),
                                f"c{k}_bc": np.array(bc), f"c{k}_char": np.array(char),
                                f"c{k}_lim": np.array(lim), f"c{k}_mthlim": np.array(mth)})
                    k += 1
    out["n"] = np.array(k)
    save("step_recon2d.npz", **out)
    rng = np.random.default_rng(8)
    st = np.stack([rng.uniform(0.5, 1.5, 4), rng.uniform(-0.5, 0.5, 4), rng.uniform(-0.5, 0.5, 4),
                   rng.uniform(0.5, 1.5, 4)], axis=1)
    quad = ["1", "0.5 0.5"] + [" ".join(repr(float(v)) for v in row) for row in st]
    runs_ = [
        # shallow-water pulse as run_sw2d_pulse, weno5_wave, SSPRK3
        ("run_sw2d_pulse_wave.npz", "shallow_water", (24, 24), 1, 2, 5, 3, (2, 2, 2, 2), 0.06,
         1e-4, 0.5, 0.45, ["2", "0.3 60.0 0.4 0.55"]),
        # (no tvd2_char run: the program aborts in free() after one, as in 1D)
        # four quadrants (moderate jumps), weno5_trans, SSPRK3
        ("run_eu2d_quad_trans.npz", "euler", (24, 24), 3, 2, 5, 3, (1, 1, 1, 1), 0.1, 1e-3, 0.5,
         0.45, quad),
        # four quadrants, tvd2_wave with van Leer, SSPRK3
        ("run_eu2d_quad_tvd2wave.npz", "euler", (24, 24), 1, 1, 3, 3, (1, 1, 1, 1), 0.1, 1e-3,
         0.5, 0.45, quad),
    ]
    for (fname, name, nx, char, lim, mth, ti, bc, tf, dt0, cmax, cdes, lines) in runs_:
        xl, xu = (0.0, 0.0), (1.0, 1.0)
        p = make_params(name, nx, xlower=xl, xupper=xu, char_decomp=char, lim_type=lim,
                        mthlim=(mth,) * 4, time_integrator=ti, mthbc=bc)
        cd = R.clawdata_for(p, nout=1, outstyle=1, tfinal=tf, dt_initial=dt0, cfl_max=cmax,
                            cfl_desired=cdes)
        fr, dtl, info = R.run(p, cd, lines)
        save(fname, t=np.array([f[0] for f in fr]), q=np.stack([f[1] for f in fr]), dtlog=dtl,
             ctl=np.array([1e99, cmax, cdes, 50000, 1]), dt_initial=np.array(dt0),
             params=np.array([char, lim, mth, ti] + list(bc)), xl=np.array(xl), xu=np.array(xu))


def kappa_aux(mx, maux, mcapa, xl, xu):
    """A smooth capacity profile in column mcapa of aux(1-mbc:mx+mbc, maux)."""
    dx = (xu - xl) / mx
    x = xl + (np.arange(-2, mx + 4) - 0.5) * dx
    aux = np.zeros((mx + 6, maux), order="F")
    for m in range(maux):
        aux[:, m] = 1.0 + 0.1 * m
    aux[:, mcapa - 1] = 1.0 + 0.5 * np.sin(2.0 * np.pi * (x - xl) / (xu - xl)) ** 2
    return aux


def aux():
    """Capacity (mcapa > 0: dtdx(i) = dt/(dx*aux(i,mcapa)), flux1.f90:56-61)
    and the radial shallow-water source of 1dRadialSymm-qwave (src1.f90),
    one step per case and a run of each."""
    rng = np.random.default_rng(20261205)
    out = {}
    k = 0
    for name in ("advection", "acoustics", "euler", "shallow_water"):
        for char in (0, 2):
            for bc in [(1, 1, 1, 1), (2, 2, 2, 2), (3, 3, 3, 3)]:
                if name == "advection" and bc[0] == 3:
                    continue
                for maux, mcapa in ((1, 1), (3, 2)):
                    p = make_params(name, (40,), char_decomp=char, mthbc=bc, maux=maux, mcapa=mcapa)
                    q = _states.grid(p, rng, smooth=True)
                    a = rng.uniform(0.5, 2.0, (46, maux))
                    dt = 0.05 * p.dx[0]
                    qb, dq, cfl = R.step(p, q, dt, aux=a)
                    out.update({f"c{k}_name": np.array(name), f"c{k}_q": q, f"c{k}_aux": a,
                                f"c{k}_qbc": qb, f"c{k}_dq": dq, f"c{k}_cfl": np.array(cfl),
                                f"c{k}_dt": np.array(dt), f"c{k}_bc": np.array(bc),
                                f"c{k}_char": np.array(char), f"c{k}_maux": np.array(maux),
                                f"c{k}_mcapa": np.array(mcapa), f"c{k}_src": np.array(0),
                                f"c{k}_tfluct": np.array(0)})
                    k += 1
    for char in (0, 2):
        for bc in [(3, 1, 1, 1), (1, 1, 1, 1)]:
            for tf in (0, 1):
                p = make_params("shallow_water", (40,), xlower=(0.0,), xupper=(2.5,), char_decomp=char,
                                mthbc=bc, src_term=1, tfluct_solver=tf)
                q = _states.grid(p, rng, smooth=True)
                dt = 0.05 * p.dx[0]
                qb, dq, cfl = R.step(p, q, dt)
                out.update({f"c{k}_name": np.array("shallow_water"), f"c{k}_q": q,
                            f"c{k}_aux": np.zeros((46, 1)), f"c{k}_qbc": qb, f"c{k}_dq": dq,
                            f"c{k}_cfl": np.array(cfl), f"c{k}_dt": np.array(dt),
                            f"c{k}_bc": np.array(bc), f"c{k}_char": np.array(char),
                            f"c{k}_maux": np.array(0), f"c{k}_mcapa": np.array(0),
                            f"c{k}_src": np.array(1), f"c{k}_tfluct": np.array(tf)})
                k += 1
    out["n"] = np.array(k)
    save("step_aux1d.npz", **out)
    # radial dam break (1dRadialSymm-qwave: rp1sw + src1, wall at r = 0), SSP(10,4)
    p = make_params("shallow_water", (400,), xlower=(0.0,), xupper=(2.5,), char_decomp=2,
                    time_integrator=4, mthbc=(3, 1, 1, 1), src_term=1)
    cd = R.clawdata_for(p, nout=2, outstyle=1, tfinal=0.3, dt_initial=1e-4, cfl_max=1.0,
                        cfl_desired=0.9)
    fr, dtl, info = R.run(p, cd, ["3", "2.0 0.5 0.5"])
    save("run_sw1d_radial_src.npz", t=np.array([f[0] for f in fr]), q=np.stack([f[1] for f in fr]),
         dtlog=dtl, ctl=np.array([1e99, 1.0, 0.9, 50000, 1]), dt_initial=np.array(1e-4))
    # Shu-Osher on a capacity-weighted grid (maux 2, mcapa 2), SSPRK3
    p = make_params("euler", (400,), xlower=(-5.0,), xupper=(5.0,), char_decomp=2,
                    time_integrator=3, mthbc=(1, 1, 1, 1), maux=2, mcapa=2)
    a = kappa_aux(400, 2, 2, -5.0, 5.0)
    cd = R.clawdata_for(p, nout=2, outstyle=1, tfinal=0.2, dt_initial=1e-3, cfl_max=0.5,
                        cfl_desired=0.45)
    fr, dtl, info = R.run(p, cd, ["2", "3.857143 2.629369 10.333333 -4.0"], aux=a)
    save("run_eu1d_capacity.npz", t=np.array([f[0] for f in fr]), q=np.stack([f[1] for f in fr]),
         dtlog=dtl, ctl=np.array([1e99, 0.5, 0.45, 50000, 1]), dt_initial=np.array(1e-3), aux=a)


def kappa_aux2(nx, ny, maux, mcapa):
    """A smooth capacity field in aux(:,:,mcapa) of aux(1-mbc:nx+mbc, 1-mbc:ny+mbc, maux)."""
    x = (np.arange(-2, nx + 4) - 0.5) / nx
    y = (np.arange(-2, ny + 4) - 0.5) / ny
    X, Y = np.meshgrid(x, y, indexing="ij")
    aux = np.zeros((nx + 6, ny + 6, maux), order="F")
    for m in range(maux):
        aux[:, :, m] = 1.0 + 0.1 * m
    aux[:, :, mcapa - 1] = 1.0 + 0.4 * np.sin(2.0 * np.pi * X) ** 2 * np.cos(np.pi * Y) ** 2
    return aux


def aux2d():
    """Capacity on 2D grids (flux2.f90 passes the slice aux(:,j,:) /
    aux(i,:,:) into flux1, whose dtdx(i) = dt/(dx(ixy)*aux(i,mcapa)) scales
    the slice's update and CFL, flux1.f90:56-61): Euler and shallow water,
    one step per case on random capacities, and a four-quadrant Euler run.
    Square grids (see recon2d)."""
    rng = np.random.default_rng(20261208)
    out = {}
    k = 0
    for name in ("euler", "shallow_water"):
        for char in (0, 2):
            for bc in [(1, 1, 1, 1), (2, 2, 2, 2), (3, 3, 3, 3), (1, 3, 2, 2)]:
                for maux, mcapa in ((1, 1), (3, 2)):
                    p = make_params(name, (20, 20), char_decomp=char, mthbc=bc, maux=maux, mcapa=mcapa)
                    q = _states.grid(p, rng, smooth=True)
                    a = np.asfortranarray(rng.uniform(0.5, 2.0, (26, 26, maux)))
                    if maux == 3:
                        # the smallest capacities on one side's ghost column /
                        # row: the CFL comes from that sweep
                        kap = a[:, :, mcapa - 1]
                        kap[:] = rng.uniform(1.0, 2.0, kap.shape)
                        side = (k // 2) % 4
                        sl = {0: (2, slice(None)), 1: (23, slice(None)), 2: (slice(None), 2),
                              3: (slice(None), 23)}[side]
                        kap[sl] = rng.uniform(0.2, 0.3, kap[sl].shape)
                    dt = 0.05 * p.dx[0]
                    qb, dq, cfl = R.step(p, q, dt, aux=a)
                    out.update({f"c{k}_name": np.array(name), f"c{k}_q": q, f"c{k}_aux": a,
                                f"c{k}_qbc": qb, f"c{k}_dq": dq, f"c{k}_cfl": np.array(cfl),
                                f"c{k}_dt": np.array(dt), f"c{k}_bc": np.array(bc),
                                f"c{k}_char": np.array(char), f"c{k}_maux": np.array(maux),
                                f"c{k}_mcapa": np.array(mcapa)})
                    k += 1
    out["n"] = np.array(k)
    save("step_aux2d.npz", **out)
    # four Euler quadrants on a capacity-weighted 32x32 grid, SSPRK3, 16 steps
    p = make_params("euler", (32, 32), char_decomp=2, time_integrator=3, mthbc=(1, 1, 1, 1), maux=2,
                    mcapa=2)
    a = kappa_aux2(32, 32, 2, 2)
    st = np.stack([rng.uniform(0.1, 1.5, 4), rng.uniform(-1, 1, 4), rng.uniform(-1, 1, 4),
                   rng.uniform(0.03, 1.5, 4)], axis=1)
    lines = ["1", "0.5 0.5"] + [" ".join(repr(float(v)) for v in row) for row in st]
    cd = R.clawdata_for(p, nout=16, outstyle=3, nstepout=8, nstop=16, dt_initial=1e-3,
                        cfl_max=0.5, cfl_desired=0.45)
    fr, dtl, info = R.run(p, cd, lines, aux=a)
    save("run_eu2d_capacity.npz", t=np.array([f[0] for f in fr]), q=np.stack([f[1] for f in fr]),
         dtlog=dtl, states=st, ctl=np.array([1e99, 0.5, 0.45, 16, 1]), dt_initial=np.array(1e-3),
         aux=a)


def swirl():
    """2D advection with the swirl app's edge velocities in aux (rpn2ad1.f,
    setaux.f/psi.f) and its time-dependent b4step: steps on random data and
    two whole runs of the app itself (tperiod 0 and 1.5)."""
    rng = np.random.default_rng(20261206)
    out = {}
    k = 0
    for shape in ((24, 24), (19, 23)):
        for bc in [(1, 1, 1, 1), (2, 2, 2, 2), (1, 2, 2, 1)]:
            if (bc[0] == 2) != (bc[1] == 2) or (bc[2] == 2) != (bc[3] == 2):
                bc = (2, 2, 1, 1)
            p = make_params("advection", shape, char_decomp=0, mthbc=bc, maux=2)
            q = _states.grid(p, rng, smooth=True)
            a = O.swirl_setaux(p)
            dt = 0.02 * p.dx[0]
            qb, dq, cfl = R.step(p, q, dt, aux=a)
            out.update({f"c{k}_q": q, f"c{k}_aux": a, f"c{k}_qbc": qb, f"c{k}_dq": dq,
                        f"c{k}_cfl": np.array(cfl), f"c{k}_dt": np.array(dt),
                        f"c{k}_bc": np.array(bc), f"c{k}_shape": np.array(shape)})
            k += 1
    out["n"] = np.array(k)
    save("step_swirl2d.npz", **out)
    for T, fname in ((0.0, "run_swirl2d.npz"), (1.5, "run_swirl2d_tperiod.npz")):
        p = make_params("advection", (50, 50), char_decomp=0, mthbc=(1, 1, 1, 1), maux=2,
                        time_integrator=4, tperiod=T)
        cd = R.clawdata_for(p, nout=3, outstyle=1, tfinal=1.5, dt_initial=0.025, cfl_max=2.5,
                            cfl_desired=2.45)
        fr, dtl, info = R.run(p, cd, [])
        save(fname, t=np.array([f[0] for f in fr]), q=np.stack([f[1] for f in fr]),
             ctl=np.array([1e99, 2.5, 2.45, 50000, 1]), dt_initial=np.array(0.025),
             tperiod=np.array(T))


PARTS = {"routines": routines, "steps": steps, "runs": runs, "tfluct": tfluct, "flush": flush,
         "recon": recon, "recon2d": recon2d, "aux": aux, "aux2d": aux2d, "swirl": swirl}

if __name__ == "__main__":
    if not R.ref_available():
        sys.exit("oracle/_ref is not built (make -C oracle ref)")
    # python make_golden.py [part ...]   (default: every part)
    for part in (sys.argv[1:] or list(PARTS)):
        PARTS[part]()
