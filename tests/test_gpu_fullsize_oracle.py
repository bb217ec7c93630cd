"""Oracle parity at the full sizes of the measured workloads.

2D Euler 1024 x 1024 (characteristic WENO5 + Roe, SSPRK3, the benchmark's
configuration): one fused SSP step of the device (sc_ts) against one step of
the oracle (orc_ts, the reference's ts_ssp33 over flux2) from the same q, to
relative L-inf <= 1e-12 with the same CFL number.  The states cover what
the tile kernel does differently at this size: the seeded four-quadrant
data (most tiles exactly constant: the constant-tile branch), the states the
device reaches after 10 and 40 steps (fronts crossing tiles; the partial
last tile row, 1024 = 68 * 15 + 4 for 16 x 15 tiles; 4416 tiles under the
XCD remap; CFL words with 32 copies), walls on every side, and a smooth
state where no tile is constant.  Every tile shape gives the oracle's
result.  The whole time loop (sc_advance) over a few steps, the four
reconstructions that run as x / y slice sweeps, and the float32 option
(relative L1 <= 1e-5 against the float64 oracle) are checked at the same
size.  The oracle takes about 3 s per 1024^2 step.

Batched (4096 shallow-water dam breaks, mx up to 4096) and served workloads
(2D shallow water up to 1024^2, acoustics up to 4096 cells) at the top of
their size ranges: dam breaks with mx in [2048, 4096] to t = 0.2 in float32
(relative L1 <= 1e-5) and float64 (the whole-run tolerance: max(1e-10, 8x
the oracle's own 1-ulp sensitivity), with the oracle's step counts); a 512^2
shallow-water request and 4096-cell acoustics requests through the server.
The oracle runs of one test go to a thread pool (ctypes releases the GIL;
the 1D and aux-free oracle paths keep no global state).
"""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from conftest import rel_linf
from oracle import oracle as O
from sharpclaw_amd import (SC_FP32, SC_FP64, Batch, ScRunCtl, Server, Solver, interior,
                           make_params)
from sharpclaw_amd.workloads import SERVED_CTL, dambreak_batch, dambreak_q
from test_gpu_fullsize import quadrants, quadrants_mild, smooth_periodic

pytestmark = pytest.mark.gpu

N = 1024
TOL_STEP = 1e-12
TOL_RUN = 1e-10
CTL = ScRunCtl(dt_max=1e99, cfl_max=0.5, cfl_desired=0.45, max_steps=100000, dt_variable=1)


def euler(mthbc=(1, 1, 1, 1), **kw):
    kw.setdefault("char_decomp", 2)
    return make_params("euler", (N, N), time_integrator=3, mthbc=mthbc, **kw)


def stable_dt(p, q, cfl=0.4):
    """dt with the given Courant number for the state's max(|u|+c, |v|+c)."""
    qi = interior(p, q)
    rho = qi[..., 0]
    u, v = qi[..., 1] / rho, qi[..., 2] / rho
    pr = (p.gamma - 1.0) * (qi[..., 3] - 0.5 * rho * (u * u + v * v))
    c = np.sqrt(p.gamma * pr / rho)
    smax = max(np.max(np.abs(u) + c), np.max(np.abs(v) + c))
    return cfl * p.dx[0] / smax


def gpu_ts(p, q, dt, tile=-1):
    with Solver(p) as s:
        s.set_tile(tile)
        s.set_q(q)
        s.save_qold()
        cfl = s.ts(dt)
        return s.get_q(), cfl, s.get_tile()


@pytest.fixture(scope="module")
def trajectory():
    """Four-quadrant data (extrapolation) and the device's states after 10 and 40 steps."""
    p = euler()
    q0 = quadrants(p)
    out = {"initial": q0}
    with Solver(p) as s:
        s.set_q(q0)
        dt = 1e-4
        for n in (10, 30):
            ctl = ScRunCtl(dt_max=1e99, cfl_max=0.5, cfl_desired=0.45, max_steps=n, dt_variable=1)
            dt, st = s.advance(ctl, 0.0, 1e9, dt)
            assert st.steps_taken == n
            out[f"steps{10 if n == 10 else 40}"] = s.get_q()
    return p, out


def _check_step(p, q, dt, tiles=(-1,), label=""):
    oq, ocfl, _ = O.ts(p, q, dt)
    ref = interior(p, oq)
    for tile in tiles:
        gq, gcfl, (k, shape) = gpu_ts(p, q, dt, tile)
        err = rel_linf(interior(p, gq), ref)
        print(f"{label} tile {k} {shape}: rel L-inf {err:.3e}, cfl {gcfl!r} vs {ocfl!r}")
        assert np.all(np.isfinite(interior(p, gq)))
        assert err <= TOL_STEP, (label, tile, err)
        assert abs(gcfl - ocfl) <= 1e-13 * ocfl, (label, tile, gcfl, ocfl)


@pytest.mark.parametrize("state", ["initial", "steps10", "steps40"])
def test_c3_step_vs_oracle(trajectory, state):
    p, states = trajectory
    q = states[state]
    _check_step(p, q, stable_dt(p, q), label=state)


def test_c3_step_every_tile_shape(trajectory):
    p, states = trajectory
    q = states["steps10"]
    _check_step(p, q, stable_dt(p, q), tiles=(0, 1, 2, 3), label="steps10")


def test_c3_step_walls(trajectory):
    _, states = trajectory
    p = euler(mthbc=(3, 3, 3, 3))
    q = states["steps10"]
    _check_step(p, q, stable_dt(p, q), label="walls")


def test_c3_step_smooth_state():
    """No tile of a smooth state is constant: every tile takes the full path."""
    p = euler()
    q = smooth_periodic(p)
    _check_step(p, q, stable_dt(p, q), label="smooth")


def test_c3_advance_vs_oracle(trajectory):
    """The time loop (dt from each step's CFL, accept/reject) over 4 steps
    from the 10-step state, against the oracle's sharpclaw loop."""
    p, states = trajectory
    q = states["steps10"]
    ctl = ScRunCtl(dt_max=1e99, cfl_max=0.5, cfl_desired=0.45, max_steps=4, dt_variable=1)
    dt0 = stable_dt(p, q, 0.6)  # above cfl_max: the first step is rejected and retaken
    oq, odt, ost = O.sharpclaw(p, ctl, q, 0.0, 1e9, dt0)
    with Solver(p) as s:
        s.set_q(q)
        gdt, gst = s.advance(ctl, 0.0, 1e9, dt0)
        gq = s.get_q()
    assert ost.rejected >= 1
    assert (gst.steps_taken, gst.rejected, gst.stage_evals) == (ost.steps_taken, ost.rejected, ost.stage_evals)
    assert gst.t == pytest.approx(ost.t, rel=1e-13)
    assert gdt == pytest.approx(odt, rel=1e-13)
    err = rel_linf(interior(p, gq), interior(p, oq))
    print(f"advance 4 steps: rel L-inf {err:.3e}")
    assert err <= 4 * TOL_STEP


@pytest.mark.parametrize("lim,char", [(2, 1), (2, 3), (1, 1), (1, 2)])
def test_c3_sweeps_step_vs_oracle(lim, char):
    """weno5_wave, weno5_trans, tvd2_wave, tvd2_char (MC limiter): the x and
    y slice-sweep kernels at 1024^2, walls."""
    p = euler(mthbc=(3, 3, 3, 3), char_decomp=char, lim_type=lim, mthlim=(4, 4, 4, 4))
    q = quadrants_mild(p)
    with Solver(p) as s:  # a few steps so the fronts have spread over several cells
        s.set_q(q)
        s.advance(ScRunCtl(dt_max=1e99, cfl_max=0.5, cfl_desired=0.45, max_steps=5, dt_variable=1),
                  0.0, 1e9, 1e-4)
        q = s.get_q()
    _check_step(p, q, stable_dt(p, q), label=f"sweeps {lim}/{char}")


def test_c3_fp32_option_l1_vs_oracle(trajectory):
    """float32 reconstruction and Riemann arithmetic over 5 steps from the
    four-quadrant data against the float64 oracle's loop."""
    p, states = trajectory
    q = states["initial"]
    ctl = ScRunCtl(dt_max=1e99, cfl_max=0.5, cfl_desired=0.45, max_steps=5, dt_variable=1)
    oq, _, ost = O.sharpclaw(p, ctl, q, 0.0, 1e9, 1e-4)
    with Solver(euler(precision=SC_FP32)) as s:
        s.set_q(q)
        _, gst = s.advance(ctl, 0.0, 1e9, 1e-4)
        gq = s.get_q()
    assert gst.steps_taken == ost.steps_taken
    a, b = interior(p, gq), interior(p, oq)
    err = np.sum(np.abs(a - b)) / np.sum(np.abs(b))
    print(f"1024^2 float32 relative L1 vs the float64 oracle: {err:.3e}")
    assert err <= 1e-5


# ------------------------------------------------------------ batched (C4)

def _dambreak_set(n=6):
    """Dam breaks of the batched workload's distribution with mx spread over
    [2048, 4096] (the largest solves of the 4096-solve batch)."""
    _, hl, hr, x0 = dambreak_batch(n, seed=77)
    mx = np.rint(np.geomspace(2048, 4096, n)).astype(np.int32)
    return mx, hl, hr, x0


def _oracle_with_tol(p, ctl, q0, tstart, tend, dt, pool):
    """The oracle's run and its whole-run tolerance (futures)."""
    qp = q0.copy(order="F")
    interior(p, qp)[:] = np.nextafter(interior(p, qp), np.inf)
    return (pool.submit(O.sharpclaw, p, ctl, q0, tstart, tend, dt),
            pool.submit(O.sharpclaw, p, ctl, qp, tstart, tend, dt))


def _tol(p, f, fp):
    q, _, _ = f.result()
    q2, _, _ = fp.result()
    return max(TOL_RUN, 8.0 * rel_linf(interior(p, q2), interior(p, q)))


def test_c4_largest_dambreaks_vs_oracle():
    mx, hl, hr, x0 = _dambreak_set()
    n = len(mx)
    base = make_params("shallow_water", (8,), char_decomp=2, time_integrator=3, grav=9.81)
    qs = [dambreak_q(int(mx[i]), hl[i], hr[i], x0[i]) for i in range(n)]
    ps = [make_params("shallow_water", (int(m),), char_decomp=2, time_integrator=3, grav=9.81) for m in mx]
    with ThreadPoolExecutor(max_workers=2 * n) as pool:
        futs = [_oracle_with_tol(ps[i], CTL, qs[i], 0.0, 0.2, 1e-4, pool) for i in range(n)]
        got = {}
        for prec in (SC_FP32, SC_FP64):
            base.precision = prec
            with Batch(base, mx, 0.0, 1.0) as b:
                for i in range(n):
                    b.set_q(i, qs[i])
                res, ms = b.run(CTL, 0.0, 0.2, 1e-4)
                print(f"batch precision {prec}: {ms:.1f} ms, block {b.block}")
                got[prec] = [(b.get_q(i), res[i]) for i in range(n)]
        for i in range(n):
            p = ps[i]
            oq, odt, ost = futs[i][0].result()
            tol = _tol(p, *futs[i])
            q64, r64 = got[SC_FP64][i]
            err = rel_linf(interior(p, q64), interior(p, oq))
            q32, r32 = got[SC_FP32][i]
            a, b_ = interior(p, q32), interior(p, oq)
            l1 = np.sum(np.abs(a - b_)) / np.sum(np.abs(b_))
            print(f"mx {mx[i]}: steps {ost.steps_taken}, fp64 rel L-inf {err:.3e} (tol {tol:.1e}), "
                  f"fp32 rel L1 {l1:.3e}")
            assert err <= tol, (int(mx[i]), err, tol)
            assert (r64.steps_taken, r64.rejected, r64.info) == (ost.steps_taken, ost.rejected, ost.info)
            assert r64.t == pytest.approx(0.2, abs=1e-14) and r32.t == pytest.approx(0.2, abs=1e-14)
            assert l1 <= 1e-5, (int(mx[i]), l1)


# ------------------------------------------------------------- served (C5)

def _served_large():
    """A 512^2 shallow-water request (to t = 0.01) and a 4096-cell acoustics
    request (to t = 0.25), each in float64 and float32."""
    out = []
    n = 512
    for prec in (SC_FP64, SC_FP32):
        p = make_params("shallow_water", (n, n), mthbc=(1, 1, 1, 1), grav=9.81, precision=prec)
        xc = (np.arange(n) + 0.5) / n
        X, Y = np.meshgrid(xc, xc, indexing="ij")
        q = O.new_grid(p)
        interior(p, q)[..., 0] = 1.0 + 0.6 * np.exp(-((X - 0.45) ** 2 + (Y - 0.55) ** 2) / 0.05 ** 2)
        out.append({"kind": "sw2d", "params": p, "q0": q, "tend": 0.01, "dt0": 1e-5})
        mx = 4096
        p = make_params("acoustics", (mx,), mthbc=(3, 3), rho=1.0, bulk=4.0, precision=prec)
        xc = (np.arange(mx) + 0.5) / mx
        q = O.new_grid(p)
        interior(p, q)[:, 0] = 0.7 * np.exp(-((xc - 0.4) / 0.05) ** 2)
        out.append({"kind": "acoustics1d", "params": p, "q0": q, "tend": 0.25, "dt0": 1e-5})
    return out


def test_c5_large_requests_vs_oracle():
    reqs = _served_large()
    ctl = ScRunCtl(**SERVED_CTL)
    with ThreadPoolExecutor(max_workers=4) as pool:
        ref = {}
        for r in reqs:
            if r["params"].precision == SC_FP64:
                ref[r["kind"]] = (r["params"], _oracle_with_tol(r["params"], ctl, r["q0"], 0.0, r["tend"],
                                                                r["dt0"], pool))
        with Server(hbm_bytes=4 << 30, max_active=8) as srv:
            ids = [srv.submit(r["params"], ctl, r["q0"], 0.0, r["tend"], r["dt0"]) for r in reqs]
            srv.drain(300e3)
            out = [srv.result(rid) for rid in ids]
        for r, (gq, gst) in zip(reqs, out):
            p64, (f, fp) = ref[r["kind"]]
            oq, _, ost = f.result()
            p = r["params"]
            assert gst.run.info == 0 and gst.run.t == pytest.approx(r["tend"], abs=1e-14)
            if p.precision == SC_FP64:
                tol = _tol(p64, f, fp)
                err = rel_linf(interior(p, gq), interior(p, oq))
                print(f"{r['kind']} fp64: steps {ost.steps_taken}, rel L-inf {err:.3e} (tol {tol:.1e})")
                assert err <= tol, (r["kind"], err, tol)
                assert gst.run.steps_taken == ost.steps_taken
            else:
                a, b = interior(p, gq), interior(p, oq)
                l1 = np.sum(np.abs(a - b)) / np.sum(np.abs(b))
                print(f"{r['kind']} fp32: rel L1 {l1:.3e}")
                assert l1 <= 1e-5, (r["kind"], l1)
