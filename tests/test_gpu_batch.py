"""Batched solves (sc_batch_*) through the C ABI against the oracle.

Every solve of a batch runs the sharpclaw() time loop (src/sharpclaw.f90:35-235)
on the device; each is compared with the oracle's loop on the same grid:
float64 to the whole-run tolerance of test_gpu_parity (max(1e-10, 8x the
oracle's own 1-ulp sensitivity)) with identical step counts, rejections and
info codes; float32 by relative L1 <= 1e-5.  A solve's result does not depend
on the batch it runs in or on the workgroup size: those are bitwise checks.
"""
import os

import numpy as np
import pytest

import _states
from conftest import rel_linf
from oracle import oracle as O
from sharpclaw_amd import PHYSICS, SC_FP32, Batch, ScRunCtl, interior, make_params
from sharpclaw_amd.workloads import dambreak_batch, dambreak_q

pytestmark = pytest.mark.gpu

TOL_RUN = 1e-10


def _solve_params(base, mx, xl, xu):
    kw = dict(char_decomp=base.char_decomp, time_integrator=base.time_integrator,
              mthbc=tuple(base.mthbc), gamma=base.gamma, grav=base.grav, u=base.u, rho=base.rho,
              bulk=base.bulk, efix=base.efix, lim_type=base.lim_type, mthlim=tuple(base.mthlim),
              precision=base.precision)
    name = {v: k for k, v in PHYSICS.items()}[base.physics]
    return make_params(name, (int(mx),), xlower=(xl,), xupper=(xu,), **kw)


def _oracle_runs(base, mx, xl, xu, qs, ctl, tstart, tend, dt):
    out = []
    for i in range(len(mx)):
        p = _solve_params(base, mx[i], xl[i], xu[i])
        q, dtl, st = O.sharpclaw(p, ctl, qs[i], tstart, tend, dt)
        qp = qs[i].copy(order="F")
        interior(p, qp)[:] = np.nextafter(interior(p, qp), np.inf)
        q2, _, _ = O.sharpclaw(p, ctl, qp, tstart, tend, dt)
        tol = max(TOL_RUN, 8.0 * rel_linf(interior(p, q2), interior(p, q)))
        out.append((p, q, dtl, st, tol))
    return out


def _check_fp64(b, res, ref, label):
    for i, (p, q, dtl, st, tol) in enumerate(ref):
        r = res[i]
        gq = b.get_q(i)
        err = rel_linf(interior(p, gq), interior(p, q))
        assert err <= tol, (label, i, int(b.mx[i]), err, tol)
        assert (r.steps_taken, r.rejected, r.info) == (st.steps_taken, st.rejected, st.info), \
            (label, i, r.as_dict(), st.as_dict())
        assert r.t == pytest.approx(st.t, rel=1e-12, abs=1e-14)
        assert r.dt_last == pytest.approx(dtl, rel=1e-6)
        assert r.cfl_maxused == pytest.approx(st.cfl_maxused, rel=1e-6)
        assert r.stage_evals == st.stage_evals


MX_EDGE = [3, 4, 5, 17, 62, 63, 64, 65, 126, 127, 128, 129, 254, 255, 256, 257, 300, 511, 513, 700]


def test_batch_dambreaks_fp64_match_oracle():
    """Uneven sizes around every chunk boundary (63/127/255 cells), first
    steps rejected on the larger grids (dt_initial too big), extrapolation."""
    n = len(MX_EDGE)
    rng = np.random.default_rng(5)
    mx = np.array(MX_EDGE, dtype=np.int32)
    hl, hr, x0 = rng.uniform(1, 3, n), rng.uniform(0.1, 1, n), rng.uniform(0.3, 0.7, n)
    base = make_params("shallow_water", (8,), char_decomp=2, time_integrator=3)
    qs = [dambreak_q(int(mx[i]), hl[i], hr[i], x0[i]) for i in range(n)]
    ctl = ScRunCtl(dt_max=1e99, cfl_max=0.5, cfl_desired=0.45, max_steps=100000, dt_variable=1)
    xl, xu = np.zeros(n), np.ones(n)
    ref = _oracle_runs(base, mx, xl, xu, qs, ctl, 0.0, 0.05, 2e-3)
    assert sum(st.rejected for _, _, _, st, _ in ref) > 0
    with Batch(base, mx, xl, xu) as b:
        for i in range(n):
            b.set_q(i, qs[i])
        res, ms = b.run(ctl, 0.0, 0.05, 2e-3)
        print("batch block", b.block, "ms", ms)
        _check_fp64(b, res, ref, "dambreak")


CASES = [
    # physics, char, integrator, mthbc, cfl (desired, max), domain, data amplitude
    ("advection", 0, 1, (2, 2), (0.45, 0.5), (0.0, 1.0), 1.0),
    ("acoustics", 2, 2, (3, 3), (0.9, 1.0), (-1.0, 1.0), 1.0),
    ("euler", 2, 4, (1, 1), (2.45, 2.5), (-5.0, 5.0), 1.0),
    ("shallow_water", 0, 3, (3, 1), (0.45, 0.5), (0.0, 2.0), 1.0),
    ("euler", 0, 3, (3, 3), (0.45, 0.5), (0.0, 1.0), 0.3),
]


@pytest.mark.parametrize("name,chr_,ti,bc,cfl,dom,amp", CASES)
def test_batch_physics_integrators_match_oracle(name, chr_, ti, bc, cfl, dom, amp):
    rng = np.random.default_rng(17)
    mx = np.array([3, 40, 64, 129, 260, 333], dtype=np.int32)
    n = len(mx)
    xl = np.full(n, dom[0]) + rng.uniform(-0.2, 0.2, n)
    xu = xl + (dom[1] - dom[0]) * rng.uniform(0.5, 1.5, n)
    base = make_params(name, (8,), char_decomp=chr_, time_integrator=ti, mthbc=bc + bc)
    qs = []
    for i in range(n):
        p = _solve_params(base, mx[i], xl[i], xu[i])
        q = O.new_grid(p)
        interior(p, q)[:] = _states.random_cells(p, (int(mx[i]),), rng, smooth=True,
                                                 amp=min(amp, 0.5) if mx[i] < 10 else amp)
        qs.append(q)
    ctl = ScRunCtl(dt_max=1e99, cfl_max=cfl[1], cfl_desired=cfl[0], max_steps=400, dt_variable=1)
    ref = _oracle_runs(base, mx, xl, xu, qs, ctl, 0.0, 0.08, 1e-3)
    with Batch(base, mx, xl, xu) as b:
        for i in range(n):
            b.set_q(i, qs[i])
        res, _ = b.run(ctl, 0.0, 0.08, 1e-3)
        _check_fp64(b, res, ref, name)


def test_batch_fixed_dt_and_info_codes():
    """dt_variable = 0: a fixed number of steps, info 12 when the CFL limit
    is exceeded (sharpclaw.f90:220-230), info 2 when dt does not divide the
    interval and info 3 for cfl_desired > cfl_max (both before any step)."""
    mx = np.array([50, 100, 200], dtype=np.int32)
    n = len(mx)
    base = make_params("shallow_water", (8,), char_decomp=2, time_integrator=3)
    xl, xu = np.zeros(n), np.ones(n)
    qs = [dambreak_q(int(m), 2.0, 0.5, 0.5) for m in mx]
    for dt, tend in ((1e-4, 0.01), (1e-3, 0.01), (3e-4, 0.01)):
        ctl = ScRunCtl(dt_max=1e99, cfl_max=0.5, cfl_desired=0.45, max_steps=1000, dt_variable=0)
        ref = _oracle_runs(base, mx, xl, xu, qs, ctl, 0.0, tend, dt)
        with Batch(base, mx, xl, xu) as b:
            for i in range(n):
                b.set_q(i, qs[i])
            res, _ = b.run(ctl, 0.0, tend, dt)
            _check_fp64(b, res, ref, f"fixed dt {dt}")
            print(dt, [r.info for r in res])
    ctl = ScRunCtl(dt_max=1e99, cfl_max=0.5, cfl_desired=0.6, max_steps=10, dt_variable=1)
    with Batch(base, mx, xl, xu) as b:
        for i in range(n):
            b.set_q(i, qs[i])
        res, _ = b.run(ctl, 0.0, 0.01, 1e-4)
        assert all(r.info == 3 and r.steps_taken == 0 for r in res)
        for i in range(n):
            assert np.array_equal(interior(_solve_params(base, mx[i], 0, 1), b.get_q(i)),
                                  interior(_solve_params(base, mx[i], 0, 1), qs[i]))


def test_batch_continues_across_runs_and_fills_ghosts():
    """Two consecutive intervals without re-upload equal the oracle's two
    chained calls; get_q returns ghost cells filled as bc1 does."""
    mx, hl, hr, x0 = dambreak_batch(6, seed=9, mx_lo=20, mx_hi=400)
    n = len(mx)
    base = make_params("shallow_water", (8,), char_decomp=2, time_integrator=4, mthbc=(3, 1))
    xl, xu = np.zeros(n), np.ones(n)
    qs = [dambreak_q(int(mx[i]), hl[i], hr[i], x0[i]) for i in range(n)]
    ctl = ScRunCtl(dt_max=1e99, cfl_max=2.5, cfl_desired=2.45, max_steps=100000, dt_variable=1)
    ref1 = _oracle_runs(base, mx, xl, xu, qs, ctl, 0.0, 0.02, 1e-3)
    with Batch(base, mx, xl, xu) as b:
        for i in range(n):
            b.set_q(i, qs[i])
        res1, _ = b.run(ctl, 0.0, 0.02, 1e-3)
        _check_fp64(b, res1, ref1, "interval 1")
        # the second interval starts on the device states, dt from each solve
        dts = [r.dt_last for r in res1]
        assert len(set(dts)) > 1
        dt = dts[0]
        ref2 = _oracle_runs(base, mx, xl, xu, [b.get_q(i) for i in range(n)], ctl, 0.02, 0.04, dt)
        res2, _ = b.run(ctl, 0.02, 0.04, dt)
        _check_fp64(b, res2, ref2, "interval 2")
        for i in range(n):
            p = _solve_params(base, mx[i], 0.0, 1.0)
            g = b.get_q(i)
            assert np.array_equal(g, O.bc(p, g))


def test_batch_fp32_dambreaks_l1():
    """float32 reconstruction + Riemann, float64 accumulation, against the
    float64 oracle: relative L1 <= 1e-5 per solve."""
    mx, hl, hr, x0 = dambreak_batch(24, seed=2024, mx_lo=64, mx_hi=512)
    n = len(mx)
    base = make_params("shallow_water", (8,), char_decomp=2, time_integrator=3)
    xl, xu = np.zeros(n), np.ones(n)
    qs = [dambreak_q(int(mx[i]), hl[i], hr[i], x0[i]) for i in range(n)]
    ctl = ScRunCtl(dt_max=1e99, cfl_max=0.5, cfl_desired=0.45, max_steps=100000, dt_variable=1)
    base32 = make_params("shallow_water", (8,), char_decomp=2, time_integrator=3, precision=SC_FP32)
    worst = 0.0
    with Batch(base32, mx, xl, xu) as b:
        for i in range(n):
            b.set_q(i, qs[i])
        res, ms = b.run(ctl, 0.0, 0.2, 1e-4)
        for i in range(n):
            p = _solve_params(base, mx[i], 0.0, 1.0)
            rq, _, st = O.sharpclaw(p, ctl, qs[i], 0.0, 0.2, 1e-4)
            g = interior(p, b.get_q(i))
            r = interior(p, rq)
            err = np.sum(np.abs(g - r)) / np.sum(np.abs(r))
            worst = max(worst, err)
            assert res[i].info == 0 and res[i].t == pytest.approx(0.2, abs=1e-14)
    print("fp32 batch worst relative L1:", worst, "ms", ms)
    assert worst <= 1e-5


def _run_batch(base, mx, xl, xu, qs, ctl, tend, dt, block=None):
    old = os.environ.get("SC_BATCH_BLOCK")
    if block:
        os.environ["SC_BATCH_BLOCK"] = str(block)
    try:
        with Batch(base, mx, xl, xu) as b:
            for i in range(len(mx)):
                b.set_q(i, qs[i])
            res, _ = b.run(ctl, 0.0, tend, dt)
            return [b.get_q(i) for i in range(len(mx))], res, b.block
    finally:
        if block:
            if old is None:
                del os.environ["SC_BATCH_BLOCK"]
            else:
                os.environ["SC_BATCH_BLOCK"] = old


@pytest.mark.parametrize("fp32", [0, 1])
def test_batch_alone_vs_mixed_and_block_sizes_bitwise(fp32):
    mx, hl, hr, x0 = dambreak_batch(12, seed=4, mx_lo=64, mx_hi=700)
    n = len(mx)
    base = make_params("shallow_water", (8,), char_decomp=2, time_integrator=3,
                       precision=SC_FP32 if fp32 else 0)
    xl, xu = np.zeros(n), np.ones(n)
    qs = [dambreak_q(int(mx[i]), hl[i], hr[i], x0[i]) for i in range(n)]
    ctl = ScRunCtl(dt_max=1e99, cfl_max=0.5, cfl_desired=0.45, max_steps=100000, dt_variable=1)
    mixed, rm, _ = _run_batch(base, mx, xl, xu, qs, ctl, 0.03, 1e-3, block=128)
    for k in (0, 5, n - 1):
        alone, ra, _ = _run_batch(base, mx[k:k + 1], xl[k:k + 1], xu[k:k + 1], qs[k:k + 1], ctl, 0.03,
                                  1e-3, block=128)
        assert np.array_equal(alone[0], mixed[k])
        assert (ra[0].steps_taken, ra[0].dt_last, ra[0].t) == (rm[k].steps_taken, rm[k].dt_last, rm[k].t)
    for blk in (64, 256, 512, 1024):
        other, _, got = _run_batch(base, mx, xl, xu, qs, ctl, 0.03, 1e-3, block=blk)
        assert got == blk
        for i in range(n):
            assert np.array_equal(other[i], mixed[i]), (blk, i)
    tuned, _, got = _run_batch(base, mx, xl, xu, qs, ctl, 0.03, 1e-3)
    assert got in (64, 128, 256, 512, 1024)
    for i in range(n):
        assert np.array_equal(tuned[i], mixed[i])


@pytest.mark.parametrize("ends", ["1", "2", "3", "8"])
def test_batch_queue_ends_bitwise(ends, monkeypatch):
    """The two-ended work queue hands every solve to exactly one workgroup,
    whichever share of workgroups takes the smallest solves first (every
    ends-th one; 1 = largest first only): results equal the single-ended
    run bit for bit, and every solve reaches t_end."""
    mx, hl, hr, x0 = dambreak_batch(20, seed=11, mx_lo=16, mx_hi=600)
    n = len(mx)
    base = make_params("shallow_water", (8,), char_decomp=2, time_integrator=3)
    xl, xu = np.zeros(n), np.ones(n)
    qs = [dambreak_q(int(mx[i]), hl[i], hr[i], x0[i]) for i in range(n)]
    ctl = ScRunCtl(dt_max=1e99, cfl_max=0.5, cfl_desired=0.45, max_steps=100000, dt_variable=1)
    monkeypatch.setenv("SC_BATCH_ENDS", "1")
    ref, rr, _ = _run_batch(base, mx, xl, xu, qs, ctl, 0.02, 1e-3, block=64)
    monkeypatch.setenv("SC_BATCH_ENDS", ends)
    got, rg, _ = _run_batch(base, mx, xl, xu, qs, ctl, 0.02, 1e-3, block=64)
    for i in range(n):
        assert rg[i].info == 0 and rg[i].t == pytest.approx(0.02, abs=1e-14)
        assert rg[i].steps_taken == rr[i].steps_taken
        assert np.array_equal(got[i], ref[i]), i


@pytest.mark.parametrize("fp32", [0, 1])
def test_batch_split_solves_bitwise(fp32, monkeypatch):
    """The tuner splits the largest solves over several workgroups that meet
    at a device barrier after every stage (sc_batch_parts > 1); the results,
    step counts and dt equal the run with every solve on one workgroup
    (SC_BATCH_PARTS=1) bit for bit."""
    mx, hl, hr, x0 = dambreak_batch(60, seed=23, mx_lo=64, mx_hi=900)
    mx[:3] = [4096, 3000, 2048]
    n = len(mx)
    base = make_params("shallow_water", (8,), char_decomp=2, time_integrator=3,
                       precision=SC_FP32 if fp32 else 0)
    xl, xu = np.zeros(n), np.ones(n)
    qs = [dambreak_q(int(mx[i]), hl[i], hr[i], x0[i]) for i in range(n)]
    ctl = ScRunCtl(dt_max=1e99, cfl_max=0.5, cfl_desired=0.45, max_steps=100000, dt_variable=1)
    runs = {}
    for cap in ("1", None):
        if cap:
            monkeypatch.setenv("SC_BATCH_PARTS", cap)
        else:
            monkeypatch.delenv("SC_BATCH_PARTS", raising=False)
        monkeypatch.setenv("SC_BATCH_BLOCK", "256")
        with Batch(base, mx, xl, xu) as b:
            for i in range(n):
                b.set_q(i, qs[i])
            res, ms = b.run(ctl, 0.0, 0.01, 1e-4)
            runs[cap] = ([b.get_q(i) for i in range(n)], res, b.parts.copy())
    q1, r1, p1 = runs["1"]
    q2, r2, p2 = runs[None]
    assert (p1 == 1).all()
    assert p2.max() > 1 and p2[0] > 1, p2[:5]
    for i in range(n):
        assert r2[i].info == 0 and r2[i].t == pytest.approx(0.01, abs=1e-14)
        assert (r2[i].steps_taken, r2[i].dt_last, r2[i].cfl_maxused) == \
            (r1[i].steps_taken, r1[i].dt_last, r1[i].cfl_maxused), i
        assert np.array_equal(q2[i], q1[i]), (i, int(mx[i]), int(p2[i]))


RECON = [
    # physics, lim_type, char_decomp, integrator, mthbc, limiters, float32
    ("euler", 2, 1, 3, (3, 1), (5, 5, 5, 5), 0),          # weno5_wave
    ("shallow_water", 2, 3, 2, (1, 3), (5, 5, 5, 5), 0),  # weno5_trans
    ("acoustics", 1, 1, 1, (3, 3), (1, 2, 5, 5), 0),      # tvd2_wave, minmod / superbee
    ("euler", 1, 2, 4, (1, 1), (4, 3, 2, 5), 0),          # tvd2_char, MC / van Leer / superbee
    ("advection", 2, 1, 3, (2, 2), (5, 5, 5, 5), 0),      # weno5_wave, periodic
    ("euler", 2, 3, 3, (1, 1), (5, 5, 5, 5), 1),          # weno5_trans, float32
    ("shallow_water", 1, 2, 3, (1, 1), (3, 3, 5, 5), 1),  # tvd2_char, float32
]


@pytest.mark.parametrize("name,lim,chr_,ti,bc,mthlim,fp32", RECON)
def test_batch_reconstructions_match_oracle(name, lim, chr_, ti, bc, mthlim, fp32):
    """The reconstructions with neighbour data (weno5_wave, weno5_trans,
    tvd2_wave, tvd2_char; flux1.f90:65-117) in the persistent batch kernel:
    uneven sizes across the 255-cell chunk boundary, each solve against the
    oracle's time loop (float64: whole-run tolerance and equal step
    counts; float32: relative L1 <= 1e-5), and the tuner keeps to the
    workgroup sizes the kernel is built for with these options (<= 256)."""
    rng = np.random.default_rng(29 + lim * 4 + chr_)
    mx = np.array([3, 40, 254, 256, 333, 700], dtype=np.int32)
    n = len(mx)
    xl = np.zeros(n)
    xu = rng.uniform(0.5, 1.5, n)
    base = make_params(name, (8,), char_decomp=chr_, lim_type=lim, mthlim=mthlim, time_integrator=ti,
                       mthbc=bc + bc, precision=SC_FP32 if fp32 else 0)
    qs = []
    for i in range(n):
        p = _solve_params(base, mx[i], xl[i], xu[i])
        q = O.new_grid(p)
        interior(p, q)[:] = _states.random_cells(p, (int(mx[i]),), rng, smooth=True, amp=0.3 if mx[i] < 10 else 0.8)
        qs.append(q)
    cfl = (2.45, 2.5) if ti == 4 else (0.45, 0.5)
    ctl = ScRunCtl(dt_max=1e99, cfl_max=cfl[1], cfl_desired=cfl[0], max_steps=5000, dt_variable=1)
    with Batch(base, mx, xl, xu) as b:
        for i in range(n):
            b.set_q(i, qs[i])
        res, _ = b.run(ctl, 0.0, 0.04, 1e-3)
        assert b.block <= 256, b.block
        assert all(r.info == 0 for r in res), [r.info for r in res]
        tun = b.tuning()
        assert all((pr < 0) == (bl > 256) for bl, pr in tun.items()), tun
        if not fp32:
            ref = _oracle_runs(base, mx, xl, xu, qs, ctl, 0.0, 0.04, 1e-3)
            _check_fp64(b, res, ref, (name, lim, chr_))
        else:
            base64 = make_params(name, (8,), char_decomp=chr_, lim_type=lim, mthlim=mthlim, time_integrator=ti,
                                 mthbc=bc + bc)
            for i in range(n):
                p = _solve_params(base64, mx[i], xl[i], xu[i])
                q, _, st = O.sharpclaw(p, ctl, qs[i], 0.0, 0.04, 1e-3)
                g, o = interior(p, b.get_q(i)), interior(p, q)
                err = np.sum(np.abs(g - o)) / np.sum(np.abs(o))
                assert st.info == 0 and err <= 1e-5, (i, int(mx[i]), err)


def test_batch_split_reconstruction_bitwise(monkeypatch):
    """weno5_trans solves split over several workgroups equal the one-workgroup
    run bit for bit (the neighbour data of a chunk's edge interfaces is
    recomputed by each part)."""
    mx = np.array([2000, 1500, 300, 90, 64, 700], dtype=np.int32)
    n = len(mx)
    rng = np.random.default_rng(31)
    base = make_params("euler", (8,), char_decomp=3, time_integrator=3, mthbc=(3, 1, 3, 1))
    xl, xu = np.zeros(n), np.ones(n)
    qs = []
    for i in range(n):
        p = _solve_params(base, mx[i], xl[i], xu[i])
        q = O.new_grid(p)
        interior(p, q)[:] = _states.random_cells(p, (int(mx[i]),), rng, smooth=True, amp=0.5)
        qs.append(q)
    ctl = ScRunCtl(dt_max=1e99, cfl_max=0.5, cfl_desired=0.45, max_steps=100000, dt_variable=1)
    runs = {}
    for cap in ("1", None):
        if cap:
            monkeypatch.setenv("SC_BATCH_PARTS", cap)
        else:
            monkeypatch.delenv("SC_BATCH_PARTS", raising=False)
        monkeypatch.setenv("SC_BATCH_BLOCK", "128")
        with Batch(base, mx, xl, xu) as b:
            for i in range(n):
                b.set_q(i, qs[i])
            res, _ = b.run(ctl, 0.0, 0.01, 1e-4)
            assert b.block == 128
            runs[cap] = ([b.get_q(i) for i in range(n)], res, b.parts.copy())
    q1, r1, p1 = runs["1"]
    q2, r2, p2 = runs[None]
    assert (p1 == 1).all() and p2[0] > 1, p2
    for i in range(n):
        assert (r2[i].steps_taken, r2[i].dt_last) == (r1[i].steps_taken, r1[i].dt_last), i
        assert np.array_equal(q2[i], q1[i]), (i, int(mx[i]), int(p2[i]))
