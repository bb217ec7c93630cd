"""The device-resident 2D time loop (rounds of steps decided on the device;
eager, or rounds after the first replayed from a captured hipGraph) against
the host-driven loop (CFL read back every step): bit-identical q, step
counts, rejections and dt, over several consecutive intervals with eager
work (get_q, sums) between the rounds and graph replays."""
import os

import numpy as np
import pytest

import _states
from oracle import oracle as O
from sharpclaw_amd import SC_FP32, ScRunCtl, Solver, interior, make_params

pytestmark = pytest.mark.gpu


def _make(p, mode):
    """mode: host, device (eager rounds) or graph (device, replayed rounds);
    the solver reads SC_HOST_LOOP and SC_GRAPH when it is created."""
    env = {"host": ("1", "0"), "device": ("0", "0"), "graph": ("0", "1")}[mode]
    old = {k: os.environ.get(k) for k in ("SC_HOST_LOOP", "SC_GRAPH")}
    os.environ["SC_HOST_LOOP"], os.environ["SC_GRAPH"] = env
    try:
        return Solver(p)
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


@pytest.mark.parametrize("phys,ti,n,prec", [("euler", 3, (160, 130), 0), ("euler", 4, (96, 96), 0),
                                            ("shallow_water", 2, (130, 170), 0), ("euler", 3, (144, 144), SC_FP32),
                                            ("shallow_water", 1, (40, 50), 0)])
def test_device_loop_equals_host_loop(phys, ti, n, prec):
    rng = np.random.default_rng(41)
    p = make_params(phys, n, time_integrator=ti, mthbc=(1, 3, 2, 2), precision=prec)
    q0 = O.new_grid(p)
    interior(p, q0)[:] = _states.random_cells(p, n, rng, smooth=True, amp=0.4)
    cfl = (2.45, 2.5) if ti == 4 else (0.45, 0.5)
    ctl = ScRunCtl(dt_max=1e99, cfl_max=cfl[1], cfl_desired=cfl[0], max_steps=60, dt_variable=1)
    out = {}
    for mode in ("host", "device", "graph"):
        with _make(p, mode) as s:
            s.set_q(q0)
            rec = []
            dt, t = 0.05, 0.0          # first step rejected (dt too large)
            for k in range(4):
                dt, st = s.advance(ctl, t, t + 0.01, dt)
                t = st.t
                rec.append((st.steps_taken, st.rejected, st.info, st.t, dt, st.cfl_maxused))
                rec.append(tuple(s.sums()))           # eager work between replays
                rec.append(s.get_q().tobytes())
            out[mode] = rec
    assert out["device"] == out["host"]
    assert out["graph"] == out["host"]
    assert out["host"][0][1] > 0 or ti == 4  # the first interval retook its first step


@pytest.mark.parametrize("phys,ti,mx,prec,lim,chr_", [
    ("euler", 4, 1600, 0, 2, 2), ("euler", 3, 5000, 0, 2, 2), ("shallow_water", 2, 700, SC_FP32, 2, 0),
    ("acoustics", 1, 300, 0, 2, 0), ("advection", 3, 255, 0, 2, 0),
    # the reconstructions with neighbour data (flux1.f90:65-117)
    ("euler", 3, 900, 0, 2, 1), ("euler", 4, 1200, 0, 2, 3), ("shallow_water", 3, 600, 0, 1, 1),
    ("acoustics", 2, 500, SC_FP32, 1, 2)])
def test_device_loop_1d_equals_host_loop(phys, ti, mx, prec, lim, chr_):
    """sc_advance on a 1D grid of up to 32 chunks runs its time loop on the
    device, the grid split over workgroups that own ranges of cells, each
    computing a whole step on its range widened by the cells its remaining
    stages read, meeting at a device barrier once per step; q, step
    counts, rejections, dt and CFL equal the host-driven loop's
    (SC_HOST_LOOP=1) bit for bit over consecutive intervals, the first of
    which retakes its first step."""
    _loop1d_case(phys, ti, mx, prec, lim, chr_, (1, 1) if phys == "advection" else (3, 1))


@pytest.mark.parametrize("phys,ti,mx,prec,lim,chr_,bc,stagewise", [
    ("euler", 4, 1600, 0, 2, 2, (2, 2), False), ("euler", 3, 900, 0, 2, 1, (2, 2), False),
    ("euler", 4, 1600, 0, 2, 2, (3, 1), True), ("shallow_water", 2, 700, SC_FP32, 2, 0, (1, 3), True)])
def test_device_loop_1d_barrier_per_stage(phys, ti, mx, prec, lim, chr_, bc, stagewise):
    """The 1D device loop with a barrier after every stage, each workgroup
    owning whole chunks: the path of periodic boundaries (whose ghost cells
    come from the far end of the grid) and of SC_LOOP_STAGEWISE=1; bitwise
    the host-driven loop's results."""
    old = os.environ.get("SC_LOOP_STAGEWISE")
    if stagewise:
        os.environ["SC_LOOP_STAGEWISE"] = "1"
    try:
        _loop1d_case(phys, ti, mx, prec, lim, chr_, bc)
    finally:
        if old is None:
            os.environ.pop("SC_LOOP_STAGEWISE", None)
        else:
            os.environ["SC_LOOP_STAGEWISE"] = old


def _loop1d_case(phys, ti, mx, prec, lim, chr_, bc):
    rng = np.random.default_rng(43)
    p = make_params(phys, (mx,), time_integrator=ti, mthbc=bc, precision=prec, char_decomp=chr_,
                    lim_type=lim, mthlim=(4, 2, 3, 5))
    q0 = O.new_grid(p)
    interior(p, q0)[:] = _states.random_cells(p, (mx,), rng, smooth=True, amp=0.4)
    cfl = (2.45, 2.5) if ti == 4 else (0.45, 0.5)
    ctl = ScRunCtl(dt_max=1e99, cfl_max=cfl[1], cfl_desired=cfl[0], max_steps=100000, dt_variable=1)
    out = {}
    for mode in ("host", "device"):
        with _make(p, mode) as s:
            s.set_q(q0)
            rec = []
            dt, t = 0.05, 0.0
            for k in range(3):
                dt, st = s.advance(ctl, t, t + 0.01, dt)
                t = st.t
                rec.append((st.steps_taken, st.rejected, st.info, st.t, dt, st.cfl_maxused, st.cfl_last,
                            st.stage_evals))
                rec.append(s.get_q().tobytes())
            out[mode] = rec
    assert out["device"] == out["host"]
    assert out["host"][0][1] > 0 or ti == 4
