"""Frame output (sc_output_begin / sc_output_wait) and restart files
(sc_write_restart / sc_read_restart) through the C ABI.

Frames: the reference's writers reset |q| < 1e-99 to zero in q itself
(src/1d/out1.f90:63-68, src/2d/out2.f90:60-67) and the run continues from
that q; the device does the same to its own q.  Checked bit for bit against
the oracle's flush_small + bc.  The copy to the host overlaps the next
interval's steps: frames and final state equal a synchronous run bitwise.
Restart: stopping at a frame, writing a restart file and continuing in a
new solver reproduces the uninterrupted run bit for bit.
"""
import os

import numpy as np
import pytest

import _states
from oracle import oracle as O
from sharpclaw_amd import ScError, ScRunCtl, Solver, interior, make_params

pytestmark = pytest.mark.gpu

CTL = dict(dt_max=1e99, cfl_max=0.5, cfl_desired=0.45, max_steps=50000, dt_variable=1)


def _tiny(p, rng):
    """A grid of magnitudes 1e-130..1e-70 with random signs, plus NaN,
    +-inf, -0.0 and subnormals."""
    q = O.new_grid(p)
    qi = interior(p, q)
    qi[...] = rng.choice([-1.0, 1.0], qi.shape) * 10.0 ** rng.uniform(-130, -70, qi.shape)
    flat = qi.reshape(-1)
    flat[:6] = [np.nan, np.inf, -np.inf, -0.0, 5e-324, -1e-99]
    return q


@pytest.mark.parametrize("name,shape,bc", [("acoustics", (37,), (1, 3, 1, 1)),
                                           ("euler", (19, 23), (2, 2, 3, 1))])
def test_output_flushes_device_q_bitwise(name, shape, bc):
    rng = np.random.default_rng(41)
    p = make_params(name, shape, mthbc=bc)
    q = _tiny(p, rng)
    want = O.bc(p, O.flush_small(p, q))
    with Solver(p) as s:
        s.set_q(q)
        s.output_begin()
        got = s.output_wait()
        after = s.get_q()
    assert np.array_equal(got, want, equal_nan=True)
    assert np.array_equal(np.signbit(got), np.signbit(want))
    # the solver's own q was reset, not only the copy
    assert np.array_equal(after, want, equal_nan=True)


@pytest.mark.parametrize("name,shape", [("shallow_water", (300,)), ("euler", (48, 48))])
def test_output_overlaps_next_interval(name, shape):
    rng = np.random.default_rng(42)
    p = make_params(name, shape, mthbc=(1, 1, 1, 1))
    q0 = _states.grid(p, rng, smooth=True, amp=0.3)
    ctl = ScRunCtl(**CTL)
    times = [0.0, 0.004, 0.008, 0.012]
    # synchronous: advance, frame, advance, ...
    sync = []
    with Solver(p) as s:
        s.set_q(q0)
        dt = 1e-4
        for k in range(1, len(times)):
            dt, _ = s.advance(ctl, times[k - 1], times[k], dt)
            s.output_begin()
            sync.append(s.output_wait())
    # overlapped: each frame's copy is collected after the next interval ran
    over = []
    with Solver(p) as s:
        s.set_q(q0)
        dt = 1e-4
        for k in range(1, len(times)):
            dt, _ = s.advance(ctl, times[k - 1], times[k], dt)
            if k > 1:
                over.append(s.output_wait())
            s.output_begin()
        over.append(s.output_wait())
    for a, b in zip(sync, over):
        assert np.array_equal(a, b)


def test_output_state_errors():
    p = make_params("advection", (16,), mthbc=(2, 2, 1, 1))
    with Solver(p) as s:
        with pytest.raises(ScError):
            s.output_begin()          # no q yet
        s.set_q(O.new_grid(p))
        with pytest.raises(ScError):
            s.output_wait()           # nothing in flight
        s.output_begin()
        with pytest.raises(ScError):
            s.output_begin()          # previous frame not collected
        s.output_wait()


@pytest.mark.parametrize("name,shape", [("euler", (400,)), ("shallow_water", (40, 40))])
def test_restart_reproduces_uninterrupted_run(name, shape, tmp_path):
    rng = np.random.default_rng(43)
    p = make_params(name, shape, mthbc=(1, 1, 1, 1), time_integrator=4)
    q0 = _states.grid(p, rng, smooth=True, amp=0.3)
    ctl = ScRunCtl(**CTL)
    with Solver(p) as s:
        s.set_q(q0)
        dt, _ = s.advance(ctl, 0.0, 0.005, 1e-4)
        dt, _ = s.advance(ctl, 0.005, 0.01, dt)
        full = s.get_q()
    path = str(tmp_path / "run.restart")
    with Solver(p) as s:
        s.set_q(q0)
        dt, st = s.advance(ctl, 0.0, 0.005, 1e-4)
        s.write_restart(path, st.t, dt)
    assert not os.path.exists(path + ".tmp")
    with Solver(p) as s:
        t, dt = s.read_restart(path)
        assert t == 0.005
        s.advance(ctl, t, 0.01, dt)
        resumed = s.get_q()
    assert np.array_equal(interior(p, resumed), interior(p, full))


def test_restart_rejects_bad_files(tmp_path):
    p = make_params("acoustics", (50,), mthbc=(3, 3, 1, 1))
    path = str(tmp_path / "a.restart")
    with Solver(p) as s:
        s.set_q(_states.grid(p, np.random.default_rng(44), smooth=True))
        s.write_restart(path, 0.25, 1e-3)
        t, dt = s.read_restart(path)
        assert (t, dt) == (0.25, 1e-3)
    raw = bytearray(open(path, "rb").read())
    raw[-5] ^= 0x40
    bad = str(tmp_path / "bad.restart")
    open(bad, "wb").write(bytes(raw))
    short = str(tmp_path / "short.restart")
    open(short, "wb").write(bytes(raw[:-16]))
    with Solver(p) as s:
        for f in (bad, short, str(tmp_path / "missing.restart")):
            with pytest.raises(ScError):
                s.read_restart(f)
    with Solver(make_params("acoustics", (51,), mthbc=(3, 3, 1, 1))) as s:
        with pytest.raises(ScError):
            s.read_restart(path)
