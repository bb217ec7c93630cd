"""User hooks behind the Fortran driver, against the reference program
built with the same hook sources (oracle/_ref/xsclaw_ac1du_fmt: the
reference library with the application's bc, src and b4step linked in place
of bc1.f90, src1.f90 and the default b4step, as its applications do).

The application (sharpclaw_amd/fortran/apps/userhooks_ac1d.f90) has a
moving-wall boundary for mthbc(1) = 0, a damping source and a b4step that
negates the velocity once t reaches trtime (the hook patterns of
apps/acoustics/1d/interface/bc1.f95, src1.f90 and
apps/elasticity/stegoton/b4step1.f:19-24).

- mthbc(1) = 0 (with a wall on the right) and a user src: the driver's host-step mode (q in host memory, the
  application's bc and src before every stage, the device's flux1 through
  sc_flux, the SSP combinations of ts_ssp*), for every time integrator;
- a b4step that changes q on a device-resident run: live hooks (q fetched
  before b4step and sent back after it); without them the reversal never
  reaches the device, which the test shows as a large difference;
- host-step mode with the default hooks (the driver's own bc rules and zero
  source from libschooks1.a) against the plain acoustics reference.

Frames agree with the reference's to max(1e-7, 8x the reference's own
change when the initial amplitude moves by one ulp), normwise per frame:
1e-7 is one unit of the eighth printed digit (out1.f90:70) of values in
[1, 10); the second term is the whole-run tolerance of the other parity
tests, since the device's flux rounds differently from the reference's and
some runs amplify rounding (a moving wall that drives waves for the whole
run).  fort.info (steps, dt, CFL at 4 digits) is identical."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

from oracle import refrun as R
from sharpclaw_amd import make_params
from test_fortran_driver import FB, have_driver, have_ref, read_q, setup_case

pytestmark = pytest.mark.gpu

QINIT = ["4", "1.0 200.0 0.5"]  # acoustic pulse p = exp(-200 (x-0.5)^2), u = 0


def run_case(app, ref_app, p, cd, hooks, userhooks, tmp, name="acoustics"):
    """The driver (gpu), the reference (ref) and the reference from an
    initial amplitude one ulp larger (sens) on the same inputs."""
    dirs = {k: os.path.join(tmp, k) for k in ("gpu", "ref", "sens")}
    for k, d in dirs.items():
        os.makedirs(d)
        amp = 1.0 if k != "sens" else float(np.nextafter(1.0, 2.0))
        setup_case(d, p, cd, [QINIT[0], f"{amp!r} {QINIT[1].split(' ', 1)[1]}"], name)
        if userhooks is not None:
            with open(os.path.join(d, "userhooks.data"), "w") as f:
                f.write(" ".join(repr(float(v)) for v in userhooks) + "\n")
    with open(os.path.join(dirs["gpu"], "gpu.data"), "a") as f:
        f.write(f"0\n0.0\n{hooks}\n")
    rg = subprocess.run([os.path.join(FB, f"xsclaw_gpu_{app}")], cwd=dirs["gpu"], capture_output=True,
                        text=True, timeout=300)
    assert rg.returncode == 0, rg.stdout + rg.stderr
    for k in ("ref", "sens"):
        rr = subprocess.run([os.path.join(R.REF, f"xsclaw_{ref_app}_fmt")], cwd=dirs[k], capture_output=True,
                            text=True, timeout=300)
        assert rr.returncode == 0, rr.stdout + rr.stderr
    return dirs["gpu"], dirs["ref"], dirs["sens"]


def nrm(a, b):
    return np.max(np.abs(a - b)) / np.max(np.abs(b))


def frames_agree(dg, dr, ds, nout):
    """Frame 0 byte-identical; returns the largest per-frame normwise
    difference in units of that frame's tolerance (<= 1 passes)."""
    assert open(os.path.join(dg, "fort.q0000")).read() == open(os.path.join(dr, "fort.q0000")).read()
    worst = 0.0
    for k in range(1, nout + 1):
        f = f"fort.q{k:04d}"
        a, b, c = (read_q(os.path.join(d, f)) for d in (dg, dr, ds))
        assert a.shape == b.shape
        worst = max(worst, nrm(a, b) / max(1e-7, 8.0 * nrm(c, b)))
    return worst


@pytest.mark.skipif(not (have_driver("ac1du") and have_ref("ac1du")), reason="driver or reference build missing")
@pytest.mark.parametrize("ti", [1, 2, 3, 4])
def test_user_bc_src_b4step_host_step(ti):
    """mthbc(1) = 0 selects the host-step mode by itself (gpu.data hooks 0)."""
    p = make_params("acoustics", (400,), char_decomp=2, time_integrator=ti, mthbc=(0, 3, 1, 1))
    cfl = (0.5, 0.55) if ti != 4 else (2.45, 2.5)
    cd = R.clawdata_for(p, nout=4, outstyle=1, tfinal=1.0, dt_initial=1e-3, cfl_max=cfl[1],
                        cfl_desired=cfl[0])
    with tempfile.TemporaryDirectory() as tmp:
        dg, dr, ds = run_case("ac1du", "ac1du", p, cd, 0, (0.1, 10.0, 0.3, 0.5), tmp)
        assert frames_agree(dg, dr, ds, 4) <= 1.0
        assert open(os.path.join(dg, "fort.info")).read() == open(os.path.join(dr, "fort.info")).read()


@pytest.mark.skipif(not (have_driver("ac1du") and have_ref("ac1du")), reason="driver or reference build missing")
def test_live_b4step_on_the_device_path():
    """Library boundaries (walls) and no source (alpha = 0): the fused
    device stages run the steps; the b4step reversal at t = 0.3 needs the
    live hooks (gpu.data hooks 1)."""
    p = make_params("acoustics", (500,), char_decomp=2, time_integrator=3, mthbc=(3, 3, 1, 1))
    cd = R.clawdata_for(p, nout=3, outstyle=1, tfinal=0.6, dt_initial=1e-3, cfl_max=0.55,
                        cfl_desired=0.5)
    uh = (0.0, 0.0, 0.0, 0.3)
    with tempfile.TemporaryDirectory() as tmp:
        dg, dr, ds = run_case("ac1du", "ac1du", p, cd, 1, uh, tmp)
        assert frames_agree(dg, dr, ds, 3) <= 1.0
        assert open(os.path.join(dg, "fort.info")).read() == open(os.path.join(dr, "fort.info")).read()
    with tempfile.TemporaryDirectory() as tmp:
        dg, dr, ds = run_case("ac1du", "ac1du", p, cd, 0, uh, tmp)
        # without live hooks the frame before the reversal agrees, the later ones do not
        a, b = read_q(os.path.join(dg, "fort.q0001")), read_q(os.path.join(dr, "fort.q0001"))
        assert nrm(a, b) <= 1e-7
        a, b = read_q(os.path.join(dg, "fort.q0003")), read_q(os.path.join(dr, "fort.q0003"))
        assert nrm(a, b) > 1e-2


@pytest.mark.skipif(not (have_driver("ac1d") and have_ref("ac1d")), reason="driver or reference build missing")
@pytest.mark.parametrize("mthbc", [(3, 3, 1, 1), (2, 2, 1, 1)])
def test_host_step_with_default_hooks(mthbc):
    """gpu.data hooks 2 on the plain acoustics application: the driver's
    default bc (the rules of bc1.f90) and zero src from its hook library."""
    p = make_params("acoustics", (300,), char_decomp=2, time_integrator=1, mthbc=mthbc)
    cd = R.clawdata_for(p, nout=2, outstyle=1, tfinal=0.5, dt_initial=1e-3, cfl_max=0.55,
                        cfl_desired=0.5)
    with tempfile.TemporaryDirectory() as tmp:
        dg, dr, ds = run_case("ac1d", "ac1d", p, cd, 2, None, tmp)
        assert frames_agree(dg, dr, ds, 2) <= 1.0
        assert open(os.path.join(dg, "fort.info")).read() == open(os.path.join(dr, "fort.info")).read()


@pytest.mark.skipif(not (have_driver("eu2d") and have_ref("eu2d")), reason="driver or reference build missing")
def test_host_step_2d_default_hooks():
    """gpu.data hooks 2 on a 2D Euler grid: the driver's default bc (the
    rules of bc2.f90, x ghosts first, corners from y) and zero src on the
    host, flux2 on the device through sc_flux (ghost-row CFL included).
    Square grid: the reference's flux2 hands its y slices to flux1's dummy
    q1d(1-mbc:maxnx+mbc, meqn), which mis-strides them when ny < maxnx."""
    p = make_params("euler", (36, 36), char_decomp=2, time_integrator=3, mthbc=(1, 3, 3, 1))
    rng = np.random.default_rng(17)
    st = np.stack([rng.uniform(0.5, 1.5, 4), rng.uniform(-0.5, 0.5, 4), rng.uniform(-0.5, 0.5, 4),
                   rng.uniform(0.5, 1.5, 4)], axis=1)
    lines = ["1", "0.5 0.5"] + [" ".join(repr(float(v)) for v in row) for row in st]
    cd = R.clawdata_for(p, nout=2, outstyle=1, tfinal=0.05, dt_initial=1e-3, cfl_max=0.5,
                        cfl_desired=0.45)
    with tempfile.TemporaryDirectory() as tmp:
        dg, dr = os.path.join(tmp, "gpu"), os.path.join(tmp, "ref")
        for d in (dg, dr):
            os.makedirs(d)
            setup_case(d, p, cd, lines, "euler")
        with open(os.path.join(dg, "gpu.data"), "a") as f:
            f.write("0\n0.0\n2\n")
        rg = subprocess.run([os.path.join(FB, "xsclaw_gpu_eu2d")], cwd=dg, capture_output=True, text=True,
                            timeout=300)
        assert rg.returncode == 0, rg.stdout + rg.stderr
        rr = subprocess.run([os.path.join(R.REF, "xsclaw_eu2d_fmt")], cwd=dr, capture_output=True, text=True,
                            timeout=300)
        assert rr.returncode == 0, rr.stdout + rr.stderr
        assert open(os.path.join(dg, "fort.q0000")).read() == open(os.path.join(dr, "fort.q0000")).read()
        for f in ("fort.q0001", "fort.q0002"):
            a, b = read_q(os.path.join(dg, f)), read_q(os.path.join(dr, f))
            assert nrm(a, b) <= 1e-11, f  # 2D frames print 16 digits (out2.f90)
        assert open(os.path.join(dg, "fort.info")).read() == open(os.path.join(dr, "fort.info")).read()
