"""The 2D stage's constant-tile path against the oracle.

A tile whose staged cells (tile and halos) are all equal takes the RK
combination with dq = 0 and one interface's speeds for the CFL.  These runs
put a compact bump into an otherwise uniform field, so most tiles are
constant, some only on some stages, and tiles beside a wall see ghost cells
with the normal momentum negated (constant only where that momentum is
zero).  Whole runs are compared with the oracle at the whole-run tolerance
of test_gpu_parity.py; the steps and rejections must agree exactly.
"""
import numpy as np
import pytest

from conftest import rel_linf
from oracle import oracle as O
from sharpclaw_amd import SC_FP32, ScRunCtl, Solver, interior, make_params
from test_gpu_parity import rel_l1, run_tolerance

pytestmark = pytest.mark.gpu

N = 96  # 6 x 7 tiles of 16 x 15, the last row partial


def bump(p, u0, v0):
    """Uniform state with velocity (u0, v0) and a compact density/pressure
    bump (zero outside radius 0.12 around (0.3, 0.6))."""
    x = (np.arange(p.nx[0]) + 0.5) / p.nx[0]
    y = (np.arange(p.nx[1]) + 0.5) / p.nx[1]
    X, Y = np.meshgrid(x, y, indexing="ij")
    r2 = ((X - 0.3) ** 2 + (Y - 0.6) ** 2) / 0.12 ** 2
    b = np.where(r2 < 1.0, (1.0 - r2) ** 3, 0.0)
    q = O.new_grid(p)
    qi = interior(p, q)
    if p.meqn == 4:
        rho, pr = 1.0 + 0.5 * b, 1.0 + 0.8 * b
        qi[..., 0], qi[..., 1], qi[..., 2] = rho, rho * u0, rho * v0
        qi[..., 3] = pr / 0.4 + 0.5 * rho * (u0 * u0 + v0 * v0)
    else:
        h = 1.0 + 0.4 * b
        qi[..., 0], qi[..., 1], qi[..., 2] = h, h * u0, h * v0
    return q


CASES = [
    ("euler", (3, 3, 3, 3), 0.0, 0.0),     # walls, fluid at rest: wall tiles constant
    ("euler", (3, 3, 3, 3), 0.4, 0.0),     # walls, flow normal to the x walls
    ("euler", (2, 2, 2, 2), 0.3, -0.2),    # periodic
    ("euler", (1, 1, 1, 1), -0.5, 0.25),   # extrapolation
    ("shallow_water", (3, 3, 3, 3), 0.0, 0.0),
    ("shallow_water", (1, 1, 1, 1), 0.2, 0.3),
]


@pytest.mark.parametrize("loop", ["host", "device"])
@pytest.mark.parametrize("phys,bc,u0,v0", CASES)
def test_constant_tiles_runs_match_oracle(phys, bc, u0, v0, loop, monkeypatch):
    """loop: the host-driven time loop (stage2d_kernel) or the device-resident
    one (stage2d_batch_kernel + the control kernel)."""
    monkeypatch.setenv("SC_HOST_LOOP", "1" if loop == "host" else "0")
    p = make_params(phys, (N, N), char_decomp=2, time_integrator=3, mthbc=bc)
    q0 = bump(p, u0, v0)
    ctl = ScRunCtl(dt_max=1e99, cfl_max=0.5, cfl_desired=0.45, max_steps=30, dt_variable=1)
    rq, rdt, rst = O.sharpclaw(p, ctl, q0, 0.0, 10.0, 1e-3)
    with Solver(p) as s:
        s.set_tile(0)
        s.set_q(q0)
        gdt, gst = s.advance(ctl, 0.0, 10.0, 1e-3)
        gq = s.get_q()
    assert (gst.steps_taken, gst.rejected) == (rst.steps_taken, rst.rejected)
    assert abs(gdt - rdt) <= 1e-10 * rdt
    tol = run_tolerance(p, ctl, q0, 0.0, 10.0, 1e-3, rq)
    assert rel_linf(interior(p, gq), interior(p, rq)) <= tol


def test_constant_tiles_fp32_l1():
    p = make_params("euler", (N, N), char_decomp=2, time_integrator=3, mthbc=(3, 3, 3, 3))
    q0 = bump(p, 0.4, 0.0)
    ctl = ScRunCtl(dt_max=1e99, cfl_max=0.5, cfl_desired=0.45, max_steps=30, dt_variable=1)
    rq, _, _ = O.sharpclaw(p, ctl, q0, 0.0, 10.0, 1e-3)
    p.precision = SC_FP32
    with Solver(p) as s:
        s.set_q(q0)
        s.advance(ctl, 0.0, 10.0, 1e-3)
        gq = s.get_q()
    assert rel_l1(interior(p, gq), interior(p, rq)) <= 1e-5
