"""sc_flux: flux1 / flux2 of a whole grid on the caller's q, with the ghost
cells exactly as the caller's bc left them (step1d.f90:20-29,
step2d.f90:25-36 with a user bc and src), against the oracle's flux1 /
flux2 on the same arrays.

The ghost cells here are random states that no boundary rule produces, so
every ghost value the device reads must come from the caller's array; dq
enters holding a random source term.  In 2D the ghost rows and columns hold
faster states than the (smooth, gentle) interior, so the CFL number is the one of the
reference's ghost sweeps (flux2.f90:36,68), which the tile kernel cannot see
and sc_flux adds by solving those four slices.  1D: dq = src - dtdx*X in
the reference's order (bitwise expected, checked to 1e-13); 2D: the
reference forms (src + dq_x) + dq_y, the device src + (dq_x + dq_y), so the
bar is the per-step tolerance, 1e-12 normwise relative.  Grids with a user
boundary (mthbc 0) are refused by the fused entry points.
"""
import ctypes as C

import numpy as np
import pytest

import _states
from conftest import rel_linf
from oracle import oracle as O
from sharpclaw_amd import ScRunCtl, Solver, ScError, interior, make_params
from sharpclaw_amd._abi import SC_ERR_UNSUPPORTED

pytestmark = pytest.mark.gpu


def full_random(p, rng, ghost_scale=1.0):
    """A smooth physical state in the interior and independent random
    physical states in every ghost cell, with their velocities scaled by
    ghost_scale (Euler: pressure set to 1 there)."""
    shape = tuple(n + 6 for n in p.nx[:p.ndim])
    q = _states.random_cells(p, shape, rng, smooth=True, amp=0.3)
    g = _states.random_cells(p, shape, rng, amp=0.5)
    ghost = np.ones(shape, dtype=bool)
    ghost[(slice(3, -3),) * p.ndim] = False
    for m in range(p.meqn):
        q[..., m][ghost] = g[..., m][ghost]
    if ghost_scale != 1.0:
        for m in range(1, p.ndim + 1):
            q[..., m][ghost] *= ghost_scale
        if p.physics == 3:
            rho = q[..., 0]
            ke = 0.5 * sum(q[..., m] ** 2 for m in range(1, p.ndim + 1)) / rho
            q[..., -1] = np.where(ghost, ke + 1.0 / (p.gamma - 1.0), q[..., -1])
    return np.asfortranarray(q)


CASES_1D = [("advection", 0), ("acoustics", 0), ("acoustics", 2), ("euler", 0), ("euler", 2),
            ("shallow_water", 2)]


@pytest.mark.parametrize("phys,char", CASES_1D)
@pytest.mark.parametrize("mx", [7, 255, 600])
def test_flux_1d_user_ghosts(phys, char, mx):
    rng = np.random.default_rng(mx * 7 + char)
    p = make_params(phys, (mx,), char_decomp=char, time_integrator=3, mthbc=(0, 0, 1, 1))
    q = full_random(p, rng)
    src = np.asfortranarray(rng.uniform(-1e-2, 1e-2, q.shape))
    dt = 0.2 * p.dx[0]
    want, wcfl = O.flux1(p, 1, mx, q, dt, dq1d=src)
    assert np.isfinite(want).all()
    with Solver(p) as s:
        s.set_q(q)
        dq = src.copy(order="F")
        cfl = s.flux(q, dq, dt, 0.1)
        # the solver's own q is untouched
        assert np.array_equal(interior(p, s.get_q()), interior(p, q))
    err = rel_linf(interior(p, dq), interior(p, want))
    assert err <= 1e-13, err
    assert cfl == pytest.approx(wcfl, rel=1e-13)
    # ghost entries of dq stay the caller's
    ghost = np.ones(dq.shape[0], dtype=bool)
    ghost[3:-3] = False
    assert np.array_equal(dq[ghost], src[ghost])


CASES_2D = [("euler", 2, 2, (64, 48)), ("euler", 0, 2, (40, 40)), ("shallow_water", 2, 2, (33, 70)),
            ("euler", 1, 2, (48, 40)), ("shallow_water", 3, 2, (40, 33)), ("euler", 2, 1, (40, 40))]


@pytest.mark.parametrize("phys,char,lim,nxy", CASES_2D)
@pytest.mark.parametrize("ghost_scale", [1.0, 3.0])
def test_flux_2d_user_ghosts(phys, char, lim, nxy, ghost_scale):
    rng = np.random.default_rng(nxy[0] * 100 + nxy[1] + char)
    kw = dict(mthlim=(1, 1, 1, 1)) if lim == 1 else {}
    p = make_params(phys, nxy, char_decomp=char, lim_type=lim, time_integrator=3, mthbc=(0, 0, 0, 0), **kw)
    q = full_random(p, rng, ghost_scale)
    src = np.asfortranarray(rng.uniform(-1e-2, 1e-2, q.shape))
    dt = 0.1 * min(p.dx[0], p.dx[1])
    d2, wcfl = O.flux2(p, q, dt, with_ghost_sweeps=True)
    _, icfl = O.flux2(p, q, dt, with_ghost_sweeps=False)
    want = src + d2
    assert np.isfinite(want).all()
    with Solver(p) as s:
        s.set_q(q)
        dq = src.copy(order="F")
        cfl = s.flux(q, dq, dt, 0.0)
    err = rel_linf(interior(p, dq), interior(p, want))
    assert err <= 1e-12, err
    assert cfl == pytest.approx(wcfl, rel=1e-12)
    assert wcfl > 1.05 * icfl  # the ghost sweeps set this CFL number


def test_flux_every_tile_shape():
    rng = np.random.default_rng(11)
    p = make_params("euler", (70, 45), char_decomp=2, time_integrator=3, mthbc=(0, 1, 0, 3))
    q = full_random(p, rng, 2.0)
    dt = 0.1 * p.dx[0]
    d2, wcfl = O.flux2(p, q, dt, with_ghost_sweeps=True)
    for tile in range(4):
        with Solver(p) as s:
            s.set_tile(tile)
            dq = np.zeros(q.shape, order="F")
            cfl = s.flux(q, dq, dt)
        assert rel_linf(interior(p, dq), interior(p, d2)) <= 1e-12, tile
        assert cfl == pytest.approx(wcfl, rel=1e-12)


def test_user_bc_refused_by_fused_steps():
    p = make_params("acoustics", (64,), char_decomp=0, time_integrator=3, mthbc=(0, 1, 1, 1))
    rng = np.random.default_rng(2)
    q = full_random(p, rng)
    ctl = ScRunCtl(dt_max=1e99, cfl_max=0.55, cfl_desired=0.5, max_steps=10, dt_variable=1)
    with Solver(p) as s:
        s.set_q(q)
        s.save_qold()
        for call in (lambda: s.ts(1e-3), lambda: s.step(1e-3), lambda: s.advance(ctl, 0.0, 0.01, 1e-3)):
            with pytest.raises(ScError) as e:
                call()
            assert e.value.code == SC_ERR_UNSUPPORTED
            assert "sc_flux" in str(e.value)
