"""Invariants of the oracle that hold by mathematics, chosen so that a
plausible mistake (a dropped term, a wrong sign or index, a transposed
operand) fails one of them.  They pin in particular the project-written
pieces no reference implementation exists for: the Euler and shallow-water
eigenvectors and the 2D Euler Roe solver."""
import numpy as np
import pytest

import _states
from conftest import rel_linf
from oracle import oracle as O
from sharpclaw_amd import ScRunCtl, interior, make_params

PHYS = [("advection", 1), ("acoustics", 1), ("euler", 1), ("shallow_water", 1), ("euler", 2),
        ("shallow_water", 2)]


def flux(p, q, ixy=1):
    """Physical flux f(q) in direction ixy, for cell states q[..., m]."""
    q = np.asarray(q)
    f = np.zeros_like(q)
    if p.physics == 1:
        f[..., 0] = p.u * q[..., 0]
    elif p.physics == 2:  # acoustics: p_t + K u_x = 0, u_t + p_x / rho = 0
        f[..., 0] = p.bulk * q[..., 1]
        f[..., 1] = q[..., 0] / p.rho
    elif p.physics == 3:
        g1 = p.gamma - 1.0
        if p.ndim == 1:
            rho, mu, E = q[..., 0], q[..., 1], q[..., 2]
            u = mu / rho
            pr = g1 * (E - 0.5 * rho * u * u)
            f[..., 0], f[..., 1], f[..., 2] = mu, mu * u + pr, u * (E + pr)
        else:
            a, b = (1, 2) if ixy == 1 else (2, 1)
            rho, mn, mt, E = q[..., 0], q[..., a], q[..., b], q[..., 3]
            un = mn / rho
            pr = g1 * (E - 0.5 * (mn * mn + mt * mt) / rho)
            f[..., 0], f[..., a], f[..., b], f[..., 3] = mn, mn * un + pr, mt * un, un * (E + pr)
    else:
        g = p.grav
        if p.ndim == 1:
            h, hu = q[..., 0], q[..., 1]
            f[..., 0], f[..., 1] = hu, hu * hu / h + 0.5 * g * h * h
        else:
            a, b = (1, 2) if ixy == 1 else (2, 1)
            h, hn, ht = q[..., 0], q[..., a], q[..., b]
            f[..., 0], f[..., a], f[..., b] = hn, hn * hn / h + 0.5 * g * h * h, hn * ht / h
    return f


@pytest.mark.parametrize("name,ndim", PHYS)
def test_roe_property(name, ndim):
    """amdq + apdq = sum_w s*wave = f(ql(i)) - f(qr(i-1)) (exact for Roe solvers);
    the entropy fix only moves the split between amdq and apdq."""
    rng = np.random.default_rng(1)
    p = make_params(name, (50,) if ndim == 1 else (50, 50))
    L = 56
    for ixy in ([1] if ndim == 1 else [1, 2]):
        ql = _states.random_cells(p, (L,), rng)
        qr = _states.random_cells(p, (L,), rng)
        wave, s, am, ap = O.rp(p, ixy, 50, ql, qr)
        df = flux(p, ql[1:], ixy) - flux(p, qr[:-1], ixy)
        assert rel_linf((am + ap)[1:], df) <= 1e-12
        sw = np.einsum("iw,imw->im", s, wave)
        assert rel_linf(sw[1:], df) <= 1e-12


@pytest.mark.parametrize("name,ndim", PHYS)
def test_eigenvectors(name, ndim):
    """L R = I, and R diag(s) L (qR - qL) = f(qR) - f(qL) with the Roe speeds:
    the evec plugin returns the eigenstructure of the Roe matrix."""
    rng = np.random.default_rng(2)
    p = make_params(name, (1,) if ndim == 1 else (1, 1), efix=0)
    M = p.meqn
    for ixy in ([1] if ndim == 1 else [1, 2]):
        for _ in range(20):
            qa = _states.random_cells(p, (1,), rng)[0]
            qb = _states.random_cells(p, (1,), rng)[0]
            q = np.zeros((7, M), order="F")  # mx = 1: cells -2..4, interface i=1 between 0 and 1
            q[2], q[3] = qa, qb
            evl, evr = O.evec(p, ixy, 1, q)
            Lm, Rm = evl[3], evr[3]
            assert np.allclose(Lm @ Rm, np.eye(M), atol=1e-13)
            if name in ("euler", "shallow_water", "acoustics", "advection"):
                ql = np.zeros((7, M), order="F")
                qr = np.zeros((7, M), order="F")
                qr[2], ql[3] = qa, qb  # Riemann problem i=1: left qr(0), right ql(1)
                ql[:3] = qa
                qr[:3] = qa
                ql[4:] = qb
                qr[3:] = qb
                _, s, _, _ = O.rp(p, ixy, 1, ql, qr)
                sp = s[3]
                # fields of the 2D Euler evec: (u-c, entropy, shear, u+c); the
                # Riemann solver carries entropy and shear in one wave of speed u
                lam = np.array([sp[0], sp[1], sp[1], sp[2]]) if M == 4 else np.array(sp)
                got = Rm @ (lam * (Lm @ (qb - qa)))
                want = flux(p, qb, ixy) - flux(p, qa, ixy)
                assert rel_linf(got, want) <= 1e-11


@pytest.mark.parametrize("name", ["euler", "shallow_water"])
def test_rotational_invariance_2d(name):
    """rp with ixy=2 on states with the momenta swapped equals rp with ixy=1,
    outputs swapped back."""
    rng = np.random.default_rng(3)
    p = make_params(name, (30, 30))
    L = 36
    ql = _states.random_cells(p, (L,), rng)
    qr = _states.random_cells(p, (L,), rng)
    sw = [0, 2, 1] + ([3] if p.meqn == 4 else [])
    w1, s1, am1, ap1 = O.rp(p, 1, 30, ql, qr)
    w2, s2, am2, ap2 = O.rp(p, 2, 30, ql[:, sw], qr[:, sw])
    assert np.array_equal(s1, s2)
    assert np.array_equal(am1, am2[:, sw]) and np.array_equal(ap1, ap2[:, sw])
    assert np.array_equal(w1, w2[:, sw, :])


def test_euler2d_reduces_to_rp1eu():
    """With v = 0 the 2D Euler Roe solver (normal components 1, mu, 4)
    reproduces the reference's 1D solver rp1eu (to rounding of a*a vs a2)."""
    rng = np.random.default_rng(4)
    p1 = make_params("euler", (40,))
    p2 = make_params("euler", (40, 40))
    L = 46
    a = _states.random_cells(p1, (L,), rng)
    b = _states.random_cells(p1, (L,), rng)
    A = np.zeros((L, 4), order="F")
    B = np.zeros((L, 4), order="F")
    A[:, [0, 1, 3]] = a
    B[:, [0, 1, 3]] = b
    w1, s1, am1, ap1 = O.rp(p1, 1, 40, a, b)
    w2, s2, am2, ap2 = O.rp(p2, 1, 40, A, B)
    assert rel_linf(s2[1:], s1[1:]) <= 1e-14
    assert rel_linf(am2[1:, [0, 1, 3]], am1[1:]) <= 1e-12
    assert rel_linf(ap2[1:, [0, 1, 3]], ap1[1:]) <= 1e-12
    assert np.max(np.abs(am2[1:, 2])) == 0.0 and np.max(np.abs(ap2[1:, 2])) == 0.0


def test_weno_constant_and_quadratic():
    """Constant data: ql = qr = the constant; cell averages of a quadratic:
    every substencil is exact, so ql(i), qr(i-1) equal the point value at
    x_{i-1/2}."""
    mx2 = 40
    q = np.full((mx2, 1), 0.7, order="F")
    ql, qr = O.weno5(q)
    assert np.allclose(ql[3:mx2 - 2], 0.7, rtol=1e-15, atol=0)
    h = 1.0 / 32
    xe = np.arange(mx2 + 1) * h  # cell edges
    F = lambda x: 0.3 * x + 0.5 * x ** 2 - 0.8 * x ** 3 / 3  # antiderivative of 0.3 + x - 0.8 x^2
    qa = ((F(xe[1:]) - F(xe[:-1])) / h).reshape(-1, 1)
    ql, qr = O.weno5(np.asfortranarray(qa))
    I = np.arange(4, mx2 - 1)  # 1-based interfaces I = 4..mx2-2, at x = (I-1)h
    fx = 0.3 + xe[I - 1] - 0.8 * xe[I - 1] ** 2
    assert np.allclose(ql[I - 1, 0], fx, rtol=0, atol=1e-13)
    assert np.allclose(qr[I - 2, 0], fx, rtol=0, atol=1e-13)


def test_weno_fifth_order():
    """Interface values of a smooth monotone profile converge at 5th order."""
    errs = []
    for n in (40, 80, 160):
        h = 1.0 / n
        xe = (np.arange(n + 7) - 3) * h
        qa = ((np.exp(xe[1:]) - np.exp(xe[:-1])) / h).reshape(-1, 1)
        ql, qr = O.weno5(np.asfortranarray(qa))
        I = np.arange(4, n + 5)
        errs.append(max(np.max(np.abs(ql[I - 1, 0] - np.exp(xe[I - 1]))),
                        np.max(np.abs(qr[I - 2, 0] - np.exp(xe[I - 1])))))
    r1, r2 = errs[0] / errs[1], errs[1] / errs[2]
    assert 24 < r1 < 40 and 24 < r2 < 40, errs


def test_weno_mirror_symmetry():
    rng = np.random.default_rng(5)
    q = rng.uniform(-1, 1, (50, 2))
    ql, qr = O.weno5(np.asfortranarray(q))
    qlm, qrm = O.weno5(np.asfortranarray(q[::-1]))
    # interface I of the mirrored data is interface mx2+2-I of the original
    mx2 = 50
    for I in range(4, mx2 - 1):
        J = mx2 + 2 - I
        assert abs(qrm[I - 2, 0] - ql[J - 1, 0]) <= 1e-14
        assert abs(qlm[I - 1, 1] - qr[J - 2, 1]) <= 1e-14


def test_weno_char_identity_eigenvectors():
    """With L = R = I the characteristic reconstruction is the component-wise one."""
    rng = np.random.default_rng(6)
    q = np.asfortranarray(rng.uniform(-1, 1, (40, 3)))
    ev = np.zeros((40, 3, 3), order="F")
    ev[:, range(3), range(3)] = 1.0
    a, b = O.weno5(q)
    c, d = O.weno5_char(q, ev, ev)
    assert rel_linf(c[3:38], a[3:38]) <= 1e-14 and rel_linf(d[2:37], b[2:37]) <= 1e-14


@pytest.mark.parametrize("name", ["euler", "shallow_water"])
@pytest.mark.parametrize("char", [0, 2])
def test_ghost_row_cfl_equals_interior_cfl(name, char):
    """The reference's 2D CFL includes the x-sweeps of ghost rows 0, ny+1 and
    the y-sweeps of ghost columns 0, nx+1 (flux2.f90:36,68).  For boundary
    options 1-3 those repeat the speeds of the adjacent interior row/column,
    so sweeping only the interior gives the same number (the device path
    relies on this)."""
    rng = np.random.default_rng(7)
    for bc in ((1, 1, 1, 1), (2, 2, 2, 2), (3, 3, 3, 3), (1, 3, 2, 2), (3, 1, 1, 3)):
        p = make_params(name, (14, 11), char_decomp=char, mthbc=bc)
        q = O.bc(p, _states.grid(p, rng, smooth=True, amp=0.5))
        _, c_all = O.flux2(p, q, 1e-3, True)
        _, c_int = O.flux2(p, q, 1e-3, False)
        assert c_all == c_int, bc


@pytest.mark.parametrize("ti,order", [(1, 2), (2, 2), (3, 3), (4, 4)])
def test_ssp_temporal_order(ti, order):
    """Fixed-dt runs of smooth periodic advection: differences between the
    dt, dt/2, dt/4 solutions shrink by 2**order (the spatial error is the
    same on one grid)."""
    p = make_params("advection", (64,), char_decomp=0, time_integrator=ti, mthbc=(2, 2, 2, 2))
    x = (np.arange(64) + 0.5) / 64
    q0 = O.new_grid(p)
    interior(p, q0)[:, 0] = np.sin(2 * np.pi * x) + 0.5 * np.cos(4 * np.pi * x)
    sols = []
    base = 0.4 / 64 * (2.0 if ti == 4 else 1.0)
    for k in range(3):
        dt = base / 2 ** k
        ctl = ScRunCtl(dt_max=1e99, cfl_max=10.0, cfl_desired=1.0, max_steps=100000, dt_variable=0)
        q, _, st = O.sharpclaw(p, ctl, q0, 0.0, 0.25, dt)
        assert st.info == 0
        sols.append(interior(p, q)[:, 0])
    r = np.max(np.abs(sols[0] - sols[1])) / np.max(np.abs(sols[1] - sols[2]))
    assert 0.7 * 2 ** order < r < 1.45 * 2 ** order, r


@pytest.mark.parametrize("name,ndim", [("acoustics", 1), ("shallow_water", 1), ("euler", 2),
                                       ("shallow_water", 2)])
def test_conservation_periodic(name, ndim):
    rng = np.random.default_rng(8)
    p = make_params(name, (120,) if ndim == 1 else (24, 24), char_decomp=2, time_integrator=4,
                    mthbc=(2, 2, 2, 2))
    q = _states.grid(p, rng, smooth=True, amp=0.5)
    s0 = O.sums(p, q)
    ctl = ScRunCtl(dt_max=1e99, cfl_max=1.5, cfl_desired=1.2, max_steps=30, dt_variable=1)
    q1, _, _ = O.sharpclaw(p, ctl, q, 0.0, 10.0, 1e-4)
    s1 = O.sums(p, q1)
    assert np.allclose(s1, s0, rtol=1e-13, atol=1e-13 * np.max(np.abs(s0)))


def test_sharpclaw_info_codes_and_retake():
    """Failure handling of the time loop (sharpclaw.f90:61-64, 195-216, 225)."""
    p = make_params("acoustics", (100,), char_decomp=2, time_integrator=3, mthbc=(3, 3, 1, 1))
    rng = np.random.default_rng(9)
    q = _states.grid(p, rng, smooth=True)
    # cfl_desired > cfl_max: info 3, nothing done
    ctl = ScRunCtl(dt_max=1e99, cfl_max=0.5, cfl_desired=0.6, max_steps=100, dt_variable=1)
    q1, _, st = O.sharpclaw(p, ctl, q, 0.0, 1.0, 1e-3)
    assert st.info == 3 and st.steps_taken == 0 and np.array_equal(q1, q)
    # too few steps to reach tend: info 11
    ctl = ScRunCtl(dt_max=1e99, cfl_max=0.5, cfl_desired=0.4, max_steps=3, dt_variable=1)
    _, _, st = O.sharpclaw(p, ctl, q, 0.0, 1.0, 1e-3)
    assert st.info == 11 and st.steps_taken == 3
    # dt_initial far too large: the first attempt is rejected and retaken
    ctl = ScRunCtl(dt_max=1e99, cfl_max=0.5, cfl_desired=0.4, max_steps=100, dt_variable=1)
    _, dtl, st = O.sharpclaw(p, ctl, q, 0.0, 0.05, 0.5)
    assert st.rejected >= 1 and st.info == 0 and st.t == 0.05 and st.cfl_last <= 0.5
