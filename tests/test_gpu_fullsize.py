"""The fused 2D stage at the benchmark's full size (2D Euler 1024 x 1024,
characteristic WENO5 + Roe, SSPRK3), checked through properties that do not
depend on the grid size and hold over many steps (the per-step comparison
with the oracle at this size is test_gpu_fullsize_oracle.py; the oracle
takes about 3 s per step there, so long runs are checked this way).

  transpose   the problem with x and y exchanged (q(i,j) -> q(j,i), momenta
              swapped) gives the transposed result: the y sweep of one is the
              x sweep of the other, interface by interface, with the same
              operations (bitwise in practice; asserted to 1e-13 so that a
              different contraction choice of the compiler in the two
              sweeps would not fail it, while any index or direction mistake
              does)
  periodic    with periodic boundaries sum q dx dy of a smooth state is
              conserved to rounding (the fluctuations telescope)
  float32     the float32 option stays within relative L1 1e-5 of the
              float64 path
The first two also for the reconstructions that run as x and y slice
sweeps.
"""
import numpy as np
import pytest

from conftest import rel_linf

from sharpclaw_amd import SC_FP32, ScRunCtl, Solver, interior, make_params

pytestmark = pytest.mark.gpu

N = 1024


def quadrants(p, seed=961):
    """Four constant states (rho, u, v, p) in U[0.1,1.5] x U[-1,1]^2 x
    U[0.03,1.5], split at x = y = 0.5 (the benchmark's data: seed 961 stays
    physical over its 200 steps; seed 2024's low-pressure quadrant gives a
    negative pressure at the centre in step 23, in the oracle as well)."""
    rng = np.random.default_rng(seed)
    st = np.stack([rng.uniform(0.1, 1.5, 4), rng.uniform(-1, 1, 4), rng.uniform(-1, 1, 4),
                   rng.uniform(0.03, 1.5, 4)], axis=1)
    x = (np.arange(p.nx[0]) + 0.5) / p.nx[0]
    y = (np.arange(p.nx[1]) + 0.5) / p.nx[1]
    X, Y = np.meshgrid(x, y, indexing="ij")
    k = (X > 0.5).astype(int) + 2 * (Y > 0.5).astype(int)
    rho, u, v, pr = (st[k, c] for c in range(4))
    q = np.zeros((p.nx[0] + 6, p.nx[1] + 6, 4), order="F")
    qi = interior(p, q)
    qi[..., 0], qi[..., 1], qi[..., 2] = rho, rho * u, rho * v
    qi[..., 3] = pr / 0.4 + 0.5 * rho * (u * u + v * v)
    return q


def transposed(q):
    t = np.asfortranarray(np.transpose(q, (1, 0, 2)))
    t[..., [1, 2]] = t[..., [2, 1]]
    return np.asfortranarray(t)


def run(p, q0, steps, tile=-1):
    ctl = ScRunCtl(dt_max=1e99, cfl_max=0.5, cfl_desired=0.45, max_steps=steps, dt_variable=1)
    with Solver(p) as s:
        s.set_tile(tile)
        s.set_q(q0)
        dt, st = s.advance(ctl, 0.0, 1e9, 1e-4)
        return s.get_q(), dt, st, s.sums()


def test_transpose_symmetry():
    p = make_params("euler", (N, N), char_decomp=2, time_integrator=3, mthbc=(1, 1, 1, 1))
    q0 = quadrants(p)
    qa, dta, sta, _ = run(p, q0, 10)
    qb, dtb, stb, _ = run(p, transposed(q0), 10)
    assert sta.steps_taken == stb.steps_taken == 10 and abs(dta - dtb) <= 1e-13 * dta
    assert np.all(np.isfinite(interior(p, qa)))
    a, b = interior(p, transposed(qa)), interior(p, qb)
    print("transpose: bitwise equal" if np.array_equal(a, b) else f"transpose: rel L-inf {rel_linf(a, b):.3e}")
    assert rel_linf(a, b) <= 1e-13


def smooth_periodic(p):
    """A smooth periodic state (the four-quadrant data wrapped around a
    periodic box drives the pressure negative within a few steps, in the
    oracle as well)."""
    x = (np.arange(p.nx[0]) + 0.5) / p.nx[0]
    y = (np.arange(p.nx[1]) + 0.5) / p.nx[1]
    X, Y = np.meshgrid(x, y, indexing="ij")
    tp = 2 * np.pi
    rho = 1.0 + 0.5 * np.sin(tp * X) * np.cos(tp * Y)
    u = 0.5 * np.sin(tp * Y)
    v = 0.5 * np.cos(tp * X)
    pr = 1.0 + 0.2 * np.cos(tp * (X + Y))
    q = np.zeros((p.nx[0] + 6, p.nx[1] + 6, 4), order="F")
    qi = interior(p, q)
    qi[..., 0], qi[..., 1], qi[..., 2] = rho, rho * u, rho * v
    qi[..., 3] = pr / 0.4 + 0.5 * rho * (u * u + v * v)
    return q


def test_periodic_conservation():
    p = make_params("euler", (N, N), char_decomp=2, time_integrator=3, mthbc=(2, 2, 2, 2))
    q0 = smooth_periodic(p)
    s0 = interior(p, q0).sum(axis=(0, 1)) * p.dx[0] * p.dx[1]
    _, _, st, s1 = run(p, q0, 10)
    assert st.steps_taken == 10
    scale = np.abs(interior(p, q0)).sum(axis=(0, 1)) * p.dx[0] * p.dx[1]
    assert np.all(np.abs(s1 - s0) <= 1e-12 * scale), (s1 - s0) / scale


def test_fp32_option_l1_fullsize():
    p = make_params("euler", (N, N), char_decomp=2, time_integrator=3, mthbc=(1, 1, 1, 1))
    q0 = quadrants(p)
    q64, _, _, _ = run(p, q0, 20)
    p32 = make_params("euler", (N, N), char_decomp=2, time_integrator=3, mthbc=(1, 1, 1, 1),
                      precision=SC_FP32)
    q32, _, _, _ = run(p32, q0, 20)
    a, b = interior(p, q32), interior(p, q64)
    err = np.sum(np.abs(a - b)) / np.sum(np.abs(b))
    print("1024^2 float32 relative L1 vs float64:", err)
    assert err <= 1e-5


def test_constant_state_constant_tiles():
    """A uniform state: every tile is constant (the stage's shortcut), q stays
    uniform and equal to q0 up to the rounding of the RK combinations
    (0.75 q + 0.25 q is q only to an ulp), and dt follows the CFL of the
    state's speeds max(|u| + c, |v| + c) (cfl_desired * dx / smax)."""
    p = make_params("euler", (N, N), char_decomp=2, time_integrator=3, mthbc=(1, 1, 1, 1))
    rho, u, v, pr = 0.8, 0.3, -0.6, 1.1
    q0 = np.zeros((N + 6, N + 6, 4), order="F")
    qi = interior(p, q0)
    qi[..., 0], qi[..., 1], qi[..., 2] = rho, rho * u, rho * v
    qi[..., 3] = pr / 0.4 + 0.5 * rho * (u * u + v * v)
    q, dt, st, _ = run(p, q0, 3)
    assert st.steps_taken == 3
    qi1 = interior(p, q)
    assert np.all(qi1 == qi1[0, 0, :])
    assert rel_linf(qi1, interior(p, q0)) <= 1e-15
    c = np.sqrt(1.4 * pr / rho)
    want = 0.45 * p.dx[0] / max(abs(u) + c, abs(v) + c)
    assert abs(dt - want) <= 1e-13 * want, (dt, want)


# The reconstructions that run as x and y slice sweeps (weno5_wave,
# weno5_trans, tvd2_wave, tvd2_char): the two sweeps are different kernels
# (one row of 128 interfaces per block against 4 columns x 64), so the
# transposed problem runs every interface through the other kernel.
SWEEPS = [(2, 1), (2, 3), (1, 1), (1, 2)]


def quadrants_mild(p, seed=8):
    """Four constant states with rho, p in U[0.5, 1.5] and u, v in
    U[-0.5, 0.5], split at x = y = 0.5 (the wave-based reconstructions give
    NaN at strong shocks, in the reference as well)."""
    rng = np.random.default_rng(seed)
    st = np.stack([rng.uniform(0.5, 1.5, 4), rng.uniform(-0.5, 0.5, 4), rng.uniform(-0.5, 0.5, 4),
                   rng.uniform(0.5, 1.5, 4)], axis=1)
    x = (np.arange(p.nx[0]) + 0.5) / p.nx[0]
    y = (np.arange(p.nx[1]) + 0.5) / p.nx[1]
    X, Y = np.meshgrid(x, y, indexing="ij")
    k = (X > 0.5).astype(int) + 2 * (Y > 0.5).astype(int)
    rho, u, v, pr = (st[k, c] for c in range(4))
    q = np.zeros((p.nx[0] + 6, p.nx[1] + 6, 4), order="F")
    qi = interior(p, q)
    qi[..., 0], qi[..., 1], qi[..., 2] = rho, rho * u, rho * v
    qi[..., 3] = pr / 0.4 + 0.5 * rho * (u * u + v * v)
    return q


@pytest.mark.parametrize("lim,char", SWEEPS)
def test_sweeps_transpose_symmetry(lim, char):
    p = make_params("euler", (N, N), char_decomp=char, lim_type=lim, mthlim=(4, 4, 4, 4),
                    time_integrator=3, mthbc=(3, 3, 3, 3))
    q0 = quadrants_mild(p)
    qa, dta, sta, _ = run(p, q0, 10)
    qb, dtb, stb, _ = run(p, transposed(q0), 10)
    assert sta.steps_taken == stb.steps_taken == 10 and abs(dta - dtb) <= 1e-13 * dta
    assert np.all(np.isfinite(interior(p, qa)))
    a, b = interior(p, transposed(qa)), interior(p, qb)
    print(f"sweeps {lim}/{char} transpose: " +
          ("bitwise equal" if np.array_equal(a, b) else f"rel L-inf {rel_linf(a, b):.3e}"))
    assert rel_linf(a, b) <= 1e-13


@pytest.mark.parametrize("lim,char", SWEEPS)
def test_sweeps_periodic_conservation(lim, char):
    p = make_params("euler", (N, N), char_decomp=char, lim_type=lim, mthlim=(4, 4, 4, 4),
                    time_integrator=3, mthbc=(2, 2, 2, 2))
    q0 = smooth_periodic(p)
    s0 = interior(p, q0).sum(axis=(0, 1)) * p.dx[0] * p.dx[1]
    _, _, st, s1 = run(p, q0, 10)
    assert st.steps_taken == 10
    scale = np.abs(interior(p, q0)).sum(axis=(0, 1)) * p.dx[0] * p.dx[1]
    assert np.all(np.abs(s1 - s0) <= 1e-12 * scale), (s1 - s0) / scale


@pytest.mark.parametrize("n", [1024, 2048])
def test_tile_orders_bitwise(n, monkeypatch):
    """The order in which the workgroups take the tiles (SC_TILE_ORDER: 0
    bands of tiles per XCD; default the adaptive order, whose tile list of
    the non-constant tiles is rebuilt after every step, by one block up to
    16384 tiles and by two kernels over many blocks above) changes no
    result: every tile computes the same values wherever it runs."""
    p = make_params("euler", (n, n), char_decomp=2, time_integrator=3, mthbc=(1, 1, 1, 1))
    q0 = quadrants(p)
    out = []
    for order in ("0", "1"):
        monkeypatch.setenv("SC_TILE_ORDER", order)
        q, dt, st, _ = run(p, q0, 6)
        assert st.steps_taken == 6
        out.append((q, dt))
    assert np.array_equal(out[0][0], out[1][0]) and out[0][1] == out[1][1]
