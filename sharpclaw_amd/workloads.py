"""Seeded synthetic inputs of the measured workloads.

dambreak_batch   many 1D shallow-water dam breaks of uneven size: mx
                 log-uniform on [64, 4096], depths h_L ~ U[1, 3], h_R ~ U[0.1, 1],
                 dam at x0 ~ U[0.3, 0.7] on [0, 1], fluid at rest, g = 9.81
dambreak_q       the initial q(1-mbc:mx+mbc, 2) of one of them (cell centres
                 left of x0 take h_L)
"""
import numpy as np


def dambreak_batch(n, seed=2024, mx_lo=64, mx_hi=4096):
    rng = np.random.default_rng(seed)
    mx = np.rint(np.exp(rng.uniform(np.log(mx_lo), np.log(mx_hi), n))).astype(np.int32)
    hl = rng.uniform(1.0, 3.0, n)
    hr = rng.uniform(0.1, 1.0, n)
    x0 = rng.uniform(0.3, 0.7, n)
    return mx, hl, hr, x0


def dambreak_q(mx, hl, hr, x0, mbc=3, xlower=0.0, xupper=1.0):
    dx = (xupper - xlower) / float(mx)
    xc = xlower + (np.arange(1, mx + 1) - 0.5) * dx
    q = np.zeros((mx + 2 * mbc, 2), order="F")
    q[mbc:mbc + mx, 0] = np.where(xc < x0, hl, hr)
    return q


def dambreak_work(mx, cfl=0.45, smax=7.5, tfinal=0.2, stages=3):
    """Rough cell-stage count of one dam break (steps ~ tfinal*smax/(cfl*dx))."""
    mx = np.asarray(mx, dtype=np.float64)
    return stages * mx * np.ceil(tfinal * smax * mx / cfl)


# ---------------------------------------------------------------- served mix
#
# Poisson arrivals (seeded rate) of requests drawn uniformly from three
# kinds, float32 or float64 with equal probability, Gaussian-bump initial
# data with seeded amplitude, centre and width:
#   acoustics1d  rho = 1, K = 4 (c = 2), mx log-uniform [128, 4096] on [0, 1],
#                walls, pressure bump, to t = 0.25
#   euler1d      gamma = 1.4, mx log-uniform [128, 1999] on [0, 1],
#                extrapolation, density/pressure bump at rest, to t = 0.1
#   sw2d         g = 9.81, mx = my log-uniform [128, 1024] on [0, 1]^2,
#                extrapolation, depth bump at rest, to t = 0.05
# All characteristic WENO5, SSPRK3, cfl_desired 0.45 / cfl_max 0.5.

KINDS = ("acoustics1d", "euler1d", "sw2d")
SERVED_CTL = dict(dt_max=1e99, cfl_max=0.5, cfl_desired=0.45, max_steps=1000000, dt_variable=1)


def _loguni(rng, lo, hi):
    return int(np.rint(np.exp(rng.uniform(np.log(lo), np.log(hi)))))


def served_request(kind, rng, precision):
    """(params, q0, tend, dt0) of one request of the given kind."""
    from ._abi import grid_shape, interior, make_params
    amp, x0, w = rng.uniform(0.2, 1.0), rng.uniform(0.3, 0.7), rng.uniform(0.03, 0.1)
    if kind == "acoustics1d":
        mx = _loguni(rng, 128, 4096)
        p = make_params("acoustics", (mx,), mthbc=(3, 3), rho=1.0, bulk=4.0, precision=precision)
        xc = (np.arange(mx) + 0.5) / mx
        q = np.zeros(grid_shape(p), order="F")
        interior(p, q)[:, 0] = amp * np.exp(-((xc - x0) / w) ** 2)
        return p, q, 0.25, 1e-5
    if kind == "euler1d":
        mx = _loguni(rng, 128, 1999)
        p = make_params("euler", (mx,), mthbc=(1, 1), gamma=1.4, precision=precision)
        xc = (np.arange(mx) + 0.5) / mx
        bump = amp * np.exp(-((xc - x0) / w) ** 2)
        q = np.zeros(grid_shape(p), order="F")
        qi = interior(p, q)
        qi[:, 0] = 1.0 + bump
        qi[:, 2] = (1.0 + bump) / 0.4
        return p, q, 0.1, 1e-5
    n = _loguni(rng, 128, 1024)
    y0 = rng.uniform(0.3, 0.7)
    p = make_params("shallow_water", (n, n), mthbc=(1, 1, 1, 1), grav=9.81, precision=precision)
    xc = (np.arange(n) + 0.5) / n
    X, Y = np.meshgrid(xc, xc, indexing="ij")
    q = np.zeros(grid_shape(p), order="F")
    interior(p, q)[..., 0] = 1.0 + amp * np.exp(-((X - x0) ** 2 + (Y - y0) ** 2) / w ** 2)
    return p, q, 0.05, 1e-5


def served_mix(n, seed=2024, rate=200.0):
    """n requests: dicts with arrival (s), kind, precision and the request."""
    from ._abi import SC_FP32, SC_FP64
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, n))
    out = []
    for k in range(n):
        kind = KINDS[rng.integers(0, len(KINDS))]
        prec = SC_FP32 if rng.integers(0, 2) else SC_FP64
        p, q, tend, dt0 = served_request(kind, rng, prec)
        out.append({"arrival": float(t[k]), "kind": kind, "precision": prec, "params": p, "q0": q,
                    "tend": tend, "dt0": dt0})
    return out
