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
