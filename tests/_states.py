"""Seeded physical states for the parity tests (test infrastructure)."""
import numpy as np

from sharpclaw_amd._abi import (SC_PHYS_ACOUSTICS, SC_PHYS_ADVECTION, SC_PHYS_EULER,
                                SC_PHYS_SHALLOW_WATER)


def random_cells(p, shape, rng, smooth=False, amp=1.0):
    """Conserved variables of physical states, shape + (meqn,), Fortran order.

    smooth=True superposes a few sine modes (smooth WENO regime); otherwise
    cell values are independent (every stencil non-smooth).  amp < 1 shrinks
    every range about its midpoint (gentle data for very small grids)."""
    shape = tuple(shape)

    def field(lo, hi):
        mid = 0.5 * (lo + hi)
        lo, hi = mid - amp * (mid - lo), mid + amp * (hi - mid)
        if not smooth:
            return rng.uniform(lo, hi, size=shape)
        grids = np.meshgrid(*[np.linspace(0, 1, n) for n in shape], indexing="ij")
        f = np.zeros(shape)
        for _ in range(3):
            k = rng.uniform(1, 4, size=len(shape))
            ph = rng.uniform(0, 2 * np.pi)
            f += np.sin(sum(2 * np.pi * kk * g for kk, g in zip(k, grids)) + ph)
        f = (f - f.min()) / max(f.max() - f.min(), 1e-300)
        return lo + (hi - lo) * f

    meqn = p.meqn
    q = np.zeros(shape + (meqn,), order="F")
    if p.physics == SC_PHYS_ADVECTION:
        q[..., 0] = field(-1.0, 2.0)
    elif p.physics == SC_PHYS_ACOUSTICS:
        q[..., 0] = field(-1.0, 1.0)
        q[..., 1] = field(-1.0, 1.0)
    elif p.physics == SC_PHYS_EULER:
        g1 = p.gamma - 1.0
        rho = field(0.3, 2.0)
        u = field(-1.5, 1.5)
        pr = field(0.3, 2.0)
        if meqn == 3:
            q[..., 0] = rho
            q[..., 1] = rho * u
            q[..., 2] = pr / g1 + 0.5 * rho * u * u
        else:
            v = field(-1.5, 1.5)
            q[..., 0] = rho
            q[..., 1] = rho * u
            q[..., 2] = rho * v
            q[..., 3] = pr / g1 + 0.5 * rho * (u * u + v * v)
    elif p.physics == SC_PHYS_SHALLOW_WATER:
        h = field(0.3, 2.0)
        q[..., 0] = h
        q[..., 1] = h * field(-1.5, 1.5)
        if meqn == 3:
            q[..., 2] = h * field(-1.5, 1.5)
    return q


def grid(p, rng, smooth=False, amp=1.0):
    if p.ndim == 1:
        return random_cells(p, (p.nx[0] + 2 * p.mbc,), rng, smooth, amp)
    return random_cells(p, (p.nx[0] + 2 * p.mbc, p.nx[1] + 2 * p.mbc), rng, smooth, amp)
