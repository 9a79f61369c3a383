"""Shared helpers of the test suite: grid descriptions for both the HIP
library (mif_amd) and the CPU oracle, and comparison utilities."""
from __future__ import annotations

import numpy as np

import mif_oracle as mo

PI = np.pi


def grids(Ng, size, origin, Re, final_time, nsteps, periodic):
    """(oracle grid, mif_amd GridSpec) of one domain."""
    og = mo.make_grid(Ng, size, origin, Re, final_time, nsteps, periodic)
    try:
        import mif_amd
        gs = mif_amd.GridSpec(int(Ng[0]), int(Ng[1]), int(Ng[2]), float(size[0]), float(size[1]), float(size[2]),
                              float(origin[0]), float(origin[1]), float(origin[2]), float(Re), float(final_time),
                              int(nsteps), tuple(bool(p) for p in periodic))
    except ImportError:
        gs = None
    return og, gs


def rel_err(a: np.ndarray, b: np.ndarray) -> float:
    """max|a - b| / max|b| (0 when both vanish)."""
    den = float(np.max(np.abs(b)))
    num = float(np.max(np.abs(a - b)))
    if den == 0.0:
        return num
    return num / den


def splitmix64(seed: int):
    """SplitMix64 stream (documented host PRNG for synthetic inputs)."""
    mask = (1 << 64) - 1
    state = seed & mask
    while True:
        state = (state + 0x9E3779B97F4A7C15) & mask
        z = state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & mask
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & mask
        yield z ^ (z >> 31)


def seeded_normal(shape, seed: int) -> np.ndarray:
    """N(0,1) samples from numpy's PCG64 with a fixed seed."""
    return np.random.default_rng(seed).standard_normal(shape)


def channel_initial(o: "mo.Oracle", seed: int, amplitude: float = 1e-2):
    """Laminar Poiseuille profile u = 1 - y^2 plus a divergence-free
    perturbation: the discrete curl of a seeded vector potential on cell
    edges times a (1 - y^2)^2 wall envelope, scaled to max |u'| = amplitude.
    The curl is taken on the staggered grid itself, so the perturbation is
    exactly divergence-free under the discrete divergence."""
    g = o.g
    N = list(g.N)
    h = list(g.h)
    # potential A on the pressure-node grid extents (+1 margin), with the
    # envelope evaluated at node y
    rng = np.random.default_rng(seed)
    shp = (N[2] + 1, N[1] + 1, N[0] + 1)
    A = [rng.standard_normal(shp) for _ in range(3)]
    j = np.arange(shp[1])
    y = g.min[1] + h[1] * (g.base[1] + j)
    env = np.clip(1 - y ** 2, 0, None) ** 2
    for a in A:
        a *= env[None, :, None]
    vel = o.velocity()
    ext = [o.extents(s) for s in range(3)]
    # u = dAz/dy - dAy/dz ; v = dAx/dz - dAz/dx ; w = dAy/dx - dAx/dy
    Ax, Ay, Az = A

    def crop(arr, e):
        return arr[: e[2], : e[1], : e[0]]

    def d(arr, axis, hh):
        # forward difference (arr[q+1] - arr[q]) / h along numpy axis
        sl_hi = [slice(None)] * 3
        sl_lo = [slice(None)] * 3
        sl_hi[axis] = slice(1, None)
        sl_lo[axis] = slice(0, -1)
        out = np.zeros_like(arr)
        out[tuple(sl_lo)] = (arr[tuple(sl_hi)] - arr[tuple(sl_lo)]) / hh
        return out

    u = d(Az, 1, h[1]) - d(Ay, 0, h[2])
    v = d(Ax, 0, h[2]) - d(Az, 2, h[0])
    w = d(Ay, 2, h[0]) - d(Ax, 1, h[1])
    comps = [crop(u, ext[0]), crop(v, ext[1]), crop(w, ext[2])]
    m = max(float(np.max(np.abs(c))) for c in comps)
    for c in range(3):
        vel[c][...] = comps[c] * (amplitude / m)
    # base flow on u at its own y coordinates
    ju = np.arange(ext[0][1])
    yu = g.min[1] + h[1] * (g.base[1] + ju)
    vel.u[...] += (1 - yu ** 2)[None, :, None]
    return vel
