"""Synthetic problem set-ups of the measured configurations.

The channel: periodic x and z, no-slip walls at y = -1 and y = 1, laminar
Poiseuille profile u = 1 - y^2 plus a seeded divergence-free perturbation
(the discrete curl, on the staggered grid itself, of a Gaussian random vector
potential that is periodic in x and z and damped by a (1 - y^2)^2 wall
envelope), scaled so that max |u'| equals the requested amplitude.  Inputs
are generated on the host in FP64 with numpy's PCG64 from a fixed seed, so
every consumer sees identical bytes.
"""
from __future__ import annotations

import math

import numpy as np

from .core import GridSpec


def channel_grid(nx: int, ny: int, nz: int, Re: float = 1000.0, dt: float = 5e-4, steps: int = 50,
                 lengths=(2 * math.pi, 2.0, math.pi)) -> GridSpec:
    """Channel with nx, ny, nz intervals per axis (N_g = n + 1 nodes, the
    periodic end node included)."""
    return GridSpec(nx + 1, ny + 1, nz + 1, lengths[0], lengths[1], lengths[2], 0.0, -1.0, 0.0, Re, dt * steps, steps,
                    (True, False, True))


def node_extents(g: GridSpec):
    """Nx, Ny, Nz of the reference at one rank (src/Constants.cpp:80-82)."""
    Ng = (g.Nx_global, g.Ny_global, g.Nz_global)
    return tuple(Ng[a] + (1 if g.periodic[a] else 0) for a in range(3))


def field_extents(g: GridSpec, stag: int):
    """Extents of a staggered field (src/StaggeredTensor.cpp:5-9)."""
    e = list(node_extents(g))
    if stag == 0:
        e[0] = e[0] if g.periodic[0] else e[0] + 1
    elif stag in (1, 2):
        e[stag] += 1
    return tuple(e)


def spacing(g: GridSpec):
    Ng = (g.Nx_global, g.Ny_global, g.Nz_global)
    L = (g.x_size, g.y_size, g.z_size)
    return tuple(L[a] / (Ng[a] - 1) for a in range(3))


def channel_initial_velocity(g: GridSpec, seed: int = 20261202, amplitude: float = 1e-2):
    """(u, v, w) numpy arrays in the reference layout (shape (sz, sy, sx))."""
    h = spacing(g)
    N = node_extents(g)
    Ng = (g.Nx_global, g.Ny_global, g.Nz_global)
    base = tuple(-1 if g.periodic[a] else 0 for a in range(3))
    mins = (g.min_x, g.min_y, g.min_z)
    rng = np.random.default_rng(seed)
    # unique potential values: periodic axes have Ng-1 distinct nodes
    uniq = tuple(Ng[a] - 1 if g.periodic[a] else Ng[a] + 2 for a in range(3))
    A = [rng.standard_normal((uniq[2], uniq[1], uniq[0])) for _ in range(3)]
    # index maps from array index (0..N) to the unique potential index
    maps = []
    for a in range(3):
        idx = np.arange(N[a] + 2)
        if g.periodic[a]:
            maps.append((idx + base[a]) % uniq[a])
        else:
            maps.append(np.minimum(idx, uniq[a] - 1))
    jy = np.arange(N[1] + 2)
    y = mins[1] + h[1] * (base[1] + jy)
    env = np.clip(1.0 - y ** 2, 0.0, None) ** 2
    P = [a[np.ix_(maps[2], maps[1], maps[0])] * env[None, :, None] for a in A]
    Ax, Ay, Az = P

    def dfwd(arr, axis, hh):
        hi = [slice(None)] * 3
        lo = [slice(None)] * 3
        hi[axis] = slice(1, None)
        lo[axis] = slice(0, -1)
        out = np.zeros_like(arr)
        out[tuple(lo)] = (arr[tuple(hi)] - arr[tuple(lo)]) / hh
        return out

    # numpy axes: 0 = z, 1 = y, 2 = x
    u = dfwd(Az, 1, h[1]) - dfwd(Ay, 0, h[2])
    v = dfwd(Ax, 0, h[2]) - dfwd(Az, 2, h[0])
    w = dfwd(Ay, 2, h[0]) - dfwd(Ax, 1, h[1])
    out = []
    for stag, comp in enumerate((u, v, w)):
        e = field_extents(g, stag)
        out.append(np.ascontiguousarray(comp[: e[2], : e[1], : e[0]]))
    m = max(float(np.max(np.abs(c))) for c in out)
    out = [c * (amplitude / m) for c in out]
    eu = field_extents(g, 0)
    yu = mins[1] + h[1] * (base[1] + np.arange(eu[1]))
    out[0] = out[0] + (1.0 - yu ** 2)[None, :, None]
    return tuple(out)


def adjoint_channel(n: int = 128, Re: float = 500.0, dt: float = 1e-3, steps: int = 32):
    """The channel of the gradient workload: (grid, initial velocity, the
    seed-perturbed initial velocity whose forward run gives the target, body
    force).  The body force (2/Re, 0, 0) balances the viscous stress of the
    Poiseuille profile, so the laminar part is steady."""
    g = channel_grid(n, n, n, Re=Re, dt=dt, steps=steps)
    u0 = channel_initial_velocity(g, seed=20261202)
    u_alt = channel_initial_velocity(g, seed=20261203)
    return g, u0, u_alt, (2.0 / Re, 0.0, 0.0)
