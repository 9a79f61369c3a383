"""Boundary data from Python callables (host side, no device).

The reference hands its boundary data to the solver as std::function
callbacks (include/VectorFunction.h:15-53).  The library evaluates its known
exact solutions itself from descriptors; any other function is evaluated
here, with numpy, into the face tables of the C ABI's *_tabulated entry
points (include/mif_gpu.h):

- a velocity face table holds, for u, v, w in turn, the faces z-, z+, y-, y+,
  x-, x+ of the non-periodic axes, each a plane over the component's extents
  along the two other axes (lower axis fastest), with the value
  VelocityTensor::apply_bc writes there (src/VelocityTensor.cpp:36-233);
- a Neumann table holds the six faces of the node extents with the normal
  component of the pressure-gradient data at every wall node
  (apply_non_homogeneous_neumann, src/PressureEquation.cpp:10-57).

A vector function is a sequence of three callables f(t, x, y, z) that accept
numpy arrays.  Coordinates are formed exactly as the reference forms them
(min + h*(base + i), minus h/2 on the staggered axis;
include/StaggeredTensor.h:95-110, include/VelocityTensor.h:16-78).
"""
from __future__ import annotations

from typing import Callable, Sequence

import numpy as np

VectorCallables = Sequence[Callable]


def is_callables(f) -> bool:
    """True for a sequence of three callables (a host vector function)."""
    try:
        return len(f) == 3 and all(callable(c) for c in f)
    except TypeError:
        return False


class _Axes:
    """Per-axis constants of a grid at one rank (src/Constants.cpp:58-101)."""

    def __init__(self, gs):
        per = tuple(bool(p) for p in gs.periodic)
        ng = (gs.Nx_global, gs.Ny_global, gs.Nz_global)
        size = (gs.x_size, gs.y_size, gs.z_size)
        self.min = (gs.min_x, gs.min_y, gs.min_z)
        self.periodic = per
        self.h = tuple(size[a] / (ng[a] - 1) for a in range(3))
        self.half = tuple(h / 2 for h in self.h)
        self.inv_h = tuple(1 / h for h in self.h)
        self.base = tuple(-1 if p else 0 for p in per)
        self.nodes = tuple(ng[a] + (1 if per[a] else 0) for a in range(3))

    def extents(self, s):
        e = list(self.nodes)
        if s == 0:
            e[0] = self.nodes[0] if self.periodic[0] else self.nodes[0] + 1
        elif s in (1, 2):
            e[s] = self.nodes[s] + 1
        return e

    def coord(self, a, s, idx):
        x = self.min[a] + self.h[a] * (self.base[a] + np.asarray(idx, dtype=np.float64))
        return x - self.half[a] if a == s else x


def _other_axes(axis):
    return (1 if axis == 0 else 0), (1 if axis == 2 else 2)


def _plane(ax: _Axes, e, axis, fixed):
    """Index arrays (i, j, k) of the plane idx[axis] = fixed over extents e of
    the two other axes, lower axis fastest."""
    o1, o2 = _other_axes(axis)
    b, a = np.meshgrid(np.arange(e[o2]), np.arange(e[o1]), indexing="ij")
    idx = [None, None, None]
    idx[o1] = a.ravel()
    idx[o2] = b.ravel()
    idx[axis] = np.full(a.size, fixed)
    return idx


def _eval(ax: _Axes, f, s, t, idx):
    return np.broadcast_to(np.asarray(f(t, ax.coord(0, s, idx[0]), ax.coord(1, s, idx[1]), ax.coord(2, s, idx[2])),
                                      dtype=np.float64), idx[0].shape)


def velocity_face_table(gs, fns: VectorCallables, t: float) -> np.ndarray:
    """One velocity face table of the exact velocity fns at time t."""
    ax = _Axes(gs)
    out = []
    for q in range(3):
        e = ax.extents(q)
        for face in range(6):
            axis = 2 - face // 2
            hi = face & 1
            if ax.periodic[axis]:
                continue
            if q != axis:
                out.append(_eval(ax, fns[q], q, t, _plane(ax, e, axis, e[axis] - 1 if hi else 0)))
                continue
            idx = _plane(ax, e, axis, ax.nodes[axis] - 1 if hi else 0)
            at_boundary = _eval(ax, fns[q], 3, t, idx)
            div = None
            for o in _other_axes(axis):
                nxt = list(idx)
                nxt[o] = idx[o] + 1
                d = (_eval(ax, fns[o], o, t, nxt) - _eval(ax, fns[o], o, t, idx)) * ax.inv_h[o]
                div = d if div is None else div + d
            out.append(at_boundary - ax.half[axis] * div if hi else at_boundary + ax.half[axis] * div)
    return np.ascontiguousarray(np.concatenate(out) if out else np.zeros(0))


def neumann_face_table(gs, fns: VectorCallables, t: float) -> np.ndarray:
    """One Neumann table of the pressure-gradient data fns at time t."""
    ax = _Axes(gs)
    e = ax.extents(3)
    out = []
    for face in range(6):
        axis = 2 - face // 2
        idx = _plane(ax, e, axis, e[axis] - 1 if face & 1 else 0)
        out.append(_eval(ax, fns[axis], 3, t, idx))
    return np.ascontiguousarray(np.concatenate(out))


def difference_over_time(fns: VectorCallables, time_1: float, time_2: float):
    """TimeVectorFunction::get_difference_over_time(time_1, time_2)
    (src/VectorFunction.cpp:53-61): f(time_2) - f(time_1), as callables of
    (t, x, y, z) that ignore t."""
    return [lambda t, x, y, z, f=f: f(time_2, x, y, z) - f(time_1, x, y, z) for f in fns]


def full_step_times(dt: float, t_n: float):
    """t_n and the three stage times of FULL_TIMESTEP (src/Timestep.cpp:99-104)."""
    time_1 = t_n + 64.0 / 120.0 * dt
    time_2 = time_1 + 16.0 / 120.0 * dt
    return t_n, time_1, time_2, t_n + dt


def velocity_step_times(dt: float, t_n: float):
    """The boundary times of timestep_velocity (src/TimestepVelocity.cpp:60-90)."""
    return t_n + (8.0 / 15.0) * dt, t_n + (2.0 / 3.0) * dt, t_n + dt
