"""Pins the oracle's time integrators (src/Timestep.cpp:10-156,
src/TimestepVelocity.cpp:11-90): the RK3 stability polynomial on a discrete
eigenfunction, and second-order convergence on the reference's
manufactured solutions (test/full_test.cpp, test/velocity_test.cpp)."""
import numpy as np
import pytest

import mif_oracle as mo
from helpers import PI, grids


def _sine_mode_setup(Re, dt_steps=1):
    # fully periodic box, u = sin(y) on u's points, v = w = 0: convection
    # vanishes and the diffusion stencil has eigenvalue 2(cos h - 1)/(Re h^2)
    n = 17
    og, gs = grids((n, n, n), (2 * PI, 2 * PI, 2 * PI), (0.0, 0.0, 0.0), Re, 0.1, dt_steps, (1, 1, 1))
    o = mo.Oracle(og)
    vel = o.velocity()
    e = o.extents(0)
    y = o.g.min[1] + o.g.h[1] * (o.g.base[1] + np.arange(e[1]))
    vel.u[...] = np.sin(y)[None, :, None]
    h = o.g.h[1]
    z = o.dt * 2 * (np.cos(h) - 1) / (Re * h * h)
    return o, vel, z


@pytest.mark.parametrize("Re", [0.25, 0.5])
def test_full_step_stability_polynomial(Re):
    o, vel, z = _sine_mode_setup(Re)
    u0 = vel.u.copy()
    buf, buf2 = o.velocity(), o.velocity()
    p, pb = o.zeros(3), o.zeros(3)
    o.timestep(vel, buf, buf2, mo.make_fn(mo.FN_ZERO), 0.0, p, pb)
    R = 1 + z + z * z / 2 + z ** 3 / 6
    assert abs(z) > 0.1
    mask = np.abs(u0) > 0.1
    assert np.max(np.abs(vel.u[mask] / u0[mask] - R)) < 1e-13
    assert np.max(np.abs(vel.v)) < 1e-14 and np.max(np.abs(vel.w)) < 1e-14
    assert np.max(np.abs(p)) < 1e-14


@pytest.mark.parametrize("Re", [0.25, 0.5])
def test_velocity_step_stability_polynomial(Re):
    o, vel, z = _sine_mode_setup(Re)
    u0 = vel.u.copy()
    buf, rhs = o.velocity(), o.velocity()
    o.timestep_velocity(vel, buf, rhs, mo.make_fn(mo.FN_ZERO), mo.make_fn(mo.FN_ZERO), 0.0)
    R = 1 + z + z * z / 2 + z ** 3 / 6
    mask = np.abs(u0) > 0.1
    assert np.max(np.abs(vel.u[mask] / u0[mask] - R)) < 1e-13


def _es_error(N, steps):
    Re, T = 1e3, 1e-4
    og, _ = grids((N, N, N), (1.0, 1.0, 2.0), (0.0, 0.0, -1.0), Re, T, steps, (0, 0, 0))
    o = mo.Oracle(og)
    ev = mo.make_fn(mo.FN_ES_VELOCITY, Re=Re)
    vel, buf, buf2 = o.velocity(), o.velocity(), o.velocity()
    p, pb = o.zeros(3), o.zeros(3)
    o.set_velocity(vel, ev, 0.0)
    o.set(p, 3, mo.make_fn(mo.FN_ES_PRESSURE, Re=Re), 0, 0.0)
    for s in range(steps):
        o.timestep(vel, buf, buf2, ev, s * o.dt, p, pb)
    return o.error_norms_velocity(vel, ev, T)


def test_full_step_second_order_convergence():
    # test/full_test.cpp settings, N doubling with the step count
    # (analysis/plot_convergence.py:21-32)
    e1 = _es_error(16, 2)
    e2 = _es_error(32, 4)
    order = np.log2(e1 / e2)
    assert np.all(order > 1.8) and np.all(order < 2.4), order


def _vt_error(N, steps):
    Re, T = 1e4, 1e-4
    og, _ = grids((N, N, N), (1.0, 1.0, 1.0), (0.0, 0.0, 0.0), Re, T, steps, (0, 0, 0))
    o = mo.Oracle(og)
    ev = mo.make_fn(mo.FN_VT_VELOCITY, Re=Re)
    fo = mo.make_fn(mo.FN_VT_FORCING, Re=Re)
    vel, buf, rhs = o.velocity(), o.velocity(), o.velocity()
    o.set_velocity(vel, ev, 0.0)
    for s in range(steps):
        o.timestep_velocity(vel, buf, rhs, ev, fo, s * o.dt)
    return o.error_norms_velocity(vel, ev, T)


def test_velocity_step_second_order_convergence():
    # test/velocity_test.cpp settings (unit cube, Re = 1e4, T = 1e-4)
    e1 = _vt_error(12, 1)
    e2 = _vt_error(24, 2)
    order = np.log2(e1[:2] / e2[:2])
    assert np.all(order > 1.7) and np.all(order < 2.5), order
