"""Host face tables of boundary data (mif_amd.boundary), the input of the C
ABI's *_tabulated entry points, checked on the CPU.

The velocity table is applied the way the device pass applies it (faces
z-, z+, y-, y+, x-, x+ written in order, then the periodic copies) and must
reproduce the reference's own VelocityTensor::apply_bc (oracle/_ref) called
with the same Python functions, bit for bit: the functions are polynomials,
so the reference's scalar evaluation and numpy's vectorised one round alike.
The Neumann table must reproduce the oracle's non-homogeneous right-hand side
(apply_non_homogeneous_neumann restated) built from the closed-form gradient."""
import numpy as np
import pytest

import mif_oracle as mo
import mif_ref
from mif_amd import boundary as bd
from mif_amd.core import GridSpec

PERIODIC = [(a, b, c) for a in (0, 1) for b in (0, 1) for c in (0, 1)]


def poly(c):
    # three polynomial components of (t, x, y, z)
    return [lambda t, x, y, z: 1.0 + x * y - 0.5 * z * t + 0.25 * x * x,
            lambda t, x, y, z: x - 2.0 * y * z + t * y * y,
            lambda t, x, y, z: 0.5 + z * x * t - y * y * x][c]


def grids(Ng, per, size=(1.0, 2.0, 1.5), origin=(-0.5, -1.0, 0.25)):
    gs = GridSpec(*Ng, *size, *origin, 100.0, 1e-2, 4, tuple(bool(p) for p in per))
    og = mo.make_grid(Ng, size, origin, 100.0, 1e-2, 4, per)
    return gs, og


def apply_velocity_table(o, vel, table):
    """The device pass's use of a velocity face table, restated on the host."""
    n = 0
    for q in range(3):
        f = vel[q]
        e = o.extents(q)
        for face in range(6):
            axis = 2 - face // 2
            if o.g.periodic[axis]:
                continue
            o1, o2 = (1 if axis == 0 else 0), (1 if axis == 2 else 2)
            cnt = e[o1] * e[o2]
            plane = table[n:n + cnt].reshape(e[o2], e[o1])
            n += cnt
            pos = e[axis] - 1 if face & 1 else 0
            sl = [slice(None)] * 3
            sl[2 - axis] = pos  # numpy order (k, j, i)
            f[tuple(sl)] = plane
    assert n == len(table)
    for q in range(3):
        o.apply_periodic(vel[q], q)


@pytest.mark.skipif(not mif_ref.available(), reason="oracle/_ref not built (reference sources absent)")
@pytest.mark.parametrize("per", PERIODIC, ids=lambda p: "p" + "".join(map(str, p)))
def test_velocity_table_reproduces_reference_apply_bc(per):
    Ng = (7, 8, 6)
    gs, og = grids(Ng, per)
    o, r = mo.Oracle(og), mif_ref.Reference(og)
    fns = [poly(c) for c in range(3)]
    t = 0.37
    rng = np.random.default_rng(sum(per))
    a = o.velocity()
    for c in range(3):
        a[c][...] = rng.standard_normal(a[c].shape)
    b = a.copy()
    table = bd.velocity_face_table(gs, fns, t)
    apply_velocity_table(o, a, table)
    r.apply_bc([b.u, b.v, b.w], mif_ref.PyFunction(fns), t)
    for c in range(3):
        assert np.array_equal(a[c], b[c]), c


@pytest.mark.parametrize("per", [(0, 0, 0), (1, 0, 1), (1, 1, 1)], ids=lambda p: "p" + "".join(map(str, p)))
def test_velocity_table_matches_oracle_closed_form(per):
    """Ethier-Steinman through numpy against the oracle's closed form: equal to
    rounding (the two evaluate sin/exp with different libraries)."""
    gs, og = grids((9, 8, 10), per, size=(1.0, 1.0, 2.0), origin=(0.0, 0.0, -1.0))
    o = mo.Oracle(og)
    Re, t = 50.0, 0.2
    a, c, d = np.pi / 4, np.pi / 2, np.pi / 2

    def es(comp):
        def f(t, x, y, z):
            X = [x, y, z]
            p, q, s = X[comp], X[(comp + 1) % 3], X[(comp + 2) % 3]
            return -a * (np.exp(a * p) * np.sin(a * q + d * s) + np.exp(a * s) * np.cos(a * p + d * q)) * \
                np.exp(-d * d * t / Re)
        return f

    table = bd.velocity_face_table(gs, [es(k) for k in range(3)], t)
    u = o.velocity()
    rng = np.random.default_rng(1)
    for k in range(3):
        u[k][...] = rng.standard_normal(u[k].shape)
    ref = u.copy()
    apply_velocity_table(o, u, table)
    o.apply_bc(ref, mo.make_fn(mo.FN_ES_VELOCITY, Re=Re), t)
    for k in range(3):
        assert np.max(np.abs(u[k] - ref[k])) <= 1e-14 * np.max(np.abs(ref[k])), k


def test_neumann_table_matches_oracle_rhs():
    """The Neumann table, added as apply_non_homogeneous_neumann adds it,
    gives the oracle's right-hand side with the closed-form gradient
    difference of the stage (get_difference_over_time(new, prev))."""
    gs, og = grids((7, 6, 8), (0, 0, 0), size=(1.0, 1.0, 2.0), origin=(0.0, 0.0, -1.0))
    o = mo.Oracle(og)
    Re = 20.0
    a, d = np.pi / 4, np.pi / 2

    def dpdx_cyc(comp):
        def f(t, x, y, z):
            X = [x, y, z]
            x, y, z = X[comp], X[(comp + 1) % 3], X[(comp + 2) % 3]
            S1, C1 = np.sin(a * x + d * y), np.cos(a * z + d * x)
            S2, C2 = np.sin(a * y + d * z), np.cos(a * x + d * y)
            S3, C3 = np.sin(a * z + d * x), np.cos(a * y + d * z)
            s = (2 * a * np.exp(2 * a * x) + 2 * np.exp(a * (y + z)) * (a * C2 * C1 - d * S1 * S3)
                 + 2 * a * np.exp(a * (z + x)) * S2 * (C2 - S1) + 2 * np.exp(a * (x + y)) * C3 * (d * C1 + a * S3))
            return -a * a / 2 * s * np.exp(-2 * d * d * t / Re)
        return f

    grad = [dpdx_cyc(k) for k in range(3)]
    t_n = 0.1
    for stage in (1, 2, 3):
        times = bd.full_step_times(o.dt, t_n)
        diff = bd.difference_over_time(grad, times[stage], times[stage - 1])
        table = bd.neumann_face_table(gs, diff, 0.0)
        vel = o.velocity()
        rng = np.random.default_rng(stage)
        for k in range(3):
            vel[k][...] = rng.standard_normal(vel[k].shape)
        base = o.zeros(3)
        o.pressure_rhs(base, vel, 0.02)
        got = base.copy()
        e = o.extents(3)
        n = 0
        inv_h = list(o.g.one_over_h)
        for face in range(6):
            axis = 2 - face // 2
            o1, o2 = (1 if axis == 0 else 0), (1 if axis == 2 else 2)
            cnt = e[o1] * e[o2]
            plane = table[n:n + cnt].reshape(e[o2], e[o1])
            n += cnt
            sl = [slice(None)] * 3
            sl[2 - axis] = e[axis] - 1 if face & 1 else 0
            sign = -1.0 if face & 1 else 1.0
            got[tuple(sl)] += sign * plane * 2.0 * inv_h[axis]
        ref = o.zeros(3)
        o.pressure_rhs(ref, vel, 0.02, grad_p=o.nhn_stage_function(mo.make_fn(mo.FN_ES_GRAD_P, Re=Re), t_n, stage))
        assert np.max(np.abs(got - ref)) <= 1e-13 * np.max(np.abs(ref)), stage
