"""Pins the oracle's stencils, analytic functions, boundary conditions and
norms by closed forms (include/MomentumEquation.h:43-255,
include/VelocityTensor.h, src/VelocityTensor.cpp:36-233, src/Norms.cpp)."""
import numpy as np
import pytest
import sympy as sp

import mif_oracle as mo
from helpers import PI, grids

RNG = np.random.default_rng(123)


def stag_coords(o, s):
    """Coordinate arrays (x, y, z) of every index of a field with staggering
    s, broadcastable to its (sz, sy, sx) shape."""
    g = o.g
    e = o.extents(s)
    out = []
    for a in range(3):
        idx = np.arange(e[a])
        x = g.min[a] + g.h[a] * (g.base[a] + idx)
        if a == s:
            x = x - g.h[a] / 2
        shape = [1, 1, 1]
        shape[2 - a] = e[a]
        out.append(x.reshape(shape))
    return out


# --------------------------------------------------------------------------
# momentum right-hand side: exact on quadratic (own component) and linear
# (other components) fields
# --------------------------------------------------------------------------
@pytest.mark.parametrize("comp", [0, 1, 2])
@pytest.mark.parametrize("per", [(0, 0, 0), (1, 0, 1)])
def test_momentum_rhs_exact_on_polynomials(comp, per):
    og, _ = grids((9, 8, 10), (1.0, 2.0, 1.5), (0.0, -1.0, 0.5), 7.0, 1.0, 1, per)
    o = mo.Oracle(og)
    Re = o.g.Re
    lin = RNG.standard_normal((3, 4))        # a0 + a.x per component
    quad = RNG.standard_normal((3, 3))       # diagonal quadratic terms of component comp

    def field(c, x, y, z):
        v = lin[c, 0] + lin[c, 1] * x + lin[c, 2] * y + lin[c, 3] * z
        if c == comp:
            v = v + quad[c, 0] * x * x + quad[c, 1] * y * y + quad[c, 2] * z * z
        return v

    def grad(c, x, y, z):
        gx = lin[c, 1] + (2 * quad[c, 0] * x if c == comp else 0)
        gy = lin[c, 2] + (2 * quad[c, 1] * y if c == comp else 0)
        gz = lin[c, 3] + (2 * quad[c, 2] * z if c == comp else 0)
        return gx, gy, gz

    vel = o.velocity()
    for c in range(3):
        x, y, z = stag_coords(o, c)
        vel[c][...] = field(c, x, y, z) + 0 * (x + y + z)
    x, y, z = stag_coords(o, comp)
    e = o.extents(comp)
    lap = 2 * (quad[comp, 0] + quad[comp, 1] + quad[comp, 2])
    worst = 0.0
    for k in range(1, e[2] - 1):
        for j in range(1, e[1] - 1):
            for i in range(1, e[0] - 1):
                px, py, pz = x.reshape(-1)[i], y.reshape(-1)[j], z.reshape(-1)[k]
                us = [field(c, px, py, pz) for c in range(3)]
                gx, gy, gz = grad(comp, px, py, pz)
                exact = -(us[0] * gx + us[1] * gy + us[2] * gz) + lap / Re
                got = o.momentum_rhs(vel, comp, i, j, k)
                worst = max(worst, abs(got - exact) / (1 + abs(exact)))
    assert worst < 1e-10


# --------------------------------------------------------------------------
# analytic functions against their defining equations (sympy)
# --------------------------------------------------------------------------
def test_velocity_test_forcing_matches_definition():
    t, x, y, z, Re = sp.symbols("t x y z Re")
    u = sp.sin(x) * sp.cos(y) * sp.sin(z) * sp.sin(t)
    v = sp.cos(x) * sp.sin(y) * sp.sin(z) * sp.sin(t)
    w = 2 * sp.cos(x) * sp.cos(y) * sp.cos(z) * sp.sin(t)
    lap = lambda f: sp.diff(f, x, 2) + sp.diff(f, y, 2) + sp.diff(f, z, 2)  # noqa: E731
    forc = [sp.diff(c, t) + u * sp.diff(c, x) + v * sp.diff(c, y) + w * sp.diff(c, z) - lap(c) / Re for c in (u, v, w)]
    fl = [sp.lambdify((t, x, y, z, Re), f) for f in forc]
    vl = [sp.lambdify((t, x, y, z), c) for c in (u, v, w)]
    o = mo.Oracle(grids((4, 4, 4), (1, 1, 1), (0, 0, 0), 1, 1, 1, (0, 0, 0))[0])
    for _ in range(20):
        tt, xx, yy, zz = RNG.uniform(-2, 2, 4)
        ReV = float(RNG.uniform(1, 1e4))
        fo = mo.make_fn(mo.FN_VT_FORCING, Re=ReV)
        ve = mo.make_fn(mo.FN_VT_VELOCITY, Re=ReV)
        for c in range(3):
            assert abs(o.fn_eval(fo, c, tt, xx, yy, zz) - fl[c](tt, xx, yy, zz, ReV)) < 1e-12
            assert abs(o.fn_eval(ve, c, tt, xx, yy, zz) - vl[c](tt, xx, yy, zz)) < 1e-14


def test_ethier_steinman_is_an_exact_solution():
    t, x, y, z, Re = sp.symbols("t x y z Re")
    a, d = sp.pi / 4, sp.pi / 2
    E = sp.exp(-d * d * t / Re)
    u = -a * (sp.exp(a * x) * sp.sin(a * y + d * z) + sp.exp(a * z) * sp.cos(a * x + d * y)) * E
    v = -a * (sp.exp(a * y) * sp.sin(a * z + d * x) + sp.exp(a * x) * sp.cos(a * y + d * z)) * E
    w = -a * (sp.exp(a * z) * sp.sin(a * x + d * y) + sp.exp(a * y) * sp.cos(a * z + d * x)) * E
    p = -a * a / 2 * (sp.exp(2 * a * x) + sp.exp(2 * a * y) + sp.exp(2 * a * z)
                      + 2 * sp.sin(a * x + d * y) * sp.cos(a * z + d * x) * sp.exp(a * (y + z))
                      + 2 * sp.sin(a * y + d * z) * sp.cos(a * x + d * y) * sp.exp(a * (z + x))
                      + 2 * sp.sin(a * z + d * x) * sp.cos(a * y + d * z) * sp.exp(a * (x + y))) * E * E
    vl = [sp.lambdify((t, x, y, z, Re), c) for c in (u, v, w)]
    pl = sp.lambdify((t, x, y, z, Re), p)
    gl = [sp.lambdify((t, x, y, z, Re), sp.diff(p, s)) for s in (x, y, z)]
    o = mo.Oracle(grids((4, 4, 4), (1, 1, 1), (0, 0, 0), 1, 1, 1, (0, 0, 0))[0])
    for _ in range(20):
        tt, xx, yy, zz = RNG.uniform(-1, 1, 4)
        ReV = float(RNG.uniform(1, 1e4))
        fv = mo.make_fn(mo.FN_ES_VELOCITY, Re=ReV)
        fp = mo.make_fn(mo.FN_ES_PRESSURE, Re=ReV)
        fg = mo.make_fn(mo.FN_ES_GRAD_P, Re=ReV)
        for c in range(3):
            assert abs(o.fn_eval(fv, c, tt, xx, yy, zz) - vl[c](tt, xx, yy, zz, ReV)) < 1e-13
            assert abs(o.fn_eval(fg, c, tt, xx, yy, zz) - gl[c](tt, xx, yy, zz, ReV)) < 1e-12
        assert abs(o.fn_eval(fp, 0, tt, xx, yy, zz) - pl(tt, xx, yy, zz, ReV)) < 1e-12
    # it solves Navier-Stokes: u_t + (u.grad)u - lap(u)/Re + grad p = 0
    lap = lambda f: sp.diff(f, x, 2) + sp.diff(f, y, 2) + sp.diff(f, z, 2)  # noqa: E731
    res = [sp.diff(c, t) + u * sp.diff(c, x) + v * sp.diff(c, y) + w * sp.diff(c, z) - lap(c) / Re + sp.diff(p, s)
           for c, s in zip((u, v, w), (x, y, z))]
    rl = [sp.lambdify((t, x, y, z, Re), r) for r in res]
    for _ in range(10):
        tt, xx, yy, zz = RNG.uniform(-1, 1, 4)
        for r in rl:
            assert abs(r(tt, xx, yy, zz, 100.0)) < 1e-11


# --------------------------------------------------------------------------
# boundary conditions: closed-form face values and the face order
# --------------------------------------------------------------------------
def es_numpy(c, t, x, y, z, Re):
    a, d = np.pi / 4, np.pi / 2
    X = (x, y, z)
    p, q, r = X[c], X[(c + 1) % 3], X[(c + 2) % 3]
    return -a * (np.exp(a * p) * np.sin(a * q + d * r) + np.exp(a * r) * np.cos(a * p + d * q)) * np.exp(-d * d * t / Re)


def test_apply_bc_walls_closed_form():
    Re, t = 100.0, 0.3
    og, _ = grids((7, 6, 8), (1.0, 1.0, 2.0), (0.0, 0.0, -1.0), Re, 1.0, 1, (0, 0, 0))
    o = mo.Oracle(og)
    vel = o.velocity()
    for c in range(3):
        vel[c][...] = RNG.standard_normal(vel[c].shape)
    interior = [v[1:-1, 1:-1, 1:-1].copy() for v in vel]
    o.apply_bc(vel, mo.make_fn(mo.FN_ES_VELOCITY, Re=Re), t)
    g = o.g
    h = list(g.h)
    for c in range(3):
        # interior untouched
        assert np.array_equal(vel[c][1:-1, 1:-1, 1:-1], interior[c])
        x, y, z = stag_coords(o, c)
        exact = es_numpy(c, t, x, y, z, Re) + 0 * (x + y + z)
        f = vel[c]
        # tangential faces: the exact value at the component's own points
        for axis in range(3):
            if axis == c:
                continue
            for side in (0, -1):
                sl = [slice(None)] * 3
                sl[2 - axis] = side
                # points of this face that no later face overwrites
                mask = np.ones(f.shape, bool)
                later = [1, 0] if axis == 2 else ([0] if axis == 1 else [])
                for a2 in later:
                    for s2 in (0, -1):
                        sl2 = [slice(None)] * 3
                        sl2[2 - a2] = s2
                        mask[tuple(sl2)] = False
                sel = tuple(sl)
                assert np.allclose(f[sel][mask[sel]], exact[sel][mask[sel]], rtol=0, atol=1e-14)
    # normal component at the z- face: w = w_exact(node) + dz/2 (du/dx + dv/dy)
    # at interior x/y (later faces overwrite the edges)
    xw, yw, zw = stag_coords(o, 2)
    xn = g.min[0] + h[0] * np.arange(o.g.N[0])
    yn = g.min[1] + h[1] * np.arange(o.g.N[1])
    zn0 = g.min[2]
    for j in range(1, o.g.N[1] - 1):
        for i in range(1, o.g.N[0] - 1):
            wb = es_numpy(2, t, xn[i], yn[j], zn0, Re)
            du = (es_numpy(0, t, xn[i + 1] - h[0] / 2, yn[j], zn0, Re) - es_numpy(0, t, xn[i] - h[0] / 2, yn[j], zn0, Re)) / h[0]
            dv = (es_numpy(1, t, xn[i], yn[j + 1] - h[1] / 2, zn0, Re) - es_numpy(1, t, xn[i], yn[j] - h[1] / 2, zn0, Re)) / h[1]
            assert abs(vel.w[0, j, i] - (wb + h[2] / 2 * (du + dv))) < 1e-13
    # edge priority: at (i, 0, 0) of w the y- face (applied after z-) wins and
    # holds the exact value at w's own (ghost) point
    for i in range(1, o.g.N[0] - 1):
        assert abs(vel.w[0, 0, i] - es_numpy(2, t, xw.ravel()[i], yw.ravel()[0], zw.ravel()[0], Re)) < 1e-14


def test_periodic_copies_closed_form():
    og, _ = grids((7, 6, 8), (2 * PI, 2.0, PI), (0.0, -1.0, 0.0), 1.0, 1.0, 1, (1, 0, 1))
    o = mo.Oracle(og)
    for s in range(4):
        f = RNG.standard_normal(o.shape(s))
        before = f.copy()
        o.apply_periodic(f, s)
        # x copies first, then z (src/StaggeredTensor.cpp:221-256)
        ref = before.copy()
        ref[:, :, 0] = ref[:, :, -2]
        ref[:, :, -1] = ref[:, :, 1]
        ref[0, :, :] = ref[-2, :, :]
        ref[-1, :, :] = ref[1, :, :]
        assert np.array_equal(f, ref)


# --------------------------------------------------------------------------
# norms and adjust_pressure
# --------------------------------------------------------------------------
def test_norms_closed_form():
    og, _ = grids((6, 7, 8), (1.0, 1.0, 1.0), (0.0, 0.0, 0.0), 1.0, 1.0, 1, (0, 0, 0))
    o = mo.Oracle(og)
    cvec = (0.3, -1.2, 2.0)
    delta = np.array([1e-3, -2e-3, 5e-4])
    fn = mo.make_fn(mo.FN_CONSTANT, c=cvec)
    vel = o.velocity()
    for c in range(3):
        vel[c][...] = cvec[c] + delta[c]
    n = o.error_norms_velocity(vel, fn, 0.0)
    count = (o.g.N[0] - 2) * (o.g.N[1] - 2) * (o.g.N[2] - 2)
    vol = o.g.h[0] * o.g.h[1] * o.g.h[2]
    assert np.isclose(n[0], count * np.linalg.norm(delta) * vol, rtol=1e-12)
    assert np.isclose(n[1], np.sqrt(count * (delta ** 2).sum() * vol), rtol=1e-12)
    assert np.isclose(n[2], np.abs(delta).max(), rtol=1e-12)
    p = np.full(o.shape(3), cvec[0] + delta[0])
    pn = o.error_norms_pressure(p, fn, 0.0)
    cnt = o.npts[0] * o.npts[1] * o.npts[2]
    assert np.isclose(pn[0], cnt * abs(delta[0]) * vol, rtol=1e-12)
    assert np.isclose(pn[2], abs(delta[0]), rtol=1e-12)


def test_adjust_pressure_removes_constant():
    og, _ = grids((9, 9, 9), (2 * PI,) * 3, (0.0, 0.0, 0.0), 1.0, 1.0, 1, (0, 0, 1))
    o = mo.Oracle(og)
    pe = mo.make_fn(mo.FN_PT_PRESSURE)
    p = o.zeros(3)
    o.set(p, 3, pe, 0, 1.0)
    exact = p.copy()
    p += 3.25
    o.adjust_pressure(p, pe, 1.0)
    assert np.max(np.abs(p - exact)) < 1e-13
