"""The hot path at the full sizes of the measured workloads, through
properties that hold at any size: agreement with the oracle over a few
steps, the residual of the exact 7-point operator, the defect correction's
convergence to the FP64 solve, and the mixed form's velocity error over a
long run.

- the stand-alone Poisson solve on N_g = 257^3 (periodic x/z, Neumann y),
  right-hand side ~ N(0,1);
- the channel 256 x 128 x 256 (N_g = 257, 129, 257; Re = 1000, dt = 5e-4,
  Poiseuille flow plus a divergence-free perturbation);
- the Ethier-Steinman run on 64^3 nodes of the unit cube at Re = 1, 10
  steps, at a stable dt and at the stated dt = 1e-4.
"""
import numpy as np
import pytest

import mif_oracle as mo
from helpers import PI, channel_initial, grids, rel_err, seeded_normal
from test_oracle_poisson import apply_op, weights

pytestmark = pytest.mark.gpu

mif = pytest.importorskip("mif_amd")

TOL = 1e-11


def l2rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _case(msg):
    print(f"[case] {msg}", flush=True)


def test_poisson_257cubed_residual_and_defect_correction():
    _case("poisson 257^3 periodic x/z, Neumann y")
    Ng, per = (257, 257, 257), (1, 0, 1)
    og, gs = grids(Ng, (2 * PI, 2.0, PI), (0.0, -1.0, 0.0), 1.0, 1.0, 1, per)
    o = mo.Oracle(og)
    b = seeded_normal(tuple(reversed(o.npts)), 20261202)
    W = weights(o).reshape(b.shape)
    # the right-hand side projected onto the range of the operator: the
    # trapezoid-weighted mean removed (the solve does the same)
    bp = b - (W * b).sum() / W.sum()
    nb = np.linalg.norm(bp)

    ctx = mif.Context(gs, precision=mif.FP64)
    ctx.poisson_set_rhs(b)
    ctx.poisson_run()
    ref = ctx.poisson_get_solution()
    ctx.close()
    res64 = np.linalg.norm(bp - apply_op(o, ref)) / nb
    print("fp64 residual", res64, flush=True)
    # FP64 transforms: the residual of the exact operator at rounding level
    # (measured 8.3e-15)
    assert res64 < 1e-13

    ctx = mif.Context(gs, precision=mif.MIXED)
    ctx.poisson_set_rhs(b)
    errs, res = [], []
    for sweeps in (0, 1, 2):
        ctx.set_defect_sweeps(sweeps)
        ctx.poisson_run()
        phi = ctx.poisson_get_solution()
        errs.append(l2rel(phi, ref))
        res.append(np.linalg.norm(bp - apply_op(o, phi)) / nb)
    ctx.close()
    print("mixed errors vs fp64", errs, "residuals", res, flush=True)
    # FP32 transforms: error bounded by about eps32 * kappa (kappa ~ 1e5
    # here), each FP64 defect sweep contracts it by a like factor (measured
    # 2.9e-7, 5.4e-12, 1.8e-13; residuals 4.3e-6, 2.1e-12, 1.4e-15)
    assert errs[0] < 1e-5
    assert errs[1] < errs[0] * 1e-2
    assert errs[2] < 1e-11
    assert res[2] < 1e-13


def _channel(ny=129, steps=50):
    Ng = (257, ny, 257)
    return grids(Ng, (2 * PI, 2.0, PI), (0.0, -1.0, 0.0), 1e3, 5e-4 * steps, steps, (1, 0, 1))


def test_channel_256x128x256_fp64_matches_oracle():
    _case("timestep channel 257x129x257 fp64 vs oracle, 50 steps")
    og, gs = _channel()
    o = mo.Oracle(og)
    vel, buf, buf2 = o.velocity(), o.velocity(), o.velocity()
    p, pb = o.zeros(3), o.zeros(3)
    v0 = channel_initial(o, 20261202)
    for c in range(3):
        vel[c][...] = v0[c]
    ctx = mif.Context(gs, precision=mif.FP64)
    gvel, gbuf, gbuf2 = ctx.velocity(), ctx.velocity(), ctx.velocity()
    gp, gpb = ctx.pressure(), ctx.pressure()
    gvel.upload(vel)
    ev = mo.make_fn(mo.FN_ZERO)
    fn = mif.function(mif.FN_ZERO)
    # the configuration's 50 steps
    for s in range(50):
        o.timestep(vel, buf, buf2, ev, s * o.dt, p, pb)
        mif.timestep(gvel, gbuf, gbuf2, fn, s * o.dt, gp, gpb)
        if s % 10 == 9:
            print(f"  step {s + 1}", flush=True)
    got = gvel.download()
    errs = [rel_err(got[c], vel[c]) for c in range(3)]
    perr = rel_err(gp.download(), p)
    ctx.close()
    # condition number of the discrete Laplacian: 4/h^2 summed over the axes
    # over the smallest nonzero eigenvalue (the periodic x mode 1)
    h = (2 * PI / 256, 2.0 / 128, PI / 256)
    kappa = sum(4 / hh ** 2 for hh in h) / (4 * np.sin(PI / 256) ** 2 / h[0] ** 2)
    print("velocity", errs, "pressure", perr, "kappa", kappa, flush=True)
    for c in range(3):
        assert errs[c] < TOL, c
    # pressure carries the solves' rounding times kappa: two FP64 orders of
    # the oracle itself differ that way
    # (test_oracle_poisson.py::test_solve_rounding_scales_with_condition_number)
    assert perr < TOL * kappa / 1e3


def test_channel_256x128x256_mixed_tracks_fp64_over_50_steps():
    _case("timestep channel 257x129x257 mixed vs fp64, 50 steps")
    og, gs = _channel()
    o = mo.Oracle(og)
    v0 = channel_initial(o, 20261202)
    out = {}
    for name, prec in (("fp64", mif.FP64), ("mixed", mif.MIXED)):
        ctx = mif.Context(gs, precision=prec)
        if prec == mif.MIXED:
            ctx.set_defect_sweeps(0)
        gvel, gbuf, gbuf2 = ctx.velocity(), ctx.velocity(), ctx.velocity()
        gp, gpb = ctx.pressure(), ctx.pressure()
        gvel.upload(v0)
        fn = mif.function(mif.FN_ZERO)
        for s in range(50):
            mif.timestep(gvel, gbuf, gbuf2, fn, s * ctx.dt, gp, gpb)
        out[name] = [np.asarray(c, dtype=np.float64) for c in gvel.download()]
        ctx.close()
    num = sum(np.linalg.norm(a - b) ** 2 for a, b in zip(out["mixed"], out["fp64"]))
    den = sum(np.linalg.norm(b) ** 2 for b in out["fp64"])
    err = float(np.sqrt(num / den))
    print("mixed velocity rel L2 vs fp64 after 50 steps", err, flush=True)
    assert all(np.isfinite(c).all() for c in out["mixed"])
    assert err < 1e-6


def _es_run(dt, steps):
    # 64^3 nodes on the unit cube, all walls, Re = 1, Ethier-Steinman initial
    # field and Dirichlet data (the manufactured-solution run of
    # test/full_test.cpp at the measured configuration's size)
    Re, N = 1.0, 64
    og, gs = grids((N, N, N), (1.0, 1.0, 1.0), (0.0, 0.0, 0.0), Re, dt * steps, steps, (0, 0, 0))
    o = mo.Oracle(og)
    ev = mo.make_fn(mo.FN_ES_VELOCITY, Re=Re)
    ep = mo.make_fn(mo.FN_ES_PRESSURE, Re=Re)
    vel, buf, buf2 = o.velocity(), o.velocity(), o.velocity()
    p, pb = o.zeros(3), o.zeros(3)
    o.set_velocity(vel, ev, 0.0)
    o.set(p, 3, ep, 0, 0.0)
    ctx = mif.Context(gs, precision=mif.FP64)
    gvel, gbuf, gbuf2 = ctx.velocity(), ctx.velocity(), ctx.velocity()
    gp, gpb = ctx.pressure(), ctx.pressure()
    gvel.upload(vel)
    gp.upload(p)
    fn = mif.function(mif.FN_ES_VELOCITY, Re=Re)
    for s in range(steps):
        o.timestep(vel, buf, buf2, ev, s * o.dt, p, pb)
        mif.timestep(gvel, gbuf, gbuf2, fn, s * o.dt, gp, gpb)
    T = steps * o.dt
    n_ref = o.error_norms_velocity(vel, ev, T)
    n_gpu = mif.error_norms_velocity(gvel, fn, T)
    got = gvel.download()
    errs = [rel_err(got[c], vel[c]) for c in range(3)]
    ctx.close()
    return errs, n_gpu, n_ref


def test_ethier_steinman_64cubed_re1_stable_dt():
    # dt = 5e-5: 12 dt / (Re h^2) = 2.38, inside the RK3 stability limit 2.51
    _case("timestep ES 64^3 walls Re=1 dt=5e-5, 10 steps")
    errs, n_gpu, n_ref = _es_run(5e-5, 10)
    print("velocity", errs, "norms", n_gpu, n_ref, flush=True)
    for c in range(3):
        assert errs[c] < TOL, c
    assert np.allclose(n_gpu, n_ref, rtol=1e-9, atol=0)


def test_ethier_steinman_64cubed_re1_literal_dt():
    # the configuration as stated, dt = 1e-4: 12 dt / (Re h^2) = 4.76 exceeds
    # the RK3 limit, so the highest grid modes grow by |R(-4.76)| = 10.4 per
    # step (1.5e10 over 10 steps) and rounding differences between any two
    # evaluation orders grow with them (measured 3.4e-8 relative; the
    # stable-dt case above agrees to 6e-16)
    _case("timestep ES 64^3 walls Re=1 dt=1e-4 (beyond the RK3 limit), 10 steps")
    errs, n_gpu, n_ref = _es_run(1e-4, 10)
    print("velocity", errs, "norms", n_gpu, n_ref, flush=True)
    assert all(np.isfinite(n_gpu))
    for c in range(3):
        assert errs[c] < 1e-6, c
    assert np.allclose(n_gpu, n_ref, rtol=1e-6, atol=0)


def test_channel_512x256x512_unshifted_layout_matches_oracle():
    # fields above 256 MB (FP64 562 MB, FP32 281 MB) keep element 0 at the
    # allocation start: the stage kernel's 4/8-byte halo loads, ordinary
    # stores and the projection's unshifted slices, as in the 512^3 FP64 and
    # 768^3 mixed workloads
    _case("timestep channel 513x257x513 fp64 and mixed vs oracle, 2 steps (unshifted layout)")
    Ng, steps = (513, 257, 513), 2
    og, gs = grids(Ng, (2 * PI, 2.0, PI), (0.0, -1.0, 0.0), 1e3, 5e-4 * steps, steps, (1, 0, 1))
    o = mo.Oracle(og)
    vel, buf, buf2 = o.velocity(), o.velocity(), o.velocity()
    p, pb = o.zeros(3), o.zeros(3)
    v0 = channel_initial(o, 20261203)
    for c in range(3):
        vel[c][...] = v0[c]
    ev = mo.make_fn(mo.FN_ZERO)
    for s in range(steps):
        o.timestep(vel, buf, buf2, ev, s * o.dt, p, pb)
    del buf, buf2, pb
    h = (2 * PI / 512, 2.0 / 256, PI / 512)
    kappa = sum(4 / hh ** 2 for hh in h) / (4 * np.sin(PI / 512) ** 2 / h[0] ** 2)
    for name, prec in (("fp64", mif.FP64), ("mixed", mif.MIXED)):
        ctx = mif.Context(gs, precision=prec)
        if prec == mif.MIXED:
            ctx.set_defect_sweeps(0)
        gvel, gbuf, gbuf2 = ctx.velocity(), ctx.velocity(), ctx.velocity()
        gp, gpb = ctx.pressure(), ctx.pressure()
        gvel.upload(v0)
        fn = mif.function(mif.FN_ZERO)
        for s in range(steps):
            mif.timestep(gvel, gbuf, gbuf2, fn, s * o.dt, gp, gpb)
        got = [np.asarray(g, dtype=np.float64) for g in gvel.download()]
        errs = [rel_err(got[c], vel[c]) for c in range(3)]
        l2 = float(np.sqrt(sum(np.linalg.norm(got[c] - vel[c]) ** 2 for c in range(3)) /
                           sum(np.linalg.norm(vel[c]) ** 2 for c in range(3))))
        perr = rel_err(np.asarray(gp.download(), dtype=np.float64), p)
        ctx.close()
        del got, gvel, gbuf, gbuf2, gp, gpb
        print(name, "velocity", errs, "rel L2", l2, "pressure", perr, "kappa", kappa, flush=True)
        if prec == mif.FP64:
            for c in range(3):
                assert errs[c] < TOL, c
            assert perr < TOL * kappa / 1e3
        else:
            # FP32 storage and transforms (kappa ~ 2e5, below the FP64-transform
            # switch): the velocity's relative L2 distance over the three
            # components within the mixed form's 1e-6 bar (per component, the
            # small perturbation components v and w carry the FP32 rounding of
            # the Poiseuille u: max-norm 2e-5 of their own size)
            assert l2 < 1e-6
            assert all(np.isfinite(errs)) and np.isfinite(perr)
