"""The mixed-precision form (FP32 fields, stencils and transforms; FP64
right-hand side, null-space projection and defect correction) against the
FP64 oracle: error bounds per field, and the defect sweeps' convergence."""
import numpy as np
import pytest

import mif_oracle as mo
from helpers import PI, channel_initial, grids, rel_err, seeded_normal
from test_oracle_poisson import apply_op, weights

pytestmark = pytest.mark.gpu

mif = pytest.importorskip("mif_amd")


def l2rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("Ng,per", [((65, 33, 65), (1, 0, 1)), ((33, 33, 33), (0, 0, 0)), ((129, 65, 129), (1, 0, 1)),
                                    # FP32 radix-3 kernels: L = 192 (x, y, z) and 96, 384
                                    ((193, 97, 193), (1, 0, 1)), ((97, 193, 33), (1, 0, 1))])
def test_defect_correction_converges(Ng, per):
    print(f"[case] mixed poisson {Ng} {per}", flush=True)
    og, gs = grids(Ng, (2 * PI, 2.0, PI), (0.0, -1.0, 0.0), 1.0, 1.0, 1, per)
    o = mo.Oracle(og)
    b = seeded_normal(tuple(reversed(o.npts)), 20261202)
    W = weights(o).reshape(b.shape)
    bp = b - (W * b).sum() / W.sum()
    ref = b.copy()
    o.poisson_packed(ref)
    ctx = mif.Context(gs, precision=mif.MIXED)
    ctx.poisson_set_rhs(b)
    errs, res = [], []
    for sweeps in (0, 1, 2):
        ctx.set_defect_sweeps(sweeps)
        ctx.poisson_run()
        phi = ctx.poisson_get_solution()
        errs.append(l2rel(phi, ref))
        res.append(np.linalg.norm(bp - apply_op(o, phi)) / np.linalg.norm(bp))
    print("errors", errs, "residuals", res)
    assert errs[0] < 1e-5
    assert errs[1] < errs[0] * 1e-2 and errs[2] < 1e-10
    assert res[2] < 1e-11


@pytest.mark.parametrize("sweeps,schedule", [(0, 0), (1, 0), (1, 1)], ids=["s0", "s1", "s1_last"])
def test_mixed_timestep_channel_error(sweeps, schedule):
    print(f"[case] mixed timestep channel sweeps={sweeps} schedule={schedule}", flush=True)
    Ng, size, origin, Re, T, nsteps, per = (65, 33, 65), (2 * PI, 2.0, PI), (0.0, -1.0, 0.0), 1e3, 5e-3, 10, (1, 0, 1)
    og, gs = grids(Ng, size, origin, Re, T, nsteps, per)
    o = mo.Oracle(og)
    vel, buf, buf2 = o.velocity(), o.velocity(), o.velocity()
    p, pb = o.zeros(3), o.zeros(3)
    v0 = channel_initial(o, 9)
    for c in range(3):
        vel[c][...] = v0[c]
    ctx = mif.Context(gs, precision=mif.MIXED)
    ctx.set_defect_sweeps(sweeps)
    ctx.set_defect_schedule(schedule)
    gvel, gbuf, gbuf2 = ctx.velocity(), ctx.velocity(), ctx.velocity()
    gp, gpb = ctx.pressure(), ctx.pressure()
    gvel.upload(vel)
    ev = mo.make_fn(mo.FN_ZERO)
    fn = mif.function(mif.FN_ZERO)
    for s in range(nsteps):
        o.timestep(vel, buf, buf2, ev, s * o.dt, p, pb)
        mif.timestep(gvel, gbuf, gbuf2, fn, s * o.dt, gp, gpb)
    got = gvel.download()
    e = [l2rel(got[c], vel[c]) for c in range(3)]
    ev_all = float(np.sqrt(sum(np.linalg.norm(got[c] - vel[c]) ** 2 for c in range(3)) /
                           sum(np.linalg.norm(vel[c]) ** 2 for c in range(3))))
    print("per-component", e, "velocity", ev_all, "pressure", l2rel(gp.download(), p))
    assert ev_all < 1e-6


@pytest.mark.parametrize("Ng,per", [((65, 33, 65), (1, 0, 1)), ((33, 33, 65), (0, 0, 1)),
                                    # x lines of 512 and 768 points: two and three points per thread of
                                    # the x passes; 257 Neumann points: a partial second point
                                    ((513, 9, 17), (1, 0, 1)), ((769, 5, 9), (1, 0, 1)), ((257, 9, 17), (0, 0, 1))],
                         ids=["channel", "walls_x", "x512", "x768", "x257_walls"])
def test_fused_defect_sweep_solve_tracks_fp64(Ng, per):
    """The one-sweep form fused into the transform passes (the x forward pass
    of the correction solve forms the FP64 defect as it loads) against an FP64
    solve of the same FP32 velocity.  The mixed form stores the pressure in
    FP32, so the floor is the FP64 solution rounded to FP32: the swept forms
    (fused one-sweep, unfused two-sweep) reach it, the no-sweep form stays
    well above it."""
    print(f"[case] fused defect sweep, stand-alone solve {Ng} {per}", flush=True)
    og, gs = grids(Ng, (2 * PI, 2.0, PI), (0.0, -1.0, 0.0), 1.0, 1.0, 1, per)
    dt = 1e-3
    sols = {}
    vel32 = None
    for name, prec, sweeps in (("fp64", mif.FP64, 0), ("s0", mif.MIXED, 0), ("s1", mif.MIXED, 1),
                               ("s2", mif.MIXED, 2)):
        ctx = mif.Context(gs, precision=prec)
        if prec == mif.MIXED:
            ctx.set_defect_sweeps(sweeps)
        vel, p = ctx.velocity(), ctx.pressure()
        if vel32 is None:
            # random velocity, rounded to FP32 so that every form solves
            # for the same divergence
            shapes = [a.shape for a in vel.download()]
            vel32 = [seeded_normal(sh, 20261210 + c).astype(np.float32).astype(np.float64)
                     for c, sh in enumerate(shapes)]
        vel.upload(vel32)
        mif.solve_pressure_equation_homogeneous_periodic(p, vel, dt)
        phi = p.download()
        if name == "fp64":
            r32 = phi.astype(np.float32).astype(np.float64)
            sols["fp64_as_fp32"] = r32 - r32.mean()
        sols[name] = phi - phi.mean()
        ctx.close()
    floor = l2rel(sols["fp64_as_fp32"], sols["fp64"])
    errs = {k: l2rel(sols[k], sols["fp64"]) for k in ("s0", "s1", "s2")}
    print("errors vs fp64", errs, "FP32 storage floor", floor, flush=True)
    assert errs["s0"] > 5 * floor
    assert errs["s1"] < 1.5 * floor
    assert errs["s2"] < 1.5 * floor


def test_defect_schedule_leaves_standalone_solve():
    # the schedule selects sweeps inside mif_timestep only: a stand-alone
    # solve takes the context's sweeps under either schedule, bit for bit
    print("[case] defect schedule, stand-alone solve 65x33x65", flush=True)
    og, gs = grids((65, 33, 65), (2 * PI, 2.0, PI), (0.0, -1.0, 0.0), 1.0, 1.0, 1, (1, 0, 1))
    o = mo.Oracle(og)
    b = seeded_normal(tuple(reversed(o.npts)), 20261203)
    ctx = mif.Context(gs, precision=mif.MIXED)
    ctx.set_defect_sweeps(2)
    ctx.poisson_set_rhs(b)
    sols = []
    for sched in (mif.DEFECT_EVERY_STAGE, mif.DEFECT_LAST_STAGE):
        ctx.set_defect_schedule(sched)
        ctx.poisson_run()
        sols.append(ctx.poisson_get_solution())
    assert np.array_equal(sols[0], sols[1])
    with pytest.raises(mif.MifError):
        ctx.set_defect_schedule(2)


def _es_nhn(N, prec, sweeps, sched, nsteps, dt, Re):
    gs = mif.GridSpec(N, N, N, 1.0, 1.0, 2.0, 0.0, 0.0, -1.0, Re, dt * nsteps, nsteps, (False, False, False))
    ctx = mif.Context(gs, precision=prec)
    if prec == mif.MIXED:
        ctx.set_defect_sweeps(sweeps)
        ctx.set_defect_schedule(sched)
    vel, buf, buf2 = ctx.velocity(), ctx.velocity(), ctx.velocity()
    p, pb = ctx.pressure(), ctx.pressure()
    ev, gp = mif.function(mif.FN_ES_VELOCITY, Re=Re), mif.function(mif.FN_ES_GRAD_P, Re=Re)
    vel.set(ev, 0.0)
    p.set(mif.function(mif.FN_ES_PRESSURE, Re=Re), 0, 0.0)
    for s in range(nsteps):
        mif.timestep_nhn(vel, buf, buf2, ev, gp, s * ctx.dt, p, pb)
    out = [c.astype(np.float64) for c in vel.download()], p.download().astype(np.float64)
    ctx.close()
    return out


@pytest.mark.parametrize("N", [33, 65])
def test_mixed_pressure_error_es(N):
    """Ethier-Steinman on [0,1]x[0,1]x[-1,1], all walls, Re = 1e3, with its
    O(1) pressure (rms 1.07): every mixed form (no sweep, a sweep per stage,
    a sweep in the last stage) against the FP64 form, 10 steps.  The
    pressure difference, ~1e-5 relative, is the FP32 velocity's storage
    rounding carried into each stage's pressure increment (order
    eps32 |u| h / dt_s); it is the same with and without the FP64 sweeps
    (measured 0.84e-5 .. 1.1e-5 at N = 33, 1.1e-5 .. 1.6e-5 at N = 65)."""
    print(f"[case] mixed ES pressure N={N}", flush=True)
    nsteps, dt, Re = 10, 1e-3 * 33 / N, 1e3
    rv, rp = _es_nhn(N, mif.FP64, 0, 0, nsteps, dt, Re)
    rp = rp - rp.mean()
    for name, sw, sc in (("s0", 0, 0), ("s1", 1, 0), ("s1_last", 1, 1)):
        v, p = _es_nhn(N, mif.MIXED, sw, sc, nsteps, dt, Re)
        ev = float(np.sqrt(sum(np.linalg.norm(a - b) ** 2 for a, b in zip(v, rv)) /
                           sum(np.linalg.norm(b) ** 2 for b in rv)))
        ep = l2rel(p - p.mean(), rp)
        print(f"  {name}: velocity {ev:.3e} pressure {ep:.3e}", flush=True)
        assert ev < 1e-6, name
        assert ep < 5e-5, name


def test_mixed_timestep_velocity_error():
    print("[case] mixed timestep_velocity", flush=True)
    Re = 1e4
    og, gs = grids((33, 33, 33), (1.0, 1.0, 1.0), (0.0, 0.0, 0.0), Re, 1e-3, 5, (0, 0, 0))
    o = mo.Oracle(og)
    ev = mo.make_fn(mo.FN_VT_VELOCITY, Re=Re)
    fo = mo.make_fn(mo.FN_VT_FORCING, Re=Re)
    vel, buf, rhs = o.velocity(), o.velocity(), o.velocity()
    o.set_velocity(vel, ev, 0.0)
    ctx = mif.Context(gs, precision=mif.MIXED)
    gvel, gbuf, grhs = ctx.velocity(), ctx.velocity(), ctx.velocity()
    gvel.upload(vel)
    gev, gfo = mif.function(mif.FN_VT_VELOCITY, Re=Re), mif.function(mif.FN_VT_FORCING, Re=Re)
    for s in range(5):
        o.timestep_velocity(vel, buf, rhs, ev, fo, s * o.dt)
        mif.timestep_velocity(gvel, gbuf, grhs, gev, gfo, s * o.dt)
    got = gvel.download()
    for c in range(3):
        assert rel_err(got[c], vel[c]) < 1e-6, c


def test_solve_precision_selection():
    """The condition-number estimate of the discrete Laplacian and the
    automatic FP32/FP64 choice of the mixed form's transforms."""
    import math

    from mif_amd.problems import channel_grid

    n = 64
    gs = channel_grid(n, n, n)
    ctx = mif.Context(gs, precision=mif.MIXED)
    eff, kappa = ctx.solve_precision()
    # per axis: max |lambda| = 4/h^2 (k = N/2 periodic, k = n-1 Neumann);
    # smallest non-zero |lambda| = 2(1 - cos(2 pi/N))/h^2 on a periodic axis
    # of N points, 2(1 - cos(pi/(n-1)))/h^2 on a wall axis of n nodes
    hx, hy, hz = 2 * math.pi / n, 2.0 / n, math.pi / n
    lam_max = 4 / hx ** 2 + 4 / hy ** 2 + 4 / hz ** 2
    lam_min = min(2 * (1 - math.cos(2 * math.pi / n)) / hx ** 2, 2 * (1 - math.cos(math.pi / n)) / hy ** 2,
                  2 * (1 - math.cos(2 * math.pi / n)) / hz ** 2)
    assert eff == mif.SOLVE_FP32
    assert abs(kappa / (lam_max / lam_min) - 1) < 1e-9, (kappa, lam_max / lam_min)
    ctx.set_solve_precision(mif.SOLVE_FP64)
    assert ctx.solve_precision()[0] == mif.SOLVE_FP64
    ctx.close()


@pytest.mark.parametrize("nx", [32, 64])
def test_mixed_with_fp64_transforms_tracks_fp64(nx):
    """FP32 fields and stencils with FP64 transforms stay within FP32
    rounding of the FP64 form: x length 32 (separate divergence and unpack
    kernels) and 64 (both fused into the FP64 x passes, which read and write
    the FP32 fields)."""
    from mif_amd.problems import channel_grid, channel_initial_velocity

    print(f"[case] mixed with FP64 transforms, {nx}x32x32", flush=True)
    gs = channel_grid(nx, 32, 32, Re=1000.0, dt=5e-4, steps=5)
    init = channel_initial_velocity(gs)
    out = {}
    for name, prec in (("fp64", mif.FP64), ("mixed64", mif.MIXED)):
        ctx = mif.Context(gs, precision=prec)
        if prec == mif.MIXED:
            ctx.set_solve_precision(mif.SOLVE_FP64)
        vel, buf, buf2 = ctx.velocity(), ctx.velocity(), ctx.velocity()
        p, pb = ctx.pressure(), ctx.pressure()
        vel.upload(init)
        for s in range(5):
            mif.timestep(vel, buf, buf2, mif.function(mif.FN_ZERO), s * ctx.dt, p, pb)
        out[name] = vel.download()
        ctx.close()
    num = sum(float(np.linalg.norm(a - b)) ** 2 for a, b in zip(out["mixed64"], out["fp64"])) ** 0.5
    den = sum(float(np.linalg.norm(b)) ** 2 for b in out["fp64"]) ** 0.5
    assert num / den < 1e-6, num / den
