"""Adjoint of timestep() on the GPU.

1. Parity: mif_timestep_adjoint against the oracle's adjoint (itself pinned
   by central differences in test_oracle_adjoint.py) on every boundary type,
   same seeded inputs, 1e-11 relative.
2. Gradient: the checkpointed torch.autograd rollout (mif_amd.autograd)
   against central finite differences of the GPU forward map along 16 seeded
   directions (eps = 1e-6, FP64), 1e-5 relative.
"""
import numpy as np
import pytest

import mif_oracle as mo
from helpers import PI, grids, rel_err

pytestmark = pytest.mark.gpu

mif = pytest.importorskip("mif_amd")
torch = pytest.importorskip("torch")

CASES = [
    ((9, 7, 9), (2 * PI, 2.0, PI), (0.0, -1.0, 0.0), (1, 0, 1)),
    ((7, 6, 8), (1.0, 1.0, 2.0), (0.0, 0.0, -1.0), (0, 0, 0)),
    ((7, 8, 6), (2 * PI, 2 * PI, 1.0), (0.0, 0.0, 0.0), (1, 1, 0)),
    ((6, 7, 8), (2 * PI, 2 * PI, 2 * PI), (0.0, 0.0, 0.0), (1, 1, 1)),
    ((33, 17, 33), (2 * PI, 2.0, PI), (0.0, -1.0, 0.0), (1, 0, 1)),
]


def _print_case(name):
    print(f"[case] {name}", flush=True)


def _oracle_adjoint(o, u0, p0, f, wu, wp, nsteps):
    o.set_body_force(f)
    vel, p = u0.copy(), p0.copy()
    buf, buf2 = o.velocity(), o.velocity()
    pb = o.zeros(3)
    ev = mo.make_fn(mo.FN_ZERO)
    rec = []
    for s in range(nsteps):
        un = vel.copy()
        o.timestep(vel, buf, buf2, ev, s * o.dt, p, pb)
        rec.append((un, buf.copy(), buf2.copy()))
    ubar = mo.Fields(*[w.copy() for w in wu])
    pbar = wp.copy()
    fbar = np.zeros(3)
    for un, y2, y3 in reversed(rec):
        o.timestep_adjoint(un, y2, y3, ubar, pbar, fbar)
    return vel, p, ubar, pbar, fbar


def _gpu_adjoint(gs, u0, p0, f, wu, wp, nsteps, device_force=False):
    ctx = mif.Context(gs)
    ctx.set_body_force(f)
    ctx.set_stage_records(True)
    vel, buf, buf2 = ctx.velocity(), ctx.velocity(), ctx.velocity()
    p, pb = ctx.pressure(), ctx.pressure()
    vel.upload([u0[c] for c in range(3)])
    p.upload(p0)
    ev = mif.function(mif.FN_ZERO)
    rec = []
    for s in range(nsteps):
        un = ctx.velocity()
        un.copy_from(vel)
        mif.timestep(vel, buf, buf2, ev, s * ctx.dt, p, pb)
        y2, y3 = ctx.velocity(), ctx.velocity()
        y2.copy_from(buf)
        y3.copy_from(buf2)
        rec.append((un, y2, y3))
    ubar, pbar = ctx.velocity(), ctx.pressure()
    ubar.upload(wu)
    pbar.upload(wp)
    fbar = np.zeros(3)
    if device_force:
        mif.adjoint_force_take(ctx)
    for un, y2, y3 in reversed(rec):
        mif.timestep_adjoint(un, y2, y3, ubar, pbar, None if device_force else fbar)
    if device_force:
        fbar = mif.adjoint_force_take(ctx)
        assert np.all(mif.adjoint_force_take(ctx) == 0.0), "the take resets the accumulator"
    return vel.download(), p.download(), ubar.download(), pbar.download(), fbar


@pytest.mark.parametrize("case", CASES, ids=lambda c: "x".join(map(str, c[0])) + "-p" + "".join(map(str, c[3])))
def test_adjoint_matches_oracle(case):
    Ng, size, origin, per = case
    _print_case(f"adjoint {Ng} periodic={per}")
    nsteps = 2
    rng = np.random.default_rng(7)
    og, gs = grids(Ng, size, origin, 50.0, 0.02, nsteps, per)
    o = mo.Oracle(og)
    u0 = o.velocity()
    for c in range(3):
        u0[c][...] = rng.standard_normal(u0[c].shape)
    p0 = rng.standard_normal(o.shape(3))
    f = rng.standard_normal(3)
    wu = [rng.standard_normal(u0[c].shape) for c in range(3)]
    wp = rng.standard_normal(o.shape(3))
    rv, rp, rub, rpb, rfb = _oracle_adjoint(o, u0, p0, f, wu, wp, nsteps)
    gv, gp, gub, gpb, gfb = _gpu_adjoint(gs, u0, p0, f, wu, wp, nsteps)
    for c in range(3):
        assert rel_err(gv[c], rv[c]) < 1e-11, ("forward u", c)
    assert rel_err(gp, rp) < 1e-11, "forward p"
    for c in range(3):
        assert rel_err(gub[c], rub[c]) < 1e-11, ("u_bar", c, rel_err(gub[c], rub[c]))
    assert rel_err(gpb, rpb) < 1e-11, ("p_bar", rel_err(gpb, rpb))
    assert np.max(np.abs(gfb - rfb)) <= 1e-11 * np.max(np.abs(rfb)), (gfb, rfb)
    # the device accumulator (force_bar = NULL, then a take) sums the same
    # per-step values in the same order as the host
    dv, dp, dub, dpb, dfb = _gpu_adjoint(gs, u0, p0, f, wu, wp, nsteps, device_force=True)
    assert np.array_equal(dfb, gfb), (dfb, gfb)
    for c in range(3):
        assert np.array_equal(dub[c], gub[c]), ("u_bar, device force accumulation", c)


def test_rollout_gradient_matches_central_differences():
    from mif_amd.autograd import Rollout

    Ng, size, origin, per = (17, 17, 17), (2 * PI, 2.0, PI), (0.0, -1.0, 0.0), (1, 0, 1)
    nsteps, k = 5, 2
    _print_case(f"rollout FD {Ng} steps={nsteps} checkpoint={k}")
    og, gs = grids(Ng, size, origin, 100.0, 0.05, nsteps, per)
    ctx = mif.Context(gs)
    roll = Rollout(ctx, nsteps, checkpoint_every=k)
    rng = np.random.default_rng(11)
    dev = torch.device("cuda")
    u0 = [torch.tensor(0.3 * rng.standard_normal(s), device=dev, requires_grad=True) for s in roll.shapes]
    f = torch.tensor(rng.standard_normal(3), dtype=torch.float64, requires_grad=True)
    w = [torch.tensor(rng.standard_normal(s), device=dev) for s in roll.shapes]

    def J(us, ff):
        out = roll(us, ff)
        return sum((o_ * w_).sum() for o_, w_ in zip(out, w))

    val = J(u0, f)
    val.backward()
    g_u = [t.grad.detach().cpu().numpy() for t in u0]
    g_f = f.grad.detach().numpy()
    eps = 1e-6
    worst = 0.0
    with torch.no_grad():
        for d in range(16):
            du = [rng.standard_normal(s) for s in roll.shapes]
            df = rng.standard_normal(3)
            up = [t.detach() + eps * torch.tensor(x, device=dev) for t, x in zip(u0, du)]
            um = [t.detach() - eps * torch.tensor(x, device=dev) for t, x in zip(u0, du)]
            fp = f.detach() + eps * torch.tensor(df)
            fm = f.detach() - eps * torch.tensor(df)
            fd = (float(J(up, fp)) - float(J(um, fm))) / (2 * eps)
            ad = sum(float((g * x).sum()) for g, x in zip(g_u, du)) + float(g_f @ df)
            err = abs(ad - fd) / max(abs(fd), 1e-30)
            worst = max(worst, err)
            assert err <= 1e-5, (d, ad, fd, err)
    print(f"[rollout FD] worst relative error over 16 directions: {worst:.3e}", flush=True)


def test_rollout_matches_uncheckpointed_adjoint():
    """Checkpoint interval must not change the gradient (recompute is exact)."""
    from mif_amd.autograd import Rollout

    Ng, size, origin, per = (17, 9, 17), (2 * PI, 2.0, PI), (0.0, -1.0, 0.0), (1, 0, 1)
    nsteps = 7
    _print_case(f"rollout checkpoint invariance {Ng} steps={nsteps}")
    og, gs = grids(Ng, size, origin, 100.0, 0.05, nsteps, per)
    ctx = mif.Context(gs)
    rng = np.random.default_rng(5)
    dev = torch.device("cuda")
    shapes = None
    grads = []
    for k in (1, 3, 8):
        roll = Rollout(ctx, nsteps, checkpoint_every=k)
        shapes = roll.shapes
        r = np.random.default_rng(5)
        u0 = [torch.tensor(0.3 * r.standard_normal(s), device=dev, requires_grad=True) for s in shapes]
        f = torch.tensor(r.standard_normal(3), dtype=torch.float64, requires_grad=True)
        w = [torch.tensor(r.standard_normal(s), device=dev) for s in shapes]
        out = roll(u0, f)
        sum((o_ * w_).sum() for o_, w_ in zip(out, w)).backward()
        grads.append(([t.grad.cpu().numpy() for t in u0], f.grad.numpy().copy()))
    del rng
    for gu, gf in grads[1:]:
        for c in range(3):
            assert np.array_equal(gu[c], grads[0][0][c])
        assert np.array_equal(gf, grads[0][1])


def test_channel_128_gradient_fd_and_fp32():
    """The gradient workload at its full size: 128^3 channel, Re = 500, 32
    steps, checkpoint every 8, loss 0.5 ||u(T) - u_target||^2; FP64 gradient
    against central differences on 16 seeded directions (eps = 1e-6, 1e-5
    relative), and the FP32 adjoint against the FP64 one."""
    from mif_amd.autograd import Rollout
    from mif_amd.problems import adjoint_channel

    n, steps, k = 128, 32, 8
    _print_case(f"channel gradient {n}^3 steps={steps} checkpoint={k}")
    gs, u0_np, ualt_np, f_np = adjoint_channel(n, steps=steps)
    f_np = np.asarray(f_np)
    dev = torch.device("cuda")
    out = {}
    for name, prec in (("fp64", mif.FP64), ("fp32", mif.MIXED)):
        ctx = mif.Context(gs, precision=prec)
        roll = Rollout(ctx, steps, checkpoint_every=k)
        target = [t.clone() for t in roll.run_forward([torch.tensor(a, device=dev) for a in ualt_np], f_np)]

        def loss_of(us, ff, roll=roll, target=target):
            o = roll(us, ff)
            return 0.5 * sum(((a - b) ** 2).sum() for a, b in zip(o, target))

        u0 = [torch.tensor(a, device=dev, requires_grad=True) for a in u0_np]
        f = torch.tensor(f_np, requires_grad=True)
        loss_of(u0, f).backward()
        out[name] = ([t.grad.detach().clone() for t in u0], f.grad.detach().numpy().copy(), roll, loss_of, u0)
    gu, gf, roll, loss_of, u0 = out["fp64"]
    rng = np.random.default_rng(20261202)
    eps = 1e-6
    worst = 0.0
    with torch.no_grad():
        for d in range(16):
            du = [torch.tensor(rng.standard_normal(s), device=dev) for s in roll.shapes]
            df = rng.standard_normal(3)
            jp = float(loss_of([a.detach() + eps * x for a, x in zip(u0, du)], torch.tensor(f_np + eps * df)))
            jm = float(loss_of([a.detach() - eps * x for a, x in zip(u0, du)], torch.tensor(f_np - eps * df)))
            fd = (jp - jm) / (2 * eps)
            ad = sum(float((g * x).sum()) for g, x in zip(gu, du)) + float(gf @ df)
            err = abs(ad - fd) / abs(fd)
            worst = max(worst, err)
            assert err <= 1e-5, (d, ad, fd)
    gu32, gf32 = out["fp32"][0], out["fp32"][1]
    num = sum(float(((a - b) ** 2).sum()) for a, b in zip(gu32, gu))
    den = sum(float((b ** 2).sum()) for b in gu)
    rel32 = (num / den) ** 0.5
    relf = float(np.linalg.norm(gf32 - gf) / np.linalg.norm(gf))
    print(f"[channel gradient] FD worst {worst:.3e}; fp32 vs fp64: u0 {rel32:.3e}, f {relf:.3e} "
          f"(fp64 df {gf}, fp32 df {gf32})", flush=True)
    # The FP32 variant stores the adjoint fields in FP32: the u0 gradient is a
    # field and keeps ~1e-4; the force gradient is the sum of the adjoint
    # velocity over all points and steps, which cancels to a small fraction of
    # its terms, so its relative error is that of the field over the
    # cancellation ratio.
    assert rel32 < 1e-3
    assert relf < 5e-2
