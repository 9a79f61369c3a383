"""The large single-GPU workloads at their full sizes (all fields resident in
HBM): the 768^3 channel in the mixed form (FP32 fields, FP64 transforms above
the condition-number switch) against the same channel in FP64 over the 20
steps the workload runs, and the 512^3 channel in FP64 against the oracle.
Both take a minute or two and tens of GB of host memory."""
import numpy as np
import pytest

import mif_oracle as mo
from helpers import PI, grids, rel_err

pytestmark = pytest.mark.gpu

mif = pytest.importorskip("mif_amd")

TOL = 1e-11


def _case(msg):
    print(f"[case] {msg}", flush=True)


def _kappa(n, ny):
    h = (2 * PI / n, 2.0 / ny, PI / n)
    return sum(4 / hh ** 2 for hh in h) / (4 * np.sin(PI / n) ** 2 / h[0] ** 2)


def _run(gs, v0, prec, steps):
    ctx = mif.Context(gs, precision=prec, device=0)
    if prec == mif.MIXED:
        ctx.set_defect_sweeps(0)
    vel, buf, buf2 = ctx.velocity(), ctx.velocity(), ctx.velocity()
    p, pb = ctx.pressure(), ctx.pressure()
    vel.upload(v0)
    fn = mif.function(mif.FN_ZERO)
    for s in range(steps):
        mif.timestep(vel, buf, buf2, fn, s * ctx.dt, p, pb)
        if s % 5 == 4:
            ctx.synchronize()
            print(f"  step {s + 1}", flush=True)
    out = [np.asarray(c, dtype=np.float64) for c in vel.download()]
    pres = np.asarray(p.download(), dtype=np.float64)
    solve = ctx.solve_precision()
    ctx.close()
    return out, pres, solve


def test_channel_768cubed_mixed_tracks_fp64_over_20_steps():
    from mif_amd.problems import channel_grid, channel_initial_velocity

    _case("timestep channel 768^3 mixed vs fp64, 20 steps")
    n, steps = 768, 20
    gs = channel_grid(n, n, n, Re=1000.0, dt=5e-4, steps=steps)
    v0 = channel_initial_velocity(gs)
    mixed, _, solve = _run(gs, v0, mif.MIXED, steps)
    assert solve[0] == mif.SOLVE_FP64, "768^3 is above the FP32-transform condition-number limit"
    fp64, _, _ = _run(gs, v0, mif.FP64, steps)
    del v0
    assert all(np.isfinite(c).all() for c in mixed)
    num = sum(np.linalg.norm(a - b) ** 2 for a, b in zip(mixed, fp64))
    den = sum(np.linalg.norm(b) ** 2 for b in fp64)
    err = float(np.sqrt(num / den))
    print("768^3 mixed velocity rel L2 vs fp64 after 20 steps", err, "kappa", solve[1], flush=True)
    assert err < 1e-6


def test_channel_512cubed_fp64_matches_oracle():
    from mif_amd.problems import channel_grid, channel_initial_velocity

    _case("timestep channel 512^3 fp64 vs oracle, 2 steps")
    n, steps = 512, 2
    gs = channel_grid(n, n, n, Re=1000.0, dt=5e-4, steps=steps)
    og = mo.make_grid((gs.Nx_global, gs.Ny_global, gs.Nz_global), (gs.x_size, gs.y_size, gs.z_size),
                      (gs.min_x, gs.min_y, gs.min_z), gs.Re, gs.final_time, gs.num_time_steps, gs.periodic)
    v0 = channel_initial_velocity(gs)
    got, gp, _ = _run(gs, v0, mif.FP64, steps)
    o = mo.Oracle(og)
    vel = mo.Fields(*[np.ascontiguousarray(a) for a in v0])
    del v0
    buf, buf2 = o.velocity(), o.velocity()
    p, pb = o.zeros(3), o.zeros(3)
    ev = mo.make_fn(mo.FN_ZERO)
    for s in range(steps):
        o.timestep(vel, buf, buf2, ev, s * o.dt, p, pb)
        print(f"  oracle step {s + 1}", flush=True)
    errs = [rel_err(got[c], vel[c]) for c in range(3)]
    perr = rel_err(gp, p)
    kappa = _kappa(n, n)
    print("512^3 fp64 velocity", errs, "pressure", perr, "kappa", kappa, flush=True)
    for c in range(3):
        assert errs[c] < TOL, c
    assert perr < TOL * kappa / 1e3
