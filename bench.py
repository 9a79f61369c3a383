#!/usr/bin/env python3
"""Timestep throughput on one MI355X.

Workload: the channel at 256^3 (256 intervals per axis, N_g = 257 nodes,
periodic x and z, no-slip walls in y, domain [0,2pi]x[-1,1]x[0,pi], Re = 1000,
dt = 5e-4), Poiseuille profile plus a seeded divergence-free perturbation of
amplitude 1e-2, all fields resident in HBM before the timed region.  One step
is a full RK3 timestep with a pressure projection per stage.  The metric is
cell-updates/s = pressure unknowns (256 x 257 x 256) x steps / seconds.

Prints one JSON line.  `--gpus N` with N > 1 exits with an explanation: this
build runs the whole domain on one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "mpi-incompressible-fluid_amd"))

METRIC = "cell-updates/s per timestep @256^3: mixed vs FP64, equal error; adjoint/fwd cost"
HBM_PEAK_GBS = 8000.0

# Algorithmic words per pressure unknown per launch of each kernel class:
# each logical array read once and written once per kernel (stencil
# neighbour re-reads and ghost planes counted once; O(N^2) boundary work
# excluded).  Stage launches alternate 10, 13, 10 words (mean 11).
WORDS_PER_LAUNCH = {
    "stage": 11.0,
    "div_rhs": 4.0,
    "fft_x": 2.0,
    "fft_y": 2.0,
    "fft_z_fused": 2.0,
    "unpack": 2.0,
    "project": 9.0,
    "fft_x_div": 4.0,
    "fft_x_unpack": 2.0,
}


def run_gpu(n: int, steps: int, warmup: int, precision: int):
    import mif_amd as mif
    from mif_amd.problems import channel_grid, channel_initial_velocity

    gs = channel_grid(n, n, n, Re=1000.0, dt=5e-4, steps=max(steps + warmup, 1))
    ctx = mif.Context(gs, precision=precision, device=0)
    vel, buf, buf2 = ctx.velocity(), ctx.velocity(), ctx.velocity()
    p, pb = ctx.pressure(), ctx.pressure()
    vel.upload(channel_initial_velocity(gs))
    ev = mif.function(mif.FN_ZERO)
    t = 0.0
    for _ in range(warmup):
        mif.timestep(vel, buf, buf2, ev, t, p, pb)
        t += ctx.dt
    ctx.synchronize()
    ctx.set_timing(True)
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        mif.timestep(vel, buf, buf2, ev, t, p, pb)
        t += ctx.dt
    ctx.synchronize()
    t1 = time.perf_counter()
    kernels = {}
    for cls, name in enumerate(mif.KERNEL_CLASSES):
        ms, cnt = ctx.kernel_time(cls)
        if cnt:
            kernels[name] = {"total_ms": ms, "launches": cnt, "avg_ms": ms / cnt}
    ctx.set_timing(False)
    unknowns = n * (n + 1) * n
    esize = 8 if precision == mif.FP64 else 4
    u = vel.u.download()
    finite = bool(abs(u).max() < 1e30)
    return {
        "seconds": t1 - t0,
        "unknowns": unknowns,
        "esize": esize,
        "kernels": kernels,
        "finite": finite,
    }


def roofline_of(res):
    """Roofline of the kernel class with the largest total time."""
    best = None
    for name, k in res["kernels"].items():
        if name not in WORDS_PER_LAUNCH:
            continue
        if best is None or k["total_ms"] > res["kernels"][best]["total_ms"]:
            best = name
    if best is None:
        return None, {}
    table = {}
    for name, k in res["kernels"].items():
        if name not in WORDS_PER_LAUNCH:
            continue
        bytes_per_launch = WORDS_PER_LAUNCH[name] * res["unknowns"] * res["esize"]
        gbs = bytes_per_launch / (k["avg_ms"] * 1e-3) / 1e9
        table[name] = {"avg_us": round(k["avg_ms"] * 1e3, 2), "launches": k["launches"],
                       "alg_bytes": int(bytes_per_launch), "GB/s": round(gbs, 1),
                       "frac": round(gbs / HBM_PEAK_GBS, 4)}
    k = table[best]
    roof = {"kernel": best, "bound": "hbm", "achieved": k["GB/s"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": k["frac"], "traffic": None, "alg_bytes_per_launch": k["alg_bytes"]}
    return roof, table


def cpu_baseline(n: int, steps: int):
    """The CPU oracle (this repository's C restatement of the reference
    algorithm, OpenMP over the host cores) on a bounded sample of the same
    workload."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np

    import mif_oracle as mo
    from mif_amd.problems import channel_grid, channel_initial_velocity

    gs = channel_grid(n, n, n, Re=1000.0, dt=5e-4, steps=steps)
    og = mo.make_grid((gs.Nx_global, gs.Ny_global, gs.Nz_global), (gs.x_size, gs.y_size, gs.z_size),
                      (gs.min_x, gs.min_y, gs.min_z), gs.Re, gs.final_time, gs.num_time_steps, gs.periodic)
    o = mo.Oracle(og)
    vel = mo.Fields(*[np.ascontiguousarray(a) for a in channel_initial_velocity(gs)])
    buf, buf2 = o.velocity(), o.velocity()
    p, pb = o.zeros(3), o.zeros(3)
    ev = mo.make_fn(mo.FN_ZERO)
    t0 = time.perf_counter()
    for s in range(steps):
        o.timestep(vel, buf, buf2, ev, s * o.dt, p, pb)
    t1 = time.perf_counter()
    unknowns = n * (n + 1) * n
    return {"value": unknowns * steps / (t1 - t0), "unit": "cell-updates/s", "cores": mo.num_threads(),
            "kind": "port", "sample": f"{steps} FP64 steps of the same {n}^3 channel on the host CPU "
                                      f"({t1 - t0:.1f} s)"}


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=256, help="intervals per axis")
    ap.add_argument("--cpu-steps", type=int, default=2, help="steps of the CPU baseline sample (0 = skip)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: --gpus {args.gpus} is not supported: this build holds the whole domain on one GPU "
              f"(no ranks, no collectives); run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import mif_amd as mif

    fp64 = run_gpu(args.n, args.steps, args.warmup, mif.FP64)
    value = fp64["unknowns"] * args.steps / fp64["seconds"]
    roof, table = roofline_of(fp64)
    step_bytes = 108 * fp64["unknowns"] * fp64["esize"]
    step_gbs = step_bytes * args.steps / fp64["seconds"] / 1e9
    cpu = cpu_baseline(args.n, args.cpu_steps) if args.cpu_steps > 0 else None
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": fp64["seconds"] / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"channel_timestep_{args.n}^3", "grid_nodes": [args.n + 1] * 3,
                   "pressure_unknowns": fp64["unknowns"], "periodic": "x,z", "walls": "y", "Re": 1000.0,
                   "dt": 5e-4, "form": "fp64"},
        "roofline": roof,
        "cpu_baseline": cpu,
        "step_roofline": {"alg_bytes_per_step": step_bytes, "achieved": round(step_gbs, 1), "unit": "GB/s",
                          "frac": round(step_gbs / HBM_PEAK_GBS, 4)},
        "kernels": table,
        "finite": fp64["finite"],
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
