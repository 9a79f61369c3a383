"""The timestep with the projections of stages 1 and 2 fused into the next
stage kernel (the default) against the unfused sequence (stage records on:
every projection its own pass, as src/Timestep.cpp:97-141 orders them), on
the same inputs.  Both evaluate the same expressions in the same order, so
FP64 results agree to a few units of rounding; the oracle parity of the fused
path itself is tested in test_gpu_parity.py (1e-11)."""
import numpy as np
import pytest

import mif_oracle as mo
from helpers import PI, channel_initial, grids, rel_err

pytestmark = pytest.mark.gpu

mif = pytest.importorskip("mif_amd")

CASES = [
    # (Ng, size, origin, periodic, init, nsteps)
    ((33, 17, 33), (2 * PI, 2.0, PI), (0.0, -1.0, 0.0), (1, 0, 1), "channel", 3),  # fused x passes
    ((65, 33, 65), (2 * PI, 2.0, PI), (0.0, -1.0, 0.0), (1, 0, 1), "channel", 2),
    ((16, 16, 16), (1.0, 1.0, 2.0), (0.0, 0.0, -1.0), (0, 0, 0), "es", 3),          # all walls
    ((17, 13, 11), (1.0, 1.0, 1.0), (0.0, 0.0, 0.0), (0, 0, 1), "es", 2),          # generic lengths
]


def _run(gs, o, init, nsteps, fn_kind, records, precision=None, sweeps=None):
    kw = {} if precision is None else {"precision": precision}
    ctx = mif.Context(gs, **kw)
    if sweeps is not None:
        ctx.set_defect_sweeps(sweeps)
    ctx.set_stage_records(records)
    vel, buf, buf2 = ctx.velocity(), ctx.velocity(), ctx.velocity()
    p, pb = ctx.pressure(), ctx.pressure()
    vel.upload(init[0])
    p.upload(init[1])
    fn = mif.function(fn_kind, Re=1e3)
    for s in range(nsteps):
        mif.timestep(vel, buf, buf2, fn, s * ctx.dt, p, pb)
    out = [c for c in vel.download()], p.download()
    ctx.close()
    return out


def _initial(case):
    Ng, size, origin, per, init, nsteps = case
    og, gs = grids(Ng, size, origin, 1e3, 1e-3 if init == "channel" else 1e-4, nsteps, per)
    o = mo.Oracle(og)
    vel = o.velocity()
    p = o.zeros(3)
    if init == "es":
        o.set_velocity(vel, mo.make_fn(mo.FN_ES_VELOCITY, Re=1e3), 0.0)
        o.set(p, 3, mo.make_fn(mo.FN_ES_PRESSURE, Re=1e3), 0, 0.0)
        fn = mif.FN_ES_VELOCITY
    else:
        v0 = channel_initial(o, 13)
        for c in range(3):
            vel[c][...] = v0[c]
        fn = mif.FN_ZERO
    return gs, o, ([vel[c] for c in range(3)], p), nsteps, fn


@pytest.mark.parametrize("case", CASES, ids=lambda c: "x".join(map(str, c[0])) + "-p" + "".join(map(str, c[3])))
def test_fused_matches_unfused_fp64(case):
    print(f"[case] fused projection FP64 {case[0]} periodic={case[3]}", flush=True)
    gs, o, init, nsteps, fn = _initial(case)
    (fv, fp) = _run(gs, o, init, nsteps, fn, records=False)
    (uv, up) = _run(gs, o, init, nsteps, fn, records=True)
    for c in range(3):
        assert rel_err(fv[c], uv[c]) < 1e-13, c
    assert rel_err(fp, up) < 1e-12


@pytest.mark.parametrize("sweeps", [0, 1])
def test_fused_matches_unfused_mixed(sweeps):
    case = CASES[1]
    print(f"[case] fused projection mixed {case[0]} sweeps={sweeps}", flush=True)
    gs, o, init, nsteps, fn = _initial(case)
    (fv, fp) = _run(gs, o, init, nsteps, fn, records=False, precision=mif.MIXED, sweeps=sweeps)
    (uv, up) = _run(gs, o, init, nsteps, fn, records=True, precision=mif.MIXED, sweeps=sweeps)
    # FP32 fields (the fused kernel stages phi through LDS): the two sequences
    # may round the same expressions apart by an ulp or two per stage (fused
    # multiply-adds chosen per kernel)
    for c in range(3):
        assert rel_err(fv[c], uv[c]) < 1e-6, c
        assert np.all(np.isfinite(fv[c]))
    assert rel_err(fp, up) < 1e-4


def _run_n(gs, init, nsteps, fn_kind, precision=None, records=False, multi=True, sweeps=None, schedule=None):
    kw = {} if precision is None else {"precision": precision}
    ctx = mif.Context(gs, **kw)
    ctx.set_stage_records(records)
    if sweeps is not None:
        ctx.set_defect_sweeps(sweeps)
    if schedule is not None:
        ctx.set_defect_schedule(schedule)
    vel, buf, buf2 = ctx.velocity(), ctx.velocity(), ctx.velocity()
    p, pb = ctx.pressure(), ctx.pressure()
    vel.upload(init[0])
    p.upload(init[1])
    fn = mif.function(fn_kind, Re=1e3)
    if multi:
        mif.timestep_n(vel, buf, buf2, fn, 0.0, p, pb, nsteps)
    else:
        for s in range(nsteps):
            mif.timestep(vel, buf, buf2, fn, s * ctx.dt, p, pb)
    out = [c for c in vel.download()], p.download()
    ctx.close()
    return out


@pytest.mark.parametrize("nsteps", [1, 2, 3, 4])
@pytest.mark.parametrize("ci", [0, 2], ids=["channel", "walls_es"])
@pytest.mark.parametrize("form", ["fp64", "mixed", "mixed_s1", "mixed_s1_last"])
def test_timestep_n_matches_loop(form, ci, nsteps):
    """mif_timestep_n (each step's last projection in the next step's first
    stage kernel; the pressure ends in p for either parity of the buffer
    swaps) against the loop of single steps; the mixed forms without defect
    sweeps, with the fused sweep in every solve and in each step's last solve."""
    case = CASES[ci]
    print(f"[case] timestep_n {form} {case[0]} nsteps={nsteps}", flush=True)
    gs, o, init, _, fn = _initial(case)
    prec = None if form == "fp64" else mif.MIXED
    sweeps = {"fp64": None, "mixed": 0}.get(form, 1)
    schedule = mif.DEFECT_LAST_STAGE if form == "mixed_s1_last" else None
    (nv, np_) = _run_n(gs, init, nsteps, fn, prec, sweeps=sweeps, schedule=schedule)
    (lv, lp) = _run_n(gs, init, nsteps, fn, prec, multi=False, sweeps=sweeps, schedule=schedule)
    tv, tp = (1e-13, 1e-12) if form == "fp64" else (1e-6, 1e-4)
    for c in range(3):
        assert rel_err(nv[c], lv[c]) < tv, c
    assert rel_err(np_, lp) < tp


def test_timestep_n_with_stage_records_is_the_loop():
    case = CASES[0]
    print(f"[case] timestep_n stage records {case[0]}", flush=True)
    gs, o, init, _, fn = _initial(case)
    (nv, np_) = _run_n(gs, init, 3, fn, records=True)
    (lv, lp) = _run_n(gs, init, 3, fn, records=True, multi=False)
    for c in range(3):
        assert np.array_equal(nv[c], lv[c]), c
    assert np.array_equal(np_, lp)


def test_timestep_n_zero_and_negative():
    case = CASES[0]
    print("[case] timestep_n 0 and -1 steps", flush=True)
    gs, o, init, _, fn = _initial(case)
    ctx = mif.Context(gs)
    vel, buf, buf2 = ctx.velocity(), ctx.velocity(), ctx.velocity()
    p, pb = ctx.pressure(), ctx.pressure()
    vel.upload(init[0])
    p.upload(init[1])
    f = mif.function(fn)
    mif.timestep_n(vel, buf, buf2, f, 0.0, p, pb, 0)
    for c, got in enumerate(vel.download()):
        assert np.array_equal(got, init[0][c]), c
    with pytest.raises(mif.MifError):
        mif.timestep_n(vel, buf, buf2, f, 0.0, p, pb, -1)
    ctx.close()
