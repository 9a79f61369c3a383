"""The stand-alone Poisson solve with a cell-centred Neumann y axis (the
DCT-II/DCT-III pair; no reference counterpart) on the GPU: against the
oracle's solve on small grids of every supported length, and at the full
256^3 size (periodic x and z, 256 cells in y) through the residual of the
exact 7-point operator, FP64 and mixed (FP32 transforms with FP64 defect
correction) against FP64."""
import numpy as np
import pytest

import mif_amd as mif
import mif_oracle as mo

pytestmark = pytest.mark.gpu

CELL = (False, True, False)


def gridspec(Ng, per=(True, False, True), size=(2 * np.pi, 2.0, np.pi)):
    return mif.GridSpec(*Ng, *size, 0.0, -1.0, 0.0, 1.0, 1.0, 1, per)


def residual(phi, b, h, per, cell):
    """b - L phi with the exact operator (periodic wrap, cell-centred mirror,
    node Neumann reflection), FP64."""
    r = b.copy()
    for a in range(3):
        ax = 2 - a
        n = phi.shape[ax]
        if per[a]:
            lo, hi = np.roll(phi, 1, axis=ax), np.roll(phi, -1, axis=ax)
        else:
            idx = np.arange(n)
            if cell[a]:
                im, ip = np.maximum(idx - 1, 0), np.minimum(idx + 1, n - 1)
            else:
                im, ip = np.where(idx == 0, 1, idx - 1), np.where(idx == n - 1, n - 2, idx + 1)
            lo, hi = np.take(phi, im, axis=ax), np.take(phi, ip, axis=ax)
        r -= (lo - 2 * phi + hi) / (h[a] * h[a])
    return r


def seeded_rhs(shape, seed=20261202):
    b = np.random.default_rng(seed).standard_normal(shape)
    return b - b.mean()  # uniform weights on periodic and cell-centred axes


@pytest.mark.parametrize("Ng", [(17, 65, 9), (9, 129, 17), (33, 257, 9), (9, 513, 5), (5, 1025, 5)],
                         ids=lambda n: "x".join(map(str, n)))
@pytest.mark.parametrize("precision", [mif.FP64, mif.MIXED], ids=["fp64", "mixed"])
def test_cell_centred_solve_matches_oracle(Ng, precision):
    print(f"[case] cell-centred {Ng} precision {precision}", flush=True)
    gs = gridspec(Ng)
    ctx = mif.Context(gs, precision=precision, device=0, cell_centred=CELL)
    if precision == mif.MIXED:
        ctx.set_defect_sweeps(2)
    n = ctx.npts
    assert n == (Ng[0] - 1, Ng[1] - 1, Ng[2] - 1)
    b = seeded_rhs((n[2], n[1], n[0]), seed=sum(Ng))
    ctx.poisson_set_rhs(b)
    ctx.poisson_run()
    got = ctx.poisson_get_solution()
    ctx.close()
    o = mo.Oracle(mo.make_grid(Ng, (2 * np.pi, 2.0, np.pi), (0.0, -1.0, 0.0), 1.0, 1.0, 1, (1, 0, 1)))
    ref = b.copy()
    o.poisson_packed_cell(ref, CELL)
    err = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
    print(f"[result] {Ng} precision {precision}: {err:.3e}", flush=True)
    assert err <= (1e-11 if precision == mif.FP64 else 1e-9)


def test_cell_centred_256_cubed_residual_and_mixed_error():
    """The 256^3 cell-centred workload: residual of the exact operator at
    rounding level in FP64; the mixed form (FP32 transforms, two FP64 defect
    sweeps) against the FP64 solve."""
    Ng = (257, 257, 257)
    gs = gridspec(Ng)
    out = {}
    for name, prec, sweeps in (("fp64", mif.FP64, 0), ("mixed_s0", mif.MIXED, 0), ("mixed_s2", mif.MIXED, 2)):
        print(f"[case] cell-centred 256^3 {name}", flush=True)
        ctx = mif.Context(gs, precision=prec, device=0, cell_centred=CELL)
        if prec == mif.MIXED:
            ctx.set_defect_sweeps(sweeps)
        n = ctx.npts
        b = seeded_rhs((n[2], n[1], n[0]))
        ctx.poisson_set_rhs(b)
        ctx.poisson_run()
        out[name] = ctx.poisson_get_solution()
        ctx.close()
    h = (2 * np.pi / 256, 2.0 / 256, np.pi / 256)
    r64 = np.linalg.norm(residual(out["fp64"], b, h, (1, 0, 1), CELL)) / np.linalg.norm(b)
    e0 = np.linalg.norm(out["mixed_s0"] - out["fp64"]) / np.linalg.norm(out["fp64"])
    e2 = np.linalg.norm(out["mixed_s2"] - out["fp64"]) / np.linalg.norm(out["fp64"])
    r2 = np.linalg.norm(residual(out["mixed_s2"], b, h, (1, 0, 1), CELL)) / np.linalg.norm(b)
    print(f"[result] 256^3 cell-centred: fp64 residual {r64:.3e}; mixed s0 error {e0:.3e}, "
          f"s2 error {e2:.3e}, s2 residual {r2:.3e}", flush=True)
    assert r64 <= 1e-13
    assert abs(out["fp64"].mean()) <= 1e-13 * np.max(np.abs(out["fp64"]))
    assert e0 <= 1e-5
    assert e2 <= 1e-11


def test_cell_centred_context_has_no_fields():
    ctx = mif.Context(gridspec((9, 65, 9)), device=0, cell_centred=CELL)
    with pytest.raises(mif.MifError):
        ctx.pressure()
    ctx.close()
    with pytest.raises(mif.MifError):
        mif.Context(gridspec((9, 18, 9)), device=0, cell_centred=CELL)  # 17 cells: no kernel
    with pytest.raises(mif.MifError):
        mif.Context(gridspec((9, 65, 9)), device=0, cell_centred=(True, False, False))  # x is periodic
