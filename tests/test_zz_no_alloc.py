"""mif_field_gather_noalloc (the allocation-free variant of mif_field_gather)
against mif_field_gather, called through the C ABI: two calls in a row on one
field, the second with fewer points, in both precisions; and the size limit."""
import ctypes

import numpy as np
import pytest

import mif_amd as mif
from mif_amd.problems import channel_grid

MAX_POINTS = 1 << 20


def _gather(lib, name, field, ijk):
    idx = np.ascontiguousarray(np.asarray(ijk, dtype=np.int32).reshape(-1, 3))
    out = np.full(idx.shape[0], np.nan)
    rc = getattr(lib, name)(field.ptr, idx.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), idx.shape[0],
                            out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return rc, out


def _points(rng, shape, n):
    sz, sy, sx = shape
    return np.stack([rng.integers(0, sx, n), rng.integers(0, sy, n), rng.integers(0, sz, n)], axis=1)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [mif.FP64, mif.MIXED], ids=["fp64", "mixed"])
def test_gather_noalloc_matches_gather(precision):
    print(f"gather_noalloc: precision {precision}", flush=True)
    ctx = mif.Context(channel_grid(48, 32, 40), precision=precision, device=0)
    lib = ctx.lib
    rng = np.random.default_rng(77)
    fields = [ctx.pressure(), ctx.velocity().w]
    for f in fields:
        f.upload(rng.standard_normal(f.shape))
    for f in fields:
        for n in (200_000, 3_000):  # second call smaller than the first
            ijk = _points(rng, f.shape, n)
            rc_a, ref = _gather(lib, "mif_field_gather", f, ijk)
            rc_b, got = _gather(lib, "mif_field_gather_noalloc", f, ijk)
            assert rc_a == 0 and rc_b == 0, mif.load().mif_last_error(ctx.ptr)
            assert np.array_equal(got, ref), (n, f.staggering)
            # against the downloaded field too
            host = f.download()
            assert np.array_equal(got, host[ijk[:, 2], ijk[:, 1], ijk[:, 0]].astype(np.float64))
    # the largest allowed request, then one point too many (rejected before any device work)
    big = _points(rng, fields[0].shape, MAX_POINTS)
    rc, got = _gather(lib, "mif_field_gather_noalloc", fields[0], big)
    assert rc == 0
    rc_a, ref = _gather(lib, "mif_field_gather", fields[0], big)
    assert rc_a == 0 and np.array_equal(got, ref)
    over = np.zeros((MAX_POINTS + 1, 3), dtype=np.int32)
    rc, _ = _gather(lib, "mif_field_gather_noalloc", fields[0], over)
    assert rc == 1  # MIF_ERR_INVALID_ARGUMENT
    assert b"at most" in mif.load().mif_last_error(ctx.ptr)
    for f in fields:
        f.close()
    ctx.close()
