"""Ownership rules of the C ABI: a field that outlives its context holds no
device memory, can still be destroyed, and is rejected by every other call."""
import numpy as np
import pytest

from helpers import PI, grids

pytestmark = pytest.mark.gpu

mif = pytest.importorskip("mif_amd")


def test_field_outliving_its_context():
    print("[case] field outliving its context", flush=True)
    _, gs = grids((9, 9, 9), (2 * PI, 2.0, PI), (0.0, -1.0, 0.0), 100.0, 0.01, 1, (1, 0, 1))
    ctx = mif.Context(gs)
    vel = ctx.velocity()
    p = ctx.pressure()
    p.upload(np.ones(p.shape))
    ctx.close()
    assert p.ptr is None and all(c.ptr is None for c in vel.components)
    # the raw C path: a field left alive when its context is destroyed
    ctx2 = mif.Context(gs)
    lib = ctx2.lib
    import ctypes
    raw = ctypes.c_void_p()
    assert lib.mif_field_create(ctx2.ptr, 3, ctypes.byref(raw)) == 0
    lib.mif_ctx_destroy(ctx2.ptr)
    ctx2.ptr = None
    buf = np.zeros((10, 9, 10))
    rc = lib.mif_field_upload(raw, buf.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    assert rc == 1  # MIF_ERR_INVALID_ARGUMENT
    assert lib.mif_field_destroy(raw) == 0


def test_stream_probe():
    # the streaming-ceiling probe: a copy of 64 MiB moves 128 MiB per launch
    # at a rate below the device's 8 TB/s peak; arguments out of range are
    # rejected
    print("[case] stream probe", flush=True)
    _, gs = grids((9, 9, 9), (2 * PI, 2.0, PI), (0.0, -1.0, 0.0), 100.0, 0.01, 1, (1, 0, 1))
    ctx = mif.Context(gs)
    nbytes = 64 << 20
    us = ctx.stream_probe(1, 1, nbytes, reps=5)
    tbs = 2 * nbytes / (us * 1e-6) / 1e12
    print("copy", us, "us", tbs, "TB/s", flush=True)
    assert 0.5 < tbs < 8.5
    for bad in ((9, 1, nbytes, 5), (0, 0, nbytes, 5), (1, 1, 8, 5), (1, 1, nbytes, 0)):
        with pytest.raises(mif.MifError):
            ctx.stream_probe(*bad)
    ctx.close()
