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
