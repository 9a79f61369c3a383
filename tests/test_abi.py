"""The C-ABI library loads and exports every entry point declared in
include/*.h; without a device the product path fails loudly (no CPU
fallback)."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    syms = set()
    inc = os.path.join(ROOT, "include")
    for name in os.listdir(inc):
        if not name.endswith(".h"):
            continue
        text = open(os.path.join(inc, name)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?\w+\s*\*?\s*(mif_\w+)\s*\(", text, flags=re.M):
            syms.add(m.group(1))
    return sorted(syms)


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert "mif_timestep" in syms and "mif_ctx_create" in syms and len(syms) >= 25


def test_library_exports_every_declared_symbol():
    import mif_amd

    lib_path = mif_amd.LIB_PATH
    assert os.path.exists(lib_path), "build the HIP library first"
    out = subprocess.run(["nm", "-D", "--defined-only", lib_path], capture_output=True, text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    missing = [s for s in declared_symbols() if s not in exported]
    assert not missing, missing
    lib = mif_amd.load()
    for s in declared_symbols():
        assert hasattr(lib, s)


def test_binding_covers_every_declared_symbol():
    from mif_amd import _lib

    assert set(declared_symbols()) == set(_lib.SIGNATURES)


def test_no_device_fails_loudly():
    import mif_amd

    gs = mif_amd.GridSpec(8, 8, 8, 1.0, 1.0, 1.0, 0.0, 0.0, 0.0, 1.0, 1.0, 1)
    try:
        ctx = mif_amd.Context(gs)
    except mif_amd.MifError as e:
        assert e.status == 6  # MIF_ERR_NO_DEVICE
        assert "device" in str(e)
        return
    ctx.close()
    pytest.skip("a HIP device is present")


def test_invalid_grid_rejected_before_device_use():
    import mif_amd

    with pytest.raises(mif_amd.MifError) as ei:
        mif_amd.Context(mif_amd.GridSpec(1, 8, 8, 1.0, 1.0, 1.0, 0.0, 0.0, 0.0, 1.0, 1.0, 1))
    assert ei.value.status == 1
    with pytest.raises(mif_amd.MifError) as ei:
        mif_amd.Context(mif_amd.GridSpec(8, 8, 8, 1.0, 1.0, 1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 1))
    assert ei.value.status == 1
    # a wall axis of 2050 nodes: complex length 4098 has no specialised
    # transform, and the generic kernel would need 128 KB of LDS per workgroup
    with pytest.raises(mif_amd.MifError) as ei:
        mif_amd.Context(mif_amd.GridSpec(2050, 8, 8, 1.0, 1.0, 1.0, 0.0, 0.0, 0.0, 1.0, 1.0, 1))
    assert ei.value.status == 5  # MIF_ERR_UNSUPPORTED
    assert b"LDS" in mif_amd.load().mif_last_error(None)
    # 2049 periodic nodes (length 2048, 64 KB) pass the check and reach the device query
    with pytest.raises(mif_amd.MifError) as ei:
        mif_amd.Context(mif_amd.GridSpec(2049, 8, 8, 1.0, 1.0, 1.0, 0.0, 0.0, 0.0, 1.0, 1.0, 1, (True, False, False)))
    assert ei.value.status == 6  # MIF_ERR_NO_DEVICE on a host without a GPU
