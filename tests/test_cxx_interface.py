"""The C++ host interface (the reference's class and function signatures
over the C ABI) builds, links against the HIP library, and fails loudly
without a device."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BUILD = os.path.join(ROOT, "mpi-incompressible-fluid_amd", "_build")
DRIVERS = ["full_test", "velocity_test", "pressure_test"]


@pytest.mark.parametrize("name", DRIVERS)
def test_driver_built_and_linked(name):
    exe = os.path.join(BUILD, name)
    assert os.access(exe, os.X_OK), f"{exe} missing: run __graft_entry__.build()"
    out = subprocess.run(["ldd", exe], capture_output=True, text=True).stdout
    assert "libmif_gpu.so" in out


def test_driver_usage_message():
    r = subprocess.run([os.path.join(BUILD, "full_test")], capture_output=True, text=True)
    assert r.returncode == 2 and "usage" in r.stderr


def test_driver_without_device_fails_loudly():
    env = dict(os.environ, HIP_VISIBLE_DEVICES="-1")
    try:
        import torch

        if torch.cuda.device_count() > 0:
            pytest.skip("a GPU is visible")
    except ImportError:
        pass
    r = subprocess.run([os.path.join(BUILD, "velocity_test"), "8", "1"], capture_output=True, text=True, env=env,
                       timeout=120)
    assert r.returncode == 1
    assert "no HIP device" in r.stderr or "hip" in r.stderr.lower()
