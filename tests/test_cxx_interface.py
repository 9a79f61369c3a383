"""The C++ host interface (the reference's class and function signatures
over the C ABI) builds, links against the HIP library, and fails loudly
without a device."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BUILD = os.path.join(ROOT, "mpi-incompressible-fluid_amd", "_build")
DRIVERS = ["full_test", "velocity_test", "pressure_test", "io_test", "mif"]


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


# ---- the `mif` driver's input file (src/InputParser.cpp:8-77): every error
# is reported before the device is touched, so these run without a GPU

VALID = {"Nt": "1000", "dt": "1e-3", "Nx": "16", "Ny": "16", "Nz": "16", "Py": "1", "Pz": "1", "test_case_2": "false"}


def _mif(tmp_path, lines):
    f = tmp_path / "input.txt"
    f.write_text("\n".join(lines))
    env = dict(os.environ, HIP_VISIBLE_DEVICES="-1")
    return subprocess.run([os.path.join(BUILD, "mif"), str(f)], capture_output=True, text=True, env=env, timeout=120)


def _lines(d):
    return [f"{k} : {v}" for k, v in d.items()]


def test_mif_usage():
    r = subprocess.run([os.path.join(BUILD, "mif")], capture_output=True, text=True)
    assert r.returncode == 1 and "Usage: ./mif [input parameter file]" in r.stderr


def test_mif_missing_input_file(tmp_path):
    r = subprocess.run([os.path.join(BUILD, "mif"), str(tmp_path / "nope.txt")], capture_output=True, text=True)
    assert r.returncode == 1
    assert "Error parsing input file: Error opening input file: " in r.stderr


@pytest.mark.parametrize("edit,message", [
    (lambda d: d.pop("Nz"), "Missing one or more inputs. Required inputs: Nx, Ny, Nz, dt, Nt, Py, Pz, test_case_2."),
    (lambda d: d.update(Nw="3"), "Unknown or duplicate key: Nw"),
    (lambda d: d.update(Nx="sixteen"), "Error parsing input file: stoul"),
])
def test_mif_input_errors(tmp_path, edit, message):
    d = dict(VALID)
    edit(d)
    r = _mif(tmp_path, _lines(d))
    assert r.returncode == 1 and message in r.stderr, r.stderr


def test_mif_duplicate_key(tmp_path):
    r = _mif(tmp_path, _lines(VALID) + ["Nx : 8"])
    assert r.returncode == 1 and "Unknown or duplicate key: Nx" in r.stderr


def test_mif_processor_counts_checked(tmp_path):
    r = _mif(tmp_path, _lines(dict(VALID, Py="0")))
    assert r.returncode == 1
    assert "The number of processors in each direction must be at least 1." in r.stderr


def test_mif_parses_reference_style_input(tmp_path):
    """The keys in any order, blanks around ':', lines without ':' skipped,
    Py/Pz of an MPI run accepted (one GPU holds the domain): the driver gets
    past parsing and stops only for want of a device."""
    try:
        import torch

        if torch.cuda.device_count() > 0:
            pytest.skip("a GPU is visible")
    except ImportError:
        pass
    lines = ["Nt : 20000", "dt : 1e-3", "Nx : 64", "# a comment line", "Ny:64", "Nz   :   64", "Py : 3", "Pz : 4",
             "test_case_2 : true"]
    r = _mif(tmp_path, lines)
    assert r.returncode == 1
    assert "Error parsing input file" not in r.stderr
    assert "Py = 3, Pz = 4 ignored" in r.stderr
