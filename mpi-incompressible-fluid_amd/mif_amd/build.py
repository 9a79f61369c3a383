"""Build recipe of the HIP library (gfx950) and of its C++ host interface.

Compiles every HIP source of csrc/ to an object in parallel and links
_build/libmif_gpu.so in-tree, so the built library travels with the source
tree.  Then compiles the C++ host interface of cxx/ (the reference's class
and function signatures over the C ABI) into _build/libmif_cxx.a and links
against both the `mif` program (the reference's executable) and the C++
driver programs of tests/cpp/.  Invoked by
__graft_entry__.build() and usable as `python -m mif_amd.build` from the
package directory.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(PKG_DIR, "csrc")
OUT = os.path.join(PKG_DIR, "_build")
LIB = os.path.join(OUT, "libmif_gpu.so")
SOURCES = ["mif_api.hip", "mif_stencil.hip", "mif_fft.hip", "mif_fft_pow2.hip", "mif_mixed.hip", "mif_adjoint.hip",
           "mif_probe.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")
ROOT = os.path.dirname(PKG_DIR)
CXX_DIR = os.path.join(PKG_DIR, "cxx")
CXX_SOURCES = ["shim.cpp", "functions.cpp", "InputParser.cpp", "VTKDatExport.cpp"]
CXX_LIB = os.path.join(OUT, "libmif_cxx.a")
DRIVER_DIR = os.path.join(ROOT, "tests", "cpp")
DRIVERS = ["full_test", "velocity_test", "pressure_test", "io_test", "callback_test"]
# programs of the package itself (the reference's `mif` executable)
PROGRAMS = {"mif": "mif_main.cpp"}
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-Wextra",
             "-I", os.path.join(CXX_DIR, "include"), "-I", os.path.join(ROOT, "include")]
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-Wall",
    # no packed-FP32 instructions in generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _compile(src: str) -> str:
    obj = os.path.join(OUT, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj + ".tmp"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{r.stderr}")
    os.replace(obj + ".tmp", obj)
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(OUT, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", LIB + ".tmp", *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(LIB + ".tmp", LIB)
    build_cxx()
    if verbose:
        print("built", LIB)
    return LIB


def _run(cmd, what):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"{what} failed:\n{r.stderr}")


def build_cxx() -> str:
    """libmif_cxx.a (the host interface) and the driver programs."""
    objs = []
    for src in CXX_SOURCES:
        obj = os.path.join(OUT, "cxx_" + os.path.splitext(src)[0] + ".o")
        _run([CXX, *CXX_FLAGS, "-c", os.path.join(CXX_DIR, "src", src), "-o", obj], f"{CXX} {src}")
        objs.append(obj)
    if os.path.exists(CXX_LIB):
        os.remove(CXX_LIB)
    _run(["ar", "rcs", CXX_LIB, *objs], "ar")
    sources = {name: os.path.join(DRIVER_DIR, name + ".cpp") for name in DRIVERS}
    sources.update({name: os.path.join(CXX_DIR, "src", src) for name, src in PROGRAMS.items()})
    for name, src in sources.items():
        exe = os.path.join(OUT, name)
        _run([CXX, *CXX_FLAGS, src, "-o", exe + ".tmp", CXX_LIB, "-L", OUT, "-lmif_gpu", "-Wl,-rpath,$ORIGIN"],
             f"{CXX} {name}")
        os.replace(exe + ".tmp", exe)
    return CXX_LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
