"""Boundary data from callbacks the device cannot evaluate (host face
tables, the *_tabulated entry points) against the descriptor path on the
GPU, through the C++ interface (lambdas the shim cannot recognise) and
through the Python binding (numpy callables).  Host and device evaluate
sin/exp with different libraries, so the two paths agree to rounding, not
bit for bit."""
import os
import subprocess

import numpy as np
import pytest

import mif_amd as mif
from mif_amd.problems import channel_grid

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BUILD = os.path.join(ROOT, "mpi-incompressible-fluid_amd", "_build")
VEL_TOL = 1e-14
P_TOL = 1e-12


@pytest.mark.parametrize("mode,n,steps", [("walls", 16, 3), ("channel", 17, 3), ("nhn", 16, 3),
                                          ("velocity", 16, 3), ("apply_bc", 16, 1)])
def test_cxx_lambda_matches_descriptor(mode, n, steps):
    print(f"[case] callback_test {mode} {n} {steps}", flush=True)
    r = subprocess.run([os.path.join(BUILD, "callback_test"), mode, str(n), str(steps)], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    dv, dp = (float(x) for x in r.stdout.split())
    print(f"[result] {mode}: velocity {dv:.3e} pressure {dp:.3e}", flush=True)
    assert dv <= VEL_TOL
    assert dp <= P_TOL


def es_callables(Re):
    a, d = np.pi / 4, np.pi / 2

    def comp(c):
        def f(t, x, y, z):
            X = [x, y, z]
            p, q, s = X[c], X[(c + 1) % 3], X[(c + 2) % 3]
            return -a * (np.exp(a * p) * np.sin(a * q + d * s) + np.exp(a * s) * np.cos(a * p + d * q)) * \
                np.exp(-d * d * t / Re)
        return f

    return [comp(c) for c in range(3)]


@pytest.mark.parametrize("precision", [mif.FP64, mif.MIXED], ids=["fp64", "mixed"])
def test_python_callables_match_descriptor(precision):
    """timestep with the Ethier-Steinman velocity as numpy callables against
    the same run with its descriptor, on an all-wall box and on the channel
    (periodic x, z)."""
    Re = 1e3
    for gs in (mif.GridSpec(16, 16, 16, 1.0, 1.0, 2.0, 0.0, 0.0, -1.0, Re, 3e-4, 3, (False, False, False)),
               channel_grid(16, 16, 16, Re=Re, dt=1e-4, steps=3)):
        print(f"[case] python callables precision={precision} periodic={gs.periodic}", flush=True)
        out = []
        for ev in (mif.function(mif.FN_ES_VELOCITY, Re=Re), es_callables(Re)):
            ctx = mif.Context(gs, precision=precision, device=0)
            vel, buf, buf2 = ctx.velocity(), ctx.velocity(), ctx.velocity()
            p, pb = ctx.pressure(), ctx.pressure()
            vel.set(mif.function(mif.FN_ES_VELOCITY, Re=Re), 0.0)
            for s in range(3):
                mif.timestep(vel, buf, buf2, ev, s * ctx.dt, p, pb)
            out.append((vel.download(), p.download()))
            ctx.close()
        (va, pa), (vb, pb_) = out
        tol = VEL_TOL if precision == mif.FP64 else 1e-6
        for c in range(3):
            assert np.max(np.abs(vb[c] - va[c])) <= tol * np.max(np.abs(va[c])), c
        ptol = P_TOL if precision == mif.FP64 else 1e-4
        assert np.max(np.abs(pb_ - pa)) <= ptol * np.max(np.abs(pa))
