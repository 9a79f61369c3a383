"""ctypes binding of oracle/_ref/libmif_ref.so (TEST INFRASTRUCTURE ONLY).

The library is the reference solver's own FFTW-free sources
(src/Constants.cpp, StaggeredTensor.cpp, VelocityTensor.cpp,
VectorFunction.cpp, Norms.cpp and the header-only stencils) compiled by
`make -C oracle ref` with the harness oracle/ref_harness.cpp.  Tests use it to
check the oracle's restatement against the reference itself.  It exists only
where the reference sources were present at build time; `available()` says
whether it was built.

Analytic functions are passed as the oracle's own evaluator (orc_fn_eval) and
a mif_function descriptor, so both sides see identical function values.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

import mif_oracle as mo

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_ref", "libmif_ref.so")

_lib = None


def available() -> bool:
    return os.path.exists(LIB_PATH)


def lib():
    global _lib
    if _lib is None:
        L = ctypes.CDLL(LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        G = ctypes.POINTER(mo.MifGrid)
        vp = ctypes.c_void_p
        L.ref_constants.argtypes = [G, ctypes.POINTER(ctypes.c_long), ctypes.POINTER(ctypes.c_long), dp]
        L.ref_set.argtypes = [G, ctypes.c_int, dp, vp, vp, ctypes.c_int, ctypes.c_double, ctypes.c_int]
        L.ref_apply_bc.argtypes = [G, dp, dp, dp, vp, vp, ctypes.c_double]
        L.ref_apply_periodic.argtypes = [G, ctypes.c_int, dp]
        L.ref_momentum_rhs.argtypes = [G, dp, dp, dp, ctypes.c_int, dp]
        L.ref_pressure_gradient.argtypes = [G, dp, ctypes.c_int, dp]
        L.ref_divergence_rhs.argtypes = [G, dp, dp, dp, ctypes.c_double, dp]
        L.ref_time_difference.argtypes = [vp, vp, ctypes.c_double, ctypes.c_double, ctypes.c_long, dp, dp]
        L.ref_nhn_stage_difference.argtypes = [vp, vp, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                               ctypes.c_long, dp, dp]
        L.ref_lid.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_long, dp, dp]
        L.ref_norms_velocity.argtypes = [G, dp, dp, dp, vp, vp, ctypes.c_double, dp]
        L.ref_norms_pressure.argtypes = [G, dp, vp, vp, ctypes.c_double, dp]
        L.ref_dump.argtypes = [G, ctypes.c_int, dp, ctypes.c_char_p]
        for name in ("ref_constants", "ref_set", "ref_apply_bc", "ref_apply_periodic", "ref_momentum_rhs",
                     "ref_pressure_gradient", "ref_divergence_rhs", "ref_time_difference",
                     "ref_nhn_stage_difference", "ref_lid", "ref_norms_velocity", "ref_norms_pressure", "ref_dump"):
            getattr(L, name).restype = ctypes.c_int
        _lib = L
    return _lib


def _dp(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _evaluator():
    """Address of the oracle's orc_fn_eval(const mif_function*, int, t, x, y, z)."""
    return ctypes.cast(mo.lib().orc_fn_eval, ctypes.c_void_p)


EVAL_FN = ctypes.CFUNCTYPE(ctypes.c_double, ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                           ctypes.c_double, ctypes.c_double)


class PyFunction:
    """A vector function given as three Python callables f(t, x, y, z),
    passed to the reference code in place of a descriptor: `fn` for the
    context argument and `evaluator` for the function pointer."""

    def __init__(self, fns):
        self.fns = fns
        self._cb = EVAL_FN(lambda _ctx, c, t, x, y, z: float(self.fns[c](t, x, y, z)))
        self.evaluator = ctypes.cast(self._cb, ctypes.c_void_p)


class Reference:
    """The reference's own code at one rank on one grid descriptor."""

    def __init__(self, grid: "mo.MifGrid"):
        self.grid = grid
        self._ext = None

    def _check(self, f: np.ndarray, s: int):
        """The reference copies whole tensors of its own extents: refuse an
        array of any other shape before handing it over."""
        if self._ext is None:
            self._ext = self.constants()[0]
        want = tuple(reversed(self._ext[s]))
        if f.shape != want:
            raise ValueError(f"array of shape {f.shape} for a staggering-{s} tensor of shape {want}")
        return f

    def constants(self):
        ext = (ctypes.c_long * 12)()
        ints = (ctypes.c_long * 9)()
        reals = (ctypes.c_double * 22)()
        lib().ref_constants(ctypes.byref(self.grid), ext, ints, reals)
        extents = [tuple(int(ext[3 * s + a]) for a in range(3)) for s in range(4)]
        return extents, [int(v) for v in ints], [float(v) for v in reals]

    def set(self, f: np.ndarray, s: int, fn, comp: int, t: float, include_border=True):
        lib().ref_set(ctypes.byref(self.grid), s, _dp(self._check(f, s)), _evaluator(), ctypes.addressof(fn),
                      comp, t, 1 if include_border else 0)

    def apply_bc(self, vel, fn, t: float):
        """fn: an oracle descriptor (mif_function) or a PyFunction."""
        if isinstance(fn, PyFunction):
            ev, ctx = fn.evaluator, None
        else:
            ev, ctx = _evaluator(), ctypes.addressof(fn)
        lib().ref_apply_bc(ctypes.byref(self.grid), *[_dp(self._check(vel[c], c)) for c in range(3)], ev, ctx, t)

    def apply_periodic(self, f: np.ndarray, s: int):
        lib().ref_apply_periodic(ctypes.byref(self.grid), s, _dp(self._check(f, s)))

    def momentum_rhs(self, vel, comp: int) -> np.ndarray:
        out = np.zeros_like(vel[comp])
        lib().ref_momentum_rhs(ctypes.byref(self.grid), *[_dp(self._check(vel[c], c)) for c in range(3)], comp,
                               _dp(out))
        return out

    def pressure_gradient(self, p: np.ndarray, comp: int, shape) -> np.ndarray:
        out = np.zeros(shape)
        lib().ref_pressure_gradient(ctypes.byref(self.grid), _dp(self._check(p, 3)), comp,
                                    _dp(self._check(out, comp)))
        return out

    def divergence_rhs(self, vel, dt: float, out: np.ndarray):
        lib().ref_divergence_rhs(ctypes.byref(self.grid), *[_dp(self._check(vel[c], c)) for c in range(3)], dt,
                                 _dp(self._check(out, 3)))

    def norms_velocity(self, vel, fn, t: float) -> np.ndarray:
        out = np.zeros(3)
        lib().ref_norms_velocity(ctypes.byref(self.grid), *[_dp(self._check(vel[c], c)) for c in range(3)],
                                 _evaluator(), ctypes.addressof(fn), t, _dp(out))
        return out

    def norms_pressure(self, p: np.ndarray, fn, t: float) -> np.ndarray:
        out = np.zeros(3)
        lib().ref_norms_pressure(ctypes.byref(self.grid), _dp(self._check(p, 3)), _evaluator(), ctypes.addressof(fn),
                                 t, _dp(out))
        return out

    def dump(self, f: np.ndarray, s: int, path: str):
        rc = lib().ref_dump(ctypes.byref(self.grid), s, _dp(self._check(f, s)), path.encode())
        if rc != 0:
            raise OSError(f"reference dump to {path} failed")


def time_difference(fn, time_1: float, time_2: float, xyz: np.ndarray) -> np.ndarray:
    """TimeVectorFunction::get_difference_over_time(time_1, time_2) at points xyz (n, 3)."""
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    out = np.zeros_like(xyz)
    lib().ref_time_difference(_evaluator(), ctypes.addressof(fn), time_1, time_2, len(xyz), _dp(xyz), _dp(out))
    return out


def nhn_stage_difference(fn, dt: float, t_n: float, stage: int, xyz: np.ndarray) -> np.ndarray:
    """The function timestep_nhn's stage `stage` hands to its pressure solve, at points xyz (n, 3)."""
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    out = np.zeros_like(xyz)
    lib().ref_nhn_stage_difference(_evaluator(), ctypes.addressof(fn), dt, t_n, stage, len(xyz), _dp(xyz),
                                   _dp(out))
    return out


def lid(tcase: int, comp: int, txyz: np.ndarray) -> np.ndarray:
    """exact_{u,v,w}_t{1,2} of include/TestCaseBoundaries.h at points txyz (n, 4)."""
    txyz = np.ascontiguousarray(txyz, dtype=np.float64)
    out = np.zeros(len(txyz))
    lib().ref_lid(tcase, comp, len(txyz), _dp(txyz), _dp(out))
    return out
