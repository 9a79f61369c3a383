"""ctypes binding of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, and only as the checker or the timed CPU baseline; the product
(the HIP library) never calls it.  Fields are numpy float64 arrays with the
reference's shape convention: a field of extents (sx, sy, sz) is stored flat
with x fastest, so the numpy array has shape (sz, sy, sx).
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libmif_oracle.so")

# mif_fn_kind (include/mif_types.h)
FN_ZERO = 0
FN_ES_VELOCITY = 1
FN_ES_PRESSURE = 2
FN_ES_GRAD_P = 3
FN_VT_VELOCITY = 4
FN_VT_FORCING = 5
FN_PT_VELOCITY = 6
FN_PT_PRESSURE = 7
FN_PT_GRAD_P = 8
FN_LID_1 = 9
FN_LID_2 = 10
FN_CONSTANT = 11
FN_TIME_DIFFERENCE = 12


class MifGrid(ctypes.Structure):
    _fields_ = [
        ("Nx_global", ctypes.c_uint64), ("Ny_global", ctypes.c_uint64), ("Nz_global", ctypes.c_uint64),
        ("x_size", ctypes.c_double), ("y_size", ctypes.c_double), ("z_size", ctypes.c_double),
        ("min_x", ctypes.c_double), ("min_y", ctypes.c_double), ("min_z", ctypes.c_double),
        ("Re", ctypes.c_double), ("final_time", ctypes.c_double), ("num_time_steps", ctypes.c_uint32),
        ("periodic", ctypes.c_int32 * 3),
    ]


class MifFunction(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int32), ("base", ctypes.c_int32), ("Re", ctypes.c_double),
        ("c", ctypes.c_double * 3), ("t0", ctypes.c_double), ("t1", ctypes.c_double),
    ]


class OrcGrid(ctypes.Structure):
    _fields_ = [
        ("Ng", ctypes.c_long * 3), ("periodic", ctypes.c_int * 3), ("L", ctypes.c_double * 3),
        ("min", ctypes.c_double * 3), ("Re", ctypes.c_double), ("dt", ctypes.c_double),
        ("N", ctypes.c_long * 3), ("base", ctypes.c_long * 3), ("h", ctypes.c_double * 3),
        ("one_over_2h", ctypes.c_double * 3), ("one_over_8h", ctypes.c_double * 3),
        ("one_over_h2Re", ctypes.c_double * 3), ("h_over_2", ctypes.c_double * 3),
        ("one_over_h", ctypes.c_double * 3), ("npts", ctypes.c_long * 3), ("force", ctypes.c_double * 3),
    ]


def make_grid(Ng, size, origin, Re, final_time, nsteps, periodic) -> MifGrid:
    g = MifGrid()
    g.Nx_global, g.Ny_global, g.Nz_global = (int(n) for n in Ng)
    g.x_size, g.y_size, g.z_size = (float(s) for s in size)
    g.min_x, g.min_y, g.min_z = (float(o) for o in origin)
    g.Re = float(Re)
    g.final_time = float(final_time)
    g.num_time_steps = int(nsteps)
    for a in range(3):
        g.periodic[a] = 1 if periodic[a] else 0
    return g


def make_fn(kind, Re=1.0, c=(0.0, 0.0, 0.0), base=0, t0=0.0, t1=0.0) -> MifFunction:
    f = MifFunction()
    f.kind = kind
    f.base = base
    f.Re = float(Re)
    for a in range(3):
        f.c[a] = float(c[a])
    f.t0 = float(t0)
    f.t1 = float(t1)
    return f


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"oracle library missing: {_LIB_PATH} (run `make -C oracle`)")
        L = ctypes.CDLL(_LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        dpp = ctypes.POINTER(dp)
        G = ctypes.POINTER(OrcGrid)
        F = ctypes.POINTER(MifFunction)
        L.orc_grid_init.argtypes = [G, ctypes.POINTER(MifGrid)]
        L.orc_grid_init.restype = ctypes.c_int
        L.orc_fn_eval.argtypes = [F, ctypes.c_int] + [ctypes.c_double] * 4
        L.orc_fn_eval.restype = ctypes.c_double
        L.orc_set.argtypes = [G, dp, ctypes.c_int, F, ctypes.c_int, ctypes.c_double, ctypes.c_int]
        L.orc_apply_bc.argtypes = [G, dp, dp, dp, F, ctypes.c_double]
        L.orc_apply_periodic.argtypes = [G, dp, ctypes.c_int]
        L.orc_timestep.argtypes = [G, dpp, dpp, dpp, F, F, ctypes.c_double, dp, dp]
        L.orc_timestep_velocity.argtypes = [G, dpp, dpp, dpp, F, F, ctypes.c_double]
        L.orc_solve_pressure.argtypes = [G, dp, dpp, F, ctypes.c_double]
        L.orc_poisson_packed.argtypes = [G, dp]
        L.orc_adjust_pressure.argtypes = [G, dp, F, ctypes.c_double]
        L.orc_error_norms_velocity.argtypes = [G, dpp, F, ctypes.c_double, dp]
        L.orc_error_norms_pressure.argtypes = [G, dp, F, ctypes.c_double, dp]
        L.orc_pressure_gradient.argtypes = [G, dp, dpp]
        L.orc_momentum_rhs.argtypes = [G, dpp, ctypes.c_int, ctypes.c_long, ctypes.c_long, ctypes.c_long]
        L.orc_momentum_rhs.restype = ctypes.c_double
        for name in ("orc_r2hc", "orc_hc2r", "orc_redft00", "orc_r2hc_direct", "orc_hc2r_direct",
                     "orc_redft00_direct", "orc_dct2_direct", "orc_dct3_direct"):
            getattr(L, name).argtypes = [ctypes.c_long, dp, dp]
        L.orc_timestep_adjoint.argtypes = [G, dpp, dpp, dpp, dpp, dp, dp]
        L.orc_extents.argtypes = [G, ctypes.c_int, ctypes.POINTER(ctypes.c_long)]
        L.orc_nhn_stage_function.argtypes = [G, F, ctypes.c_double, ctypes.c_int, F]
        L.orc_pressure_rhs.argtypes = [G, dp, dpp, F, ctypes.c_double]
        L.orc_poisson_packed_cell.argtypes = [G, ctypes.POINTER(ctypes.c_int), dp]
        L.orc_poisson_packed_direct.argtypes = [G, dp]
        L.orc_set_direct_transforms.argtypes = [ctypes.c_int]
        L.orc_num_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def _dp(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _dpp(arrs):
    ptrs = (ctypes.POINTER(ctypes.c_double) * 3)(*[_dp(a) for a in arrs])
    return ptrs


@dataclass
class Fields:
    """One velocity triple (u, v, w) as numpy arrays."""
    u: np.ndarray
    v: np.ndarray
    w: np.ndarray

    def __iter__(self):
        return iter((self.u, self.v, self.w))

    def __getitem__(self, i):
        return (self.u, self.v, self.w)[i]

    def copy(self):
        return Fields(self.u.copy(), self.v.copy(), self.w.copy())


class Oracle:
    """Single-rank CPU restatement of the reference timestep."""

    def __init__(self, grid: MifGrid):
        self.desc = grid
        self.g = OrcGrid()
        rc = lib().orc_grid_init(ctypes.byref(self.g), ctypes.byref(grid))
        if rc != 0:
            raise ValueError(f"invalid grid (status {rc})")

    # shapes ---------------------------------------------------------------
    def extents(self, s):
        """(sx, sy, sz) of a field of staggering s (0..2 = u, v, w; 3 = pressure),
        from the oracle's orc_extents."""
        e = (ctypes.c_long * 3)()
        lib().orc_extents(ctypes.byref(self.g), s, e)
        return tuple(int(v) for v in e)

    def shape(self, s):
        e = self.extents(s)
        return (e[2], e[1], e[0])

    def zeros(self, s):
        return np.zeros(self.shape(s), dtype=np.float64)

    def velocity(self):
        return Fields(self.zeros(0), self.zeros(1), self.zeros(2))

    @property
    def dt(self):
        return self.g.dt

    @property
    def npts(self):
        return tuple(self.g.npts)

    # functions ------------------------------------------------------------
    def fn_eval(self, fn, comp, t, x, y, z):
        return lib().orc_fn_eval(ctypes.byref(fn), comp, t, x, y, z)

    def set(self, f, s, fn, comp, t, include_border=True):
        lib().orc_set(ctypes.byref(self.g), _dp(f), s, ctypes.byref(fn), comp, t, 1 if include_border else 0)

    def set_velocity(self, vel: Fields, fn, t, include_border=True):
        for c in range(3):
            self.set(vel[c], c, fn, c, t, include_border)

    def apply_bc(self, vel: Fields, fn, t):
        lib().orc_apply_bc(ctypes.byref(self.g), _dp(vel.u), _dp(vel.v), _dp(vel.w), ctypes.byref(fn), t)

    def apply_periodic(self, f, s):
        lib().orc_apply_periodic(ctypes.byref(self.g), _dp(f), s)

    # solver ---------------------------------------------------------------
    def timestep(self, vel, buf, buf2, exact_velocity, t_n, p, p_buf, grad_p=None):
        lib().orc_timestep(ctypes.byref(self.g), _dpp(vel), _dpp(buf), _dpp(buf2), ctypes.byref(exact_velocity),
                           ctypes.byref(grad_p) if grad_p is not None else None, t_n, _dp(p), _dp(p_buf))

    def set_body_force(self, f):
        """Constant body force added to the momentum right-hand side of
        timestep() (zero reproduces the reference)."""
        for a in range(3):
            self.g.force[a] = float(f[a])

    def timestep_adjoint(self, un, y2, y3, ubar, pbar, fbar):
        """Adjoint of one timestep(): (ubar, pbar) hold the adjoints of
        (u^{n+1}, p^{n+1}) on entry and of (u^n, p^n) on return; fbar (3) is
        accumulated."""
        assert fbar.dtype == np.float64 and fbar.shape == (3,)
        lib().orc_timestep_adjoint(ctypes.byref(self.g), _dpp(un), _dpp(y2), _dpp(y3), _dpp(ubar), _dp(pbar),
                                   _dp(fbar))

    def timestep_velocity(self, vel, buf, rhs, exact_velocity, forcing, t_n):
        lib().orc_timestep_velocity(ctypes.byref(self.g), _dpp(vel), _dpp(buf), _dpp(rhs),
                                    ctypes.byref(exact_velocity), ctypes.byref(forcing), t_n)

    def solve_pressure(self, p, vel, dt, grad_p=None):
        lib().orc_solve_pressure(ctypes.byref(self.g), _dp(p), _dpp(vel),
                                 ctypes.byref(grad_p) if grad_p is not None else None, dt)

    def poisson_packed(self, b: np.ndarray):
        """In-place spectral solve on a packed (nz, ny, nx) owner array."""
        assert b.shape == tuple(reversed(self.npts))
        lib().orc_poisson_packed(ctypes.byref(self.g), _dp(b))

    def poisson_packed_direct(self, b: np.ndarray):
        """The same solve with O(N^2) direct-sum transforms: another FP64
        evaluation order of the same mathematics."""
        assert b.shape == tuple(reversed(self.npts))
        lib().orc_poisson_packed_direct(ctypes.byref(self.g), _dp(b))

    def poisson_packed_cell(self, b: np.ndarray, cell):
        """In-place spectral solve with cell-centred Neumann axes (cell[a] != 0):
        b is (nz, ny, nx) with N_g - 1 points on those axes."""
        c = (ctypes.c_int * 3)(*[1 if x else 0 for x in cell])
        n = [self.g.Ng[a] - 1 if cell[a] else self.g.npts[a] for a in range(3)]
        assert b.shape == (n[2], n[1], n[0])
        lib().orc_poisson_packed_cell(ctypes.byref(self.g), c, _dp(b))

    def adjust_pressure(self, p, fn, t):
        lib().orc_adjust_pressure(ctypes.byref(self.g), _dp(p), ctypes.byref(fn), t)

    def error_norms_velocity(self, vel, fn, t):
        out = np.zeros(3)
        lib().orc_error_norms_velocity(ctypes.byref(self.g), _dpp(vel), ctypes.byref(fn), t, _dp(out))
        return out

    def error_norms_pressure(self, p, fn, t):
        out = np.zeros(3)
        lib().orc_error_norms_pressure(ctypes.byref(self.g), _dp(p), ctypes.byref(fn), t, _dp(out))
        return out

    def pressure_gradient(self, p):
        out = self.velocity()
        lib().orc_pressure_gradient(ctypes.byref(self.g), _dp(p), _dpp(out))
        return out

    def momentum_rhs(self, vel, comp, i, j, k):
        return lib().orc_momentum_rhs(ctypes.byref(self.g), _dpp(vel), comp, i, j, k)

    def nhn_stage_function(self, grad_p, t_n, stage) -> MifFunction:
        """The boundary function timestep_nhn passes to stage 1..3's solve."""
        out = MifFunction()
        lib().orc_nhn_stage_function(ctypes.byref(self.g), ctypes.byref(grad_p), t_n, stage, ctypes.byref(out))
        return out

    def pressure_rhs(self, p, vel, dt, grad_p=None):
        """div(vel)/dt (plus Neumann data) on the owner points of p, in place."""
        lib().orc_pressure_rhs(ctypes.byref(self.g), _dp(p), _dpp(vel),
                               ctypes.byref(grad_p) if grad_p is not None else None, dt)


def transform(name: str, x: np.ndarray) -> np.ndarray:
    """Run one 1-D oracle transform (r2hc, hc2r, redft00 and *_direct)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.zeros_like(x)
    getattr(lib(), "orc_" + name)(len(x), _dp(x), _dp(out))
    return out


def set_direct_transforms(on: bool):
    """Run every later solve with O(N^2) direct-sum transforms (True) or the
    FFTs (False, the default)."""
    lib().orc_set_direct_transforms(1 if on else 0)


def num_threads() -> int:
    return lib().orc_num_threads()
