"""Python host side over the C ABI (include/mif_gpu.h).

Mirrors the reference's interfaces for the timestep path with the same names
and argument meaning: Constants (src/Constants.cpp) + PressureSolverStructures
(src/PressureSolverStructures.cpp) become a Context; StaggeredTensor /
VelocityTensor become Field / VelocityField; timestep, timestep_nhn,
timestep_velocity, solve_pressure_equation_*, adjust_pressure and the error
norms keep their reference names.  Every call goes to the HIP library; a
failure raises MifError with the library's message.
"""
from __future__ import annotations

import ctypes
import sys
import weakref
from dataclasses import dataclass
from typing import Sequence

import numpy as np

from . import _lib as L
from . import boundary as _bd
from .restart import format_text_dump


class MifError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"{L.STATUS_NAMES.get(status, status)}: {msg}")
        self.status = status


def _check(rc: int, ctx_ptr=None):
    if rc != L.OK:
        msg = L.load().mif_last_error(ctx_ptr)
        raise MifError(rc, msg.decode() if msg else "")


def function(kind: int, Re: float = 1.0, c=(0.0, 0.0, 0.0), base: int = 0, t0: float = 0.0,
             t1: float = 0.0) -> L.MifFunction:
    """A known analytic function (the device-side stand-in for the reference's
    std::function callbacks, include/VectorFunction.h:15-53)."""
    f = L.MifFunction()
    f.kind = kind
    f.base = base
    f.Re = float(Re)
    for a in range(3):
        f.c[a] = float(c[a])
    f.t0 = float(t0)
    f.t1 = float(t1)
    return f


@dataclass(frozen=True)
class GridSpec:
    """Arguments of the reference's Constants constructor at one rank
    (include/Constants.h:86-90)."""
    Nx_global: int
    Ny_global: int
    Nz_global: int
    x_size: float
    y_size: float
    z_size: float
    min_x: float
    min_y: float
    min_z: float
    Re: float
    final_time: float
    num_time_steps: int
    periodic: tuple = (False, False, False)

    def to_c(self) -> L.MifGrid:
        g = L.MifGrid()
        g.Nx_global, g.Ny_global, g.Nz_global = self.Nx_global, self.Ny_global, self.Nz_global
        g.x_size, g.y_size, g.z_size = self.x_size, self.y_size, self.z_size
        g.min_x, g.min_y, g.min_z = self.min_x, self.min_y, self.min_z
        g.Re = self.Re
        g.final_time = self.final_time
        g.num_time_steps = self.num_time_steps
        for a in range(3):
            g.periodic[a] = 1 if self.periodic[a] else 0
        return g


class Context:
    """One domain on one GPU (Constants + PressureSolverStructures)."""

    def __init__(self, grid: GridSpec, precision: int = L.FP64, device: int = -1, cell_centred=None):
        """cell_centred: per-axis flags of cell-centred Neumann axes, for a
        context of stand-alone Poisson solves (mif_ctx_create_cell_centred):
        N_g - 1 unknowns at the cell centres, DCT-II/DCT-III transforms."""
        self.lib = L.load()
        self.grid = grid
        self.precision = precision
        self.cell_centred = tuple(bool(c) for c in cell_centred) if cell_centred is not None else (False,) * 3
        ptr = ctypes.c_void_p()
        if any(self.cell_centred):
            flags = (ctypes.c_int * 3)(*[1 if c else 0 for c in self.cell_centred])
            _check(self.lib.mif_ctx_create_cell_centred(ctypes.byref(grid.to_c()), flags, precision, device,
                                                        ctypes.byref(ptr)), None)
        else:
            _check(self.lib.mif_ctx_create(ctypes.byref(grid.to_c()), precision, device, ctypes.byref(ptr)), None)
        self.ptr = ptr
        dt = ctypes.c_double()
        ext = (ctypes.c_uint64 * 3)()
        _check(self.lib.mif_ctx_info(self.ptr, ctypes.byref(dt), ext), self.ptr)
        self.dt = dt.value
        self.node_extents = tuple(int(e) for e in ext)
        self._fields = weakref.WeakSet()

    def close(self):
        """Destroy the context.  Its fields are released first; a Field
        object that outlives the context holds no device memory."""
        if self.ptr:
            for f in list(self._fields):
                f.close()
            self.lib.mif_ctx_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        # at interpreter exit the HIP runtime may already be torn down
        if sys.is_finalizing():
            return
        try:
            self.close()
        except Exception:
            pass

    def check(self, rc):
        _check(rc, self.ptr)

    def synchronize(self):
        self.check(self.lib.mif_ctx_synchronize(self.ptr))

    def set_defect_sweeps(self, n: int):
        self.check(self.lib.mif_ctx_set_defect_sweeps(self.ptr, n))

    def set_defect_schedule(self, schedule: int):
        """Which solves of timestep() take the defect sweeps:
        L.DEFECT_EVERY_STAGE (all three, the default) or L.DEFECT_LAST_STAGE
        (the solve of stage 3 only; stages 1 and 2 use FP32 transforms
        alone).  Stand-alone solves always take them."""
        self.check(self.lib.mif_ctx_set_defect_schedule(self.ptr, schedule))

    def set_solve_precision(self, mode: int):
        """Transforms of the mixed form's pressure solve: L.SOLVE_AUTO (FP32
        up to condition number 6e5, FP64 above), L.SOLVE_FP32, L.SOLVE_FP64."""
        self.check(self.lib.mif_ctx_set_solve_precision(self.ptr, mode))

    def solve_precision(self):
        """(effective transform precision, condition-number estimate)."""
        eff = ctypes.c_int()
        kappa = ctypes.c_double()
        self.check(self.lib.mif_ctx_get_solve_precision(self.ptr, ctypes.byref(eff), ctypes.byref(kappa)))
        return eff.value, kappa.value

    def set_stage_records(self, keep: bool):
        """keep: timestep() leaves the projected stage states Y2, Y3 in its
        two buffers (what timestep_adjoint reads); otherwise the projections
        of stages 1 and 2 are fused into the next stage kernel (the default)."""
        self.check(self.lib.mif_ctx_set_stage_records(self.ptr, 1 if keep else 0))

    def set_body_force(self, f):
        """Constant body force added to N in every stage of timestep()
        (zero = the reference's equation)."""
        arr = (ctypes.c_double * 3)(*[float(x) for x in f])
        self.check(self.lib.mif_ctx_set_body_force(self.ptr, arr))

    # fields -----------------------------------------------------------------
    def field(self, staggering: int) -> "Field":
        return Field(self, staggering)

    def velocity(self) -> "VelocityField":
        return VelocityField(self)

    def pressure(self) -> "Field":
        return Field(self, L.STAGGER_NONE)

    # stand-alone spectral solve on a packed owner-point right-hand side ---------
    @property
    def npts(self):
        g = self.grid
        Ng = (g.Nx_global, g.Ny_global, g.Nz_global)
        return tuple(Ng[a] - (1 if g.periodic[a] or self.cell_centred[a] else 0) for a in range(3))

    def poisson_set_rhs(self, b: np.ndarray):
        n = self.npts
        b = np.ascontiguousarray(b, dtype=np.float64)
        if b.shape != (n[2], n[1], n[0]):
            raise ValueError(f"shape {b.shape} != {(n[2], n[1], n[0])}")
        self.check(self.lib.mif_poisson_set_rhs(self.ptr, b.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))

    def poisson_run(self):
        self.check(self.lib.mif_poisson_run(self.ptr))

    def poisson_get_solution(self) -> np.ndarray:
        n = self.npts
        out = np.empty((n[2], n[1], n[0]), dtype=np.float64)
        self.check(self.lib.mif_poisson_get_solution(self.ptr, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return out

    # timing -----------------------------------------------------------------
    def set_timing(self, on: bool):
        self.check(self.lib.mif_ctx_set_timing(self.ptr, 1 if on else 0))

    def stream_probe(self, nread: int, nwrite: int, nbytes: int, reps: int = 10) -> float:
        """Average microseconds of a kernel streaming nread arrays in and
        nwrite arrays out, nbytes each (mif_stream_probe): the ceiling that
        read/write mix can reach on this device."""
        us = ctypes.c_double()
        self.check(self.lib.mif_stream_probe(self.ptr, int(nread), int(nwrite), int(nbytes), int(reps),
                                             ctypes.byref(us)))
        return us.value

    def kernel_time(self, cls: int):
        ms = ctypes.c_double()
        n = ctypes.c_uint64()
        self.check(self.lib.mif_ctx_kernel_time(self.ptr, cls, ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value


class Field:
    """A staggered tensor on the device (StaggeredTensor)."""

    def __init__(self, ctx: Context, staggering: int):
        self.ctx = ctx
        self.staggering = staggering
        ptr = ctypes.c_void_p()
        ctx.check(ctx.lib.mif_field_create(ctx.ptr, staggering, ctypes.byref(ptr)))
        self.ptr = ptr
        ctx._fields.add(self)
        ext = (ctypes.c_uint64 * 3)()
        ctx.check(ctx.lib.mif_field_extents(self.ptr, ext))
        self.extents = tuple(int(e) for e in ext)

    @property
    def shape(self):
        """numpy shape (sz, sy, sx) of the reference layout."""
        return (self.extents[2], self.extents[1], self.extents[0])

    def close(self):
        if self.ptr:
            self.ctx.lib.mif_field_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        if sys.is_finalizing():
            return
        try:
            self.close()
        except Exception:
            pass

    def upload(self, a: np.ndarray):
        a = np.ascontiguousarray(a, dtype=np.float64)
        if a.shape != self.shape:
            raise ValueError(f"shape {a.shape} != {self.shape}")
        self.ctx.check(self.ctx.lib.mif_field_upload(self.ptr, a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))

    def download(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=np.float64)
        self.ctx.check(self.ctx.lib.mif_field_download(self.ptr, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return out

    # ---- restart files (Tensor::dump / Tensor::load, include/Tensor.h:433-476,
    # and the exact binary variant of the C++ interface's StaggeredTensor) ----
    _BIN_MAGIC = b"MIFTNSR1"

    def dump(self, path: str):
        """Text dump in the reference's format: the dimensions line, then the
        values in storage order (x fastest) with 6 significant digits."""
        with open(path, "w") as f:
            f.write(format_text_dump(self.download()))

    def load(self, path: str):
        """Values from a text dump or a plain whitespace-separated list, in
        storage order; points the file does not reach keep their values."""
        text = open(path).read()
        if text.startswith("Tensor"):
            close = text.find("]")
            start = text.find("[", close) if close >= 0 else -1
            text = text[start + 1:] if start >= 0 else ""
        vals = []
        for tok in text.split():
            try:
                vals.append(float(tok))
            except ValueError:
                break
        cur = self.download().ravel()
        n = min(len(vals), cur.size)
        cur[:n] = vals[:n]
        self.upload(cur.reshape(self.shape))

    def dump_binary(self, path: str):
        """Exact restart file: 'MIFTNSR1', the extents (x, y, z), the
        staggering and the value size as little-endian int64, then the values
        as little-endian doubles in storage order."""
        sx = self.shape[::-1]
        head = np.array([sx[0], sx[1], sx[2], self.staggering, 8], dtype="<i8")
        with open(path, "wb") as f:
            f.write(self._BIN_MAGIC)
            f.write(head.tobytes())
            f.write(self.download().astype("<f8").tobytes())

    def load_binary(self, path: str):
        data = open(path, "rb").read()
        sx = self.shape[::-1]
        head = np.frombuffer(data, dtype="<i8", count=5, offset=8) if len(data) >= 48 else None
        if data[:8] != self._BIN_MAGIC or head is None or list(head) != [sx[0], sx[1], sx[2], self.staggering, 8]:
            raise ValueError(f"{path}: not a restart file of a field with these extents and staggering")
        n = int(np.prod(self.shape))
        if len(data) != 48 + 8 * n:
            raise ValueError(f"{path}: truncated restart file")
        self.upload(np.frombuffer(data, dtype="<f8", count=n, offset=48).reshape(self.shape))

    def upload_device(self, ptr: int):
        """Copy from a device array (reference layout, FP64) at address ptr."""
        self.ctx.check(self.ctx.lib.mif_field_upload_device(self.ptr, ctypes.c_void_p(ptr)))

    def download_device(self, ptr: int):
        """Copy into a device array (reference layout, FP64) at address ptr."""
        self.ctx.check(self.ctx.lib.mif_field_download_device(self.ptr, ctypes.c_void_p(ptr)))

    def copy_from(self, other: "Field"):
        self.ctx.check(self.ctx.lib.mif_field_copy(self.ptr, other.ptr))

    def dot(self, other: "Field") -> float:
        out = ctypes.c_double()
        self.ctx.check(self.ctx.lib.mif_field_dot(self.ptr, other.ptr, ctypes.byref(out)))
        return out.value

    def gather(self, ijk) -> np.ndarray:
        """Values at points (i, j, k) of the extents (an (n, 3) integer array),
        fetched from the device without downloading the field."""
        idx = np.ascontiguousarray(np.asarray(ijk, dtype=np.int32).reshape(-1, 3))
        out = np.empty(idx.shape[0])
        self.ctx.check(self.ctx.lib.mif_field_gather(
            self.ptr, idx.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), idx.shape[0],
            out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return out

    def set(self, fn: L.MifFunction, comp: int = 0, t: float = 0.0, include_border: bool = True):
        self.ctx.check(self.ctx.lib.mif_field_set(self.ptr, ctypes.byref(fn), comp, t, 1 if include_border else 0))

    def apply_periodic_bc(self):
        self.ctx.check(self.ctx.lib.mif_field_apply_periodic_bc(self.ptr))

    def swap_data(self, other: "Field"):
        self.ctx.check(self.ctx.lib.mif_field_swap(self.ptr, other.ptr))


class VelocityField:
    """u, v, w on their staggered grids (VelocityTensor)."""

    def __init__(self, ctx: Context):
        self.ctx = ctx
        self.u = Field(ctx, L.STAGGER_X)
        self.v = Field(ctx, L.STAGGER_Y)
        self.w = Field(ctx, L.STAGGER_Z)
        self.components = (self.u, self.v, self.w)

    def c_array(self):
        return L._V3(self.u.ptr, self.v.ptr, self.w.ptr)

    def set(self, fn: L.MifFunction, t: float = 0.0, include_border: bool = True):
        for c, f in enumerate(self.components):
            f.set(fn, c, t, include_border)

    def upload(self, arrays: Sequence[np.ndarray]):
        for f, a in zip(self.components, arrays):
            f.upload(a)

    def download(self):
        return tuple(f.download() for f in self.components)

    def apply_bc(self, exact_velocity, t: float):
        """VelocityTensor::apply_bc; exact_velocity is a descriptor or three
        callables f(t, x, y, z) evaluated on the host into a face table."""
        if _bd.is_callables(exact_velocity):
            tab = _bd.velocity_face_table(self.ctx.grid, exact_velocity, t)
            self.ctx.check(self.ctx.lib.mif_velocity_apply_bc_tabulated(self.c_array(), _dptr(tab)))
            return
        self.ctx.check(self.ctx.lib.mif_velocity_apply_bc(self.c_array(), ctypes.byref(exact_velocity), t))

    def swap_data(self, other: "VelocityField"):
        for a, b in zip(self.components, other.components):
            a.swap_data(b)

    def copy_from(self, other: "VelocityField"):
        for a, b in zip(self.components, other.components):
            a.copy_from(b)

    def dot(self, other: "VelocityField") -> float:
        return sum(a.dot(b) for a, b in zip(self.components, other.components))


# ---- reference-named free functions -------------------------------------------

def _dptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def timestep(velocity: VelocityField, velocity_buffer: VelocityField, velocity_buffer_2: VelocityField,
             exact_velocity, t_n: float, pressure: Field, pressure_buffer: Field):
    """timestep (include/Timestep.h:16-19).  exact_velocity is a descriptor
    (function()) or three callables f(t, x, y, z), evaluated on the host at
    the stage times into face tables."""
    ctx = velocity.ctx
    if _bd.is_callables(exact_velocity):
        times = _bd.full_step_times(ctx.dt, t_n)
        faces = np.concatenate([_bd.velocity_face_table(ctx.grid, exact_velocity, t) for t in times[1:]])
        ctx.check(ctx.lib.mif_timestep_tabulated(velocity.c_array(), velocity_buffer.c_array(),
                                                 velocity_buffer_2.c_array(), _dptr(faces), t_n, pressure.ptr,
                                                 pressure_buffer.ptr))
        return
    ctx.check(ctx.lib.mif_timestep(velocity.c_array(), velocity_buffer.c_array(), velocity_buffer_2.c_array(),
                                   ctypes.byref(exact_velocity), t_n, pressure.ptr, pressure_buffer.ptr))


def timestep_n(velocity: VelocityField, velocity_buffer: VelocityField, velocity_buffer_2: VelocityField,
               exact_velocity, t_n: float, pressure: Field, pressure_buffer: Field, nsteps: int):
    """nsteps timestep() calls at t_n, t_n + dt, ... in one call
    (mif_timestep_n): each step's last projection rides in the next step's
    first stage kernel.  exact_velocity is a descriptor (function())."""
    ctx = velocity.ctx
    if _bd.is_callables(exact_velocity):
        raise TypeError("timestep_n takes a function() descriptor; call timestep() per step for callables")
    ctx.check(ctx.lib.mif_timestep_n(velocity.c_array(), velocity_buffer.c_array(), velocity_buffer_2.c_array(),
                                     ctypes.byref(exact_velocity), t_n, pressure.ptr, pressure_buffer.ptr,
                                     int(nsteps)))


def timestep_nhn(velocity, velocity_buffer, velocity_buffer_2, exact_velocity, exact_pressure_gradient, t_n,
                 pressure, pressure_buffer):
    """timestep_nhn (include/Timestep.h:23-27).  Descriptors, or (both) three
    callables each: the Neumann data of stage s are then
    get_difference_over_time(new_time, prev_time) of the gradient
    (src/Timestep.cpp:89-95), tabulated on the host."""
    ctx = velocity.ctx
    if _bd.is_callables(exact_velocity) or _bd.is_callables(exact_pressure_gradient):
        if not (_bd.is_callables(exact_velocity) and _bd.is_callables(exact_pressure_gradient)):
            raise TypeError("timestep_nhn: give both functions as descriptors or both as callables")
        times = _bd.full_step_times(ctx.dt, t_n)
        faces = np.concatenate([_bd.velocity_face_table(ctx.grid, exact_velocity, t) for t in times[1:]])
        neumann = np.concatenate([
            _bd.neumann_face_table(ctx.grid, _bd.difference_over_time(exact_pressure_gradient, times[s],
                                                                      times[s - 1]), 0.0)
            for s in (1, 2, 3)])
        ctx.check(ctx.lib.mif_timestep_nhn_tabulated(velocity.c_array(), velocity_buffer.c_array(),
                                                     velocity_buffer_2.c_array(), _dptr(faces), _dptr(neumann), t_n,
                                                     pressure.ptr, pressure_buffer.ptr))
        return
    ctx.check(ctx.lib.mif_timestep_nhn(velocity.c_array(), velocity_buffer.c_array(), velocity_buffer_2.c_array(),
                                       ctypes.byref(exact_velocity), ctypes.byref(exact_pressure_gradient), t_n,
                                       pressure.ptr, pressure_buffer.ptr))


def timestep_velocity(velocity, velocity_buffer, rhs_buffer, exact_velocity, forcing, t_n):
    """timestep_velocity (include/TimestepVelocity.h:15-16); swaps velocity and
    velocity_buffer storage like the reference.  exact_velocity is a
    descriptor or three callables (host face tables); the forcing is the
    manufactured forcing descriptor, as the reference hard-wires it."""
    ctx = velocity.ctx
    if _bd.is_callables(exact_velocity):
        faces = np.concatenate([_bd.velocity_face_table(ctx.grid, exact_velocity, t)
                                for t in _bd.velocity_step_times(ctx.dt, t_n)])
        ctx.check(ctx.lib.mif_timestep_velocity_tabulated(velocity.c_array(), velocity_buffer.c_array(),
                                                          rhs_buffer.c_array(), _dptr(faces), ctypes.byref(forcing),
                                                          t_n))
        return
    ctx.check(ctx.lib.mif_timestep_velocity(velocity.c_array(), velocity_buffer.c_array(), rhs_buffer.c_array(),
                                            ctypes.byref(exact_velocity), ctypes.byref(forcing), t_n))


def solve_pressure_equation_homogeneous_periodic(pressure: Field, velocity: VelocityField, dt: float):
    """include/PressureEquation.h:10-13 (the PressureTensor buffer lives in the context)."""
    ctx = pressure.ctx
    ctx.check(ctx.lib.mif_solve_pressure_homogeneous_periodic(pressure.ptr, velocity.c_array(), dt))


def solve_pressure_equation_non_homogeneous_neumann(pressure, velocity, exact_pressure_gradient, t, dt):
    """include/PressureEquation.h:16-21; the gradient (a descriptor, or three
    callables tabulated on the host) is evaluated at time t."""
    ctx = pressure.ctx
    if _bd.is_callables(exact_pressure_gradient):
        tab = _bd.neumann_face_table(ctx.grid, exact_pressure_gradient, t)
        ctx.check(ctx.lib.mif_solve_pressure_non_homogeneous_neumann_tabulated(pressure.ptr, velocity.c_array(),
                                                                              _dptr(tab), dt))
        return
    ctx.check(ctx.lib.mif_solve_pressure_non_homogeneous_neumann(pressure.ptr, velocity.c_array(),
                                                                ctypes.byref(exact_pressure_gradient), t, dt))


def poisson_solve(pressure: Field):
    """The spectral solve alone on the owner points of `pressure`."""
    pressure.ctx.check(pressure.ctx.lib.mif_poisson_solve(pressure.ptr))


def adjust_pressure(pressure: Field, exact_pressure: L.MifFunction, t: float):
    """include/PressureEquation.h:25-26."""
    pressure.ctx.check(pressure.ctx.lib.mif_adjust_pressure(pressure.ptr, ctypes.byref(exact_pressure), t))


def error_norms_velocity(velocity: VelocityField, exact_velocity: L.MifFunction, t: float):
    """(ErrorL1Norm, ErrorL2Norm, ErrorLInfNorm) of the velocity (include/Norms.h:19-29)."""
    out = (ctypes.c_double * 3)()
    velocity.ctx.check(velocity.ctx.lib.mif_error_norms_velocity(velocity.c_array(), ctypes.byref(exact_velocity),
                                                                 t, out))
    return np.array(list(out))


def error_norms_pressure(pressure: Field, exact_pressure: L.MifFunction, t: float):
    """(ErrorL1Norm, ErrorL2Norm, ErrorLInfNorm) of a scalar field (include/Norms.h:32-34)."""
    out = (ctypes.c_double * 3)()
    pressure.ctx.check(pressure.ctx.lib.mif_error_norms_pressure(pressure.ptr, ctypes.byref(exact_pressure), t, out))
    return np.array(list(out))


def timestep_adjoint(u_n: VelocityField, y2: VelocityField, y3: VelocityField, u_bar: VelocityField,
                     p_bar: Field, force_bar):
    """Adjoint of one timestep(): (u_bar, p_bar) go from the adjoints of
    (u^{n+1}, p^{n+1}) to those of (u^n, p^n); force_bar (3) accumulates.
    u_n, y2, y3 are the stencil sources of the forward step (u^n and the
    buffers `velocity_buffer`, `velocity_buffer_2` as timestep() leaves them).
    force_bar=None: the force adjoint accumulates on the device (no
    synchronisation) until adjoint_force_take()."""
    ctx = u_n.ctx
    if force_bar is None:
        ctx.check(ctx.lib.mif_timestep_adjoint(u_n.c_array(), y2.c_array(), y3.c_array(), u_bar.c_array(), p_bar.ptr,
                                               None))
        return
    fb = (ctypes.c_double * 3)(*[float(x) for x in force_bar])
    ctx.check(ctx.lib.mif_timestep_adjoint(u_n.c_array(), y2.c_array(), y3.c_array(), u_bar.c_array(), p_bar.ptr, fb))
    for a in range(3):
        force_bar[a] = fb[a]


def adjoint_force_take(ctx: "Context") -> np.ndarray:
    """The force adjoint accumulated by timestep_adjoint(..., None) calls
    since the last take; resets the accumulator."""
    fb = (ctypes.c_double * 3)(0.0, 0.0, 0.0)
    ctx.check(ctx.lib.mif_adjoint_force_take(ctx.ptr, fb))
    return np.array(list(fb))


def pressure_gradient(pressure: Field, out: VelocityField):
    """Pressure gradient on interior points (include/PressureGradient.h:9-25)."""
    pressure.ctx.check(pressure.ctx.lib.mif_pressure_gradient(pressure.ptr, out.c_array()))
