"""MI355X-native incompressible Navier-Stokes timestep (host side).

The compute path is the HIP library built from ../csrc into ../_build
(libmif_gpu.so); this package is its Python host binding, mirroring the
reference solver's interfaces (see core.py).
"""
from ._lib import (DEFECT_EVERY_STAGE, DEFECT_LAST_STAGE, FP64, MIXED, SOLVE_AUTO, SOLVE_FP32, SOLVE_FP64, STAGGER_NONE, STAGGER_X, STAGGER_Y, STAGGER_Z, FN_CONSTANT, FN_ES_GRAD_P,
                   FN_ES_PRESSURE, FN_ES_VELOCITY, FN_LID_1, FN_LID_2, FN_PT_GRAD_P, FN_PT_PRESSURE,
                   FN_PT_VELOCITY, FN_TIME_DIFFERENCE, FN_VT_FORCING, FN_VT_VELOCITY, FN_ZERO, KERNEL_CLASSES,
                   LIB_PATH, load)
from .core import (Context, Field, GridSpec, MifError, VelocityField, adjoint_force_take, adjust_pressure,
                   error_norms_pressure,
                   error_norms_velocity, function, poisson_solve, pressure_gradient,
                   solve_pressure_equation_homogeneous_periodic, solve_pressure_equation_non_homogeneous_neumann,
                   timestep, timestep_adjoint, timestep_n, timestep_nhn, timestep_velocity)

__all__ = [n for n in dir() if not n.startswith("_")]
