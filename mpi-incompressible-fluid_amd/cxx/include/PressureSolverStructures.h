// Pressure-solver set-up (PressureSolverStructures, include/
// PressureSolverStructures.h:14-50 of the reference).  The transform plans,
// eigenvalue tables and scratch buffers it holds in the reference live in the
// device context of the domain; this object keeps the reference's
// problem-size members for callers that read them.
#ifndef MIF_SHIM_PRESSURE_SOLVER_STRUCTURES_H
#define MIF_SHIM_PRESSURE_SOLVER_STRUCTURES_H

#include "Constants.h"

namespace mif {

class PressureSolverStructures {
 public:
  explicit PressureSolverStructures(const Constants &constants);
  PressureSolverStructures(const PressureSolverStructures &) = delete;
  PressureSolverStructures &operator=(const PressureSolverStructures &) = delete;

  const Constants &constants;
  bool periodic_bc[3];
  // Unknowns per axis: the node count on a wall axis, one less on a
  // periodic axis (src/PressureSolverStructures.cpp:15-17).
  const int Nx_points;
  const int Ny_points;
  const int Nz_points;
};

}  // namespace mif

#endif  // MIF_SHIM_PRESSURE_SOLVER_STRUCTURES_H
