// Domain constants (include/Constants.h, src/Constants.cpp:58-120 of the
// reference) for one rank holding the whole domain (Py = Pz = 1, rank 0),
// plus the device context that holds every field of the domain on one GPU.
// The public members keep the reference's names and values so that callers
// reading them (loop bounds, spacings, time step) are unchanged.
#ifndef MIF_SHIM_CONSTANTS_H
#define MIF_SHIM_CONSTANTS_H

#include <array>
#include <cstddef>

#include "mif_gpu.h"
#include "Real.h"

namespace mif {

class Constants {
 public:
  const size_t Nx_global;
  const size_t Ny_global;
  const size_t Nz_global;
  const Real x_size;
  const Real y_size_global;
  const Real z_size_global;
  const Real min_x_global;
  const Real min_y_global;
  const Real min_z_global;
  const Real Re;
  const Real final_time;
  const unsigned int num_time_steps;
  const std::array<bool, 3> periodic_bc;

  // Decomposition: always one rank.
  const int Py;
  const int Pz;
  const int rank;
  const int y_rank;
  const int z_rank;

  const Real dt;
  const size_t Nx_domains;
  const size_t Ny_domains_global;
  const size_t Nz_domains_global;
  const Real dx;
  const Real dy;
  const Real dz;
  const Real one_over_2_dx;
  const Real one_over_2_dy;
  const Real one_over_2_dz;
  const Real one_over_8_dx;
  const Real one_over_8_dy;
  const Real one_over_8_dz;
  const Real one_over_dx2_Re;
  const Real one_over_dy2_Re;
  const Real one_over_dz2_Re;
  const Real dx_over_2;
  const Real dy_over_2;
  const Real dz_over_2;
  const Real one_over_dx;
  const Real one_over_dy;
  const Real one_over_dz;

  const int P;
  const size_t Ny_owner;
  const size_t Nz_owner;
  const size_t Nx;
  const size_t Ny;
  const size_t Nz;
  const size_t Nx_staggered;
  const size_t Ny_staggered;
  const size_t Nz_staggered;
  const int base_i;
  const int base_j;
  const int base_k;

  // No neighbouring ranks.
  const int prev_proc_y;
  const int next_proc_y;
  const int prev_proc_z;
  const int next_proc_z;

  // The reference's constructor.  Py, Pz and rank must describe a single
  // rank (1, 1, 0); anything else throws std::invalid_argument.  precision
  // selects the device form (MIF_PRECISION_FP64 or MIF_PRECISION_MIXED) and
  // device the HIP device (-1: the current one).  Throws std::runtime_error
  // when the device context cannot be created.
  Constants(size_t Nx_global, size_t Ny_global, size_t Nz_global, Real x_size, Real y_size_global,
            Real z_size_global, Real min_x_global, Real min_y_global, Real min_z_global, Real Re, Real final_time,
            unsigned int num_time_steps, int Py, int Pz, int rank, const std::array<bool, 3> &periodic_bc,
            int precision = MIF_PRECISION_FP64, int device = -1);
  Constants(const Constants &) = delete;
  Constants &operator=(const Constants &) = delete;
  ~Constants();

  // The device context of this domain.
  mif_ctx *context() const { return ctx_; }

 private:
  mif_ctx *ctx_ = nullptr;
};

// Throws std::runtime_error with the library's message when rc is not MIF_OK.
void check_status(int rc, const mif_ctx *ctx, const char *what);

}  // namespace mif

#endif  // MIF_SHIM_CONSTANTS_H
