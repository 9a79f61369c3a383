// mif_device.h -- device-side data model shared by the HIP kernels.
//
// Every field of one context lives in the same padded x-fastest layout:
// element (i,j,k) of any staggering is at i + PX*(j + SY*k), where PX is the
// largest x extent rounded up to 64 bytes and SY the largest y extent.  The
// logical extents of each staggering are those of the reference tensors
// (src/StaggeredTensor.cpp:5-9, src/Constants.cpp:80-85 at one rank); the
// padding is never read as data.  A common layout lets a stencil thread
// address u, v, w and p at one offset.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/mif_types.h"

namespace mif {

// Single-rank domain constants (src/Constants.cpp:58-120 at Py = Pz = 1).
struct DGrid {
  int N[3];        // Nx, Ny, Nz: node extents incl. periodic ghosts
  int Ng[3];       // global node counts
  int base[3];     // base_i/j/k
  int periodic[3];
  int npts[3];     // Poisson unknowns per axis
  int olo[3];      // first owner index per axis (1 on periodic axes)
  int cell[3];     // 1: cell-centred Neumann axis (stand-alone Poisson solves)
  double h[3], min[3];
  double one_over_2h[3], one_over_8h[3], one_over_h2Re[3], h_over_2[3], one_over_h[3];
  double Re, dt;
};

struct Layout {
  int PX, SY, SZ;  // padded pitch, rows per plane, planes
  long plane;      // PX * SY
  long elems;      // PX * SY * SZ
  int lead;        // elements between the (aligned) allocation start and element 0
  int nt;          // nonzero: the streaming kernels store their outputs non-temporally
  __host__ __device__ inline long idx(int i, int j, int k) const {
    return (long)i + (long)PX * ((long)j + (long)SY * (long)k);
  }
};

// Extents of a field with staggering s (0 = x, 1 = y, 2 = z, 3 = none).
__host__ __device__ inline void extents(const DGrid& g, int s, int e[3]) {
  e[0] = g.N[0];
  e[1] = g.N[1];
  e[2] = g.N[2];
  if (s == 0) e[0] = g.periodic[0] ? g.N[0] : g.N[0] + 1;
  if (s == 1) e[1] = g.N[1] + 1;
  if (s == 2) e[2] = g.N[2] + 1;
}

// ---------------------------------------------------------------------------
// Analytic functions (device closed forms of the reference's generated
// callbacks; see include/mif_types.h for the list).
// ---------------------------------------------------------------------------
// kind kFnFaceTable: boundary data evaluated on the host, for callbacks that
// have no closed form here.  tab holds, for each component q (of a velocity;
// only q = 0 for the Neumann data of a node field) and each face
// f = z-, z+, y-, y+, x-, x+, a plane of values over the field's extents
// along the two other axes, lower axis fastest (face_plane_index), starting
// at tab + toff[q][f]; faces of periodic axes have none.  fn_eval never
// reads it; the boundary and Neumann-data kernels look values up instead.
constexpr int kFnFaceTable = 1000;

struct FnDesc {
  int kind, base;
  double Re, c[3], t0, t1;
  const double* tab;
  long toff[3][6];
};

// Index of point (i, j, k) in the plane of face f (0,1 = z-,z+; 2,3 =
// y-,y+; 4,5 = x-,x+) of a field with extents e.
__host__ __device__ inline long face_plane_index(int face, const int e[3], int i, int j, int k) {
  const int axis = 2 - (face >> 1);
  if (axis == 2) return (long)i + (long)e[0] * j;
  if (axis == 1) return (long)i + (long)e[0] * k;
  return (long)j + (long)e[1] * k;
}

__host__ inline FnDesc make_fn(const mif_function* f) {
  FnDesc d{};
  if (f) {
    d.kind = f->kind;
    d.base = f->base;
    d.Re = f->Re;
    d.c[0] = f->c[0];
    d.c[1] = f->c[1];
    d.c[2] = f->c[2];
    d.t0 = f->t0;
    d.t1 = f->t1;
  }
  return d;
}

#define MIF_PI 3.14159265358979323846

// Ethier-Steinman u (generators/manufsol.py:36-44); v, w by cyclic argument
// permutation.
__device__ inline double es_u(double t, double x, double y, double z, double Re) {
  const double a = MIF_PI / 4, d = MIF_PI / 2;
  return -a * (exp(a * x) * sin(a * y + d * z) + exp(a * z) * cos(a * x + d * y)) * exp(-d * d * t / Re);
}

// Ethier-Steinman p (generators/manufsol.py:55-68).
__device__ inline double es_p(double t, double x, double y, double z, double Re) {
  const double a = MIF_PI / 4, d = MIF_PI / 2;
  const double s = exp(2 * a * x) + exp(2 * a * y) + exp(2 * a * z) +
                   2 * sin(a * x + d * y) * cos(a * z + d * x) * exp(a * (y + z)) +
                   2 * sin(a * y + d * z) * cos(a * x + d * y) * exp(a * (z + x)) +
                   2 * sin(a * z + d * x) * cos(a * y + d * z) * exp(a * (x + y));
  return -a * a / 2 * s * exp(-2 * d * d * t / Re);
}

// d p / d x of the Ethier-Steinman pressure (term-by-term derivative).
__device__ inline double es_dpdx(double t, double x, double y, double z, double Re) {
  const double a = MIF_PI / 4, d = MIF_PI / 2;
  const double S1 = sin(a * x + d * y), C1 = cos(a * z + d * x);
  const double S2 = sin(a * y + d * z), C2 = cos(a * x + d * y);
  const double S3 = sin(a * z + d * x), C3 = cos(a * y + d * z);
  const double s = 2 * a * exp(2 * a * x) + 2 * exp(a * (y + z)) * (a * C2 * C1 - d * S1 * S3) +
                   2 * a * exp(a * (z + x)) * S2 * (C2 - S1) + 2 * exp(a * (x + y)) * C3 * (d * C1 + a * S3);
  return -a * a / 2 * s * exp(-2 * d * d * t / Re);
}

// Forcing of the velocity-only manufactured field: f = u_t + (u.grad)u -
// lap(u)/Re of MIF_FN_VT_VELOCITY (generators/manufsol_velocity.py:40-50),
// simplified, from the sines and cosines of the coordinates and the time.
__host__ __device__ inline double vt_forcing(int c, double sx, double cx, double sy, double cy, double sz, double cz,
                                             double st, double ct, double Re) {
  if (c == 0) return sx * cx * st * st * (2 - 2 * sy * sy - sz * sz) + sx * sz * cy * (ct + 3 * st / Re);
  if (c == 1) return sy * cy * st * st * (2 - 2 * sx * sx - sz * sz) + sy * sz * cx * (ct + 3 * st / Re);
  return sz * cz * st * st * (2 * sx * sx + 2 * sy * sy - 4) + cx * cy * cz * (2 * ct + 6 * st / Re);
}

// Kept out of line: the closed forms are large and only the boundary,
// forcing and set-up kernels evaluate them.
__device__ __noinline__ double fn_base(const FnDesc& f, int kind, int c, double t, double x, double y, double z) {
  switch (kind) {
    case MIF_FN_ZERO:
      return 0.0;
    case MIF_FN_ES_VELOCITY:
      return c == 0 ? es_u(t, x, y, z, f.Re) : (c == 1 ? es_u(t, y, z, x, f.Re) : es_u(t, z, x, y, f.Re));
    case MIF_FN_ES_PRESSURE:
      return es_p(t, x, y, z, f.Re);
    case MIF_FN_ES_GRAD_P:
      return c == 0 ? es_dpdx(t, x, y, z, f.Re) : (c == 1 ? es_dpdx(t, y, z, x, f.Re) : es_dpdx(t, z, x, y, f.Re));
    case MIF_FN_VT_VELOCITY:
      // generators/manufsol_velocity.py:55-57
      if (c == 0) return sin(x) * cos(y) * sin(z) * sin(t);
      if (c == 1) return cos(x) * sin(y) * sin(z) * sin(t);
      return 2 * cos(x) * cos(y) * cos(z) * sin(t);
    case MIF_FN_VT_FORCING:
      return vt_forcing(c, sin(x), cos(x), sin(y), cos(y), sin(z), cos(z), sin(t), cos(t), f.Re);
    case MIF_FN_PT_VELOCITY:
    case MIF_FN_PT_GRAD_P:
      // generators/manufsol_pressure.py:36-38 (grad p equals the velocity)
      if (c == 0) return -sin(x) * cos(y) * cos(z) * t;
      if (c == 1) return -cos(x) * sin(y) * cos(z) * t;
      return -cos(x) * cos(y) * sin(z) * t;
    case MIF_FN_PT_PRESSURE:
      return cos(x) * cos(y) * cos(z) * t;
    case MIF_FN_LID_1:
    case MIF_FN_LID_2: {
      // include/TestCaseBoundaries.h:18-56: v = 1 on the lid plane x = 1
      // (test 1) or x = -0.5 (test 2), zero elsewhere.
      const double xl = kind == MIF_FN_LID_1 ? 1.0 : -0.5;
      const double eps = 1e-12;
      return (c == 1 && x < xl + eps && x > xl - eps) ? 1.0 : 0.0;
    }
    case MIF_FN_CONSTANT:
      return f.c[c];
    default:
      return 0.0;
  }
}

__device__ inline double fn_eval(const FnDesc& f, int c, double t, double x, double y, double z) {
  if (f.kind == MIF_FN_TIME_DIFFERENCE)
    return fn_base(f, f.base, c, f.t1, x, y, z) - fn_base(f, f.base, c, f.t0, x, y, z);
  return fn_base(f, f.kind, c, t, x, y, z);
}

// Coordinate of index i along axis a for staggering s: min + h*(base+i),
// minus h/2 on the staggered axis (include/StaggeredTensor.h:95-110,
// include/VelocityTensor.h:16-78).  Contraction is off so that the
// coordinates round exactly as the reference computes them.
__device__ inline double coord(const DGrid& g, int a, int s, int i) {
#pragma clang fp contract(off)
  double x = g.min[a] + g.h[a] * (double)(g.base[a] + i);
  if (a == s) x = x - g.h_over_2[a];
  return x;
}

__device__ inline double eval_at(const DGrid& g, const FnDesc& f, int s, int c, double t, int i, int j, int k) {
  return fn_eval(f, c, t, coord(g, 0, s, i), coord(g, 1, s, j), coord(g, 2, s, k));
}

// Weight of owner index i along axis a in the left null vector of the 1-D
// operator the solve inverts: uniform on periodic and cell-centred axes,
// trapezoid (1/2 at the two end nodes) on node Neumann axes.
__device__ inline double owner_weight(const DGrid& g, int a, int i) {
  if (g.periodic[a] || g.cell[a]) return 1.0;
  return (i == 0 || i == g.npts[a] - 1) ? 0.5 : 1.0;
}

// Workgroups of a launch are dispatched round-robin over the 8 XCDs of the
// device, each with its own L2.  Mapping the dispatch index b of n
// workgroups to this tile index gives every XCD one contiguous range of
// tiles, so tiles that share halo lines run on the same L2.  A bijection on
// [0, n).
__device__ inline int xcd_tile(int b, int n) {
  constexpr int NXCD = 8;
  const int per = n / NXCD, rem = n % NXCD, x = b % NXCD;
  return x * per + (x < rem ? x : rem) + b / NXCD;
}

}  // namespace mif
