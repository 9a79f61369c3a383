// mif_adjoint.hip -- adjoint of one RK3 + projection step (no counterpart in
// the reference, which has no adjoint).
//
// The forward step (src/Timestep.cpp:97-141) is, per stage: a stencil update
// with the momentum right-hand side N (quadratic in the velocity,
// include/MomentumEquation.h:43-255) and the pressure gradient; the boundary
// map (Dirichlet faces are constants, periodic ghosts copy interior points,
// src/VelocityTensor.cpp:36-233); the divergence, the spectral solve M, the
// unpack with periodic images; p += phi; u -= dt_s grad(phi) on interior
// points.  Each kernel below is the transpose of a run of those linear maps,
// written as a gather (each output point sums the contributions that the
// forward scatters into it), so no atomics are needed and results are
// deterministic.  One adjoint stage is three kernels around the solve:
//   k_adj_rhs   phi_bar = p_bar - dt_s G^T y_bar, then U^T and W^-1: the
//               packed right-hand side of M^T = W M W^-1 (W = trapezoid
//               weights of the owner nodes)
//   (spectral solve)
//   k_adj_zb    z_bar = B^T (y_bar + D^T W psi / dt_s): the adjoint of the
//               stage output before its boundary map; also N_bar, the adjoint
//               of N's output, and per-block sums of z_bar (body force)
//   k_adj_xb    x_bar = z_bar + J_N(X)^T N_bar (the transposed linearised
//               convection plus the diffusion stencil), and p_bar += b G^T z_bar
//               for the stage's pressure-gradient term
// z_bar and N_bar vanish outside the component interiors (the forward
// computes N and writes the stage update only there), which lets k_adj_xb
// read neighbours without interior tests.
#include "mif_kernels.h"

namespace mif {

__device__ inline bool adj_interior(const DGrid& g, int c, int i, int j, int k) {
  int e[3];
  extents(g, c, e);
  return i >= 1 && j >= 1 && k >= 1 && i <= e[0] - 2 && j <= e[1] - 2 && k <= e[2] - 2;
}

__device__ inline bool in_field(const DGrid& g, int s, int i, int j, int k) {
  int e[3];
  extents(g, s, e);
  return i >= 0 && j >= 0 && k >= 0 && i < e[0] && j < e[1] && k < e[2];
}

__device__ inline double trap_weight(const DGrid& g, int a, int i) {
  if (g.periodic[a]) return 1.0;
  return (i == 0 || i == g.npts[a] - 1) ? 0.5 : 1.0;
}

// Points of a periodic axis that hold copies of index P (P itself first):
// 1 -> also s-1, s-2 -> also 0, on an axis of extent s.
__device__ inline int images(const DGrid& g, int a, int P, int s, int out[2]) {
  out[0] = P;
  int n = 1;
  if (g.periodic[a]) {
    if (P == 1) out[n++] = s - 1;
    if (P == s - 2) out[n++] = 0;
  }
  return n;
}

// ---------------------------------------------------------------------------
// k_adj_rhs: for each owner node o of the packed solve,
//   B(o) = sum over the pressure points R holding o (o and its periodic
//          images) of phi_bar(R), divided by W(o),
//   phi_bar(R) = p_bar(R) + dt_s sum_c (y_bar_c(R + e_c) - y_bar_c(R)) / h_c
// with y_bar_c taken on the interior of component c only.
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256) k_adj_rhs(DGrid g, Layout L, const T* __restrict__ P, AdjVec Y, double dts,
                                                 T* __restrict__ B, long bp0, long bplane, long total) {
  // flat indices fit in 32 bits (the context bounds every field below 2^31
  // elements), so the decomposition uses 32-bit division
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= total) return;
  const int a = q % g.npts[0], rq = q / g.npts[0], b = rq % g.npts[1], c = rq / g.npts[1];
  const int O[3] = {a + g.olo[0], b + g.olo[1], c + g.olo[2]};
  const long off[3] = {1, L.PX, L.plane};
  // R and R + e_c lie inside the padded array for every pressure point R,
  // so all seven loads are issued unconditionally and the interior masks
  // apply to the values
  auto phib = [&](int r0, int r1, int r2) -> double {
    const int R[3] = {r0, r1, r2};
    const long o = L.idx(r0, r1, r2);
    double hi[3], lo[3];
#pragma unroll
    for (int cc = 0; cc < 3; cc++) {
      const T* F = (const T*)Y.f[cc];
      hi[cc] = (double)F[o + off[cc]];
      lo[cc] = (double)F[o];
    }
    const double pv = (double)P[o];
    double acc = 0;
#pragma unroll
    for (int cc = 0; cc < 3; cc++) {
      int Rp[3] = {r0, r1, r2};
      Rp[cc] += 1;
      double d = 0;
      if (adj_interior(g, cc, Rp[0], Rp[1], Rp[2])) d += hi[cc];
      if (adj_interior(g, cc, R[0], R[1], R[2])) d -= lo[cc];
      acc += d * g.one_over_h[cc];
    }
    return pv + dts * acc;
  };
  int opts[3][2], nopt[3];
  bool single = true;
  for (int ax = 0; ax < 3; ax++) {
    nopt[ax] = images(g, ax, O[ax], g.N[ax], opts[ax]);
    single = single && nopt[ax] == 1;
  }
  double s = 0;
  if (single) {
    s = phib(O[0], O[1], O[2]);
  } else {
    for (int qa = 0; qa < nopt[0]; qa++)
      for (int qb = 0; qb < nopt[1]; qb++)
        for (int qc = 0; qc < nopt[2]; qc++) s += phib(opts[0][qa], opts[1][qb], opts[2][qc]);
  }
  const double w = trap_weight(g, 0, a) * trap_weight(g, 1, b) * trap_weight(g, 2, c);
  B[a + bp0 * b + bplane * c] = (T)(s / w);
}

// ---------------------------------------------------------------------------
// k_adj_zb: component c = blockIdx.y, interior points Q of c (grid-stride):
//   z_bar_c(Q) = sum over the points g holding Q (Q and its periodic ghost
//                copies) of y_bar_c(g) + (Wpsi(g - e_c) - Wpsi(g)) / (dt_s h_c)
//   N_bar_c(Q) = alpha z_bar_c(Q) + beta Zo_c(Q)
// with Wpsi = W * psi on owner nodes (zero elsewhere), and per workgroup the
// partial sum of z_bar_c weighted by fcoef, stored (accumulate = 0) or added
// to the workgroup's slot (accumulate = 1: the later stages of one adjoint
// step, launched with the same grid, so slot b always belongs to workgroup b).
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256) k_adj_zb(DGrid g, Layout L, AdjVec Y, const T* __restrict__ Bpsi, long bp0,
                                                long bplane, double dts, AdjVec Zb, AdjVec Nb, AdjVec Zo, double alpha,
                                                double beta, double fcoef, int accumulate, double* __restrict__ partial) {
  const int c = blockIdx.y;
  int e[3];
  extents(g, c, e);
  const int n0 = e[0] - 2, n1 = e[1] - 2, n2 = e[2] - 2;
  const int total = (n0 > 0 && n1 > 0 && n2 > 0) ? n0 * n1 * n2 : 0;
  const T* y = (const T*)Y.f[c];
  T* zb = (T*)Zb.f[c];
  T* nb = (T*)Nb.f[c];
  const T* zo = (const T*)Zo.f[c];
  // W * psi at pressure point P (zero off the owner nodes); the load reads
  // a valid address either way
  auto wpsi = [&](int pi, int pj, int pk) -> double {
    const int q0 = pi - g.olo[0], q1 = pj - g.olo[1], q2 = pk - g.olo[2];
    const bool ok = (unsigned)q0 < (unsigned)g.npts[0] && (unsigned)q1 < (unsigned)g.npts[1] &&
                    (unsigned)q2 < (unsigned)g.npts[2];
    const double v = (double)Bpsi[ok ? q0 + bp0 * q1 + bplane * q2 : 0];
    const double w = trap_weight(g, 0, q0) * trap_weight(g, 1, q1) * trap_weight(g, 2, q2);
    return ok ? w * v : 0.0;
  };
  double part = 0;
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < total; q += gridDim.x * blockDim.x) {
    const int rq = q / n0;
    const int Q[3] = {q % n0 + 1, rq % n1 + 1, rq / n1 + 1};
    auto term = [&](int g0, int g1, int g2) -> double {
      int Gm[3] = {g0, g1, g2};
      Gm[c] -= 1;
      const double d = (wpsi(Gm[0], Gm[1], Gm[2]) - wpsi(g0, g1, g2)) * g.one_over_h[c] / dts;
      return (double)(T)((double)y[L.idx(g0, g1, g2)] + d);
    };
    int opts[3][2], nopt[3];
    bool single = true;
    for (int ax = 0; ax < 3; ax++) {
      nopt[ax] = images(g, ax, Q[ax], e[ax], opts[ax]);
      single = single && nopt[ax] == 1;
    }
    double s = 0;
    if (single) {
      s = term(Q[0], Q[1], Q[2]);
    } else {
      for (int qa = 0; qa < nopt[0]; qa++)
        for (int qb = 0; qb < nopt[1]; qb++)
          for (int qc = 0; qc < nopt[2]; qc++) s += term(opts[0][qa], opts[1][qb], opts[2][qc]);
    }
    const long o = L.idx(Q[0], Q[1], Q[2]);
    const T z = (T)s;
    zb[o] = z;
    // the later stage's z_bar enters only with beta != 0 (stage 2): the
    // other stages skip its load (a uniform branch)
    double nbv = alpha * (double)z;
    if (beta != 0.0) nbv += beta * (double)zo[o];
    nb[o] = (T)nbv;
    part += (double)z;
  }
  __shared__ double sh[4];
  for (int off = 32; off > 0; off >>= 1) part += __shfl_down(part, off, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = part;
  __syncthreads();
  if (threadIdx.x == 0) {
    const double v = fcoef * (sh[0] + sh[1] + sh[2] + sh[3]);
    double* slot = &partial[c * gridDim.x + blockIdx.x];
    *slot = accumulate ? *slot + v : v;
  }
}

// out[c] (= or +=) sum of the nblk partials of component c: one workgroup of
// 1024 threads, each loading its kAdjSumPer values per component at once,
// then a fixed-order wave and workgroup reduction (deterministic).
constexpr int kAdjSumThreads = 1024, kAdjSumPer = 4;
static_assert(kAdjSumThreads * kAdjSumPer >= kAdjMaxBlocks, "k_adj_sum covers every partial");
__global__ void __launch_bounds__(kAdjSumThreads) k_adj_sum(const double* __restrict__ partial, int nblk,
                                                            int accumulate, double* __restrict__ out) {
  double v[3][kAdjSumPer];
#pragma unroll
  for (int c = 0; c < 3; c++)
#pragma unroll
    for (int q = 0; q < kAdjSumPer; q++) {
      const int b = threadIdx.x + kAdjSumThreads * q;
      const double x = partial[c * nblk + (b < nblk ? b : 0)];
      v[c][q] = b < nblk ? x : 0.0;
    }
  __shared__ double sh[3][kAdjSumThreads / 64];
#pragma unroll
  for (int c = 0; c < 3; c++) {
    double s = (v[c][0] + v[c][1]) + (v[c][2] + v[c][3]);
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if ((threadIdx.x & 63) == 0) sh[c][threadIdx.x >> 6] = s;
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    const int c = threadIdx.x;
    double s = 0;
    for (int w = 0; w < kAdjSumThreads / 64; w++) s += sh[c][w];
    out[c] = accumulate ? out[c] + s : s;
  }
}

// ---------------------------------------------------------------------------
// k_adj_xb: for every point Q of the three components,
//   x_bar_c(Q) = z_bar_c(Q) + (J_N(X)^T N_bar)_c(Q)
// and, at pressure points, p_bar(Q) += b sum_c (z_bar_c(Q + e_c) - z_bar_c(Q)) / h_c.
// N_bar and z_bar are zero outside the component interiors and in the
// padding, and every tap index is clamped into the array, so taps need no
// tests: a clamped N_bar tap reads a boundary or padding cell (zero) and the
// X value it multiplies is finite.
// ---------------------------------------------------------------------------
// x_bar at one point Q = (i, j, k) from tap providers: tp(f, di, dj, dk)
// returns X_f (f = 0..2) or N_bar_{f-3} (f = 3..5) at Q + (di, dj, dk), and
// zt(c, di, dj, dk) returns z_bar_c there.  Offsets are literals at every
// call site, so a provider can pick its source at compile time.
template <typename T, typename TP, typename ZT>
__device__ inline void adj_xb_point(const DGrid& g, long o, int i, int j, int k, TP tp, ZT zt, T* __restrict__ XB0,
                                    T* __restrict__ XB1, T* __restrict__ XB2, T* __restrict__ P, double bcoef) {
  const double A2 = g.one_over_2h[0], B2 = g.one_over_2h[1], C2 = g.one_over_2h[2];
  const double A8 = g.one_over_8h[0], B8 = g.one_over_8h[1], C8 = g.one_over_8h[2];
  const double Dx = g.one_over_h2Re[0], Dy = g.one_over_h2Re[1], Dz = g.one_over_h2Re[2];
  const double Dsum = Dx + Dy + Dz;

  // u at Q
  if (in_field(g, 0, i, j, k)) {
    double acc = zt(0, 0, 0, 0);
    acc += tp(3, 0, 0, 0) * (-(tp(0, 1, 0, 0) - tp(0, -1, 0, 0)) * A2 - 2 * Dsum);
    acc += tp(3, -1, 0, 0) * (-tp(0, -1, 0, 0) * A2 + Dx);
    acc += tp(3, 1, 0, 0) * (tp(0, 1, 0, 0) * A2 + Dx);
    // Vs_u(P) = v(P+y) + v(P) + v(P-x+y) + v(P-x), at P = Q -+ y
    const double vsm = tp(1, 0, 0, 0) + tp(1, 0, -1, 0) + tp(1, -1, 0, 0) + tp(1, -1, -1, 0);
    const double vsp = tp(1, 0, 2, 0) + tp(1, 0, 1, 0) + tp(1, -1, 2, 0) + tp(1, -1, 1, 0);
    acc += tp(3, 0, -1, 0) * (-vsm * B8 + Dy);
    acc += tp(3, 0, 1, 0) * (vsp * B8 + Dy);
    // Ws_u(P) = w(P+z) + w(P) + w(P-x+z) + w(P-x), at P = Q -+ z
    const double wsm = tp(2, 0, 0, 0) + tp(2, 0, 0, -1) + tp(2, -1, 0, 0) + tp(2, -1, 0, -1);
    const double wsp = tp(2, 0, 0, 2) + tp(2, 0, 0, 1) + tp(2, -1, 0, 2) + tp(2, -1, 0, 1);
    acc += tp(3, 0, 0, -1) * (-wsm * C8 + Dz);
    acc += tp(3, 0, 0, 1) * (wsp * C8 + Dz);
    // from the v equation at P = Q-x, Q, Q-x+y, Q+y: -(v(P+x) - v(P-x)) A8
    acc += tp(4, -1, 0, 0) * (-(tp(1, 0, 0, 0) - tp(1, -2, 0, 0)) * A8);
    acc += tp(4, 0, 0, 0) * (-(tp(1, 1, 0, 0) - tp(1, -1, 0, 0)) * A8);
    acc += tp(4, -1, 1, 0) * (-(tp(1, 0, 1, 0) - tp(1, -2, 1, 0)) * A8);
    acc += tp(4, 0, 1, 0) * (-(tp(1, 1, 1, 0) - tp(1, -1, 1, 0)) * A8);
    // from the w equation at P = Q-x, Q, Q-x+z, Q+z: -(w(P+x) - w(P-x)) A8
    acc += tp(5, -1, 0, 0) * (-(tp(2, 0, 0, 0) - tp(2, -2, 0, 0)) * A8);
    acc += tp(5, 0, 0, 0) * (-(tp(2, 1, 0, 0) - tp(2, -1, 0, 0)) * A8);
    acc += tp(5, -1, 0, 1) * (-(tp(2, 0, 0, 1) - tp(2, -2, 0, 1)) * A8);
    acc += tp(5, 0, 0, 1) * (-(tp(2, 1, 0, 1) - tp(2, -1, 0, 1)) * A8);
    XB0[o] = (T)acc;
  }
  // v at Q
  if (in_field(g, 1, i, j, k)) {
    double acc = zt(1, 0, 0, 0);
    // from the u equation at P = Q-y, Q, Q+x-y, Q+x: -(u(P+y) - u(P-y)) B8
    acc += tp(3, 0, -1, 0) * (-(tp(0, 0, 0, 0) - tp(0, 0, -2, 0)) * B8);
    acc += tp(3, 0, 0, 0) * (-(tp(0, 0, 1, 0) - tp(0, 0, -1, 0)) * B8);
    acc += tp(3, 1, -1, 0) * (-(tp(0, 1, 0, 0) - tp(0, 1, -2, 0)) * B8);
    acc += tp(3, 1, 0, 0) * (-(tp(0, 1, 1, 0) - tp(0, 1, -1, 0)) * B8);
    acc += tp(4, 0, 0, 0) * (-(tp(1, 0, 1, 0) - tp(1, 0, -1, 0)) * B2 - 2 * Dsum);
    // Us_v(P) = u(P+x) + u(P) + u(P+x-y) + u(P-y), at P = Q -+ x
    const double usm = tp(0, 0, 0, 0) + tp(0, -1, 0, 0) + tp(0, 0, -1, 0) + tp(0, -1, -1, 0);
    const double usp = tp(0, 2, 0, 0) + tp(0, 1, 0, 0) + tp(0, 2, -1, 0) + tp(0, 1, -1, 0);
    acc += tp(4, -1, 0, 0) * (-usm * A8 + Dx);
    acc += tp(4, 1, 0, 0) * (usp * A8 + Dx);
    acc += tp(4, 0, -1, 0) * (-tp(1, 0, -1, 0) * B2 + Dy);
    acc += tp(4, 0, 1, 0) * (tp(1, 0, 1, 0) * B2 + Dy);
    // Ws_v(P) = w(P+z) + w(P) + w(P-y+z) + w(P-y), at P = Q -+ z
    const double wsm = tp(2, 0, 0, 0) + tp(2, 0, 0, -1) + tp(2, 0, -1, 0) + tp(2, 0, -1, -1);
    const double wsp = tp(2, 0, 0, 2) + tp(2, 0, 0, 1) + tp(2, 0, -1, 2) + tp(2, 0, -1, 1);
    acc += tp(4, 0, 0, -1) * (-wsm * C8 + Dz);
    acc += tp(4, 0, 0, 1) * (wsp * C8 + Dz);
    // from the w equation at P = Q-y, Q, Q-y+z, Q+z: -(w(P+y) - w(P-y)) B8
    acc += tp(5, 0, -1, 0) * (-(tp(2, 0, 0, 0) - tp(2, 0, -2, 0)) * B8);
    acc += tp(5, 0, 0, 0) * (-(tp(2, 0, 1, 0) - tp(2, 0, -1, 0)) * B8);
    acc += tp(5, 0, -1, 1) * (-(tp(2, 0, 0, 1) - tp(2, 0, -2, 1)) * B8);
    acc += tp(5, 0, 0, 1) * (-(tp(2, 0, 1, 1) - tp(2, 0, -1, 1)) * B8);
    XB1[o] = (T)acc;
  }
  // w at Q
  if (in_field(g, 2, i, j, k)) {
    double acc = zt(2, 0, 0, 0);
    // from the u equation at P = Q-z, Q, Q+x-z, Q+x: -(u(P+z) - u(P-z)) C8
    acc += tp(3, 0, 0, -1) * (-(tp(0, 0, 0, 0) - tp(0, 0, 0, -2)) * C8);
    acc += tp(3, 0, 0, 0) * (-(tp(0, 0, 0, 1) - tp(0, 0, 0, -1)) * C8);
    acc += tp(3, 1, 0, -1) * (-(tp(0, 1, 0, 0) - tp(0, 1, 0, -2)) * C8);
    acc += tp(3, 1, 0, 0) * (-(tp(0, 1, 0, 1) - tp(0, 1, 0, -1)) * C8);
    // from the v equation at P = Q-z, Q, Q+y-z, Q+y: -(v(P+z) - v(P-z)) C8
    acc += tp(4, 0, 0, -1) * (-(tp(1, 0, 0, 0) - tp(1, 0, 0, -2)) * C8);
    acc += tp(4, 0, 0, 0) * (-(tp(1, 0, 0, 1) - tp(1, 0, 0, -1)) * C8);
    acc += tp(4, 0, 1, -1) * (-(tp(1, 0, 1, 0) - tp(1, 0, 1, -2)) * C8);
    acc += tp(4, 0, 1, 0) * (-(tp(1, 0, 1, 1) - tp(1, 0, 1, -1)) * C8);
    acc += tp(5, 0, 0, 0) * (-(tp(2, 0, 0, 1) - tp(2, 0, 0, -1)) * C2 - 2 * Dsum);
    // Us_w(P) = u(P+x) + u(P) + u(P+x-z) + u(P-z), at P = Q -+ x
    const double usm = tp(0, 0, 0, 0) + tp(0, -1, 0, 0) + tp(0, 0, 0, -1) + tp(0, -1, 0, -1);
    const double usp = tp(0, 2, 0, 0) + tp(0, 1, 0, 0) + tp(0, 2, 0, -1) + tp(0, 1, 0, -1);
    acc += tp(5, -1, 0, 0) * (-usm * A8 + Dx);
    acc += tp(5, 1, 0, 0) * (usp * A8 + Dx);
    // Vs_w(P) = v(P+y) + v(P) + v(P+y-z) + v(P-z), at P = Q -+ y
    const double vsm = tp(1, 0, 0, 0) + tp(1, 0, -1, 0) + tp(1, 0, 0, -1) + tp(1, 0, -1, -1);
    const double vsp = tp(1, 0, 2, 0) + tp(1, 0, 1, 0) + tp(1, 0, 2, -1) + tp(1, 0, 1, -1);
    acc += tp(5, 0, -1, 0) * (-vsm * B8 + Dy);
    acc += tp(5, 0, 1, 0) * (vsp * B8 + Dy);
    acc += tp(5, 0, 0, -1) * (-tp(2, 0, 0, -1) * C2 + Dz);
    acc += tp(5, 0, 0, 1) * (tp(2, 0, 0, 1) * C2 + Dz);
    XB2[o] = (T)acc;
  }
  // p_bar += b G^T z_bar at pressure points
  if (i < g.N[0] && j < g.N[1] && k < g.N[2]) {
    double acc = 0;
    acc += (zt(0, 1, 0, 0) - zt(0, 0, 0, 0)) * g.one_over_h[0];
    acc += (zt(1, 0, 1, 0) - zt(1, 0, 0, 0)) * g.one_over_h[1];
    acc += (zt(2, 0, 0, 1) - zt(2, 0, 0, 0)) * g.one_over_h[2];
    P[o] = (T)((double)P[o] + bcoef * acc);
  }
}

// A 64 x 4 workgroup owns an (i, j) tile of the union of the component
// extents and marches KB planes along k.  Planes k-1, k, k+1 of X and N_bar
// (one-point halo in i and j) live in an LDS ring; plane k+2 is fetched
// into registers while plane k is computed.  The few reach-2 taps of X and
// the z_bar, p_bar accesses go to global memory.  Ring cells outside the
// array hold zero: wherever N_bar is non-zero (component interiors) every X
// tap of its term is an array point, so those zeros only meet N_bar = 0.
// (256^3 FP64, 8 adjoint steps: 64 x 4 tiles of 4 planes 54.9 ms, 8 planes
// 53.9, 16 planes 53.7; 64 x 8 tiles 55.8 (4 planes) and 55.1 (8 planes):
// the kernel is not bound by its halo loads.  Every tap read from global
// memory instead of the ring measured slower.)
constexpr int AXT = 64, AYT = 4, AKB = 4;
constexpr int AHX = AXT + 2, AHY = AYT + 2;
constexpr int APL = AHY * AHX;                         // one field of a ring plane
constexpr int ARH = (APL + AXT * AYT - 1) / (AXT * AYT);  // elements per thread and field

template <typename T>
__global__ void __launch_bounds__(AXT * AYT) k_adj_xb(DGrid g, Layout L, AdjVec Xv, AdjVec Nbv, AdjVec Zv, AdjVec Xbv,
                                                T* __restrict__ P, double bcoef, int ntx, int nty, int ntiles) {
  __shared__ T ring[3][6][AHY][AHX];
  const int tx = threadIdx.x % AXT, ty = threadIdx.x / AXT, tid = threadIdx.x;
  const int tile = xcd_tile(blockIdx.x, ntiles);
  const int bx = tile % ntx, by = (tile / ntx) % nty, bz = tile / (ntx * nty);
  const int i0 = bx * AXT, j0 = by * AYT, k0 = bz * AKB;
  const int kend = min(k0 + AKB, g.N[2] + 1);
  const int i = i0 + tx, j = j0 + ty;
  const T* F[6] = {(const T*)Xv.f[0], (const T*)Xv.f[1], (const T*)Xv.f[2],
                   (const T*)Nbv.f[0], (const T*)Nbv.f[1], (const T*)Nbv.f[2]};
  // this thread's elements of every field of a ring plane: the same (ii, jj)
  // positions in each of the six fields, so the field pointer is uniform
  int off[ARH], lds[ARH];
  bool ok[ARH];
#pragma unroll
  for (int h = 0; h < ARH; h++) {
    const int q = tid + AXT * AYT * h;
    const int jj = q / AHX, ii = q - jj * AHX;
    const int gi = i0 - 1 + ii, gj = j0 - 1 + jj;
    ok[h] = q < APL && gi >= 0 && gj >= 0 && gi < L.PX && gj < L.SY;
    off[h] = ok[h] ? (int)L.idx(gi, gj, 0) : 0;
    lds[h] = q < APL ? q : -1;
  }
  // planes outside the array are fetched at plane 0 and zeroed, like the
  // cells outside it, when committed (so no instruction waits for the loads
  // before the commit)
  auto fetch = [&](int k, T (&v)[6][ARH]) {
    const long kk = k >= 0 && k < L.SZ ? (long)k * L.plane : 0;
#pragma unroll
    for (int f = 0; f < 6; f++)
#pragma unroll
      for (int h = 0; h < ARH; h++) v[f][h] = F[f][kk + off[h]];
  };
  auto slot_of = [](int k) { return (k + 3) % 3; };
  auto commit = [&](int k, const T (&v)[6][ARH]) {
    const int sl = slot_of(k);
    const bool kin = k >= 0 && k < L.SZ;
#pragma unroll
    for (int f = 0; f < 6; f++)
#pragma unroll
      for (int h = 0; h < ARH; h++)
        if (lds[h] >= 0) (&ring[sl][f][0][0])[lds[h]] = ok[h] && kin ? v[f][h] : (T)0;
  };
  {
    T a[6][ARH];
    fetch(k0 - 1, a);
    commit(k0 - 1, a);
    fetch(k0, a);
    commit(k0, a);
    fetch(k0 + 1, a);
    commit(k0 + 1, a);
  }
  T pre[6][ARH];
  if (k0 + 2 <= kend) fetch(k0 + 2, pre);
  __syncthreads();

  const bool in_tile = i <= g.N[0] && j <= g.N[1];
  const long last = L.elems - 1;
  const long PX = L.PX, PL = L.plane;
  for (int k = k0; k < kend; k++) {
    if (in_tile) {
      const long o = L.idx(i, j, k);
      const int sl[3] = {slot_of(k - 1), slot_of(k), slot_of(k + 1)};
      auto tp = [&](int f, int di, int dj, int dk) -> double {
        if (di >= -1 && di <= 1 && dj >= -1 && dj <= 1 && dk >= -1 && dk <= 1)
          return (double)ring[sl[dk + 1]][f][ty + 1 + dj][tx + 1 + di];
        long idx = o + di + PX * dj + PL * dk;
        idx = idx < 0 ? 0 : (idx > last ? last : idx);
        return (double)F[f][idx];
      };
      auto zt = [&](int c, int di, int dj, int dk) -> double {
        long idx = o + di + PX * dj + PL * dk;
        idx = idx < 0 ? 0 : (idx > last ? last : idx);
        return (double)((const T*)Zv.f[c])[idx];
      };
      adj_xb_point<T>(g, o, i, j, k, tp, zt, (T*)Xbv.f[0], (T*)Xbv.f[1], (T*)Xbv.f[2], P, bcoef);
    }
    __syncthreads();
    // plane k+2 replaces plane k-1; plane k+3 starts loading
    if (k + 2 <= kend) commit(k + 2, pre);
    if (k + 3 <= kend) fetch(k + 3, pre);
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
template <typename T>
void launch_adj_rhs(const DGrid& g, const Layout& L, const T* P, const AdjVec& Y, double dts, T* B, long bp0,
                    long bplane, hipStream_t st) {
  const long total = (long)g.npts[0] * g.npts[1] * g.npts[2];
  k_adj_rhs<T><<<(unsigned)((total + 255) / 256), 256, 0, st>>>(g, L, P, Y, dts, B, bp0, bplane, total);
}

template <typename T>
void launch_adj_zb(const DGrid& g, const Layout& L, const AdjVec& Y, const T* Bpsi, long bp0, long bplane, double dts,
                   const AdjVec& Zb, const AdjVec& Nb, const AdjVec& Zo, double alpha, double beta, double fcoef,
                   int accumulate, double* partial,
                   int nblk, hipStream_t st) {
  k_adj_zb<T><<<dim3(nblk, 3, 1), 256, 0, st>>>(g, L, Y, Bpsi, bp0, bplane, dts, Zb, Nb, Zo, alpha, beta, fcoef,
                                                accumulate, partial);
}

void launch_adj_sum(const double* partial, int nblk, int accumulate, double* out, hipStream_t st) {
  k_adj_sum<<<1, kAdjSumThreads, 0, st>>>(partial, nblk, accumulate, out);
}

template <typename T>
void launch_adj_xb(const DGrid& g, const Layout& L, const AdjVec& X, const AdjVec& Nb, const AdjVec& Zb,
                   const AdjVec& Xb, T* P, double bcoef, hipStream_t st) {
  // tiles over points 0..N per axis
  const int ntx = (g.N[0] + 1 + AXT - 1) / AXT, nty = (g.N[1] + 1 + AYT - 1) / AYT;
  const int ntz = (g.N[2] + 1 + AKB - 1) / AKB;
  const int ntiles = ntx * nty * ntz;
  k_adj_xb<T><<<ntiles, AXT * AYT, 0, st>>>(g, L, X, Nb, Zb, Xb, P, bcoef, ntx, nty, ntiles);
}

#define MIF_ADJ_INST(T)                                                                                              \
  template void launch_adj_rhs<T>(const DGrid&, const Layout&, const T*, const AdjVec&, double, T*, long, long,       \
                                  hipStream_t);                                                                      \
  template void launch_adj_zb<T>(const DGrid&, const Layout&, const AdjVec&, const T*, long, long, double,            \
                                 const AdjVec&, const AdjVec&, const AdjVec&, double, double, double, int, double*,   \
                                 int, hipStream_t);                                                                  \
  template void launch_adj_xb<T>(const DGrid&, const Layout&, const AdjVec&, const AdjVec&, const AdjVec&,            \
                                 const AdjVec&, T*, double, hipStream_t);

MIF_ADJ_INST(double)
MIF_ADJ_INST(float)

}  // namespace mif
