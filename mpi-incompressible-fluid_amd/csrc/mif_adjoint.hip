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
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= total) return;
  const int a = (int)(q % g.npts[0]), b = (int)((q / g.npts[0]) % g.npts[1]), c = (int)(q / ((long)g.npts[0] * g.npts[1]));
  const int O[3] = {a + g.olo[0], b + g.olo[1], c + g.olo[2]};
  int opts[3][2], nopt[3];
  for (int ax = 0; ax < 3; ax++) nopt[ax] = images(g, ax, O[ax], g.N[ax], opts[ax]);
  const long off[3] = {1, L.PX, L.plane};
  double s = 0;
  for (int qa = 0; qa < nopt[0]; qa++)
    for (int qb = 0; qb < nopt[1]; qb++)
      for (int qc = 0; qc < nopt[2]; qc++) {
        const int R[3] = {opts[0][qa], opts[1][qb], opts[2][qc]};
        const long o = L.idx(R[0], R[1], R[2]);
        double acc = 0;
        for (int cc = 0; cc < 3; cc++) {
          const T* F = (const T*)Y.f[cc];
          int Rp[3] = {R[0], R[1], R[2]};
          Rp[cc] += 1;
          double d = 0;
          if (adj_interior(g, cc, Rp[0], Rp[1], Rp[2])) d += (double)F[o + off[cc]];
          if (adj_interior(g, cc, R[0], R[1], R[2])) d -= (double)F[o];
          acc += d * g.one_over_h[cc];
        }
        s += (double)P[o] + dts * acc;
      }
  const double w = trap_weight(g, 0, a) * trap_weight(g, 1, b) * trap_weight(g, 2, c);
  B[a + bp0 * b + bplane * c] = (T)(s / w);
}

// ---------------------------------------------------------------------------
// k_adj_zb: component c = blockIdx.y, interior points Q of c (grid-stride):
//   z_bar_c(Q) = sum over the points g holding Q (Q and its periodic ghost
//                copies) of y_bar_c(g) + (Wpsi(g - e_c) - Wpsi(g)) / (dt_s h_c)
//   N_bar_c(Q) = alpha z_bar_c(Q) + beta Zo_c(Q)
// with Wpsi = W * psi on owner nodes (zero elsewhere), and one partial sum
// of z_bar_c per workgroup.
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256) k_adj_zb(DGrid g, Layout L, AdjVec Y, const T* __restrict__ Bpsi, long bp0,
                                                long bplane, double dts, AdjVec Zb, AdjVec Nb, AdjVec Zo, double alpha,
                                                double beta, double* __restrict__ partial) {
  const int c = blockIdx.y;
  int e[3];
  extents(g, c, e);
  const long n0 = e[0] - 2, n1 = e[1] - 2, n2 = e[2] - 2;
  const long total = (n0 > 0 && n1 > 0 && n2 > 0) ? n0 * n1 * n2 : 0;
  const T* y = (const T*)Y.f[c];
  T* zb = (T*)Zb.f[c];
  T* nb = (T*)Nb.f[c];
  const T* zo = (const T*)Zo.f[c];
  auto wpsi = [&](int pi, int pj, int pk) -> double {
    const int qq[3] = {pi - g.olo[0], pj - g.olo[1], pk - g.olo[2]};
    for (int ax = 0; ax < 3; ax++)
      if (qq[ax] < 0 || qq[ax] >= g.npts[ax]) return 0.0;
    const double w = trap_weight(g, 0, qq[0]) * trap_weight(g, 1, qq[1]) * trap_weight(g, 2, qq[2]);
    return w * (double)Bpsi[qq[0] + bp0 * qq[1] + bplane * qq[2]];
  };
  double part = 0;
  for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < total; q += (long)gridDim.x * blockDim.x) {
    const int Q[3] = {(int)(q % n0) + 1, (int)((q / n0) % n1) + 1, (int)(q / (n0 * n1)) + 1};
    int opts[3][2], nopt[3];
    for (int ax = 0; ax < 3; ax++) nopt[ax] = images(g, ax, Q[ax], e[ax], opts[ax]);
    double s = 0;
    for (int qa = 0; qa < nopt[0]; qa++)
      for (int qb = 0; qb < nopt[1]; qb++)
        for (int qc = 0; qc < nopt[2]; qc++) {
          const int G[3] = {opts[0][qa], opts[1][qb], opts[2][qc]};
          int Gm[3] = {G[0], G[1], G[2]};
          Gm[c] -= 1;
          const double d = (wpsi(Gm[0], Gm[1], Gm[2]) - wpsi(G[0], G[1], G[2])) * g.one_over_h[c] / dts;
          s += (double)(T)((double)y[L.idx(G[0], G[1], G[2])] + d);
        }
    const long o = L.idx(Q[0], Q[1], Q[2]);
    const T z = (T)s;
    zb[o] = z;
    nb[o] = (T)(alpha * (double)z + (beta != 0.0 ? beta * (double)zo[o] : 0.0));
    part += (double)z;
  }
  __shared__ double sh[4];
  for (int off = 32; off > 0; off >>= 1) part += __shfl_down(part, off, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = part;
  __syncthreads();
  if (threadIdx.x == 0) partial[c * gridDim.x + blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
}

// out[c] += coef * (sum of the nblk partials of component c), one workgroup,
// fixed summation order.
__global__ void __launch_bounds__(256) k_adj_sum(const double* __restrict__ partial, int nblk, double coef,
                                                 double* __restrict__ out) {
  __shared__ double sh[256];
  for (int c = 0; c < 3; c++) {
    double s = 0;
    for (int q = threadIdx.x; q < nblk; q += blockDim.x) s += partial[c * nblk + q];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int stp = blockDim.x / 2; stp > 0; stp >>= 1) {
      if (threadIdx.x < stp) sh[threadIdx.x] += sh[threadIdx.x + stp];
      __syncthreads();
    }
    if (threadIdx.x == 0) out[c] += coef * sh[0];
    __syncthreads();
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
template <typename T>
__global__ void __launch_bounds__(256) k_adj_xb(DGrid g, Layout L, AdjVec Xv, AdjVec Nbv, AdjVec Zv, AdjVec Xbv,
                                                T* __restrict__ P, double bcoef, int nqt) {
  const int qt = blockIdx.x % nqt, k = blockIdx.x / nqt;
  const int q = qt * 256 + threadIdx.x;
  const int i = q % L.PX, j = q / L.PX;
  if (j > g.N[1] || i > g.N[0] || k > g.N[2]) return;
  const long o = (long)q + L.plane * k;
  const long last = L.elems - 1;
  const long PX = L.PX, PL = L.plane;
  auto at = [&](const void* F, long d) -> double {
    long idx = o + d;
    idx = idx < 0 ? 0 : (idx > last ? last : idx);
    return (double)((const T*)F)[idx];
  };
  const void* X0 = Xv.f[0];
  const void* X1 = Xv.f[1];
  const void* X2 = Xv.f[2];
  const void* N0 = Nbv.f[0];
  const void* N1 = Nbv.f[1];
  const void* N2 = Nbv.f[2];
  const double A2 = g.one_over_2h[0], B2 = g.one_over_2h[1], C2 = g.one_over_2h[2];
  const double A8 = g.one_over_8h[0], B8 = g.one_over_8h[1], C8 = g.one_over_8h[2];
  const double Dx = g.one_over_h2Re[0], Dy = g.one_over_h2Re[1], Dz = g.one_over_h2Re[2];
  const double Dsum = Dx + Dy + Dz;
  // offsets of (di, dj, dk) from Q
  auto D = [&](int di, int dj, int dk) -> long { return (long)di + PX * dj + PL * dk; };

  // u at Q
  if (in_field(g, 0, i, j, k)) {
    double acc = at(Zv.f[0], 0);
    acc += at(N0, 0) * (-(at(X0, D(1, 0, 0)) - at(X0, D(-1, 0, 0))) * A2 - 2 * Dsum);
    acc += at(N0, D(-1, 0, 0)) * (-at(X0, D(-1, 0, 0)) * A2 + Dx);
    acc += at(N0, D(1, 0, 0)) * (at(X0, D(1, 0, 0)) * A2 + Dx);
    // Vs_u(P) = v(P+y) + v(P) + v(P-x+y) + v(P-x), at P = Q -+ y
    const double vsm = at(X1, D(0, 0, 0)) + at(X1, D(0, -1, 0)) + at(X1, D(-1, 0, 0)) + at(X1, D(-1, -1, 0));
    const double vsp = at(X1, D(0, 2, 0)) + at(X1, D(0, 1, 0)) + at(X1, D(-1, 2, 0)) + at(X1, D(-1, 1, 0));
    acc += at(N0, D(0, -1, 0)) * (-vsm * B8 + Dy);
    acc += at(N0, D(0, 1, 0)) * (vsp * B8 + Dy);
    // Ws_u(P) = w(P+z) + w(P) + w(P-x+z) + w(P-x), at P = Q -+ z
    const double wsm = at(X2, D(0, 0, 0)) + at(X2, D(0, 0, -1)) + at(X2, D(-1, 0, 0)) + at(X2, D(-1, 0, -1));
    const double wsp = at(X2, D(0, 0, 2)) + at(X2, D(0, 0, 1)) + at(X2, D(-1, 0, 2)) + at(X2, D(-1, 0, 1));
    acc += at(N0, D(0, 0, -1)) * (-wsm * C8 + Dz);
    acc += at(N0, D(0, 0, 1)) * (wsp * C8 + Dz);
    // from the v equation at P = Q-x, Q, Q-x+y, Q+y: -(v(P+x) - v(P-x)) A8
    acc += at(N1, D(-1, 0, 0)) * (-(at(X1, D(0, 0, 0)) - at(X1, D(-2, 0, 0))) * A8);
    acc += at(N1, D(0, 0, 0)) * (-(at(X1, D(1, 0, 0)) - at(X1, D(-1, 0, 0))) * A8);
    acc += at(N1, D(-1, 1, 0)) * (-(at(X1, D(0, 1, 0)) - at(X1, D(-2, 1, 0))) * A8);
    acc += at(N1, D(0, 1, 0)) * (-(at(X1, D(1, 1, 0)) - at(X1, D(-1, 1, 0))) * A8);
    // from the w equation at P = Q-x, Q, Q-x+z, Q+z: -(w(P+x) - w(P-x)) A8
    acc += at(N2, D(-1, 0, 0)) * (-(at(X2, D(0, 0, 0)) - at(X2, D(-2, 0, 0))) * A8);
    acc += at(N2, D(0, 0, 0)) * (-(at(X2, D(1, 0, 0)) - at(X2, D(-1, 0, 0))) * A8);
    acc += at(N2, D(-1, 0, 1)) * (-(at(X2, D(0, 0, 1)) - at(X2, D(-2, 0, 1))) * A8);
    acc += at(N2, D(0, 0, 1)) * (-(at(X2, D(1, 0, 1)) - at(X2, D(-1, 0, 1))) * A8);
    ((T*)Xbv.f[0])[o] = (T)acc;
  }
  // v at Q
  if (in_field(g, 1, i, j, k)) {
    double acc = at(Zv.f[1], 0);
    // from the u equation at P = Q-y, Q, Q+x-y, Q+x: -(u(P+y) - u(P-y)) B8
    acc += at(N0, D(0, -1, 0)) * (-(at(X0, D(0, 0, 0)) - at(X0, D(0, -2, 0))) * B8);
    acc += at(N0, D(0, 0, 0)) * (-(at(X0, D(0, 1, 0)) - at(X0, D(0, -1, 0))) * B8);
    acc += at(N0, D(1, -1, 0)) * (-(at(X0, D(1, 0, 0)) - at(X0, D(1, -2, 0))) * B8);
    acc += at(N0, D(1, 0, 0)) * (-(at(X0, D(1, 1, 0)) - at(X0, D(1, -1, 0))) * B8);
    acc += at(N1, 0) * (-(at(X1, D(0, 1, 0)) - at(X1, D(0, -1, 0))) * B2 - 2 * Dsum);
    // Us_v(P) = u(P+x) + u(P) + u(P+x-y) + u(P-y), at P = Q -+ x
    const double usm = at(X0, D(0, 0, 0)) + at(X0, D(-1, 0, 0)) + at(X0, D(0, -1, 0)) + at(X0, D(-1, -1, 0));
    const double usp = at(X0, D(2, 0, 0)) + at(X0, D(1, 0, 0)) + at(X0, D(2, -1, 0)) + at(X0, D(1, -1, 0));
    acc += at(N1, D(-1, 0, 0)) * (-usm * A8 + Dx);
    acc += at(N1, D(1, 0, 0)) * (usp * A8 + Dx);
    acc += at(N1, D(0, -1, 0)) * (-at(X1, D(0, -1, 0)) * B2 + Dy);
    acc += at(N1, D(0, 1, 0)) * (at(X1, D(0, 1, 0)) * B2 + Dy);
    // Ws_v(P) = w(P+z) + w(P) + w(P-y+z) + w(P-y), at P = Q -+ z
    const double wsm = at(X2, D(0, 0, 0)) + at(X2, D(0, 0, -1)) + at(X2, D(0, -1, 0)) + at(X2, D(0, -1, -1));
    const double wsp = at(X2, D(0, 0, 2)) + at(X2, D(0, 0, 1)) + at(X2, D(0, -1, 2)) + at(X2, D(0, -1, 1));
    acc += at(N1, D(0, 0, -1)) * (-wsm * C8 + Dz);
    acc += at(N1, D(0, 0, 1)) * (wsp * C8 + Dz);
    // from the w equation at P = Q-y, Q, Q-y+z, Q+z: -(w(P+y) - w(P-y)) B8
    acc += at(N2, D(0, -1, 0)) * (-(at(X2, D(0, 0, 0)) - at(X2, D(0, -2, 0))) * B8);
    acc += at(N2, D(0, 0, 0)) * (-(at(X2, D(0, 1, 0)) - at(X2, D(0, -1, 0))) * B8);
    acc += at(N2, D(0, -1, 1)) * (-(at(X2, D(0, 0, 1)) - at(X2, D(0, -2, 1))) * B8);
    acc += at(N2, D(0, 0, 1)) * (-(at(X2, D(0, 1, 1)) - at(X2, D(0, -1, 1))) * B8);
    ((T*)Xbv.f[1])[o] = (T)acc;
  }
  // w at Q
  if (in_field(g, 2, i, j, k)) {
    double acc = at(Zv.f[2], 0);
    // from the u equation at P = Q-z, Q, Q+x-z, Q+x: -(u(P+z) - u(P-z)) C8
    acc += at(N0, D(0, 0, -1)) * (-(at(X0, D(0, 0, 0)) - at(X0, D(0, 0, -2))) * C8);
    acc += at(N0, D(0, 0, 0)) * (-(at(X0, D(0, 0, 1)) - at(X0, D(0, 0, -1))) * C8);
    acc += at(N0, D(1, 0, -1)) * (-(at(X0, D(1, 0, 0)) - at(X0, D(1, 0, -2))) * C8);
    acc += at(N0, D(1, 0, 0)) * (-(at(X0, D(1, 0, 1)) - at(X0, D(1, 0, -1))) * C8);
    // from the v equation at P = Q-z, Q, Q+y-z, Q+y: -(v(P+z) - v(P-z)) C8
    acc += at(N1, D(0, 0, -1)) * (-(at(X1, D(0, 0, 0)) - at(X1, D(0, 0, -2))) * C8);
    acc += at(N1, D(0, 0, 0)) * (-(at(X1, D(0, 0, 1)) - at(X1, D(0, 0, -1))) * C8);
    acc += at(N1, D(0, 1, -1)) * (-(at(X1, D(0, 1, 0)) - at(X1, D(0, 1, -2))) * C8);
    acc += at(N1, D(0, 1, 0)) * (-(at(X1, D(0, 1, 1)) - at(X1, D(0, 1, -1))) * C8);
    acc += at(N2, 0) * (-(at(X2, D(0, 0, 1)) - at(X2, D(0, 0, -1))) * C2 - 2 * Dsum);
    // Us_w(P) = u(P+x) + u(P) + u(P+x-z) + u(P-z), at P = Q -+ x
    const double usm = at(X0, D(0, 0, 0)) + at(X0, D(-1, 0, 0)) + at(X0, D(0, 0, -1)) + at(X0, D(-1, 0, -1));
    const double usp = at(X0, D(2, 0, 0)) + at(X0, D(1, 0, 0)) + at(X0, D(2, 0, -1)) + at(X0, D(1, 0, -1));
    acc += at(N2, D(-1, 0, 0)) * (-usm * A8 + Dx);
    acc += at(N2, D(1, 0, 0)) * (usp * A8 + Dx);
    // Vs_w(P) = v(P+y) + v(P) + v(P+y-z) + v(P-z), at P = Q -+ y
    const double vsm = at(X1, D(0, 0, 0)) + at(X1, D(0, -1, 0)) + at(X1, D(0, 0, -1)) + at(X1, D(0, -1, -1));
    const double vsp = at(X1, D(0, 2, 0)) + at(X1, D(0, 1, 0)) + at(X1, D(0, 2, -1)) + at(X1, D(0, 1, -1));
    acc += at(N2, D(0, -1, 0)) * (-vsm * B8 + Dy);
    acc += at(N2, D(0, 1, 0)) * (vsp * B8 + Dy);
    acc += at(N2, D(0, 0, -1)) * (-at(X2, D(0, 0, -1)) * C2 + Dz);
    acc += at(N2, D(0, 0, 1)) * (at(X2, D(0, 0, 1)) * C2 + Dz);
    ((T*)Xbv.f[2])[o] = (T)acc;
  }
  // p_bar += b G^T z_bar at pressure points
  if (i < g.N[0] && j < g.N[1] && k < g.N[2]) {
    double acc = 0;
    acc += (at(Zv.f[0], D(1, 0, 0)) - at(Zv.f[0], 0)) * g.one_over_h[0];
    acc += (at(Zv.f[1], D(0, 1, 0)) - at(Zv.f[1], 0)) * g.one_over_h[1];
    acc += (at(Zv.f[2], D(0, 0, 1)) - at(Zv.f[2], 0)) * g.one_over_h[2];
    P[o] = (T)((double)P[o] + bcoef * acc);
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
                   const AdjVec& Zb, const AdjVec& Nb, const AdjVec& Zo, double alpha, double beta, double* partial,
                   int nblk, hipStream_t st) {
  k_adj_zb<T><<<dim3(nblk, 3, 1), 256, 0, st>>>(g, L, Y, Bpsi, bp0, bplane, dts, Zb, Nb, Zo, alpha, beta, partial);
}

void launch_adj_sum(const double* partial, int nblk, double coef, double* out, hipStream_t st) {
  k_adj_sum<<<1, 256, 0, st>>>(partial, nblk, coef, out);
}

template <typename T>
void launch_adj_xb(const DGrid& g, const Layout& L, const AdjVec& X, const AdjVec& Nb, const AdjVec& Zb,
                   const AdjVec& Xb, T* P, double bcoef, hipStream_t st) {
  const int nqt = (int)((L.plane + 255) / 256);
  k_adj_xb<T><<<(unsigned)(nqt * (g.N[2] + 1)), 256, 0, st>>>(g, L, X, Nb, Zb, Xb, P, bcoef, nqt);
}

#define MIF_ADJ_INST(T)                                                                                              \
  template void launch_adj_rhs<T>(const DGrid&, const Layout&, const T*, const AdjVec&, double, T*, long, long,       \
                                  hipStream_t);                                                                      \
  template void launch_adj_zb<T>(const DGrid&, const Layout&, const AdjVec&, const T*, long, long, double,            \
                                 const AdjVec&, const AdjVec&, const AdjVec&, double, double, double*, int,           \
                                 hipStream_t);                                                                       \
  template void launch_adj_xb<T>(const DGrid&, const Layout&, const AdjVec&, const AdjVec&, const AdjVec&,            \
                                 const AdjVec&, T*, double, hipStream_t);

MIF_ADJ_INST(double)
MIF_ADJ_INST(float)

}  // namespace mif
