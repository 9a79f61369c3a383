// mif_adjoint.hip -- adjoint of one RK3 + projection step (no counterpart in
// the reference, which has no adjoint).
//
// The forward step (src/Timestep.cpp:97-141) is, per stage: a stencil update
// with the momentum right-hand side N (quadratic in the velocity,
// include/MomentumEquation.h:43-255) and the pressure gradient; the boundary
// map (Dirichlet faces are constants, periodic ghosts copy interior points,
// src/VelocityTensor.cpp:36-233); the divergence, the spectral solve M, the
// unpack with periodic images; p += phi; u -= dt_s grad(phi) on interior
// points.  Every kernel below is the transpose of one of those linear maps,
// written as a gather (each output point sums the contributions that the
// forward scatters into it), so no atomics are needed and results are
// deterministic:
//   k_adj_grad      G^T (stage pressure term and projection)
//   k_adj_unpack    U^T then W^-1   (M^T = W M W^-1, W = trapezoid weights)
//   k_adj_div       W, then D^T / dt_s
//   k_adj_bc        B^T (ghost adjoints return to their interior source)
//   k_adj_nbar      adjoint of N's output per stage (+ body-force adjoint)
//   k_adj_convdiff  J_N(X)^T, the transposed linearised convection plus the
//                   (self-adjoint on its interior) diffusion stencil
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

// out(R) = in(R) + coef * sum_c (Ob_c(R + e_c) - Ob_c(R)) / h_c over the
// interior points of each component; R runs over the pressure extents.
template <typename T>
__global__ void __launch_bounds__(256) k_adj_grad(DGrid g, Layout L, const T* in, T* out, AdjVec Ob, double coef) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y, k = blockIdx.z;
  if (i >= g.N[0]) return;
  const long o = L.idx(i, j, k);
  const long off[3] = {1, L.PX, L.plane};
  const int R[3] = {i, j, k};
  double acc = 0;
  for (int c = 0; c < 3; c++) {
    const T* F = (const T*)Ob.f[c];
    double s = 0;
    int Rp[3] = {R[0], R[1], R[2]};
    Rp[c] += 1;
    if (adj_interior(g, c, Rp[0], Rp[1], Rp[2])) s += (double)F[o + off[c]];
    if (adj_interior(g, c, i, j, k)) s -= (double)F[o];
    acc += s * g.one_over_h[c];
  }
  out[o] = (T)((double)in[o] + coef * acc);
}

__device__ inline double trap_weight(const DGrid& g, int a, int i) {
  if (g.periodic[a]) return 1.0;
  return (i == 0 || i == g.npts[a] - 1) ? 0.5 : 1.0;
}

// B(o) = (sum over the pressure points unpacked from owner o of phib) / W(o)
template <typename T>
__global__ void __launch_bounds__(256) k_adj_unpack(DGrid g, Layout L, const T* __restrict__ phib, T* __restrict__ B,
                                                    long bp0, long bplane) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y, c = blockIdx.z;
  if (a >= g.npts[0]) return;
  const int P[3] = {a + g.olo[0], b + g.olo[1], c + g.olo[2]};
  int opts[3][3], nopt[3];
  for (int ax = 0; ax < 3; ax++) {
    nopt[ax] = 1;
    opts[ax][0] = P[ax];
    if (g.periodic[ax]) {
      if (P[ax] == 1) opts[ax][nopt[ax]++] = g.N[ax] - 1;
      if (P[ax] == g.N[ax] - 2) opts[ax][nopt[ax]++] = 0;
    }
  }
  double s = 0;
  for (int qa = 0; qa < nopt[0]; qa++)
    for (int qb = 0; qb < nopt[1]; qb++)
      for (int qc = 0; qc < nopt[2]; qc++) s += (double)phib[L.idx(opts[0][qa], opts[1][qb], opts[2][qc])];
  const double w = trap_weight(g, 0, a) * trap_weight(g, 1, b) * trap_weight(g, 2, c);
  B[a + bp0 * b + bplane * c] = (T)(s / w);
}

// yb_c(Q) += (Wb(Q - e_c) - Wb(Q)) / (dts h_c), Wb = W * B on owner nodes.
template <typename T>
__global__ void __launch_bounds__(256) k_adj_div(DGrid g, Layout L, const T* __restrict__ B, long bp0, long bplane,
                                                 AdjVec yb, double dts) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y, k = blockIdx.z;
  if (i > g.N[0]) return;
  const long o = L.idx(i, j, k);
  auto wb = [&](int pi, int pj, int pk) -> double {
    const int q[3] = {pi - g.olo[0], pj - g.olo[1], pk - g.olo[2]};
    for (int ax = 0; ax < 3; ax++)
      if (q[ax] < 0 || q[ax] >= g.npts[ax]) return 0.0;
    const double w = trap_weight(g, 0, q[0]) * trap_weight(g, 1, q[1]) * trap_weight(g, 2, q[2]);
    return w * (double)B[q[0] + bp0 * q[1] + bplane * q[2]];
  };
  const int Q[3] = {i, j, k};
  for (int c = 0; c < 3; c++) {
    if (!in_field(g, c, i, j, k)) continue;
    int Pm[3] = {Q[0], Q[1], Q[2]};
    Pm[c] -= 1;
    const double d = (wb(Pm[0], Pm[1], Pm[2]) - wb(i, j, k)) * g.one_over_h[c] / dts;
    T* F = (T*)yb.f[c];
    F[o] = (T)((double)F[o] + d);
  }
}

// Yb = B^T yb over all points of the three components: an interior point
// collects itself and the periodic ghosts copied from it; the shell is 0.
template <typename T>
__global__ void __launch_bounds__(256) k_adj_bc(DGrid g, Layout L, AdjVec yb, AdjVec Yb) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y, k = blockIdx.z;
  for (int c = 0; c < 3; c++) {
    if (!in_field(g, c, i, j, k)) continue;
    int e[3];
    extents(g, c, e);
    double s = 0;
    if (adj_interior(g, c, i, j, k)) {
      const int P[3] = {i, j, k};
      int opts[3][3], nopt[3];
      for (int ax = 0; ax < 3; ax++) {
        nopt[ax] = 1;
        opts[ax][0] = P[ax];
        if (g.periodic[ax]) {
          if (P[ax] == 1) opts[ax][nopt[ax]++] = e[ax] - 1;
          if (P[ax] == e[ax] - 2) opts[ax][nopt[ax]++] = 0;
        }
      }
      const T* y = (const T*)yb.f[c];
      for (int qa = 0; qa < nopt[0]; qa++)
        for (int qb = 0; qb < nopt[1]; qb++)
          for (int qc = 0; qc < nopt[2]; qc++) s += (double)y[L.idx(opts[0][qa], opts[1][qb], opts[2][qc])];
    }
    ((T*)Yb.f[c])[L.idx(i, j, k)] = (T)s;
  }
}

// Adjoint of N's output at interior points, per stage:
//   stage 3: Nb = a3 Y,          Sb = -Y
//   stage 2: Nb = a2 (Y + Sb),   R1b = a1 Y
//   stage 1: Nb = a1 Y + R1b
// and per-block partial sums of Nb per component (the body-force adjoint).
template <typename T, int STAGE>
__global__ void __launch_bounds__(256) k_adj_nbar(DGrid g, Layout L, AdjVec Y, AdjVec Nb, AdjVec Sb, AdjVec R1b,
                                                  double a1, double a2, double a3, double* partial) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y, k = blockIdx.z;
  double s[3] = {0, 0, 0};
  for (int c = 0; c < 3; c++) {
    if (!adj_interior(g, c, i, j, k)) continue;
    const long o = L.idx(i, j, k);
    const double y = (double)((const T*)Y.f[c])[o];
    double n;
    if (STAGE == 3) {
      n = a3 * y;
      ((T*)Sb.f[c])[o] = (T)(-y);
    } else if (STAGE == 2) {
      n = a2 * (y + (double)((const T*)Sb.f[c])[o]);
      ((T*)R1b.f[c])[o] = (T)(a1 * y);
    } else {
      n = a1 * y + (double)((const T*)R1b.f[c])[o];
    }
    ((T*)Nb.f[c])[o] = (T)n;
    s[c] = n;
  }
  __shared__ double sh[3][4];
  for (int c = 0; c < 3; c++) {
    double v = s[c];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if ((threadIdx.x & 63) == 0) sh[c][threadIdx.x >> 6] = v;
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    double v = 0;
    for (int q = 0; q < (int)(blockDim.x + 63) / 64; q++) v += sh[threadIdx.x][q];
    const long blk = (long)blockIdx.x + (long)gridDim.x * ((long)blockIdx.y + (long)gridDim.y * blockIdx.z);
    partial[3 * blk + threadIdx.x] = v;
  }
}

// Xb(Q) = id(Q) [interior] + (J_N(X)^T Nb)(Q), for every point Q of the
// three components.  Nb is zero outside the interiors.
template <typename T>
__global__ void __launch_bounds__(256) k_adj_convdiff(DGrid g, Layout L, AdjVec Xv, AdjVec Nbv, AdjVec idv,
                                                      AdjVec Xbv) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y, k = blockIdx.z;
  const T* X[3] = {(const T*)Xv.f[0], (const T*)Xv.f[1], (const T*)Xv.f[2]};
  const T* Nb[3] = {(const T*)Nbv.f[0], (const T*)Nbv.f[1], (const T*)Nbv.f[2]};
  const double A2 = g.one_over_2h[0], B2 = g.one_over_2h[1], C2 = g.one_over_2h[2];
  const double A8 = g.one_over_8h[0], B8 = g.one_over_8h[1], C8 = g.one_over_8h[2];
  const double Dx = g.one_over_h2Re[0], Dy = g.one_over_h2Re[1], Dz = g.one_over_h2Re[2];
  const double Dsum = Dx + Dy + Dz;
  auto x = [&](int f, int a, int b, int c) -> double { return (double)X[f][L.idx(a, b, c)]; };
  auto nb = [&](int f, int a, int b, int c) -> double {
    return adj_interior(g, f, a, b, c) ? (double)Nb[f][L.idx(a, b, c)] : 0.0;
  };
  // coefficients of the u, v, w outputs at P (valid for interior P)
  auto Vs_u = [&](int a, int b, int c) { return x(1, a, b + 1, c) + x(1, a, b, c) + x(1, a - 1, b + 1, c) + x(1, a - 1, b, c); };
  auto Ws_u = [&](int a, int b, int c) { return x(2, a, b, c + 1) + x(2, a, b, c) + x(2, a - 1, b, c + 1) + x(2, a - 1, b, c); };
  auto Us_v = [&](int a, int b, int c) { return x(0, a + 1, b, c) + x(0, a, b, c) + x(0, a + 1, b - 1, c) + x(0, a, b - 1, c); };
  auto Ws_v = [&](int a, int b, int c) { return x(2, a, b, c + 1) + x(2, a, b, c) + x(2, a, b - 1, c + 1) + x(2, a, b - 1, c); };
  auto Us_w = [&](int a, int b, int c) { return x(0, a + 1, b, c) + x(0, a, b, c) + x(0, a + 1, b, c - 1) + x(0, a, b, c - 1); };
  auto Vs_w = [&](int a, int b, int c) { return x(1, a, b + 1, c) + x(1, a, b, c) + x(1, a, b + 1, c - 1) + x(1, a, b, c - 1); };

  // u at Q
  if (in_field(g, 0, i, j, k)) {
    double acc = adj_interior(g, 0, i, j, k) ? (double)((const T*)idv.f[0])[L.idx(i, j, k)] : 0.0;
    double n;
    if ((n = nb(0, i, j, k)) != 0.0) acc += n * (-(x(0, i + 1, j, k) - x(0, i - 1, j, k)) * A2 - 2 * Dsum);
    if ((n = nb(0, i - 1, j, k)) != 0.0) acc += n * (-x(0, i - 1, j, k) * A2 + Dx);
    if ((n = nb(0, i + 1, j, k)) != 0.0) acc += n * (x(0, i + 1, j, k) * A2 + Dx);
    if ((n = nb(0, i, j - 1, k)) != 0.0) acc += n * (-Vs_u(i, j - 1, k) * B8 + Dy);
    if ((n = nb(0, i, j + 1, k)) != 0.0) acc += n * (Vs_u(i, j + 1, k) * B8 + Dy);
    if ((n = nb(0, i, j, k - 1)) != 0.0) acc += n * (-Ws_u(i, j, k - 1) * C8 + Dz);
    if ((n = nb(0, i, j, k + 1)) != 0.0) acc += n * (Ws_u(i, j, k + 1) * C8 + Dz);
    // from v outputs: cu_v(P) = -(v(P+x) - v(P-x)) A8 at P = Q-x, Q, Q-x+y, Q+y
    const int pv[4][3] = {{i - 1, j, k}, {i, j, k}, {i - 1, j + 1, k}, {i, j + 1, k}};
    for (int q = 0; q < 4; q++)
      if ((n = nb(1, pv[q][0], pv[q][1], pv[q][2])) != 0.0)
        acc += n * (-(x(1, pv[q][0] + 1, pv[q][1], pv[q][2]) - x(1, pv[q][0] - 1, pv[q][1], pv[q][2])) * A8);
    // from w outputs: cu_w(P) = -(w(P+x) - w(P-x)) A8 at P = Q-x, Q, Q-x+z, Q+z
    const int pw[4][3] = {{i - 1, j, k}, {i, j, k}, {i - 1, j, k + 1}, {i, j, k + 1}};
    for (int q = 0; q < 4; q++)
      if ((n = nb(2, pw[q][0], pw[q][1], pw[q][2])) != 0.0)
        acc += n * (-(x(2, pw[q][0] + 1, pw[q][1], pw[q][2]) - x(2, pw[q][0] - 1, pw[q][1], pw[q][2])) * A8);
    ((T*)Xbv.f[0])[L.idx(i, j, k)] = (T)acc;
  }
  // v at Q
  if (in_field(g, 1, i, j, k)) {
    double acc = adj_interior(g, 1, i, j, k) ? (double)((const T*)idv.f[1])[L.idx(i, j, k)] : 0.0;
    double n;
    // from u outputs: cv_u(P) = -(u(P+y) - u(P-y)) B8 at P = Q-y, Q, Q+x-y, Q+x
    const int pu[4][3] = {{i, j - 1, k}, {i, j, k}, {i + 1, j - 1, k}, {i + 1, j, k}};
    for (int q = 0; q < 4; q++)
      if ((n = nb(0, pu[q][0], pu[q][1], pu[q][2])) != 0.0)
        acc += n * (-(x(0, pu[q][0], pu[q][1] + 1, pu[q][2]) - x(0, pu[q][0], pu[q][1] - 1, pu[q][2])) * B8);
    if ((n = nb(1, i, j, k)) != 0.0) acc += n * (-(x(1, i, j + 1, k) - x(1, i, j - 1, k)) * B2 - 2 * Dsum);
    if ((n = nb(1, i - 1, j, k)) != 0.0) acc += n * (-Us_v(i - 1, j, k) * A8 + Dx);
    if ((n = nb(1, i + 1, j, k)) != 0.0) acc += n * (Us_v(i + 1, j, k) * A8 + Dx);
    if ((n = nb(1, i, j - 1, k)) != 0.0) acc += n * (-x(1, i, j - 1, k) * B2 + Dy);
    if ((n = nb(1, i, j + 1, k)) != 0.0) acc += n * (x(1, i, j + 1, k) * B2 + Dy);
    if ((n = nb(1, i, j, k - 1)) != 0.0) acc += n * (-Ws_v(i, j, k - 1) * C8 + Dz);
    if ((n = nb(1, i, j, k + 1)) != 0.0) acc += n * (Ws_v(i, j, k + 1) * C8 + Dz);
    // from w outputs: cv_w(P) = -(w(P+y) - w(P-y)) B8 at P = Q-y, Q, Q-y+z, Q+z
    const int pw[4][3] = {{i, j - 1, k}, {i, j, k}, {i, j - 1, k + 1}, {i, j, k + 1}};
    for (int q = 0; q < 4; q++)
      if ((n = nb(2, pw[q][0], pw[q][1], pw[q][2])) != 0.0)
        acc += n * (-(x(2, pw[q][0], pw[q][1] + 1, pw[q][2]) - x(2, pw[q][0], pw[q][1] - 1, pw[q][2])) * B8);
    ((T*)Xbv.f[1])[L.idx(i, j, k)] = (T)acc;
  }
  // w at Q
  if (in_field(g, 2, i, j, k)) {
    double acc = adj_interior(g, 2, i, j, k) ? (double)((const T*)idv.f[2])[L.idx(i, j, k)] : 0.0;
    double n;
    // from u outputs: cw_u(P) = -(u(P+z) - u(P-z)) C8 at P = Q-z, Q, Q+x-z, Q+x
    const int pu[4][3] = {{i, j, k - 1}, {i, j, k}, {i + 1, j, k - 1}, {i + 1, j, k}};
    for (int q = 0; q < 4; q++)
      if ((n = nb(0, pu[q][0], pu[q][1], pu[q][2])) != 0.0)
        acc += n * (-(x(0, pu[q][0], pu[q][1], pu[q][2] + 1) - x(0, pu[q][0], pu[q][1], pu[q][2] - 1)) * C8);
    // from v outputs: cw_v(P) = -(v(P+z) - v(P-z)) C8 at P = Q-z, Q, Q+y-z, Q+y
    const int pv[4][3] = {{i, j, k - 1}, {i, j, k}, {i, j + 1, k - 1}, {i, j + 1, k}};
    for (int q = 0; q < 4; q++)
      if ((n = nb(1, pv[q][0], pv[q][1], pv[q][2])) != 0.0)
        acc += n * (-(x(1, pv[q][0], pv[q][1], pv[q][2] + 1) - x(1, pv[q][0], pv[q][1], pv[q][2] - 1)) * C8);
    if ((n = nb(2, i, j, k)) != 0.0) acc += n * (-(x(2, i, j, k + 1) - x(2, i, j, k - 1)) * C2 - 2 * Dsum);
    if ((n = nb(2, i - 1, j, k)) != 0.0) acc += n * (-Us_w(i - 1, j, k) * A8 + Dx);
    if ((n = nb(2, i + 1, j, k)) != 0.0) acc += n * (Us_w(i + 1, j, k) * A8 + Dx);
    if ((n = nb(2, i, j - 1, k)) != 0.0) acc += n * (-Vs_w(i, j - 1, k) * B8 + Dy);
    if ((n = nb(2, i, j + 1, k)) != 0.0) acc += n * (Vs_w(i, j + 1, k) * B8 + Dy);
    if ((n = nb(2, i, j, k - 1)) != 0.0) acc += n * (-x(2, i, j, k - 1) * C2 + Dz);
    if ((n = nb(2, i, j, k + 1)) != 0.0) acc += n * (x(2, i, j, k + 1) * C2 + Dz);
    ((T*)Xbv.f[2])[L.idx(i, j, k)] = (T)acc;
  }
}

// ---------------------------------------------------------------------------
// Launchers: grids cover the union of all component extents (N + 1 per axis).
// ---------------------------------------------------------------------------
static dim3 full_grid(const DGrid& g) { return dim3((g.N[0] + 1 + 255) / 256, g.N[1] + 1, g.N[2] + 1); }

template <typename T>
void launch_adj_grad(const DGrid& g, const Layout& L, const T* in, T* out, const AdjVec& Ob, double coef,
                     hipStream_t st) {
  dim3 grd((g.N[0] + 255) / 256, g.N[1], g.N[2]);
  k_adj_grad<T><<<grd, 256, 0, st>>>(g, L, in, out, Ob, coef);
}

template <typename T>
void launch_adj_unpack(const DGrid& g, const Layout& L, const T* phib, T* B, long bp0, long bplane, hipStream_t st) {
  dim3 grd((g.npts[0] + 255) / 256, g.npts[1], g.npts[2]);
  k_adj_unpack<T><<<grd, 256, 0, st>>>(g, L, phib, B, bp0, bplane);
}

template <typename T>
void launch_adj_div(const DGrid& g, const Layout& L, const T* B, long bp0, long bplane, const AdjVec& yb, double dts,
                    hipStream_t st) {
  k_adj_div<T><<<full_grid(g), 256, 0, st>>>(g, L, B, bp0, bplane, yb, dts);
}

template <typename T>
void launch_adj_bc(const DGrid& g, const Layout& L, const AdjVec& yb, const AdjVec& Yb, hipStream_t st) {
  k_adj_bc<T><<<full_grid(g), 256, 0, st>>>(g, L, yb, Yb);
}

template <typename T>
int launch_adj_nbar(int stage, const DGrid& g, const Layout& L, const AdjVec& Y, const AdjVec& Nb, const AdjVec& Sb,
                    const AdjVec& R1b, double a1, double a2, double a3, double* partial, hipStream_t st) {
  const dim3 grd = full_grid(g);
  if (stage == 3)
    k_adj_nbar<T, 3><<<grd, 256, 0, st>>>(g, L, Y, Nb, Sb, R1b, a1, a2, a3, partial);
  else if (stage == 2)
    k_adj_nbar<T, 2><<<grd, 256, 0, st>>>(g, L, Y, Nb, Sb, R1b, a1, a2, a3, partial);
  else
    k_adj_nbar<T, 1><<<grd, 256, 0, st>>>(g, L, Y, Nb, Sb, R1b, a1, a2, a3, partial);
  return (int)(grd.x * grd.y * grd.z);
}

template <typename T>
void launch_adj_convdiff(const DGrid& g, const Layout& L, const AdjVec& X, const AdjVec& Nb, const AdjVec& id,
                         const AdjVec& Xb, hipStream_t st) {
  k_adj_convdiff<T><<<full_grid(g), 256, 0, st>>>(g, L, X, Nb, id, Xb);
}

// sum of per-block partial triples into out[0..2] (one block, fixed order)
__global__ void __launch_bounds__(256) k_sum3(const double* partial, long nblocks, double* out) {
  __shared__ double sh[3][256];
  for (int c = 0; c < 3; c++) {
    double s = 0;
    for (long q = threadIdx.x; q < nblocks; q += blockDim.x) s += partial[3 * q + c];
    sh[c][threadIdx.x] = s;
  }
  __syncthreads();
  for (int stp = blockDim.x / 2; stp > 0; stp >>= 1) {
    if (threadIdx.x < stp)
      for (int c = 0; c < 3; c++) sh[c][threadIdx.x] += sh[c][threadIdx.x + stp];
    __syncthreads();
  }
  if (threadIdx.x < 3) out[threadIdx.x] += sh[threadIdx.x][0];
}

void launch_sum3(const double* partial, long nblocks, double* out, hipStream_t st) {
  k_sum3<<<1, 256, 0, st>>>(partial, nblocks, out);
}

#define MIF_ADJ_INST(T)                                                                                           \
  template void launch_adj_grad<T>(const DGrid&, const Layout&, const T*, T*, const AdjVec&, double, hipStream_t); \
  template void launch_adj_unpack<T>(const DGrid&, const Layout&, const T*, T*, long, long, hipStream_t);         \
  template void launch_adj_div<T>(const DGrid&, const Layout&, const T*, long, long, const AdjVec&, double,       \
                                  hipStream_t);                                                                   \
  template void launch_adj_bc<T>(const DGrid&, const Layout&, const AdjVec&, const AdjVec&, hipStream_t);         \
  template int launch_adj_nbar<T>(int, const DGrid&, const Layout&, const AdjVec&, const AdjVec&, const AdjVec&,  \
                                  const AdjVec&, double, double, double, double*, hipStream_t);                   \
  template void launch_adj_convdiff<T>(const DGrid&, const Layout&, const AdjVec&, const AdjVec&, const AdjVec&,  \
                                       const AdjVec&, hipStream_t);

MIF_ADJ_INST(double)
MIF_ADJ_INST(float)

}  // namespace mif
