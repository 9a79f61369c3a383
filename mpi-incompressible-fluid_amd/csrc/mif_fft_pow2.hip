// mif_fft_pow2.hip -- transform passes for complex lengths 2^k and 3 * 2^k.
//
// Same passes and slot order as mif_fft.hip (see there for the algorithm
// and the reference it follows), specialised at compile time for
// L = 64, 128, 256, 512, 1024 and 96, 192, 384, 768, 1536 (periodic
// n = L, Neumann n = L/2 + 1),
// the lengths of the 2^k- and 3*2^k-interval grids:
//   * radix-8 Stockham stages (then a radix-4 or -2 stage, and a radix-3
//     stage for 3 * 2^k), compile-time index arithmetic, data staged in
//     registers so one LDS buffer suffices;
//   * twiddles in LDS; one complex pad element every 8 (the radix-8 store
//     stride) so each group of 8 lanes of a 16-byte LDS access hits
//     distinct banks;
//   * global loads go straight into the packed complex lines and the
//     post-processing writes global memory directly (no staging pass);
//   * NP line pairs per 256-thread workgroup: 2*NP adjacent x positions on
//     the strided axes, so every global access is a run of 2*NP*sizeof(T)
//     contiguous bytes.
#include "mif_kernels.h"

namespace mif {

namespace p2 {

template <typename T>
struct Cx {
  T re, im;
};

template <typename T>
__device__ inline Cx<T> add(Cx<T> a, Cx<T> b) {
  return {a.re + b.re, a.im + b.im};
}
template <typename T>
__device__ inline Cx<T> sub(Cx<T> a, Cx<T> b) {
  return {a.re - b.re, a.im - b.im};
}
template <typename T>
__device__ inline Cx<T> mul(Cx<T> a, Cx<T> b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <int SIGN, typename T>
__device__ inline Cx<T> mul_si(Cx<T> a) {  // a * (SIGN * i)
  return SIGN < 0 ? Cx<T>{a.im, -a.re} : Cx<T>{-a.im, a.re};
}

template <int SIGN, int R, typename T>
__device__ inline void dft(Cx<T>* v) {
  if constexpr (R == 2) {
    const Cx<T> a = v[0], b = v[1];
    v[0] = add(a, b);
    v[1] = sub(a, b);
  } else if constexpr (R == 4) {
    const Cx<T> t0 = add(v[0], v[2]), t1 = sub(v[0], v[2]);
    const Cx<T> t2 = add(v[1], v[3]), t3 = mul_si<SIGN>(sub(v[1], v[3]));
    v[0] = add(t0, t2);
    v[2] = sub(t0, t2);
    v[1] = add(t1, t3);
    v[3] = sub(t1, t3);
  } else if constexpr (R == 3) {
    // w = exp(SIGN 2 pi i / 3): X1 = a - (b+c)/2 + (sqrt3/2) SIGN i (b - c)
    const T hs = (T)0.86602540378443864676;
    const Cx<T> a = v[0], t1 = add(v[1], v[2]);
    const Cx<T> t2 = {a.re - (T)0.5 * t1.re, a.im - (T)0.5 * t1.im};
    const Cx<T> d = mul_si<SIGN>(sub(v[1], v[2]));
    const Cx<T> t3 = {hs * d.re, hs * d.im};
    v[0] = add(a, t1);
    v[1] = add(t2, t3);
    v[2] = sub(t2, t3);
  } else {
    static_assert(R == 8, "radix");
    Cx<T> e[4] = {v[0], v[2], v[4], v[6]};
    Cx<T> o[4] = {v[1], v[3], v[5], v[7]};
    dft<SIGN, 4>(e);
    dft<SIGN, 4>(o);
    const T s = (T)0.70710678118654752440;
    const Cx<T> o1 = mul(o[1], Cx<T>{s, (T)SIGN * s});
    const Cx<T> o2 = mul_si<SIGN>(o[2]);
    const Cx<T> o3 = mul(o[3], Cx<T>{-s, (T)SIGN * s});
    v[0] = add(e[0], o[0]);
    v[4] = sub(e[0], o[0]);
    v[1] = add(e[1], o1);
    v[5] = sub(e[1], o1);
    v[2] = add(e[2], o2);
    v[6] = sub(e[2], o2);
    v[3] = add(e[3], o3);
    v[7] = sub(e[3], o3);
  }
}

constexpr int ilog2(int x) { return x <= 1 ? 0 : 1 + ilog2(x >> 1); }
// L = 2^m or 3 * 2^m: stages of radix 8, 8, ..., then 4 or 2, then 3
constexpr bool has3(int L) { return L % 3 == 0; }
constexpr int pow2_part(int L) { return has3(L) ? L / 3 : L; }
constexpr int radix_of(int L, int s) {
  const int m = ilog2(pow2_part(L)), full = m / 3;
  if (s < full) return 8;
  if (m % 3 && s == full) return 1 << (m % 3);
  return 3;
}
constexpr int nstages(int L) {
  return ilog2(pow2_part(L)) / 3 + (ilog2(pow2_part(L)) % 3 ? 1 : 0) + (has3(L) ? 1 : 0);
}
constexpr int ns_of(int L, int s) { return s == 0 ? 1 : ns_of(L, s - 1) * radix_of(L, s - 1); }
constexpr int min_radix(int L, int s = 0) {
  return s + 1 >= nstages(L) ? radix_of(L, s)
                             : (radix_of(L, s) < min_radix(L, s + 1) ? radix_of(L, s) : min_radix(L, s + 1));
}

// padded LDS index: one pad element every 8 (one every 16, which a bank
// model of the radix stages rates conflict-free, measured no faster: 256^3
// mixed step 1.574 -> 1.582 ms)
__device__ constexpr inline int pd(int e) { return e + (e >> 3); }
// LDS pitch of one line pair: the padded length plus two elements, so that
// the lanes of a wave that address neighbouring elements of different pairs
// (load, store and scaling phases of the strided passes: lanes run over the
// lines first, then over the elements) fall in different banks.  Measured
// at 256^3 against one element: FP32 y pass 38.4 -> 37.0 us, mixed step
// 1.576 -> 1.561 ms (four elements: the same; eight: the z pass slower)
// (Two, four and six extra elements, the latter two conflict-free in a
// 64-bank model of the load and store phases, measured within 1 % of each
// other on the 256^3 y and z passes, FP64 and FP32.)
constexpr int lds_pitch(int L) { return L + L / 8 + 2; }

// Twiddle entries the stages read: stage S reads tw[k * L / (Ns R)] for
// k < Ns, all below L / R, so the table needs L / (smallest radix) entries.
constexpr int tw_count(int L) { return L / min_radix(L); }

template <typename T, int L, int NP, int SIGN, int S>
__device__ inline void stage(Cx<T> (*Z)[lds_pitch(L)], const Cx<T>* tw) {
  constexpr int R = radix_of(L, S), Ns = ns_of(L, S), m = L / R;
  constexpr int items = NP * m;
  constexpr int per = (items + 255) / 256;
  constexpr int tstep = L / (Ns * R);
  Cx<T> v[per][R];
  const int t = threadIdx.x;
#pragma unroll
  for (int b = 0; b < per; b++) {
    const int it = t + 256 * b;
    if (items % 256 == 0 || it < items) {
      const int p = it / m, jb = it % m, k = jb % Ns;
#pragma unroll
      for (int r = 0; r < R; r++) v[b][r] = Z[p][pd(jb + r * m)];
      if (Ns > 1) {
        // one twiddle from the table, the others as its powers (products at
        // most three deep: w2 = w1 w1, w4 = w2 w2, w3 = w1 w2, w5 = w1 w4,
        // w6 = w2 w4, w7 = w3 w4), one LDS read per butterfly instead of R-1
        Cx<T> w[R];
        w[1] = tw[k * tstep];
        if (SIGN > 0) w[1].im = -w[1].im;
        if constexpr (R >= 3) w[2] = mul(w[1], w[1]);
        if constexpr (R >= 4) w[3] = mul(w[1], w[2]);
        if constexpr (R == 8) {
          w[4] = mul(w[2], w[2]);
          w[5] = mul(w[1], w[4]);
          w[6] = mul(w[2], w[4]);
          w[7] = mul(w[3], w[4]);
        }
#pragma unroll
        for (int r = 1; r < R; r++) v[b][r] = mul(v[b][r], w[r]);
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int b = 0; b < per; b++) {
    const int it = t + 256 * b;
    if (items % 256 == 0 || it < items) {
      const int p = it / m, jb = it % m, k = jb % Ns;
      dft<SIGN, R>(v[b]);
      const int idx = (jb / Ns) * Ns * R + k;
#pragma unroll
      for (int r = 0; r < R; r++) Z[p][pd(idx + r * Ns)] = v[b][r];
    }
  }
  __syncthreads();
}

template <typename T, int L, int NP, int SIGN, int S = 0>
__device__ inline void fft(Cx<T> (*Z)[lds_pitch(L)], const Cx<T>* tw) {
  if constexpr (S < nstages(L)) {
    stage<T, L, NP, SIGN, S>(Z, tw);
    fft<T, L, NP, SIGN, S + 1>(Z, tw);
  }
}

}  // namespace p2

using p2::Cx;
using p2::pd;

// Stores and loads of the transform passes on the packed buffer stay
// ordinary: the next pass reads the buffer back from the Infinity Cache
// (non-temporal stores there took the 256^3 FP32 y pass from 39 to 43 us,
// non-temporal loads to 53 us).
#define MIF_ST_F(addr, val) (*(addr) = (val))
// The x inverse pass stores into the pressure field non-temporally at every
// size (256^3 step with the stage and projection stores non-temporal too:
// FP64 3.166 -> 3.110 ms, mixed 1.654 -> 1.632 ms; 768^3: unpack 1256 ->
// 1249 us, the projection after it 3578 -> 3427 us).  A run-time choice
// here is not kept: the compiler merges the two store branches into plain
// stores.
#define MIF_ST_U(addr, val) __builtin_nontemporal_store((val), (addr))

// LDS bytes of one transform workgroup (the static __shared__ limit of a
// kernel), and the bound under which the twiddle table joins the lines in LDS
constexpr int kLdsMax = 65536;
constexpr int kTwlMax = 65536;

// MODE 0: forward, 1: inverse, 2: forward + scale + inverse (axis 2 only).
// FUSED (axis 0 only): MODE 0 reads div(u)/dt from the velocity instead of
// B (compute_rhs_homogeneous_periodic, src/PressureEquation.cpp:59-61);
// MODE 1 writes the solution into the pressure field, owner points and
// periodic ghost images (PressureTensor::copy_to_staggered,
// src/PressureTensor.cpp:21-34), instead of B.
// TV: element type of the velocity and pressure fields of the fused passes.
// DC (fused x passes of the mixed form's one-sweep defect correction, FP32):
// see FusedIO::dc.
// RES (x forward pass, FP32): the lines are the defect r = (b - mean) - L phi
// of the first FP32 solution (FusedIO::res), formed in FP64 as they load.
template <typename T, int L, int NP, int AXIS, int MODE, bool FUSED, typename TV = T, bool DC = false,
          bool RES = false>
__global__ void __launch_bounds__(256) k_fft_p2(DGrid g, const T* __restrict__ W, SpectralScale sc, T* __restrict__ B,
                                                long bp0, long bplane, int periodic, FusedIO io) {
  static_assert(!DC || (FUSED && AXIS == 0), "defect-correction outputs belong to the fused x passes");
  static_assert(!RES || (AXIS == 0 && MODE == 0 && !FUSED), "the defect is formed by the x forward pass");
  // twiddles live in LDS when the workgroup stays within kTwlMax
  // bytes, otherwise they are read from the (L1-resident) global table
  constexpr int LP = p2::lds_pitch(L);
  constexpr int TWN = p2::tw_count(L);
  constexpr bool TWL = (NP * LP + TWN) * sizeof(Cx<T>) <= kTwlMax;
  static_assert(NP * LP * sizeof(Cx<T>) + (L / 2 + 1) * sizeof(double) <= kLdsMax - (TWL ? TWN * sizeof(Cx<T>) : 0),
                "workgroup LDS above the limit");
  __shared__ Cx<T> Z[NP][LP];
  __shared__ Cx<T> tws[TWL ? TWN : 1];
  __shared__ long lbase[2 * NP];
  const Cx<T>* tw = TWL ? tws : (const Cx<T>*)W;
  constexpr int NL = 2 * NP;
  const int t = threadIdx.x;
  // tile (bx, by) from the XCD-contiguous order: neighbouring column tiles
  // of a strided pass and neighbouring row groups of the x pass share L2
  int bx, by;
  {
    const int tile = xcd_tile(blockIdx.x + gridDim.x * blockIdx.y, gridDim.x * gridDim.y);
    bx = tile % gridDim.x;
    by = tile / gridDim.x;
  }
  const int n = periodic ? L : L / 2 + 1;
  const int n0 = g.npts[0], n1 = g.npts[1], n2 = g.npts[2];
  const long ms = AXIS == 0 ? 1 : (AXIS == 1 ? bp0 : bplane);
  if (TWL)
    for (int e = t; e < TWN; e += 256) tws[e] = Cx<T>{W[2 * e], W[2 * e + 1]};
  // eigenvalues of the scaling step (fused z pass): the z table, this
  // workgroup's x columns and its y row, staged once
  __shared__ double lamz[MODE == 2 ? L / 2 + 1 : 1];
  __shared__ double lamx[MODE == 2 ? 2 * NP : 1];
  __shared__ double lamy;
  if (MODE == 2) {
    for (int e = t; e <= L / 2; e += 256) lamz[e] = sc.lam[2][e];
    if (t < NL) {
      const int ic = bx * NL + t;
      lamx[t] = ic < g.npts[0] ? sc.lam[0][ic] : 0.0;
    }
    if (t == 0) lamy = sc.lam[1][by];
  }
  // fused x passes, per line: the owner row's first point (divergence), or
  // the rows the line is written to, owner row and its periodic images in y
  // and z (unpack); -1: none.  Computed once per workgroup, out of the
  // per-point loops.
  __shared__ long prow[AXIS == 0 && FUSED ? 2 * NP : 1][4];
  if (AXIS == 0 && FUSED && MODE == 0 && t < NL) {
    const int r = bx * NL + t;
    prow[t][0] = r < n1 * n2 ? io.L.idx(g.olo[0], r % n1 + g.olo[1], r / n1 + g.olo[2]) : -1;
  }
  if (AXIS == 0 && FUSED && MODE == 1 && t < NL) {
    const int r = bx * NL + t;
    long rows[4] = {-1, -1, -1, -1};
    if (r < n1 * n2) {
      const int N1 = g.N[1], N2 = g.N[2];
      const int j = r % n1 + g.olo[1], k = r / n1 + g.olo[2];
      int js[2] = {j, -1}, ks[2] = {k, -1};
      if (g.periodic[1]) js[1] = j == 1 ? N1 - 1 : (j == N1 - 2 ? 0 : -1);
      if (g.periodic[2]) ks[1] = k == 1 ? N2 - 1 : (k == N2 - 2 ? 0 : -1);
      for (int qk = 0; qk < 2; qk++)
        for (int qj = 0; qj < 2; qj++)
          if (js[qj] >= 0 && ks[qk] >= 0) rows[2 * qk + qj] = io.L.idx(0, js[qj], ks[qk]);
    }
    for (int q = 0; q < 4; q++) prow[t][q] = rows[q];
  }
  // DC: each line's owner weight along y and z, formed once per workgroup
  __shared__ double lwt[DC ? 2 * NP : 1];
  if (DC && t < NL) {
    const int r = bx * NL + t;
    lwt[t] = r < n1 * n2 ? owner_weight(g, 1, r % n1) * owner_weight(g, 2, r / n1) : 0.0;
  }
  // defect lines (RES): the rows of each line's y and z neighbours
  __shared__ long yzrow[RES ? 2 * NP : 1][4];
  if (RES && t < NL) {
    const int r = bx * NL + t;
    const int j = r < n1 * n2 ? r % n1 : 0, k = r < n1 * n2 ? r / n1 : 0;
    yzrow[t][0] = bp0 * nb(g, 1, j, -1) + bplane * k;
    yzrow[t][1] = bp0 * nb(g, 1, j, 1) + bplane * k;
    yzrow[t][2] = bp0 * j + bplane * nb(g, 2, k, -1);
    yzrow[t][3] = bp0 * j + bplane * nb(g, 2, k, 1);
  }
  if (t < NL) {
    long b = -1;
    if (AXIS == 0) {
      const int r = bx * NL + t;
      if (r < n1 * n2) b = bp0 * (r % n1) + bplane * (r / n1);
    } else {
      const int i = bx * NL + t;
      if (i < n0) b = AXIS == 1 ? i + bplane * by : i + bp0 * by;
    }
    lbase[t] = b;
  }
  __syncthreads();

  // ---- load: line l -> pair l % NP, real part (l < NP) or imaginary part
  auto put = [&](int l, int m, T val) {
    const int p = l % NP;
    T* z = (T*)&Z[p][pd(m)];
    z[l < NP ? 0 : 1] = val;
    if (!periodic && m >= 1 && m <= n - 2) ((T*)&Z[p][pd(L - m)])[l < NP ? 0 : 1] = val;
  };
  // Global loads are issued in unrolled chunks from always-valid addresses
  // (an element outside the lines reads index 0 and is discarded), so a
  // chunk's loads are in flight together before its values go to LDS.
  constexpr int R0 = (L + 255) / 256;  // elements per thread and line on axis 0
  // DC: this thread's owner-weighted sums (W.b or W.phi, W.1)
  double dc_s = 0, dc_w = 0;
  auto line_weight = [&](int l) -> double { return lwt[l]; };
  if (AXIS == 0 && FUSED && MODE == 0) {
    const TV* u = (const TV*)io.u;
    const TV* v = (const TV*)io.v;
    const TV* w = (const TV*)io.w;
    const long PX = io.L.PX, PL = io.L.plane;
    // lines per chunk (6 loads per value): 8 in FP32 (256^3 FP32 pass 53.6 ->
    // 52.7 us against 4; 2 took 54.4), 4 in FP64
    constexpr int CK = sizeof(T) == 4 ? 8 : 4;
    constexpr int CL0 = R0 >= CK ? 1 : CK / R0;
    constexpr int CL = CL0 < NL ? CL0 : NL;
    for (int l0 = 0; l0 < NL; l0 += CL) {
      T val[CL][R0];
#pragma unroll
      for (int a = 0; a < CL; a++) {
        const int l = l0 + a;
        const long rb = l < NL ? prow[l][0] : -1;
        const bool okl = rb >= 0;
#pragma unroll
        for (int q = 0; q < R0; q++) {
          const int m = t + 256 * q;
          const bool okm = okl && m < n;
          const long o = okm ? rb + m : 0;
          const TV u0 = u[o], u1 = u[o + 1], v0 = v[o], v1 = v[o + PX], w0 = w[o], w1 = w[o + PL];
          const double du_dx = ((double)u1 - (double)u0) * g.one_over_h[0];
          const double dv_dy = ((double)v1 - (double)v0) * g.one_over_h[1];
          const double dw_dz = ((double)w1 - (double)w0) * g.one_over_h[2];
          const double bv = (du_dx + dv_dy + dw_dz) / io.dt;
          val[a][q] = okm ? (T)bv : (T)0;
          if constexpr (DC) {
            if (okm) {
              io.b64[lbase[l] + m] = bv;
              const double wgt = owner_weight(g, 0, m) * line_weight(l);
              dc_s += wgt * bv;
              dc_w += wgt;
            }
          }
        }
      }
#pragma unroll
      for (int a = 0; a < CL; a++)
#pragma unroll
        for (int q = 0; q < R0; q++) {
          const int l = l0 + a, m = t + 256 * q;
          if (l < NL && m < n) put(l, m, val[a][q]);
        }
    }
  } else if (AXIS == 0) {
    constexpr int CL0 = R0 >= 16 ? 1 : 16 / R0;  // lines per chunk: 16 values in flight
    constexpr int CL = CL0 < NL ? CL0 : NL;
    // RES: b's weighted mean and the 7-point operator's coefficients
    const double rmean = RES ? io.bmean->s1 / io.bmean->s2 : 0.0;
    const double ihx2 = 1.0 / (g.h[0] * g.h[0]), ihy2 = 1.0 / (g.h[1] * g.h[1]), ihz2 = 1.0 / (g.h[2] * g.h[2]);
    // RES: each thread's points and their x neighbours in a line (the same
    // for every line; points past the line read point 0)
    int xo[RES ? R0 : 1], xm[RES ? R0 : 1], xp[RES ? R0 : 1];
    if constexpr (RES) {
#pragma unroll
      for (int q = 0; q < R0; q++) {
        const int m = t + 256 * q;
        xo[q] = m < n ? m : 0;
        xm[q] = (int)nb(g, 0, xo[q], -1);
        xp[q] = (int)nb(g, 0, xo[q], 1);
      }
    }
    for (int l0 = 0; l0 < NL; l0 += CL) {
      T val[CL][R0];
#pragma unroll
      for (int a = 0; a < CL; a++) {
        const int l = l0 + a;
        const long b = l < NL ? lbase[l] : -1;
        if constexpr (RES) {
          // r = (b - mean) - (lx + ly + lz) in FP64, the 7-point operator
          // scaled by 1/h^2; a line outside the grid reads row 0 and is dropped
          const float* P = io.phi32;
          const int lc = l < NL ? l : 0;
          const long row = b >= 0 ? b : 0;
          const long rym = yzrow[lc][0], ryp = yzrow[lc][1], rzm = yzrow[lc][2], rzp = yzrow[lc][3];
#pragma unroll
          for (int q = 0; q < R0; q++) {
            const int m = t + 256 * q;
            const bool okm = b >= 0 && m < n;
            const long o = row + xo[q];
            const double c = P[o];
            const double lx = ((double)P[row + xm[q]] - 2 * c + (double)P[row + xp[q]]) * ihx2;
            const double ly = ((double)P[rym + xo[q]] - 2 * c + (double)P[ryp + xo[q]]) * ihy2;
            const double lz = ((double)P[rzm + xo[q]] - 2 * c + (double)P[rzp + xo[q]]) * ihz2;
            const T x = (T)((io.b64[o] - rmean) - (lx + ly + lz));
            val[a][q] = okm ? x : (T)0;
          }
          continue;
        }
#pragma unroll
        for (int q = 0; q < R0; q++) {
          const int m = t + 256 * q;
          const bool okm = b >= 0 && m < n;
          const T x = B[okm ? b + m : 0];
          val[a][q] = okm ? x : (T)0;
        }
      }
#pragma unroll
      for (int a = 0; a < CL; a++)
#pragma unroll
        for (int q = 0; q < R0; q++) {
          const int l = l0 + a, m = t + 256 * q;
          if (l < NL && m < n) put(l, m, val[a][q]);
        }
    }
  } else {
    const int l = t % NL;
    const long b = lbase[l];
    constexpr int STEP = 256 / NL;        // rows per pass of the workgroup
    constexpr int MQ = (L + STEP - 1) / STEP;  // rows per thread for n <= L
    constexpr int CH = MQ < 16 ? MQ : 16;      // rows per chunk
    for (int q0 = 0; q0 < MQ && (t / NL) + q0 * STEP < n; q0 += CH) {
      T val[CH];
#pragma unroll
      for (int q = 0; q < CH; q++) {
        const int m = t / NL + (q0 + q) * STEP;
        const bool okm = b >= 0 && m < n;
        const T x = B[okm ? b + m * ms : 0];
        val[q] = okm ? x : (T)0;
      }
#pragma unroll
      for (int q = 0; q < CH; q++) {
        const int m = t / NL + (q0 + q) * STEP;
        if (m < n) put(l, m, val[q]);
      }
    }
  }
  __syncthreads();

  if (MODE == 1 && periodic) {
    // halfcomplex (h1, h2) held as Z[m] = (h1[m], h2[m]) -> Z_k = X1_k + i X2_k
    for (int it = t; it < NP * (L / 2 + 1); it += 256) {
      const int p = it / (L / 2 + 1), k = it % (L / 2 + 1);
      if (k == 0 || k == L / 2) continue;  // Z_0 and Z_{L/2} are already (X1, X2)
      const Cx<T> a = Z[p][pd(k)], b = Z[p][pd(L - k)];
      // X1_k = h1[k] + i h1[L-k], X2_k = h2[k] + i h2[L-k]
      Z[p][pd(k)] = Cx<T>{a.re - b.im, b.re + a.im};
      Z[p][pd(L - k)] = Cx<T>{a.re + b.im, -b.re + a.im};
    }
    __syncthreads();
  }

  if (MODE == 1)
    p2::fft<T, L, NP, +1>(Z, tw);
  else
    p2::fft<T, L, NP, -1>(Z, tw);

  if (MODE == 2) {
    // scale every mode by 1/((lx+ly+lz) * norm); constant mode -> 0
    const int j = by;
    for (int it = t; it < NP * (L / 2 + 1); it += 256) {
      const int p = it % NP, k = it / NP;
      const int i1 = bx * NL + p, i2 = i1 + NP;
      const double l1 = lamx[p] + lamy, l2 = lamx[p + NP] + lamy;
      if (periodic) {
        const double lz = lamz[k];
        const T s1 = (i1 == 0 && j == 0 && k == 0) ? (T)0 : (T)((1.0 / (l1 + lz)) * sc.inv_norm);
        const T s2 = (i2 == 0 && j == 0 && k == 0) ? (T)0 : (T)((1.0 / (l2 + lz)) * sc.inv_norm);
        const Cx<T> a = Z[p][pd(k)];
        if (k == 0 || k == L / 2) {
          Z[p][pd(k)] = Cx<T>{a.re * s1, a.im * s2};
        } else {
          const Cx<T> b = Z[p][pd(L - k)];
          // X1 = (Z_k + conj Z_{L-k})/2, X2 = (Z_k - conj Z_{L-k})/(2i)
          const T x1r = (a.re + b.re) * (T)0.5 * s1, x1i = (a.im - b.im) * (T)0.5 * s1;
          const T x2r = (a.im + b.im) * (T)0.5 * s2, x2i = (b.re - a.re) * (T)0.5 * s2;
          Z[p][pd(k)] = Cx<T>{x1r - x2i, x1i + x2r};
          Z[p][pd(L - k)] = Cx<T>{x1r + x2i, -x1i + x2r};
        }
      } else {
        // DCT-I coefficients k = 0..n-1 (n = L/2 + 1), even extension kept
        const double lz = lamz[k];
        const T s1 = (i1 == 0 && j == 0 && k == 0) ? (T)0 : (T)((1.0 / (l1 + lz)) * sc.inv_norm);
        const T s2 = (i2 == 0 && j == 0 && k == 0) ? (T)0 : (T)((1.0 / (l2 + lz)) * sc.inv_norm);
        const Cx<T> a = Z[p][pd(k)];
        const Cx<T> sa = Cx<T>{a.re * s1, a.im * s2};
        Z[p][pd(k)] = sa;
        if (k >= 1 && k <= L / 2 - 1) Z[p][pd(L - k)] = sa;
      }
    }
    __syncthreads();
    p2::fft<T, L, NP, +1>(Z, tw);
  }

  // ---- store
  auto out_index = [&](int l, int s) -> long { return lbase[l] + (long)s * ms; };
  if (MODE == 0 && periodic) {
    // halfcomplex spectra of both lines of each pair
    auto emit = [&](int p, int k) {
      const long b1 = lbase[p], b2 = lbase[p + NP];
      const Cx<T> a = Z[p][pd(k)];
      if (k == 0 || k == L / 2) {
        if (b1 >= 0) MIF_ST_F(&B[b1 + k * ms], a.re);
        if (b2 >= 0) MIF_ST_F(&B[b2 + k * ms], a.im);
        return;
      }
      const Cx<T> b = Z[p][pd(L - k)];
      if (b1 >= 0) {
        MIF_ST_F(&B[b1 + k * ms], (a.re + b.re) * (T)0.5);
        MIF_ST_F(&B[b1 + (L - k) * ms], (a.im - b.im) * (T)0.5);
      }
      if (b2 >= 0) {
        MIF_ST_F(&B[b2 + k * ms], (a.im + b.im) * (T)0.5);
        MIF_ST_F(&B[b2 + (L - k) * ms], (b.re - a.re) * (T)0.5);
      }
    };
    if (AXIS == 0) {
      for (int p = 0; p < NP; p++)
        for (int k = t; k <= L / 2; k += 256) emit(p, k);
    } else {
      for (int it = t; it < NP * (L / 2 + 1); it += 256) emit(it % NP, it / NP);
    }
  } else {
    // DCT-I coefficients, or the real lines of an inverse transform
    if (AXIS == 0 && FUSED && MODE == 1) {
      // all lines' values of a point first, then the stores
      TV* P = (TV*)io.P;
      const int N0 = g.N[0];
      for (int m = t; m < n; m += 256) {
        const int i = m + g.olo[0];
        const int i2 = g.periodic[0] ? (i == 1 ? N0 - 1 : (i == N0 - 2 ? 0 : -1)) : -1;
        TV val[NL];
#pragma unroll
        for (int l = 0; l < NL; l++) {
          const Cx<T> a = Z[l % NP][pd(m)];
          if constexpr (DC) {
            // phi = phi_0 + delta in FP64, stored in the field's precision
            const long b = lbase[l];
            const double phi = (double)(l < NP ? a.re : a.im) + (b >= 0 ? (double)io.add32[b + m] : 0.0);
            val[l] = (TV)phi;
            if (b >= 0) {
              const double wgt = owner_weight(g, 0, m) * line_weight(l);
              dc_s += wgt * (double)val[l];
              dc_w += wgt;
            }
          } else {
            val[l] = (TV)(l < NP ? a.re : a.im);
          }
        }
#pragma unroll
        for (int l = 0; l < NL; l++)
#pragma unroll
          for (int q = 0; q < 4; q++) {
            const long rb = prow[l][q];
            if (rb >= 0) {
              MIF_ST_U(&P[rb + i], val[l]);
              if (i2 >= 0) MIF_ST_U(&P[rb + i2], val[l]);
            }
          }
      }
    } else if (AXIS == 0) {
      for (int l = 0; l < NL; l++) {
        const long b = lbase[l];
        if (b < 0) continue;
        for (int m = t; m < n; m += 256) {
          const Cx<T> a = Z[l % NP][pd(m)];
          MIF_ST_F(&B[b + m], l < NP ? a.re : a.im);
        }
      }
    } else {
      const int l = t % NL;
      const long b = lbase[l];
      if (b >= 0)
        for (int m = t / NL; m < n; m += 256 / NL) {
          const Cx<T> a = Z[l % NP][pd(m)];
          MIF_ST_F(&B[out_index(l, m)], l < NP ? a.re : a.im);
        }
    }
  }
  if constexpr (DC) {
    // the workgroup's (W.x, W.1) into its slot of io.partial
    for (int off = 32; off > 0; off >>= 1) {
      dc_s += __shfl_down(dc_s, off, 64);
      dc_w += __shfl_down(dc_w, off, 64);
    }
    __shared__ double red[2][4];
    if ((t & 63) == 0) {
      red[0][t >> 6] = dc_s;
      red[1][t >> 6] = dc_w;
    }
    __syncthreads();
    if (t == 0)
      io.partial[bx + gridDim.x * by] = Red3{red[0][0] + red[0][1] + red[0][2] + red[0][3],
                                              red[1][0] + red[1][1] + red[1][2] + red[1][3], 0};
  }
}

// Neumann y pass at half length.  The DCT-I (FFTW_REDFT00, unnormalised;
// it is its own inverse up to the normalisation the z pass folds in) of a
// line x_0..x_N of n = N + 1 points, through an N-point FFT:
//   y_0 = (x_0 + x_N) / 2,
//   y_j = (x_j + x_{N-j}) / 2 - sin(pi j/N) (x_j - x_{N-j}),   j = 1..N-1,
//   R = DFT_N(y),  F_2k = 2 Re R_k (k = 0..N/2),
//   F_1 = 2 S0, S0 = sum_m c_m x_m (c_0 = 1/2, c_N = -1/2, c_m = cos(pi m/N)),
//   F_2k+1 = F_2k-1 - 2 Im R_k (k = 1..N/2-1): a prefix sum.
// Two lines a, b go through one complex FFT as y^a + i y^b, so a pair of
// lines costs an N-point transform instead of the 2N-point transform of
// their even extensions, and a workgroup holds twice the lines in the same
// LDS.  Each line's values live in LDS as a real array (line a of pair p at
// real offset 0 of Z[p], line b at real offset LP), which the FFT buffer
// Z[p] overlays; every phase reads into registers, synchronises, then
// writes.  S0 and the prefix sum use TPL lanes per line (a group of
// consecutive lanes inside one wave).
template <typename T, int N, int NP>
__global__ void __launch_bounds__(256) k_dct1_y(DGrid g, const T* __restrict__ W, T* __restrict__ B, long bp0,
                                                long bplane) {
  constexpr int LP = p2::lds_pitch(N);
  constexpr int NL = 2 * NP;
  constexpr int STEP = 256 / NL;
  constexpr int n = N + 1;
  constexpr int MQ = (n + STEP - 1) / STEP;
  constexpr int NH = N / 2;
  constexpr int TPL = 256 / NL < 64 ? 256 / NL : 64;
  constexpr int SEG = NH / TPL;
  static_assert(N % 2 == 0 && NH % TPL == 0, "scan segments");
  constexpr int TWN = p2::tw_count(N);
  static_assert((NP * LP + TWN + NH + 1) * sizeof(Cx<T>) <= kLdsMax, "workgroup LDS above the limit");
  constexpr int NI = NP * (NH + 1);         // pre/post items: (pair, j <= N/2)
  constexpr int PER = (NI + 255) / 256;
  __shared__ Cx<T> Z[NP][LP];
  __shared__ Cx<T> tws[TWN];     // exp(-2 pi i e / N), e < TWN
  __shared__ Cx<T> cs[NH + 1];   // exp(-i pi j / N)
  __shared__ T s0[NL];
  const int t = threadIdx.x;
  int bx, by;
  {
    const int tile = xcd_tile(blockIdx.x + gridDim.x * blockIdx.y, gridDim.x * gridDim.y);
    bx = tile % gridDim.x;
    by = tile / gridDim.x;
  }
  // W is the table of the 2N-point plan, exp(-2 pi i e / 2N): its even
  // entries are the N-point twiddles, its first N/2 + 1 entries cos/sin(pi j/N)
  for (int e = t; e < TWN; e += 256) tws[e] = Cx<T>{W[4 * e], W[4 * e + 1]};
  for (int j = t; j <= NH; j += 256) cs[j] = Cx<T>{W[2 * j], W[2 * j + 1]};
  auto line = [&](int l) -> T* { return (T*)&Z[l % NP][0] + (l < NP ? 0 : LP); };

  // ---- load: thread t serves line l = t % NL, rows t / NL + q * STEP
  const int l = t % NL;
  const int i = bx * NL + l;
  const long b = i < g.npts[0] ? i + bplane * by : -1;
  {
    T val[MQ];
#pragma unroll
    for (int q = 0; q < MQ; q++) {
      const int m = t / NL + q * STEP;
      const bool okm = b >= 0 && m < n;
      const T x = B[okm ? b + m * bp0 : 0];
      val[q] = okm ? x : (T)0;
    }
    T* xr = line(l);
#pragma unroll
    for (int q = 0; q < MQ; q++) {
      const int m = t / NL + q * STEP;
      if (m < n) xr[m] = val[q];
    }
  }
  __syncthreads();

  // ---- S0 of each line and the pre-processing, read phase
  const int gl = t / TPL, gs = t % TPL;
  T s0v = 0;
  if (gl < NL) {
    const T* x = line(gl);
    for (int m = gs; m <= N; m += TPL) {
      const T c = m == 0 ? (T)0.5 : (m == N ? (T)-0.5 : (m <= NH ? cs[m].re : -cs[N - m].re));
      s0v += c * x[m];
    }
  }
  for (int off = TPL / 2; off > 0; off >>= 1) s0v += __shfl_xor(s0v, off, 64);
  Cx<T> zj[PER], zn[PER];
#pragma unroll
  for (int r = 0; r < PER; r++) {
    const int it = t + 256 * r;
    if (NI % 256 == 0 || it < NI) {
      const int p = it / (NH + 1), j = it % (NH + 1);
      const T* xa = line(p);
      const T* xb = line(p + NP);
      const T aj = xa[j], an = xa[N - j], bj = xb[j], bn = xb[N - j];
      const T sn = -cs[j].im;  // sin(pi j / N) = sin(pi (N - j) / N)
      zj[r] = Cx<T>{(T)0.5 * (aj + an) - sn * (aj - an), (T)0.5 * (bj + bn) - sn * (bj - bn)};
      zn[r] = Cx<T>{(T)0.5 * (aj + an) + sn * (aj - an), (T)0.5 * (bj + bn) + sn * (bj - bn)};
    }
  }
  __syncthreads();
  if (gl < NL && gs == 0) s0[gl] = s0v;
#pragma unroll
  for (int r = 0; r < PER; r++) {
    const int it = t + 256 * r;
    if (NI % 256 == 0 || it < NI) {
      const int p = it / (NH + 1), j = it % (NH + 1);
      Z[p][pd(j)] = zj[r];
      if (j > 0 && j < NH) Z[p][pd(N - j)] = zn[r];
    }
  }
  __syncthreads();

  p2::fft<T, N, NP, -1>(Z, tws);

  // ---- post-processing: even coefficients, and the odd increments in the
  // odd slots (D_0 = 2 S0 in slot 1, -2 Im R_k in slot 2k+1)
  T fa[PER], fb[PER], da[PER], db[PER];
#pragma unroll
  for (int r = 0; r < PER; r++) {
    const int it = t + 256 * r;
    if (NI % 256 == 0 || it < NI) {
      const int p = it / (NH + 1), k = it % (NH + 1);
      const Cx<T> zk = Z[p][pd(k)], zr = Z[p][pd(k == 0 ? 0 : N - k)];
      fa[r] = zk.re + zr.re;
      fb[r] = zk.im + zr.im;
      da[r] = zr.im - zk.im;
      db[r] = zk.re - zr.re;
    }
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < PER; r++) {
    const int it = t + 256 * r;
    if (NI % 256 == 0 || it < NI) {
      const int p = it / (NH + 1), k = it % (NH + 1);
      T* xa = line(p);
      T* xb = line(p + NP);
      xa[2 * k] = fa[r];
      xb[2 * k] = fb[r];
      if (k == 0) {
        xa[1] = (T)2 * s0[p];
        xb[1] = (T)2 * s0[p + NP];
      } else if (k < NH) {
        xa[2 * k + 1] = da[r];
        xb[2 * k + 1] = db[r];
      }
    }
  }
  __syncthreads();

  // ---- odd coefficients: inclusive prefix sum over the odd slots of each
  // line, SEG consecutive slots per lane, then across the TPL lanes
  if (gl < NL) {
    T* x = line(gl);
    T v[SEG];
    T sum = 0;
#pragma unroll
    for (int q = 0; q < SEG; q++) {
      v[q] = x[2 * (gs * SEG + q) + 1];
      sum += v[q];
    }
    T incl = sum;
    for (int off = 1; off < TPL; off <<= 1) {
      const T o = __shfl_up(incl, off, TPL);
      if (gs >= off) incl += o;
    }
    T run = incl - sum;
#pragma unroll
    for (int q = 0; q < SEG; q++) {
      run += v[q];
      x[2 * (gs * SEG + q) + 1] = run;
    }
  }
  __syncthreads();

  // ---- store
  if (b >= 0) {
    const T* xr = line(l);
    for (int m = t / NL; m < n; m += STEP) MIF_ST_F(&B[b + m * bp0], xr[m]);
  }
}

// Cell-centred Neumann y axis (no counterpart in the reference, whose
// Neumann axes are node-centred DCT-I): N unknowns at the cell centres, the
// ghost reflection phi_{-1} = phi_0, phi_N = phi_{N-1}, so the 1-D operator
// has rows (-1, 1), (1, -2, 1), (1, -1) over h^2.  Its eigenvectors are
// cos(pi k (j + 1/2) / N) with eigenvalues 2 (cos(pi k / N) - 1) / h^2, so
// the transform pair is the DCT-II
//   X_k = sum_j x_j cos(pi k (2j + 1) / (2N)),           k = 0..N-1,
// and the DCT-III
//   y_j = X_0 + 2 sum_{k>=1} X_k cos(pi k (2j + 1) / (2N)) = N x_j.
// Both go through one N-point complex FFT per pair of lines (Makhoul's
// reordering): v_m = x_{2m}, v_{N-1-m} = x_{2m+1} (m < N/2), V = DFT(v),
// X_k = Re(exp(-i pi k / (2N)) V_k); backwards V_k = exp(i pi k / (2N))
// (X_k - i X_{N-k}) with X_N = 0, and the unnormalised inverse DFT of V is N
// times v.  Two lines a, b share the FFT as v^a + i v^b.  H holds
// (cos, sin)(pi k / (2N)), k < N.
template <typename T, int N, int NP, int INV>
__global__ void __launch_bounds__(256) k_dct2_y(DGrid g, const T* __restrict__ W, const T* __restrict__ H,
                                                T* __restrict__ B, long bp0, long bplane) {
  constexpr int LP = p2::lds_pitch(N);
  constexpr int NL = 2 * NP;
  constexpr int STEP = 256 / NL;
  constexpr int MQ = (N + STEP - 1) / STEP;
  constexpr int TWN = p2::tw_count(N);
  static_assert((NP * LP + TWN + N) * sizeof(Cx<T>) <= kLdsMax, "workgroup LDS above the limit");
  constexpr int NI = NP * N;  // post/pre items: (pair, k)
  constexpr int PER = (NI + 255) / 256;
  __shared__ Cx<T> Z[NP][LP];
  __shared__ Cx<T> tws[TWN];  // exp(-2 pi i e / N), e < TWN
  __shared__ Cx<T> hc[N];     // (cos, sin)(pi k / (2N))
  const int t = threadIdx.x;
  int bx, by;
  {
    const int tile = xcd_tile(blockIdx.x + gridDim.x * blockIdx.y, gridDim.x * gridDim.y);
    bx = tile % gridDim.x;
    by = tile / gridDim.x;
  }
  for (int e = t; e < TWN; e += 256) tws[e] = Cx<T>{W[2 * e], W[2 * e + 1]};
  for (int k = t; k < N; k += 256) hc[k] = Cx<T>{H[2 * k], H[2 * k + 1]};
  // line l: pair l % NP, real part (l < NP) or imaginary part; thread t
  // serves line t % NL, rows t / NL + q * STEP
  const int l = t % NL;
  const int i = bx * NL + l;
  const long b = i < g.npts[0] ? i + bplane * by : -1;
  auto line = [&](int ll) -> T* { return (T*)&Z[ll % NP][0] + (ll < NP ? 0 : LP); };
  T val[MQ];
#pragma unroll
  for (int q = 0; q < MQ; q++) {
    const int m = t / NL + q * STEP;
    const bool okm = b >= 0 && m < N;
    const T x = B[okm ? b + m * bp0 : 0];
    val[q] = okm ? x : (T)0;
  }
  __syncthreads();  // twiddle tables staged
  if (!INV) {
    // x_j -> v at the reordered position
#pragma unroll
    for (int q = 0; q < MQ; q++) {
      const int j = t / NL + q * STEP;
      if (j < N) {
        const int m = (j & 1) ? N - 1 - (j >> 1) : (j >> 1);
        ((T*)&Z[l % NP][pd(m)])[l < NP ? 0 : 1] = val[q];
      }
    }
    __syncthreads();
    p2::fft<T, N, NP, -1>(Z, tws);
    // X_k of both lines, straight to the packed buffer
#pragma unroll
    for (int r = 0; r < PER; r++) {
      const int it = t + 256 * r;
      if (NI % 256 == 0 || it < NI) {
        const int p = it % NP, k = it / NP;
        const Cx<T> A = Z[p][pd(k)], Bk = Z[p][pd(k == 0 ? 0 : N - k)];
        const T c = hc[k].re, sn = hc[k].im;
        const T xa = c * (A.re + Bk.re) * (T)0.5 + sn * (A.im - Bk.im) * (T)0.5;
        const T xb = c * (A.im + Bk.im) * (T)0.5 + sn * (Bk.re - A.re) * (T)0.5;
        const long b1 = bx * NL + p < g.npts[0] ? bx * NL + p + bplane * by : -1;
        const long b2 = bx * NL + p + NP < g.npts[0] ? bx * NL + p + NP + bplane * by : -1;
        if (b1 >= 0) MIF_ST_F(&B[b1 + k * bp0], xa);
        if (b2 >= 0) MIF_ST_F(&B[b2 + k * bp0], xb);
      }
    }
  } else {
    // X of each line as a real array overlaying its pair's buffer
    T* xr = line(l);
#pragma unroll
    for (int q = 0; q < MQ; q++) {
      const int m = t / NL + q * STEP;
      if (m < N) xr[m] = val[q];
    }
    __syncthreads();
    Cx<T> zk[PER];
#pragma unroll
    for (int r = 0; r < PER; r++) {
      const int it = t + 256 * r;
      if (NI % 256 == 0 || it < NI) {
        const int p = it % NP, k = it / NP;
        const T* xa = line(p);
        const T* xb = line(p + NP);
        const T c = hc[k].re, sn = hc[k].im;
        const T a0 = xa[k], a1 = k == 0 ? (T)0 : xa[N - k];
        const T b0 = xb[k], b1 = k == 0 ? (T)0 : xb[N - k];
        // V = (c + i s)(X_k - i X_{N-k}); Z = V^a + i V^b
        const Cx<T> va = {c * a0 + sn * a1, sn * a0 - c * a1};
        const Cx<T> vb = {c * b0 + sn * b1, sn * b0 - c * b1};
        zk[r] = Cx<T>{va.re - vb.im, va.im + vb.re};
      }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < PER; r++) {
      const int it = t + 256 * r;
      if (NI % 256 == 0 || it < NI) Z[it % NP][pd(it / NP)] = zk[r];
    }
    __syncthreads();
    p2::fft<T, N, NP, +1>(Z, tws);
    // x_j = v at the reordered position (N times the line)
    if (b >= 0)
      for (int j = t / NL; j < N; j += STEP) {
        const int m = (j & 1) ? N - 1 - (j >> 1) : (j >> 1);
        const Cx<T> z = Z[l % NP][pd(m)];
        MIF_ST_F(&B[b + j * bp0], l < NP ? z.re : z.im);
      }
  }
}

template <typename T, int N, int NP>
static void launch_dct2_y(const DGrid& g, const FftPlan& plan, int inverse, T* B, long bp0, long bplane,
                          hipStream_t st) {
  const T* W = sizeof(T) == 8 ? (const T*)plan.W64 : (const T*)plan.W32;
  const T* H = sizeof(T) == 8 ? (const T*)plan.H64 : (const T*)plan.H32;
  const dim3 grd((unsigned)((g.npts[0] + 2 * NP - 1) / (2 * NP)), (unsigned)g.npts[2], 1);
  if (inverse)
    k_dct2_y<T, N, NP, 1><<<grd, 256, 0, st>>>(g, W, H, B, bp0, bplane);
  else
    k_dct2_y<T, N, NP, 0><<<grd, 256, 0, st>>>(g, W, H, B, bp0, bplane);
}

// The cell-centred y pass; false for lengths without a kernel.
template <typename T>
bool launch_dct2_y_any(const DGrid& g, const FftPlan& plan, int inverse, T* B, long bp0, long bplane,
                       hipStream_t st) {
  constexpr bool d = sizeof(T) == 8;
  switch (plan.L) {
    case 64:
      launch_dct2_y<T, 64, (d ? 8 : 16)>(g, plan, inverse, B, bp0, bplane, st);
      return true;
    case 128:
      launch_dct2_y<T, 128, (d ? 8 : 8)>(g, plan, inverse, B, bp0, bplane, st);
      return true;
    case 256:
      launch_dct2_y<T, 256, (d ? 4 : 8)>(g, plan, inverse, B, bp0, bplane, st);
      return true;
    case 512:
      launch_dct2_y<T, 512, (d ? 2 : 4)>(g, plan, inverse, B, bp0, bplane, st);
      return true;
    case 1024:
      launch_dct2_y<T, 1024, (d ? 1 : 2)>(g, plan, inverse, B, bp0, bplane, st);
      return true;
    default:
      return false;
  }
}

template bool launch_dct2_y_any<double>(const DGrid&, const FftPlan&, int, double*, long, long, hipStream_t);
template bool launch_dct2_y_any<float>(const DGrid&, const FftPlan&, int, float*, long, long, hipStream_t);

// Line pairs per 256-thread workgroup, per element type, length and pass:
// 2*NP adjacent lines, i.e. runs of 2*NP*sizeof(T) contiguous bytes on the
// strided (y, z) passes; the LDS footprint grows with NP*L, and fewer pairs
// per workgroup give more workgroups in flight.  The x pass reads whole
// rows, so it takes the smallest NP that keeps 256 threads busy.
// (Chosen by measurement on the 256^3 and 128^3 channels.)
constexpr int np_strided(bool d, int L) {
  switch (L) {
    case 64: return 16;
    case 128: return d ? 8 : 16;
    case 256: return d ? 4 : 8;
    // FP64 too: 256^3 y pass 71.7 -> 68.4 us, 512^3 z pass 660 -> 527 us
    case 512: return 4;
    case 768: return d ? 2 : 4;
    case 1024: return d ? 1 : 2;
    case 1536: return d ? 1 : 2;
    // lengths 3 * 2^k below 768, as their power-of-two neighbours
    case 96: return d ? 8 : 16;
    case 192: return d ? 4 : 8;
    case 384: return d ? 2 : 4;
    default: return 1;
  }
}
// x pass (lengths from 192 on; below, one kernel serves all three passes)
constexpr int np_x(bool d, int L) {
  switch (L) {
    case 192: case 256: case 384: return d ? 2 : 4;
    case 512: return d ? 2 : 4;
    case 768: return d ? 1 : 2;
    default: return np_strided(d, L);
  }
}
// x inverse passes of length 256: four pairs in FP64 too (256^3 FP64 x
// inverse + unpack 54.0 -> 49.5 us; the forward + divergence pass measured
// 110.0 -> 113.5 us with four, so it keeps two)
constexpr int np_x_inv(bool d, int L) { return L == 256 ? 4 : np_x(d, L); }
// the half-length Neumann y pass (k_dct1_y, N = L / 2)
constexpr int np_dct1(bool d, int N) { return N >= 1024 ? (d ? 1 : 2) : (d ? 2 : 4); }

template <typename T, int N, int NP>
static void launch_dct1_y(const DGrid& g, const FftPlan& plan, T* B, long bp0, long bplane, hipStream_t st) {
  const T* W = sizeof(T) == 8 ? (const T*)plan.W64 : (const T*)plan.W32;
  const dim3 grd((unsigned)((g.npts[0] + 2 * NP - 1) / (2 * NP)), (unsigned)g.npts[2], 1);
  k_dct1_y<T, N, NP><<<grd, 256, 0, st>>>(g, W, B, bp0, bplane);
}

// The half-length Neumann y pass for lines of N + 1 = 513, 769 and 1025
// points; false otherwise.  Below that the 2N-point kernel measured as fast
// or faster (256^3 channel: equal in FP64, 8% faster in FP32), while from
// 513 points on the halved LDS footprint doubles the lines per workgroup
// (512^3 FP64 y pass 926 -> 683 us, 768^3 mixed 4153 -> 2558 us).
template <typename T>
static bool launch_dct1_y_any(const DGrid& g, const FftPlan& plan, T* B, long bp0, long bplane, hipStream_t st) {
  constexpr bool d = sizeof(T) == 8;
  switch (plan.L / 2) {
    case 512:
      launch_dct1_y<T, 512, np_dct1(d, 512)>(g, plan, B, bp0, bplane, st);
      return true;
    case 768:
      launch_dct1_y<T, 768, np_dct1(d, 768)>(g, plan, B, bp0, bplane, st);
      return true;
    case 1024:
      launch_dct1_y<T, 1024, np_dct1(d, 1024)>(g, plan, B, bp0, bplane, st);
      return true;
    default:
      return false;
  }
}

// PASSES: 1 = the x pass only, 2 = the y and z passes only, 3 = all.
template <typename T, int L, int NP, int PASSES = 3>
static void launch_p2(const DGrid& g, int axis, int mode, const FftPlan& plan, const SpectralScale& sc, T* B, long bp0,
                      long bplane, const FusedIO* io, hipStream_t st) {
  FusedIO none{};
  const T* W;
  if constexpr (sizeof(T) == 8)
    W = (const T*)plan.W64;
  else
    W = (const T*)plan.W32;
  dim3 grd;
  if (axis == 0) {
    const long lines = (long)g.npts[1] * g.npts[2];
    grd = dim3((unsigned)((lines + 2 * NP - 1) / (2 * NP)), 1, 1);
  } else {
    grd = dim3((unsigned)((g.npts[0] + 2 * NP - 1) / (2 * NP)), axis == 1 ? g.npts[2] : g.npts[1], 1);
  }
  const int per = plan.periodic;
#define MIF_P2(AX, MD) k_fft_p2<T, L, NP, AX, MD, false><<<grd, 256, 0, st>>>(g, W, sc, B, bp0, bplane, per, none)
  if constexpr (PASSES & 1) {
    if (axis == 0 && io && io->res) {
      if constexpr (sizeof(T) == 4)
        k_fft_p2<T, L, NP, 0, 0, false, T, false, true><<<grd, 256, 0, st>>>(g, W, sc, B, bp0, bplane, per, *io);
    } else if (axis == 0 && io && io->dc) {
      if constexpr (sizeof(T) == 4) {
        if (mode == 0)
          k_fft_p2<T, L, NP, 0, 0, true, T, true><<<grd, 256, 0, st>>>(g, W, sc, B, bp0, bplane, per, *io);
        else
          k_fft_p2<T, L, NP, 0, 1, true, T, true><<<grd, 256, 0, st>>>(g, W, sc, B, bp0, bplane, per, *io);
      }
    } else if (axis == 0 && io && sizeof(T) == 8 && io->fields_fp32) {
      if (mode == 0)
        k_fft_p2<T, L, NP, 0, 0, true, float><<<grd, 256, 0, st>>>(g, W, sc, B, bp0, bplane, per, *io);
      else
        k_fft_p2<T, L, NP, 0, 1, true, float><<<grd, 256, 0, st>>>(g, W, sc, B, bp0, bplane, per, *io);
    } else if (axis == 0 && io) {
      if (mode == 0)
        k_fft_p2<T, L, NP, 0, 0, true><<<grd, 256, 0, st>>>(g, W, sc, B, bp0, bplane, per, *io);
      else
        k_fft_p2<T, L, NP, 0, 1, true><<<grd, 256, 0, st>>>(g, W, sc, B, bp0, bplane, per, *io);
    } else if (axis == 0) {
      if (mode == 0) MIF_P2(0, 0);
      else MIF_P2(0, 1);
    }
  }
  if constexpr (PASSES & 2) {
    if (axis == 1) {
      if (mode == 0) MIF_P2(1, 0);
      else MIF_P2(1, 1);
    } else if (axis == 2) {
      if (mode == 0) MIF_P2(2, 0);
      else if (mode == 1) MIF_P2(2, 1);
      else MIF_P2(2, 2);
    }
  }
#undef MIF_P2
}

bool fft_fast_len(int L) {
  switch (L) {
    case 64: case 128: case 256: case 512: case 1024: case 96: case 192: case 384: case 768: case 1536:
      return true;
    default:
      return false;
  }
}

// Returns false when the plan's length has no specialised kernel.
template <typename T>
bool launch_fft_pow2(const DGrid& g, int axis, int mode, const FftPlan& plan, const SpectralScale& sc, T* B, long bp0,
                     long bplane, const FusedIO* io, hipStream_t st) {
  constexpr bool d = sizeof(T) == 8;
  // Neumann y: both directions are the same DCT-I
  if (axis == 1 && !plan.periodic && mode != 2 &&
      launch_dct1_y_any<T>(g, plan, B, bp0, bplane, st))
    return true;
  switch (plan.L) {
    case 64:
      launch_p2<T, 64, np_strided(d, 64)>(g, axis, mode, plan, sc, B, bp0, bplane, io, st);
      return true;
    case 128:
      launch_p2<T, 128, np_strided(d, 128)>(g, axis, mode, plan, sc, B, bp0, bplane, io, st);
      return true;
    case 256:
      if (axis == 0 && mode == 1)
        launch_p2<T, 256, np_x_inv(d, 256), 1>(g, axis, mode, plan, sc, B, bp0, bplane, io, st);
      else if (axis == 0)
        launch_p2<T, 256, np_x(d, 256), 1>(g, axis, mode, plan, sc, B, bp0, bplane, io, st);
      else
        launch_p2<T, 256, np_strided(d, 256), 2>(g, axis, mode, plan, sc, B, bp0, bplane, io, st);
      return true;
    case 512:
      if (axis == 0)
        launch_p2<T, 512, np_x(d, 512), 1>(g, axis, mode, plan, sc, B, bp0, bplane, io, st);
      else
        launch_p2<T, 512, np_strided(d, 512), 2>(g, axis, mode, plan, sc, B, bp0, bplane, io, st);
      return true;
    case 1024:
      if (axis == 0)
        launch_p2<T, 1024, 1, 1>(g, axis, mode, plan, sc, B, bp0, bplane, io, st);
      else
        launch_p2<T, 1024, np_strided(d, 1024), 2>(g, axis, mode, plan, sc, B, bp0, bplane, io, st);
      return true;
    case 96:
      launch_p2<T, 96, np_strided(d, 96)>(g, axis, mode, plan, sc, B, bp0, bplane, io, st);
      return true;
    case 192:
      if (axis == 0)
        launch_p2<T, 192, np_x(d, 192), 1>(g, axis, mode, plan, sc, B, bp0, bplane, io, st);
      else
        launch_p2<T, 192, np_strided(d, 192), 2>(g, axis, mode, plan, sc, B, bp0, bplane, io, st);
      return true;
    case 384:
      if (axis == 0)
        launch_p2<T, 384, np_x(d, 384), 1>(g, axis, mode, plan, sc, B, bp0, bplane, io, st);
      else
        launch_p2<T, 384, np_strided(d, 384), 2>(g, axis, mode, plan, sc, B, bp0, bplane, io, st);
      return true;
    case 768:
      if (axis == 0)
        launch_p2<T, 768, np_x(d, 768), 1>(g, axis, mode, plan, sc, B, bp0, bplane, io, st);
      else
        launch_p2<T, 768, np_strided(d, 768), 2>(g, axis, mode, plan, sc, B, bp0, bplane, io, st);
      return true;
    case 1536:
      if (axis == 0)
        launch_p2<T, 1536, 1, 1>(g, axis, mode, plan, sc, B, bp0, bplane, io, st);
      else
        launch_p2<T, 1536, np_strided(d, 1536), 2>(g, axis, mode, plan, sc, B, bp0, bplane, io, st);
      return true;
    default:
      return false;
  }
}

// Workgroups of the x pass launch_fft_pow2 runs for a plan (its line pairs
// per workgroup follow the same table); 0 when the length has no kernel.
int fused_x_blocks(const DGrid& g, const FftPlan& plan, bool fp64) {
  int np;
  switch (plan.L) {
    case 64: np = np_strided(fp64, 64); break;
    case 128: np = np_strided(fp64, 128); break;
    case 96: np = np_strided(fp64, 96); break;
    case 256: np = np_x(fp64, 256); break;
    case 512: np = np_x(fp64, 512); break;
    case 192: np = np_x(fp64, 192); break;
    case 384: np = np_x(fp64, 384); break;
    case 768: np = np_x(fp64, 768); break;
    case 1024: case 1536: np = 1; break;
    default: return 0;
  }
  const long lines = (long)g.npts[1] * g.npts[2];
  return (int)((lines + 2 * np - 1) / (2 * np));
}

template bool launch_fft_pow2<double>(const DGrid&, int, int, const FftPlan&, const SpectralScale&, double*, long, long,
                                      const FusedIO*, hipStream_t);
template bool launch_fft_pow2<float>(const DGrid&, int, int, const FftPlan&, const SpectralScale&, float*, long, long,
                                     const FusedIO*, hipStream_t);

}  // namespace mif
