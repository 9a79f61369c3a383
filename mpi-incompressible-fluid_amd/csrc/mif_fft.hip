// mif_fft.hip -- LDS transform kernels of the spectral Poisson solve.
//
// The reference solves L phi = b by diagonalisation: a real transform along
// every axis (FFTW_REDFT00 = DCT-I on Neumann axes, FFTW_R2HC / FFTW_HC2R on
// periodic axes), multiplication by 1/(lx+ly+lz) with the constant mode
// zeroed, and the inverse transforms with a division by 2(N-1) or N after
// each axis (src/PressureEquation.cpp:65-264, src/PressureSolverStructures.
// cpp:5-69).  Between axes it transposes the data with 2DECOMP so that the
// active axis is contiguous (deps/2Decomp_C/Transpose*.cpp).
//
// Here every axis is transformed in place in the packed x-fastest buffer:
// an x pass reads whole rows, a y or z pass reads a tile of 2P adjacent x
// positions times the full line, so the loads stay coalesced and no
// transpose pass is needed.  Inside a workgroup, two real lines are packed
// into one complex line and transformed by a mixed-radix Stockham FFT held
// in LDS (ping-pong buffers, radix 2/4/8 butterflies in registers, any other
// factor by direct sums).  Outputs keep FFTW's definitions and slot order
// (halfcomplex r0..r_{n/2}, i_{(n+1)/2-1}..i_1 on periodic axes, DCT-I
// coefficients on Neumann axes), so the eigenvalue of slot k is the
// reference's compute_eigenvalue_{periodic,neumann}(k).  The z pass fuses
// forward transform, scaling (with all three inverse normalisations folded
// in) and inverse transform.
#include "mif_kernels.h"

namespace mif {

template <typename T>
struct Cx {
  T re, im;
};

template <typename T>
__device__ inline Cx<T> cadd(Cx<T> a, Cx<T> b) {
  return {a.re + b.re, a.im + b.im};
}
template <typename T>
__device__ inline Cx<T> csub(Cx<T> a, Cx<T> b) {
  return {a.re - b.re, a.im - b.im};
}
template <typename T>
__device__ inline Cx<T> cmul(Cx<T> a, Cx<T> b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
// multiply by sign * i  (sign = -1: -i, sign = +1: +i)
template <typename T>
__device__ inline Cx<T> cmul_si(Cx<T> a, int sign) {
  return sign < 0 ? Cx<T>{a.im, -a.re} : Cx<T>{-a.im, a.re};
}

template <typename T>
__device__ inline Cx<T> twiddle(const FftPlan& P, long e, int sign);
template <>
__device__ inline Cx<double> twiddle<double>(const FftPlan& P, long e, int sign) {
  const double re = P.W64[2 * e], im = P.W64[2 * e + 1];
  return {re, sign < 0 ? im : -im};
}
template <>
__device__ inline Cx<float> twiddle<float>(const FftPlan& P, long e, int sign) {
  const float re = P.W32[2 * e], im = P.W32[2 * e + 1];
  return {re, sign < 0 ? im : -im};
}

// In-register DFT of size R (R = 2, 4, 8 specialised; others by direct sums
// with table twiddles).
template <typename T, int R>
__device__ inline void dft_reg(Cx<T>* v, int sign, const FftPlan& P) {
  if constexpr (R == 2) {
    const Cx<T> a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  } else if constexpr (R == 4) {
    const Cx<T> t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
    const Cx<T> t2 = cadd(v[1], v[3]), t3 = cmul_si(csub(v[1], v[3]), sign);
    v[0] = cadd(t0, t2);
    v[2] = csub(t0, t2);
    v[1] = cadd(t1, t3);
    v[3] = csub(t1, t3);
  } else if constexpr (R == 8) {
    Cx<T> e[4] = {v[0], v[2], v[4], v[6]};
    Cx<T> o[4] = {v[1], v[3], v[5], v[7]};
    dft_reg<T, 4>(e, sign, P);
    dft_reg<T, 4>(o, sign, P);
    const T s = (T)0.70710678118654752440;
    // o[k] *= w8^k, w8 = exp(sign * 2 pi i / 8)
    const Cx<T> o1 = cmul(o[1], Cx<T>{s, (T)sign * s});
    const Cx<T> o2 = cmul_si(o[2], sign);
    const Cx<T> o3 = cmul(o[3], Cx<T>{-s, (T)sign * s});
    v[0] = cadd(e[0], o[0]);
    v[4] = csub(e[0], o[0]);
    v[1] = cadd(e[1], o1);
    v[5] = csub(e[1], o1);
    v[2] = cadd(e[2], o2);
    v[6] = csub(e[2], o2);
    v[3] = cadd(e[3], o3);
    v[7] = csub(e[3], o3);
  } else {
    Cx<T> out[R];
    const int step = P.L / R;
    for (int r = 0; r < R; r++) {
      Cx<T> acc = v[0];
      for (int q = 1; q < R; q++) acc = cadd(acc, cmul(v[q], twiddle<T>(P, (long)((r * q) % R) * step, sign)));
      out[r] = acc;
    }
    for (int r = 0; r < R; r++) v[r] = out[r];
  }
}

// One Stockham stage of radix R on P line-pairs: src -> dst.
template <typename T, int R>
__device__ inline void stage_reg(const FftPlan& P, const Cx<T>* src, Cx<T>* dst, int npairs, int Ns, int sign) {
  const int L = P.L, m = L / R;
  const int tstep = L / (Ns * R);
  const int items = npairs * m;
  for (int it = threadIdx.x; it < items; it += blockDim.x) {
    const int p = it / m, jb = it - p * m;
    const Cx<T>* s = src + (long)p * L;
    Cx<T>* d = dst + (long)p * L;
    const int k = jb % Ns;
    Cx<T> v[R];
#pragma unroll
    for (int r = 0; r < R; r++) v[r] = s[jb + r * m];
    if (Ns > 1) {
#pragma unroll
      for (int r = 1; r < R; r++) v[r] = cmul(v[r], twiddle<T>(P, (long)r * k * tstep, sign));
    }
    dft_reg<T, R>(v, sign, P);
    const int idx = (jb / Ns) * Ns * R + k;
#pragma unroll
    for (int r = 0; r < R; r++) d[idx + r * Ns] = v[r];
  }
}

// Stage of an arbitrary radix by direct sums (one output per work item).
template <typename T>
__device__ inline void stage_generic(const FftPlan& P, const Cx<T>* src, Cx<T>* dst, int npairs, int Ns, int R,
                                     int sign) {
  const int L = P.L, m = L / R;
  const long tstep = L / (Ns * R);
  const int items = npairs * L;
  for (int it = threadIdx.x; it < items; it += blockDim.x) {
    const int p = it / L, rem = it - p * L;
    const int jb = rem / R, ro = rem - jb * R;
    const Cx<T>* s = src + (long)p * L;
    Cx<T>* d = dst + (long)p * L;
    const int k = jb % Ns;
    const long ex = (long)(k + ro * Ns) * tstep;
    Cx<T> acc = s[jb];
    for (int q = 1; q < R; q++) acc = cadd(acc, cmul(s[jb + q * m], twiddle<T>(P, (q * ex) % L, sign)));
    const int idx = (jb / Ns) * Ns * R + k;
    d[idx + ro * Ns] = acc;
  }
}

// Full complex FFT of npairs lines of length L; returns the buffer holding
// the result (a or b).
template <typename T>
__device__ inline Cx<T>* fft_lds(const FftPlan& P, Cx<T>* a, Cx<T>* b, int npairs, int sign) {
  int Ns = 1;
  for (int s = 0; s < P.nst; s++) {
    const int R = P.radix[s];
    if (R == 8)
      stage_reg<T, 8>(P, a, b, npairs, Ns, sign);
    else if (R == 4)
      stage_reg<T, 4>(P, a, b, npairs, Ns, sign);
    else if (R == 2)
      stage_reg<T, 2>(P, a, b, npairs, Ns, sign);
    else if (R == 3)
      stage_reg<T, 3>(P, a, b, npairs, Ns, sign);
    else if (R == 5)
      stage_reg<T, 5>(P, a, b, npairs, Ns, sign);
    else
      stage_generic<T>(P, a, b, npairs, Ns, R, sign);
    __syncthreads();
    Cx<T>* t = a;
    a = b;
    b = t;
    Ns *= R;
  }
  return a;
}

// Pack two real lines (Rr[p][0][:], Rr[p][1][:], length n) into complex lines
// of length L: periodic z = x1 + i x2; Neumann: even extension of both.
template <typename T>
__device__ inline void pack_lines(const FftPlan& P, const T* Rr, Cx<T>* Z, int npairs) {
  const int n = P.n, L = P.L;
  for (int it = threadIdx.x; it < npairs * n; it += blockDim.x) {
    const int p = it / n, m = it - p * n;
    const Cx<T> z = {Rr[(2 * p) * n + m], Rr[(2 * p + 1) * n + m]};
    Z[(long)p * L + m] = z;
    if (!P.periodic && m >= 1 && m <= n - 2) Z[(long)p * L + (L - m)] = z;
  }
}

// Halfcomplex slot k of a real line's spectrum -> complex X_k (any k < n).
template <typename T>
__device__ inline Cx<T> hc_get(const T* h, int n, int k) {
  if (k == 0) return {h[0], (T)0};
  if (2 * k == n) return {h[k], (T)0};
  if (2 * k < n) return {h[k], h[n - k]};
  return {h[n - k], -h[k]};
}

// Forward post-processing: spectra of the two real lines from Z, written to
// Rr in FFTW slot order.
template <typename T>
__device__ inline void unpack_forward(const FftPlan& P, const Cx<T>* Z, T* Rr, int npairs) {
  const int n = P.n, L = P.L;
  for (int it = threadIdx.x; it < npairs * n; it += blockDim.x) {
    const int p = it / n, s = it - p * n;
    const Cx<T>* z = Z + (long)p * L;
    T* h1 = Rr + (2 * p) * n;
    T* h2 = Rr + (2 * p + 1) * n;
    if (!P.periodic) {
      h1[s] = z[s].re;
      h2[s] = z[s].im;
    } else if (2 * s <= n) {
      const int k = s, kk = (L - k) % L;
      h1[s] = (z[k].re + z[kk].re) * (T)0.5;
      h2[s] = (z[k].im + z[kk].im) * (T)0.5;
    } else {
      const int k = n - s, kk = L - k;
      h1[s] = (z[k].im - z[kk].im) * (T)0.5;
      h2[s] = (z[kk].re - z[k].re) * (T)0.5;
    }
  }
}

// Inverse pre-processing: Z_k = X1_k + i X2_k from the two halfcomplex lines
// (periodic) or the even extension (Neumann).
template <typename T>
__device__ inline void pack_inverse(const FftPlan& P, const T* Rr, Cx<T>* Z, int npairs) {
  if (!P.periodic) {
    pack_lines<T>(P, Rr, Z, npairs);
    return;
  }
  const int n = P.n, L = P.L;
  for (int it = threadIdx.x; it < npairs * n; it += blockDim.x) {
    const int p = it / n, k = it - p * n;
    const Cx<T> x1 = hc_get(Rr + (2 * p) * n, n, k);
    const Cx<T> x2 = hc_get(Rr + (2 * p + 1) * n, n, k);
    Z[(long)p * L + k] = {x1.re - x2.im, x1.im + x2.re};
  }
}

// Inverse post-processing: real lines from the complex result.
template <typename T>
__device__ inline void unpack_inverse(const FftPlan& P, const Cx<T>* Z, T* Rr, int npairs) {
  const int n = P.n, L = P.L;
  for (int it = threadIdx.x; it < npairs * n; it += blockDim.x) {
    const int p = it / n, m = it - p * n;
    const Cx<T> z = Z[(long)p * L + m];
    Rr[(2 * p) * n + m] = z.re;
    Rr[(2 * p + 1) * n + m] = z.im;
  }
}

// Line addressing: on axis 0 the lines are rows (j,k) and pair q holds rows
// 2q and 2q+1 of the block; on axes 1 and 2 a block covers 2*npairs adjacent
// x positions at a fixed outer index (k for axis 1, j for axis 2) and pair p
// holds x positions i0+p and i0+p+npairs.
template <typename T, int AXIS, int MODE>
__global__ void __launch_bounds__(256) k_fft(DGrid g, FftPlan P, SpectralScale sc, T* __restrict__ B, long bp0,
                                             long bplane, int npairs) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int n = P.n, L = P.L;
  Cx<T>* bufA = (Cx<T>*)smem;
  Cx<T>* bufB = bufA + (long)npairs * L;
  const int nl = 2 * npairs;
  const int n0 = g.npts[0], n1 = g.npts[1], n2 = g.npts[2];

  // element address of (line l, position m); -1 if the line is padding
  auto addr = [&](int l, int m) -> long {
    if (AXIS == 0) {
      const int r = blockIdx.x * nl + l;
      if (r >= n1 * n2) return -1;
      const int j = r % n1, k = r / n1;
      return m + bp0 * j + bplane * k;
    } else {
      const int p = l % npairs, part = l / npairs;
      const int i = blockIdx.x * nl + p + part * npairs;
      if (i >= n0) return -1;
      const int outer = blockIdx.y;
      return AXIS == 1 ? i + bp0 * m + bplane * outer : i + bp0 * outer + bplane * m;
    }
  };
  // staging index of (line l, position m): pair-major, part, position
  auto sidx = [&](int l, int m) -> int {
    if (AXIS == 0) return l * n + m;  // l = 2p + part
    const int p = l % npairs, part = l / npairs;
    return (2 * p + part) * n + m;
  };

  // load: staging reals live in bufB
  T* Rs = (T*)bufB;
  for (int it = threadIdx.x; it < nl * n; it += blockDim.x) {
    int l, m;
    if (AXIS == 0) {
      l = it / n;
      m = it - l * n;
    } else {
      m = it / nl;
      l = it - m * nl;
    }
    const long a = addr(l, m);
    Rs[sidx(l, m)] = a >= 0 ? B[a] : (T)0;
  }
  __syncthreads();
  const int sign = MODE == 1 ? +1 : -1;
  if (MODE == 1)
    pack_inverse<T>(P, Rs, bufA, npairs);
  else
    pack_lines<T>(P, Rs, bufA, npairs);
  __syncthreads();
  Cx<T>* res = fft_lds<T>(P, bufA, bufB, npairs, sign);
  Cx<T>* other = res == bufA ? bufB : bufA;
  T* Ro = (T*)other;
  if (MODE == 1)
    unpack_inverse<T>(P, res, Ro, npairs);
  else
    unpack_forward<T>(P, res, Ro, npairs);
  __syncthreads();
  if (MODE == 2) {
    // scale by 1/((lx+ly+lz) * norm) with the constant mode removed
    // (src/PressureEquation.cpp:158-163); only AXIS 2 runs fused
    const int j = blockIdx.y;
    for (int it = threadIdx.x; it < nl * n; it += blockDim.x) {
      const int l = it / n, k = it - l * n;
      const int p = l % npairs, part = l / npairs;
      const int i = blockIdx.x * nl + p + part * npairs;
      if (i >= n0) continue;
      const double lam = (sc.lam[0][i] + sc.lam[1][j]) + sc.lam[2][k];
      const double s = (i == 0 && j == 0 && k == 0) ? 0.0 : (1.0 / lam) * sc.inv_norm;
      Ro[(2 * p + part) * n + k] *= (T)s;
    }
    __syncthreads();
    pack_inverse<T>(P, Ro, res, npairs);
    __syncthreads();
    Cx<T>* res2 = fft_lds<T>(P, res, other, npairs, +1);
    Cx<T>* other2 = res2 == res ? other : res;
    Ro = (T*)other2;
    unpack_inverse<T>(P, res2, Ro, npairs);
    __syncthreads();
  }
  // store
  for (int it = threadIdx.x; it < nl * n; it += blockDim.x) {
    int l, m;
    if (AXIS == 0) {
      l = it / n;
      m = it - l * n;
    } else {
      m = it / nl;
      l = it - m * nl;
    }
    const long a = addr(l, m);
    if (a >= 0) B[a] = Ro[sidx(l, m)];
  }
}

static int pairs_for(int L, size_t csize) {
  const size_t budget = 64 * 1024;
  int np = (int)(budget / (2 * (size_t)L * csize));
  if (np < 1) np = 1;
  if (np > 16) np = 16;
  return np;
}

template <typename T, int AXIS, int MODE>
static void launch_one(const DGrid& g, const FftPlan& plan, const SpectralScale& sc, T* B, long bp0, long bplane,
                       hipStream_t st) {
  const int np = pairs_for(plan.L, sizeof(Cx<T>));
  const size_t lds = 2 * (size_t)np * plan.L * sizeof(Cx<T>);
  dim3 grd;
  if (AXIS == 0) {
    const long lines = (long)g.npts[1] * g.npts[2];
    grd = dim3((unsigned)((lines + 2 * np - 1) / (2 * np)), 1, 1);
  } else {
    grd = dim3((unsigned)((g.npts[0] + 2 * np - 1) / (2 * np)), AXIS == 1 ? g.npts[2] : g.npts[1], 1);
  }
  k_fft<T, AXIS, MODE><<<grd, 256, lds, st>>>(g, plan, sc, B, bp0, bplane, np);
}

template <typename T>
void launch_fft_axis(const DGrid& g, int axis, int dir, int fused, const FftPlan& plan, const SpectralScale& sc, T* B,
                     long bp0, long bplane, hipStream_t st) {
  const int mode = fused ? 2 : (dir < 0 ? 0 : 1);
  if (plan.dct2) {
    // cell-centred axes: y only, lengths with a kernel (checked when the
    // context is created)
    (void)launch_dct2_y_any<T>(g, plan, dir > 0, B, bp0, bplane, st);
    return;
  }
  if (plan.L >= 2 && launch_fft_pow2<T>(g, axis, mode, plan, sc, B, bp0, bplane, nullptr, st)) return;
  if (axis == 0) {
    if (mode == 0) launch_one<T, 0, 0>(g, plan, sc, B, bp0, bplane, st);
    else launch_one<T, 0, 1>(g, plan, sc, B, bp0, bplane, st);
  } else if (axis == 1) {
    if (mode == 0) launch_one<T, 1, 0>(g, plan, sc, B, bp0, bplane, st);
    else launch_one<T, 1, 1>(g, plan, sc, B, bp0, bplane, st);
  } else {
    if (mode == 0) launch_one<T, 2, 0>(g, plan, sc, B, bp0, bplane, st);
    else if (mode == 1) launch_one<T, 2, 1>(g, plan, sc, B, bp0, bplane, st);
    else launch_one<T, 2, 2>(g, plan, sc, B, bp0, bplane, st);
  }
}

template void launch_fft_axis<double>(const DGrid&, int, int, int, const FftPlan&, const SpectralScale&, double*, long,
                                      long, hipStream_t);
template void launch_fft_axis<float>(const DGrid&, int, int, int, const FftPlan&, const SpectralScale&, float*, long,
                                     long, hipStream_t);

}  // namespace mif
