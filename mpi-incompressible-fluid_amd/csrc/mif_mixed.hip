// mif_mixed.hip -- FP64 pieces of the mixed-precision pressure solve.
//
// On a cell-centred Neumann axis (stand-alone Poisson solves) the weights
// are uniform and the ghost mirrors the first/last cell (phi_{-1} = phi_0).
//
// The mixed form stores and advances the velocity in FP32 and runs the
// spectral transforms in FP32.  What stays FP64 (no counterpart in the
// reference, which is uniformly FP64 or uniformly FP32, include/Real.h:9-17):
//   * the right-hand side b = div(u)/dt, and its projection onto the range of
//     the Laplacian: b -= (W.b)/(W.1) with W the trapezoid weights (1/2 at
//     the two end nodes of every Neumann axis), the left null vector of the
//     node Neumann / periodic operator the solve inverts;
//   * defect correction: phi accumulated in FP64, r = b - L phi with the exact
//     7-point operator (Neumann rows reflect the ghost: phi_{-1} = phi_1,
//     phi_n = phi_{n-2}; periodic rows wrap), FP32 re-solve of L d = r,
//     phi += d;
//   * the final null-space projection of phi (zero weighted mean, the
//     normalisation of the reference's zeroed constant mode).
#include "mif_kernels.h"

namespace mif {

__device__ inline double axis_weight(const DGrid& g, int a, int i) { return owner_weight(g, a, i); }

// Partial sums of (W.b, W.1) over the packed buffer.
__global__ void __launch_bounds__(256) k_wsum(DGrid g, const double* __restrict__ B, long bp0, long bplane,
                                              Red3* partial) {
  const int n0 = g.npts[0], n1 = g.npts[1], n2 = g.npts[2];
  const int total = n0 * n1 * n2;
  double s = 0, sw = 0;
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < total; q += gridDim.x * blockDim.x) {
    const int rq = q / n0;
    const int i = q % n0, j = rq % n1, k = rq / n1;
    const double w = axis_weight(g, 0, i) * axis_weight(g, 1, j) * axis_weight(g, 2, k);
    s += w * B[i + bp0 * j + bplane * k];
    sw += w;
  }
  __shared__ double sh[2][4];
  for (int off = 32; off > 0; off >>= 1) {
    s += __shfl_down(s, off, 64);
    sw += __shfl_down(sw, off, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    sh[0][threadIdx.x >> 6] = s;
    sh[1][threadIdx.x >> 6] = sw;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    Red3 r{0, 0, 0};
    for (int q = 0; q < 4; q++) {
      r.s1 += sh[0][q];
      r.s2 += sh[1][q];
    }
    partial[blockIdx.x] = r;
  }
}

// b -= mean (mean = total.s1 / total.s2 read on the device); B32 = (float) b.
__global__ void __launch_bounds__(256) k_sub_mean_to32(DGrid g, double* __restrict__ B, float* __restrict__ B32,
                                                       long bp0, long bplane, const Red3* __restrict__ total) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y, k = blockIdx.z;
  if (i >= g.npts[0]) return;
  const double mean = total->s1 / total->s2;
  const long o = i + bp0 * j + bplane * k;
  const double v = B[o] - mean;
  B[o] = v;
  if (B32) B32[o] = (float)v;
}

__global__ void __launch_bounds__(256) k_to64(DGrid g, const float* __restrict__ B32, double* __restrict__ P, long bp0,
                                              long bplane) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y, k = blockIdx.z;
  if (i >= g.npts[0]) return;
  const long o = i + bp0 * j + bplane * k;
  P[o] = (double)B32[o];
}

// B32 = (float)(b - L phi) in FP64.
__global__ void __launch_bounds__(256) k_residual(DGrid g, const double* __restrict__ b, const double* __restrict__ P,
                                                  float* __restrict__ B32, long bp0, long bplane) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y, k = blockIdx.z;
  if (i >= g.npts[0]) return;
  const long o = i + bp0 * j + bplane * k;
  const double c = P[o];
  const double lx = (P[nb(g, 0, i, -1) + bp0 * j + bplane * k] - 2 * c + P[nb(g, 0, i, 1) + bp0 * j + bplane * k]) /
                    (g.h[0] * g.h[0]);
  const double ly = (P[i + bp0 * nb(g, 1, j, -1) + bplane * k] - 2 * c + P[i + bp0 * nb(g, 1, j, 1) + bplane * k]) /
                    (g.h[1] * g.h[1]);
  const double lz = (P[i + bp0 * j + bplane * nb(g, 2, k, -1)] - 2 * c + P[i + bp0 * j + bplane * nb(g, 2, k, 1)]) /
                    (g.h[2] * g.h[2]);
  B32[o] = (float)(b[o] - (lx + ly + lz));
}

// Subtract total->s1 / total->s2 from every point of a node field.
__global__ void __launch_bounds__(256) k_sub_mean_field(DGrid g, Layout L, float* __restrict__ F,
                                                        const Red3* __restrict__ total) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y, k = blockIdx.z;
  if (i >= g.N[0]) return;
  F[L.idx(i, j, k)] -= (float)(total->s1 / total->s2);
}

__global__ void __launch_bounds__(256) k_accumulate(DGrid g, double* __restrict__ P, const float* __restrict__ B32,
                                                    long bp0, long bplane) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y, k = blockIdx.z;
  if (i >= g.npts[0]) return;
  const long o = i + bp0 * j + bplane * k;
  P[o] += (double)B32[o];
}

// (s1, s2) summed over n per-workgroup partials by one workgroup of
// kReduceThreads threads, each with four independent accumulations in flight
// (256^3: 8192 partials; 256 threads with one chain each took 12 us)
constexpr int kReduceThreads = 1024;
__global__ void __launch_bounds__(kReduceThreads) k_reduce_sum2(const Red3* partial, int n, Red3* out) {
  __shared__ double sh[2][kReduceThreads / 64];
  double s[4] = {0, 0, 0, 0}, sw[4] = {0, 0, 0, 0};
  int q = threadIdx.x;
  for (; q + 3 * kReduceThreads < n; q += 4 * kReduceThreads) {
#pragma unroll
    for (int u = 0; u < 4; u++) {
      s[u] += partial[q + u * kReduceThreads].s1;
      sw[u] += partial[q + u * kReduceThreads].s2;
    }
  }
  for (; q < n; q += kReduceThreads) {
    s[0] += partial[q].s1;
    sw[0] += partial[q].s2;
  }
  double a = (s[0] + s[1]) + (s[2] + s[3]), b = (sw[0] + sw[1]) + (sw[2] + sw[3]);
  for (int off = 32; off > 0; off >>= 1) {
    a += __shfl_down(a, off, 64);
    b += __shfl_down(b, off, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    sh[0][threadIdx.x >> 6] = a;
    sh[1][threadIdx.x >> 6] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double ta = 0, tb = 0;
    for (int w = 0; w < kReduceThreads / 64; w++) {
      ta += sh[0][w];
      tb += sh[1][w];
    }
    *out = Red3{ta, tb, 0};
  }
}

void launch_project_rhs(const DGrid& g, double* B64, float* B32, long bp0, long bplane, Red3* partial, int nblocks,
                        Red3* total, hipStream_t st) {
  k_wsum<<<nblocks, 256, 0, st>>>(g, B64, bp0, bplane, partial);
  k_reduce_sum2<<<1, kReduceThreads, 0, st>>>(partial, nblocks, total);
  dim3 grd((g.npts[0] + 255) / 256, g.npts[1], g.npts[2]);
  k_sub_mean_to32<<<grd, 256, 0, st>>>(g, B64, B32, bp0, bplane, total);
}

// Remove the weighted mean of an FP64 packed array (the constant-mode
// component, which the FP32 solves leave at FP32 rounding level).
void launch_remove_mean(const DGrid& g, double* P, long bp0, long bplane, Red3* partial, int nblocks, Red3* total,
                        hipStream_t st) {
  k_wsum<<<nblocks, 256, 0, st>>>(g, P, bp0, bplane, partial);
  k_reduce_sum2<<<1, kReduceThreads, 0, st>>>(partial, nblocks, total);
  dim3 grd((g.npts[0] + 255) / 256, g.npts[1], g.npts[2]);
  k_sub_mean_to32<<<grd, 256, 0, st>>>(g, P, nullptr, bp0, bplane, total);
}

void launch_to64(const DGrid& g, const float* B32, double* P, long bp0, long bplane, hipStream_t st) {
  dim3 grd((g.npts[0] + 255) / 256, g.npts[1], g.npts[2]);
  k_to64<<<grd, 256, 0, st>>>(g, B32, P, bp0, bplane);
}

void launch_residual(const DGrid& g, const double* b, const double* P, float* B32, long bp0, long bplane,
                     hipStream_t st) {
  dim3 grd((g.npts[0] + 255) / 256, g.npts[1], g.npts[2]);
  k_residual<<<grd, 256, 0, st>>>(g, b, P, B32, bp0, bplane);
}


void launch_reduce_partials(const Red3* partial, int n, Red3* total, hipStream_t st) {
  k_reduce_sum2<<<1, kReduceThreads, 0, st>>>(partial, n, total);
}

void launch_sub_mean_field(const DGrid& g, const Layout& L, float* F, const Red3* total, hipStream_t st) {
  dim3 grd((g.N[0] + 255) / 256, g.N[1], g.N[2]);
  k_sub_mean_field<<<grd, 256, 0, st>>>(g, L, F, total);
}

void launch_accumulate(const DGrid& g, double* P, const float* B32, long bp0, long bplane, hipStream_t st) {
  dim3 grd((g.npts[0] + 255) / 256, g.npts[1], g.npts[2]);
  k_accumulate<<<grd, 256, 0, st>>>(g, P, B32, bp0, bplane);
}

}  // namespace mif
