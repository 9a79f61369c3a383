// mif_probe.hip -- streaming-ceiling probe (a measurement aid; no counterpart
// in the reference).  k_stream reads every element of nr arrays and writes
// every element of nw arrays, 16 bytes per lane and array per iteration,
// grid-stride.  Its time for the read/write mix of a production kernel is
// what that mix can stream on the device at all, the ceiling the kernel's
// own bandwidth is compared with.
#include "mif_kernels.h"

namespace mif {

__global__ void __launch_bounds__(256) k_stream(StreamArgs a, int nr, int nw, long n4) {
  for (long q = blockIdx.x * 256L + threadIdx.x; q < n4; q += (long)gridDim.x * 256) {
    float4 v[kStreamMax];
#pragma unroll
    for (int r = 0; r < kStreamMax; r++)
      if (r < nr) v[r] = a.in[r][q];
    float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
    for (int r = 0; r < kStreamMax; r++)
      if (r < nr) {
        acc.x += v[r].x;
        acc.y += v[r].y;
        acc.z += v[r].z;
        acc.w += v[r].w;
      }
#pragma unroll
    for (int w = 0; w < kStreamMax; w++)
      if (w < nw) a.out[w][q] = make_float4(acc.x + (float)w, acc.y, acc.z, acc.w);
  }
}

void launch_stream(const StreamArgs& a, int nr, int nw, long n4, int blocks, hipStream_t st) {
  k_stream<<<blocks, 256, 0, st>>>(a, nr, nw, n4);
}

}  // namespace mif
