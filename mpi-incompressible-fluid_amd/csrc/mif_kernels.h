// mif_kernels.h -- launch interfaces of the HIP kernels (host side).
#pragma once

#include "mif_device.h"

namespace mif {

// Arguments of the fused three-component RK3 stage kernel of the full step
// (src/Timestep.cpp:10-54).  X is the stencil source, A the pointwise
// operand, O1/O2 the outputs; a1, a2, b are the stage coefficients (stage 3
// passes its a3 in a2).
struct Red3;
struct StageArgs {
  const void* X[3];
  const void* A[3];
  void* O1[3];
  void* O2[3];
  const void* p;
  double a1, a2, b;
  double force[3];  // constant body force added to N (zero in the reference)
  // The previous stage's projection, fused into this stage (kinds 2 and 3):
  // when phi is set, X holds that stage's output before its projection and p
  // the pressure before its increment; the kernel reads them as
  // X - pdts grad(phi) (component interiors) and p + phi - mean (mean from
  // pmean, or 0), and stores the incremented p at its tile's points into pout
  // (not p, which neighbouring threads and workgroups read).
  const void* phi;
  double pdts;
  const Red3* pmean;
  void* pout;
};

// Arguments of the velocity-only stage kernel (src/TimestepVelocity.cpp:20-51).
struct VelStageArgs {
  const void* X[3];
  void* V[3];
  void* R[3];
  void* B[3];
  double t, dt;
  FnDesc forcing;
  // MIF_FN_VT_FORCING from tables of sin/cos of the coordinates (see
  // launch_trig_tables); null: the forcing is evaluated point by point
  const double* trig;
  int tn;
};

struct Red3 {
  double s1, s2, mx;
};

// One axis of the spectral Poisson solve: complex transform length L (= n on
// a periodic axis, 2(n-1) on a Neumann axis), its radix factorisation and
// the twiddle table W[e] = exp(-2 pi i e / L).
struct FftPlan {
  int n, L, periodic;
  int nst;
  int radix[24];
  int Ns[24];
  const double* W64;
  const float* W32;
  // cell-centred Neumann axis: DCT-II/DCT-III of length n = L, with the
  // half-angle table (cos, sin)(pi k / (2L)), k < L
  int dct2;
  const double* H64;
  const float* H32;
};

// Scaling applied between the forward and inverse transform of the last
// axis: 1 / ((lx[i] + ly[j] + lz[k]) * norm), zero for the constant mode.
struct SpectralScale {
  const double* lam[3];
  double inv_norm;
};

// Neighbour index along axis a of packed index i (d = -1 or +1): periodic
// wrap, cell-centred reflection of the first/last cell, or the node-centred
// even reflection (the 7-point operator of the packed solve).
__device__ inline long nb(const DGrid& g, int a, int i, int d) {
  const int n = g.npts[a];
  int q = i + d;
  if (g.periodic[a]) {
    q = q < 0 ? q + n : (q >= n ? q - n : q);
  } else if (g.cell[a]) {
    // cell-centred: the ghost mirrors the first/last cell
    if (q < 0) q = 0;
    if (q >= n) q = n - 1;
  } else {
    if (q < 0) q = n > 1 ? 1 : 0;
    if (q >= n) q = n > 1 ? n - 2 : 0;
  }
  return q;
}

// Operands of the x passes fused with the divergence (forward) or with the
// unpack into the pressure field (inverse).
struct FusedIO {
  const void* u;
  const void* v;
  const void* w;
  void* P;
  Layout L;
  double dt;
  // the velocity and pressure fields hold FP32 values while the transform
  // runs in FP64 (the mixed form with FP64 transforms)
  int fields_fp32 = 0;
  // mixed form with one FP64 defect sweep (FP32 transforms): the forward
  // pass also stores the FP64 divergence b (packed layout) and per-workgroup
  // owner-weighted sums (W.b, W.1) into partial; the inverse pass adds the
  // first FP32 solution add32 (packed layout) to the correction in FP64,
  // stores the sum, and leaves (W.phi, W.1) of the stored values in partial
  int dc = 0;
  double* b64 = nullptr;
  const float* add32 = nullptr;
  Red3* partial = nullptr;
  // x forward pass of the defect (res): the pass loads r = (b - mean) - L phi
  // in FP64 (b64, bmean, the first FP32 solution phi32, packed layout) instead
  // of its buffer, and transforms it into its buffer
  int res = 0;
  const float* phi32 = nullptr;
  const Red3* bmean = nullptr;
};

template <typename T>
void launch_stage(int stage, const DGrid& g, const Layout& L, const StageArgs& s, hipStream_t st);
// pout = pin + (phi - mean) on the pressure points a stage kernel with a fused
// projection leaves out (outside the union of the components' interiors):
// on its own (launch_p_rim) or in the boundary pass that follows the stage
// (launch_bc with a PressureRim)
struct PressureRim {
  const void* pin = nullptr;  // null: no rim
  const void* phi = nullptr;
  void* pout = nullptr;
  const Red3* mean = nullptr;
};
template <typename T>
void launch_p_rim(const DGrid& g, const Layout& L, const T* pin, const T* phi, T* pout, const Red3* mean,
                  hipStream_t st);
template <typename T>
void launch_vel_stage(int stage, const DGrid& g, const Layout& L, const VelStageArgs& s, hipStream_t st);
// trig[((a * 2 + s) * 2 + q) * tn + i] = sin (q = 0) or cos (q = 1) of the
// coordinate of index i on axis a, unshifted (s = 0) or shifted by -h/2
// (s = 1, the staggered axis of a component), i < tn.
void launch_trig_tables(const DGrid& g, double* trig, int tn, hipStream_t st);
template <typename T>
void launch_bc(const DGrid& g, const Layout& L, void* const F[3], const FnDesc& f, double t, hipStream_t st,
               const PressureRim& rim = PressureRim{});
template <typename T>
void launch_periodic(const DGrid& g, const Layout& L, T* F, int stag, hipStream_t st);
template <typename T>
void launch_set(const DGrid& g, const Layout& L, T* F, int stag, int comp, const FnDesc& f, double t, int border,
                hipStream_t st);
template <typename T, typename S>
void launch_div_rhs(const DGrid& g, const Layout& L, const T* u, const T* v, const T* w, S* B, long bp0, long bplane,
                    double dt, const FnDesc& gp, double tg, int nhn, hipStream_t st);
template <typename T, typename S>
void launch_pack(const DGrid& g, const Layout& L, const T* P, S* B, long bp0, long bplane, hipStream_t st);
template <typename T, typename S>
void launch_unpack(const DGrid& g, const Layout& L, const S* B, long bp0, long bplane, T* P, hipStream_t st);
template <typename T>
// p += phi (minus mean->s1 / mean->s2 when mean is not null), velocity -= dts grad phi
void launch_project(const DGrid& g, const Layout& L, T* u, T* v, T* w, T* p, const T* phi, double dts, hipStream_t st,
                    const Red3* mean = nullptr, T* pout = nullptr);
template <typename T>
void launch_pgrad(const DGrid& g, const Layout& L, const T* p, T* u, T* v, T* w, hipStream_t st);
template <typename T>
void launch_from_dense(const Layout& L, const int e[3], const double* src, T* dst, hipStream_t st);
template <typename T>
void launch_to_dense(const Layout& L, const int e[3], const T* src, double* dst, hipStream_t st);
template <typename T>
void launch_copy(long n, const T* src, T* dst, hipStream_t st);
template <typename T>
void launch_dot(const Layout& L, const int e[3], const T* a, const T* b, Red3* partial, int nblocks, Red3* out,
                hipStream_t st);
// out[q] = F(ijk[3q], ijk[3q+1], ijk[3q+2]) for q < n (indices checked by the caller).
template <typename T>
void launch_gather(const Layout& L, const T* F, const int* ijk, long n, double* out, hipStream_t st);
template <typename T>
void launch_add_const(const DGrid& g, const Layout& L, T* P, double c, hipStream_t st);
// Exact-solution tables for the norms and adjust_pressure: node-shaped
// fields (element type of the context) holding the exact values; f[0] null
// selects the closed-form descriptor instead.
struct ExactTables {
  const void* f[3];
};
template <typename T>
void launch_norm_velocity(const DGrid& g, const Layout& L, const T* u, const T* v, const T* w, const FnDesc& f,
                          double t, const ExactTables& tab, Red3* partial, int nblocks, Red3* out, hipStream_t st);
template <typename T>
void launch_norm_pressure(const DGrid& g, const Layout& L, const T* p, const FnDesc& f, double t, int adjust,
                          const ExactTables& tab, Red3* partial, int nblocks, Red3* out, hipStream_t st);

// Spectral solve passes on the packed buffer B (extents npts, row pitch bp0,
// plane pitch bplane).  dir: -1 forward, +1 inverse; axis 2 with fused = 1
// runs forward, scale, inverse in one pass.
template <typename T>
void launch_fft_axis(const DGrid& g, int axis, int dir, int fused, const FftPlan& plan, const SpectralScale& sc, T* B,
                     long bp0, long bplane, hipStream_t st);
// Specialised transform pass (complex lengths 2^k and 3 * 2^k, see
// fft_fast_len); on axis 0 with io set, the forward pass is fused with the
// divergence (mode 0) and the inverse with the unpack into the pressure field
// (mode 1).  Returns false when the length has no specialised kernel.
bool fft_fast_len(int L);
template <typename T>
bool launch_fft_pow2(const DGrid& g, int axis, int mode, const FftPlan& plan, const SpectralScale& sc, T* B, long bp0,
                     long bplane, const FusedIO* io, hipStream_t st);
// Workgroups of the specialised x pass of a plan (0 without a kernel): the
// number of partial sums a fused x pass with FusedIO::dc leaves.
int fused_x_blocks(const DGrid& g, const FftPlan& plan, bool fp64);
// DCT-II (forward) or DCT-III (inverse) pass of a cell-centred Neumann y
// axis; false when the length (64, 128, 256, 512 or 1024) has no kernel.
template <typename T>
bool launch_dct2_y_any(const DGrid& g, const FftPlan& plan, int inverse, T* B, long bp0, long bplane, hipStream_t st);

// Three component arrays of one velocity-shaped quantity.
struct AdjVec {
  void* f[3];
};

// Adjoint of one step (mif_adjoint.hip).
// Packed right-hand side of the transposed solve from p_bar and y_bar.
template <typename T>
void launch_adj_rhs(const DGrid& g, const Layout& L, const T* P, const AdjVec& Y, double dts, T* B, long bp0,
                    long bplane, hipStream_t st);
// z_bar and N_bar = alpha z_bar + beta Zo of a stage, plus nblk partial sums
// of z_bar per component (partial holds 3 * nblk doubles).
template <typename T>
void launch_adj_zb(const DGrid& g, const Layout& L, const AdjVec& Y, const T* Bpsi, long bp0, long bplane, double dts,
                   const AdjVec& Zb, const AdjVec& Nb, const AdjVec& Zo, double alpha, double beta, double fcoef,
                   int accumulate, double* partial, int nblk, hipStream_t st);
// Workgroups per component of k_adj_zb (one point per thread up to this
// many, grid-stride beyond), and the number of partials k_adj_sum reduces.
constexpr int kAdjMaxBlocks = 4096;
// out[c] = (accumulate: out[c] +) the sum of the nblk <= kAdjMaxBlocks
// partials of component c.
void launch_adj_sum(const double* partial, int nblk, int accumulate, double* out, hipStream_t st);
// x_bar = z_bar + J_N(X)^T N_bar; p_bar += bcoef * G^T z_bar.
template <typename T>
void launch_adj_xb(const DGrid& g, const Layout& L, const AdjVec& X, const AdjVec& Nb, const AdjVec& Zb,
                   const AdjVec& Xb, T* P, double bcoef, hipStream_t st);

// Mixed form: FP64 right-hand side projection and defect correction
// (mif_mixed.hip).
void launch_project_rhs(const DGrid& g, double* B64, float* B32, long bp0, long bplane, Red3* partial, int nblocks,
                        Red3* total, hipStream_t st);
void launch_to64(const DGrid& g, const float* B32, double* P, long bp0, long bplane, hipStream_t st);
void launch_remove_mean(const DGrid& g, double* P, long bp0, long bplane, Red3* partial, int nblocks, Red3* total,
                        hipStream_t st);
void launch_residual(const DGrid& g, const double* b, const double* P, float* B32, long bp0, long bplane,
                     hipStream_t st);
void launch_accumulate(const DGrid& g, double* P, const float* B32, long bp0, long bplane, hipStream_t st);
// One-sweep defect correction with FP32 transforms: the sum of n partial
// (W.x, W.1) pairs into total; a node field minus the mean total holds.
void launch_reduce_partials(const Red3* partial, int n, Red3* total, hipStream_t st);
void launch_sub_mean_field(const DGrid& g, const Layout& L, float* F, const Red3* total, hipStream_t st);

// Streaming-ceiling probe (mif_probe.hip): nr arrays read and nw arrays
// written, n4 16-byte elements each.
constexpr int kStreamMax = 8;
struct StreamArgs {
  const float4* in[kStreamMax];
  float4* out[kStreamMax];
};
void launch_stream(const StreamArgs& a, int nr, int nw, long n4, int blocks, hipStream_t st);

}  // namespace mif
