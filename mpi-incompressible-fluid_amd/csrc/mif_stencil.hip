// mif_stencil.hip -- stencil, boundary and reduction kernels of the timestep.
//
// Each kernel cites the reference loop it replaces.  All kernels take the
// padded common layout of mif_device.h; T is the storage/arithmetic type of
// the fields (double for the plain form, float for the mixed form).
#include <type_traits>

#include "mif_kernels.h"

namespace mif {

// ---------------------------------------------------------------------------
// Momentum right-hand side at one point (include/MomentumEquation.h:43-255,
// the OPT_CPU_CACHE branch).  X(f, di, dj, dk) returns component f (0 = u,
// 1 = v, 2 = w) at offset (di, dj, dk) from the point.
// ---------------------------------------------------------------------------
template <typename T, typename Tap>
__device__ inline T rhs_u(const DGrid& g, Tap X) {
  const T u13 = X(0, -1, 0, 0), u11 = X(0, 0, -1, 0), u5 = X(0, 0, 0, -1), u14 = X(0, 0, 0, 0), u23 = X(0, 0, 0, 1),
          u17 = X(0, 0, 1, 0), u15 = X(0, 1, 0, 0);
  const T v13 = X(1, -1, 0, 0), v16 = X(1, -1, 1, 0), v14 = X(1, 0, 0, 0), v17 = X(1, 0, 1, 0);
  const T w13 = X(2, -1, 0, 0), w22 = X(2, -1, 0, 1), w14 = X(2, 0, 0, 0), w23 = X(2, 0, 0, 1);
  const T conv = -u14 * (u15 - u13) * (T)g.one_over_2h[0] - (v17 + v14 + v16 + v13) * (u17 - u11) * (T)g.one_over_8h[1] -
                 (w23 + w14 + w22 + w13) * (u23 - u5) * (T)g.one_over_8h[2];
  const T diff = (u15 - 2 * u14 + u13) * (T)g.one_over_h2Re[0] + (u17 - 2 * u14 + u11) * (T)g.one_over_h2Re[1] +
                 (u23 - 2 * u14 + u5) * (T)g.one_over_h2Re[2];
  return conv + diff;
}

template <typename T, typename Tap>
__device__ inline T rhs_v(const DGrid& g, Tap X) {
  const T u11 = X(0, 0, -1, 0), u14 = X(0, 0, 0, 0), u12 = X(0, 1, -1, 0), u15 = X(0, 1, 0, 0);
  const T v13 = X(1, -1, 0, 0), v11 = X(1, 0, -1, 0), v5 = X(1, 0, 0, -1), v14 = X(1, 0, 0, 0), v23 = X(1, 0, 0, 1),
          v17 = X(1, 0, 1, 0), v15 = X(1, 1, 0, 0);
  const T w11 = X(2, 0, -1, 0), w20 = X(2, 0, -1, 1), w14 = X(2, 0, 0, 0), w23 = X(2, 0, 0, 1);
  const T conv = -(u15 + u14 + u12 + u11) * (v15 - v13) * (T)g.one_over_8h[0] - v14 * (v17 - v11) * (T)g.one_over_2h[1] -
                 (w23 + w14 + w20 + w11) * (v23 - v5) * (T)g.one_over_8h[2];
  const T diff = (v15 - 2 * v14 + v13) * (T)g.one_over_h2Re[0] + (v17 - 2 * v14 + v11) * (T)g.one_over_h2Re[1] +
                 (v23 - 2 * v14 + v5) * (T)g.one_over_h2Re[2];
  return conv + diff;
}

template <typename T, typename Tap>
__device__ inline T rhs_w(const DGrid& g, Tap X) {
  const T u5 = X(0, 0, 0, -1), u14 = X(0, 0, 0, 0), u6 = X(0, 1, 0, -1), u15 = X(0, 1, 0, 0);
  const T v5 = X(1, 0, 0, -1), v14 = X(1, 0, 0, 0), v8 = X(1, 0, 1, -1), v17 = X(1, 0, 1, 0);
  const T w13 = X(2, -1, 0, 0), w11 = X(2, 0, -1, 0), w5 = X(2, 0, 0, -1), w14 = X(2, 0, 0, 0), w23 = X(2, 0, 0, 1),
          w17 = X(2, 0, 1, 0), w15 = X(2, 1, 0, 0);
  const T conv = -(u15 + u14 + u6 + u5) * (w15 - w13) * (T)g.one_over_8h[0] -
                 (v17 + v14 + v8 + v5) * (w17 - w11) * (T)g.one_over_8h[1] - w14 * (w23 - w5) * (T)g.one_over_2h[2];
  const T diff = (w15 - 2 * w14 + w13) * (T)g.one_over_h2Re[0] + (w17 - 2 * w14 + w11) * (T)g.one_over_h2Re[1] +
                 (w23 - 2 * w14 + w5) * (T)g.one_over_h2Re[2];
  return conv + diff;
}

// Interior test of component c (STAGGERED_TENSOR_ITERATE_OVER_ALL_POINTS with
// include_border = false, include/StaggeredTensorMacros.h:6-36).
__device__ inline bool interior(const DGrid& g, int c, int i, int j, int k) {
  int e[3];
  extents(g, c, e);
  return i >= 1 && j >= 1 && k >= 1 && i <= e[0] - 2 && j <= e[1] - 2 && k <= e[2] - 2;
}

// ---------------------------------------------------------------------------
// RK3 stage kernels (src/Timestep.cpp:10-54 and src/TimestepVelocity.cpp:
// 20-51), all three components in one pass.
//
// A 64 x 8 (FP64) or 128 x 4 (FP32) workgroup owns an (i, j) tile and
// marches along k.  The planes
// k-1, k, k+1 of the stencil source (u, v, w with a one-point halo in i and
// j) live in an LDS ring buffer, so each source value is fetched from
// memory about once; every tap of the 15-point stencils is an LDS read.
// Plane k+2 is fetched into registers while plane k is computed and moved
// to LDS afterwards, and the pointwise operands of all three components are
// loaded together before the stencil work, so memory latency overlaps the
// arithmetic instead of being paid once per load.  Tiles are numbered x
// fastest, then y, then k chunk, and handed out per XCD in contiguous ranges
// (xcd_tile), so tiles that share halo rows share an L2.
//
// Full step (KIND 1..3):
//   stage 1: O1 = X + a1*N(X) - b*Gp,          O2 = N(X)        (X = u^n)
//   stage 2: O1 = X + a1*A + a2*N(X) - b*Gp,   O2 = a2*N(X)     (A = R1)
//   stage 3: O1 = X - A + a3*N(X) - b*Gp                        (A = S)
// Velocity-only step with forcing f (KIND 4..6):
//   Y2: R = N(X)+f(t_n);  B = X + dt*a21*R                       (X = u)
//   Y3: R' = V + dt*(b1*R); V = V + dt*(a31*R + a32*(N(X)+f))    (X = buf)
//   U*: B = R + dt*(b3*(N(X)+f))                                 (X = u)
// Outputs may alias pointwise operands (same point, same thread); the
// stencil source is never written.
// ---------------------------------------------------------------------------
// Stores of the stage and projection kernels: non-temporal when NT (the
// layout's `nt`, set for fields that fit the Infinity Cache).
// With NT, the operands each point reads once (the stage's A operand, the
// projection's p, u, v, w) are loaded non-temporally too (256^3 step: FP64
// 3.10 -> 2.99 ms, mixed 1.636 -> 1.565 ms).  Loads that neighbouring points
// or tiles read again stay ordinary: non-temporal halo-plane loads took the
// mixed stage kernel from 180 to 196 us.
constexpr bool kNtLoads = true;
template <bool NT, typename T>
__device__ inline T load_in(const T* addr) {
  if constexpr (NT)
    return __builtin_nontemporal_load(addr);
  else
    return *addr;
}
template <bool NT, typename T>
__device__ inline void store_out(T* addr, T val) {
  if constexpr (NT)
    __builtin_nontemporal_store(val, addr);
  else
    *addr = val;
}
// Stage-kernel tile: STX x STY points of a plane per 512-thread workgroup,
// per element type.  Measured at 256^3 against 64 x 8: 128 x 4 took the FP32
// stage kernel from 182.5 to 177.7 us (mixed step 1.562 -> 1.552 ms) and the
// FP64 one from 335 to 345 us.
template <typename T>
struct StageTile {
  static constexpr int STX = sizeof(T) == 4 ? 128 : 64;
  static constexpr int STY = sizeof(T) == 4 ? 4 : 8;
  static constexpr int SHX = STX + 2, SHY = STY + 2;
  static constexpr int STN = STX * STY;                      // threads per workgroup
  static constexpr int PLANE_ELEMS = 3 * SHY * SHX;          // one halo plane of u, v, w
  static constexpr int PQ = (PLANE_ELEMS + STN - 1) / STN;   // plane elements per thread
  static_assert(STN == 512, "the stage kernel's launch bounds");
};
// planes per workgroup.  FP32 on 128 x 4 tiles at 256^3: 8 planes 180.4 us,
// 16 176.5 us, 24 178.0 us, 32 192.5 us.  FP64 on 64 x 8 tiles, with the
// plane order of k_stage: 4 planes 329 us, 6 328, 8 320, 10 320,
// 12 323, 16 336, 24 357, 32 342 (four ring slots instead of three: 8 planes
// 318 us, 16 330 us, so three slots stay).  Fields beyond the Infinity Cache
// (no non-temporal path, NT false) keep 16 planes in the full step: 512^3
// FP64 stage kernel 2763 us at 16 planes, 2831 us at 8.  The velocity-only
// kinds (4..6) take 8 in FP64 (256^3 step 1.53 -> 1.46 ms with the request
// order above).
// The FP64 kernels with a fused projection take three ring slots: with the
// LDS phi ring (71 KB of LDS, two workgroups per CU) a fourth would leave one.
// Their chunk is 12 planes (256^3 stage kernels: 6 planes 381 us, 8 377,
// 12 372, 16 383); the FP32 ones keep 16 (8 planes 200 us, 16 196, 24 199).
template <typename T, int KIND, bool NT, bool PROJ = false>
constexpr int stage_planes() {
  return sizeof(T) == 8 && PROJ ? 12 : (sizeof(T) == 8 && (NT || KIND >= 4) ? 8 : 16);
}
constexpr int kStageSlotsFused64 = 3;

// ring slots of the FP32 and FP64 stage kernels (see k_stage)
constexpr int kStageSlots32 = 4, kStageSlots64 = 3;
// halo planes fetched with 16-byte loads where the storage shift aligns the
// rows (see the launcher)
constexpr bool kStageVec = true;

// FT (velocity-only step): MIF_FN_VT_FORCING read from the trig tables.
// VEC: halo planes fetched with 16-byte loads (element 1 of every row
// 16-byte aligned: the layout's storage shift).
// PROJ (kinds 2 and 3): the previous stage's projection fused in (StageArgs
// phi): each halo element of X is corrected by -pdts grad(phi) as it is
// committed to LDS (k_project's expression, on the component's interior),
// the pressure gradient uses p + phi - mean, and that incremented pressure is
// stored into pout at the tile's points (another buffer than p: neighbouring
// threads and tiles read p).
template <typename T, int KIND, bool FT = false, bool NT = false, bool VEC = false, bool PROJ = false>
// (FP64 full step: at most 80 VGPRs, i.e. three 512-thread workgroups per CU;
// with the fused projection 128, two workgroups, which its LDS allows anyway;
// FP32 with the fused projection, and the FP32 velocity-only kinds with the
// tabulated forcing: at most 80 VGPRs, three workgroups (256^3 mixed
// velocity-only step 0.926 -> 0.908 ms; the FP64 ones spill under that cap,
// 1.47 -> 1.86 ms))
__global__ void __launch_bounds__(512)
__attribute__((amdgpu_waves_per_eu(KIND <= 3 && (sizeof(T) == 8 || PROJ) ? (sizeof(T) == 8 && PROJ ? 4 : 6)
                                                                      : (FT && sizeof(T) == 4 ? 6 : 1))))
k_stage(DGrid g, Layout L, StageArgs s, VelStageArgs vs, int imax, int jmax, int kmax, int ntx, int nty, int ntiles) {
  static_assert(!PROJ || KIND <= 3, "a fused projection belongs to the full-step kinds");
  // ring slots: 3 (planes k-1, k, k+1; two barriers per plane) or 4 and
  // more (plane k+2 is committed into a free slot while plane k is
  // computed; one barrier per plane)
  constexpr int NS = PROJ && sizeof(T) == 8 ? kStageSlotsFused64 : (sizeof(T) == 4 ? kStageSlots32 : kStageSlots64);
  // p at plane k-1 carried in a register from plane to plane instead of
  // loaded (FP64: 256^3 stage kernel 348 -> 341 us, step 2.955 -> 2.928 ms)
  constexpr bool PCARRY = sizeof(T) == 8 || KIND <= 3;
  // fused projection: the phi planes are staged through an LDS ring, so that
  // phi is read once per plane instead of at every halo element and below it
  constexpr bool PL = PROJ;
  using TT = StageTile<T>;
  constexpr int STX = TT::STX, STY = TT::STY, SHX = TT::SHX, SHY = TT::SHY, STN = TT::STN;
  constexpr int PLANE_ELEMS = TT::PLANE_ELEMS, PQ = TT::PQ;
  __shared__ T pl[NS][3][SHY][SHX];
  // PL: phi on planes k .. k+3 over the halo plane's points plus one more row
  // and column below (the lower neighbours along x and y)
  constexpr int NSP = 4, PHX = SHX + 1, PHE = (SHY + 1) * PHX, PHQ = (PHE + STN - 1) / STN;
  __shared__ T phl[PL ? NSP : 1][PL ? PHE : 1];
  const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * STX + tx;
  const int tile = xcd_tile(blockIdx.x, ntiles);
  const int bx = tile % ntx, by = (tile / ntx) % nty, bz = tile / (ntx * nty);
  constexpr int SKB = stage_planes<T, KIND, NT, PROJ>();
  const int i0 = bx * STX + 1, j0 = by * STY + 1, k0 = bz * SKB + 1;
  const int k1 = min(k0 + SKB, kmax + 1);
  const int i = i0 + tx, j = j0 + ty;
  const T* X[3];
  for (int f = 0; f < 3; f++) X[f] = (const T*)(KIND <= 3 ? s.X[f] : vs.X[f]);

  // this thread's share of a halo plane: source column and LDS offset.
  // Scalar form: PQ elements.  VEC form: QV 16-byte loads of VW tile-row
  // elements (row element i0 + VW*iq, 16-byte aligned when element 1 of a
  // row is) plus at most one halo-column element.
  constexpr int VW = 16 / (int)sizeof(T);
  constexpr int RQ = STX / VW;                    // 16-byte loads per tile row
  constexpr int NIV = 3 * SHY * RQ;               // 16-byte loads per plane
  constexpr int QV = (NIV + STN - 1) / STN;
  constexpr int NHV = 3 * SHY * 2;                // halo-column elements per plane
  static_assert(NHV <= STN, "one halo-column element per thread");
  // PL: the halo plane is fetched field by field: slot s of a thread is
  // point q = tid + STN s of the SHY x SHX plane, the same point of u, v and
  // w, so that the three share its phi values and its index arithmetic
  constexpr int FQ = (SHY * SHX + STN - 1) / STN;
  // field-major slots: the fused kernels and every FP32 kernel (the FP64
  // kernels without a projection keep their 16-byte loads: the FP64
  // velocity-only step 1.49 -> 1.54 ms with field-major slots).  With them
  // the FP32 kernels took the request order of the FP64 ones (see the plane
  // loop), which with their earlier element-major slots had measured slower
  // (stage 1 175 -> 181 us): 256^3 mixed fused kinds about 231 -> 211 us,
  // stage 1 about 177 -> 166 us, velocity-only step 1.007 -> 0.931 ms.
  constexpr bool FM = PL || sizeof(T) == 4;
  constexpr int PR = FM ? 3 * FQ : (VEC ? QV * VW + 1 : PQ);  // plane values a thread holds
  typedef T Vec __attribute__((ext_vector_type(VW)));
  // halo loads go through global-address-space pointers: generic (flat)
  // loads also count as LDS operations, so every wait for the LDS ring would
  // wait for the plane in flight as well
  typedef const __attribute__((address_space(1))) T GT;
  typedef const __attribute__((address_space(1))) Vec GV;
  GT* src[VEC ? QV + 1 : PQ];
  int dst[VEC ? QV + 1 : PQ];
  bool ok[VEC ? QV + 1 : PQ];
  if constexpr (VEC) {
#pragma unroll
    for (int r = 0; r < QV; r++) {
      const int q = tid + STN * r;
      const int f = q / (SHY * RQ), rr = q - f * (SHY * RQ);
      const int jj = rr / RQ, iq = rr - jj * RQ;
      const int gi = i0 + VW * iq, gj = j0 - 1 + jj;
      // a load that starts inside the row may run up to VW - 1 elements past
      // its end (into the next row, or the allocation's tail); those lanes
      // feed no output point
      ok[r] = q < NIV && gi < L.PX && gj < L.SY;
      src[r] = (GT*)(ok[r] ? X[f] + L.idx(gi, gj, 0) : X[0] + 1);
      dst[r] = q < NIV ? (f * SHY + jj) * SHX + 1 + VW * iq : -1;
    }
    const int f = tid / (SHY * 2), rr = tid - f * (SHY * 2);
    const int jj = rr >> 1, side = rr & 1;
    const int gi = side ? i0 + STX : i0 - 1, gj = j0 - 1 + jj;
    ok[QV] = tid < NHV && gi < L.PX && gj < L.SY;
    src[QV] = (GT*)(ok[QV] ? X[f] + L.idx(gi, gj, 0) : X[0]);
    dst[QV] = tid < NHV ? (f * SHY + jj) * SHX + (side ? SHX - 1 : 0) : -1;
  } else {
#pragma unroll
    for (int r = 0; r < PQ; r++) {
      const int q = tid + STN * r;
      const int f = q / (SHY * SHX), rr = q - f * (SHY * SHX);
      const int jj = rr / SHX, ii = rr - jj * SHX;
      const int gi = i0 - 1 + ii, gj = j0 - 1 + jj;
      ok[r] = q < PLANE_ELEMS && gi < L.PX && gj < L.SY;
      src[r] = (GT*)(ok[r] ? X[f] + L.idx(gi, gj, 0) : X[0]);
      dst[r] = q < PLANE_ELEMS ? q : -1;
    }
  }
  const T* phi = (const T*)s.phi;
  // FM: per slot the source offset of its point (a point outside the
  // arrays reads element 0: its value is masked), whether it is inside, its
  // (i, j) and (PL) its index in the LDS phi plane
  constexpr int FQS = FM ? FQ : 1;
  int xo[FQS], xgi[FQS], xgj[FQS], xix[FQS];
  bool xok[FQS];
  if constexpr (FM) {
#pragma unroll
    for (int r = 0; r < FQ; r++) {
      const int q = tid + STN * r;
      const int jj = q / SHX, ii = q - jj * SHX;
      xgi[r] = i0 - 1 + ii;
      xgj[r] = j0 - 1 + jj;
      xok[r] = q < SHY * SHX && xgi[r] < L.PX && xgj[r] < L.SY;
      xo[r] = xok[r] ? (int)L.idx(xgi[r], xgj[r], 0) : 0;
      xix[r] = (jj + 1) * PHX + ii + 1;
    }
  }
  // PL: this thread's share of a phi plane, source column and LDS index tid +
  // STN r (a column outside the array reads element 0 of the plane: it is
  // the lower neighbour of no interior element)
  constexpr int PQS = PL ? PHQ : 1;
  int poff[PQS];
  if constexpr (PL) {
#pragma unroll
    for (int r = 0; r < PHQ; r++) {
      const int q = tid + STN * r;
      const int jj = q / PHX, ii = q - jj * PHX;
      const int gi = i0 - 2 + ii, gj = j0 - 2 + jj;
      const bool okp = q < PHE && gi >= 0 && gj >= 0 && gi < L.PX && gj < L.SY;
      poff[r] = okp ? (int)L.idx(gi, gj, 0) : 0;
    }
  }
  // PL: phi plane k (plane 0 for k = -1, which feeds no interior element)
  // into registers, and from there into its LDS slot
  auto pfetch = [&](int k, T (&v)[PQS]) {
    if constexpr (PL) {
      GT* pk = (GT*)phi + (long)max(k, 0) * L.plane;
#pragma unroll
      for (int r = 0; r < PHQ; r++) v[r] = pk[poff[r]];
    }
  };
  auto pcommit = [&](int k, const T (&v)[PQS]) {
    if constexpr (PL) {
      T* base = &phl[(k + NSP) % NSP][0];
#pragma unroll
      for (int r = 0; r < PHQ; r++)
        if (tid + STN * r < PHE) base[tid + STN * r] = v[r];
    }
  };
  // every plane read lies in 0 .. kmax + 1 <= SZ - 1, so the load address is
  // valid for every lane and the loads need no branch.  Lanes outside the
  // field are zeroed when the values are committed to LDS, so that no
  // instruction waits for the loads before the commit.
  auto mask = [&](bool keep, T x) { return keep ? x : (T)0; };
  auto fetch = [&](int k, T (&v)[PR]) {
    if constexpr (FM) {
#pragma unroll
      for (int r = 0; r < FQ; r++)
#pragma unroll
        for (int f = 0; f < 3; f++) v[f * FQ + r] = ((GT*)X[f])[xo[r] + (long)k * L.plane];
    } else if constexpr (VEC) {
#pragma unroll
      for (int r = 0; r < QV; r++) {
        const Vec x = *(GV*)(src[r] + (long)k * L.plane);
#pragma unroll
        for (int e = 0; e < VW; e++) v[r * VW + e] = x[e];
      }
      const T h = src[QV][(long)k * L.plane];
      v[QV * VW] = h;
    } else {
#pragma unroll
      for (int r = 0; r < PQ; r++) {
        const T x = src[r][(long)k * L.plane];
        v[r] = x;
      }
    }
  };
  auto commit = [&](int k, const T (&v)[PR]) {
    T* base = &pl[k % NS][0][0][0];
    if constexpr (FM && !PL) {
#pragma unroll
      for (int r = 0; r < FQ; r++) {
        if (tid + STN * r >= SHY * SHX) continue;
#pragma unroll
        for (int f = 0; f < 3; f++) base[f * (SHY * SHX) + tid + STN * r] = mask(xok[r], v[f * FQ + r]);
      }
    } else if constexpr (PL) {
      // projected values X - pdts grad(phi) on each field's interior, phi at
      // the point and below it from the LDS planes k and k-1 (k >= 0 here)
      const T* cur = &phl[k % NSP][0];
      const T* prv = &phl[(k + NSP - 1) % NSP][0];
#pragma unroll
      for (int r = 0; r < FQ; r++) {
        if (tid + STN * r >= SHY * SHX) continue;
        const int ix = xix[r];
        const T fc = cur[ix], lo[3] = {cur[ix - 1], cur[ix - PHX], prv[ix]};
#pragma unroll
        for (int f = 0; f < 3; f++) {
          int e[3];
          extents(g, f, e);
          const bool in = xok[r] && xgi[r] >= 1 && xgj[r] >= 1 && xgi[r] <= e[0] - 2 && xgj[r] <= e[1] - 2 && k >= 1 &&
                          k <= e[2] - 2;
          const T x = mask(xok[r], v[f * FQ + r]);
          const T xp = x - (fc - lo[f]) * (T)g.one_over_h[f] * (T)s.pdts;
          base[f * (SHY * SHX) + tid + STN * r] = in ? xp : x;
        }
      }
    } else if constexpr (VEC) {
#pragma unroll
      for (int r = 0; r < QV; r++)
        if (dst[r] >= 0)
#pragma unroll
          for (int e = 0; e < VW; e++)
            base[dst[r] + e] = mask(ok[r], v[r * VW + e]);
      if (dst[QV] >= 0) base[dst[QV]] = mask(ok[QV], v[QV * VW]);
    } else {
#pragma unroll
      for (int r = 0; r < PQ; r++)
        if (dst[r] >= 0) base[dst[r]] = mask(ok[r], v[r]);
    }
  };
  if constexpr (PL) {
    // phi planes k0-2 .. k0+1, which the projections of the first three
    // planes read
    T q[4][PQS];
#pragma unroll
    for (int d = 0; d < 4; d++) pfetch(k0 - 2 + d, q[d]);
#pragma unroll
    for (int d = 0; d < 4; d++) pcommit(k0 - 2 + d, q[d]);
    __syncthreads();
  }
  {
    T a[PR], b[PR], c[PR];
    fetch(k0 - 1, a);
    fetch(k0, b);
    fetch(k0 + 1, c);
    commit(k0 - 1, a);
    commit(k0, b);
    commit(k0 + 1, c);
  }
  T pre[PR];
  T pq[PQS];
  if constexpr (PL) {
    // phi plane k0+2 takes the slot of plane k0-2 once every projection
    // above is done
    if (k0 + 2 <= k1) pfetch(k0 + 2, pq);
    __syncthreads();
    if (k0 + 2 <= k1) pcommit(k0 + 2, pq);
  }
  __syncthreads();

  const bool in_tile = i <= imax && j <= jmax;

  const long doff[3] = {1, L.PX, L.plane};
  const T mv = PROJ && s.pmean ? (T)(s.pmean->s1 / s.pmean->s2) : (T)0;
  T p_prev = (T)0;
  if (KIND <= 3 && PCARRY && in_tile) {
    const long ob = L.idx(i, j, k0 - 1);
    p_prev = ((const T*)s.p)[ob];
    if constexpr (PROJ) p_prev = p_prev + (phi[ob] - mv);
  }
  // Order of the memory requests per plane k: the pointwise operands of
  // plane k first, then the halo plane k+2 into registers; the stencil work
  // of plane k waits for the pointwise operands only (loads complete in issue
  // order), so plane k+2 stays in flight during the work and is committed to
  // LDS after it (256^3 FP64 stage kernel: 356 -> 343 us against committing
  // plane k+2 and requesting plane k+3 before the pointwise operands).
  // The pointwise loads are unconditional: threads outside the tile load at
  // a clamped column (every address lies inside the allocation: i, j, k >= 1
  // and i <= imax, j <= jmax, k <= kmax) and their values are not used, so
  // the number of loads in flight is fixed for the compiler's waits.
  const long oc = L.idx(min(i, imax), min(j, jmax), 0);
  for (int k = k0; k < k1; k++) {
    const long o = L.idx(i, j, k);
    const long ol = oc + (long)k * L.plane;
    T p0 = (T)0, pm[3] = {(T)0, (T)0, (T)0}, av[3] = {(T)0, (T)0, (T)0}, bv[3] = {(T)0, (T)0, (T)0};
    T f0 = (T)0, fm[3] = {(T)0, (T)0, (T)0};  // fused projection: phi at the point and below it
    // pointwise operands of all three components, loaded together
    auto load_pointwise = [&](const bool (&inside)[3]) {
      if constexpr (KIND <= 3) {
        const T* p = (const T*)s.p + ol;
        p0 = p[0];
#pragma unroll
        for (int c = 0; c < 2; c++)
          if (inside[c]) pm[c] = p[-doff[c]];
        if constexpr (!PCARRY)
          if (inside[2]) pm[2] = p[-doff[2]];
        if constexpr (KIND >= 2) {
#pragma unroll
          for (int c = 0; c < 3; c++)
            if (inside[c]) av[c] = load_in<NT && kNtLoads>(&((const T*)s.A[c])[ol]);
        }
      } else if constexpr (KIND >= 5) {
#pragma unroll
        for (int c = 0; c < 3; c++)
          if (inside[c]) {
            av[c] = ((const T*)vs.R[c])[ol];
            if constexpr (KIND == 5) bv[c] = ((const T*)vs.V[c])[ol];
          }
      }
    };
    {
      const bool every[3] = {true, true, true};
      load_pointwise(every);
    }
    // plane k+2 is read by plane k+1's stencil; past the chunk's last plane
    // the fetch repeats plane k1 (a cache hit) and is not committed
    const bool more = k + 2 <= k1;
    fetch(min(k + 2, k1), pre);
    // PL: phi plane k+3, which the projection of plane k+3 reads one plane
    // later (and past the chunk's last plane a repeat, not committed)
    pfetch(min(k + 3, k1), pq);
    // the full step computes every component on every thread (the taps of a
    // thread outside the field still lie in the LDS tile; only the stores
    // are masked), so that the compiler keeps the loads above unconditional
    // instead of sinking them into the branches
    constexpr bool EVERY = KIND <= 3;
    if (EVERY || in_tile) {
      bool inside[3];
      for (int c = 0; c < 3; c++) inside[c] = interior(g, c, i, j, k);
      // fused projection: the pressure after the previous stage's increment,
      // p + (phi - mean) (k_project's expression), at the point and below it
      T pc = p0;
      if constexpr (PL) {
        // phi at the point and below it along x and y (along z the carried
        // pressure has it)
        static_assert(!PL || PCARRY, "the fused kernels carry p + phi from plane to plane");
        const T* cur = &phl[k % NSP][0];
        const int ix = (ty + 2) * PHX + tx + 2;
        f0 = cur[ix];
        fm[0] = cur[ix - 1];
        fm[1] = cur[ix - PHX];
      }
      if constexpr (PROJ) {
        pc = p0 + (f0 - mv);
#pragma unroll
        for (int c = 0; c < (PCARRY ? 2 : 3); c++) pm[c] = pm[c] + (fm[c] - mv);
      }
      // p at plane k-1 of this column: the previous plane's (loaded before
      // the loop for the first plane)
      if constexpr (KIND <= 3 && PCARRY) {
        pm[2] = p_prev;
        p_prev = pc;
      }
      const int slot[3] = {(k + NS - 1) % NS, k % NS, (k + 1) % NS};
      auto tap = [&](int f, int di, int dj, int dk) -> T { return pl[slot[dk + 1]][f][ty + 1 + dj][tx + 1 + di]; };
#pragma unroll
      for (int c = 0; c < 3; c++) {
        if (!EVERY && !inside[c]) continue;
        T rhs = c == 0 ? rhs_u<T>(g, tap) : (c == 1 ? rhs_v<T>(g, tap) : rhs_w<T>(g, tap));
        const T x = tap(c, 0, 0, 0);
        if constexpr (KIND <= 3) {
          rhs += (T)s.force[c];
          // pressure_gradient_{u,v,w} (include/PressureGradient.h:9-25)
          const T pg = (pc - pm[c]) * (T)g.one_over_h[c];
          T o1, o2 = rhs;
          if constexpr (KIND == 1) {
            o1 = x + (T)s.a1 * rhs - (T)s.b * pg;
          } else if constexpr (KIND == 2) {
            o2 = (T)s.a2 * rhs;
            o1 = x + (T)s.a1 * av[c] + o2 - (T)s.b * pg;
          } else {
            o1 = x + (-av[c]) + (T)s.a2 * rhs - (T)s.b * pg;
          }
          if constexpr (EVERY) {
            // unconditional stores, so that the loop has no branches: a
            // point outside component c's interior is a boundary-shell
            // point, which the boundary pass rewrites after the stage (O1)
            // or nothing reads (O2); threads outside the tile store at the
            // shell point (0, 0, k)
            const long os = in_tile ? o : L.idx(0, 0, k);
            store_out<NT>(&((T*)s.O1[c])[os], o1);
            if constexpr (KIND <= 2) store_out<NT>(&((T*)s.O2[c])[os], o2);
          } else {
            store_out<NT>(&((T*)s.O1[c])[o], o1);
            if constexpr (KIND <= 2) store_out<NT>(&((T*)s.O2[c])[o], o2);
          }
        } else {
          // calculate_momentum_rhs_with_forcing_* (include/MomentumEquationForcing.h:10-31)
          const T dt = (T)vs.dt;
          double fc;
          if constexpr (FT) {
            // component c sits at coordinates shifted on axis c
            const double* tb = vs.trig;
            const int tn = vs.tn;
            auto tab = [&](int a, int q, int idx) { return tb[((a * 2 + (a == c ? 1 : 0)) * 2 + q) * tn + idx]; };
            fc = vt_forcing(c, tab(0, 0, i), tab(0, 1, i), tab(1, 0, j), tab(1, 1, j), tab(2, 0, k), tab(2, 1, k),
                            sin(vs.t), cos(vs.t), vs.forcing.Re);
          } else {
            fc = eval_at(g, vs.forcing, c, c, vs.t, i, j, k);
          }
          const T f = rhs + (T)fc;
          if constexpr (KIND == 4) {
            ((T*)vs.R[c])[o] = f;
            ((T*)vs.B[c])[o] = x + dt * (T)(8.0 / 15.0) * f;
          } else if constexpr (KIND == 5) {
            ((T*)vs.R[c])[o] = bv[c] + dt * ((T)(1.0 / 4.0) * av[c]);
            ((T*)vs.V[c])[o] = bv[c] + dt * ((T)(1.0 / 4.0) * av[c] + (T)(5.0 / 12.0) * f);
          } else {
            ((T*)vs.B[c])[o] = av[c] + dt * ((T)(3.0 / 4.0) * f);
          }
        }
      }
      if constexpr (PROJ) {
        // threads outside the tile store into the padding column i = N0 of
        // the pressure, beyond its extents
        const long op = EVERY && !in_tile ? L.idx(g.N[0], 0, k) : o;
        store_out<NT>(&((T*)s.pout)[op], pc);
      }
    }
    // plane k+2 goes to the slot of plane k-1 (three slots: once every
    // thread is done with plane k) or of plane k-2 (four and more slots: no
    // thread has read it since the previous barrier)
    if constexpr (NS == 3) __syncthreads();
    if (more) commit(k + 2, pre);
    // PL: phi plane k+3 into the slot of plane k-1 (the pressure carried
    // from plane to plane leaves plane k-1 of phi unread here)
    if constexpr (PL)
      if (k + 3 <= k1) pcommit(k + 3, pq);
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Boundary conditions: VelocityTensor::apply_bc (src/VelocityTensor.cpp:
// 36-233) followed by apply_periodic_bc (src/StaggeredTensor.cpp:221-256),
// in one pass over the boundary shell of the three components.
//
// The reference writes the six Dirichlet faces in the order z-, z+, y-, y+,
// x-, x+ (so at an edge the later face wins) and then copies ghost planes
// x, then y, then z (plane 0 <- plane s-2, plane s-1 <- plane 1).  A ghost
// point therefore ends up holding the value of the point whose periodic
// coordinates are mapped (0 -> s-2, s-1 -> 1) on every periodic axis; that
// source point is either interior (untouched by the BC) or on a Dirichlet
// face (its value is a function evaluation), so every shell point is
// computed independently and the pass needs no ordering between threads.
// ---------------------------------------------------------------------------
__host__ __device__ inline int rim_out_count(int n, int mx) { return 1 + max(0, n - 1 - mx); }
__device__ inline int rim_out_index(int t, int mx) { return t == 0 ? 0 : mx + t; }

// The extents of the stage kernels' tiles (the union of the components'
// interiors) and the rim's point counts per group.
struct RimShape {
  int imax, jmax, kmax;
  long count[3];
};
inline RimShape rim_shape(const DGrid& g) {
  int eu[3], ev[3], ew[3];
  extents(g, 0, eu);
  extents(g, 1, ev);
  extents(g, 2, ew);
  RimShape r;
  r.imax = max(eu[0], max(ev[0], ew[0])) - 2;
  r.jmax = max(eu[1], max(ev[1], ew[1])) - 2;
  r.kmax = max(eu[2], max(ev[2], ew[2])) - 2;
  const int n0 = g.N[0], n1 = g.N[1], n2 = g.N[2];
  r.count[0] = (long)rim_out_count(n2, r.kmax) * n0 * n1;
  r.count[1] = (long)rim_out_count(n1, r.jmax) * n0 * min(r.kmax, n2 - 1);
  r.count[2] = (long)rim_out_count(n0, r.imax) * min(r.jmax, n1 - 1) * min(r.kmax, n2 - 1);
  return r;
}

// Point q of rim group grp: pout = pin + (phi - mean).
template <typename T>
__device__ inline void p_rim_point(const DGrid& g, const Layout& L, int grp, long q, const T* __restrict__ pin,
                                   const T* __restrict__ phi, T* __restrict__ pout, const Red3* __restrict__ mean,
                                   int imax, int jmax, int kmax) {
  const int n0 = g.N[0], n1 = g.N[1], n2 = g.N[2];
  const int ki = min(kmax, n2 - 1), ji = min(jmax, n1 - 1);  // inner ranges [1, ki], [1, ji]
  int i, j, k;
  if (grp == 0) {
    if (q >= (long)rim_out_count(n2, kmax) * n0 * n1) return;
    i = (int)(q % n0);
    j = (int)((q / n0) % n1);
    k = rim_out_index((int)(q / ((long)n0 * n1)), kmax);
  } else if (grp == 1) {
    if (q >= (long)rim_out_count(n1, jmax) * n0 * ki) return;
    i = (int)(q % n0);
    const long r = q / n0;
    j = rim_out_index((int)(r % rim_out_count(n1, jmax)), jmax);
    k = 1 + (int)(r / rim_out_count(n1, jmax));
  } else {
    const int no = rim_out_count(n0, imax);
    if (q >= (long)no * ji * ki) return;
    i = rim_out_index((int)(q % no), imax);
    const long r = q / no;
    j = 1 + (int)(r % ji);
    k = 1 + (int)(r / ji);
  }
  const T mv = mean ? (T)(mean->s1 / mean->s2) : (T)0;
  const long o = L.idx(i, j, k);
  pout[o] = pin[o] + (phi[o] - mv);
}

template <typename T>
__global__ void __launch_bounds__(256) k_p_rim(DGrid g, Layout L, const T* __restrict__ pin, const T* __restrict__ phi,
                                               T* __restrict__ pout, const Red3* __restrict__ mean, int imax, int jmax,
                                               int kmax) {
  p_rim_point<T>(g, L, blockIdx.y, (long)blockIdx.x * blockDim.x + threadIdx.x, pin, phi, pout, mean, imax, jmax,
                 kmax);
}

__device__ inline double face_value(const DGrid& g, const FnDesc& f, double t, int c, int face, int i, int j, int k) {
  // faces 0,1 = z-,z+; 2,3 = y-,y+; 4,5 = x-,x+
  const int axis = 2 - (face >> 1);
  if (axis != c) return eval_at(g, f, c, c, t, i, j, k);  // tangential: exact value
  const bool hi = face & 1;
  if (c == 2) {
    const int kn = hi ? g.N[2] - 1 : 0;
    const double wb = eval_at(g, f, 3, 2, t, i, j, kn);
    const double du_dx = (eval_at(g, f, 0, 0, t, i + 1, j, kn) - eval_at(g, f, 0, 0, t, i, j, kn)) * g.one_over_h[0];
    const double dv_dy = (eval_at(g, f, 1, 1, t, i, j + 1, kn) - eval_at(g, f, 1, 1, t, i, j, kn)) * g.one_over_h[1];
    return hi ? wb - g.h_over_2[2] * (du_dx + dv_dy) : wb + g.h_over_2[2] * (du_dx + dv_dy);
  }
  if (c == 1) {
    const int jn = hi ? g.N[1] - 1 : 0;
    const double vb = eval_at(g, f, 3, 1, t, i, jn, k);
    const double du_dx = (eval_at(g, f, 0, 0, t, i + 1, jn, k) - eval_at(g, f, 0, 0, t, i, jn, k)) * g.one_over_h[0];
    const double dw_dz = (eval_at(g, f, 2, 2, t, i, jn, k + 1) - eval_at(g, f, 2, 2, t, i, jn, k)) * g.one_over_h[2];
    return hi ? vb - g.h_over_2[1] * (du_dx + dw_dz) : vb + g.h_over_2[1] * (du_dx + dw_dz);
  }
  // c == 0; the x+ wall node is Nx_domains = Nx_global - 1, derivatives at Nx - 1
  const int in = hi ? g.N[0] - 1 : 0;
  const double ub = eval_at(g, f, 3, 0, t, hi ? g.Ng[0] - 1 : 0, j, k);
  const double dv_dy = (eval_at(g, f, 1, 1, t, in, j + 1, k) - eval_at(g, f, 1, 1, t, in, j, k)) * g.one_over_h[1];
  const double dw_dz = (eval_at(g, f, 2, 2, t, in, j, k + 1) - eval_at(g, f, 2, 2, t, in, j, k)) * g.one_over_h[2];
  return hi ? ub - g.h_over_2[0] * (dv_dy + dw_dz) : ub + g.h_over_2[0] * (dv_dy + dw_dz);
}

struct BCArgs {
  void* F[3];
  FnDesc f;
  double t;
  int count[3][3];  // threads (opposite shell points) per (component, face pair)
};

// Face data of a boundary pass: BC_ZERO homogeneous data (no-slip walls),
// every face value is 0 and the function evaluation code is compiled out,
// which keeps the kernel small; BC_FN closed-form descriptors evaluated on
// the device; BC_TABLE values evaluated on the host (FnDesc kFnFaceTable).
enum BCMode { BC_ZERO = 0, BC_FN = 1, BC_TABLE = 2 };

// Value of shell point (i, j, k) of component c: a Dirichlet face value, or
// the value at its periodic source, which is then an interior point that
// this pass does not write.
template <typename T, int MODE>
__device__ inline T shell_value(const DGrid& g, const Layout& L, const BCArgs& a, int c, const int e[3], int i, int j,
                                int k, const T* F) {
  // periodic ghost source
  int s[3] = {i, j, k};
  for (int ax = 0; ax < 3; ax++) {
    if (!g.periodic[ax]) continue;
    if (s[ax] == 0)
      s[ax] = e[ax] - 2;
    else if (s[ax] == e[ax] - 1)
      s[ax] = 1;
  }
  // last Dirichlet face containing the source point
  int face = -1;
  if (!g.periodic[2]) {
    if (s[2] == 0) face = 0;
    if (s[2] == e[2] - 1) face = 1;
  }
  if (!g.periodic[1]) {
    if (s[1] == 0) face = 2;
    if (s[1] == e[1] - 1) face = 3;
  }
  if (!g.periodic[0]) {
    if (s[0] == 0) face = 4;
    if (s[0] == e[0] - 1) face = 5;
  }
  if (face >= 0) {
    if constexpr (MODE == BC_ZERO)
      return (T)0;
    else if constexpr (MODE == BC_TABLE)
      return (T)a.f.tab[a.f.toff[c][face] + face_plane_index(face, e, s[0], s[1], s[2])];
    else
      return (T)face_value(g, a.f, a.t, c, face, s[0], s[1], s[2]);
  }
  return F[L.idx(s[0], s[1], s[2])];
}

// Each thread owns the two opposite shell points of one (pair, position):
// pair 0 = the k planes (all i, j), pair 1 = the j planes (all i, inner k),
// pair 2 = the i planes (inner j, inner k).  On a periodic axis each of the
// two points is the other's source-side neighbour (ghost 0 and point 1 share
// a row segment, as do e-2 and ghost e-1), so both reads come first and the
// two writes then land on lines the same thread has just read.
// RIM: blocks with blockIdx.y >= 9 write the pressure rim of the fused
// stage kernel before this pass (PressureRim), in the same launch.
template <typename T, int MODE, bool RIM = false>
__global__ void __launch_bounds__(256) k_bc(DGrid g, Layout L, BCArgs a, PressureRim rim, RimShape rs) {
  if constexpr (RIM) {
    if (blockIdx.y >= 9) {
      p_rim_point<T>(g, L, blockIdx.y - 9, (long)blockIdx.x * blockDim.x + threadIdx.x, (const T*)rim.pin,
                     (const T*)rim.phi, (T*)rim.pout, rim.mean, rs.imax, rs.jmax, rs.kmax);
      return;
    }
  }
  const int c = blockIdx.y / 3, pair = blockIdx.y % 3;
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= a.count[c][pair]) return;
  int e[3];
  extents(g, c, e);
  int i, j, k, axis;
  if (pair == 0) {
    i = q % e[0];
    j = q / e[0];
    k = 0;
    axis = 2;
  } else if (pair == 1) {
    i = q % e[0];
    k = q / e[0] + 1;
    j = 0;
    axis = 1;
  } else {
    j = q % (e[1] - 2) + 1;
    k = q / (e[1] - 2) + 1;
    i = 0;
    axis = 0;
  }
  int p1[3] = {i, j, k};
  p1[axis] = e[axis] - 1;
  T* F = (T*)a.F[c];
  const T v0 = shell_value<T, MODE>(g, L, a, c, e, i, j, k, F);
  const T v1 = shell_value<T, MODE>(g, L, a, c, e, p1[0], p1[1], p1[2], F);
  F[L.idx(i, j, k)] = v0;
  F[L.idx(p1[0], p1[1], p1[2])] = v1;
}

// Periodic ghost copies of one field (StaggeredTensor::apply_periodic_bc,
// src/StaggeredTensor.cpp:221-256), same composition argument as above.
template <typename T>
__global__ void __launch_bounds__(256) k_periodic(DGrid g, Layout L, T* F, int stag, int count0, int count1, int count2) {
  const int pair = blockIdx.y;
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  const int cnt = pair == 0 ? count0 : (pair == 1 ? count1 : count2);
  if (q >= cnt) return;
  int e[3];
  extents(g, stag, e);
  int i, j, k;
  if (pair == 0) {
    const int area = e[0] * e[1];
    const int side = q / area, r = q % area;
    i = r % e[0];
    j = r / e[0];
    k = side ? e[2] - 1 : 0;
  } else if (pair == 1) {
    const int area = e[0] * (e[2] - 2);
    const int side = q / area, r = q % area;
    i = r % e[0];
    k = r / e[0] + 1;
    j = side ? e[1] - 1 : 0;
  } else {
    const int area = (e[1] - 2) * (e[2] - 2);
    const int side = q / area, r = q % area;
    j = r % (e[1] - 2) + 1;
    k = r / (e[1] - 2) + 1;
    i = side ? e[0] - 1 : 0;
  }
  int s[3] = {i, j, k};
  bool ghost = false;
  for (int ax = 0; ax < 3; ax++) {
    if (!g.periodic[ax]) continue;
    if (s[ax] == 0) {
      s[ax] = e[ax] - 2;
      ghost = true;
    } else if (s[ax] == e[ax] - 1) {
      s[ax] = 1;
      ghost = true;
    }
  }
  if (ghost) F[L.idx(i, j, k)] = F[L.idx(s[0], s[1], s[2])];
}

// ---------------------------------------------------------------------------
// Field fill from a function (StaggeredTensor::set, src/StaggeredTensor.cpp:
// 205-219).
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256) k_set(DGrid g, Layout L, T* F, int stag, int comp, FnDesc f, double t, int border) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  const int k = blockIdx.z;
  int e[3];
  extents(g, stag, e);
  if (i >= e[0]) return;
  const int lo = border ? 0 : 1;
  if (i < lo || j < lo || k < lo || i > e[0] - 1 - lo || j > e[1] - 1 - lo || k > e[2] - 1 - lo) return;
  F[L.idx(i, j, k)] = (T)eval_at(g, f, stag, comp, t, i, j, k);
}

// ---------------------------------------------------------------------------
// Pressure right-hand side: div(u)/dt on owner points (compute_rhs_
// homogeneous_periodic, src/PressureEquation.cpp:59-61, divergence per
// include/VelocityDivergence.h:9-20), plus the Neumann boundary data of
// apply_non_homogeneous_neumann (src/PressureEquation.cpp:10-57) when
// gp.kind != ZERO, written packed into the spectral buffer (owner point
// (i,j,k) -> (i-olo0, j-olo1, k-olo2), PressureTensor::copy_from_staggered,
// src/PressureTensor.cpp:16-19).
// ---------------------------------------------------------------------------
template <typename T, typename S, int NHN>
__global__ void __launch_bounds__(256) k_div_rhs(DGrid g, Layout L, const T* __restrict__ u, const T* __restrict__ v,
                                                 const T* __restrict__ w, S* __restrict__ B, long bp0, long bplane,
                                                 double dt, FnDesc gp, double tg) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;  // packed owner coordinates
  const int b = blockIdx.y;
  const int c = blockIdx.z;
  if (a >= g.npts[0]) return;
  const int i = a + g.olo[0], j = b + g.olo[1], k = c + g.olo[2];
  const long o = L.idx(i, j, k);
  const double du_dx = ((double)u[o + 1] - (double)u[o]) * g.one_over_h[0];
  const double dv_dy = ((double)v[o + L.PX] - (double)v[o]) * g.one_over_h[1];
  const double dw_dz = ((double)w[o + L.plane] - (double)w[o]) * g.one_over_h[2];
  double r = (du_dx + dv_dy + dw_dz) / dt;
  if (NHN == 1) {
    const int sx = g.N[0], sy = g.N[1], sz = g.N[2];
    if (k == 0) r += eval_at(g, gp, 3, 2, tg, i, j, k) * 2.0 * g.one_over_h[2];
    if (k == sz - 1) r -= eval_at(g, gp, 3, 2, tg, i, j, k) * 2.0 * g.one_over_h[2];
    if (j == 0) r += eval_at(g, gp, 3, 1, tg, i, j, k) * 2.0 * g.one_over_h[1];
    if (j == sy - 1) r -= eval_at(g, gp, 3, 1, tg, i, j, k) * 2.0 * g.one_over_h[1];
    if (i == 0) r += eval_at(g, gp, 3, 0, tg, i, j, k) * 2.0 * g.one_over_h[0];
    if (i == sx - 1) r -= eval_at(g, gp, 3, 0, tg, i, j, k) * 2.0 * g.one_over_h[0];
  } else if (NHN == 2) {
    // host-evaluated normal gradient components at the wall nodes
    const int e[3] = {g.N[0], g.N[1], g.N[2]};
    const double* t = gp.tab;
    if (k == 0) r += t[gp.toff[0][0] + face_plane_index(0, e, i, j, k)] * 2.0 * g.one_over_h[2];
    if (k == e[2] - 1) r -= t[gp.toff[0][1] + face_plane_index(1, e, i, j, k)] * 2.0 * g.one_over_h[2];
    if (j == 0) r += t[gp.toff[0][2] + face_plane_index(2, e, i, j, k)] * 2.0 * g.one_over_h[1];
    if (j == e[1] - 1) r -= t[gp.toff[0][3] + face_plane_index(3, e, i, j, k)] * 2.0 * g.one_over_h[1];
    if (i == 0) r += t[gp.toff[0][4] + face_plane_index(4, e, i, j, k)] * 2.0 * g.one_over_h[0];
    if (i == e[0] - 1) r -= t[gp.toff[0][5] + face_plane_index(5, e, i, j, k)] * 2.0 * g.one_over_h[0];
  }
  B[a + bp0 * b + bplane * c] = (S)r;
}

// Copy owner values of a pressure field into the packed buffer (used when
// the right-hand side is given directly, as in the stand-alone Poisson solve).
template <typename T, typename S>
__global__ void __launch_bounds__(256) k_pack(DGrid g, Layout L, const T* __restrict__ P, S* __restrict__ B, long bp0,
                                              long bplane) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y, c = blockIdx.z;
  if (a >= g.npts[0]) return;
  B[a + bp0 * b + bplane * c] = (S)P[L.idx(a + g.olo[0], b + g.olo[1], c + g.olo[2])];
}

// Unpack the packed solution into every point of a pressure field: owner
// points and, through the periodic map, the ghost planes
// (PressureTensor::copy_to_staggered, src/PressureTensor.cpp:21-34).
template <typename T, typename S>
__global__ void __launch_bounds__(256) k_unpack(DGrid g, Layout L, const S* __restrict__ B, long bp0, long bplane,
                                                T* __restrict__ P) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y, k = blockIdx.z;
  if (i >= g.N[0]) return;
  int s[3] = {i, j, k};
  for (int ax = 0; ax < 3; ax++) {
    if (!g.periodic[ax]) continue;
    if (s[ax] == 0) s[ax] = g.N[ax] - 2;
    else if (s[ax] == g.N[ax] - 1) s[ax] = 1;
    s[ax] -= 1;
  }
  P[L.idx(i, j, k)] = (T)B[s[0] + bp0 * s[1] + bplane * s[2]];
}

// ---------------------------------------------------------------------------
// Projection: p += phi on all points (update_pressure, src/Timestep.cpp:
// 79-81) and u -= dt_s * grad(phi) on interior points
// (update_velocity_pressure_gradient, src/Timestep.cpp:66-76).
// ---------------------------------------------------------------------------
// One thread per (i, j) position of the padded plane marches a chunk of k
// (the launcher's chunk is one plane): phi at k-1 stays in a register, and
// the operands of plane k+1 are loaded while plane k is updated.  Loads read
// always-valid addresses and the stores carry the point tests, so a plane's
// loads are issued together.
template <typename T, bool NT>
// The incremented pressure goes to pout: p itself, or another buffer (every
// pressure point is written).
__global__ void __launch_bounds__(256) k_project(DGrid g, Layout L, T* __restrict__ u, T* __restrict__ v,
                                                 T* __restrict__ w, const T* p, const T* __restrict__ phi,
                                                 double dts, int nqt, int kchunk, const Red3* __restrict__ mean,
                                                 T* pout) {
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const int qt = tile % nqt, kt = tile / nqt;
  // slices start at the allocation's aligned origin, L.lead elements
  // before element 0
  const int q = qt * 256 + threadIdx.x - L.lead;
  if (q < 0) return;
  const int i = q % L.PX, j = q / L.PX;
  if (j > g.N[1] || i > g.N[0]) return;
  const int k0 = kt * kchunk, k1 = min(k0 + kchunk, g.N[2] + 1);
  struct Pt {
    T f, fx, fy, pp, a[3];
  };
  auto fetch = [&](int k, Pt& d) {
    const long o = (long)q + L.plane * k;
    d.f = phi[o];
    d.fx = phi[o >= 1 ? o - 1 : o];
    d.fy = phi[o >= L.PX ? o - L.PX : o];
    d.pp = load_in<NT && kNtLoads>(&p[o]);
    d.a[0] = load_in<NT && kNtLoads>(&u[o]);
    d.a[1] = load_in<NT && kNtLoads>(&v[o]);
    d.a[2] = load_in<NT && kNtLoads>(&w[o]);
  };
  T fz = k0 >= 1 ? phi[(long)q + L.plane * (k0 - 1)] : (T)0;
  // the increment's weighted mean, when the solve left it in phi (the mixed
  // form's one-sweep defect correction): p += phi - mean
  const T pm = mean ? (T)(mean->s1 / mean->s2) : (T)0;
  Pt cur;
  fetch(k0, cur);
  T* F[3] = {u, v, w};
  for (int k = k0; k < k1; k++) {
    Pt nxt = cur;
    if (k + 1 < k1) fetch(k + 1, nxt);
    const long o = (long)q + L.plane * k;
    if (i < g.N[0] && j < g.N[1] && k < g.N[2]) store_out<NT>(&pout[o], cur.pp + (cur.f - pm));
    const T d[3] = {cur.f - cur.fx, cur.f - cur.fy, cur.f - fz};
#pragma unroll
    for (int c = 0; c < 3; c++)
      if (interior(g, c, i, j, k)) store_out<NT>(&F[c][o], cur.a[c] - d[c] * (T)g.one_over_h[c] * (T)dts);
    fz = cur.f;
    cur = nxt;
  }
}

// Pressure-gradient field on interior points (test/full_test.cpp:141-142).
template <typename T>
__global__ void __launch_bounds__(256) k_pgrad(DGrid g, Layout L, const T* __restrict__ p, T* u, T* v, T* w, int imax,
                                               int jmax, int kmax) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  const int k = blockIdx.z;
  if (i > imax || j > jmax || k > kmax) return;
  const long o = L.idx(i, j, k);
  const long doff[3] = {1, L.PX, L.plane};
  T* F[3] = {u, v, w};
#pragma unroll
  for (int c = 0; c < 3; c++) {
    if (!interior(g, c, i, j, k)) continue;
    F[c][o] = (p[o] - p[o - doff[c]]) * (T)g.one_over_h[c];
  }
}

// ---------------------------------------------------------------------------
// Host-layout transfers: dense reference layout (x fastest, exact extents,
// FP64) <-> padded device layout of type T.
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256) k_from_dense(Layout L, int e0, int e1, const double* __restrict__ src, T* dst) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y, k = blockIdx.z;
  if (i >= e0) return;
  dst[L.idx(i, j, k)] = (T)src[i + (long)e0 * (j + (long)e1 * k)];
}

template <typename T>
__global__ void __launch_bounds__(256) k_to_dense(Layout L, int e0, int e1, const T* __restrict__ src, double* dst) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y, k = blockIdx.z;
  if (i >= e0) return;
  dst[i + (long)e0 * (j + (long)e1 * k)] = (double)src[L.idx(i, j, k)];
}

template <typename T>
__global__ void __launch_bounds__(256) k_copy(long n, const T* __restrict__ src, T* __restrict__ dst) {
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) dst[q] = src[q];
}

template <typename T>
__global__ void __launch_bounds__(256) k_add_const(DGrid g, Layout L, T* P, double c) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y, k = blockIdx.z;
  if (i >= g.N[0]) return;
  P[L.idx(i, j, k)] += (T)c;
}

// ---------------------------------------------------------------------------
// Reductions: wave64 shuffles, then LDS across the block, then one partial
// per block; a single block folds the partials in a fixed order so results
// are deterministic.
// ---------------------------------------------------------------------------
__device__ inline Red3 red_combine(Red3 a, Red3 b) { return {a.s1 + b.s1, a.s2 + b.s2, fmax(a.mx, b.mx)}; }

__device__ inline Red3 block_reduce(Red3 v) {
  __shared__ Red3 sh[16];
  for (int off = 32; off > 0; off >>= 1) {
    Red3 o;
    o.s1 = __shfl_down(v.s1, off, 64);
    o.s2 = __shfl_down(v.s2, off, 64);
    o.mx = __shfl_down(v.mx, off, 64);
    v = red_combine(v, o);
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  Red3 r{0, 0, 0};
  if (threadIdx.x == 0) {
    const int nw = (blockDim.x + 63) >> 6;
    for (int q = 0; q < nw; q++) r = red_combine(r, sh[q]);
  }
  return r;
}

// ErrorL1/L2/LInfNorm of the velocity (src/Norms.cpp:11-85): errors after
// interpolating each component to the interior nodes 1..N-2.  The exact
// value comes from f, or, when tab.f[0] is not null, from three node-shaped
// tables holding the exact components at the nodes (callers whose exact
// solution is a host callback tabulate it).
template <typename T>
__global__ void __launch_bounds__(256) k_norm_velocity(DGrid g, Layout L, const T* __restrict__ u,
                                                       const T* __restrict__ v, const T* __restrict__ w, FnDesc f,
                                                       double t, ExactTables tab, Red3* partial) {
  const long n0 = g.N[0] - 2, n1 = g.N[1] - 2, n2 = g.N[2] - 2;
  const long total = n0 * n1 * n2;
  Red3 acc{0, 0, 0};
  for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < total; q += (long)gridDim.x * blockDim.x) {
    const int i = (int)(q % n0) + 1, j = (int)((q / n0) % n1) + 1, k = (int)(q / (n0 * n1)) + 1;
    const double x = g.min[0] + (double)(g.base[0] + i) * g.h[0];
    const double y = g.min[1] + (double)(g.base[1] + j) * g.h[1];
    const double z = g.min[2] + (double)(g.base[2] + k) * g.h[2];
    const long o = L.idx(i, j, k);
    const double iu = ((double)u[o] + (double)u[o + 1]) / 2.0;
    const double iv = ((double)v[o] + (double)v[o + L.PX]) / 2.0;
    const double iw = ((double)w[o] + (double)w[o + L.plane]) / 2.0;
    double ex[3];
    if (tab.f[0]) {
      for (int c = 0; c < 3; c++) ex[c] = (double)((const T*)tab.f[c])[o];
    } else {
      for (int c = 0; c < 3; c++) ex[c] = fn_eval(f, c, t, x, y, z);
    }
    const double eu = ex[0] - iu;
    const double ev = ex[1] - iv;
    const double ew = ex[2] - iw;
    const double e2 = eu * eu + ev * ev + ew * ew;
    acc.s1 += sqrt(e2);
    acc.s2 += e2;
    acc.mx = fmax(acc.mx, fmax(fabs(eu), fmax(fabs(ev), fabs(ew))));
  }
  const Red3 r = block_reduce(acc);
  if (threadIdx.x == 0) partial[blockIdx.x] = r;
}

// Pressure error over owner points (src/Norms.cpp:88-120); in mode 1 it
// accumulates sum(f - p) for adjust_pressure (src/PressureEquation.cpp:
// 288-295).  With tab.f[0] not null the exact values come from that
// pressure-shaped table instead of f.
template <typename T>
__global__ void __launch_bounds__(256) k_norm_pressure(DGrid g, Layout L, const T* __restrict__ p, FnDesc f, double t,
                                                       int adjust, ExactTables tab, Red3* partial) {
  const long n0 = g.npts[0], n1 = g.npts[1], n2 = g.npts[2];
  const long total = n0 * n1 * n2;
  Red3 acc{0, 0, 0};
  for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < total; q += (long)gridDim.x * blockDim.x) {
    const int i = (int)(q % n0) + g.olo[0], j = (int)((q / n0) % n1) + g.olo[1], k = (int)(q / (n0 * n1)) + g.olo[2];
    double err;
    if (tab.f[0]) {
      err = (double)((const T*)tab.f[0])[L.idx(i, j, k)] - (double)p[L.idx(i, j, k)];
    } else if (adjust) {
      err = eval_at(g, f, 3, 0, t, i, j, k) - (double)p[L.idx(i, j, k)];
    } else {
      const double x = g.min[0] + (double)(g.base[0] + i) * g.h[0];
      const double y = g.min[1] + (double)(g.base[1] + j) * g.h[1];
      const double z = g.min[2] + (double)(g.base[2] + k) * g.h[2];
      err = fn_eval(f, 0, t, x, y, z) - (double)p[L.idx(i, j, k)];
    }
    acc.s1 += adjust ? err : fabs(err);
    acc.s2 += err * err;
    acc.mx = fmax(acc.mx, fabs(err));
  }
  const Red3 r = block_reduce(acc);
  if (threadIdx.x == 0) partial[blockIdx.x] = r;
}

// sum over a field's extents of a * b (FP64 accumulation)
template <typename T>
__global__ void __launch_bounds__(256) k_dot(Layout L, int e0, int e1, int e2, const T* __restrict__ a,
                                             const T* __restrict__ b, Red3* partial) {
  const long total = (long)e0 * e1 * e2;
  Red3 acc{0, 0, 0};
  for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < total; q += (long)gridDim.x * blockDim.x) {
    const int i = (int)(q % e0), j = (int)((q / e0) % e1), k = (int)(q / ((long)e0 * e1));
    const long o = L.idx(i, j, k);
    acc.s1 += (double)a[o] * (double)b[o];
  }
  const Red3 r = block_reduce(acc);
  if (threadIdx.x == 0) partial[blockIdx.x] = r;
}

__global__ void __launch_bounds__(256) k_reduce_final(const Red3* partial, int n, Red3* out) {
  Red3 acc{0, 0, 0};
  for (int q = threadIdx.x; q < n; q += blockDim.x) acc = red_combine(acc, partial[q]);
  const Red3 r = block_reduce(acc);
  if (threadIdx.x == 0) *out = r;
}

// ---------------------------------------------------------------------------
// Host launchers.
// ---------------------------------------------------------------------------
template <typename T, int KIND, bool FT = false, bool NT = false>
static void launch_stage_kind(const DGrid& g, const Layout& L, const StageArgs& s, const VelStageArgs& vs,
                              hipStream_t st) {
  // union of the three components' interiors
  int eu[3], ev[3], ew[3];
  extents(g, 0, eu);
  extents(g, 1, ev);
  extents(g, 2, ew);
  const int imax = max(eu[0], max(ev[0], ew[0])) - 2;
  const int jmax = max(eu[1], max(ev[1], ew[1])) - 2;
  const int kmax = max(eu[2], max(ev[2], ew[2])) - 2;
  constexpr int STX = StageTile<T>::STX, STY = StageTile<T>::STY;
  const int ntx = (imax + STX - 1) / STX, nty = (jmax + STY - 1) / STY;
  const dim3 blk(STX, STY, 1);
  auto go = [&](auto vec, auto proj) {
    constexpr bool V = decltype(vec)::value, P = decltype(proj)::value;
    constexpr int SKB = stage_planes<T, KIND, NT, P>();
    const int ntiles = ntx * nty * ((kmax + SKB - 1) / SKB);
    k_stage<T, KIND, FT, NT, V, P><<<ntiles, blk, 0, st>>>(g, L, s, vs, imax, jmax, kmax, ntx, nty, ntiles);
  };
  using yes = std::true_type;
  using no = std::false_type;
  // the full-step kinds take the previous stage's projection when s.phi is
  // set (kind 1: the last projection of the previous step, multi-step calls)
  constexpr bool kFuse = KIND <= 3;
  // the 16-byte loads need element 1 of every row on a 16-byte boundary:
  // the storage shift puts it on a 128-byte one, and rows are 64-byte
  // multiples.  Measured against 4- and 8-byte loads (step times): 256^3
  // FP64 3.000 -> 2.965 ms (stage kernel 362 -> 350 us), 192^3 FP64 1.400 ->
  // 1.359 ms, 128^3 FP64 0.523 -> 0.509 ms, 192^3 mixed 0.787 -> 0.775 ms;
  // the 256^3 mixed stage kernel (non-temporal stores) 181 -> 185 us (182
  // with the LDS ring padded to six slots, i.e. three workgroups per CU as
  // with scalar loads), so FP32 with non-temporal stores kept the scalar
  // loads (on the FP32 128 x 4 tiles: 176.2 -> 175.9 us, no gain either;
  // the FP32 kernels now take field-major slots, which ignore VEC).  The
  // FP64 velocity-only kinds take them too: 256^3-node step 1.467 -> 1.432
  // ms.
  if constexpr (kStageVec && (KIND <= 3 || sizeof(T) == 8) && (sizeof(T) == 8 || !NT)) {
    if (L.lead * (long)sizeof(T) == 128 - (long)sizeof(T) && (L.PX * sizeof(T)) % 16 == 0) {
      if constexpr (kFuse) {
        if (s.phi) return go(yes{}, yes{});
      }
      return go(yes{}, no{});
    }
  }
  if constexpr (kFuse) {
    if (s.phi) return go(no{}, yes{});
  }
  go(no{}, no{});
}

// ---------------------------------------------------------------------------
// The pressure increment of a projection fused into the next stage
// (update_pressure, src/Timestep.cpp:76-81: p += phi at every point), on the
// points the stage kernel's tiles leave out: those of the pressure's extents
// outside [1, imax] x [1, jmax] x [1, kmax] (imax, jmax, kmax: the union of
// the velocity components' interiors).  pout = pin + (phi - mean).
// Three groups of points: the k planes outside [1, kmax]; within it the j
// rows outside [1, jmax]; within both the i columns outside [1, imax].
// ---------------------------------------------------------------------------
template <typename T>
void launch_p_rim(const DGrid& g, const Layout& L, const T* pin, const T* phi, T* pout, const Red3* mean,
                  hipStream_t st) {
  const RimShape rs = rim_shape(g);
  const long mx = max(rs.count[0], max(rs.count[1], rs.count[2]));
  k_p_rim<T><<<dim3((unsigned)((mx + 255) / 256), 3, 1), 256, 0, st>>>(g, L, pin, phi, pout, mean, rs.imax, rs.jmax,
                                                                      rs.kmax);
}

template <typename T>
void launch_stage(int stage, const DGrid& g, const Layout& L, const StageArgs& s, hipStream_t st) {
  VelStageArgs vs{};
  if (stage == 1)
    L.nt ? launch_stage_kind<T, 1, false, true>(g, L, s, vs, st) : launch_stage_kind<T, 1>(g, L, s, vs, st);
  else if (stage == 2)
    L.nt ? launch_stage_kind<T, 2, false, true>(g, L, s, vs, st) : launch_stage_kind<T, 2>(g, L, s, vs, st);
  else
    L.nt ? launch_stage_kind<T, 3, false, true>(g, L, s, vs, st) : launch_stage_kind<T, 3>(g, L, s, vs, st);
}

template <typename T>
void launch_vel_stage(int stage, const DGrid& g, const Layout& L, const VelStageArgs& vs, hipStream_t st) {
  StageArgs s{};
  const bool ft = vs.trig && vs.forcing.kind == MIF_FN_VT_FORCING;
  if (stage == 1)
    ft ? launch_stage_kind<T, 4, true>(g, L, s, vs, st) : launch_stage_kind<T, 4>(g, L, s, vs, st);
  else if (stage == 2)
    ft ? launch_stage_kind<T, 5, true>(g, L, s, vs, st) : launch_stage_kind<T, 5>(g, L, s, vs, st);
  else
    ft ? launch_stage_kind<T, 6, true>(g, L, s, vs, st) : launch_stage_kind<T, 6>(g, L, s, vs, st);
}

__global__ void __launch_bounds__(256) k_trig_tables(DGrid g, double* __restrict__ trig, int tn) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int a = blockIdx.y / 2, sh = blockIdx.y % 2;
  if (i >= tn) return;
  const double x = coord(g, a, sh ? a : 3, i);
  trig[((a * 2 + sh) * 2 + 0) * tn + i] = sin(x);
  trig[((a * 2 + sh) * 2 + 1) * tn + i] = cos(x);
}

void launch_trig_tables(const DGrid& g, double* trig, int tn, hipStream_t st) {
  k_trig_tables<<<dim3((tn + 255) / 256, 6, 1), 256, 0, st>>>(g, trig, tn);
}

static void shell_counts(const DGrid& g, int stag, int cnt[3]) {
  int e[3];
  extents(g, stag, e);
  cnt[0] = 2 * e[0] * e[1];
  cnt[1] = 2 * e[0] * (e[2] - 2);
  cnt[2] = 2 * (e[1] - 2) * (e[2] - 2);
}

template <typename T>
void launch_bc(const DGrid& g, const Layout& L, void* const F[3], const FnDesc& f, double t, hipStream_t st,
               const PressureRim& rim) {
  BCArgs a{};
  int mx = 0;
  for (int c = 0; c < 3; c++) {
    a.F[c] = F[c];
    shell_counts(g, c, a.count[c]);
    // one thread per pair of opposite shell points
    for (int q = 0; q < 3; q++) {
      a.count[c][q] /= 2;
      mx = max(mx, a.count[c][q]);
    }
  }
  a.f = f;
  a.t = t;
  const RimShape rs = rim_shape(g);
  if (rim.pin) {
    // three more block rows for the pressure rim, as many blocks as either needs
    const long rmx = max(rs.count[0], max(rs.count[1], rs.count[2]));
    const unsigned nbx = (unsigned)max((long)(mx + 255) / 256, (rmx + 255) / 256);
    const dim3 grd(nbx, 12, 1);
    if (f.kind == MIF_FN_ZERO)
      k_bc<T, BC_ZERO, true><<<grd, 256, 0, st>>>(g, L, a, rim, rs);
    else if (f.kind == kFnFaceTable)
      k_bc<T, BC_TABLE, true><<<grd, 256, 0, st>>>(g, L, a, rim, rs);
    else
      k_bc<T, BC_FN, true><<<grd, 256, 0, st>>>(g, L, a, rim, rs);
    return;
  }
  dim3 grd((mx + 255) / 256, 9, 1);
  if (f.kind == MIF_FN_ZERO)
    k_bc<T, BC_ZERO><<<grd, 256, 0, st>>>(g, L, a, rim, rs);
  else if (f.kind == kFnFaceTable)
    k_bc<T, BC_TABLE><<<grd, 256, 0, st>>>(g, L, a, rim, rs);
  else
    k_bc<T, BC_FN><<<grd, 256, 0, st>>>(g, L, a, rim, rs);
}

template <typename T>
void launch_periodic(const DGrid& g, const Layout& L, T* F, int stag, hipStream_t st) {
  if (!g.periodic[0] && !g.periodic[1] && !g.periodic[2]) return;
  int cnt[3];
  shell_counts(g, stag, cnt);
  const int mx = max(cnt[0], max(cnt[1], cnt[2]));
  dim3 grd((mx + 255) / 256, 3, 1);
  k_periodic<T><<<grd, 256, 0, st>>>(g, L, F, stag, cnt[0], cnt[1], cnt[2]);
}

template <typename T>
void launch_set(const DGrid& g, const Layout& L, T* F, int stag, int comp, const FnDesc& f, double t, int border,
                hipStream_t st) {
  int e[3];
  extents(g, stag, e);
  dim3 grd((e[0] + 255) / 256, e[1], e[2]);
  k_set<T><<<grd, 256, 0, st>>>(g, L, F, stag, comp, f, t, border);
}

template <typename T, typename S>
void launch_div_rhs(const DGrid& g, const Layout& L, const T* u, const T* v, const T* w, S* B, long bp0, long bplane,
                    double dt, const FnDesc& gp, double tg, int nhn, hipStream_t st) {
  dim3 grd((g.npts[0] + 255) / 256, g.npts[1], g.npts[2]);
  if (nhn && gp.kind == kFnFaceTable)
    k_div_rhs<T, S, 2><<<grd, 256, 0, st>>>(g, L, u, v, w, B, bp0, bplane, dt, gp, tg);
  else if (nhn)
    k_div_rhs<T, S, 1><<<grd, 256, 0, st>>>(g, L, u, v, w, B, bp0, bplane, dt, gp, tg);
  else
    k_div_rhs<T, S, 0><<<grd, 256, 0, st>>>(g, L, u, v, w, B, bp0, bplane, dt, gp, tg);
}

template <typename T, typename S>
void launch_pack(const DGrid& g, const Layout& L, const T* P, S* B, long bp0, long bplane, hipStream_t st) {
  dim3 grd((g.npts[0] + 255) / 256, g.npts[1], g.npts[2]);
  k_pack<T, S><<<grd, 256, 0, st>>>(g, L, P, B, bp0, bplane);
}

template <typename T, typename S>
void launch_unpack(const DGrid& g, const Layout& L, const S* B, long bp0, long bplane, T* P, hipStream_t st) {
  dim3 grd((g.N[0] + 255) / 256, g.N[1], g.N[2]);
  k_unpack<T, S><<<grd, 256, 0, st>>>(g, L, B, bp0, bplane, P);
}

template <typename T>
void launch_project(const DGrid& g, const Layout& L, T* u, T* v, T* w, T* p, const T* phi, double dts, hipStream_t st,
                    const Red3* mean, T* pout) {
  if (!pout) pout = p;
  // points 0..N per axis (the union of the extents; u's x+ ghost included,
  // the interior tests filter), as 256-position slices of the padded plane
  // times one plane per workgroup: the most workgroups in flight (256^3 FP64:
  // 258 us with chunks of 16 planes, 246 with 8, 235 with 4, 224 with 2, 216
  // with 1; FP32 112, 108, 104, 99, 99)
  constexpr int kchunk = 1;
  const int nqt = (int)((L.plane + L.lead + 255) / 256);
  const int nkt = (g.N[2] + 1 + kchunk - 1) / kchunk;
  if (L.nt)
    k_project<T, true><<<nqt * nkt, 256, 0, st>>>(g, L, u, v, w, p, phi, dts, nqt, kchunk, mean, pout);
  else
    k_project<T, false><<<nqt * nkt, 256, 0, st>>>(g, L, u, v, w, p, phi, dts, nqt, kchunk, mean, pout);
}

template <typename T>
void launch_pgrad(const DGrid& g, const Layout& L, const T* p, T* u, T* v, T* w, hipStream_t st) {
  const int imax = g.N[0], jmax = g.N[1], kmax = g.N[2];
  dim3 blk(64, 4, 1);
  dim3 grd((imax + 1 + 63) / 64, (jmax + 1 + 3) / 4, kmax + 1);
  k_pgrad<T><<<grd, blk, 0, st>>>(g, L, p, u, v, w, imax, jmax, kmax);
}

template <typename T>
void launch_from_dense(const Layout& L, const int e[3], const double* src, T* dst, hipStream_t st) {
  dim3 grd((e[0] + 255) / 256, e[1], e[2]);
  k_from_dense<T><<<grd, 256, 0, st>>>(L, e[0], e[1], src, dst);
}

template <typename T>
void launch_to_dense(const Layout& L, const int e[3], const T* src, double* dst, hipStream_t st) {
  dim3 grd((e[0] + 255) / 256, e[1], e[2]);
  k_to_dense<T><<<grd, 256, 0, st>>>(L, e[0], e[1], src, dst);
}

template <typename T>
void launch_copy(long n, const T* src, T* dst, hipStream_t st) {
  k_copy<T><<<(unsigned)((n + 255) / 256), 256, 0, st>>>(n, src, dst);
}

template <typename T>
void launch_add_const(const DGrid& g, const Layout& L, T* P, double c, hipStream_t st) {
  dim3 grd((g.N[0] + 255) / 256, g.N[1], g.N[2]);
  k_add_const<T><<<grd, 256, 0, st>>>(g, L, P, c);
}

template <typename T>
void launch_norm_velocity(const DGrid& g, const Layout& L, const T* u, const T* v, const T* w, const FnDesc& f,
                          double t, const ExactTables& tab, Red3* partial, int nblocks, Red3* out,
                          hipStream_t st) {
  k_norm_velocity<T><<<nblocks, 256, 0, st>>>(g, L, u, v, w, f, t, tab, partial);
  k_reduce_final<<<1, 256, 0, st>>>(partial, nblocks, out);
}

template <typename T>
void launch_norm_pressure(const DGrid& g, const Layout& L, const T* p, const FnDesc& f, double t, int adjust,
                          const ExactTables& tab, Red3* partial, int nblocks, Red3* out, hipStream_t st) {
  k_norm_pressure<T><<<nblocks, 256, 0, st>>>(g, L, p, f, t, adjust, tab, partial);
  k_reduce_final<<<1, 256, 0, st>>>(partial, nblocks, out);
}

template <typename T>
void launch_dot(const Layout& L, const int e[3], const T* a, const T* b, Red3* partial, int nblocks, Red3* out,
                hipStream_t st) {
  k_dot<T><<<nblocks, 256, 0, st>>>(L, e[0], e[1], e[2], a, b, partial);
  k_reduce_final<<<1, 256, 0, st>>>(partial, nblocks, out);
}

// Values of one field at a list of points (i, j, k), as FP64.
template <typename T>
__global__ void __launch_bounds__(256) k_gather(Layout L, const T* __restrict__ F, const int* __restrict__ ijk,
                                                long n, double* __restrict__ out) {
  const long q = blockIdx.x * 256L + threadIdx.x;
  if (q >= n) return;
  out[q] = (double)F[L.idx(ijk[3 * q], ijk[3 * q + 1], ijk[3 * q + 2])];
}

template <typename T>
void launch_gather(const Layout& L, const T* F, const int* ijk, long n, double* out, hipStream_t st) {
  k_gather<T><<<(unsigned)((n + 255) / 256), 256, 0, st>>>(L, F, ijk, n, out);
}

#define MIF_INSTANTIATE(T)                                                                                          \
  template void launch_stage<T>(int, const DGrid&, const Layout&, const StageArgs&, hipStream_t);                   \
  template void launch_p_rim<T>(const DGrid&, const Layout&, const T*, const T*, T*, const Red3*, hipStream_t);    \
  template void launch_vel_stage<T>(int, const DGrid&, const Layout&, const VelStageArgs&, hipStream_t);            \
  template void launch_bc<T>(const DGrid&, const Layout&, void* const*, const FnDesc&, double, hipStream_t,         \
                             const PressureRim&);                                                                   \
  template void launch_periodic<T>(const DGrid&, const Layout&, T*, int, hipStream_t);                              \
  template void launch_set<T>(const DGrid&, const Layout&, T*, int, int, const FnDesc&, double, int, hipStream_t);  \
  template void launch_project<T>(const DGrid&, const Layout&, T*, T*, T*, T*, const T*, double, hipStream_t,       \
                                  const Red3*, T*);                                                                 \
  template void launch_pgrad<T>(const DGrid&, const Layout&, const T*, T*, T*, T*, hipStream_t);                    \
  template void launch_from_dense<T>(const Layout&, const int*, const double*, T*, hipStream_t);                    \
  template void launch_to_dense<T>(const Layout&, const int*, const T*, double*, hipStream_t);                      \
  template void launch_copy<T>(long, const T*, T*, hipStream_t);                                                    \
  template void launch_dot<T>(const Layout&, const int*, const T*, const T*, Red3*, int, Red3*, hipStream_t);       \
  template void launch_gather<T>(const Layout&, const T*, const int*, long, double*, hipStream_t);                 \
  template void launch_add_const<T>(const DGrid&, const Layout&, T*, double, hipStream_t);                          \
  template void launch_norm_velocity<T>(const DGrid&, const Layout&, const T*, const T*, const T*, const FnDesc&,   \
                                        double, const ExactTables&, Red3*, int, Red3*, hipStream_t);                \
  template void launch_norm_pressure<T>(const DGrid&, const Layout&, const T*, const FnDesc&, double, int,          \
                                        const ExactTables&, Red3*, int, Red3*, hipStream_t);

MIF_INSTANTIATE(double)
MIF_INSTANTIATE(float)

template void launch_div_rhs<double, double>(const DGrid&, const Layout&, const double*, const double*, const double*,
                                             double*, long, long, double, const FnDesc&, double, int, hipStream_t);
template void launch_div_rhs<float, float>(const DGrid&, const Layout&, const float*, const float*, const float*,
                                           float*, long, long, double, const FnDesc&, double, int, hipStream_t);
template void launch_div_rhs<float, double>(const DGrid&, const Layout&, const float*, const float*, const float*,
                                            double*, long, long, double, const FnDesc&, double, int, hipStream_t);
template void launch_pack<double, double>(const DGrid&, const Layout&, const double*, double*, long, long, hipStream_t);
template void launch_pack<float, float>(const DGrid&, const Layout&, const float*, float*, long, long, hipStream_t);
template void launch_unpack<double, double>(const DGrid&, const Layout&, const double*, long, long, double*,
                                            hipStream_t);
template void launch_unpack<float, float>(const DGrid&, const Layout&, const float*, long, long, float*, hipStream_t);
template void launch_unpack<float, double>(const DGrid&, const Layout&, const double*, long, long, float*, hipStream_t);
template void launch_unpack<double, float>(const DGrid&, const Layout&, const float*, long, long, double*, hipStream_t);
template void launch_pack<float, double>(const DGrid&, const Layout&, const float*, double*, long, long, hipStream_t);

}  // namespace mif
