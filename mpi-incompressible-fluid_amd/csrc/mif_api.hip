// mif_api.hip -- context, fields and the C ABI entry points (include/mif_gpu.h).
//
// The solver sequence of every entry point follows the reference function it
// replaces; the file:line citations are on the declarations in mif_gpu.h and
// on the steps below.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <unordered_set>
#include <vector>

#include "../../include/mif_gpu.h"
#include "mif_kernels.h"

using namespace mif;

// Device scratch of mif_field_gather_noalloc: the point indices (three int32
// per point) at offset 0, the gathered FP64 values at kGatherNoallocValues,
// for up to kGatherNoallocMaxPoints points.  Every GPU the library runs on
// holds its own copy until the process ends.  Host code reaches a device's
// copy only through the address hipGetSymbolAddress returns with that device
// current (kept per device in g_scratch_addr and in each context); the
// array's name is a host-side stand-in that must never be passed to a kernel
// or a copy.
constexpr int64_t kGatherNoallocMaxPoints = int64_t(1) << 20;
constexpr size_t kGatherNoallocValues = size_t(12) << 20;  // 12 bytes per point, 256-byte aligned
constexpr size_t kScratchBytes = kGatherNoallocValues + (size_t(8) << 20);
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

thread_local std::string g_create_error;

constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_addr[kMaxScratchDevices] = {};

// Device address of `scratch` on `device`, which must be current; looked up
// once per device.
hipError_t scratch_address(int device, void** out) {
  *out = nullptr;
  if (device < 0 || device >= kMaxScratchDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_addr[device]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_addr[device] = ptr;
  }
  *out = g_scratch_addr[device];
  return hipSuccess;
}

struct TimerRec {
  int cls;
  hipEvent_t a, b;
};

}  // namespace

struct mif_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  mif_grid desc{};
  DGrid g{};
  Layout L{};
  // field storage starts `soff` elements into its allocation, so that point
  // i = 1 of a row (the first interior point, where the stage kernels' tiles
  // start) sits on a 128-byte boundary
  long soff = 0;
  int precision = 0;
  size_t esize = 8;
  int sweeps = 1;
  // mixed form: transforms of the pressure solve in FP32 or FP64
  // (mif_solve_precision), and the condition number of the discrete
  // Laplacian that the automatic choice reads
  int solve_mode = MIF_SOLVE_AUTO;
  double kappa = 0;
  // spectral solve
  FftPlan plan[3]{};
  double* W64[3]{};
  float* W32[3]{};
  double* H64[3]{};  // half-angle tables of cell-centred axes
  float* H32[3]{};
  double* lam[3]{};
  double inv_norm = 1;
  long bp0 = 0, bplane = 0, belems = 0;
  void* B = nullptr;       // packed working buffer (field precision)
  void* B2 = nullptr;      // mixed form: second FP32 packed buffer (the correction of the one-sweep path)
  Red3* xpart = nullptr;   // mixed form: per-workgroup sums of the fused x passes
  int nxpart = 0;
  Red3* red_b = nullptr;   // mixed form: (W.b, W.1) of the right-hand side
  Red3* red_phi = nullptr; // mixed form: (W.phi, W.1) of the stored increment
  // the increment the last solve stored still holds its weighted mean
  // (red_phi); the next projection subtracts it from the pressure update
  const Red3* phi_mean = nullptr;
  // timestep: the projections of stages 1 and 2 are fused into the next
  // stage kernel unless the caller asked for the stage records (the
  // projected Y2, Y3 in the step's buffers, which mif_timestep_adjoint reads);
  // pswap is the second pressure buffer the fused stages write into
  bool stage_records = false;
  int defect_schedule = MIF_DEFECT_EVERY_STAGE;
  void* pswap = nullptr;
  double* B64 = nullptr;   // mixed form: FP64 right-hand side
  double* P64 = nullptr;   // mixed form: FP64 accumulated solution
  double* R64 = nullptr;   // stand-alone Poisson runs: stored FP64 right-hand side
  double force[3]{};       // body force of timestep() (zero: the reference)
  // adjoint workspace (allocated on first use): two z_bar triples, N_bar,
  // x_bar, per-block partial sums and the force-adjoint accumulator
  void* adj[4][3]{};
  double* adj_partial = nullptr;
  double* adj_fbar = nullptr;
  bool adj_ready = false;  // every adjoint buffer allocated
  // velocity-only step: sin/cos of the coordinates of every axis,
  // unshifted and staggered (launch_trig_tables), filled on first use
  double* trig = nullptr;
  int trig_n = 0;
  // transfers and reductions
  double* staging = nullptr;
  long staging_elems = 0;
  Red3* partial = nullptr;
  Red3* red_out = nullptr;
  int nred = 1024;
  std::string err;
  // this device's copy of the no-allocation scratch array (nullptr if its
  // address could not be obtained; scratch_err then says why)
  void* scratch = nullptr;
  std::string scratch_err;
  // host-evaluated boundary data: three velocity face tables, then three
  // Neumann tables (allocated on first use)
  double* ftab = nullptr;
  long ftab_elems = 0;
  // kernel timing
  bool timing = false;
  std::vector<TimerRec> pending;
  std::vector<hipEvent_t> pool;
  double total_ms[MIF_KC_COUNT]{};
  uint64_t launches[MIF_KC_COUNT]{};
  // live fields: destroying the context frees their storage and orphans them
  std::unordered_set<mif_field*> fields;
};

struct mif_field {
  mif_ctx* ctx;  // nullptr once the owning context has been destroyed
  int stag;
  void* data;
};

namespace {

int fail(mif_ctx* c, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  if (c)
    c->err = buf;
  else
    g_create_error = buf;
  return code;
}

#define HIPCHK(c, expr)                                                                          \
  do {                                                                                           \
    hipError_t e_ = (expr);                                                                      \
    if (e_ != hipSuccess)                                                                        \
      return fail((c), e_ == hipErrorOutOfMemory ? MIF_ERR_OUT_OF_MEMORY : MIF_ERR_HIP, "%s: %s", \
                  #expr, hipGetErrorString(e_));                                                 \
  } while (0)

#define LAUNCHCHK(c)                                                                                        \
  do {                                                                                                      \
    hipError_t e_ = hipGetLastError();                                                                      \
    if (e_ != hipSuccess) return fail((c), MIF_ERR_HIP, "kernel launch failed: %s", hipGetErrorString(e_)); \
  } while (0)

// Event pair around the launches of one kernel class.
struct KTimer {
  mif_ctx* c;
  int cls;
  hipEvent_t a = nullptr;
  KTimer(mif_ctx* c_, int cls_) : c(c_), cls(cls_) {
    if (!c->timing) return;
    a = get_event();
    (void)hipEventRecord(a, c->stream);
  }
  ~KTimer() {
    if (!c->timing || !a) return;
    hipEvent_t b = get_event();
    (void)hipEventRecord(b, c->stream);
    c->pending.push_back({cls, a, b});
  }
  hipEvent_t get_event() {
    if (!c->pool.empty()) {
      hipEvent_t e = c->pool.back();
      c->pool.pop_back();
      return e;
    }
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    return e;
  }
};

void drain_timers(mif_ctx* c) {
  if (c->pending.empty()) return;
  (void)hipStreamSynchronize(c->stream);
  for (auto& r : c->pending) {
    float ms = 0;
    (void)hipEventElapsedTime(&ms, r.a, r.b);
    c->total_ms[r.cls] += ms;
    c->launches[r.cls] += 1;
    c->pool.push_back(r.a);
    c->pool.push_back(r.b);
  }
  c->pending.clear();
}

// Factor a transform length into Stockham radices (8, 4, 2 first, then odd
// primes).
void factor_radices(int L, FftPlan& P) {
  P.nst = 0;
  int n = L;
  while (n % 8 == 0) {
    P.radix[P.nst++] = 8;
    n /= 8;
  }
  while (n % 4 == 0) {
    P.radix[P.nst++] = 4;
    n /= 4;
  }
  while (n % 2 == 0) {
    P.radix[P.nst++] = 2;
    n /= 2;
  }
  for (int p = 3; n > 1; p += 2)
    while (n % p == 0) {
      P.radix[P.nst++] = p;
      n /= p;
    }
  int Ns = 1;
  for (int s = 0; s < P.nst; s++) {
    P.Ns[s] = Ns;
    Ns *= P.radix[s];
  }
}

// compute_eigenvalue_neumann / _periodic (src/PressureSolverStructures.cpp:5-11)
double eig_neumann(size_t index, double delta, size_t N) {
  return 2.0 * (std::cos(M_PI * index / (N - 1)) - 1.0) / (delta * delta);
}
double eig_periodic(size_t index, double delta, size_t N) {
  return 2.0 * (std::cos(2.0 * M_PI * index / N) - 1.0) / (delta * delta);
}
// Cell-centred Neumann axis of N cells (no reference counterpart): the
// eigenvalue of cos(pi k (j + 1/2) / N).
double eig_cell_neumann(size_t index, double delta, size_t N) {
  return 2.0 * (std::cos(M_PI * index / N) - 1.0) / (delta * delta);
}

template <typename T>
T* fptr(mif_field* f) {
  return (T*)f->data;
}

bool is_fp64(const mif_ctx* c) { return c->precision == MIF_PRECISION_FP64; }

// A field's storage pointer is soff elements into its allocation.
// Bytes allocated past a field's last element: the stage kernel's 16-byte
// row loads that start inside the last row may run up to 12 bytes past it.
constexpr long kFieldTailBytes = 256;

hipError_t alloc_field_storage(const mif_ctx* c, void** data) {
  void* raw = nullptr;
  const hipError_t e = hipMalloc(&raw, (c->L.elems + c->soff) * c->esize + kFieldTailBytes);
  *data = e == hipSuccess ? (char*)raw + c->soff * c->esize : nullptr;
  return e;
}
void free_field_storage(const mif_ctx* c, void* data) {
  if (data) (void)hipFree((char*)data - c->soff * c->esize);
}

int check_vel(mif_field* const v[3], mif_ctx** out) {
  if (!v || !v[0] || !v[1] || !v[2]) return fail(nullptr, MIF_ERR_INVALID_ARGUMENT, "null velocity field");
  mif_ctx* c = v[0]->ctx;
  if (!c) return fail(nullptr, MIF_ERR_INVALID_ARGUMENT, "field used after its context was destroyed");
  for (int q = 0; q < 3; q++) {
    if (v[q]->ctx != c) return fail(c, MIF_ERR_INVALID_ARGUMENT, "velocity components belong to different contexts");
    if (v[q]->stag != q) return fail(c, MIF_ERR_INVALID_ARGUMENT, "velocity component %d has staggering %d", q, v[q]->stag);
  }
  *out = c;
  return MIF_OK;
}

// ---------------------------------------------------------------------------
// The spectral solve on the packed buffer (src/PressureEquation.cpp:65-264):
// forward x, forward y, fused z (forward, scale, inverse), inverse y,
// inverse x.
// ---------------------------------------------------------------------------
int reduce_blocks(long total) {
  long b = (total + 255) / 256;
  if (b < 1) b = 1;
  if (b > 1024) b = 1024;
  return (int)b;
}

template <typename T>
SpectralScale scale_of(const mif_ctx* c) {
  SpectralScale sc;
  sc.lam[0] = c->lam[0];
  sc.lam[1] = c->lam[1];
  sc.lam[2] = c->lam[2];
  sc.inv_norm = c->inv_norm;
  return sc;
}

// Forward y, fused z (forward, scale, inverse), inverse y.
template <typename T>
void spectral_middle(mif_ctx* c, T* B) {
  const SpectralScale sc = scale_of<T>(c);
  {
    KTimer t(c, MIF_KC_FFT_Y);
    launch_fft_axis<T>(c->g, 1, -1, 0, c->plan[1], sc, B, c->bp0, c->bplane, c->stream);
  }
  {
    KTimer t(c, MIF_KC_FFT_Z);
    launch_fft_axis<T>(c->g, 2, 0, 1, c->plan[2], sc, B, c->bp0, c->bplane, c->stream);
  }
  {
    KTimer t(c, MIF_KC_FFT_Y);
    launch_fft_axis<T>(c->g, 1, +1, 0, c->plan[1], sc, B, c->bp0, c->bplane, c->stream);
  }
}

// ---------------------------------------------------------------------------
// The spectral solve on the packed buffer (src/PressureEquation.cpp:65-264):
// forward x, forward y, fused z (forward, scale, inverse), inverse y,
// inverse x.
// ---------------------------------------------------------------------------
template <typename T>
int spectral_solve(mif_ctx* c, T* B) {
  const SpectralScale sc = scale_of<T>(c);
  {
    KTimer t(c, MIF_KC_FFT_X);
    launch_fft_axis<T>(c->g, 0, -1, 0, c->plan[0], sc, B, c->bp0, c->bplane, c->stream);
  }
  spectral_middle<T>(c, B);
  {
    KTimer t(c, MIF_KC_FFT_X);
    launch_fft_axis<T>(c->g, 0, +1, 0, c->plan[0], sc, B, c->bp0, c->bplane, c->stream);
  }
  LAUNCHCHK(c);
  return MIF_OK;
}

// Right-hand side from the velocity, solve, and unpack into the pressure
// field (solve_pressure_equation_* , src/PressureEquation.cpp:266-286).  On
// x lengths with a specialised transform the divergence is computed inside the x forward
// pass and the x inverse pass writes the pressure field directly.
template <typename T>
int pressure_solve(mif_ctx* c, mif_field* p, mif_field* const vel[3], double dt, const FnDesc& gp, double tg,
                   int nhn) {
  T* B = (T*)c->B;
  const SpectralScale sc = scale_of<T>(c);
  if (!nhn && fft_fast_len(c->plan[0].L)) {
    FusedIO io{vel[0]->data, vel[1]->data, vel[2]->data, p->data, c->L, dt};
    {
      KTimer t(c, MIF_KC_FFT_X_DIV);
      launch_fft_pow2<T>(c->g, 0, 0, c->plan[0], sc, B, c->bp0, c->bplane, &io, c->stream);
    }
    spectral_middle<T>(c, B);
    {
      KTimer t(c, MIF_KC_FFT_X_UNPACK);
      launch_fft_pow2<T>(c->g, 0, 1, c->plan[0], sc, B, c->bp0, c->bplane, &io, c->stream);
    }
    LAUNCHCHK(c);
    return MIF_OK;
  }
  {
    KTimer t(c, MIF_KC_DIV);
    launch_div_rhs<T, T>(c->g, c->L, fptr<T>(vel[0]), fptr<T>(vel[1]), fptr<T>(vel[2]), B, c->bp0, c->bplane, dt, gp,
                         tg, nhn, c->stream);
  }
  int rc = spectral_solve<T>(c, B);
  if (rc) return rc;
  {
    KTimer t(c, MIF_KC_UNPACK);
    launch_unpack<T, T>(c->g, c->L, B, c->bp0, c->bplane, fptr<T>(p), c->stream);
  }
  LAUNCHCHK(c);
  return MIF_OK;
}

// Mixed form with an FP64 right-hand side in c->B64 (owner points, packed):
// project it onto the range of the Laplacian, solve in FP32, then `sweeps`
// FP64 defect corrections; the result is left in c->P64 (sweeps > 0) or in
// the FP32 buffer c->B (sweeps == 0).  Returns which one holds it.
enum SolutionIn { IN_B32, IN_P64 };

// FP32 transforms amplify their rounding by the condition number of the
// discrete Laplacian: an FP32 solve returns phi with an error of order
// C eps32 kappa relative to the right-hand side's scale.  Both the
// incremental projection (which hands each stage's pressure error to the
// next solve) and the defect correction contract only while that factor
// stays below one.  The 512^3 channel (kappa ~4.0e5) is stable with FP32
// transforms and the 768^3 channel (kappa ~8.9e5) is not, so the automatic
// choice moves the transforms to FP64 above kappa = 6e5.
constexpr double kMaxKappaFp32 = 6.0e5;

bool mixed_solve_fp64(const mif_ctx* c) {
  if (c->solve_mode == MIF_SOLVE_FP64) return true;
  if (c->solve_mode == MIF_SOLVE_FP32) return false;
  return c->kappa > kMaxKappaFp32;
}

int mixed_solve_from_b64(mif_ctx* c, SolutionIn* where) {
  float* B32 = (float*)c->B;
  const long total = (long)c->g.npts[0] * c->g.npts[1] * c->g.npts[2];
  const int nb = reduce_blocks(total);
  {
    KTimer t(c, MIF_KC_DEFECT);
    launch_project_rhs(c->g, c->B64, B32, c->bp0, c->bplane, c->partial, nb, c->red_out, c->stream);
  }
  int rc = spectral_solve<float>(c, B32);
  if (rc) return rc;
  if (c->sweeps == 0) {
    *where = IN_B32;
    return MIF_OK;
  }
  {
    KTimer t(c, MIF_KC_DEFECT);
    launch_to64(c->g, B32, c->P64, c->bp0, c->bplane, c->stream);
  }
  for (int s = 0; s < c->sweeps; s++) {
    {
      KTimer t(c, MIF_KC_DEFECT);
      launch_residual(c->g, c->B64, c->P64, B32, c->bp0, c->bplane, c->stream);
    }
    if ((rc = spectral_solve<float>(c, B32))) return rc;
    {
      KTimer t(c, MIF_KC_DEFECT);
      launch_accumulate(c->g, c->P64, B32, c->bp0, c->bplane, c->stream);
    }
  }
  {
    KTimer t(c, MIF_KC_DEFECT);
    launch_remove_mean(c->g, c->P64, c->bp0, c->bplane, c->partial, nb, c->red_out, c->stream);
  }
  LAUNCHCHK(c);
  *where = IN_P64;
  return MIF_OK;
}

// Pressure solve of the mixed form: FP64 divergence (plus Neumann data),
// then mixed_solve_from_b64, then unpack into the FP32 pressure field.
int pressure_solve_mixed(mif_ctx* c, mif_field* p, mif_field* const vel[3], double dt, const FnDesc& gp, double tg,
                         int nhn) {
  {
    KTimer t(c, MIF_KC_DIV);
    launch_div_rhs<float, double>(c->g, c->L, fptr<float>(vel[0]), fptr<float>(vel[1]), fptr<float>(vel[2]), c->B64,
                                  c->bp0, c->bplane, dt, gp, tg, nhn, c->stream);
  }
  SolutionIn where;
  int rc = mixed_solve_from_b64(c, &where);
  if (rc) return rc;
  {
    KTimer t(c, MIF_KC_UNPACK);
    if (where == IN_P64)
      launch_unpack<float, double>(c->g, c->L, c->P64, c->bp0, c->bplane, fptr<float>(p), c->stream);
    else
      launch_unpack<float, float>(c->g, c->L, (float*)c->B, c->bp0, c->bplane, fptr<float>(p), c->stream);
  }
  LAUNCHCHK(c);
  return MIF_OK;
}

// Mixed form with FP64 transforms: FP64 divergence of the FP32 velocity,
// the FP64 spectral solve on the FP64 packed buffer, unpack into the FP32
// increment.  On x lengths with a specialised transform the divergence and
// the unpack are fused into the FP64 x passes, which read and write the
// FP32 fields.
int pressure_solve_mixed64(mif_ctx* c, mif_field* p, mif_field* const vel[3], double dt, const FnDesc& gp, double tg,
                           int nhn) {
  if (!nhn && fft_fast_len(c->plan[0].L)) {
    FusedIO io{vel[0]->data, vel[1]->data, vel[2]->data, p->data, c->L, dt};
    io.fields_fp32 = 1;
    const SpectralScale sc = scale_of<double>(c);
    {
      KTimer t(c, MIF_KC_FFT_X_DIV);
      launch_fft_pow2<double>(c->g, 0, 0, c->plan[0], sc, c->B64, c->bp0, c->bplane, &io, c->stream);
    }
    spectral_middle<double>(c, c->B64);
    {
      KTimer t(c, MIF_KC_FFT_X_UNPACK);
      launch_fft_pow2<double>(c->g, 0, 1, c->plan[0], sc, c->B64, c->bp0, c->bplane, &io, c->stream);
    }
    LAUNCHCHK(c);
    return MIF_OK;
  }
  {
    KTimer t(c, MIF_KC_DIV);
    launch_div_rhs<float, double>(c->g, c->L, fptr<float>(vel[0]), fptr<float>(vel[1]), fptr<float>(vel[2]), c->B64,
                                  c->bp0, c->bplane, dt, gp, tg, nhn, c->stream);
  }
  int rc = spectral_solve<double>(c, c->B64);
  if (rc) return rc;
  {
    KTimer t(c, MIF_KC_UNPACK);
    launch_unpack<float, double>(c->g, c->L, c->B64, c->bp0, c->bplane, fptr<float>(p), c->stream);
  }
  LAUNCHCHK(c);
  return MIF_OK;
}

// Mixed form with one FP64 defect sweep and FP32 transforms, fused into the
// transform passes (homogeneous data, an x length with a fused kernel):
//   x forward + divergence: b = div(u)/dt in FP64, stored (b64) with its
//     weighted sums (W.b, W.1), and its x transform in FP32;
//   y, z (scale), y, x inverse: phi_0 in FP32 (B);
//   x forward of the defect r = (b - mean_b) - L phi_0, formed in FP64 as
//     the pass loads its lines (FP32 phi_0 read exactly), into B2;
//   y, z, y transforms of r; x inverse + unpack: phi = phi_0 + delta in
//     FP64, stored as the FP32 increment with its periodic images, and its
//     weighted sums (W.phi, W.1).
// The increment's weighted mean (FP32 rounding of the zeroed constant mode)
// is left in the field: the projection subtracts it from the pressure update
// (c->phi_mean), stand-alone solves subtract it from the field.
int pressure_solve_mixed_dc1(mif_ctx* c, mif_field* p, mif_field* const vel[3], double dt) {
  float* B32 = (float*)c->B;
  float* R32 = (float*)c->B2;
  const SpectralScale sc = scale_of<float>(c);
  FusedIO io{vel[0]->data, vel[1]->data, vel[2]->data, p->data, c->L, dt};
  io.dc = 1;
  io.b64 = c->B64;
  io.partial = c->xpart;
  {
    KTimer t(c, MIF_KC_FFT_X_DIV);
    launch_fft_pow2<float>(c->g, 0, 0, c->plan[0], sc, B32, c->bp0, c->bplane, &io, c->stream);
  }
  {
    KTimer t(c, MIF_KC_DEFECT);
    launch_reduce_partials(c->xpart, c->nxpart, c->red_b, c->stream);
  }
  spectral_middle<float>(c, B32);
  {
    KTimer t(c, MIF_KC_FFT_X);
    launch_fft_axis<float>(c->g, 0, +1, 0, c->plan[0], sc, B32, c->bp0, c->bplane, c->stream);
  }
  {
    // the x forward pass forms the defect as it loads its lines
    KTimer t(c, MIF_KC_FFT_X);
    FusedIO rio{};
    rio.res = 1;
    rio.phi32 = B32;
    rio.b64 = c->B64;
    rio.bmean = c->red_b;
    launch_fft_pow2<float>(c->g, 0, 0, c->plan[0], sc, R32, c->bp0, c->bplane, &rio, c->stream);
  }
  spectral_middle<float>(c, R32);
  io.add32 = B32;
  {
    KTimer t(c, MIF_KC_FFT_X_UNPACK);
    launch_fft_pow2<float>(c->g, 0, 1, c->plan[0], sc, R32, c->bp0, c->bplane, &io, c->stream);
  }
  {
    KTimer t(c, MIF_KC_DEFECT);
    launch_reduce_partials(c->xpart, c->nxpart, c->red_phi, c->stream);
  }
  LAUNCHCHK(c);
  c->phi_mean = c->red_phi;
  return MIF_OK;
}

bool dc1_fused(const mif_ctx* c, int nhn) {
  return !is_fp64(c) && !mixed_solve_fp64(c) && c->sweeps == 1 && !nhn && c->nxpart > 0 &&
         fft_fast_len(c->plan[0].L);
}

int pressure_solve_any(mif_ctx* c, mif_field* p, mif_field* const vel[3], double dt, const FnDesc& gp, double tg,
                       int nhn) {
  c->phi_mean = nullptr;
  if (is_fp64(c)) return pressure_solve<double>(c, p, vel, dt, gp, tg, nhn);
  if (mixed_solve_fp64(c)) return pressure_solve_mixed64(c, p, vel, dt, gp, tg, nhn);
  if (dc1_fused(c, nhn)) return pressure_solve_mixed_dc1(c, p, vel, dt);
  if (c->sweeps > 0 || nhn) return pressure_solve_mixed(c, p, vel, dt, gp, tg, nhn);
  return pressure_solve<float>(c, p, vel, dt, gp, tg, nhn);
}

// A stand-alone solve hands the increment to the caller: its weighted mean,
// if the solve left one, is removed from the field here.
int finish_standalone_solve(mif_ctx* c, mif_field* p) {
  if (c->phi_mean) {
    launch_sub_mean_field(c->g, c->L, fptr<float>(p), c->phi_mean, c->stream);
    c->phi_mean = nullptr;
    LAUNCHCHK(c);
  }
  return MIF_OK;
}

int pressure_solve_any(mif_ctx* c, mif_field* p, mif_field* const vel[3], double dt, const FnDesc& gp, double tg,
                       int nhn);

// phi (not null): the previous stage's projection fused in, with time step
// pdts and the increment's mean c->phi_mean; the incremented pressure goes to
// pout, here and on the points the stage's tiles leave out: by the boundary
// pass that follows (rim, for apply_bc), or by a pass of its own (rim null).
template <typename T>
int stage(mif_ctx* c, int st, mif_field* const X[3], mif_field* const A[3], mif_field* const O1[3],
          mif_field* const O2[3], const void* p, double a1, double a2, double b, const void* phi = nullptr,
          double pdts = 0, void* pout = nullptr, PressureRim* rim = nullptr) {
  StageArgs s{};
  for (int q = 0; q < 3; q++) {
    s.X[q] = X[q]->data;
    s.A[q] = A ? A[q]->data : nullptr;
    s.O1[q] = O1[q]->data;
    s.O2[q] = O2 ? O2[q]->data : nullptr;
  }
  s.p = p;
  s.a1 = a1;
  s.a2 = a2;
  s.b = b;
  for (int q = 0; q < 3; q++) s.force[q] = c->force[q];
  s.phi = phi;
  s.pdts = pdts;
  s.pmean = phi ? c->phi_mean : nullptr;
  s.pout = pout;
  {
    KTimer t(c, MIF_KC_STAGE);
    launch_stage<T>(st, c->g, c->L, s, c->stream);
  }
  if (phi && rim) {
    *rim = PressureRim{p, phi, pout, c->phi_mean};
  } else if (phi) {
    KTimer t(c, MIF_KC_PRESSURE_RIM);
    launch_p_rim<T>(c->g, c->L, (const T*)p, (const T*)phi, (T*)pout, c->phi_mean, c->stream);
  }
  if (phi) c->phi_mean = nullptr;
  return MIF_OK;
}

// rim: the pressure rim of the fused stage kernel just launched, written in
// the same launch
template <typename T>
int apply_bc(mif_ctx* c, mif_field* const v[3], const FnDesc& f, double t, const PressureRim& rim = PressureRim{}) {
  void* F[3] = {v[0]->data, v[1]->data, v[2]->data};
  KTimer tm(c, MIF_KC_BC);
  launch_bc<T>(c->g, c->L, F, f, t, c->stream, rim);
  return MIF_OK;
}

// pout: where p + phi goes (default: p itself)
template <typename T>
int project(mif_ctx* c, mif_field* const v[3], void* p, mif_field* phi, double dts, void* pout = nullptr) {
  KTimer tm(c, MIF_KC_PROJECT);
  launch_project<T>(c->g, c->L, fptr<T>(v[0]), fptr<T>(v[1]), fptr<T>(v[2]), (T*)p, fptr<T>(phi), dts, c->stream,
                    c->phi_mean, (T*)pout);
  c->phi_mean = nullptr;
  return MIF_OK;
}
template <typename T>
int project(mif_ctx* c, mif_field* const v[3], mif_field* p, mif_field* phi, double dts) {
  return project<T>(c, v, p->data, phi, dts);
}

// Stage times of FULL_TIMESTEP (src/Timestep.cpp:99-104): the boundary data
// of stage s are taken at stage_time[s] and its Neumann data are
// f(prev_time[s]) - f(stage_time[s]).
struct StageTimes {
  double dt_s[4];
  double time[4];  // [0] = t_n
};

StageTimes stage_times(double dt, double t_n) {
  StageTimes st{};
  st.dt_s[1] = 64.0 / 120.0 * dt;
  st.dt_s[2] = 16.0 / 120.0 * dt;
  st.dt_s[3] = 40.0 / 120.0 * dt;
  st.time[0] = t_n;
  st.time[1] = t_n + st.dt_s[1];
  st.time[2] = st.time[1] + st.dt_s[2];
  st.time[3] = t_n + dt;
  return st;
}

// State carried from step to step by a multi-step call (fused form): the
// pressure buffer that holds p (0: the field, 1: the context's second
// pressure buffer) and, when pending_dts > 0, the previous step's stage-3
// projection (phi in pb, its mean in c->phi_mean), which the next step's
// stage-1 kernel applies.
struct StepCarry {
  int cur = 0;
  double pending_dts = 0;
};

// FULL_TIMESTEP (src/Timestep.cpp:97-141).  ev[s - 1] is the boundary data
// of stage s (a descriptor evaluated at the stage time, or a host table);
// gd, when not null, the Neumann data of the three stages' solves.  carry,
// last: a multi-step call (mif_timestep_n); every step but the last leaves
// its stage-3 projection to the next step's stage-1 kernel.
template <typename T>
int full_timestep(mif_ctx* c, mif_field* const vel[3], mif_field* const buf[3], mif_field* const buf2[3],
                  const FnDesc ev[3], const FnDesc* gd, double t_n, mif_field* p, mif_field* pb,
                  StepCarry* carry = nullptr, bool last = true) {
  const double dt = c->g.dt;
  const StageTimes st = stage_times(dt, t_n);
  const int nhn = gd != nullptr;
  const FnDesc none{};
  int rc;
  // Fused form: the projections of stages 1 and 2 (velocity correction and
  // pressure increment) are applied by the stage kernels of stages 2 and 3
  // as they read Y2, Y3 and p (one pass over the velocity and pressure fewer
  // per fused projection).  Each fused stage reads p from one of two buffers
  // (p and the context's second pressure buffer) and writes the incremented
  // pressure into the other; the last projection of the call writes p.
  const bool fuse = !c->stage_records;
  if (fuse && !c->pswap) HIPCHK(c, alloc_field_storage(c, &c->pswap));
  StepCarry single;
  StepCarry& cs = carry ? *carry : single;
  void* const pbuf[2] = {p->data, c->pswap};
  // the pressure buffer a fused stage reads, then (after it) the one it wrote
  auto pin = [&]() { return pbuf[cs.cur]; };
  auto pnext = [&]() { return pbuf[cs.cur ^ 1]; };
  // Pressure solve of stage k; with MIF_DEFECT_LAST_STAGE the mixed form's
  // defect sweeps run in the solve of stage 3 only.
  auto solve = [&](int k, mif_field* const* u) {
    const int sweeps = c->sweeps;
    if (k < 3 && c->defect_schedule == MIF_DEFECT_LAST_STAGE) c->sweeps = 0;
    const int r = pressure_solve_any(c, pb, u, st.dt_s[k], nhn ? gd[k - 1] : none, 0.0, nhn);
    c->sweeps = sweeps;
    return r;
  };
  // the pressure rim of each fused stage kernel is written by the boundary
  // pass that follows it (one launch)
  PressureRim rim;
  // Stage 1 (src/Timestep.cpp:107-117), with the previous step's last
  // projection when a multi-step call left it pending
  {
    const double a1 = 64.0 / 120.0 * dt;
    if (fuse && cs.pending_dts > 0) {
      rc = stage<T>(c, 1, vel, nullptr, buf, buf2, pin(), a1, 0.0, a1, pb->data, cs.pending_dts, pnext(), &rim);
      cs.cur ^= 1;
      cs.pending_dts = 0;
    } else {
      rc = stage<T>(c, 1, vel, nullptr, buf, buf2, pin(), a1, 0.0, a1);
    }
    if (rc) return rc;
  }
  apply_bc<T>(c, buf, ev[0], st.time[1], rim);
  if ((rc = solve(1, buf))) return rc;
  if (!fuse) project<T>(c, buf, p, pb, st.dt_s[1]);
  // Stage 2 (src/Timestep.cpp:119-129)
  {
    const double a1 = -34.0 / 120.0 * dt, a2 = 50.0 / 120.0 * dt;
    if (fuse) {
      rc = stage<T>(c, 2, buf, buf2, buf2, vel, pin(), a1, a2, a1 + a2, pb->data, st.dt_s[1], pnext(), &rim);
      cs.cur ^= 1;
    } else {
      rc = stage<T>(c, 2, buf, buf2, buf2, vel, pin(), a1, a2, a1 + a2);
    }
    if (rc) return rc;
  }
  apply_bc<T>(c, buf2, ev[1], st.time[2], rim);
  if ((rc = solve(2, buf2))) return rc;
  if (!fuse) project<T>(c, buf2, p, pb, st.dt_s[2]);
  // Stage 3 (src/Timestep.cpp:131-140)
  {
    const double a2 = -50.0 / 120.0 * dt, a3 = 90.0 / 120.0 * dt;
    if (fuse) {
      rc = stage<T>(c, 3, buf2, vel, vel, nullptr, pin(), 0.0, a3, a2 + a3, pb->data, st.dt_s[2], pnext(), &rim);
      cs.cur ^= 1;
    } else {
      rc = stage<T>(c, 3, buf2, vel, vel, nullptr, pin(), 0.0, a3, a2 + a3);
    }
    if (rc) return rc;
  }
  apply_bc<T>(c, vel, ev[2], st.time[3], rim);
  if ((rc = solve(3, vel))) return rc;
  if (fuse && !last) {
    cs.pending_dts = st.dt_s[3];
  } else {
    // the step's last projection, its pressure written into p
    project<T>(c, vel, pin(), pb, st.dt_s[3], p->data);
    cs.cur = 0;
  }
  LAUNCHCHK(c);
  return MIF_OK;
}

// timestep_velocity (src/TimestepVelocity.cpp:60-90).
template <typename T>
int velocity_timestep(mif_ctx* c, mif_field* const vel[3], mif_field* const buf[3], mif_field* const rhs[3],
                      const FnDesc ev[3], const FnDesc& fo, double t_n) {
  const double dt = c->g.dt;
  const double time_1 = t_n + (8.0 / 15.0) * dt;
  const double time_2 = t_n + (2.0 / 3.0) * dt;
  const double final_time = t_n + dt;
  VelStageArgs s{};
  s.dt = dt;
  s.forcing = fo;
  if (fo.kind == MIF_FN_VT_FORCING) {
    if (!c->trig) {
      const int tn = std::max(c->g.N[0], std::max(c->g.N[1], c->g.N[2])) + 1;
      HIPCHK(c, hipMalloc(&c->trig, 12 * (size_t)tn * sizeof(double)));
      c->trig_n = tn;
      launch_trig_tables(c->g, c->trig, tn, c->stream);
    }
    s.trig = c->trig;
    s.tn = c->trig_n;
  }
  for (int q = 0; q < 3; q++) {
    s.V[q] = vel[q]->data;
    s.R[q] = rhs[q]->data;
    s.B[q] = buf[q]->data;
  }
  for (int q = 0; q < 3; q++) s.X[q] = vel[q]->data;
  s.t = t_n;
  {
    KTimer t(c, MIF_KC_STAGE);
    launch_vel_stage<T>(1, c->g, c->L, s, c->stream);
  }
  apply_bc<T>(c, buf, ev[0], time_1);
  for (int q = 0; q < 3; q++) s.X[q] = buf[q]->data;
  s.t = time_1;
  {
    KTimer t(c, MIF_KC_STAGE);
    launch_vel_stage<T>(2, c->g, c->L, s, c->stream);
  }
  apply_bc<T>(c, vel, ev[1], time_2);
  for (int q = 0; q < 3; q++) s.X[q] = vel[q]->data;
  s.t = time_2;
  {
    KTimer t(c, MIF_KC_STAGE);
    launch_vel_stage<T>(3, c->g, c->L, s, c->stream);
  }
  apply_bc<T>(c, buf, ev[2], final_time);
  // velocity.swap_data(velocity_buffer) (src/TimestepVelocity.cpp:89)
  for (int q = 0; q < 3; q++) std::swap(vel[q]->data, buf[q]->data);
  LAUNCHCHK(c);
  return MIF_OK;
}


// ---------------------------------------------------------------------------
// Adjoint of one step (no reference counterpart).  The reverse sweep of
// FULL_TIMESTEP (src/Timestep.cpp:97-141): per stage, in reverse order,
//   phib = pbar + dt_s G^T ubar                  (projection, p += phi)
//   ubar += D^T W M W^-1 U^T phib / dt_s         (Poisson solve transposed)
//   Y    = B^T ubar                              (boundary map transposed)
//   Nb, Sb/R1b from Y; Xb = Y + J_N(X)^T Nb      (stage update transposed)
//   pbar += b_s G^T Y;  fbar += sum Nb
// ---------------------------------------------------------------------------
// Allocates the adjoint workspace on first use.  Every buffer is allocated
// into a local first; the context takes them only once all succeeded, so a
// failed call leaves the context without a workspace (and frees what it got)
// and the next call starts over.
int adj_workspace(mif_ctx* c) {
  if (c->adj_ready) return MIF_OK;
  const size_t bytes = c->L.elems * c->esize;
  void* adj[4][3] = {};
  double* partial = nullptr;
  double* fbar = nullptr;
  hipError_t e = hipSuccess;
  // z_bar of the current and of a later stage, N_bar, x_bar: zero outside
  // the component interiors is an invariant the kernels rely on
  for (int q = 0; q < 4 && e == hipSuccess; q++)
    for (int cc = 0; cc < 3 && e == hipSuccess; cc++) {
      e = alloc_field_storage(c, &adj[q][cc]);
      if (e == hipSuccess) e = hipMemsetAsync(adj[q][cc], 0, bytes, c->stream);
    }
  if (e == hipSuccess) e = hipMalloc(&partial, 3 * kAdjMaxBlocks * sizeof(double));
  // [0..2]: the force adjoint of the last step; [3..5]: the accumulator of
  // the steps called without a host force_bar (mif_adjoint_force_take)
  if (e == hipSuccess) e = hipMalloc(&fbar, 6 * sizeof(double));
  if (e == hipSuccess) e = hipMemsetAsync(fbar, 0, 6 * sizeof(double), c->stream);
  if (e != hipSuccess) {
    (void)hipStreamSynchronize(c->stream);
    for (auto& row : adj)
      for (void* p : row)
        if (p) free_field_storage(c, p);
    (void)hipFree(partial);
    (void)hipFree(fbar);
    return fail(c, e == hipErrorOutOfMemory ? MIF_ERR_OUT_OF_MEMORY : MIF_ERR_HIP, "adjoint workspace: %s",
                hipGetErrorString(e));
  }
  std::memcpy(c->adj, adj, sizeof(adj));
  c->adj_partial = partial;
  c->adj_fbar = fbar;
  c->adj_ready = true;
  return MIF_OK;
}

AdjVec vec_of(mif_field* const f[3]) { return AdjVec{{f[0]->data, f[1]->data, f[2]->data}}; }
AdjVec vec_ws(mif_ctx* c, int q) { return AdjVec{{c->adj[q][0], c->adj[q][1], c->adj[q][2]}}; }

// Adjoint of full_timestep, stages 3, 2, 1.  Per stage s with forward
// source X_s, step dts and pressure-gradient coefficient b_s:
//   B     = W^-1 U^T (p_bar - dts G^T y_bar)        (k_adj_rhs)
//   psi   = M B                                      (the forward solve)
//   z_bar = B^T (y_bar + D^T W psi / dts)            (k_adj_zb)
//   N_bar = stage 3: a3 z3;  stage 2: a2 (z2 - z3);  stage 1: a1 z1 + c21 z2
//   x_bar = z_bar + J_N(X_s)^T N_bar, p_bar += b_s G^T z_bar   (k_adj_xb)
// where y_bar is u_bar (stage 3) or the previous x_bar.  The force adjoint
// is sum(N_bar) over the stages = (a3 - a2) S3 + (a2 + c21) S2 + a1 S1 with
// S_s = sum(z_bar_s): k_adj_zb accumulates the weighted block sums of the
// three stages and one reduction per step adds them up.  force_bar == NULL:
// the result is added to the context's device accumulator, with no
// synchronisation.
template <typename T>
int step_adjoint(mif_ctx* c, mif_field* const un[3], mif_field* const y2[3], mif_field* const y3[3],
                 mif_field* const ubar[3], mif_field* pbar, double fbar[3]) {
  int rc = adj_workspace(c);
  if (rc) return rc;
  const DGrid& g = c->g;
  const Layout& L = c->L;
  hipStream_t st = c->stream;
  const double dt = g.dt;
  const double dts[4] = {0, 64.0 / 120.0 * dt, 16.0 / 120.0 * dt, 40.0 / 120.0 * dt};
  const double s1_a1 = 64.0 / 120.0 * dt, s1_b = s1_a1;
  const double s2_a1 = -34.0 / 120.0 * dt, s2_a2 = 50.0 / 120.0 * dt, s2_b = s2_a1 + s2_a2;
  const double s3_a2 = -50.0 / 120.0 * dt, s3_a3 = 90.0 / 120.0 * dt, s3_b = s3_a2 + s3_a3;
  const double bcoef[4] = {0, s1_b, s2_b, s3_b};
  // N_bar = alpha z_s + beta z_later, and the force-adjoint weight of S_s
  const double alpha[4] = {0, s1_a1, s2_a2, s3_a3};
  const double beta[4] = {0, s2_a1, -s2_a2, 0};
  const double fcoef[4] = {0, s1_a1, s2_a2 + s2_a1, s3_a3 - s2_a2};
  const AdjVec ZA = vec_ws(c, 0), ZB = vec_ws(c, 1), NB = vec_ws(c, 2), XT = vec_ws(c, 3);
  const AdjVec UB = vec_of(ubar);
  const AdjVec src[4] = {AdjVec{}, vec_of(un), vec_of(y2), vec_of(y3)};
  // z_bar of each stage: stage 3 -> ZA, stage 2 -> ZB, stage 1 -> ZA (z3 is
  // no longer needed); the later stage's z_bar feeding N_bar
  const AdjVec zcur[4] = {AdjVec{}, ZA, ZB, ZA};
  const AdjVec zlater[4] = {AdjVec{}, ZB, ZA, ZA};
  T* P = fptr<T>(pbar);
  T* B = (T*)c->B;
  const long interior_pts = (long)(g.N[0] + 1) * (g.N[1] + 1) * (g.N[2] + 1);
  const int nblk = (int)std::min<long>(kAdjMaxBlocks, (interior_pts + 255) / 256);
  KTimer tm(c, MIF_KC_ADJOINT);
  for (int s = 3; s >= 1; s--) {
    const AdjVec Yin = s == 3 ? UB : XT;
    launch_adj_rhs<T>(g, L, P, Yin, dts[s], B, c->bp0, c->bplane, st);
    if ((rc = spectral_solve<T>(c, B))) return rc;
    launch_adj_zb<T>(g, L, Yin, B, c->bp0, c->bplane, dts[s], zcur[s], NB, zlater[s], alpha[s], beta[s], fcoef[s],
                     s == 3 ? 0 : 1, c->adj_partial, nblk, st);
    launch_adj_xb<T>(g, L, src[s], NB, zcur[s], s == 1 ? UB : XT, P, bcoef[s], st);
  }
  if (!fbar) {
    launch_adj_sum(c->adj_partial, nblk, 1, c->adj_fbar + 3, st);
    LAUNCHCHK(c);
    return MIF_OK;
  }
  launch_adj_sum(c->adj_partial, nblk, 0, c->adj_fbar, st);
  LAUNCHCHK(c);
  double fb[3];
  HIPCHK(c, hipMemcpyAsync(fb, c->adj_fbar, 3 * sizeof(double), hipMemcpyDeviceToHost, st));
  HIPCHK(c, hipStreamSynchronize(st));
  for (int q = 0; q < 3; q++) fbar[q] += fb[q];
  return MIF_OK;
}

// ---------------------------------------------------------------------------
// Boundary data evaluated on the host (callbacks with no closed form).  A
// velocity face table holds, for u, v, w in turn, the faces z-, z+, y-, y+,
// x-, x+ of non-periodic axes, each a plane over the component's extents
// along the two other axes (lower axis fastest) with the value
// VelocityTensor::apply_bc writes there (src/VelocityTensor.cpp:36-233).  A
// Neumann table holds the six wall faces of the node extents with the
// normal component of the pressure-gradient data at each wall node
// (apply_non_homogeneous_neumann, src/PressureEquation.cpp:10-57).
// ---------------------------------------------------------------------------
long face_table_layout(const DGrid& g, bool velocity, long toff[3][6]) {
  long n = 0;
  for (int q = 0; q < 3; q++)
    for (int f = 0; f < 6; f++) toff[q][f] = -1;
  for (int q = 0; q < (velocity ? 3 : 1); q++) {
    int e[3];
    extents(g, velocity ? q : MIF_STAGGER_NONE, e);
    for (int f = 0; f < 6; f++) {
      const int axis = 2 - (f >> 1);
      if (g.periodic[axis]) continue;
      toff[q][f] = n;
      n += axis == 2 ? (long)e[0] * e[1] : (axis == 1 ? (long)e[0] * e[2] : (long)e[1] * e[2]);
    }
  }
  return n;
}

// Copies `count` host tables of `velocity` (or Neumann) layout into the
// context's table buffer (allocated on first use) and returns their
// descriptors in out[0..count).  The copy is ordered after the work already
// on the stream and complete at return, so the caller may reuse its arrays.
int upload_face_tables(mif_ctx* c, bool velocity, const double* host, int count, FnDesc* out) {
  long vtoff[3][6], ntoff[3][6];
  const long nv = face_table_layout(c->g, true, vtoff);
  const long nn = face_table_layout(c->g, false, ntoff);
  if (!c->ftab) {
    c->ftab_elems = 3 * (nv + nn);
    HIPCHK(c, hipMalloc(&c->ftab, std::max<long>(c->ftab_elems, 1) * sizeof(double)));
  }
  const long per = velocity ? nv : nn;
  double* dst = c->ftab + (velocity ? 0 : 3 * nv);
  if (per > 0) {
    HIPCHK(c, hipMemcpyAsync(dst, host, count * per * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
  }
  for (int q = 0; q < count; q++) {
    FnDesc d{};
    d.kind = kFnFaceTable;
    d.tab = dst + q * per;
    std::memcpy(d.toff, velocity ? vtoff : ntoff, sizeof(d.toff));
    out[q] = d;
  }
  return MIF_OK;
}

// Constants, layout, transform plans and scratch of a new context.
int ctx_init(mif_ctx* c, const mif_grid* grid, const uint64_t Ng[3], int precision, int device,
             const int cell[3]) {
  if (device < 0) (void)hipGetDevice(&device);
  c->device = device;
  if (hipSetDevice(device) != hipSuccess) return fail(nullptr, MIF_ERR_NO_DEVICE, "hipSetDevice(%d) failed", device);
  c->desc = *grid;
  c->precision = precision;
  c->esize = precision == MIF_PRECISION_FP64 ? 8 : 4;

  // Constants (src/Constants.cpp:58-120 at Py = Pz = 1)
  DGrid& g = c->g;
  const double size[3] = {grid->x_size, grid->y_size, grid->z_size};
  const double mn[3] = {grid->min_x, grid->min_y, grid->min_z};
  g.Re = grid->Re;
  g.dt = grid->final_time / grid->num_time_steps;
  for (int a = 0; a < 3; a++) {
    g.Ng[a] = (int)Ng[a];
    g.periodic[a] = grid->periodic[a] ? 1 : 0;
    g.N[a] = g.Ng[a] + g.periodic[a];
    g.base[a] = g.periodic[a] ? -1 : 0;
    g.cell[a] = cell && cell[a] ? 1 : 0;
    // cell-centred axis: the N_g - 1 cells between the nodes
    g.npts[a] = g.Ng[a] - (g.periodic[a] || g.cell[a] ? 1 : 0);
    g.olo[a] = g.periodic[a];
    const double h = size[a] / (double)(Ng[a] - 1);
    g.h[a] = h;
    g.min[a] = mn[a];
    g.one_over_2h[a] = 1 / (2 * h);
    g.one_over_8h[a] = 1 / (8 * h);
    g.one_over_h2Re[a] = 1 / (grid->Re * h * h);
    g.h_over_2[a] = h / 2;
    g.one_over_h[a] = 1 / h;
  }
  // common padded layout
  const int align = 64 / (int)c->esize;
  const int maxx = g.N[0] + 1, maxy = g.N[1] + 1, maxz = g.N[2] + 1;
  c->L.PX = (maxx + align - 1) / align * align;
  c->L.SY = maxy;
  c->L.SZ = maxz;
  c->L.plane = (long)c->L.PX * c->L.SY;
  c->L.elems = c->L.plane * c->L.SZ;
  // Storage shift: element 1 of every row on a 128-byte line, so that the
  // interior of a row starts a cache line.  Measured step times with and without the shift (FP32/FP64 fields of
  // 67/135 MB at 256^3: 1.728 -> 1.701 ms mixed, 3.303 -> 3.286 ms FP64;
  // 1.09 GB FP64 fields at 512^3: 24.80 -> 25.06 ms; 1.8 GB FP32 fields at
  // 768^3: 57.98 -> 58.85 ms): the shift pays while a field fits the 256 MB
  // Infinity Cache, so it is applied up to that size
  constexpr long kShiftMaxFieldBytes = 256L << 20;
  const bool cached = c->L.elems * (long)c->esize <= kShiftMaxFieldBytes;
  const bool shift = cached;
  c->soff = shift ? 128 / (long)c->esize - 1 : 0;
  c->L.lead = (int)c->soff;
  // Non-temporal stores (and loads of single-use operands) in the stage and
  // projection kernels, for fields of kNtMinFieldBytes up to the shift's
  // limit (the x inverse pass stores non-temporally at every size).
  // Measured steps, plain -> non-temporal: 256^3 FP64 3.287 -> 3.110 ms, mixed
  // 1.703 -> 1.632 ms (the stage kernel's 10-13 arrays exceed the Infinity
  // Cache, and streaming its outputs past it keeps the next kernel's lines
  // resident); 768^3 mixed 58.0 -> 59.3 ms (stage kernel +13%), 512^3 FP64
  // unchanged.  Below the lower limit the arrays a stage kernel streams fit
  // the cache together, so the next kernel finds its operands there:
  // non-temporal access took the 128^3 FP64 step (18 MB fields) from 0.52 to
  // 0.62 ms and the mixed one (9 MB) from 0.38 to 0.44 ms.  A 192^3 FP32
  // field (31 MB) measured the same either way.
  constexpr long kNtMinFieldBytes = 32L << 20;
  const bool streamed = c->L.elems * (long)c->esize > kNtMinFieldBytes;
  c->L.nt = cached && streamed ? 1 : 0;
  // kernels decompose flat indices in 32 bits
  if (c->L.elems >= (1L << 31))
    return fail(nullptr, MIF_ERR_UNSUPPORTED, "grid too large: a field would hold %ld padded elements (limit 2^31 - 1)",
                c->L.elems);

  HIPCHK(nullptr, hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));

  // pressure-solver structures (src/PressureSolverStructures.cpp:13-69)
  c->bp0 = (g.npts[0] + align - 1) / align * align;
  c->bplane = c->bp0 * g.npts[1];
  c->belems = c->bplane * g.npts[2];
  HIPCHK(nullptr, hipMalloc(&c->B, c->belems * c->esize));
  HIPCHK(nullptr, hipMemsetAsync(c->B, 0, c->belems * c->esize, c->stream));
  if (precision == MIF_PRECISION_MIXED) {
    HIPCHK(nullptr, hipMalloc(&c->B64, c->belems * sizeof(double)));
    HIPCHK(nullptr, hipMalloc(&c->P64, c->belems * sizeof(double)));
    HIPCHK(nullptr, hipMalloc(&c->B2, c->belems * sizeof(float)));
    HIPCHK(nullptr, hipMemsetAsync(c->B2, 0, c->belems * sizeof(float), c->stream));
    HIPCHK(nullptr, hipMalloc(&c->red_b, 2 * sizeof(Red3)));
    c->red_phi = c->red_b + 1;
  }
  double norm = 1.0;
  double lam_max = 0, lam_min = 0;  // |lambda| extremes of the 3-D operator
  for (int a = 0; a < 3; a++) {
    const int n = g.npts[a];
    FftPlan& P = c->plan[a];
    P.n = n;
    P.periodic = g.periodic[a];
    P.dct2 = g.cell[a];
    P.L = g.periodic[a] || g.cell[a] ? n : 2 * (n - 1);
    factor_radices(P.L, P);
    if (P.dct2) {
      // (cos, sin)(pi k / (2n)) of the DCT-II/III pre- and post-processing
      std::vector<double> h64(2 * (size_t)n);
      std::vector<float> h32(2 * (size_t)n);
      for (int k = 0; k < n; k++) {
        const long double ang = 3.141592653589793238462643383279502884L * (long double)k / (2.0L * n);
        h64[2 * k] = (double)cosl(ang);
        h64[2 * k + 1] = (double)sinl(ang);
        h32[2 * k] = (float)cosl(ang);
        h32[2 * k + 1] = (float)sinl(ang);
      }
      HIPCHK(nullptr, hipMalloc(&c->H64[a], h64.size() * sizeof(double)));
      HIPCHK(nullptr, hipMalloc(&c->H32[a], h32.size() * sizeof(float)));
      HIPCHK(nullptr, hipMemcpy(c->H64[a], h64.data(), h64.size() * sizeof(double), hipMemcpyHostToDevice));
      HIPCHK(nullptr, hipMemcpy(c->H32[a], h32.data(), h32.size() * sizeof(float), hipMemcpyHostToDevice));
      P.H64 = c->H64[a];
      P.H32 = c->H32[a];
    }
    std::vector<double> w64(2 * (size_t)P.L);
    std::vector<float> w32(2 * (size_t)P.L);
    for (int e = 0; e < P.L; e++) {
      const long double ang = 2.0L * 3.141592653589793238462643383279502884L * (long double)e / (long double)P.L;
      w64[2 * e] = (double)cosl(ang);
      w64[2 * e + 1] = (double)-sinl(ang);
      w32[2 * e] = (float)cosl(ang);
      w32[2 * e + 1] = (float)-sinl(ang);
    }
    HIPCHK(nullptr, hipMalloc(&c->W64[a], w64.size() * sizeof(double)));
    HIPCHK(nullptr, hipMalloc(&c->W32[a], w32.size() * sizeof(float)));
    HIPCHK(nullptr, hipMemcpy(c->W64[a], w64.data(), w64.size() * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(nullptr, hipMemcpy(c->W32[a], w32.data(), w32.size() * sizeof(float), hipMemcpyHostToDevice));
    P.W64 = c->W64[a];
    P.W32 = c->W32[a];
    std::vector<double> lam(n);
    for (int k = 0; k < n; k++)
      lam[k] = g.periodic[a] ? eig_periodic(k, g.h[a], n)
                             : (g.cell[a] ? eig_cell_neumann(k, g.h[a], n) : eig_neumann(k, g.h[a], n));
    HIPCHK(nullptr, hipMalloc(&c->lam[a], n * sizeof(double)));
    HIPCHK(nullptr, hipMemcpy(c->lam[a], lam.data(), n * sizeof(double), hipMemcpyHostToDevice));
    double amax = 0, amin = 0;
    for (int k = 0; k < n; k++) {
      const double v = std::fabs(lam[k]);
      amax = std::max(amax, v);
      if (v > 0 && (amin == 0 || v < amin)) amin = v;
    }
    lam_max += amax;
    if (amin > 0 && (lam_min == 0 || amin < lam_min)) lam_min = amin;
    // normalisation after each inverse axis (src/PressureEquation.cpp:167,200,234);
    // the DCT-III of a cell-centred axis returns n times the line
    norm *= (double)(Ng[a] - 1) * (g.periodic[a] || g.cell[a] ? 1.0 : 2.0);
  }
  c->inv_norm = 1.0 / norm;
  c->kappa = lam_min > 0 ? lam_max / lam_min : 0;
  if (precision == MIF_PRECISION_MIXED) {
    c->nxpart = fused_x_blocks(g, c->plan[0], false);
    if (c->nxpart > 0) HIPCHK(nullptr, hipMalloc(&c->xpart, c->nxpart * sizeof(Red3)));
  }
  // transfers and reductions
  long mx = 0;
  for (int s = 0; s < 4; s++) {
    int e[3];
    extents(g, s, e);
    mx = std::max(mx, (long)e[0] * e[1] * e[2]);
  }
  c->staging_elems = mx;
  HIPCHK(nullptr, hipMalloc(&c->staging, mx * sizeof(double)));
  HIPCHK(nullptr, hipMalloc(&c->partial, c->nred * sizeof(Red3)));
  HIPCHK(nullptr, hipMalloc(&c->red_out, sizeof(Red3)));
  HIPCHK(nullptr, hipStreamSynchronize(c->stream));
  // the no-allocation entry point's scratch on this device
  const hipError_t se = scratch_address(device, &c->scratch);
  if (se != hipSuccess) c->scratch_err = hipGetErrorString(se);
  return MIF_OK;
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

const char* mif_last_error(const mif_ctx* ctx) { return ctx ? ctx->err.c_str() : g_create_error.c_str(); }

static int ctx_create(const mif_grid* grid, int precision, int device, const int cell[3], mif_ctx** out);

int mif_ctx_create(const mif_grid* grid, int precision, int device, mif_ctx** out) {
  return ctx_create(grid, precision, device, nullptr, out);
}

int mif_ctx_create_cell_centred(const mif_grid* grid, const int cell_centred[3], int precision, int device,
                                mif_ctx** out) {
  if (!grid || !cell_centred || !out) return fail(nullptr, MIF_ERR_INVALID_ARGUMENT, "null argument");
  for (int a = 0; a < 3; a++) {
    if (!cell_centred[a]) continue;
    if (grid->periodic[a])
      return fail(nullptr, MIF_ERR_INVALID_ARGUMENT, "axis %d: a cell-centred axis has walls, not periodic ends", a);
    if (a != 1)
      return fail(nullptr, MIF_ERR_UNSUPPORTED, "axis %d: cell-centred transforms run on the y axis only", a);
    const uint64_t n = grid->Ny_global - 1;
    if (n != 64 && n != 128 && n != 256 && n != 512 && n != 1024)
      return fail(nullptr, MIF_ERR_UNSUPPORTED,
                  "cell-centred y axis of %llu cells: 64, 128, 256, 512 or 1024 cells are supported",
                  (unsigned long long)n);
  }
  return ctx_create(grid, precision, device, cell_centred, out);
}

// LDS bytes one workgroup of the generic transform kernel needs on an axis
// (two line buffers of the complex length, one line pair, FP64 so that the
// mixed form's FP64 transforms fit too); 0 where a specialised kernel runs.
static size_t generic_transform_lds(int axis, uint64_t Ng, bool periodic, bool cell) {
  if (cell) return 0;
  const long n = (long)Ng - (periodic ? 1 : 0);
  const long L = periodic ? n : 2 * (n - 1);
  if (L < 2 || fft_fast_len((int)L)) return 0;
  // the half-length DCT-I of 513, 769 and 1025-node wall-normal y axes
  if (axis == 1 && !periodic && (L / 2 == 512 || L / 2 == 768 || L / 2 == 1024)) return 0;
  return 2 * (size_t)L * 2 * sizeof(double);
}
constexpr size_t kGenericTransformLdsMax = 64 * 1024;

static int ctx_create(const mif_grid* grid, int precision, int device, const int cell[3], mif_ctx** out) {
  if (!grid || !out) return fail(nullptr, MIF_ERR_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  if (precision != MIF_PRECISION_FP64 && precision != MIF_PRECISION_MIXED)
    return fail(nullptr, MIF_ERR_INVALID_ARGUMENT, "unknown precision %d", precision);
  const uint64_t Ng[3] = {grid->Nx_global, grid->Ny_global, grid->Nz_global};
  for (int a = 0; a < 3; a++) {
    if (Ng[a] < 2 || Ng[a] > (1u << 20)) return fail(nullptr, MIF_ERR_INVALID_ARGUMENT, "axis %d: %llu nodes", a, (unsigned long long)Ng[a]);
    if (grid->periodic[a] && Ng[a] < 3)
      return fail(nullptr, MIF_ERR_INVALID_ARGUMENT, "periodic axis %d needs at least 3 nodes", a);
  }
  if (grid->num_time_steps == 0 || !(grid->final_time > 0))
    return fail(nullptr, MIF_ERR_INVALID_ARGUMENT, "final_time and num_time_steps must be positive");
  if (!(grid->Re > 0)) return fail(nullptr, MIF_ERR_INVALID_ARGUMENT, "Re must be positive");
  for (int a = 0; a < 3; a++) {
    const size_t lds = generic_transform_lds(a, Ng[a], grid->periodic[a] != 0, cell && cell[a]);
    if (lds > kGenericTransformLdsMax)
      return fail(nullptr, MIF_ERR_UNSUPPORTED,
                  "axis %d: %llu nodes give a transform length with no specialised kernel whose generic kernel "
                  "needs %zu bytes of LDS per workgroup (limit %zu)",
                  a, (unsigned long long)Ng[a], lds, kGenericTransformLdsMax);
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(nullptr, MIF_ERR_NO_DEVICE, "no HIP device");
  mif_ctx* c = new (std::nothrow) mif_ctx();
  if (!c) return fail(nullptr, MIF_ERR_OUT_OF_MEMORY, "host allocation failed");
  const int rc = ctx_init(c, grid, Ng, precision, device, cell);
  if (rc != MIF_OK) {
    mif_ctx_destroy(c);
    return rc;
  }
  *out = c;
  return MIF_OK;
}

int mif_ctx_destroy(mif_ctx* c) {
  if (!c) return MIF_OK;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (auto& r : c->pending) {
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
  }
  for (auto e : c->pool) (void)hipEventDestroy(e);
  for (mif_field* f : c->fields) {
    free_field_storage(c, f->data);
    f->data = nullptr;
    f->ctx = nullptr;
  }
  c->fields.clear();
  free_field_storage(c, c->pswap);
  (void)hipFree(c->B);
  (void)hipFree(c->B64);
  (void)hipFree(c->P64);
  (void)hipFree(c->B2);
  (void)hipFree(c->xpart);
  (void)hipFree(c->red_b);
  (void)hipFree(c->R64);
  for (int q = 0; q < 4; q++)
    for (int cc = 0; cc < 3; cc++) free_field_storage(c, c->adj[q][cc]);
  (void)hipFree(c->ftab);
  (void)hipFree(c->adj_partial);
  (void)hipFree(c->adj_fbar);
  (void)hipFree(c->trig);
  for (int a = 0; a < 3; a++) {
    (void)hipFree(c->W64[a]);
    (void)hipFree(c->W32[a]);
    (void)hipFree(c->H64[a]);
    (void)hipFree(c->H32[a]);
    (void)hipFree(c->lam[a]);
  }
  (void)hipFree(c->staging);
  (void)hipFree(c->partial);
  (void)hipFree(c->red_out);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return MIF_OK;
}

int mif_ctx_synchronize(mif_ctx* c) {
  if (!c) return fail(nullptr, MIF_ERR_INVALID_ARGUMENT, "null context");
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return MIF_OK;
}

int mif_ctx_info(const mif_ctx* c, double* dt, uint64_t ext[3]) {
  if (!c) return fail(nullptr, MIF_ERR_INVALID_ARGUMENT, "null context");
  if (dt) *dt = c->g.dt;
  if (ext)
    for (int a = 0; a < 3; a++) ext[a] = (uint64_t)c->g.N[a];
  return MIF_OK;
}

int mif_ctx_set_stage_records(mif_ctx* c, int keep) {
  if (!c) return fail(nullptr, MIF_ERR_INVALID_ARGUMENT, "null context");
  c->stage_records = keep != 0;
  return MIF_OK;
}

int mif_ctx_set_defect_sweeps(mif_ctx* c, int sweeps) {
  if (!c) return fail(nullptr, MIF_ERR_INVALID_ARGUMENT, "null context");
  if (sweeps < 0 || sweeps > 4) return fail(c, MIF_ERR_INVALID_ARGUMENT, "sweeps must be in 0..4");
  c->sweeps = sweeps;
  return MIF_OK;
}

int mif_ctx_set_defect_schedule(mif_ctx* c, int schedule) {
  if (!c) return fail(nullptr, MIF_ERR_INVALID_ARGUMENT, "null context");
  if (schedule != MIF_DEFECT_EVERY_STAGE && schedule != MIF_DEFECT_LAST_STAGE)
    return fail(c, MIF_ERR_INVALID_ARGUMENT, "unknown defect schedule %d", schedule);
  c->defect_schedule = schedule;
  return MIF_OK;
}

int mif_ctx_set_solve_precision(mif_ctx* c, int mode) {
  if (!c) return fail(nullptr, MIF_ERR_INVALID_ARGUMENT, "null context");
  if (mode != MIF_SOLVE_AUTO && mode != MIF_SOLVE_FP32 && mode != MIF_SOLVE_FP64)
    return fail(c, MIF_ERR_INVALID_ARGUMENT, "unknown solve precision %d", mode);
  c->solve_mode = mode;
  return MIF_OK;
}

int mif_ctx_get_solve_precision(const mif_ctx* c, int* effective, double* kappa) {
  if (!c || !effective) return fail(nullptr, MIF_ERR_INVALID_ARGUMENT, "null argument");
  *effective = (is_fp64(c) || mixed_solve_fp64(c)) ? MIF_SOLVE_FP64 : MIF_SOLVE_FP32;
  if (kappa) *kappa = c->kappa;
  return MIF_OK;
}

int mif_ctx_get_defect_sweeps(const mif_ctx* c, int* sweeps) {
  if (!c || !sweeps) return fail(nullptr, MIF_ERR_INVALID_ARGUMENT, "null argument");
  *sweeps = c->sweeps;
  return MIF_OK;
}

int mif_field_create(mif_ctx* c, int stag, mif_field** out) {
  if (!c || !out) return fail(c, MIF_ERR_INVALID_ARGUMENT, "null argument");
  if (stag < 0 || stag > 3) return fail(c, MIF_ERR_INVALID_ARGUMENT, "unknown staggering %d", stag);
  if (c->g.cell[0] || c->g.cell[1] || c->g.cell[2])
    return fail(c, MIF_ERR_UNSUPPORTED,
                "a context with a cell-centred axis runs stand-alone Poisson solves only "
                "(mif_poisson_set_rhs / mif_poisson_run / mif_poisson_get_solution); it has no staggered fields");
  mif_field* f = new (std::nothrow) mif_field{c, stag, nullptr};
  if (!f) return fail(c, MIF_ERR_OUT_OF_MEMORY, "host allocation failed");
  hipError_t e = alloc_field_storage(c, &f->data);
  if (e != hipSuccess) {
    delete f;
    return fail(c, MIF_ERR_OUT_OF_MEMORY, "field allocation of %ld bytes failed",
                (long)((c->L.elems + c->soff) * c->esize + kFieldTailBytes));
  }
  e = hipMemsetAsync(f->data, 0, c->L.elems * c->esize, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) {
    free_field_storage(c, f->data);
    delete f;
    return fail(c, MIF_ERR_HIP, "field clear failed: %s", hipGetErrorString(e));
  }
  try {
    c->fields.insert(f);
  } catch (...) {
    free_field_storage(c, f->data);
    delete f;
    return fail(c, MIF_ERR_OUT_OF_MEMORY, "host allocation failed");
  }
  *out = f;
  return MIF_OK;
}

int mif_field_destroy(mif_field* f) {
  if (!f) return MIF_OK;
  if (mif_ctx* c = f->ctx) {
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    free_field_storage(c, f->data);
    c->fields.erase(f);
  }
  delete f;
  return MIF_OK;
}

int mif_field_extents(const mif_field* f, uint64_t ext[3]) {
  if (!f || !f->ctx || !ext) return fail(nullptr, MIF_ERR_INVALID_ARGUMENT, "null argument");
  int e[3];
  extents(f->ctx->g, f->stag, e);
  for (int a = 0; a < 3; a++) ext[a] = (uint64_t)e[a];
  return MIF_OK;
}

int mif_field_upload(mif_field* f, const double* host) {
  if (!f || !f->ctx || !host) return fail(f ? f->ctx : nullptr, MIF_ERR_INVALID_ARGUMENT, "null argument");
  mif_ctx* c = f->ctx;
  int e[3];
  extents(c->g, f->stag, e);
  const long n = (long)e[0] * e[1] * e[2];
  HIPCHK(c, hipMemcpyAsync(c->staging, host, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  if (is_fp64(c))
    launch_from_dense<double>(c->L, e, c->staging, fptr<double>(f), c->stream);
  else
    launch_from_dense<float>(c->L, e, c->staging, fptr<float>(f), c->stream);
  LAUNCHCHK(c);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return MIF_OK;
}

int mif_field_download(const mif_field* f, double* host) {
  if (!f || !f->ctx || !host) return fail(f ? f->ctx : nullptr, MIF_ERR_INVALID_ARGUMENT, "null argument");
  mif_ctx* c = f->ctx;
  int e[3];
  extents(c->g, f->stag, e);
  const long n = (long)e[0] * e[1] * e[2];
  if (is_fp64(c))
    launch_to_dense<double>(c->L, e, (const double*)f->data, c->staging, c->stream);
  else
    launch_to_dense<float>(c->L, e, (const float*)f->data, c->staging, c->stream);
  LAUNCHCHK(c);
  HIPCHK(c, hipMemcpyAsync(host, c->staging, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return MIF_OK;
}

int mif_field_upload_device(mif_field* f, const double* dev) {
  if (!f || !f->ctx || !dev) return fail(f ? f->ctx : nullptr, MIF_ERR_INVALID_ARGUMENT, "null argument");
  mif_ctx* c = f->ctx;
  int e[3];
  extents(c->g, f->stag, e);
  if (is_fp64(c))
    launch_from_dense<double>(c->L, e, dev, fptr<double>(f), c->stream);
  else
    launch_from_dense<float>(c->L, e, dev, fptr<float>(f), c->stream);
  LAUNCHCHK(c);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return MIF_OK;
}

int mif_field_download_device(const mif_field* f, double* dev) {
  if (!f || !f->ctx || !dev) return fail(f ? f->ctx : nullptr, MIF_ERR_INVALID_ARGUMENT, "null argument");
  mif_ctx* c = f->ctx;
  int e[3];
  extents(c->g, f->stag, e);
  if (is_fp64(c))
    launch_to_dense<double>(c->L, e, (const double*)f->data, dev, c->stream);
  else
    launch_to_dense<float>(c->L, e, (const float*)f->data, dev, c->stream);
  LAUNCHCHK(c);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return MIF_OK;
}

int mif_field_swap(mif_field* a, mif_field* b) {
  if (!a || !b || !a->ctx) return fail(nullptr, MIF_ERR_INVALID_ARGUMENT, "null field or destroyed context");
  if (a->ctx != b->ctx) return fail(a->ctx, MIF_ERR_INVALID_ARGUMENT, "fields belong to different contexts");
  if (a->stag != b->stag) return fail(a->ctx, MIF_ERR_INVALID_ARGUMENT, "fields have different staggering");
  std::swap(a->data, b->data);
  return MIF_OK;
}

int mif_field_set(mif_field* f, const mif_function* fn, int comp, double t, int include_border) {
  if (!f || !f->ctx || !fn) return fail(f ? f->ctx : nullptr, MIF_ERR_INVALID_ARGUMENT, "null argument");
  mif_ctx* c = f->ctx;
  if (comp < 0 || comp > 2) return fail(c, MIF_ERR_INVALID_ARGUMENT, "component %d", comp);
  const FnDesc d = make_fn(fn);
  if (is_fp64(c))
    launch_set<double>(c->g, c->L, fptr<double>(f), f->stag, comp, d, t, include_border, c->stream);
  else
    launch_set<float>(c->g, c->L, fptr<float>(f), f->stag, comp, d, t, include_border, c->stream);
  LAUNCHCHK(c);
  return MIF_OK;
}

int mif_field_apply_periodic_bc(mif_field* f) {
  if (!f || !f->ctx) return fail(nullptr, MIF_ERR_INVALID_ARGUMENT, "null field");
  mif_ctx* c = f->ctx;
  if (is_fp64(c))
    launch_periodic<double>(c->g, c->L, fptr<double>(f), f->stag, c->stream);
  else
    launch_periodic<float>(c->g, c->L, fptr<float>(f), f->stag, c->stream);
  LAUNCHCHK(c);
  return MIF_OK;
}

int mif_velocity_apply_bc(mif_field* const vel[3], const mif_function* fn, double t) {
  mif_ctx* c = nullptr;
  int rc = check_vel(vel, &c);
  if (rc) return rc;
  if (!fn) return fail(c, MIF_ERR_INVALID_ARGUMENT, "null function");
  const FnDesc d = make_fn(fn);
  if (is_fp64(c))
    apply_bc<double>(c, vel, d, t);
  else
    apply_bc<float>(c, vel, d, t);
  LAUNCHCHK(c);
  return MIF_OK;
}

int mif_bc_table_size(const mif_ctx* c, int64_t* velocity_faces, int64_t* neumann_faces) {
  if (!c) return fail(nullptr, MIF_ERR_INVALID_ARGUMENT, "null context");
  long toff[3][6];
  if (velocity_faces) *velocity_faces = face_table_layout(c->g, true, toff);
  if (neumann_faces) *neumann_faces = face_table_layout(c->g, false, toff);
  return MIF_OK;
}

int mif_velocity_apply_bc_tabulated(mif_field* const vel[3], const double* faces) {
  mif_ctx* c = nullptr;
  int rc = check_vel(vel, &c);
  if (rc) return rc;
  if (!faces) return fail(c, MIF_ERR_INVALID_ARGUMENT, "null face table");
  FnDesc d;
  if ((rc = upload_face_tables(c, true, faces, 1, &d))) return rc;
  if (is_fp64(c))
    apply_bc<double>(c, vel, d, 0.0);
  else
    apply_bc<float>(c, vel, d, 0.0);
  LAUNCHCHK(c);
  return MIF_OK;
}

// Argument checks of the timestep entry points; c receives the context.
static int check_timestep(mif_field* const vel[3], mif_field* const buf[3], mif_field* const buf2[3], mif_field* p,
                          mif_field* pb, bool nhn, mif_ctx** out) {
  mif_ctx *c = nullptr, *c2 = nullptr, *c3 = nullptr;
  int rc;
  if ((rc = check_vel(vel, &c)) || (rc = check_vel(buf, &c2)) || (rc = check_vel(buf2, &c3))) return rc;
  if (c2 != c || c3 != c || !p || !pb || p->ctx != c || pb->ctx != c)
    return fail(c, MIF_ERR_INVALID_ARGUMENT, "inconsistent timestep arguments");
  if (p->stag != MIF_STAGGER_NONE || pb->stag != MIF_STAGGER_NONE)
    return fail(c, MIF_ERR_INVALID_ARGUMENT, "pressure fields must be unstaggered");
  if (nhn && (c->g.periodic[0] || c->g.periodic[1] || c->g.periodic[2]))
    return fail(c, MIF_ERR_INVALID_ARGUMENT, "non-homogeneous Neumann data needs walls on every axis");
  *out = c;
  return MIF_OK;
}

static int run_timestep(mif_ctx* c, mif_field* const vel[3], mif_field* const buf[3], mif_field* const buf2[3],
                        const FnDesc ev[3], const FnDesc* gd, double t_n, mif_field* p, mif_field* pb) {
  if (is_fp64(c)) return full_timestep<double>(c, vel, buf, buf2, ev, gd, t_n, p, pb);
  return full_timestep<float>(c, vel, buf, buf2, ev, gd, t_n, p, pb);
}

int mif_timestep(mif_field* const vel[3], mif_field* const buf[3], mif_field* const buf2[3],
                 const mif_function* exact_velocity, double t_n, mif_field* p, mif_field* pb) {
  mif_ctx* c = nullptr;
  int rc = check_timestep(vel, buf, buf2, p, pb, false, &c);
  if (rc) return rc;
  if (!exact_velocity) return fail(c, MIF_ERR_INVALID_ARGUMENT, "null exact velocity");
  const FnDesc d = make_fn(exact_velocity);
  const FnDesc ev[3] = {d, d, d};
  return run_timestep(c, vel, buf, buf2, ev, nullptr, t_n, p, pb);
}

int mif_timestep_n(mif_field* const vel[3], mif_field* const buf[3], mif_field* const buf2[3],
                   const mif_function* exact_velocity, double t_n, mif_field* p, mif_field* pb, int nsteps) {
  mif_ctx* c = nullptr;
  int rc = check_timestep(vel, buf, buf2, p, pb, false, &c);
  if (rc) return rc;
  if (!exact_velocity) return fail(c, MIF_ERR_INVALID_ARGUMENT, "null exact velocity");
  if (nsteps < 0) return fail(c, MIF_ERR_INVALID_ARGUMENT, "negative step count %d", nsteps);
  const FnDesc d = make_fn(exact_velocity);
  const FnDesc ev[3] = {d, d, d};
  StepCarry cs;
  for (int s = 0; s < nsteps; s++) {
    const double t = t_n + s * c->g.dt;
    const bool last = s == nsteps - 1;
    rc = is_fp64(c) ? full_timestep<double>(c, vel, buf, buf2, ev, nullptr, t, p, pb, &cs, last)
                    : full_timestep<float>(c, vel, buf, buf2, ev, nullptr, t, p, pb, &cs, last);
    if (rc) return rc;
  }
  return MIF_OK;
}

int mif_timestep_nhn(mif_field* const vel[3], mif_field* const buf[3], mif_field* const buf2[3],
                     const mif_function* exact_velocity, const mif_function* exact_grad_p, double t_n, mif_field* p,
                     mif_field* pb) {
  mif_ctx* c = nullptr;
  int rc = check_timestep(vel, buf, buf2, p, pb, true, &c);
  if (rc) return rc;
  if (!exact_velocity || !exact_grad_p) return fail(c, MIF_ERR_INVALID_ARGUMENT, "null function");
  const FnDesc d = make_fn(exact_velocity);
  const FnDesc ev[3] = {d, d, d};
  // TimeVectorFunction::get_difference_over_time(new_time, prev_time) yields
  // f(prev_time) - f(new_time) (src/Timestep.cpp:89-95, src/VectorFunction.cpp:53-61)
  const StageTimes st = stage_times(c->g.dt, t_n);
  FnDesc gd[3];
  for (int s = 0; s < 3; s++) {
    gd[s] = make_fn(exact_grad_p);
    gd[s].base = exact_grad_p->kind;
    gd[s].kind = MIF_FN_TIME_DIFFERENCE;
    gd[s].t0 = st.time[s + 1];
    gd[s].t1 = st.time[s];
  }
  return run_timestep(c, vel, buf, buf2, ev, gd, t_n, p, pb);
}

int mif_timestep_tabulated(mif_field* const vel[3], mif_field* const buf[3], mif_field* const buf2[3],
                           const double* faces, double t_n, mif_field* p, mif_field* pb) {
  mif_ctx* c = nullptr;
  int rc = check_timestep(vel, buf, buf2, p, pb, false, &c);
  if (rc) return rc;
  if (!faces) return fail(c, MIF_ERR_INVALID_ARGUMENT, "null face tables");
  FnDesc ev[3];
  if ((rc = upload_face_tables(c, true, faces, 3, ev))) return rc;
  return run_timestep(c, vel, buf, buf2, ev, nullptr, t_n, p, pb);
}

int mif_timestep_nhn_tabulated(mif_field* const vel[3], mif_field* const buf[3], mif_field* const buf2[3],
                               const double* faces, const double* neumann, double t_n, mif_field* p, mif_field* pb) {
  mif_ctx* c = nullptr;
  int rc = check_timestep(vel, buf, buf2, p, pb, true, &c);
  if (rc) return rc;
  if (!faces || !neumann) return fail(c, MIF_ERR_INVALID_ARGUMENT, "null tables");
  FnDesc ev[3], gd[3];
  if ((rc = upload_face_tables(c, true, faces, 3, ev))) return rc;
  if ((rc = upload_face_tables(c, false, neumann, 3, gd))) return rc;
  return run_timestep(c, vel, buf, buf2, ev, gd, t_n, p, pb);
}

static int velocity_timestep_any(mif_field* const vel[3], mif_field* const buf[3], mif_field* const rhs[3],
                                 const mif_function* ev, const double* faces, const mif_function* fo, double t_n) {
  mif_ctx *c = nullptr, *c2 = nullptr, *c3 = nullptr;
  int rc;
  if ((rc = check_vel(vel, &c)) || (rc = check_vel(buf, &c2)) || (rc = check_vel(rhs, &c3))) return rc;
  if (c2 != c || c3 != c || !(ev || faces) || !fo) return fail(c, MIF_ERR_INVALID_ARGUMENT, "inconsistent arguments");
  FnDesc d[3];
  if (faces) {
    if ((rc = upload_face_tables(c, true, faces, 3, d))) return rc;
  } else {
    d[0] = d[1] = d[2] = make_fn(ev);
  }
  const FnDesc f = make_fn(fo);
  if (is_fp64(c)) return velocity_timestep<double>(c, vel, buf, rhs, d, f, t_n);
  return velocity_timestep<float>(c, vel, buf, rhs, d, f, t_n);
}

int mif_timestep_velocity(mif_field* const vel[3], mif_field* const buf[3], mif_field* const rhs[3],
                          const mif_function* ev, const mif_function* fo, double t_n) {
  return velocity_timestep_any(vel, buf, rhs, ev, nullptr, fo, t_n);
}

int mif_timestep_velocity_tabulated(mif_field* const vel[3], mif_field* const buf[3], mif_field* const rhs[3],
                                    const double* faces, const mif_function* fo, double t_n) {
  if (!faces) return fail(vel && vel[0] ? vel[0]->ctx : nullptr, MIF_ERR_INVALID_ARGUMENT, "null face tables");
  return velocity_timestep_any(vel, buf, rhs, nullptr, faces, fo, t_n);
}

int mif_solve_pressure_homogeneous_periodic(mif_field* p, mif_field* const vel[3], double dt) {
  mif_ctx* c = nullptr;
  int rc = check_vel(vel, &c);
  if (rc) return rc;
  if (!p || p->ctx != c || p->stag != MIF_STAGGER_NONE) return fail(c, MIF_ERR_INVALID_ARGUMENT, "bad pressure field");
  FnDesc none{};
  if ((rc = pressure_solve_any(c, p, vel, dt, none, 0.0, 0))) return rc;
  return finish_standalone_solve(c, p);
}

int mif_solve_pressure_non_homogeneous_neumann(mif_field* p, mif_field* const vel[3], const mif_function* gp,
                                               double t, double dt) {
  mif_ctx* c = nullptr;
  int rc = check_vel(vel, &c);
  if (rc) return rc;
  if (!p || p->ctx != c || p->stag != MIF_STAGGER_NONE || !gp)
    return fail(c, MIF_ERR_INVALID_ARGUMENT, "bad pressure field or gradient");
  if (c->g.periodic[0] || c->g.periodic[1] || c->g.periodic[2])
    return fail(c, MIF_ERR_INVALID_ARGUMENT, "non-homogeneous Neumann data needs walls on every axis");
  const FnDesc d = make_fn(gp);
  return pressure_solve_any(c, p, vel, dt, d, t, 1);
}

int mif_solve_pressure_non_homogeneous_neumann_tabulated(mif_field* p, mif_field* const vel[3], const double* neumann,
                                                         double dt) {
  mif_ctx* c = nullptr;
  int rc = check_vel(vel, &c);
  if (rc) return rc;
  if (!p || p->ctx != c || p->stag != MIF_STAGGER_NONE || !neumann)
    return fail(c, MIF_ERR_INVALID_ARGUMENT, "bad pressure field or Neumann table");
  if (c->g.periodic[0] || c->g.periodic[1] || c->g.periodic[2])
    return fail(c, MIF_ERR_INVALID_ARGUMENT, "non-homogeneous Neumann data needs walls on every axis");
  FnDesc d;
  if ((rc = upload_face_tables(c, false, neumann, 1, &d))) return rc;
  return pressure_solve_any(c, p, vel, dt, d, 0.0, 1);
}

int mif_poisson_solve(mif_field* p) {
  if (!p || !p->ctx || p->stag != MIF_STAGGER_NONE) return fail(p ? p->ctx : nullptr, MIF_ERR_INVALID_ARGUMENT, "bad field");
  mif_ctx* c = p->ctx;
  if (is_fp64(c)) {
    launch_pack<double, double>(c->g, c->L, fptr<double>(p), (double*)c->B, c->bp0, c->bplane, c->stream);
    int rc = spectral_solve<double>(c, (double*)c->B);
    if (rc) return rc;
    launch_unpack<double, double>(c->g, c->L, (double*)c->B, c->bp0, c->bplane, fptr<double>(p), c->stream);
  } else {
    launch_pack<float, float>(c->g, c->L, fptr<float>(p), (float*)c->B, c->bp0, c->bplane, c->stream);
    int rc = spectral_solve<float>(c, (float*)c->B);
    if (rc) return rc;
    launch_unpack<float, float>(c->g, c->L, (float*)c->B, c->bp0, c->bplane, fptr<float>(p), c->stream);
  }
  LAUNCHCHK(c);
  return MIF_OK;
}

int mif_poisson_set_rhs(mif_ctx* c, const double* b) {
  if (!c || !b) return fail(c, MIF_ERR_INVALID_ARGUMENT, "null argument");
  const long n0 = c->g.npts[0], n1 = c->g.npts[1], n2 = c->g.npts[2];
  if (!c->R64) HIPCHK(c, hipMalloc(&c->R64, c->belems * sizeof(double)));
  HIPCHK(c, hipMemsetAsync(c->R64, 0, c->belems * sizeof(double), c->stream));
  HIPCHK(c, hipMemcpy2DAsync(c->R64, c->bp0 * sizeof(double), b, n0 * sizeof(double), n0 * sizeof(double), n1 * n2,
                             hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return MIF_OK;
}

int mif_poisson_run(mif_ctx* c) {
  if (!c) return fail(nullptr, MIF_ERR_INVALID_ARGUMENT, "null context");
  if (!c->R64) return fail(c, MIF_ERR_INVALID_ARGUMENT, "no right-hand side set");
  const size_t bytes = c->belems * sizeof(double);
  if (is_fp64(c)) {
    HIPCHK(c, hipMemcpyAsync(c->B, c->R64, bytes, hipMemcpyDeviceToDevice, c->stream));
    return spectral_solve<double>(c, (double*)c->B);
  }
  if (mixed_solve_fp64(c)) {
    HIPCHK(c, hipMemcpyAsync(c->P64, c->R64, bytes, hipMemcpyDeviceToDevice, c->stream));
    return spectral_solve<double>(c, c->P64);
  }
  HIPCHK(c, hipMemcpyAsync(c->B64, c->R64, bytes, hipMemcpyDeviceToDevice, c->stream));
  SolutionIn where;
  int rc = mixed_solve_from_b64(c, &where);
  if (rc) return rc;
  if (where == IN_B32) {
    KTimer t(c, MIF_KC_DEFECT);
    launch_to64(c->g, (const float*)c->B, c->P64, c->bp0, c->bplane, c->stream);
  }
  LAUNCHCHK(c);
  return MIF_OK;
}

int mif_poisson_get_solution(mif_ctx* c, double* phi) {
  if (!c || !phi) return fail(c, MIF_ERR_INVALID_ARGUMENT, "null argument");
  const long n0 = c->g.npts[0], n1 = c->g.npts[1], n2 = c->g.npts[2];
  const double* src = is_fp64(c) ? (const double*)c->B : c->P64;
  HIPCHK(c, hipMemcpy2DAsync(phi, n0 * sizeof(double), src, c->bp0 * sizeof(double), n0 * sizeof(double), n1 * n2,
                             hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return MIF_OK;
}

namespace {

// Shared bodies of adjust_pressure and the error norms: the exact solution
// comes from a closed-form descriptor (fn) or from device tables (tab).
int adjust_pressure_impl(mif_ctx* c, mif_field* p, const FnDesc& d, double t, const ExactTables& tab) {
  const long total = (long)c->g.npts[0] * c->g.npts[1] * c->g.npts[2];
  const int nb = reduce_blocks(total);
  if (is_fp64(c))
    launch_norm_pressure<double>(c->g, c->L, fptr<double>(p), d, t, 1, tab, c->partial, nb, c->red_out, c->stream);
  else
    launch_norm_pressure<float>(c->g, c->L, fptr<float>(p), d, t, 1, tab, c->partial, nb, c->red_out, c->stream);
  LAUNCHCHK(c);
  Red3 r;
  HIPCHK(c, hipMemcpyAsync(&r, c->red_out, sizeof(Red3), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  // the mean over unique nodes (src/PressureEquation.cpp:312-315), added to all points
  const double difference = r.s1 / (double)((size_t)c->g.npts[0] * c->g.npts[1] * c->g.npts[2]);
  if (is_fp64(c))
    launch_add_const<double>(c->g, c->L, fptr<double>(p), difference, c->stream);
  else
    launch_add_const<float>(c->g, c->L, fptr<float>(p), difference, c->stream);
  LAUNCHCHK(c);
  return MIF_OK;
}

int finish_norms(mif_ctx* c, double out[3], const char* what) {
  Red3 r;
  HIPCHK(c, hipMemcpyAsync(&r, c->red_out, sizeof(Red3), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const double vol = c->g.h[0] * c->g.h[1] * c->g.h[2];
  out[0] = r.s1 * c->g.h[0] * c->g.h[1] * c->g.h[2];
  out[1] = std::sqrt(r.s2 * vol);
  out[2] = r.mx;
  if (!std::isfinite(out[0]) || !std::isfinite(out[1]) || !std::isfinite(out[2]))
    return fail(c, MIF_ERR_NON_FINITE, "non-finite %s error norm", what);
  return MIF_OK;
}

int norms_velocity_impl(mif_ctx* c, mif_field* const vel[3], const FnDesc& d, double t, const ExactTables& tab,
                        double out[3]) {
  const long total = (long)(c->g.N[0] - 2) * (c->g.N[1] - 2) * (c->g.N[2] - 2);
  const int nb = reduce_blocks(total);
  if (is_fp64(c))
    launch_norm_velocity<double>(c->g, c->L, fptr<double>(vel[0]), fptr<double>(vel[1]), fptr<double>(vel[2]), d, t,
                                 tab, c->partial, nb, c->red_out, c->stream);
  else
    launch_norm_velocity<float>(c->g, c->L, fptr<float>(vel[0]), fptr<float>(vel[1]), fptr<float>(vel[2]), d, t,
                                tab, c->partial, nb, c->red_out, c->stream);
  LAUNCHCHK(c);
  return finish_norms(c, out, "velocity");
}

int norms_pressure_impl(mif_ctx* c, const mif_field* p, const FnDesc& d, double t, const ExactTables& tab,
                        double out[3]) {
  const long total = (long)c->g.npts[0] * c->g.npts[1] * c->g.npts[2];
  const int nb = reduce_blocks(total);
  if (is_fp64(c))
    launch_norm_pressure<double>(c->g, c->L, (const double*)p->data, d, t, 0, tab, c->partial, nb, c->red_out,
                                 c->stream);
  else
    launch_norm_pressure<float>(c->g, c->L, (const float*)p->data, d, t, 0, tab, c->partial, nb, c->red_out,
                                c->stream);
  LAUNCHCHK(c);
  return finish_norms(c, out, "pressure");
}

bool is_node_field_of(const mif_field* f, const mif_ctx* c) {
  return f && f->ctx == c && f->stag == MIF_STAGGER_NONE;
}

}  // namespace

int mif_adjust_pressure(mif_field* p, const mif_function* fn, double t) {
  if (!p || !p->ctx || !fn || p->stag != MIF_STAGGER_NONE)
    return fail(p ? p->ctx : nullptr, MIF_ERR_INVALID_ARGUMENT, "bad arguments");
  return adjust_pressure_impl(p->ctx, p, make_fn(fn), t, ExactTables{});
}

int mif_adjust_pressure_tabulated(mif_field* p, const mif_field* exact) {
  if (!p || !p->ctx || p->stag != MIF_STAGGER_NONE || !is_node_field_of(exact, p->ctx))
    return fail(p ? p->ctx : nullptr, MIF_ERR_INVALID_ARGUMENT, "bad arguments");
  ExactTables tab{};
  tab.f[0] = exact->data;
  return adjust_pressure_impl(p->ctx, p, FnDesc{}, 0.0, tab);
}

int mif_error_norms_velocity(mif_field* const vel[3], const mif_function* fn, double t, double out[3]) {
  mif_ctx* c = nullptr;
  int rc = check_vel(vel, &c);
  if (rc) return rc;
  if (!fn || !out) return fail(c, MIF_ERR_INVALID_ARGUMENT, "null argument");
  return norms_velocity_impl(c, vel, make_fn(fn), t, ExactTables{}, out);
}

int mif_error_norms_velocity_tabulated(mif_field* const vel[3], mif_field* const exact[3], double out[3]) {
  mif_ctx* c = nullptr;
  int rc = check_vel(vel, &c);
  if (rc) return rc;
  if (!exact || !out) return fail(c, MIF_ERR_INVALID_ARGUMENT, "null argument");
  ExactTables tab{};
  for (int q = 0; q < 3; q++) {
    if (!is_node_field_of(exact[q], c))
      return fail(c, MIF_ERR_INVALID_ARGUMENT, "exact table %d is not a node field of this context", q);
    tab.f[q] = exact[q]->data;
  }
  return norms_velocity_impl(c, vel, FnDesc{}, 0.0, tab, out);
}

int mif_error_norms_pressure(const mif_field* p, const mif_function* fn, double t, double out[3]) {
  if (!p || !p->ctx || !fn || !out || p->stag != MIF_STAGGER_NONE)
    return fail(p ? p->ctx : nullptr, MIF_ERR_INVALID_ARGUMENT, "bad arguments");
  return norms_pressure_impl(p->ctx, p, make_fn(fn), t, ExactTables{}, out);
}

int mif_error_norms_pressure_tabulated(const mif_field* p, const mif_field* exact, double out[3]) {
  if (!p || !p->ctx || !out || p->stag != MIF_STAGGER_NONE || !is_node_field_of(exact, p->ctx))
    return fail(p ? p->ctx : nullptr, MIF_ERR_INVALID_ARGUMENT, "bad arguments");
  ExactTables tab{};
  tab.f[0] = exact->data;
  return norms_pressure_impl(p->ctx, p, FnDesc{}, 0.0, tab, out);
}

int mif_pressure_gradient(const mif_field* p, mif_field* const out[3]) {
  mif_ctx* c = nullptr;
  int rc = check_vel(out, &c);
  if (rc) return rc;
  if (!p || p->ctx != c || p->stag != MIF_STAGGER_NONE) return fail(c, MIF_ERR_INVALID_ARGUMENT, "bad pressure");
  if (is_fp64(c))
    launch_pgrad<double>(c->g, c->L, (const double*)p->data, fptr<double>(out[0]), fptr<double>(out[1]),
                         fptr<double>(out[2]), c->stream);
  else
    launch_pgrad<float>(c->g, c->L, (const float*)p->data, fptr<float>(out[0]), fptr<float>(out[1]),
                        fptr<float>(out[2]), c->stream);
  LAUNCHCHK(c);
  return MIF_OK;
}

int mif_ctx_set_body_force(mif_ctx* c, const double force[3]) {
  if (!c || !force) return fail(c, MIF_ERR_INVALID_ARGUMENT, "null argument");
  for (int q = 0; q < 3; q++) c->force[q] = force[q];
  return MIF_OK;
}

int mif_timestep_adjoint(mif_field* const un[3], mif_field* const y2[3], mif_field* const y3[3],
                         mif_field* const ubar[3], mif_field* pbar, double force_bar[3]) {
  mif_ctx *c = nullptr, *c2 = nullptr, *c3 = nullptr, *c4 = nullptr;
  int rc;
  if ((rc = check_vel(un, &c)) || (rc = check_vel(y2, &c2)) || (rc = check_vel(y3, &c3)) ||
      (rc = check_vel(ubar, &c4)))
    return rc;
  if (c2 != c || c3 != c || c4 != c || !pbar || pbar->ctx != c || pbar->stag != MIF_STAGGER_NONE)
    return fail(c, MIF_ERR_INVALID_ARGUMENT, "inconsistent adjoint arguments");
  if (is_fp64(c)) return step_adjoint<double>(c, un, y2, y3, ubar, pbar, force_bar);
  return step_adjoint<float>(c, un, y2, y3, ubar, pbar, force_bar);
}

int mif_adjoint_force_take(mif_ctx* c, double force_bar[3]) {
  if (!c || !force_bar) return fail(c, MIF_ERR_INVALID_ARGUMENT, "null argument");
  if (!c->adj_fbar) return MIF_OK;  // no adjoint step has run
  HIPCHK(c, hipSetDevice(c->device));
  double fb[3];
  HIPCHK(c, hipMemcpyAsync(fb, c->adj_fbar + 3, 3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemsetAsync(c->adj_fbar + 3, 0, 3 * sizeof(double), c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (int q = 0; q < 3; q++) force_bar[q] += fb[q];
  return MIF_OK;
}

int mif_field_copy(mif_field* dst, const mif_field* src) {
  if (!dst || !src || !dst->ctx || dst->ctx != src->ctx || dst->stag != src->stag)
    return fail(dst ? dst->ctx : nullptr, MIF_ERR_INVALID_ARGUMENT, "bad fields");
  mif_ctx* c = dst->ctx;
  HIPCHK(c, hipMemcpyAsync(dst->data, src->data, c->L.elems * c->esize, hipMemcpyDeviceToDevice, c->stream));
  return MIF_OK;
}

int mif_field_dot(const mif_field* a, const mif_field* b, double* out) {
  if (!a || !b || !out || !a->ctx || a->ctx != b->ctx || a->stag != b->stag)
    return fail(a ? a->ctx : nullptr, MIF_ERR_INVALID_ARGUMENT, "bad fields");
  mif_ctx* c = a->ctx;
  int e[3];
  extents(c->g, a->stag, e);
  if (is_fp64(c))
    launch_dot<double>(c->L, e, (const double*)a->data, (const double*)b->data, c->partial, c->nred, c->red_out,
                       c->stream);
  else
    launch_dot<float>(c->L, e, (const float*)a->data, (const float*)b->data, c->partial, c->nred, c->red_out,
                      c->stream);
  LAUNCHCHK(c);
  Red3 r;
  HIPCHK(c, hipMemcpyAsync(&r, c->red_out, sizeof(Red3), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  *out = r.s1;
  return MIF_OK;
}

int mif_field_gather(const mif_field* f, const int32_t* ijk, int64_t n, double* out) {
  if (!f || !f->ctx || n < 0 || (n > 0 && (!ijk || !out)))
    return fail(f ? f->ctx : nullptr, MIF_ERR_INVALID_ARGUMENT, "bad gather arguments");
  mif_ctx* c = f->ctx;
  if (n == 0) return MIF_OK;
  int e[3];
  extents(c->g, f->stag, e);
  for (int64_t q = 0; q < n; q++)
    for (int a = 0; a < 3; a++)
      if (ijk[3 * q + a] < 0 || ijk[3 * q + a] >= e[a])
        return fail(c, MIF_ERR_INVALID_ARGUMENT, "gather point %lld outside the field", (long long)q);
  HIPCHK(c, hipSetDevice(c->device));
  int* dijk = nullptr;
  double* dout = nullptr;
  hipError_t err = hipMalloc(&dijk, 3 * n * sizeof(int));
  if (err == hipSuccess) err = hipMalloc(&dout, n * sizeof(double));
  if (err == hipSuccess) err = hipMemcpyAsync(dijk, ijk, 3 * n * sizeof(int), hipMemcpyHostToDevice, c->stream);
  if (err == hipSuccess) {
    if (is_fp64(c))
      launch_gather<double>(c->L, (const double*)f->data, dijk, (long)n, dout, c->stream);
    else
      launch_gather<float>(c->L, (const float*)f->data, dijk, (long)n, dout, c->stream);
    err = hipGetLastError();
  }
  if (err == hipSuccess) err = hipMemcpyAsync(out, dout, n * sizeof(double), hipMemcpyDeviceToHost, c->stream);
  if (err == hipSuccess) err = hipStreamSynchronize(c->stream);
  (void)hipFree(dijk);
  (void)hipFree(dout);
  if (err != hipSuccess) return fail(c, MIF_ERR_HIP, "gather: %s", hipGetErrorString(err));
  return MIF_OK;
}

int mif_field_gather_noalloc(const mif_field* f, const int32_t* ijk, int64_t n, double* out) {
  if (!f || !f->ctx || n < 0 || (n > 0 && (!ijk || !out)))
    return fail(f ? f->ctx : nullptr, MIF_ERR_INVALID_ARGUMENT, "bad gather arguments");
  mif_ctx* c = f->ctx;
  if (n > kGatherNoallocMaxPoints)
    return fail(c, MIF_ERR_INVALID_ARGUMENT, "gather_noalloc: %lld points requested, at most %lld", (long long)n,
                (long long)kGatherNoallocMaxPoints);
  if (n == 0) return MIF_OK;
  if (!c->scratch)
    return fail(c, MIF_ERR_HIP, "gather_noalloc: no device scratch on device %d (%s)", c->device,
                c->scratch_err.c_str());
  int e[3];
  extents(c->g, f->stag, e);
  for (int64_t q = 0; q < n; q++)
    for (int a = 0; a < 3; a++)
      if (ijk[3 * q + a] < 0 || ijk[3 * q + a] >= e[a])
        return fail(c, MIF_ERR_INVALID_ARGUMENT, "gather point %lld outside the field", (long long)q);
  HIPCHK(c, hipSetDevice(c->device));
  int* dijk = (int*)c->scratch;
  double* dout = (double*)((char*)c->scratch + kGatherNoallocValues);
  HIPCHK(c, hipMemcpyAsync(dijk, ijk, 3 * n * sizeof(int), hipMemcpyHostToDevice, c->stream));
  if (is_fp64(c))
    launch_gather<double>(c->L, (const double*)f->data, dijk, (long)n, dout, c->stream);
  else
    launch_gather<float>(c->L, (const float*)f->data, dijk, (long)n, dout, c->stream);
  LAUNCHCHK(c);
  HIPCHK(c, hipMemcpyAsync(out, dout, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return MIF_OK;
}

int mif_ctx_set_timing(mif_ctx* c, int enable) {
  if (!c) return fail(nullptr, MIF_ERR_INVALID_ARGUMENT, "null context");
  drain_timers(c);
  c->timing = enable != 0;
  for (int q = 0; q < MIF_KC_COUNT; q++) {
    c->total_ms[q] = 0;
    c->launches[q] = 0;
  }
  return MIF_OK;
}

int mif_ctx_kernel_time(mif_ctx* c, int cls, double* total_ms, uint64_t* launches) {
  if (!c || cls < 0 || cls >= MIF_KC_COUNT) return fail(c, MIF_ERR_INVALID_ARGUMENT, "bad kernel class");
  drain_timers(c);
  if (total_ms) *total_ms = c->total_ms[cls];
  if (launches) *launches = c->launches[cls];
  return MIF_OK;
}

int mif_stream_probe(mif_ctx* c, int nread, int nwrite, int64_t bytes, int reps, double* us_per_launch) {
  if (!c) return fail(nullptr, MIF_ERR_INVALID_ARGUMENT, "null context");
  if (nread < 0 || nwrite < 0 || nread > kStreamMax || nwrite > kStreamMax || nread + nwrite == 0 || reps < 1 ||
      bytes < 16 || !us_per_launch)
    return fail(c, MIF_ERR_INVALID_ARGUMENT, "stream probe: 0..%d reads and writes, bytes >= 16, reps >= 1",
                kStreamMax);
  const long n4 = (long)(bytes / 16);
  StreamArgs a{};
  void* buf[2 * kStreamMax] = {};
  const int nb = nread + nwrite;
  hipError_t err = hipSuccess;
  for (int q = 0; q < nb && err == hipSuccess; q++) {
    err = hipMalloc(&buf[q], (size_t)n4 * 16);
    if (err == hipSuccess) err = hipMemsetAsync(buf[q], 0, (size_t)n4 * 16, c->stream);
  }
  int ncu = 0;
  if (err == hipSuccess) err = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, c->device);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (err == hipSuccess) err = hipEventCreate(&e0);
  if (err == hipSuccess) err = hipEventCreate(&e1);
  // the fastest of four grid sizes (4 to 32 workgroups per CU)
  float best = 0;
  if (err == hipSuccess) {
    for (int r = 0; r < nread; r++) a.in[r] = (const float4*)buf[r];
    for (int w = 0; w < nwrite; w++) a.out[w] = (float4*)buf[nread + w];
    for (int per_cu = 4; per_cu <= 32 && err == hipSuccess; per_cu *= 2) {
      const int blocks = per_cu * (ncu > 0 ? ncu : 256);
      launch_stream(a, nread, nwrite, n4, blocks, c->stream);  // warm-up
      (void)hipEventRecord(e0, c->stream);
      for (int i = 0; i < reps; i++) launch_stream(a, nread, nwrite, n4, blocks, c->stream);
      (void)hipEventRecord(e1, c->stream);
      err = hipGetLastError();
      float ms = 0;
      if (err == hipSuccess) err = hipEventSynchronize(e1);
      if (err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
      if (err == hipSuccess && (best == 0 || ms < best)) best = ms;
    }
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipStreamSynchronize(c->stream);
  for (int q = 0; q < nb; q++) (void)hipFree(buf[q]);
  if (err != hipSuccess)
    return fail(c, err == hipErrorOutOfMemory ? MIF_ERR_OUT_OF_MEMORY : MIF_ERR_HIP, "stream probe: %s",
                hipGetErrorString(err));
  *us_per_launch = (double)best * 1e3 / reps;
  return MIF_OK;
}

const char* mif_kernel_class_name(int cls) {
  static const char* names[MIF_KC_COUNT] = {"stage",  "bc",      "div_rhs", "fft_x",     "fft_y",        "fft_z_fused",
                                            "unpack", "project", "defect",  "fft_x_div", "fft_x_unpack", "adjoint"};
  if (cls < 0 || cls >= MIF_KC_COUNT) return "unknown";
  return names[cls];
}

}  // extern "C"
