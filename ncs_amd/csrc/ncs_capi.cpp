// ncs_capi.cpp -- C ABI of libncs_hip.so: context, OTF tables, launches,
// host/device frame batches, the C-library drop-in, graphs and autotuning.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "ncs_common.h"

namespace ncs {
int solve_waves_per_cu(int m, bool general, bool pad, bool f16, int wpg);
hipError_t launch_solve(int m, bool general, bool pad, bool f16, unsigned int waves, int wpg,
                        hipStream_t s, const SolveArgs &a);
hipError_t launch_adu_to_photons(const uint16_t *adu, const float *offset, const float *gain,
                                 float *out, long long n, hipStream_t s);
hipError_t launch_gamma_from_camera(const float *var, const float *gain, float *gamma,
                                    long long n, hipStream_t s);
hipError_t launch_adu_levels(const uint16_t *adu, const float *offset, const float *gains,
                             const float *gammas, long long stride, const int *t, float *out,
                             float *gamma, long long n, hipStream_t s);
hipError_t net_pack(const float *P, void *packs, hipStream_t s);
size_t net_packs_bytes();
hipError_t net_forward(const float *x, float *y, void *acts, const void *packs, const float *P,
                       long long n, int grid, hipStream_t s);
hipError_t net_backward(const float *x, const void *acts, const float *gout, float *gP,
                        const void *packs, long long n, int grid, hipStream_t s);
int net_forward_grid(long long n, int cus);
hipError_t net_warm(const FrameDev *frames, int n_frames, long long total, float *u0,
                    const void *packs, const float *P, int grid, hipStream_t s);
hipError_t launch_eval(bool general, const float *data, const float *gamma, const float *u,
                       const float *tables, float alpha, int n, float *cost, float *grad,
                       hipStream_t s);
}  // namespace ncs

using namespace ncs;

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------

static thread_local std::string g_last_error;

static int fail(int code, const std::string &msg) {
  g_last_error = msg;
  return code;
}

namespace ncs {
// For the library's other translation units (the sub-region ABI).
void set_last_error(const char *msg) { g_last_error = msg; }
}  // namespace ncs

#define HIPCHK(x)                                                              \
  do {                                                                         \
    hipError_t e_ = (x);                                                       \
    if (e_ != hipSuccess)                                                      \
      return fail(NCS_E_HIP, std::string(#x) + ": " + hipGetErrorString(e_));  \
  } while (0)

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------

struct DevBuf {
  void *p = nullptr;
  size_t cap = 0;
  hipError_t ensure(size_t bytes) {
    if (bytes <= cap) return hipSuccess;
    if (p) {
      hipError_t e = hipFree(p);
      if (e != hipSuccess) return e;
      p = nullptr;
      cap = 0;
    }
    size_t want = std::max<size_t>(bytes, 256);
    hipError_t e = hipMalloc(&p, want);
    if (e == hipSuccess) cap = want;
    return e;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
  }
  template <class T> T *as() const { return reinterpret_cast<T *>(p); }
};

// A persistent-solve schedule: history depth, grid, workgroup size and the
// arithmetic of the basis changes.
struct Schedule {
  int history = 5;
  int waves_per_cu = 0;  // 0: the occupancy limit
  int wpg = 1;
  bool f16 = false;      // FP16-split basis changes (else FP32)
};

// Key of an autotuned schedule: mask kind, padded geometry, precision.
typedef std::array<bool, 3> TuneKey;

struct ncs_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  int cus = 0;
  char arch[64] = {0};
  DevBuf tables, frames, queue, data, gamma, out, iters, status, cost, tiles_u;
  bool has_mask = false;
  bool general = false;
  int win = 16;  // window side of the reference geometry (the mask's size)
  float mask_fft[256];
  // host copy of the per-lane tables (mask part and linear-start part)
  std::vector<float> tab_host = std::vector<float>(kTableFloats, 0.0f);
  bool has_linear = false;
  float lin_c = 0.0f;
  float lin_c1 = 0.0f, lin_aref = 0.1f;  // the linear start's alpha slope
  bool has_slope = false;
  float k_norm = 0.0f;  // max row sum |2 q_ee| + |2 q_oo| of the mask tables
  // autotuned schedules, keyed by (general mask, padded geometry, FP16 split)
  std::map<TuneKey, Schedule> tuned;
  // last uploaded frame descriptors (pinned staging + completion event)
  std::vector<FrameDev> last_frames;
  bool frames_valid = false;
  void *frames_staging = nullptr;
  size_t frames_staging_cap = 0;
  hipEvent_t frames_event = nullptr;
  bool frames_event_pending = false;
  // conv net (learned warm start): FP32 parameters and BF16 packs
  DevBuf net_params, net_packs, net_tmp_packs, u0_buf;
  bool has_net = false;
  // The frame table, work-queue counters and start buffer above are shared
  // by every solve of the context: a solve enqueued on another stream than
  // the previous one first waits on the device for that one's completion
  // event.
  hipEvent_t solve_done = nullptr;
  hipStream_t solve_stream = nullptr;
  bool solve_pending = false;
  // device address of the allocation-free entry point's scratch on this GPU
  void *scratch = nullptr;
};

// Orders a solve on stream s after the context's previous one.
static hipError_t order_after_previous(ncs_ctx *ctx, hipStream_t s) {
  if (ctx->solve_pending && ctx->solve_stream != s)
    return hipStreamWaitEvent(s, ctx->solve_done, 0);
  return hipSuccess;
}

static hipError_t mark_solve(ncs_ctx *ctx, hipStream_t s) {
  hipError_t e = hipEventRecord(ctx->solve_done, s);
  if (e == hipSuccess) {
    ctx->solve_stream = s;
    ctx->solve_pending = true;
  }
  return e;
}


static hipError_t set_device(ncs_ctx *ctx) { return hipSetDevice(ctx->device); }

// ---------------------------------------------------------------------------
// device scratch of the allocation-free entry point
// ---------------------------------------------------------------------------

// ncs_hip_reduce_frames_device_noalloc keeps its work-queue counters and its
// frame table in this fixed array instead of context buffers: kQueueBytes
// of counters, then up to kScratchFrames frame descriptors (21,504 bytes).
// Each GPU holds its own copy until the process ends; every call on a GPU
// shares it.
constexpr int kScratchFrames = 256;
constexpr size_t kScratchFrameOffset = 1024;
constexpr size_t kScratchBytes = kScratchFrameOffset + kScratchFrames * sizeof(FrameDev);
static_assert(kQueueBytes <= kScratchFrameOffset, "queue counters overflow their slot");
static_assert(kScratchFrameOffset % alignof(FrameDev) == 0, "misaligned frame table");
static_assert(kScratchBytes == 21504, "the documented scratch size");
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void *g_scratch[kMaxDevices] = {};

// Device address of `scratch` on `device`, which must be current; looked up
// once per device.  Host code uses this address, never the array's name.
static int scratch_on_device(int device, void **out) {
  if (device < 0 || device >= kMaxDevices) return fail(NCS_E_ARG, "device index out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch[device]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(NCS_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch[device] = p;
  }
  *out = g_scratch[device];
  return NCS_OK;
}

// Frame descriptors passed by value as a kernel argument and written to the
// scratch frame table: no host staging buffer and no copy from host memory.
constexpr int kFramesPerWrite = 24;
struct FrameChunk {
  int first, n;
  FrameDev f[kFramesPerWrite];
};
static_assert(sizeof(FrameChunk) <= 4096, "kernel argument too large");

__global__ void write_frames_kernel(FrameChunk c, FrameDev *dst) {
  const int i = threadIdx.x;
  if (i < c.n) dst[c.first + i] = c.f[i];
}

extern "C" void ncs_hip_default_params(ncs_params *p) {
  p->max_iterations = 200;
  p->history = 5;
  p->epsilon = 1.0e-5f;
  p->linesearch = NCS_LS_MORE_THUENTE;
  p->max_linesearch = 20;
  p->ftol = 1.0e-4f;
  p->gtol = 0.9f;
  p->xtol = 1.0e-7f;
  p->min_step = 1.0e-20f;
  p->max_step = 1.0e20f;
  p->stepm = 0.5f;
  p->start = NCS_START_PLAIN;
  p->fallback = 1;
  p->waves_per_cu = 0;
  p->precond = 2;
  p->waves_per_group = 0;
  p->precision = NCS_PRECISION_FP32;
}

extern "C" const char *ncs_hip_last_error(void) { return g_last_error.c_str(); }

extern "C" int ncs_hip_create(int device, ncs_ctx **out) {
  if (!out) return fail(NCS_E_ARG, "out is NULL");
  *out = nullptr;
  int n = 0;
  HIPCHK(hipGetDeviceCount(&n));
  if (device < 0 || device >= n) return fail(NCS_E_ARG, "no such device");
  std::unique_ptr<ncs_ctx> c(new ncs_ctx());
  c->device = device;
  HIPCHK(hipSetDevice(device));
  hipDeviceProp_t prop;
  HIPCHK(hipGetDeviceProperties(&prop, device));
  c->cus = prop.multiProcessorCount;
  std::snprintf(c->arch, sizeof(c->arch), "%s", prop.gcnArchName);
  HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  HIPCHK(hipEventCreateWithFlags(&c->frames_event, hipEventDisableTiming));
  HIPCHK(hipEventCreateWithFlags(&c->solve_done, hipEventDisableTiming));
  HIPCHK(c->queue.ensure(kQueueBytes));
  const int rc = scratch_on_device(device, &c->scratch);
  if (rc) return rc;
  *out = c.release();
  return NCS_OK;
}

extern "C" void ncs_hip_destroy(ncs_ctx *ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  DevBuf *bufs[] = {&ctx->tables, &ctx->frames, &ctx->queue, &ctx->data, &ctx->gamma,
                    &ctx->out, &ctx->iters, &ctx->status, &ctx->cost, &ctx->tiles_u,
                    &ctx->net_params, &ctx->net_packs, &ctx->net_tmp_packs, &ctx->u0_buf};
  for (DevBuf *b : bufs) b->release();
  if (ctx->frames_event) (void)hipEventDestroy(ctx->frames_event);
  if (ctx->solve_done) (void)hipEventDestroy(ctx->solve_done);
  if (ctx->frames_staging) (void)hipHostFree(ctx->frames_staging);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

extern "C" int ncs_hip_device_info(ncs_ctx *ctx, char *buf, int buflen) {
  if (!ctx || !buf || buflen <= 0) return fail(NCS_E_ARG, "bad arguments");
  std::snprintf(buf, buflen, "libncs_hip device=%d arch=%s cus=%d", ctx->device,
                ctx->arch, ctx->cus);
  return NCS_OK;
}

// ---------------------------------------------------------------------------
// OTF tables
// ---------------------------------------------------------------------------

// Orthonormal real DFT basis of length 16: rows cos k (k = 0..8) then
// sin k (k = 1..7).
// Orthonormal real DFT basis of length r (even, <= 16) in the 16 slots of
// the tile layout: slot k holds cos k (k = 0..r/2), slot 8 + k holds sin k
// (k = 1..r/2 - 1), so a cosine row and its sine partner are always 8 slots
// apart; the other slots, and pixels i >= r, are zero.
static bool slot_used(int a, int r) { return a <= 8 ? a <= r / 2 : a - 8 < r / 2; }
static double rbasis(int a, int i, int r = 16) {
  const double PI = 3.14159265358979323846;
  if (i >= r || !slot_used(a, r)) return 0.0;
  const double n0 = 1.0 / std::sqrt((double)r), n1 = std::sqrt(2.0 / r);
  if (a == 0) return n0;
  if (a == r / 2) return (i & 1) ? -n0 : n0;
  if (a <= 8) return n1 * std::cos(2.0 * PI * a * i / r);
  return n1 * std::sin(2.0 * PI * (a - 8) * i / r);
}

static int freq_of(int a) { return a <= 8 ? a : a - 8; }

// v = hi + lo in FP16 (round to nearest even), |v - hi - lo| <~ 2^-22 |v|.
static void split_f16(double v, uint16_t &hi, uint16_t &lo) {
  const _Float16 h = (_Float16)(float)v;
  const _Float16 l = (_Float16)(float)(v - (double)(float)h);
  std::memcpy(&hi, &h, 2);
  std::memcpy(&lo, &l, 2);
}

extern "C" int ncs_hip_set_otf_mask(ncs_ctx *ctx, const float *mask) {
  return ncs_hip_set_otf_mask_window(ctx, mask, 16);
}

extern "C" int ncs_hip_set_otf_mask_window(ncs_ctx *ctx, const float *mask, int win) {
  if (!ctx || !mask) return fail(NCS_E_ARG, "bad arguments");
  if (win < 4 || win > 16 || (win & 1)) return fail(NCS_E_ARG, "window side must be even, 4..16");
  double q[16][16] = {}, qcs[16][16] = {};
  double qmax = 0.0;
  for (int k = 0; k < win; k++)
    for (int l = 0; l < win; l++) {
      double m = mask[k * win + l];
      if (!std::isfinite(m)) return fail(NCS_E_MASK, "mask has non-finite values");
      q[k][l] = m * m;
    }
  for (int k = 0; k < win; k++)
    for (int l = 0; l < win; l++) {
      qcs[k][l] = 0.5 * (q[k][l] + q[(win - k) % win][(win - l) % win]);
      qmax = std::max(qmax, std::fabs(qcs[k][l]));
    }
  std::vector<float> &tab = ctx->tab_host;
  double qo_max = 0.0, k_norm = 0.0;
  for (int lane = 0; lane < 64; lane++) {
    int g = lane >> 4, x = lane & 15;
    for (int r = 0; r < 4; r++) {
      int b = 4 * g + r, a = x;
      tab[lane * kTableLaneFloats + r] = (float)rbasis(x, 4 * g + r, win);
      tab[lane * kTableLaneFloats + 4 + r] = (float)rbasis(4 * g + r, x, win);
      // FP16 hi/lo pairs of the same entries, each duplicated into both
      // halves of an 8-wide FP16 operand (slots r and 4 + r)
      for (int t = 0; t < 2; t++) {
        const double v = t == 0 ? rbasis(x, 4 * g + r, win) : rbasis(4 * g + r, x, win);
        uint16_t hi, lo;
        split_f16(v, hi, lo);
        uint16_t *h = reinterpret_cast<uint16_t *>(&tab[lane * kTableLaneFloats + 24 + 8 * t]);
        h[r] = h[4 + r] = hi;      // entry 6 (t = 0) or 8 (t = 1): hi
        h[8 + r] = h[12 + r] = lo;  // entry 7 or 9: lo
      }
      int k = freq_of(a), l = freq_of(b);
      const bool used = slot_used(a, win) && slot_used(b, win);
      double qee = used ? 0.5 * (qcs[k][l] + qcs[(win - k) % win][l]) : 0.0;
      double qoo = used ? 0.5 * (qcs[k][l] - qcs[(win - k) % win][l]) : 0.0;
      double sgn = ((a <= 8) == (b <= 8)) ? -1.0 : 1.0;
      tab[lane * kTableLaneFloats + 8 + r] = (float)(2.0 * qee);
      tab[lane * kTableLaneFloats + 12 + r] = (float)(sgn * 2.0 * qoo);
      qo_max = std::max(qo_max, std::fabs(qoo));
      k_norm = std::max(k_norm, 2.0 * std::fabs(qee) + 2.0 * std::fabs(qoo));
    }
  }
  HIPCHK(set_device(ctx));
  HIPCHK(ctx->tables.ensure(tab.size() * sizeof(float)));
  HIPCHK(hipMemcpy(ctx->tables.p, tab.data(), tab.size() * sizeof(float),
                   hipMemcpyHostToDevice));
  ctx->general = qo_max > 1e-12 * std::max(qmax, 1e-300);
  ctx->k_norm = (float)k_norm;
  std::memset(ctx->mask_fft, 0, sizeof(ctx->mask_fft));
  std::memcpy(ctx->mask_fft, mask, (size_t)win * win * sizeof(float));
  ctx->win = win;
  ctx->has_mask = true;
  return NCS_OK;
}

extern "C" int ncs_hip_set_linear_start(ncs_ctx *ctx, const float *coef) {
  if (!ctx || !coef) return fail(NCS_E_ARG, "bad arguments");
  for (int i = 0; i < NCS_LINEAR_START_PARAMS; i++)
    if (!std::isfinite(coef[i])) return fail(NCS_E_ARG, "non-finite coefficient");
  // lane (g, j) slot r holds Z^T[4g + r][j] = Z[j][4g + r]
  std::vector<float> &tab = ctx->tab_host;
  for (int lane = 0; lane < 64; lane++) {
    const int g = lane >> 4, j = lane & 15;
    for (int r = 0; r < 4; r++) {
      const int k = j * 16 + 4 * g + r;
      tab[lane * kTableLaneFloats + 16 + r] = coef[k];
      tab[lane * kTableLaneFloats + 20 + r] = coef[256 + k];
      tab[lane * kTableLaneFloats + 4 * kTableSlope + r] = 0.0f;  // slopes cleared
      tab[lane * kTableLaneFloats + 4 * kTableSlope + 4 + r] = 0.0f;
    }
  }
  HIPCHK(set_device(ctx));
  HIPCHK(ctx->tables.ensure(tab.size() * sizeof(float)));
  HIPCHK(hipMemcpy(ctx->tables.p, tab.data(), tab.size() * sizeof(float),
                   hipMemcpyHostToDevice));
  ctx->lin_c = coef[512];
  ctx->lin_c1 = 0.0f;
  ctx->has_slope = false;
  ctx->has_linear = true;
  return NCS_OK;
}

extern "C" int ncs_hip_set_linear_start_alpha(ncs_ctx *ctx, const float *slope,
                                             float alpha_ref) {
  if (!ctx || !slope) return fail(NCS_E_ARG, "bad arguments");
  if (!ctx->has_linear) return fail(NCS_E_NOWEIGHTS, "set the linear start's coefficients first");
  if (!(alpha_ref > 0.0f) || !std::isfinite(alpha_ref))
    return fail(NCS_E_ARG, "alpha_ref must be positive and finite");
  for (int i = 0; i < NCS_LINEAR_START_PARAMS; i++)
    if (!std::isfinite(slope[i])) return fail(NCS_E_ARG, "non-finite slope");
  std::vector<float> &tab = ctx->tab_host;
  for (int lane = 0; lane < 64; lane++) {
    const int g = lane >> 4, j = lane & 15;
    for (int r = 0; r < 4; r++) {
      const int k = j * 16 + 4 * g + r;
      tab[lane * kTableLaneFloats + 4 * kTableSlope + r] = slope[k];
      tab[lane * kTableLaneFloats + 4 * kTableSlope + 4 + r] = slope[256 + k];
    }
  }
  HIPCHK(set_device(ctx));
  HIPCHK(ctx->tables.ensure(tab.size() * sizeof(float)));
  HIPCHK(hipMemcpy(ctx->tables.p, tab.data(), tab.size() * sizeof(float),
                   hipMemcpyHostToDevice));
  ctx->lin_c1 = slope[512];
  ctx->lin_aref = alpha_ref;
  ctx->has_slope = true;
  return NCS_OK;
}

// ---------------------------------------------------------------------------
// tiling
// ---------------------------------------------------------------------------

// Windows along an axis of n pixels: the reference geometry with side r
// steps by s = r - 2 and needs n >= s; 1 + ceil((n - s) / s) windows
// (ncs.c:71-143).  Blocks of side r: ceil(n / r).
static long long tiles_axis(int n, int geom, int r = 16) {
  if (geom == NCS_GEOM_REFERENCE) {
    const int s = r - 2;
    if (n < s) return -1;
    return 1 + (n - s + s - 1) / s;
  }
  if (n < 1) return -1;
  return (n + r - 1) / r;
}

extern "C" long long ncs_hip_tile_count_window(int height, int width, int geometry, int r) {
  if (geometry != NCS_GEOM_REFERENCE && geometry != NCS_GEOM_BLOCKS) return -1;
  if (r < 4 || r > 16 || (r & 1)) return -1;
  long long a = tiles_axis(height, geometry, r), b = tiles_axis(width, geometry, r);
  if (a < 0 || b < 0) return -1;
  return a * b;
}

extern "C" long long ncs_hip_tile_count(int height, int width, int geometry) {
  return ncs_hip_tile_count_window(height, width, geometry, 16);
}

// A window side other than 16 (a smaller mask) takes the reference
// geometry and the plain start: the block geometry, tile entry points and
// the learned starts are built on 16 x 16 tiles.
static int check_window(const ncs_ctx *ctx, const ncs_params *p) {
  if (ctx->win != 16 && p->start != NCS_START_PLAIN)
    return fail(NCS_E_ARG, "a " + std::to_string(ctx->win) + " x " + std::to_string(ctx->win) +
                               " mask takes the plain start");
  return NCS_OK;
}

static int check_params(const ncs_params *p) {
  if (!p) return fail(NCS_E_ARG, "params is NULL");
  if (p->history < 0 || p->history > 8)
    return fail(NCS_E_ARG, "history must be 1..8 (or 0 for the autotuned depth)");
  if (p->max_iterations < 0 || p->max_iterations > 100000)
    return fail(NCS_E_ARG, "max_iterations out of range");
  if (p->max_linesearch < 1 || p->max_linesearch > 1000)
    return fail(NCS_E_ARG, "max_linesearch out of range");
  if (!(p->epsilon >= 0.0f)) return fail(NCS_E_ARG, "epsilon must be >= 0");
  if (p->linesearch != NCS_LS_MORE_THUENTE && p->linesearch != NCS_LS_ARMIJO)
    return fail(NCS_E_ARG, "unknown line search");
  if (!(p->min_step > 0.0f) || !(p->max_step > p->min_step))
    return fail(NCS_E_ARG, "bad step bounds");
  if (!(p->ftol > 0.0f && p->ftol < 0.5f)) return fail(NCS_E_ARG, "ftol must be in (0, 0.5)");
  if (p->linesearch == NCS_LS_MORE_THUENTE && !(p->gtol > p->ftol && p->gtol < 1.0f))
    return fail(NCS_E_ARG, "gtol must be in (ftol, 1)");
  if (p->linesearch == NCS_LS_ARMIJO && !(p->stepm > 0.0f && p->stepm < 1.0f))
    return fail(NCS_E_ARG, "stepm must be in (0, 1)");
  if (!(p->xtol >= 0.0f)) return fail(NCS_E_ARG, "xtol must be >= 0");
  if (p->waves_per_cu < 0 || p->waves_per_cu > 64) return fail(NCS_E_ARG, "waves_per_cu out of range");
  if (p->precond < 0 || p->precond > 2) return fail(NCS_E_ARG, "precond must be 0, 1 or 2");
  if (p->waves_per_group < 0 || p->waves_per_group > kMaxWavesPerGroup)
    return fail(NCS_E_ARG, "waves_per_group must be 0..4");
  if (p->start < NCS_START_PLAIN || p->start > NCS_START_LINEAR)
    return fail(NCS_E_ARG, "unknown start");
  if (p->precision != NCS_PRECISION_FP32 && p->precision != NCS_PRECISION_F16_SPLIT)
    return fail(NCS_E_ARG, "unknown precision");
  return NCS_OK;
}

// Builds kernel frame descriptors (device pointers) and the tile total.
static int describe_frames(const ncs_ctx *ctx, const ncs_frame *frames, int n, FrameDev *fd,
                           long long &total) {
  const int win = ctx->win;
  total = 0;
  for (int i = 0; i < n; i++) {
    const ncs_frame &f = frames[i];
    if (!f.data || !f.gamma) return fail(NCS_E_ARG, "frame without data or gamma");
    long long nt = ncs_hip_tile_count_window(f.height, f.width, f.geometry, win);
    if (nt < 0)
      return fail(NCS_E_ARG, "frame " + std::to_string(i) + ": bad size or geometry (the "
                  "reference tiling needs height, width >= " + std::to_string(win - 2) + ")");
    if (!std::isfinite(f.alpha)) return fail(NCS_E_ARG, "non-finite alpha");
    FrameDev &d = fd[i];
    d.data = f.data;
    d.gamma = f.gamma;
    d.out = f.out;
    d.H = f.height;
    d.W = f.width;
    d.ds = f.data_stride > 0 ? f.data_stride : f.width;
    d.gs = f.gamma_stride > 0 ? f.gamma_stride : f.width;
    d.os = f.out_stride > 0 ? f.out_stride : f.width;
    if (d.ds < d.W || d.gs < d.W || d.os < d.W) return fail(NCS_E_ARG, "stride < width");
    const long long rows = d.H;
    if (rows * std::max(d.ds, std::max(d.gs, d.os)) >= (1LL << 31))
      return fail(NCS_E_ARG, "frame " + std::to_string(i) + " spans 2^31 or more elements");
    d.alpha = f.alpha;
    d.geom = f.geometry;
    d.ty = (int)tiles_axis(f.width, f.geometry, win);
    d.win = win;
    d.first = total;
    d.slot = total;
    total += nt;
  }
  if (total > 0xF0000000LL) return fail(NCS_E_ARG, "too many tiles in one batch");
  return NCS_OK;
}

static int build_frames(const ncs_ctx *ctx, const ncs_frame *frames, int n,
                        std::vector<FrameDev> &fd, long long &total) {
  fd.resize(n);
  return describe_frames(ctx, frames, n, fd.data(), total);
}

// Frame order of the schedule.  The solver claims tiles in frame-table order:
// the first half in static per-wave blocks, the rest from per-XCD queues, so
// the last tiles claimed set the launch's tail.  Frames are scheduled smallest
// first (stable, by tile count): small frames and ROIs carry the tiles that
// run long (edge blocks that are mostly padding, few valid pixels), and this
// order gives them to many waves at the start, leaving ordinary full tiles of
// the large frames for the end.  On a batch of equal frames it is the
// caller's order.  Results do not depend on it: each tile is solved by one
// wave on its own, and per-tile arrays stay in the caller's order (`slot`).
// `fd` holds the frames in the caller's order (first == slot); `order` is the
// schedule (indices into fd).
static long long frame_tiles(const FrameDev *fd, int n, long long total, int i) {
  return (i + 1 < n ? fd[i + 1].slot : total) - fd[i].slot;
}

static void schedule_small_first(std::vector<FrameDev> &fd, long long total) {
  const int n = (int)fd.size();
  std::vector<long long> nt(n);
  for (int i = 0; i < n; i++) nt[i] = frame_tiles(fd.data(), n, total, i);
  std::vector<int> order(n);
  for (int i = 0; i < n; i++) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return nt[a] < nt[b]; });
  std::vector<FrameDev> out(n);
  long long first = 0;
  for (int k = 0; k < n; k++) {
    out[k] = fd[order[k]];
    out[k].first = first;
    first += nt[order[k]];
  }
  fd.swap(out);
}

// The same order without allocating: a stable insertion sort over at most
// kScratchFrames frames into the caller-provided `order`; the caller places
// frame order[k] at schedule position k.
static void schedule_small_first_fixed(const FrameDev *fd, int n, long long total, int *order) {
  for (int k = 0; k < n; k++) {
    int j = k;
    const long long key = frame_tiles(fd, n, total, k);
    while (j > 0 && frame_tiles(fd, n, total, order[j - 1]) > key) {
      order[j] = order[j - 1];
      j--;
    }
    order[j] = k;
  }
}

// Frame descriptors go to the device through a pinned staging copy; a batch
// identical to the previous one (the common streaming / benchmark case) is
// not uploaded again.  The staging buffer is rewritten only after the event of
// its last copy has completed.
static int upload_frames(ncs_ctx *ctx, const std::vector<FrameDev> &fd, hipStream_t s) {
  const size_t bytes = fd.size() * sizeof(FrameDev);
  if (ctx->frames_valid && ctx->last_frames.size() == fd.size() &&
      std::memcmp(ctx->last_frames.data(), fd.data(), bytes) == 0)
    return NCS_OK;
  if (ctx->frames_event_pending) {
    HIPCHK(hipEventSynchronize(ctx->frames_event));
    ctx->frames_event_pending = false;
  }
  if (bytes > ctx->frames_staging_cap) {
    if (ctx->frames_staging) HIPCHK(hipHostFree(ctx->frames_staging));
    ctx->frames_staging = nullptr;
    ctx->frames_staging_cap = 0;
    HIPCHK(hipHostMalloc(&ctx->frames_staging, bytes, hipHostMallocDefault));
    ctx->frames_staging_cap = bytes;
  }
  HIPCHK(ctx->frames.ensure(bytes));
  std::memcpy(ctx->frames_staging, fd.data(), bytes);
  ctx->frames_valid = false;
  HIPCHK(hipMemcpyAsync(ctx->frames.p, ctx->frames_staging, bytes, hipMemcpyHostToDevice, s));
  HIPCHK(hipEventRecord(ctx->frames_event, s));
  ctx->frames_event_pending = true;
  ctx->last_frames = fd;
  ctx->frames_valid = true;
  return NCS_OK;
}

static bool any_pad(const ncs_ctx *ctx, const ncs_frame *frames, int n) {
  if (ctx->win != 16) return true;  // pixels outside the window are padding
  for (int i = 0; i < n; i++)
    if (frames[i].geometry != NCS_GEOM_REFERENCE) return true;
  return false;
}

// The schedule of a solve: explicit params win, then the autotuned schedule
// of this mask kind and geometry, then history 5, one wave per workgroup and
// the occupancy limit.
static Schedule resolve_schedule(const ncs_ctx *ctx, const ncs_params *p, bool pad) {
  const bool f16 = p->precision == NCS_PRECISION_F16_SPLIT;
  Schedule d;
  auto it = ctx->tuned.find(TuneKey{ctx->general, pad, f16});
  if (it != ctx->tuned.end()) d = it->second;
  Schedule s;
  s.f16 = f16;
  s.history = p->history > 0 ? p->history : d.history;
  s.wpg = p->waves_per_group > 0 ? p->waves_per_group : d.wpg;
  // a tuned grid only applies to the depth and workgroup size it was tuned for
  const bool same = s.history == d.history && s.wpg == d.wpg;
  s.waves_per_cu = p->waves_per_cu > 0 ? p->waves_per_cu : (same ? d.waves_per_cu : 0);
  if (s.waves_per_cu <= 0) {
    int w = solve_waves_per_cu(s.history, ctx->general, pad, s.f16, s.wpg);
    s.waves_per_cu = w > 0 ? w : 8;
  }
  return s;
}

static SolveArgs make_args(const ncs_ctx *ctx, const FrameDev *d_frames, int n_frames,
                           long long total, const ncs_params *p, const float *u0,
                           float *u_tiles, int *iters, int *status, float *cost,
                           unsigned int *queue) {
  SolveArgs a;
  a.frames = d_frames;
  a.n_frames = n_frames;
  a.total_tiles = total;
  a.queue = queue;
  a.tables = ctx->tables.as<float>();
  a.u0 = u0;
  a.u_tiles = u_tiles;
  a.iterations = iters;
  a.status = status;
  a.cost = cost;
  a.evaluations = nullptr;
  a.max_it = p->max_iterations;
  a.eps = p->epsilon;
  a.ls = p->linesearch;
  a.max_ls = p->max_linesearch;
  a.ftol = p->ftol;
  a.gtol = p->gtol;
  a.xtol = p->xtol;
  a.min_step = p->min_step;
  a.max_step = p->max_step;
  a.stepm = p->stepm;
  a.fallback = p->fallback;
  a.precond = p->precond;
  a.lin_start = p->start == NCS_START_LINEAR && u0 == nullptr;
  a.lin_c = ctx->lin_c;
  a.lin_c1 = ctx->lin_c1;
  a.lin_aref = ctx->lin_aref;
  a.lin_slope = ctx->has_slope ? 1 : 0;
  a.k_norm = ctx->k_norm;
  return a;
}

static hipError_t launch_schedule(const ncs_ctx *ctx, const Schedule &sc, bool pad,
                                  long long total, hipStream_t s, const SolveArgs &a) {
  long long waves = (long long)sc.waves_per_cu * ctx->cus;
  waves = std::max<long long>(1, std::min(waves, total));
  return launch_solve(sc.history, ctx->general, pad, sc.f16, (unsigned int)waves, sc.wpg, s, a);
}

// Enqueue one persistent solve over device frames already described in
// d_frames (device copy of fd).
static int enqueue_solve(ncs_ctx *ctx, const FrameDev *d_frames, int n_frames, bool pad,
                         long long total, const ncs_params *p, const float *u0,
                         float *u_tiles, int *iters, int *status, float *cost,
                         unsigned int *queue, hipStream_t s) {
  if (!ctx->has_mask) return fail(NCS_E_NOMASK, "no OTF mask set");
  if (total <= 0) return NCS_OK;
  SolveArgs a = make_args(ctx, d_frames, n_frames, total, p, u0, u_tiles, iters, status, cost,
                          queue);
  HIPCHK(launch_schedule(ctx, resolve_schedule(ctx, p, pad), pad, total, s, a));
  return NCS_OK;
}

// ---------------------------------------------------------------------------
// device-pointer entry points
// ---------------------------------------------------------------------------

extern "C" int ncs_hip_reduce_frames_device(ncs_ctx *ctx, const ncs_frame *frames,
                                            int n_frames, const ncs_params *params,
                                            const float *u0_tiles, int *iterations,
                                            int *status, float *cost, void *stream) {
  if (!ctx || !frames || n_frames <= 0) return fail(NCS_E_ARG, "bad arguments");
  int rc = check_params(params);
  if (rc) return rc;
  rc = check_window(ctx, params);
  if (rc) return rc;
  if (params->start == NCS_START_TILES && !u0_tiles)
    return fail(NCS_E_ARG, "NCS_START_TILES needs u0_tiles");
  if (params->start == NCS_START_LEARNED && !ctx->has_net)
    return fail(NCS_E_NOWEIGHTS, "learned start requested but no network weights are set");
  if (params->start == NCS_START_LINEAR && !ctx->has_linear)
    return fail(NCS_E_NOWEIGHTS, "linear start requested but no coefficients are set");
  std::vector<FrameDev> fd;
  long long total = 0;
  rc = build_frames(ctx, frames, n_frames, fd, total);
  if (rc) return rc;
  HIPCHK(set_device(ctx));
  hipStream_t s = (hipStream_t)stream;  // 0 = the default (NULL) stream
  HIPCHK(order_after_previous(ctx, s));
  schedule_small_first(fd, total);
  rc = upload_frames(ctx, fd, s);
  if (rc) return rc;
  const float *u0 = params->start == NCS_START_TILES ? u0_tiles : nullptr;
  if (params->start == NCS_START_LEARNED) {
    HIPCHK(ctx->u0_buf.ensure((size_t)total * kTilePix * sizeof(float)));
    const int grid = net_forward_grid(total, ctx->cus);
    HIPCHK(net_warm(ctx->frames.as<FrameDev>(), n_frames, total, ctx->u0_buf.as<float>(),
                    ctx->net_packs.p, ctx->net_params.as<float>(), grid, s));
    u0 = ctx->u0_buf.as<float>();
  }
  rc = enqueue_solve(ctx, ctx->frames.as<FrameDev>(), n_frames, any_pad(ctx, frames, n_frames),
                     total, params, u0, nullptr,
                     iterations, status, cost, ctx->queue.as<unsigned int>(), s);
  if (rc) return rc;
  HIPCHK(mark_solve(ctx, s));
  return NCS_OK;
}

// ncs_hip_reduce_frames_device without allocation: the frame table and the
// work-queue counters live in the device scratch array, the frame table is
// written by kernels from their arguments, and nothing is allocated.  Calls
// that overlap on one GPU race on the scratch, whatever their context.
extern "C" int ncs_hip_reduce_frames_device_noalloc(ncs_ctx *ctx, const ncs_frame *frames,
                                                    int n_frames, const ncs_params *params,
                                                    const float *u0_tiles, int *iterations,
                                                    int *status, float *cost, void *stream) {
  if (!ctx || !frames || n_frames <= 0) return fail(NCS_E_ARG, "bad arguments");
  if (n_frames > kScratchFrames)
    return fail(NCS_E_ARG, "the allocation-free entry point takes at most " +
                               std::to_string(kScratchFrames) + " frames per call (got " +
                               std::to_string(n_frames) + ")");
  int rc = check_params(params);
  if (rc) return rc;
  rc = check_window(ctx, params);
  if (rc) return rc;
  if (params->start == NCS_START_LEARNED)
    return fail(NCS_E_ARG, "the learned start needs the context's start buffer: use "
                           "ncs_hip_reduce_frames_device");
  if (params->start == NCS_START_TILES && !u0_tiles)
    return fail(NCS_E_ARG, "NCS_START_TILES needs u0_tiles");
  if (params->start == NCS_START_LINEAR && !ctx->has_linear)
    return fail(NCS_E_NOWEIGHTS, "linear start requested but no coefficients are set");
  if (!ctx->scratch) return fail(NCS_E_HIP, "no device scratch on this context's GPU");
  FrameDev fd[kScratchFrames];
  int order[kScratchFrames];
  long long total = 0;
  rc = describe_frames(ctx, frames, n_frames, fd, total);
  if (rc) return rc;
  schedule_small_first_fixed(fd, n_frames, total, order);
  HIPCHK(set_device(ctx));
  hipStream_t s = (hipStream_t)stream;
  unsigned int *dq = static_cast<unsigned int *>(ctx->scratch);
  FrameDev *dfr =
      reinterpret_cast<FrameDev *>(static_cast<unsigned char *>(ctx->scratch) + kScratchFrameOffset);
  long long first = 0;  // schedule position of the next frame's first tile
  for (int i0 = 0; i0 < n_frames; i0 += kFramesPerWrite) {
    FrameChunk c;
    c.first = i0;
    c.n = std::min(kFramesPerWrite, n_frames - i0);
    for (int k = 0; k < c.n; k++) {
      const int i = order[i0 + k];
      c.f[k] = fd[i];
      c.f[k].first = first;
      first += frame_tiles(fd, n_frames, total, i);
    }
    hipLaunchKernelGGL(write_frames_kernel, dim3(1), dim3(64), 0, s, c, dfr);
    HIPCHK(hipGetLastError());
  }
  // the counters are reset by a kernel before the solve (launch_solve)
  return enqueue_solve(ctx, dfr, n_frames, any_pad(ctx, frames, n_frames), total, params,
                       params->start == NCS_START_TILES ? u0_tiles : nullptr, nullptr,
                       iterations, status, cost, dq, s);
}

extern "C" int ncs_hip_solve_tiles_device(ncs_ctx *ctx, const float *data_in,
                                          const float *gamma, const float *u0,
                                          float *data_out, int *iterations, int *status,
                                          float *cost, int num_tiles, float alpha,
                                          const ncs_params *params, void *stream) {
  if (!ctx || !data_in || !gamma || num_tiles < 0) return fail(NCS_E_ARG, "bad arguments");
  if (num_tiles == 0) return NCS_OK;
  int rc = check_params(params);
  if (rc) return rc;
  rc = check_window(ctx, params);
  if (rc) return rc;
  const int r = ctx->win;  // tiles of the mask's side
  if ((long long)r * num_tiles >= (1LL << 31) / r) return fail(NCS_E_ARG, "too many tiles");
  ncs_frame f;
  f.data = data_in;
  f.gamma = gamma;
  f.out = data_out;
  f.height = r * num_tiles;
  f.width = r;
  f.data_stride = f.gamma_stride = f.out_stride = r;
  f.alpha = alpha;
  f.geometry = NCS_GEOM_BLOCKS;
  ncs_params p2 = *params;
  if (u0) p2.start = NCS_START_TILES;
  else if (p2.start != NCS_START_LINEAR) p2.start = NCS_START_PLAIN;
  return ncs_hip_reduce_frames_device(ctx, &f, 1, &p2, u0, iterations, status, cost,
                                      stream);
}

// ---------------------------------------------------------------------------
// host-memory entry points
// ---------------------------------------------------------------------------

extern "C" int ncs_hip_reduce_frames(ncs_ctx *ctx, const ncs_frame *frames, int n_frames,
                                     const ncs_params *params, int *iterations,
                                     int *status, float *cost) {
  if (!ctx || !frames || n_frames <= 0) return fail(NCS_E_ARG, "bad arguments");
  int rc = check_params(params);
  if (rc) return rc;
  rc = check_window(ctx, params);
  if (rc) return rc;
  if (params->start != NCS_START_PLAIN && params->start != NCS_START_LINEAR)
    return fail(NCS_E_ARG, "host-memory batches use the plain or the linear start");
  if (params->start == NCS_START_LINEAR && !ctx->has_linear)
    return fail(NCS_E_NOWEIGHTS, "linear start requested but no coefficients are set");
  std::vector<FrameDev> fd;
  long long total = 0;
  rc = build_frames(ctx, frames, n_frames, fd, total);
  if (rc) return rc;
  HIPCHK(set_device(ctx));
  hipStream_t s = ctx->stream;
  HIPCHK(order_after_previous(ctx, s));
  // device layout: each frame's data / out packed [H][W]; gamma deduplicated
  size_t pix = 0, gpix = 0;
  std::vector<size_t> off(n_frames), goff(n_frames);
  std::vector<int> gsrc(n_frames, -1);
  for (int i = 0; i < n_frames; i++) {
    off[i] = pix;
    pix += (size_t)frames[i].height * frames[i].width;
    for (int j = 0; j < i; j++)
      if (frames[j].gamma == frames[i].gamma && frames[j].height == frames[i].height &&
          frames[j].width == frames[i].width &&
          fd[j].gs == fd[i].gs) { gsrc[i] = j; break; }
    if (gsrc[i] < 0) {
      goff[i] = gpix;
      gpix += (size_t)frames[i].height * frames[i].width;
    } else {
      goff[i] = goff[gsrc[i]];
    }
  }
  HIPCHK(ctx->data.ensure(pix * 4));
  HIPCHK(ctx->out.ensure(pix * 4));
  HIPCHK(ctx->gamma.ensure(gpix * 4));
  HIPCHK(ctx->iters.ensure(total * 4));
  HIPCHK(ctx->status.ensure(total * 4));
  HIPCHK(ctx->cost.ensure(total * 4));
  float *dd = ctx->data.as<float>(), *dg = ctx->gamma.as<float>(), *dout = ctx->out.as<float>();
  for (int i = 0; i < n_frames; i++) {
    const ncs_frame &f = frames[i];
    HIPCHK(hipMemcpy2DAsync(dd + off[i], f.width * 4, f.data, fd[i].ds * 4, f.width * 4,
                            f.height, hipMemcpyHostToDevice, s));
    if (gsrc[i] < 0)
      HIPCHK(hipMemcpy2DAsync(dg + goff[i], f.width * 4, f.gamma, fd[i].gs * 4, f.width * 4,
                              f.height, hipMemcpyHostToDevice, s));
    fd[i].data = dd + off[i];
    fd[i].gamma = dg + goff[i];
    fd[i].out = f.out ? dout + off[i] : nullptr;
    fd[i].ds = fd[i].gs = fd[i].os = f.width;
  }
  schedule_small_first(fd, total);
  rc = upload_frames(ctx, fd, s);
  if (rc) return rc;
  rc = enqueue_solve(ctx, ctx->frames.as<FrameDev>(), n_frames, any_pad(ctx, frames, n_frames),
                     total, params, nullptr, nullptr, ctx->iters.as<int>(), ctx->status.as<int>(),
                     ctx->cost.as<float>(), ctx->queue.as<unsigned int>(), s);
  if (rc) {
    (void)hipStreamSynchronize(s);
    return rc;
  }
  for (int i = 0; i < n_frames; i++) {
    const ncs_frame &f = frames[i];
    if (!f.out) continue;
    int os = f.out_stride > 0 ? f.out_stride : f.width;
    HIPCHK(hipMemcpy2DAsync(f.out, os * 4, dout + off[i], f.width * 4, f.width * 4, f.height,
                            hipMemcpyDeviceToHost, s));
  }
  if (iterations)
    HIPCHK(hipMemcpyAsync(iterations, ctx->iters.p, total * 4, hipMemcpyDeviceToHost, s));
  if (status) HIPCHK(hipMemcpyAsync(status, ctx->status.p, total * 4, hipMemcpyDeviceToHost, s));
  if (cost) HIPCHK(hipMemcpyAsync(cost, ctx->cost.p, total * 4, hipMemcpyDeviceToHost, s));
  HIPCHK(mark_solve(ctx, s));
  HIPCHK(hipStreamSynchronize(s));
  return NCS_OK;
}

extern "C" int ncs_hip_solve_tiles(ncs_ctx *ctx, const float *data_in, const float *gamma,
                                   float *data_out, int *iterations, int *status, float *cost,
                                   int num_tiles, float alpha, const ncs_params *params) {
  if (!ctx || !data_in || !gamma || !data_out || num_tiles < 0)
    return fail(NCS_E_ARG, "bad arguments");
  if (num_tiles == 0) return NCS_OK;
  const int r = ctx->win;  // tiles of the mask's side
  if ((long long)r * num_tiles >= (1LL << 31) / r) return fail(NCS_E_ARG, "too many tiles");
  ncs_frame f;
  f.data = data_in;
  f.gamma = gamma;
  f.out = data_out;
  f.height = r * num_tiles;
  f.width = r;
  f.data_stride = f.gamma_stride = f.out_stride = r;
  f.alpha = alpha;
  f.geometry = NCS_GEOM_BLOCKS;
  return ncs_hip_reduce_frames(ctx, &f, 1, params, iterations, status, cost);
}

extern "C" int ncs_hip_eval_tiles(ncs_ctx *ctx, const float *data, const float *gamma,
                                  const float *u, float *cost, float *grad, int num_tiles,
                                  float alpha) {
  if (!ctx || !data || !gamma || !u || num_tiles < 0) return fail(NCS_E_ARG, "bad arguments");
  if (!ctx->has_mask) return fail(NCS_E_NOMASK, "no OTF mask set");
  if (ctx->win != 16) return fail(NCS_E_ARG, "tiles are 16 x 16: set a 16 x 16 mask");
  if (num_tiles == 0) return NCS_OK;
  HIPCHK(set_device(ctx));
  hipStream_t s = ctx->stream;
  size_t n = (size_t)num_tiles * kTilePix;
  DevBuf b;
  HIPCHK(b.ensure(n * 4 * 4 + (size_t)num_tiles * 4));
  float *dd = b.as<float>(), *dg = dd + n, *du = dg + n, *dgr = du + n, *dc = dgr + n;
  hipError_t e = hipMemcpyAsync(dd, data, n * 4, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipMemcpyAsync(dg, gamma, n * 4, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipMemcpyAsync(du, u, n * 4, hipMemcpyHostToDevice, s);
  if (e == hipSuccess)
    e = launch_eval(ctx->general, dd, dg, du, ctx->tables.as<float>(), alpha, num_tiles, dc,
                    dgr, s);
  if (e == hipSuccess && cost)
    e = hipMemcpyAsync(cost, dc, (size_t)num_tiles * 4, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess && grad) e = hipMemcpyAsync(grad, dgr, n * 4, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  b.release();
  HIPCHK(e);
  return NCS_OK;
}

// ---------------------------------------------------------------------------
// learned warm start
// ---------------------------------------------------------------------------

extern "C" int ncs_hip_net_set_weights(ncs_ctx *ctx, const float *params_f32) {
  if (!ctx || !params_f32) return fail(NCS_E_ARG, "bad arguments");
  for (int i = 0; i < NCS_NET_PARAMS; i++)
    if (!std::isfinite(params_f32[i])) return fail(NCS_E_ARG, "non-finite network parameter");
  HIPCHK(set_device(ctx));
  HIPCHK(ctx->net_params.ensure(NCS_NET_PARAMS * sizeof(float)));
  HIPCHK(ctx->net_packs.ensure(net_packs_bytes()));
  HIPCHK(hipMemcpyAsync(ctx->net_params.p, params_f32, NCS_NET_PARAMS * sizeof(float),
                        hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(net_pack(ctx->net_params.as<float>(), ctx->net_packs.p, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  ctx->has_net = true;
  return NCS_OK;
}

// Packs `params` (device) into the context's scratch packs, or returns the
// context's own packs when params is NULL.
static int packs_for(ncs_ctx *ctx, const float *params, hipStream_t s, const void **packs,
                     const float **P) {
  if (params) {
    HIPCHK(ctx->net_tmp_packs.ensure(net_packs_bytes()));
    HIPCHK(net_pack(params, ctx->net_tmp_packs.p, s));
    *packs = ctx->net_tmp_packs.p;
    *P = params;
    return NCS_OK;
  }
  if (!ctx->has_net) return fail(NCS_E_NOWEIGHTS, "no network weights set");
  *packs = ctx->net_packs.p;
  *P = ctx->net_params.as<float>();
  return NCS_OK;
}

extern "C" int ncs_hip_net_predict_frames_device(ncs_ctx *ctx, const ncs_frame *frames,
                                                 int n_frames, float *u0_tiles, void *stream) {
  if (!ctx || !frames || n_frames <= 0 || !u0_tiles) return fail(NCS_E_ARG, "bad arguments");
  if (!ctx->has_net) return fail(NCS_E_NOWEIGHTS, "no network weights set");
  if (ctx->win != 16) return fail(NCS_E_ARG, "the network predicts 16 x 16 tiles");
  std::vector<FrameDev> fd;
  long long total = 0;
  int rc = build_frames(ctx, frames, n_frames, fd, total);
  if (rc) return rc;
  HIPCHK(set_device(ctx));
  hipStream_t s = (hipStream_t)stream;
  schedule_small_first(fd, total);
  rc = upload_frames(ctx, fd, s);
  if (rc) return rc;
  const int grid = net_forward_grid(total, ctx->cus);
  HIPCHK(net_warm(ctx->frames.as<FrameDev>(), n_frames, total, u0_tiles, ctx->net_packs.p,
                  ctx->net_params.as<float>(), grid, s));
  return NCS_OK;
}

extern "C" int ncs_hip_net_forward(ncs_ctx *ctx, const float *params, const float *x, float *out,
                                   void *acts, int n, void *stream) {
  if (!ctx || !x || !out || n < 0) return fail(NCS_E_ARG, "bad arguments");
  HIPCHK(set_device(ctx));
  hipStream_t s = (hipStream_t)stream;  // 0 = the default (NULL) stream
  const void *packs;
  const float *P;
  int rc = packs_for(ctx, params, s, &packs, &P);
  if (rc) return rc;
  const int grid = net_forward_grid(n, ctx->cus);
  HIPCHK(net_forward(x, out, acts, packs, P, n, grid, s));
  return NCS_OK;
}

extern "C" int ncs_hip_net_backward(ncs_ctx *ctx, const float *params, const float *x,
                                    const void *acts, const float *grad_out, float *grad_params,
                                    float *grad_x, int n, void *stream) {
  if (!ctx || !x || !acts || !grad_out || !grad_params || n < 0)
    return fail(NCS_E_ARG, "bad arguments");
  if (grad_x) return fail(NCS_E_ARG, "the input gradient is not provided (pass NULL)");
  HIPCHK(set_device(ctx));
  hipStream_t s = (hipStream_t)stream;  // 0 = the default (NULL) stream
  const void *packs;
  const float *P;
  int rc = packs_for(ctx, params, s, &packs, &P);
  if (rc) return rc;
  int grid = (int)std::max<long long>(1, std::min<long long>(n, (long long)ctx->cus * 4));
  HIPCHK(net_backward(x, acts, grad_out, grad_params, packs, n, grid, s));
  return NCS_OK;
}

// ---------------------------------------------------------------------------
// drop-in for the C library's ncsReduceNoise (clib/pyCNCS/ncs.h:42)
// ---------------------------------------------------------------------------

static std::mutex g_default_mu;
static ncs_ctx *g_default_ctx = nullptr;

extern "C" void ncsReduceNoise(double *ncs_image, double *image, double *gamma,
                               double *otf_mask, double alpha, int im_x, int im_y,
                               int r_size) {
  // the reference rejects odd sizes (ncs.c:54-58); the GPU solves sides up
  // to its 16 x 16 tile
  if (r_size < 4 || r_size > 16 || (r_size & 1)) {
    fail(NCS_E_ARG, "libncs_hip solves even sub-region sizes 4..16");
    return;
  }
  if (!ncs_image || !image || !gamma || !otf_mask || im_x < r_size - 2 || im_y < r_size - 2) {
    fail(NCS_E_ARG, "bad arguments");
    return;
  }
  std::lock_guard<std::mutex> lk(g_default_mu);
  if (!g_default_ctx && ncs_hip_create(0, &g_default_ctx) != NCS_OK) return;
  float m[256];
  for (int i = 0; i < r_size * r_size; i++) m[i] = (float)otf_mask[i];
  if (ncs_hip_set_otf_mask_window(g_default_ctx, m, r_size) != NCS_OK) return;
  size_t n = (size_t)im_x * im_y;
  std::vector<float> d(n), g(n), o(n);
  for (size_t i = 0; i < n; i++) {
    d[i] = (float)image[i];
    g[i] = (float)gamma[i];
  }
  ncs_frame f;
  f.data = d.data();
  f.gamma = g.data();
  f.out = o.data();
  f.height = im_x;
  f.width = im_y;
  f.data_stride = f.gamma_stride = f.out_stride = 0;
  f.alpha = (float)alpha;
  f.geometry = NCS_GEOM_REFERENCE;
  ncs_params p;
  ncs_hip_default_params(&p);
  if (ncs_hip_reduce_frames(g_default_ctx, &f, 1, &p, nullptr, nullptr, nullptr) != NCS_OK)
    return;
  for (size_t i = 0; i < n; i++) ncs_image[i] = o[i];
}

// ---------------------------------------------------------------------------
// autotuner
// ---------------------------------------------------------------------------

extern "C" int ncs_hip_autotune(ncs_ctx *ctx, const ncs_frame *frames, int n_frames,
                                const ncs_params *base, int tune_history,
                                ncs_tune_result *res) {
  if (!ctx || !frames || n_frames <= 0 || !res) return fail(NCS_E_ARG, "bad arguments");
  int rc = check_params(base);
  if (rc) return rc;
  rc = check_window(ctx, base);
  if (rc) return rc;
  if (!ctx->has_mask) return fail(NCS_E_NOMASK, "no OTF mask set");
  if (base->start != NCS_START_PLAIN && base->start != NCS_START_LINEAR)
    return fail(NCS_E_ARG, "the autotuner times the plain or the linear start");
  if (base->start == NCS_START_LINEAR && !ctx->has_linear)
    return fail(NCS_E_NOWEIGHTS, "linear start requested but no coefficients are set");
  std::vector<FrameDev> fd;
  long long total = 0;
  rc = build_frames(ctx, frames, n_frames, fd, total);
  if (rc) return rc;
  if (total <= 0) return fail(NCS_E_ARG, "empty sample");
  for (FrameDev &f : fd) f.out = nullptr;  // time the solve, write nothing
  schedule_small_first(fd, total);  // time the order the solves run in
  const bool pad = any_pad(ctx, frames, n_frames);
  HIPCHK(set_device(ctx));
  hipStream_t s = ctx->stream;
  DevBuf dfr, q;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  auto cleanup = [&]() {
    dfr.release();
    q.release();
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
  };
  hipError_t e = dfr.ensure(fd.size() * sizeof(FrameDev));
  if (e == hipSuccess) e = q.ensure(kQueueBytes);
  if (e == hipSuccess)
    e = hipMemcpy(dfr.p, fd.data(), fd.size() * sizeof(FrameDev), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  if (e != hipSuccess) {
    cleanup();
    return fail(NCS_E_HIP, std::string("autotune setup: ") + hipGetErrorString(e));
  }
  const SolveArgs a = make_args(ctx, dfr.as<FrameDev>(), n_frames, total, base, nullptr,
                                nullptr, nullptr, nullptr, nullptr, q.as<unsigned int>());
  // one untimed run (code load, caches), then the best of three
  auto time_one = [&](const Schedule &sc, float *ms) -> hipError_t {
    hipError_t r = launch_schedule(ctx, sc, pad, total, s, a);
    float best = 3.0e38f;
    for (int rep = 0; rep < 3 && r == hipSuccess; rep++) {
      r = hipEventRecord(e0, s);
      if (r == hipSuccess) r = launch_schedule(ctx, sc, pad, total, s, a);
      if (r == hipSuccess) r = hipEventRecord(e1, s);
      if (r == hipSuccess) r = hipEventSynchronize(e1);
      float t = 0.0f;
      if (r == hipSuccess) r = hipEventElapsedTime(&t, e0, e1);
      best = std::min(best, t);
    }
    *ms = best;
    return r;
  };
  const bool f16 = base->precision == NCS_PRECISION_F16_SPLIT;
  Schedule def;
  def.history = base->history > 0 ? base->history : 5;
  def.wpg = 1;
  def.f16 = f16;
  def.waves_per_cu = std::max(1, solve_waves_per_cu(def.history, ctx->general, pad, f16, 1));
  float ms_def = 0.0f;
  e = time_one(def, &ms_def);
  struct Cand {
    float ms;
    Schedule sc;
    int occ;
  };
  std::vector<Cand> cands = {{ms_def, def, def.waves_per_cu}};
  int tried = 1;
  std::vector<int> depths;
  if (tune_history) depths = {1, 2, 3, 5, 8};
  else depths = {def.history};
  const int groups[] = {1, 2, 4};
  for (int m : depths) {
    for (int wpg : groups) {
      if (e != hipSuccess) break;
      const int occ = solve_waves_per_cu(m, ctx->general, pad, f16, wpg);
      if (occ <= 0) continue;
      std::vector<int> grid = {occ, std::max(wpg, (occ * 3 / 4) / wpg * wpg),
                               std::max(wpg, (occ / 2) / wpg * wpg)};
      std::sort(grid.begin(), grid.end());
      grid.erase(std::unique(grid.begin(), grid.end()), grid.end());
      for (int w : grid) {
        Schedule sc;
        sc.history = m;
        sc.wpg = wpg;
        sc.waves_per_cu = w;
        sc.f16 = f16;
        if (sc.history == def.history && sc.wpg == def.wpg && sc.waves_per_cu == def.waves_per_cu)
          continue;
        float ms = 0.0f;
        e = time_one(sc, &ms);
        if (e != hipSuccess) break;
        tried++;
        cands.push_back({ms, sc, occ});
      }
    }
  }
  // The sweep's leaders are often within a few percent of each other, where
  // one best-of-three can pick the slower: the four fastest are timed again,
  // interleaved over three rounds, and the best of those timings wins.
  std::sort(cands.begin(), cands.end(),
            [](const Cand &x, const Cand &y) { return x.ms < y.ms; });
  if (cands.size() > 4) cands.resize(4);
  for (Cand &cd : cands) cd.ms = 3.0e38f;
  for (int round = 0; round < 3 && e == hipSuccess; round++) {
    for (Cand &cd : cands) {
      float ms = 0.0f;
      e = time_one(cd.sc, &ms);
      if (e != hipSuccess) break;
      cd.ms = std::min(cd.ms, ms);
    }
  }
  const Cand &win = *std::min_element(
      cands.begin(), cands.end(), [](const Cand &x, const Cand &y) { return x.ms < y.ms; });
  Schedule best = win.sc;
  const float ms_best = win.ms;
  const int best_occ = win.occ;
  cleanup();
  HIPCHK(e);
  if (!tune_history) best.history = def.history;
  ctx->tuned[TuneKey{ctx->general, pad, f16}] = best;
  res->history = best.history;
  res->waves_per_cu = best.waves_per_cu;
  res->waves_per_group = best.wpg;
  res->lds_waves_per_cu = best_occ;
  res->ms = ms_best;
  res->ms_default = ms_def;
  res->candidates = tried;
  return NCS_OK;
}

// ---------------------------------------------------------------------------
// captured streaming pipeline
// ---------------------------------------------------------------------------

// Replays of one pipeline in flight at once; each has its own completion
// event and host copy of the per-tile results.
constexpr int kReplaySlots = NCS_PIPELINE_REPLAYS;

struct ncs_pipeline {
  ncs_ctx *ctx = nullptr;
  int n = 0;
  std::vector<ncs_pipeline_frame> shape;
  ncs_params params;
  int input = NCS_INPUT_PHOTONS_F32;
  size_t elem = 4;              // bytes per input pixel
  std::vector<size_t> off;      // first pixel of each frame in the packed buffers
  size_t pix = 0;
  long long total = 0;
  bool pad = false, general = false, keep = false, learned = false;
  int win = 16;  // the context's window side when the pipeline was created
  Schedule sched;
  DevBuf raw, data, gamma, out, offset, gain, frames, queue, results, u0;
  // several gain amplifiers: per frame, the number of levels (1 = one gain
  // map) and the division points as integers; level maps [levels][pix]
  std::vector<int> levels;
  std::vector<std::array<int, NCS_MAX_GAIN_LEVELS - 1>> level_points;
  DevBuf level_gains, level_gammas;
  char *h_res[kReplaySlots] = {};  // per slot: iterations, status, cost
  std::vector<char> gamma_set;
  hipStream_t stream = nullptr;
  hipEvent_t done[kReplaySlots] = {};
  long long launched = 0;  // replays enqueued
  long long waited = 0;    // replays known to be complete
  hipStream_t last_stream = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<const void *> cur_in;
  std::vector<float *> cur_out;
};

extern "C" void ncs_hip_pipeline_destroy(ncs_pipeline *pl) {
  if (!pl) return;
  if (pl->ctx) (void)hipSetDevice(pl->ctx->device);
  if (pl->launched > pl->waited)
    (void)hipEventSynchronize(pl->done[(pl->launched - 1) % kReplaySlots]);
  if (pl->exec) (void)hipGraphExecDestroy(pl->exec);
  DevBuf *bufs[] = {&pl->raw, &pl->data, &pl->gamma, &pl->out, &pl->offset, &pl->gain,
                    &pl->frames, &pl->queue, &pl->results, &pl->u0,
                    &pl->level_gains, &pl->level_gammas};
  for (DevBuf *b : bufs) b->release();
  for (int k = 0; k < kReplaySlots; k++) {
    if (pl->h_res[k]) (void)hipHostFree(pl->h_res[k]);
    if (pl->done[k]) (void)hipEventDestroy(pl->done[k]);
  }
  if (pl->stream) (void)hipStreamDestroy(pl->stream);
  delete pl;
}

static size_t frame_pixels(const ncs_pipeline_frame &f) { return (size_t)f.height * f.width; }

extern "C" int ncs_hip_pipeline_create(ncs_ctx *ctx, const ncs_pipeline_frame *frames,
                                       int n_frames, const ncs_params *params,
                                       int input_format, int keep_tile_results,
                                       ncs_pipeline **out) {
  if (!out) return fail(NCS_E_ARG, "out is NULL");
  *out = nullptr;
  if (!ctx || !frames || n_frames <= 0) return fail(NCS_E_ARG, "bad arguments");
  int rc = check_params(params);
  if (rc) return rc;
  rc = check_window(ctx, params);
  if (rc) return rc;
  if (params->start == NCS_START_TILES)
    return fail(NCS_E_ARG, "a pipeline takes the plain or the learned start");
  if (params->start == NCS_START_LEARNED && !ctx->has_net)
    return fail(NCS_E_NOWEIGHTS, "learned start requested but no network weights are set");
  if (params->start == NCS_START_LINEAR && !ctx->has_linear)
    return fail(NCS_E_NOWEIGHTS, "linear start requested but no coefficients are set");
  if (input_format != NCS_INPUT_PHOTONS_F32 && input_format != NCS_INPUT_ADU_U16)
    return fail(NCS_E_ARG, "unknown input format");
  if (!ctx->has_mask) return fail(NCS_E_NOMASK, "no OTF mask set");
  std::unique_ptr<ncs_pipeline, void (*)(ncs_pipeline *)> pl(new ncs_pipeline(),
                                                             ncs_hip_pipeline_destroy);
  pl->ctx = ctx;
  pl->n = n_frames;
  pl->shape.assign(frames, frames + n_frames);
  pl->params = *params;
  pl->input = input_format;
  pl->elem = input_format == NCS_INPUT_ADU_U16 ? 2 : 4;
  pl->keep = keep_tile_results != 0;
  pl->learned = params->start == NCS_START_LEARNED;
  std::vector<FrameDev> fd(n_frames);
  pl->off.resize(n_frames);
  for (int i = 0; i < n_frames; i++) {
    const ncs_pipeline_frame &f = frames[i];
    long long nt = ncs_hip_tile_count_window(f.height, f.width, f.geometry, ctx->win);
    if (nt < 0) return fail(NCS_E_ARG, "frame " + std::to_string(i) + ": bad size or geometry");
    if (!std::isfinite(f.alpha)) return fail(NCS_E_ARG, "non-finite alpha");
    if ((long long)f.height * f.width >= (1LL << 31))
      return fail(NCS_E_ARG, "frame " + std::to_string(i) + " spans 2^31 or more elements");
    pl->off[i] = pl->pix;
    pl->pix += (frame_pixels(f) + 63) / 64 * 64;  // 256-byte aligned frames
    FrameDev &d = fd[i];
    d.H = f.height;
    d.W = f.width;
    d.ds = d.gs = d.os = f.width;
    d.alpha = f.alpha;
    d.geom = f.geometry;
    d.ty = (int)tiles_axis(f.width, f.geometry, ctx->win);
    d.win = ctx->win;
    d.first = pl->total;
    d.slot = pl->total;
    pl->total += nt;
    if (f.geometry != NCS_GEOM_REFERENCE || ctx->win != 16) pl->pad = true;
  }
  if (pl->total > 0xF0000000LL) return fail(NCS_E_ARG, "too many tiles in one batch");
  HIPCHK(set_device(ctx));
  const size_t fb = pl->pix * sizeof(float);
  HIPCHK(pl->data.ensure(fb));
  HIPCHK(pl->gamma.ensure(fb));
  HIPCHK(pl->out.ensure(fb));
  if (pl->input == NCS_INPUT_ADU_U16) {
    HIPCHK(pl->raw.ensure(pl->pix * 2));
    HIPCHK(pl->offset.ensure(fb));
    HIPCHK(pl->gain.ensure(fb));
  }
  HIPCHK(pl->queue.ensure(kQueueBytes));
  // iterations, status and cost back to back: one download per replay
  HIPCHK(pl->results.ensure((size_t)pl->total * 12));
  if (pl->learned) HIPCHK(pl->u0.ensure((size_t)pl->total * kTilePix * sizeof(float)));
  if (pl->keep)
    for (int k = 0; k < kReplaySlots; k++)
      HIPCHK(hipHostMalloc((void **)&pl->h_res[k], (size_t)pl->total * 12, hipHostMallocDefault));
  for (int i = 0; i < n_frames; i++) {
    fd[i].data = pl->data.as<float>() + pl->off[i];
    fd[i].gamma = pl->gamma.as<float>() + pl->off[i];
    fd[i].out = pl->out.as<float>() + pl->off[i];
  }
  schedule_small_first(fd, pl->total);
  HIPCHK(pl->frames.ensure(fd.size() * sizeof(FrameDev)));
  HIPCHK(hipMemcpy(pl->frames.p, fd.data(), fd.size() * sizeof(FrameDev), hipMemcpyHostToDevice));
  HIPCHK(hipStreamCreateWithFlags(&pl->stream, hipStreamNonBlocking));
  for (int k = 0; k < kReplaySlots; k++)
    HIPCHK(hipEventCreateWithFlags(&pl->done[k], hipEventDisableTiming));
  pl->gamma_set.assign(n_frames, 0);
  pl->levels.assign(n_frames, 1);
  pl->level_points.resize(n_frames);
  pl->sched = resolve_schedule(ctx, params, pl->pad);
  pl->win = ctx->win;
  *out = pl.release();
  return NCS_OK;
}

// Waits on the host for replay `seq` (launch order).  Replays of one
// pipeline run in launch order, so every earlier one is complete too.
static int wait_replay(ncs_pipeline *pl, long long seq) {
  if (seq < pl->waited) return NCS_OK;
  HIPCHK(hipEventSynchronize(pl->done[seq % kReplaySlots]));
  pl->waited = seq + 1;
  return NCS_OK;
}

static int pipeline_idle(ncs_pipeline *pl) {
  return pl->launched ? wait_replay(pl, pl->launched - 1) : NCS_OK;
}

// The next launch captures the graph again (the pipeline is idle).
static void drop_graph(ncs_pipeline *pl) {
  if (pl->exec) (void)hipGraphExecDestroy(pl->exec);
  pl->exec = nullptr;
}

extern "C" int ncs_hip_pipeline_set_gamma(ncs_pipeline *pl, int frame, const float *gamma) {
  if (!pl || !gamma || frame < 0 || frame >= pl->n) return fail(NCS_E_ARG, "bad arguments");
  if (pl->levels[frame] > 1)
    return fail(NCS_E_ARG, "gamma of frame " + std::to_string(frame) +
                               " comes from its gain levels at every replay");
  HIPCHK(set_device(pl->ctx));
  int rc = pipeline_idle(pl);
  if (rc) return rc;
  HIPCHK(hipMemcpy(pl->gamma.as<float>() + pl->off[frame], gamma,
                   frame_pixels(pl->shape[frame]) * sizeof(float), hipMemcpyHostToDevice));
  pl->gamma_set[frame] = 1;
  return NCS_OK;
}

extern "C" int ncs_hip_pipeline_set_camera(ncs_pipeline *pl, int frame, const float *offset,
                                           const float *gain, const float *variance) {
  if (!pl || !offset || !gain || !variance || frame < 0 || frame >= pl->n)
    return fail(NCS_E_ARG, "bad arguments");
  if (pl->input != NCS_INPUT_ADU_U16)
    return fail(NCS_E_ARG, "camera maps belong to NCS_INPUT_ADU_U16 pipelines");
  HIPCHK(set_device(pl->ctx));
  int rc = pipeline_idle(pl);
  if (rc) return rc;
  const size_t n = frame_pixels(pl->shape[frame]);
  const size_t o = pl->off[frame];
  HIPCHK(hipMemcpy(pl->offset.as<float>() + o, offset, n * 4, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(pl->gain.as<float>() + o, gain, n * 4, hipMemcpyHostToDevice));
  // the variance goes through the output buffer, which every replay rewrites
  HIPCHK(hipMemcpy(pl->out.as<float>() + o, variance, n * 4, hipMemcpyHostToDevice));
  HIPCHK(launch_gamma_from_camera(pl->out.as<float>() + o, pl->gain.as<float>() + o,
                                  pl->gamma.as<float>() + o, (long long)n, pl->stream));
  HIPCHK(hipStreamSynchronize(pl->stream));
  pl->gamma_set[frame] = 1;
  if (pl->levels[frame] != 1) {
    pl->levels[frame] = 1;
    drop_graph(pl);
  }
  return NCS_OK;
}

extern "C" int ncs_hip_pipeline_set_camera_levels(ncs_pipeline *pl, int frame, int n_levels,
                                                  const double *points, const float *offset,
                                                  const float *gains, const float *variances) {
  if (!pl || !offset || !gains || !variances || frame < 0 || frame >= pl->n)
    return fail(NCS_E_ARG, "bad arguments");
  if (pl->input != NCS_INPUT_ADU_U16)
    return fail(NCS_E_ARG, "camera maps belong to NCS_INPUT_ADU_U16 pipelines");
  if (n_levels < 1 || n_levels > NCS_MAX_GAIN_LEVELS)
    return fail(NCS_E_ARG, "n_levels must be 1.." + std::to_string(NCS_MAX_GAIN_LEVELS));
  if (n_levels > 1 && !points) return fail(NCS_E_ARG, "division points are NULL");
  // ADU is an integer, so ADU >= p exactly when ADU >= ceil(p)
  std::array<int, NCS_MAX_GAIN_LEVELS - 1> t;
  t.fill(65536);
  for (int j = 0; j + 1 < n_levels; j++) {
    const double p = points[j];
    if (std::isnan(p)) return fail(NCS_E_ARG, "a division point is NaN");
    if (j > 0 && p < points[j - 1]) return fail(NCS_E_ARG, "division points must ascend");
    t[j] = (int)std::min(65536.0, std::max(0.0, std::ceil(p)));
  }
  HIPCHK(set_device(pl->ctx));
  int rc = pipeline_idle(pl);
  if (rc) return rc;
  if (n_levels == 1)
    return ncs_hip_pipeline_set_camera(pl, frame, offset, gains, variances);
  const size_t n = frame_pixels(pl->shape[frame]);
  const size_t o = pl->off[frame];
  const size_t lb = (size_t)NCS_MAX_GAIN_LEVELS * pl->pix * sizeof(float);
  HIPCHK(pl->level_gains.ensure(lb));
  HIPCHK(pl->level_gammas.ensure(lb));
  HIPCHK(hipMemcpy(pl->offset.as<float>() + o, offset, n * 4, hipMemcpyHostToDevice));
  for (int l = 0; l < n_levels; l++) {
    float *g = pl->level_gains.as<float>() + (size_t)l * pl->pix + o;
    HIPCHK(hipMemcpy(g, gains + (size_t)l * n, n * 4, hipMemcpyHostToDevice));
    // the variance goes through the output buffer, which every replay rewrites
    HIPCHK(hipMemcpy(pl->out.as<float>() + o, variances + (size_t)l * n, n * 4,
                     hipMemcpyHostToDevice));
    HIPCHK(launch_gamma_from_camera(pl->out.as<float>() + o, g,
                                    pl->level_gammas.as<float>() + (size_t)l * pl->pix + o,
                                    (long long)n, pl->stream));
    HIPCHK(hipStreamSynchronize(pl->stream));
  }
  pl->gamma_set[frame] = 1;
  if (pl->levels[frame] != n_levels || pl->level_points[frame] != t) {
    pl->levels[frame] = n_levels;
    pl->level_points[frame] = t;
    drop_graph(pl);  // the division points are arguments of the captured kernel
  }
  return NCS_OK;
}

// The device part of a replay, captured once on the pipeline's stream as a
// single chain: camera conversion, learned start, queue reset and solve.
// The host transfers stay outside the graph: a copy node of a graph runs as
// a blit kernel on the compute units, where it holds wave slots the
// persistent solver's statically assigned waves need (a 2304x2304 solve
// overlapping one took 5x its time alone) and moves data one PCIe direction
// at a time, while a stream copy runs on the DMA engines, uploads and
// downloads of different pipelines at once.  With the host pointers out of
// the graph, new buffers need no capture or update.
static int capture_replay(ncs_pipeline *pl, hipGraph_t *out) {
  ncs_ctx *ctx = pl->ctx;
  hipStream_t s = pl->stream;
  HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed));
  hipError_t e = hipSuccess;
  if (pl->input == NCS_INPUT_ADU_U16) {
    for (int i = 0; i < pl->n && e == hipSuccess; i++) {
      const size_t o = pl->off[i];
      const long long n = (long long)frame_pixels(pl->shape[i]);
      if (pl->levels[i] > 1)
        e = launch_adu_levels(pl->raw.as<uint16_t>() + o, pl->offset.as<float>() + o,
                              pl->level_gains.as<float>() + o, pl->level_gammas.as<float>() + o,
                              (long long)pl->pix, pl->level_points[i].data(),
                              pl->data.as<float>() + o, pl->gamma.as<float>() + o, n, s);
      else
        e = launch_adu_to_photons(pl->raw.as<uint16_t>() + o, pl->offset.as<float>() + o,
                                  pl->gain.as<float>() + o, pl->data.as<float>() + o, n, s);
    }
  }
  const float *u0 = nullptr;
  if (e == hipSuccess && pl->learned) {
    const int grid = net_forward_grid(pl->total, ctx->cus);
    e = net_warm(pl->frames.as<FrameDev>(), pl->n, pl->total, pl->u0.as<float>(),
                 ctx->net_packs.p, ctx->net_params.as<float>(), grid, s);
    u0 = pl->u0.as<float>();
  }
  if (e == hipSuccess) {
    const SolveArgs a = make_args(ctx, pl->frames.as<FrameDev>(), pl->n, pl->total, &pl->params,
                                  u0, nullptr, pl->results.as<int>(),
                                  pl->results.as<int>() + pl->total,
                                  pl->results.as<float>() + 2 * pl->total,
                                  pl->queue.as<unsigned int>());
    e = launch_schedule(ctx, pl->sched, pl->pad, pl->total, s, a);
  }
  hipGraph_t g = nullptr;
  hipError_t e2 = hipStreamEndCapture(s, &g);
  if (e != hipSuccess || e2 != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    HIPCHK(e);
    HIPCHK(e2);
  }
  *out = g;
  return NCS_OK;
}

static int build_pipeline_graph(ncs_pipeline *pl) {
  hipGraph_t graph = nullptr;
  int rc = capture_replay(pl, &graph);
  if (rc) return rc;
  hipError_t e = hipGraphInstantiate(&pl->exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e != hipSuccess) pl->exec = nullptr;
  HIPCHK(e);
  pl->general = pl->ctx->general;
  return NCS_OK;
}

// Enqueues one replay on s: upload of every frame, the graph, download of
// every frame and of the per-tile results into the replay's slot.
static hipError_t enqueue_replay(ncs_pipeline *pl, hipStream_t s, int slot) {
  char *dev_in = pl->input == NCS_INPUT_ADU_U16 ? pl->raw.as<char>() : pl->data.as<char>();
  hipError_t e = hipSuccess;
  for (int i = 0; i < pl->n && e == hipSuccess; i++)
    e = hipMemcpyAsync(dev_in + pl->off[i] * pl->elem, pl->cur_in[i],
                       frame_pixels(pl->shape[i]) * pl->elem, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipGraphLaunch(pl->exec, s);
  for (int i = 0; i < pl->n && e == hipSuccess; i++)
    e = hipMemcpyAsync(pl->cur_out[i], pl->out.as<float>() + pl->off[i],
                       frame_pixels(pl->shape[i]) * sizeof(float), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess && pl->keep) {
    e = hipMemcpyAsync(pl->h_res[slot], pl->results.p, (size_t)pl->total * 12,
                       hipMemcpyDeviceToHost, s);
  }
  return e;
}

extern "C" int ncs_hip_pipeline_launch(ncs_pipeline *pl, const void *const *inputs,
                                       float *const *outputs, void *stream) {
  if (!pl) return fail(NCS_E_ARG, "bad arguments");
  if (pl->cur_in.empty() && (!inputs || !outputs))
    return fail(NCS_E_ARG, "the first launch needs inputs and outputs");
  for (int i = 0; i < pl->n; i++) {
    if ((inputs && !inputs[i]) || (outputs && !outputs[i]))
      return fail(NCS_E_ARG, "NULL frame buffer");
    if (!pl->gamma_set[i])
      return fail(NCS_E_ARG, "gamma of frame " + std::to_string(i) + " is not set");
  }
  ncs_ctx *ctx = pl->ctx;
  if (!ctx->has_mask) return fail(NCS_E_NOMASK, "no OTF mask set");
  if (pl->learned && !ctx->has_net) return fail(NCS_E_NOWEIGHTS, "no network weights set");
  HIPCHK(set_device(ctx));
  if (pl->exec && ctx->general != pl->general)
    return fail(NCS_E_MASK, "the kind of the OTF mask changed since the pipeline was "
                "captured; create a new pipeline");
  if (ctx->win != pl->win)
    return fail(NCS_E_MASK, "the mask size changed since the pipeline was created; create a "
                "new pipeline");
  int rc;
  if (!pl->exec) {
    rc = build_pipeline_graph(pl);
    if (rc) return rc;
  }
  const long long seq = pl->launched;
  const int slot = (int)(seq % kReplaySlots);
  // the slot's last replay must be complete before its event and results
  // are reused
  if (seq >= kReplaySlots) {
    rc = wait_replay(pl, seq - kReplaySlots);
    if (rc) return rc;
  }
  if (inputs) pl->cur_in.assign(inputs, inputs + pl->n);
  if (outputs) pl->cur_out.assign(outputs, outputs + pl->n);
  hipStream_t s = stream ? (hipStream_t)stream : pl->stream;
  hipError_t e = hipSuccess;
  // the device buffers are shared by every replay: one enqueued on another
  // stream waits on the device for the previous one
  if (seq > pl->waited && s != pl->last_stream)
    e = hipStreamWaitEvent(s, pl->done[(seq - 1) % kReplaySlots], 0);
  if (e == hipSuccess) e = enqueue_replay(pl, s, slot);
  // whatever was enqueued is covered by the replay's event
  hipError_t e2 = hipEventRecord(pl->done[slot], s);
  if (e2 == hipSuccess) {
    pl->launched = seq + 1;
    pl->last_stream = s;
  }
  HIPCHK(e);
  HIPCHK(e2);
  return NCS_OK;
}

extern "C" int ncs_hip_pipeline_wait(ncs_pipeline *pl) {
  if (!pl) return fail(NCS_E_ARG, "bad arguments");
  HIPCHK(set_device(pl->ctx));
  return pipeline_idle(pl);
}

extern "C" int ncs_hip_pipeline_wait_replay(ncs_pipeline *pl, long long replay) {
  if (!pl || replay < 0 || replay >= pl->launched)
    return fail(NCS_E_ARG, "bad arguments or a replay not launched");
  HIPCHK(set_device(pl->ctx));
  return wait_replay(pl, replay);
}

extern "C" int ncs_hip_pipeline_replay_results(ncs_pipeline *pl, long long replay,
                                               int *iterations, int *status, float *cost) {
  if (!pl) return fail(NCS_E_ARG, "bad arguments");
  if (!pl->keep) return fail(NCS_E_ARG, "the pipeline was created without tile results");
  if (replay < 0 || replay >= pl->launched || replay < pl->launched - kReplaySlots)
    return fail(NCS_E_ARG, "replay " + std::to_string(replay) + " is not among the last " +
                               std::to_string(kReplaySlots) + " launched");
  HIPCHK(set_device(pl->ctx));
  int rc = wait_replay(pl, replay);
  if (rc) return rc;
  const size_t b = (size_t)pl->total * 4;
  const char *h = pl->h_res[replay % kReplaySlots];
  if (iterations) std::memcpy(iterations, h, b);
  if (status) std::memcpy(status, h + b, b);
  if (cost) std::memcpy(cost, h + 2 * b, b);
  return NCS_OK;
}

extern "C" int ncs_hip_pipeline_tile_results(ncs_pipeline *pl, int *iterations, int *status,
                                             float *cost) {
  if (!pl) return fail(NCS_E_ARG, "bad arguments");
  if (!pl->launched) return fail(NCS_E_ARG, "no replay launched");
  return ncs_hip_pipeline_replay_results(pl, pl->launched - 1, iterations, status, cost);
}

extern "C" long long ncs_hip_pipeline_launched(const ncs_pipeline *pl) {
  return pl ? pl->launched : -1;
}

extern "C" long long ncs_hip_pipeline_tiles(const ncs_pipeline *pl) {
  return pl ? pl->total : -1;
}
