// ncs_capi.cpp -- C ABI of libncs_hip.so: context, OTF tables, launches,
// host/device frame batches, the C-library drop-in, graphs and autotuning.
#include <algorithm>
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
int solve_waves_per_cu(int m, bool general, bool pad);
hipError_t launch_solve(int m, bool general, bool pad, unsigned int waves, hipStream_t s,
                        const SolveArgs &a);
hipError_t net_pack(const float *P, void *packs, hipStream_t s);
size_t net_packs_bytes();
hipError_t net_forward(const float *x, float *y, void *acts, const void *packs, const float *P,
                       long long n, int grid, hipStream_t s);
hipError_t net_backward(const float *x, const void *acts, const float *gout, float *gP,
                        const void *packs, long long n, int grid, hipStream_t s);
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

struct GraphRec;

struct ncs_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  int cus = 0;
  char arch[64] = {0};
  DevBuf tables, frames, queue, data, gamma, out, iters, status, cost, tiles_u;
  bool has_mask = false;
  bool general = false;
  float mask_fft[256];
  std::map<int, int> tuned_waves;  // history -> waves per CU
  std::map<int, std::unique_ptr<GraphRec>> graphs;
  int next_graph = 1;
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
};

struct GraphRec {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<FrameDev> fdev;
  float *h_data = nullptr, *h_gamma = nullptr, *h_out = nullptr, *h_cost = nullptr;
  int *h_iters = nullptr, *h_status = nullptr;
  DevBuf d_data, d_gamma, d_out, d_frames, d_queue, d_iters, d_status, d_cost;
  size_t n_pix = 0, n_gamma = 0;
  long long n_tiles = 0;
};

static hipError_t set_device(ncs_ctx *ctx) { return hipSetDevice(ctx->device); }

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
  p->precond = 1;
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
  HIPCHK(c->queue.ensure(256));
  *out = c.release();
  return NCS_OK;
}

static void destroy_graph(GraphRec *g) {
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  if (g->h_data) (void)hipHostFree(g->h_data);
  if (g->h_gamma) (void)hipHostFree(g->h_gamma);
  if (g->h_out) (void)hipHostFree(g->h_out);
  if (g->h_iters) (void)hipHostFree(g->h_iters);
  if (g->h_status) (void)hipHostFree(g->h_status);
  if (g->h_cost) (void)hipHostFree(g->h_cost);
  g->d_data.release(); g->d_gamma.release(); g->d_out.release();
  g->d_frames.release(); g->d_queue.release(); g->d_iters.release();
  g->d_status.release(); g->d_cost.release();
}

extern "C" void ncs_hip_destroy(ncs_ctx *ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  for (auto &kv : ctx->graphs) destroy_graph(kv.second.get());
  ctx->graphs.clear();
  DevBuf *bufs[] = {&ctx->tables, &ctx->frames, &ctx->queue, &ctx->data, &ctx->gamma,
                    &ctx->out, &ctx->iters, &ctx->status, &ctx->cost, &ctx->tiles_u,
                    &ctx->net_params, &ctx->net_packs, &ctx->net_tmp_packs, &ctx->u0_buf};
  for (DevBuf *b : bufs) b->release();
  if (ctx->frames_event) (void)hipEventDestroy(ctx->frames_event);
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
static double rbasis(int a, int i) {
  const double PI = 3.14159265358979323846;
  if (a == 0) return 0.25;
  if (a == 8) return (i & 1) ? -0.25 : 0.25;
  if (a < 8) return std::sqrt(2.0) / 4.0 * std::cos(2.0 * PI * a * i / 16.0);
  return std::sqrt(2.0) / 4.0 * std::sin(2.0 * PI * (a - 8) * i / 16.0);
}

static int freq_of(int a) { return a <= 8 ? a : a - 8; }

extern "C" int ncs_hip_set_otf_mask(ncs_ctx *ctx, const float *mask) {
  if (!ctx || !mask) return fail(NCS_E_ARG, "bad arguments");
  double q[16][16], qcs[16][16];
  double qmax = 0.0;
  for (int k = 0; k < 16; k++)
    for (int l = 0; l < 16; l++) {
      double m = mask[k * 16 + l];
      if (!std::isfinite(m)) return fail(NCS_E_MASK, "mask has non-finite values");
      q[k][l] = m * m;
    }
  for (int k = 0; k < 16; k++)
    for (int l = 0; l < 16; l++) {
      qcs[k][l] = 0.5 * (q[k][l] + q[(16 - k) & 15][(16 - l) & 15]);
      qmax = std::max(qmax, std::fabs(qcs[k][l]));
    }
  std::vector<float> tab(kTableFloats);
  double qo_max = 0.0;
  for (int lane = 0; lane < 64; lane++) {
    int g = lane >> 4, x = lane & 15;
    for (int r = 0; r < 4; r++) {
      int b = 4 * g + r, a = x;
      tab[lane * 16 + r] = (float)rbasis(x, 4 * g + r);
      tab[lane * 16 + 4 + r] = (float)rbasis(4 * g + r, x);
      int k = freq_of(a), l = freq_of(b);
      double qee = 0.5 * (qcs[k][l] + qcs[(16 - k) & 15][l]);
      double qoo = 0.5 * (qcs[k][l] - qcs[(16 - k) & 15][l]);
      double sgn = ((a <= 8) == (b <= 8)) ? -1.0 : 1.0;
      tab[lane * 16 + 8 + r] = (float)(2.0 * qee);
      tab[lane * 16 + 12 + r] = (float)(sgn * 2.0 * qoo);
      qo_max = std::max(qo_max, std::fabs(qoo));
    }
  }
  HIPCHK(set_device(ctx));
  HIPCHK(ctx->tables.ensure(tab.size() * sizeof(float)));
  HIPCHK(hipMemcpy(ctx->tables.p, tab.data(), tab.size() * sizeof(float),
                   hipMemcpyHostToDevice));
  ctx->general = qo_max > 1e-12 * std::max(qmax, 1e-300);
  std::memcpy(ctx->mask_fft, mask, sizeof(ctx->mask_fft));
  ctx->has_mask = true;
  return NCS_OK;
}

// ---------------------------------------------------------------------------
// tiling
// ---------------------------------------------------------------------------

static long long tiles_axis(int n, int geom) {
  if (geom == NCS_GEOM_REFERENCE) {
    if (n < 14) return -1;
    return 1 + (n - 14 + 13) / 14;  // 1 + ceil((n - 14) / 14), ncs.c:71-143
  }
  if (n < 1) return -1;
  return (n + 15) / 16;
}

extern "C" long long ncs_hip_tile_count(int height, int width, int geometry) {
  if (geometry != NCS_GEOM_REFERENCE && geometry != NCS_GEOM_BLOCKS) return -1;
  long long a = tiles_axis(height, geometry), b = tiles_axis(width, geometry);
  if (a < 0 || b < 0) return -1;
  return a * b;
}

static int check_params(const ncs_params *p) {
  if (!p) return fail(NCS_E_ARG, "params is NULL");
  if (p->history < 1 || p->history > 8) return fail(NCS_E_ARG, "history must be 1..8");
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
  if (p->precond != 0 && p->precond != 1) return fail(NCS_E_ARG, "precond must be 0 or 1");
  return NCS_OK;
}

// Builds kernel frame descriptors (device pointers) and the tile total.
static int build_frames(const ncs_frame *frames, int n, std::vector<FrameDev> &fd,
                        long long &total) {
  fd.resize(n);
  total = 0;
  for (int i = 0; i < n; i++) {
    const ncs_frame &f = frames[i];
    if (!f.data || !f.gamma) return fail(NCS_E_ARG, "frame without data or gamma");
    long long nt = ncs_hip_tile_count(f.height, f.width, f.geometry);
    if (nt < 0)
      return fail(NCS_E_ARG, "frame " + std::to_string(i) + ": bad size or geometry (the "
                  "reference tiling needs height, width >= 14)");
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
    d.alpha = f.alpha;
    d.geom = f.geometry;
    d.ty = (int)tiles_axis(f.width, f.geometry);
    d.pad = 0;
    d.first = total;
    total += nt;
  }
  if (total > 0xF0000000LL) return fail(NCS_E_ARG, "too many tiles in one batch");
  return NCS_OK;
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

static bool any_pad(const ncs_frame *frames, int n) {
  for (int i = 0; i < n; i++)
    if (frames[i].geometry != NCS_GEOM_REFERENCE) return true;
  return false;
}

static int waves_for(ncs_ctx *ctx, const ncs_params *p, bool pad) {
  if (p->waves_per_cu > 0) return p->waves_per_cu;
  auto it = ctx->tuned_waves.find(p->history);
  if (it != ctx->tuned_waves.end()) return it->second;
  int w = solve_waves_per_cu(p->history, ctx->general, pad);
  return w > 0 ? w : 8;
}

// Enqueue one persistent solve over device frames already described in
// d_frames (device copy of fd).
static int enqueue_solve(ncs_ctx *ctx, const FrameDev *d_frames, int n_frames, bool pad,
                         long long total, const ncs_params *p, const float *u0,
                         float *u_tiles, int *iters, int *status, float *cost,
                         unsigned int *queue, hipStream_t s) {
  if (!ctx->has_mask) return fail(NCS_E_NOMASK, "no OTF mask set");
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
  if (total <= 0) return NCS_OK;
  long long waves = (long long)waves_for(ctx, p, pad) * ctx->cus;
  waves = std::max<long long>(1, std::min(waves, total));
  HIPCHK(launch_solve(p->history, ctx->general, pad, (unsigned int)waves, s, a));
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
  if (params->start == NCS_START_TILES && !u0_tiles)
    return fail(NCS_E_ARG, "NCS_START_TILES needs u0_tiles");
  if (params->start == NCS_START_LEARNED && !ctx->has_net)
    return fail(NCS_E_NOWEIGHTS, "learned start requested but no network weights are set");
  std::vector<FrameDev> fd;
  long long total = 0;
  rc = build_frames(frames, n_frames, fd, total);
  if (rc) return rc;
  HIPCHK(set_device(ctx));
  hipStream_t s = (hipStream_t)stream;  // 0 = the default (NULL) stream
  rc = upload_frames(ctx, fd, s);
  if (rc) return rc;
  const float *u0 = params->start == NCS_START_TILES ? u0_tiles : nullptr;
  if (params->start == NCS_START_LEARNED) {
    HIPCHK(ctx->u0_buf.ensure((size_t)total * kTilePix * sizeof(float)));
    int grid = (int)std::min<long long>(total, (long long)ctx->cus * 4);
    HIPCHK(net_warm(ctx->frames.as<FrameDev>(), n_frames, total, ctx->u0_buf.as<float>(),
                    ctx->net_packs.p, ctx->net_params.as<float>(), grid, s));
    u0 = ctx->u0_buf.as<float>();
  }
  return enqueue_solve(ctx, ctx->frames.as<FrameDev>(), n_frames, any_pad(frames, n_frames),
                       total, params, u0, nullptr,
                       iterations, status, cost, ctx->queue.as<unsigned int>(), s);
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
  ncs_frame f;
  f.data = data_in;
  f.gamma = gamma;
  f.out = data_out;
  f.height = 16 * num_tiles;
  f.width = 16;
  f.data_stride = f.gamma_stride = f.out_stride = 16;
  f.alpha = alpha;
  f.geometry = NCS_GEOM_BLOCKS;
  ncs_params p2 = *params;
  p2.start = u0 ? NCS_START_TILES : NCS_START_PLAIN;
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
  if (params->start != NCS_START_PLAIN)
    return fail(NCS_E_ARG, "host-memory batches use the plain start");
  std::vector<FrameDev> fd;
  long long total = 0;
  rc = build_frames(frames, n_frames, fd, total);
  if (rc) return rc;
  HIPCHK(set_device(ctx));
  hipStream_t s = ctx->stream;
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
  rc = upload_frames(ctx, fd, s);
  if (rc) return rc;
  rc = enqueue_solve(ctx, ctx->frames.as<FrameDev>(), n_frames, any_pad(frames, n_frames),
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
  HIPCHK(hipStreamSynchronize(s));
  return NCS_OK;
}

extern "C" int ncs_hip_solve_tiles(ncs_ctx *ctx, const float *data_in, const float *gamma,
                                   float *data_out, int *iterations, int *status, float *cost,
                                   int num_tiles, float alpha, const ncs_params *params) {
  if (!ctx || !data_in || !gamma || !data_out || num_tiles < 0)
    return fail(NCS_E_ARG, "bad arguments");
  if (num_tiles == 0) return NCS_OK;
  ncs_frame f;
  f.data = data_in;
  f.gamma = gamma;
  f.out = data_out;
  f.height = 16 * num_tiles;
  f.width = 16;
  f.data_stride = f.gamma_stride = f.out_stride = 16;
  f.alpha = alpha;
  f.geometry = NCS_GEOM_BLOCKS;
  return ncs_hip_reduce_frames(ctx, &f, 1, params, iterations, status, cost);
}

extern "C" int ncs_hip_eval_tiles(ncs_ctx *ctx, const float *data, const float *gamma,
                                  const float *u, float *cost, float *grad, int num_tiles,
                                  float alpha) {
  if (!ctx || !data || !gamma || !u || num_tiles < 0) return fail(NCS_E_ARG, "bad arguments");
  if (!ctx->has_mask) return fail(NCS_E_NOMASK, "no OTF mask set");
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
  std::vector<FrameDev> fd;
  long long total = 0;
  int rc = build_frames(frames, n_frames, fd, total);
  if (rc) return rc;
  HIPCHK(set_device(ctx));
  hipStream_t s = (hipStream_t)stream;
  rc = upload_frames(ctx, fd, s);
  if (rc) return rc;
  int grid = (int)std::max<long long>(1, std::min<long long>(total, (long long)ctx->cus * 4));
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
  int grid = (int)std::max<long long>(1, std::min<long long>(n, (long long)ctx->cus * 4));
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
  if (r_size != 16) {
    fail(NCS_E_ARG, "libncs_hip solves 16 x 16 sub-regions only");
    return;
  }
  if (!ncs_image || !image || !gamma || !otf_mask || im_x < 14 || im_y < 14) {
    fail(NCS_E_ARG, "bad arguments");
    return;
  }
  std::lock_guard<std::mutex> lk(g_default_mu);
  if (!g_default_ctx && ncs_hip_create(0, &g_default_ctx) != NCS_OK) return;
  float m[256];
  for (int i = 0; i < 256; i++) m[i] = (float)otf_mask[i];
  if (ncs_hip_set_otf_mask(g_default_ctx, m) != NCS_OK) return;
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
