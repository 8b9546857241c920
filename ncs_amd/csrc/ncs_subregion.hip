// ncs_subregion.hip -- the per-sub-region C ABI of the reference C library
// (clib/pyCNCS/ncs.h:42-53) on gfx950.
//
// An ncsSubRegion keeps its r x r data, gamma, iterate and squared mask in
// host memory, as the reference does (ncs.c:394-446), and owns a solver
// context and a stream on GPU 0.  ncsSRSolve is a one-tile persistent solve
// in FP32 with liblbfgs's semantics (history 6, epsilon 1e-5, More-Thuente,
// scalar initial Hessian, no iteration cap); its per-tile status is mapped
// to liblbfgs's return codes.  The ncsSRCalc* functions evaluate the cost
// terms in FP64 on the GPU (one workgroup, direct r-point DFTs with FFTW's
// r2c / c2r conventions), so they return what the reference's double
// precision C returns.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "ncs_common.h"

namespace ncs {

void set_last_error(const char *msg);

// liblbfgs return codes used here (clib/pyCNCS/lbfgs.h:72-147)
constexpr int kLbfgsSuccess = 0;
constexpr int kLbfgsAlreadyMinimized = 2;
constexpr int kLbfgsErrRoundingError = -1001;
constexpr int kLbfgsErrMinimumStep = -1000;
constexpr int kLbfgsErrMaximumIteration = -997;
constexpr int kLbfgsErrIncreaseGradient = -994;
constexpr int kLbfgsErrUnknown = -1024;

// Cost terms of one r x r sub-region in FP64, one workgroup of 256 threads,
// thread p holding pixel p = i r + j.  in: data, gamma, u and q = mask^2
// (FFT order), r * r values each.  out: [0] the log-likelihood
// (ncs.c:185-208), [1] the noise contribution (ncs.c:259-299), then the
// log-likelihood gradient (ncs.c:158-175) and the noise-contribution
// gradient (ncs.c:220-250), r * r values each.
__global__ __launch_bounds__(256) void sr_eval_f64_kernel(int r, const double *in,
                                                          double *out) {
  __shared__ double ar[256], ai[256];  // first-axis transforms
  __shared__ double gr[256], gi[256];  // 2 q U on the half spectrum
  __shared__ double part[2][256];
  const int p = threadIdx.x, n = r * r, h = r / 2;
  const bool act = p < n;
  const int i = act ? p / r : 0, j = act ? p - i * r : 0;
  const double *d = in, *gm = in + n, *u = in + 2 * n, *q = in + 3 * n;
  double ll = 0.0;
  if (act) {
    const double c = d[p] + gm[p];
    double v = u[p] + gm[p];
    if (v < 1.0e-6) v = 1.0e-6;  // negative-value guard
    ll = u[p] - c * log(v);
    out[2 + p] = 1.0 - c / v;
  }
  // forward DFT, first axis: A[k][m] = sum_t u[t][m] e^(-2 pi i k t / r),
  // thread (k = i, m = j)
  if (act) {
    double sr = 0.0, si = 0.0;
    for (int t = 0; t < r; t++) {
      double s, c;
      sincospi(2.0 * ((i * t) % r) / r, &s, &c);
      sr += u[t * r + j] * c;
      si -= u[t * r + j] * s;
    }
    ar[p] = sr;
    ai[p] = si;
  }
  __syncthreads();
  // second axis: U[k][l] = sum_m A[k][m] e^(-2 pi i l m / r), thread (k, l)
  double ur = 0.0, ui = 0.0;
  if (act) {
    for (int m = 0; m < r; m++) {
      double s, c;
      sincospi(2.0 * ((j * m) % r) / r, &s, &c);
      const double xr = ar[i * r + m], xi = ai[i * r + m];
      ur += xr * c + xi * s;
      ui += xi * c - xr * s;
    }
  }
  // the noise contribution sums q |U|^2 over the full spectrum, which is
  // what the reference's unfolding of its half spectrum adds up
  const double nc = act ? q[p] * (ur * ur + ui * ui) : 0.0;
  // gradient spectrum 2 q U on the half spectrum l = 0..r/2, with q's own
  // first r/2 + 1 columns (ncs.c:234-241)
  if (act && j <= h) {
    gr[i * (h + 1) + j] = 2.0 * q[p] * ur;
    gi[i * (h + 1) + j] = 2.0 * q[p] * ui;
  }
  __syncthreads();
  // unnormalised c2r with FFTW's conventions (ncs.c:244): the complex
  // inverse DFT along the first axis, B[m][l] = sum_k G[k][l] e^(2 pi i k m / r)
  if (act && j <= h) {
    double br = 0.0, bi = 0.0;
    for (int k = 0; k < r; k++) {
      double s, c;
      sincospi(2.0 * ((k * i) % r) / r, &s, &c);
      const double xr = gr[k * (h + 1) + j], xi = gi[k * (h + 1) + j];
      br += xr * c - xi * s;
      bi += xr * s + xi * c;
    }
    ar[i * (h + 1) + j] = br;
    ai[i * (h + 1) + j] = bi;
  }
  __syncthreads();
  // then the real inverse along the second axis, which reads only the real
  // parts of the l = 0 and l = r/2 entries:
  // x[m][t] = Re B[m][0] + (-1)^t Re B[m][r/2] + 2 sum_{0<l<r/2} Re(B[m][l] e^(2 pi i l t / r))
  if (act) {
    const double *br = ar + i * (h + 1), *bi = ai + i * (h + 1);
    double x = br[0] + ((j & 1) ? -br[h] : br[h]);
    for (int l = 1; l < h; l++) {
      double s, c;
      sincospi(2.0 * ((l * j) % r) / r, &s, &c);
      x += 2.0 * (br[l] * c - bi[l] * s);
    }
    out[2 + n + p] = x / (double)n;  // normalization 1 / r^2 (ncs.c:247-249)
  }
  part[0][p] = ll;
  part[1][p] = nc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (p < s) {
      part[0][p] += part[0][p + s];
      part[1][p] += part[1][p + s];
    }
    __syncthreads();
  }
  if (p == 0) {
    out[0] = part[0][0];
    out[1] = part[1][0] / (double)n;
  }
}

}  // namespace ncs

using namespace ncs;

// The per-tile solver state of clib/pyCNCS/ncs.h:12-36, opaque to callers.
struct ncsSubRegion {
  int r = 16;
  double alpha = 0.0;
  std::vector<double> data, gamma, q, u;  // r x r each; q = mask^2
  std::vector<float> mask32;              // FFT-order mask for the solver tables
  bool mask_pending = true;               // mask32 not yet given to ctx
  ncs_ctx *ctx = nullptr;                 // solver context (GPU 0)
  hipStream_t stream = nullptr;           // FP64 evaluations
  double *dev = nullptr;                  // [4 r^2] inputs, [2 + 2 r^2] outputs
  std::vector<double> res;                // last evaluation's outputs
};

static const int kSrDevice = 0;

static void sr_release(ncsSubRegion *sr) {
  if (!sr) return;
  (void)hipSetDevice(kSrDevice);
  if (sr->stream) (void)hipStreamSynchronize(sr->stream);
  if (sr->dev) (void)hipFree(sr->dev);
  if (sr->stream) (void)hipStreamDestroy(sr->stream);
  if (sr->ctx) ncs_hip_destroy(sr->ctx);
  delete sr;
}

// ncs.c:394-446: NULL for an odd size; the GPU takes even sizes 4..16.
extern "C" ncsSubRegion *ncsSRInitialize(int r_size) {
  if (r_size < 4 || r_size > 16 || (r_size & 1)) {
    set_last_error("ncsSRInitialize: the GPU solves even sub-region sizes 4..16");
    return nullptr;
  }
  ncsSubRegion *sr = new ncsSubRegion();
  const size_t n = (size_t)r_size * r_size;
  sr->r = r_size;
  sr->data.assign(n, 0.0);
  sr->gamma.assign(n, 0.0);
  sr->q.assign(n, 0.0);
  sr->u.assign(n, 0.0);
  sr->mask32.assign(n, 0.0f);
  sr->res.assign(2 + 2 * n, 0.0);
  if (ncs_hip_create(kSrDevice, &sr->ctx) != NCS_OK ||
      hipSetDevice(kSrDevice) != hipSuccess ||
      hipStreamCreateWithFlags(&sr->stream, hipStreamNonBlocking) != hipSuccess ||
      hipMalloc(&sr->dev, (4 * n + 2 + 2 * n) * sizeof(double)) != hipSuccess) {
    set_last_error("ncsSRInitialize: no GPU context (device 0)");
    sr_release(sr);
    return nullptr;
  }
  return sr;
}

extern "C" void ncsSRCleanup(ncsSubRegion *sr) { sr_release(sr); }

extern "C" void ncsSRNewRegion(ncsSubRegion *sr, double *image, double *gamma) {
  if (!sr || !image || !gamma) return;
  std::memcpy(sr->data.data(), image, sr->data.size() * sizeof(double));
  std::memcpy(sr->gamma.data(), gamma, sr->gamma.size() * sizeof(double));
}

// ncs.c:502-510: the squared mask (FFT order) for the evaluations, and the
// mask itself for the solver's tables.
extern "C" void ncsSRSetOTFMask(ncsSubRegion *sr, double *otf_mask) {
  if (!sr || !otf_mask) return;
  for (size_t i = 0; i < sr->q.size(); i++) {
    sr->q[i] = otf_mask[i] * otf_mask[i];
    sr->mask32[i] = (float)otf_mask[i];
  }
  sr->mask_pending = true;
}

extern "C" void ncsSRSetU(ncsSubRegion *sr, double *u) {
  if (!sr || !u) return;
  std::memcpy(sr->u.data(), u, sr->u.size() * sizeof(double));
}

extern "C" void ncsSRGetU(ncsSubRegion *sr, double *u) {
  if (!sr || !u) return;
  std::memcpy(u, sr->u.data(), sr->u.size() * sizeof(double));
}

// FP64 evaluation of every cost term at the current u into sr->res.
static bool sr_evaluate(ncsSubRegion *sr) {
  const size_t n = sr->data.size();
  std::vector<double> in(4 * n);
  std::memcpy(in.data(), sr->data.data(), n * sizeof(double));
  std::memcpy(in.data() + n, sr->gamma.data(), n * sizeof(double));
  std::memcpy(in.data() + 2 * n, sr->u.data(), n * sizeof(double));
  std::memcpy(in.data() + 3 * n, sr->q.data(), n * sizeof(double));
  hipError_t e = hipSetDevice(kSrDevice);
  if (e == hipSuccess)
    e = hipMemcpyAsync(sr->dev, in.data(), 4 * n * sizeof(double), hipMemcpyHostToDevice,
                       sr->stream);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(sr_eval_f64_kernel, dim3(1), dim3(256), 0, sr->stream, sr->r, sr->dev,
                       sr->dev + 4 * n);
    e = hipGetLastError();
  }
  if (e == hipSuccess)
    e = hipMemcpyAsync(sr->res.data(), sr->dev + 4 * n, sr->res.size() * sizeof(double),
                       hipMemcpyDeviceToHost, sr->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(sr->stream);
  if (e != hipSuccess) {
    std::fill(sr->res.begin(), sr->res.end(), std::nan(""));
    return false;
  }
  return true;
}

extern "C" double ncsSRCalcLogLikelihood(ncsSubRegion *sr) {
  if (!sr) return std::nan("");
  sr_evaluate(sr);
  return sr->res[0];
}

extern "C" void ncsSRCalcLLGradient(ncsSubRegion *sr, double *gradient) {
  if (!sr || !gradient) return;
  sr_evaluate(sr);
  std::memcpy(gradient, sr->res.data() + 2, sr->u.size() * sizeof(double));
}

extern "C" double ncsSRCalcNoiseContribution(ncsSubRegion *sr) {
  if (!sr) return std::nan("");
  sr_evaluate(sr);
  return sr->res[1];
}

// The reference transforms the u of the last ncsSRCalcNoiseContribution
// (ncs.c:214-216); this one transforms the current u, which is the same in
// the reference's call order (noise contribution first).
extern "C" void ncsSRCalcNCGradient(ncsSubRegion *sr, double *gradient) {
  if (!sr || !gradient) return;
  sr_evaluate(sr);
  std::memcpy(gradient, sr->res.data() + 2 + sr->u.size(), sr->u.size() * sizeof(double));
}

// ncs.c:539-561: u = data, then L-BFGS; returns liblbfgs's code.  The
// solver's per-tile status maps to it: converged at the start ->
// LBFGS_ALREADY_MINIMIZED, converged -> LBFGS_SUCCESS, minimum step ->
// LBFGSERR_MINIMUMSTEP, a direction that does not descend ->
// LBFGSERR_INCREASEGRADIENT, no representable progress ->
// LBFGSERR_ROUNDING_ERROR, iteration cap (100,000) ->
// LBFGSERR_MAXIMUMITERATION.  verbose prints the last iteration's norms
// (the reference's progress callback prints every iteration, ncs.c:476-491).
extern "C" int ncsSRSolve(ncsSubRegion *sr, double alpha, int verbose) {
  if (!sr) return kLbfgsErrUnknown;
  sr->alpha = alpha;
  const int r = sr->r;
  const size_t n = sr->data.size();
  if (sr->mask_pending) {
    if (ncs_hip_set_otf_mask_window(sr->ctx, sr->mask32.data(), r) != NCS_OK)
      return kLbfgsErrUnknown;
    sr->mask_pending = false;
  }
  std::vector<float> d(n), g(n), out(n);
  for (size_t i = 0; i < n; i++) {
    d[i] = (float)sr->data[i];
    g[i] = (float)sr->gamma[i];
  }
  ncs_frame f;
  f.data = d.data();
  f.gamma = g.data();
  f.out = out.data();
  f.height = r;
  f.width = r;
  f.data_stride = f.gamma_stride = f.out_stride = 0;
  f.alpha = (float)alpha;
  f.geometry = NCS_GEOM_BLOCKS;  // one block: the sub-region as given
  ncs_params p;
  ncs_hip_default_params(&p);
  p.history = 6;               // lbfgs.h default m
  p.precond = 0;               // s.y / y.y initial Hessian
  p.max_iterations = 100000;   // liblbfgs's 0 = no cap
  p.precision = NCS_PRECISION_FP32;
  int iters = 0, status = NCS_UNSTARTED;
  if (ncs_hip_reduce_frames(sr->ctx, &f, 1, &p, &iters, &status, nullptr) != NCS_OK)
    return kLbfgsErrUnknown;
  for (size_t i = 0; i < n; i++) sr->u[i] = out[i];
  if (verbose && sr_evaluate(sr)) {
    double xn = 0.0, gn = 0.0;
    for (size_t i = 0; i < n; i++) {
      const double gi = sr->res[2 + i] + alpha * sr->res[2 + n + i];
      xn += sr->u[i] * sr->u[i];
      gn += gi * gi;
    }
    std::printf("Iteration %d:\n", iters);
    std::printf("  xnorm = %f, gnorm = %f\n", std::sqrt(xn), std::sqrt(gn));
    std::printf("\n");
    std::fflush(stdout);
  }
  switch (status) {
    case NCS_SUCCESS: return iters == 0 ? kLbfgsAlreadyMinimized : kLbfgsSuccess;
    case NCS_REACHED_MAXITERS: return kLbfgsErrMaximumIteration;
    case NCS_INCREASING_GRADIENT: return kLbfgsErrIncreaseGradient;
    case NCS_MINIMUM_STEP: return kLbfgsErrMinimumStep;
    case NCS_REACHED_MAXPRECISION: return kLbfgsErrRoundingError;
    default: return kLbfgsErrUnknown;
  }
}
