/*
 * ncs_hip.h -- C ABI of libncs_hip.so, the MI355X (gfx950) NCS solver.
 *
 * Plain pointers and sizes only.  Every entry point returns 0 on success and
 * a negative NCS_E* code on error (the message is kept for
 * ncs_hip_last_error()); none of them prints.  Per-tile solver status uses the
 * codes of the reference GPU kernel (opencl/ncs.cl:66-71).
 *
 * Reference interfaces replaced (file:line in HuanglabPurdue/NCS):
 *   ncsReduceNoise (C ABI)          clib/pyCNCS/ncs.h:42, ncs.c:41-147
 *   ncsReduceNoise (GPU kernel ABI) opencl/ncs.cl:820-826, cuda/ncs.cu:873-879
 *   reduceNoise (Python host)       opencl/pyOpenCLNCS/ncs.py:47-198, 236-251
 *   setOTFMask                      opencl/pyOpenCLNCS/ncs.py:208-222,
 *                                   clib/pyCNCS/ncs.c:502-510
 */
#ifndef NCS_HIP_H
#define NCS_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* error codes */
#define NCS_OK 0
#define NCS_E_ARG (-1)      /* invalid argument */
#define NCS_E_HIP (-2)      /* HIP runtime error */
#define NCS_E_NOMASK (-3)   /* no OTF mask set on the context */
#define NCS_E_MASK (-4)     /* mask not point-symmetric */
#define NCS_E_NOWEIGHTS (-5) /* learned start requested without weights */

/* per-tile status (opencl/ncs.cl:66-71) */
#define NCS_SUCCESS 0
#define NCS_UNSTARTED (-1)
#define NCS_REACHED_MAXITERS (-2)
#define NCS_INCREASING_GRADIENT (-3)
#define NCS_MINIMUM_STEP (-4)
#define NCS_REACHED_MAXPRECISION (-5)

/* tiling of a frame */
#define NCS_GEOM_REFERENCE 0 /* 16x16 windows, stride 14, 1-px replicated halo,
                                last window clamped, interiors stitched with the
                                later window owning overlaps (ncs.c:71-143) */
#define NCS_GEOM_BLOCKS 1    /* non-overlapping blocks of the window side (16,
                                or r after ncs_hip_set_otf_mask_window); pixels
                                outside the frame are zero-padded and excluded
                                from the likelihood term */

/* line searches */
#define NCS_LS_MORE_THUENTE 0 /* liblbfgs default (lbfgs.h:704-719) */
#define NCS_LS_ARMIJO 1       /* backtracking of opencl/ncs.cl:945-1035 */

/* arithmetic of the solver's basis changes (the 16x16 real-DFT transforms) */
#define NCS_PRECISION_FP32 0      /* FP32 MFMA (v_mfma_f32_16x16x4_f32) throughout:
                                     every operand and accumulator FP32 (default) */
#define NCS_PRECISION_F16_SPLIT 1 /* each FP32 operand split into two FP16 halves
                                     on the FP16 matrix cores, FP32 accumulation
                                     (about 2^-22 per operand); FP32 MFMA once the
                                     gradient is within 1e-5 of the stop rule's
                                     scale, on overflow, and for the rest of a
                                     tile whose line search fails */

/* starting point */
#define NCS_START_PLAIN 0   /* u0 = data (ncs.c:549-551, ncs.cl:906) */
#define NCS_START_TILES 1   /* u0 from a caller tile buffer [tiles][256] */
#define NCS_START_LEARNED 2 /* u0 from the context's conv net */
#define NCS_START_LINEAR 3  /* u0 from the context's learned linear start,
                               computed inside the solve */

typedef struct {
  int max_iterations;  /* L-BFGS iteration cap (ncs.cl:61 MAXITERS = 200) */
  int history;         /* correction pairs m, 1..8 (lbfgs.h default 6, ncs.cl 8);
                          0 = the autotuned depth (5 until ncs_hip_autotune ran) */
  float epsilon;       /* stop when |g| <= epsilon * max(1, |u|); 0 disables */
  int linesearch;      /* NCS_LS_* */
  int max_linesearch;  /* trials per iteration */
  float ftol;          /* sufficient decrease (Armijo c1 = ftol) */
  float gtol;          /* curvature (More-Thuente) */
  float xtol;          /* relative interval width (More-Thuente) */
  float min_step;      /* smallest step */
  float max_step;      /* largest step */
  float stepm;         /* Armijo backtracking factor */
  int start;           /* NCS_START_* */
  int fallback;        /* warm starts: keep u0 = data where f(data) <= f(u0) */
  int waves_per_cu;    /* persistent grid size; 0 = autotuned (else the
                          occupancy limit) */
  int precond;         /* initial inverse Hessian H0 of the two-loop
                          recursion.  1: P = (kappa I + alpha K)^-1 in the
                          real DFT basis, kappa = tile mean of the
                          likelihood curvature (d+gamma)/(u+gamma)^2:
                          diagonal for masks without an odd-odd part, 2 x 2
                          blocks of coupled coefficients otherwise.
                          2 (default): P - P D P, D = the per-pixel
                          curvature less kappa (clamped at kappa / 2 from
                          above), which adds the curvature's variation over
                          the tile to first order (two more basis changes
                          per direction, fewer iterations).  0: the scalar
                          s.y/y.y of liblbfgs and ncs.cl:1123 */
  int waves_per_group; /* tiles (waves) per workgroup, 1..4; 0 = autotuned
                          (else 1) */
  int precision;       /* NCS_PRECISION_* */
} ncs_params;

typedef struct {
  const float *data;  /* photo-electrons, row-major [height][stride] */
  const float *gamma; /* read-noise variance / gain^2, same shape */
  float *out;         /* corrected image, same shape (may be NULL) */
  int height, width;
  int data_stride, gamma_stride, out_stride; /* elements; 0 = width */
  float alpha;        /* NCS alpha of this frame */
  int geometry;       /* NCS_GEOM_* */
} ncs_frame;

/* Context: device, stream, OTF-mask tables, work buffers, captured graphs,
 * autotune cache and (optionally) conv-net weights. */
typedef struct ncs_ctx ncs_ctx;

/* Fill defaults: max_iterations 200, history 5, epsilon 1e-5, More-Thuente
 * (ftol 1e-4, gtol 0.9, xtol 1e-7, max_linesearch 20, steps 1e-20..1e20),
 * Armijo stepm 0.5, plain start, fallback on, autotuned grid and workgroup
 * size, spectral H0, FP32 basis changes. */
void ncs_hip_default_params(ncs_params *p);

int ncs_hip_create(int device, ncs_ctx **out);
void ncs_hip_destroy(ncs_ctx *ctx);
/* Message of the last error on this thread. */
const char *ncs_hip_last_error(void);
/* Version string of the library and the GPU it runs on. */
int ncs_hip_device_info(ncs_ctx *ctx, char *buf, int buflen);

/* Set the 16x16 OTF mask, FFT order, unsquared (replaces ncsSRSetOTFMask,
 * ncs.c:502-510, and the otf_mask argument of ncs.cl:822).  The solver uses
 * the point-symmetric part of q = mask^2, which equals q for every mask that
 * passes the reference's checkOTFMask. */
int ncs_hip_set_otf_mask(ncs_ctx *ctx, const float *mask_fft_order);

/* An r x r mask (FFT order, r even, 4..16) for sub-regions of side r, the
 * clib's r_size (ncs.c:41-147 takes any even r; pyCNCS derives it from the
 * mask, ncs_c.py:250-276).  Frames then use windows (reference geometry:
 * stride r - 2) or blocks of side r, each solved in the top-left of a
 * 16 x 16 tile whose other pixels are left out of the cost, and the tile
 * entry points take r x r tiles; only the plain start is accepted, and the
 * evaluation and network entry points, which take 16 x 16 tiles, are
 * refused.  r = 16 is ncs_hip_set_otf_mask. */
int ncs_hip_set_otf_mask_window(ncs_ctx *ctx, const float *mask_fft_order, int r);

/* Number of tiles a frame produces with the given geometry (16 x 16
 * windows); -1 if the frame is too small.  Per-tile result arrays hold
 * this many entries. */
long long ncs_hip_tile_count(int height, int width, int geometry);
/* The same for windows or blocks of side r: the count after
 * ncs_hip_set_otf_mask_window(ctx, mask, r). */
long long ncs_hip_tile_count_window(int height, int width, int geometry, int r);

/* Whole frames, host memory in and out; synchronous.  Replaces the batched
 * host reduceNoise (pyOpenCLNCS/ncs.py:47-198): every frame may have its own
 * size, gamma, alpha and geometry and all tiles of all frames go through one
 * persistent launch.  iterations/status/cost (host, may be NULL) receive one
 * entry per tile, frames in order, tiles in raster order.  The launch
 * schedules the frames smallest first (stable, by tile count), so that the
 * long-running tiles of small frames and ROIs start early; results and the
 * order of every per-tile array are the same as in the caller's order. */
int ncs_hip_reduce_frames(ncs_ctx *ctx, const ncs_frame *frames, int n_frames,
                          const ncs_params *params, int *iterations,
                          int *status, float *cost);

/* Same with device pointers (frames[] itself in host memory), enqueued on
 * `stream` (a hipStream_t; 0 is the default NULL stream, as torch's default
 * stream) without waiting for completion.  u0_tiles: device [tiles][256] for NCS_START_TILES, else
 * NULL.  iterations/status/cost: device arrays or NULL.  The solves of one
 * context share its frame table and work queue: one enqueued on another
 * stream than the context's previous solve waits on the device for that
 * solve to complete, so calls on several streams are safe but run one after
 * the other (use one context per stream for concurrent solves). */
int ncs_hip_reduce_frames_device(ncs_ctx *ctx, const ncs_frame *frames,
                                 int n_frames, const ncs_params *params,
                                 const float *u0_tiles, int *iterations,
                                 int *status, float *cost, void *stream);

/* ncs_hip_reduce_frames_device without allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * work-queue counters and the frame table live in a fixed device array of
 * the library (21,504 bytes per GPU, held until the process ends), written
 * by kernels; at most 256 frames per call; the learned (conv-net) start,
 * which needs a start buffer, is refused.  Every call on one GPU shares that
 * array: calls that overlap on one GPU race, whatever their context. */
int ncs_hip_reduce_frames_device_noalloc(ncs_ctx *ctx, const ncs_frame *frames,
                                         int n_frames, const ncs_params *params,
                                         const float *u0_tiles, int *iterations,
                                         int *status, float *cost, void *stream);

/* Independent 16x16 tiles [num_tiles][16][16] (r x r tiles [num_tiles][r][r]
 * after an r x r mask), host memory; synchronous.  Replaces the kernel ABI
 * ncsReduceNoise(data_in, g_gamma, otf_mask, data_out, g_iterations,
 * g_status, alpha) of opencl/ncs.cl:820-826 (the mask comes from
 * ncs_hip_set_otf_mask).  cost may be NULL. */
int ncs_hip_solve_tiles(ncs_ctx *ctx, const float *data_in,
                        const float *gamma, float *data_out, int *iterations,
                        int *status, float *cost, int num_tiles, float alpha,
                        const ncs_params *params);

/* Device-pointer variant of ncs_hip_solve_tiles, enqueued on `stream`.
 * u0 (device [num_tiles][256]) or NULL for the plain start. */
int ncs_hip_solve_tiles_device(ncs_ctx *ctx, const float *data_in,
                               const float *gamma, const float *u0,
                               float *data_out, int *iterations, int *status,
                               float *cost, int num_tiles, float alpha,
                               const ncs_params *params, void *stream);

/* FP32 cost f(u) and gradient of independent tiles [num_tiles][16][16]
 * (host memory; grad or cost may be NULL).  The evaluation the solver uses,
 * exposed for parity with ncsSREvaluate (clib/pyCNCS/ncs.c:335-364) and the
 * ncsSRCalc* functions it calls (ncs.c:158-299). */
int ncs_hip_eval_tiles(ncs_ctx *ctx, const float *data, const float *gamma,
                       const float *u, float *cost, float *grad, int num_tiles,
                       float alpha);

/* Drop-in for the C library's whole-image entry (clib/pyCNCS/ncs.h:42):
 * same signature and semantics (reference tiling, FFT-order mask), solved on
 * GPU 0 in FP32 with the default parameters.  r_size must be 16; on error
 * ncs_image is left unchanged (the C prints and returns, ncs.c:55-58). */
void ncsReduceNoise(double *ncs_image, double *image, double *gamma,
                    double *otf_mask, double alpha, int im_x, int im_y,
                    int r_size);

/* ---- per-sub-region C ABI of the C library (clib/pyCNCS/ncs.h:43-53) ----
 *
 * Same names, signatures and call conventions as the reference, so that a
 * pyCNCS-style ctypes binding (clib/pyCNCS/ncs_c.py:18-60) loads this
 * library unchanged.  A sub-region keeps its r x r data, gamma, iterate and
 * mask in host memory (the caller's arrays are copied, float64 row-major)
 * and owns a solver context on GPU 0; one sub-region per host thread, as in
 * the reference.  The sizes are even, 4..16 (the reference takes any even
 * size).  The ncsSRCalc* functions evaluate in FP64 on the GPU; ncsSRSolve
 * runs the FP32 solver with liblbfgs's semantics. */
typedef struct ncsSubRegion ncsSubRegion;

/* ncs.c:394-446.  NULL for an unsupported size or without a GPU (message in
 * ncs_hip_last_error()). */
ncsSubRegion *ncsSRInitialize(int r_size);
/* ncs.c:459-468: copy the sub-region's data and gamma (r x r each). */
void ncsSRNewRegion(ncsSubRegion *sr, double *image, double *gamma);
/* ncs.c:502-510: r x r mask, FFT order, unsquared. */
void ncsSRSetOTFMask(ncsSubRegion *sr, double *otf_mask);
/* ncs.c:521-529 and 375-383: set / read the iterate u. */
void ncsSRSetU(ncsSubRegion *sr, double *u);
void ncsSRGetU(ncsSubRegion *sr, double *u);
/* ncs.c:539-561: u = data, then L-BFGS (history 6, epsilon 1e-5,
 * More-Thuente, initial Hessian s.y/y.y, no iteration cap; FP32 on the GPU).
 * Returns liblbfgs's code (lbfgs.h:72-147): 0 converged, 2 already minimal,
 * -1000 minimum step, -994 not a descent direction, -1001 no representable
 * progress, -1024 a library error.  verbose prints the final iteration's
 * norms (the reference prints every iteration, ncs.c:476-491). */
int ncsSRSolve(ncsSubRegion *sr, double alpha, int verbose);
/* ncs.c:185-208: sum of u - (data + gamma) ln max(u + gamma, 1e-6). */
double ncsSRCalcLogLikelihood(ncsSubRegion *sr);
/* ncs.c:158-175: 1 - (data + gamma) / max(u + gamma, 1e-6), r x r. */
void ncsSRCalcLLGradient(ncsSubRegion *sr, double *gradient);
/* ncs.c:259-299: (1/r^2) sum_k mask_k^2 |DFT2(u)_k|^2. */
double ncsSRCalcNoiseContribution(ncsSubRegion *sr);
/* ncs.c:220-250: c2r of 2 mask^2 DFT2(u) on the half spectrum, / r^2.  It
 * transforms the current u (the reference reuses the transform of the last
 * ncsSRCalcNoiseContribution: the same in its call order). */
void ncsSRCalcNCGradient(ncsSubRegion *sr, double *gradient);
/* ncs.c:307-327. */
void ncsSRCleanup(ncsSubRegion *sr);

/* ---- learned warm start: 4-layer 3x3 conv net on BF16 MFMA ---- */

/* The net (no counterpart in the reference): 2 -> 32 -> 32 -> 32 -> 1
 * channels, 3x3 kernels with zero padding inside each 16x16 tile, ReLU after
 * layers 1-3.  Inputs per tile: mu = mean(d), sigma = sqrt(max(mu,1) +
 * mean(gamma)), x0 = (d - mu)/sigma, x1 = gamma/sigma^2; the start is
 * u0 = d + sigma * y.  Parameters are FP32, per layer the weights
 * [cout][cin][3][3] then the bias (19,393 values); the kernels compute in
 * BF16 with FP32 accumulation. */
#define NCS_NET_PARAMS 19393

/* Sets the context's network (host FP32 parameters) for NCS_START_LEARNED. */
int ncs_hip_net_set_weights(ncs_ctx *ctx, const float *params_f32);

/* The starts u0 = d + sigma * net(x) of every window of device frames, as
 * NCS_START_LEARNED computes them, into u0_tiles (device [tiles][256]). */
int ncs_hip_net_predict_frames_device(ncs_ctx *ctx, const ncs_frame *frames,
                                      int n_frames, float *u0_tiles,
                                      void *stream);

/* Forward on device tensors: x [n][2][256] normalised inputs (FP32) ->
 * out [n][256] (FP32).  acts (device, may be NULL) receives the post-ReLU
 * activations of layers 1-3 as BF16 [3][n][256][32] for the backward pass.
 * params: device FP32 parameters, or NULL for the context's weights.
 * Enqueued on `stream`. */
int ncs_hip_net_forward(ncs_ctx *ctx, const float *params, const float *x,
                        float *out, void *acts, int n, void *stream);

/* Backward: from dL/dout [n][256] (FP32) and the saved activations, writes
 * dL/dparams (device FP32 [NCS_NET_PARAMS], overwritten).  grad_x must be
 * NULL (the input gradient is not needed for training). */
int ncs_hip_net_backward(ncs_ctx *ctx, const float *params, const float *x,
                         const void *acts, const float *grad_out,
                         float *grad_params, float *grad_x, int n,
                         void *stream);

/* ---- learned linear start (no reference counterpart) ----
 *
 * A linear predictor of the NCS solution in the orthonormal real-DFT basis
 * R of the 16x16 tile (rows cos k, k = 0..8, then sin k, k = 1..7, of length
 * 16; Z = R X R^T): with mu = mean d, sigma = sqrt(max(mu, 1) + mean gamma),
 * x0 = (d - mu) / sigma and x1 = gamma / sigma^2,
 *   Z(u0) = Z(d) + sigma (A * Z(x0) + B * Z(x1) + c * Z(1)),
 * coefficient-wise.  It is fitted in closed form on (tile, solution) pairs
 * the solver produced (ncs_amd.linstart) and evaluated inside the solve
 * kernel at the cost of three basis changes; the plain start is kept per
 * tile whenever it scores the lower cost (params->fallback). */
#define NCS_LINEAR_START_PARAMS 513

/* coef: A[16][16], then B[16][16] (indices [a][b] of Z), then c.  Clears
 * any slopes set by ncs_hip_set_linear_start_alpha. */
int ncs_hip_set_linear_start(ncs_ctx *ctx, const float *coef);

/* No reference counterpart.  Slopes of the linear start's coefficients in
 * ln(alpha), same layout as coef (NCS_LINEAR_START_PARAMS floats): a frame
 * at alpha uses coef + ln(alpha / alpha_ref) * slope, so one fitted set
 * serves a batch whose frames carry different alphas.  Frames at alpha_ref
 * (and at alpha <= 0) use coef unchanged; ln(alpha / alpha_ref) is clamped
 * to +-ln 8, so the slopes extrapolate at most 8x from alpha_ref.  Call
 * after ncs_hip_set_linear_start (NCS_E_NOWEIGHTS otherwise); alpha_ref > 0. */
int ncs_hip_set_linear_start_alpha(ncs_ctx *ctx, const float *slope, float alpha_ref);

/* ---- autotuner (no reference counterpart) ---- */

typedef struct {
  int history;         /* chosen L-BFGS depth (the given one unless tuned) */
  int waves_per_cu;    /* chosen persistent grid, waves per CU */
  int waves_per_group; /* chosen tiles (waves) per workgroup */
  int lds_waves_per_cu;/* occupancy limit of the chosen depth (LDS/VGPR) */
  float ms;            /* time of the chosen schedule on the sample */
  float ms_default;    /* time of the untuned schedule on the sample */
  int candidates;      /* schedules timed */
} ncs_tune_result;

/* Times the persistent solve on a sample of the workload (device frames, as
 * for ncs_hip_reduce_frames_device; nothing is written to their outputs) over
 * candidate schedules -- waves per CU up to the occupancy that the history's
 * LDS rings allow, 1/2/4 tiles per workgroup and, if tune_history is set,
 * history depths 1, 2, 3, 5 and 8 -- each the best of three timed launches
 * after an untimed one -- and keeps the fastest in the context's
 * cache for the mask kind and geometry of the sample.  Later calls whose
 * params leave waves_per_cu, waves_per_group or history at 0 use it.
 * Synchronous. */
int ncs_hip_autotune(ncs_ctx *ctx, const ncs_frame *frames, int n_frames,
                     const ncs_params *base, int tune_history,
                     ncs_tune_result *result);

/* ---- captured streaming pipeline (no reference counterpart) ----
 *
 * The device work of a fixed batch shape is captured once as a hipGraph:
 * camera conversion (raw input) -> learned start (if params->start is
 * NCS_START_LEARNED) -> work-queue reset -> persistent solve with the stitch
 * fused.  A replay is that graph between stream copies on the DMA engines:
 * upload of every frame from host memory before it, download of every
 * corrected frame (and, if asked, of the per-tile results) after it.  Each
 * replay may name other host buffers at no cost; pinned buffers
 * (hipHostMalloc) make the copies asynchronous.  Replays of one pipeline run
 * one after the other in launch order (they share its device buffers);
 * launching does not wait on the host for the previous replay, so the host
 * can enqueue ahead of the device, up to NCS_PIPELINE_REPLAYS replays per
 * pipeline.  Two pipelines overlap one frame's upload and solve with the
 * other's download.  Gamma (or the camera maps) are uploaded once per
 * pipeline. */

#define NCS_PIPELINE_REPLAYS 4 /* replays of one pipeline in flight at most */
#define NCS_MAX_GAIN_LEVELS 4  /* gain amplifiers of one camera at most */

#define NCS_INPUT_PHOTONS_F32 0 /* data in photo-electrons, float32 [H][W] */
#define NCS_INPUT_ADU_U16 1     /* raw camera ADU, uint16 [H][W]; photons =
                                   (ADU - offset) / gain on the device, values
                                   <= 0 set to 1e-6 (denoisetools.py:218-224) */

typedef struct {
  int height, width;
  float alpha;
  int geometry; /* NCS_GEOM_* */
} ncs_pipeline_frame;

typedef struct ncs_pipeline ncs_pipeline;

/* The pipeline keeps a pointer to ctx, which must outlive it; the context's
 * mask and network are read at every replay (a change of mask kind, even to
 * general, needs a new pipeline). */
int ncs_hip_pipeline_create(ncs_ctx *ctx, const ncs_pipeline_frame *frames,
                            int n_frames, const ncs_params *params,
                            int input_format, int keep_tile_results,
                            ncs_pipeline **out);
void ncs_hip_pipeline_destroy(ncs_pipeline *pl);
/* Gamma of frame i (host float32 [H][W], photons^2); synchronous. */
int ncs_hip_pipeline_set_gamma(ncs_pipeline *pl, int frame, const float *gamma);
/* Camera maps of frame i for NCS_INPUT_ADU_U16 (host float32 [H][W]): offset
 * (ADU), gain (ADU/e-), read variance (ADU^2); also sets gamma = variance /
 * gain^2 (denoisetools.py:266).  Synchronous. */
int ncs_hip_pipeline_set_camera(ncs_pipeline *pl, int frame, const float *offset,
                                const float *gain, const float *variance);
/* Camera maps of frame i for an sCMOS with several gain amplifiers
 * (NCS_INPUT_ADU_U16; replaces matlab/genmaps.m:29-62 and its use in
 * matlab/multimaps_example.m).  n_levels (1..NCS_MAX_GAIN_LEVELS) gain and
 * read-variance maps, host float32 [n_levels][H][W], and n_levels - 1
 * ascending ADU division points (float64; may be NULL for one level).  At
 * every replay each pixel picks level L = the number of division points
 * <= its ADU, then data = (ADU - offset)/gain_L (values <= 0 set to 1e-6)
 * and gamma = variance_L/gain_L^2.  One level is ncs_hip_pipeline_set_camera.
 * Synchronous; a change of levels or division points captures the graph
 * again at the next launch.  Division points that descend are rejected
 * (genmaps.m would keep the previous frame's maps for the pixels no level
 * covers). */
int ncs_hip_pipeline_set_camera_levels(ncs_pipeline *pl, int frame, int n_levels,
                                       const double *points, const float *offset,
                                       const float *gains, const float *variances);
/* Enqueue one replay on `stream` (0: the pipeline's own stream): inputs[i]
 * is frame i's host input, outputs[i] its host float32 output [H][W].  The
 * first launch needs both arrays; later ones may pass NULL to keep the
 * previous pointers.  The buffers must stay valid until the replay is
 * complete.  Replays are numbered 0, 1, ... in launch order.  A replay on
 * another stream than the previous one waits for it on the device.  The
 * host waits only when NCS_PIPELINE_REPLAYS replays are in flight (for the
 * oldest). */
int ncs_hip_pipeline_launch(ncs_pipeline *pl, const void *const *inputs,
                            float *const *outputs, void *stream);
/* Replays launched so far (the next replay's number); -1 for NULL. */
long long ncs_hip_pipeline_launched(const ncs_pipeline *pl);
/* Wait for every launched replay (outputs in host memory). */
int ncs_hip_pipeline_wait(ncs_pipeline *pl);
/* Wait for replay `replay` and every earlier one. */
int ncs_hip_pipeline_wait_replay(ncs_pipeline *pl, long long replay);
/* Per-tile results of replay `replay`, one of the last NCS_PIPELINE_REPLAYS
 * launched (waits for it), as for ncs_hip_reduce_frames; needs
 * keep_tile_results.  Any output may be NULL. */
int ncs_hip_pipeline_replay_results(ncs_pipeline *pl, long long replay, int *iterations,
                                    int *status, float *cost);
/* ncs_hip_pipeline_replay_results of the last replay launched. */
int ncs_hip_pipeline_tile_results(ncs_pipeline *pl, int *iterations, int *status,
                                  float *cost);
/* Tiles of one replay. */
long long ncs_hip_pipeline_tiles(const ncs_pipeline *pl);

#ifdef __cplusplus
}
#endif
#endif
