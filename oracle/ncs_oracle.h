/*
 * ncs_oracle.h -- FP64 CPU oracle for the NCS noise-correction path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library (ncs_amd/)
 * links, loads or calls this code; only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg use it, and only as the checker.
 *
 * It restates the reference C library (clib/pyCNCS/ncs.c, ncs.h) in
 * double precision.  The reference links two third-party libraries that
 * are absent from this image and from the reference tree:
 *   - FFTW3 (version not pinned by the reference; header clib/pyCNCS/fftw3.h)
 *     -> replaced by a direct 2-D real DFT in double with the same r2c/c2r
 *        conventions (unnormalised, half spectrum r x (r/2+1));
 *   - liblbfgs 1.10 (clib/pyCNCS/lbfgs.h:605-611; lbfgs.c not shipped)
 *     -> replaced by a restatement of its published algorithm: L-BFGS
 *        two-loop recursion (Nocedal 1980, Liu & Nocedal 1989) with the
 *        More-Thuente line search (More & Thuente 1994, ACM TOMS 20:286),
 *        parameters and return codes as documented in lbfgs.h:74-358.
 *
 * The exported ncsReduceNoise / ncsSR* functions keep the signatures of
 * clib/pyCNCS/ncs.h:42-53 so a pyCNCS-style ctypes wrapper binds them
 * unchanged.  ncso_* functions are batch/diagnostic additions.
 */
#ifndef NCS_ORACLE_H
#define NCS_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* liblbfgs return codes (clib/pyCNCS/lbfgs.h:72-147, enum order kept). */
enum {
  NCSO_LBFGS_SUCCESS = 0,
  NCSO_LBFGS_STOP = 1,
  NCSO_LBFGS_ALREADY_MINIMIZED = 2,
  NCSO_LBFGSERR_UNKNOWNERROR = -1024,
  NCSO_LBFGSERR_LOGICERROR,
  NCSO_LBFGSERR_OUTOFMEMORY,
  NCSO_LBFGSERR_CANCELED,
  NCSO_LBFGSERR_INVALID_N,
  NCSO_LBFGSERR_INVALID_N_SSE,
  NCSO_LBFGSERR_INVALID_X_SSE,
  NCSO_LBFGSERR_INVALID_EPSILON,
  NCSO_LBFGSERR_INVALID_TESTPERIOD,
  NCSO_LBFGSERR_INVALID_DELTA,
  NCSO_LBFGSERR_INVALID_LINESEARCH,
  NCSO_LBFGSERR_INVALID_MINSTEP,
  NCSO_LBFGSERR_INVALID_MAXSTEP,
  NCSO_LBFGSERR_INVALID_FTOL,
  NCSO_LBFGSERR_INVALID_WOLFE,
  NCSO_LBFGSERR_INVALID_GTOL,
  NCSO_LBFGSERR_INVALID_XTOL,
  NCSO_LBFGSERR_INVALID_MAXLINESEARCH,
  NCSO_LBFGSERR_INVALID_ORTHANTWISE,
  NCSO_LBFGSERR_INVALID_ORTHANTWISE_START,
  NCSO_LBFGSERR_INVALID_ORTHANTWISE_END,
  NCSO_LBFGSERR_OUTOFINTERVAL,
  NCSO_LBFGSERR_INCORRECT_TMINMAX,
  NCSO_LBFGSERR_ROUNDING_ERROR,
  NCSO_LBFGSERR_MINIMUMSTEP,
  NCSO_LBFGSERR_MAXIMUMSTEP,
  NCSO_LBFGSERR_MAXIMUMLINESEARCH,
  NCSO_LBFGSERR_MAXIMUMITERATION,
  NCSO_LBFGSERR_WIDTHTOOSMALL,
  NCSO_LBFGSERR_INVALIDPARAMETERS,
  NCSO_LBFGSERR_INCREASEGRADIENT
};

/* Subset of lbfgs_parameter_t (lbfgs.h:198-358) that the NCS path uses.
 * Defaults (ncso_param_init): m=6, epsilon=1e-5, past=0, delta=1e-5,
 * max_iterations=0 (unbounded), max_linesearch=40, min_step=1e-20,
 * max_step=1e20, ftol=1e-4, gtol=0.9, xtol=1e-16 (machine precision). */
typedef struct {
  int m;
  double epsilon;
  int past;
  double delta;
  int max_iterations;
  int max_linesearch;
  double min_step;
  double max_step;
  double ftol;
  double gtol;
  double xtol;
} ncso_param;

void ncso_param_init(ncso_param *p);

/* Opaque per-tile state (clib/pyCNCS/ncs.h:12-36). */
typedef struct ncsSubRegion ncsSubRegion;

/* ---- reference C ABI (clib/pyCNCS/ncs.h:42-53) ---- */
void ncsReduceNoise(double *ncs_image, double *image, double *gamma,
                    double *otf_mask, double alpha, int im_x, int im_y,
                    int r_size);
void ncsSRCalcLLGradient(ncsSubRegion *sr, double *gradient);
double ncsSRCalcLogLikelihood(ncsSubRegion *sr);
void ncsSRCalcNCGradient(ncsSubRegion *sr, double *gradient);
double ncsSRCalcNoiseContribution(ncsSubRegion *sr);
void ncsSRCleanup(ncsSubRegion *sr);
void ncsSRGetU(ncsSubRegion *sr, double *u);
ncsSubRegion *ncsSRInitialize(int r_size);
void ncsSRNewRegion(ncsSubRegion *sr, double *image, double *gamma);
void ncsSRSetOTFMask(ncsSubRegion *sr, double *otf_mask);
void ncsSRSetU(ncsSubRegion *sr, double *u);
int ncsSRSolve(ncsSubRegion *sr, double alpha, int verbose);

/* ---- oracle additions ---- */

/* Solver parameters of an ncsSubRegion (copied in / out). */
void ncso_sr_set_param(ncsSubRegion *sr, const ncso_param *p);
/* Iterations and evaluations used by the last ncsSRSolve, and its final cost. */
void ncso_sr_last_stats(ncsSubRegion *sr, int *iterations, int *evaluations,
                        double *final_cost);
/* Solve from the current u (ncsSRSetU) instead of u = data. */
int ncso_sr_solve_from_u(ncsSubRegion *sr, double alpha);

/* Cost f(u) = LL + alpha*NC and gradient of one r x r tile (mask in FFT
 * order, unsquared).  grad may be NULL. */
double ncso_cost_grad(const double *u, const double *data, const double *gamma,
                      const double *mask_fft, double alpha, int r,
                      double *grad);

/* Batch of independent r x r tiles, tile-major [ntiles][r*r].
 * u0 == NULL starts from data (the reference's plain start).
 * Outputs (any may be NULL): u_out, cost_out (final cost), status_out
 * (liblbfgs code), iters_out (iterations), evals_out (evaluations).
 * nthreads <= 0 uses all OpenMP threads.  Returns 0 or -1 on bad args. */
int ncso_solve_tiles(const double *data, const double *gamma,
                     const double *u0, const double *mask_fft, double alpha,
                     int r, int ntiles, const ncso_param *p, int nthreads,
                     double *u_out, double *cost_out, int *status_out,
                     int *iters_out, int *evals_out);

/* Same, with FP32 tile inputs/outputs (costs stay double). */
int ncso_solve_tiles_f32(const float *data, const float *gamma,
                         const float *u0, const float *mask_fft, double alpha,
                         int r, int ntiles, const ncso_param *p, int nthreads,
                         float *u_out, double *cost_out, int *status_out,
                         int *iters_out, int *evals_out);

/* FP32 tiles with a per-pixel likelihood weight lin in {0, 1} (zero-padded
 * pixels of the block tiling); lin = 1 everywhere is ncsSRSolve. */
int ncso_solve_tiles_masked_f32(const float *data, const float *gamma,
                                const float *lin, const float *mask_fft,
                                double alpha, int r, int ntiles,
                                const ncso_param *p, int nthreads, float *u_out,
                                double *cost_out, int *status_out,
                                int *iters_out);

/* Batch cost evaluation of FP32 tiles in FP64 (error reporting). */
void ncso_eval_cost_f32(const float *u, const float *data, const float *gamma,
                        const float *mask_fft, double alpha, int r, int ntiles,
                        int nthreads, double *cost_out, double *gnorm_out,
                        double *unorm_out);

/* Whole-image NCS with the reference tiling (ncs.c:71-143): tiles solved in
 * parallel, then stitched in raster order (last writer wins), so the result
 * equals the serial ncsReduceNoise bit for bit.  Returns the tile count, or
 * -1 on bad arguments.  tile_cost/tile_status/tile_iters may be NULL. */
int ncso_reduce_noise(double *ncs_image, const double *image,
                      const double *gamma, const double *mask_fft,
                      double alpha, int im_x, int im_y, int r_size,
                      const ncso_param *p, int nthreads, double *tile_cost,
                      int *tile_status, int *tile_iters);

/* Number of windows per axis of the reference tiling for an axis of size n. */
int ncso_tiles_per_axis(int n, int r_size);

#ifdef __cplusplus
}
#endif
#endif
