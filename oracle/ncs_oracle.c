/*
 * ncs_oracle.c -- FP64 CPU oracle for the NCS noise-correction path.
 *
 * TEST INFRASTRUCTURE ONLY (see ncs_oracle.h).  Restates, in double
 * precision, the semantics of clib/pyCNCS/ncs.c; each function cites the
 * reference lines it follows.  FFTW and liblbfgs, which the reference links
 * but which are absent here, are replaced by a direct DFT and by an own
 * L-BFGS + More-Thuente restatement (ncs_oracle.h explains both).
 */
#include "ncs_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define FITMIN 1.0e-6 /* negative-value guard, ncs.c:169-171, 198-200 */

struct ncsSubRegion {
  int r_size;        /* ncs.h:14 */
  int fft_size;      /* r/2 + 1, ncs.c:428 */
  double alpha;      /* ncs.h:17 */
  double normalization; /* 1/r^2, ncs.c:430 */
  double *data, *gamma, *otf_mask_sqr, *t1, *t2, *u;
  double *u_fft_re, *u_fft_im; /* half spectrum r x fft_size (ncs.h:33) */
  double *g_fft_re, *g_fft_im; /* half spectrum of the NC gradient (ncs.h:32) */
  double *g;                   /* c2r output (ncs.c:433) */
  double *cos_tab, *sin_tab;   /* cos/sin(2*pi*m/r), m = 0..r-1 */
  double *w_re, *w_im;         /* r x r DFT workspace */
  const double *lin;           /* per-pixel likelihood weight (NULL = all 1) */
  ncso_param param;
  int last_iters, last_evals;
  double last_cost;
};

/* ------------------------------------------------------------------ */
/* Parameters (lbfgs.h:198-358; liblbfgs 1.10 lbfgs_parameter_init).   */
/* ------------------------------------------------------------------ */

void ncso_param_init(ncso_param *p) {
  p->m = 6;
  p->epsilon = 1e-5;
  p->past = 0;
  p->delta = 1e-5;
  p->max_iterations = 0;
  p->max_linesearch = 40;
  p->min_step = 1e-20;
  p->max_step = 1e20;
  p->ftol = 1e-4;
  p->gtol = 0.9;
  p->xtol = 1e-16;
}

/* ------------------------------------------------------------------ */
/* Direct 2-D real DFT with FFTW's r2c / c2r conventions.              */
/* ------------------------------------------------------------------ */

/* U[k][l] = sum_{i,j} u[i][j] exp(-2 pi i (k i + l j) / r), l < r/2+1.
 * Replaces fftw_plan_dft_r2c_2d + fftw_execute (ncs.c:438-439, 268). */
static void dft_r2c(ncsSubRegion *sr, const double *u) {
  int r = sr->r_size, fs = sr->fft_size;
  /* along the fast axis: w[i][l] */
  for (int i = 0; i < r; i++) {
    for (int l = 0; l < fs; l++) {
      double re = 0.0, im = 0.0;
      for (int j = 0; j < r; j++) {
        int m = (l * j) % r;
        re += u[i * r + j] * sr->cos_tab[m];
        im -= u[i * r + j] * sr->sin_tab[m];
      }
      sr->w_re[i * fs + l] = re;
      sr->w_im[i * fs + l] = im;
    }
  }
  /* along the slow axis */
  for (int k = 0; k < r; k++) {
    for (int l = 0; l < fs; l++) {
      double re = 0.0, im = 0.0;
      for (int i = 0; i < r; i++) {
        int m = (k * i) % r;
        double c = sr->cos_tab[m], s = sr->sin_tab[m];
        double a = sr->w_re[i * fs + l], b = sr->w_im[i * fs + l];
        /* (a + ib)(c - is) */
        re += a * c + b * s;
        im += b * c - a * s;
      }
      sr->u_fft_re[k * fs + l] = re;
      sr->u_fft_im[k * fs + l] = im;
    }
  }
}

/* Unnormalised c2r of the half spectrum G (r x fs) into g (r x r).
 * Replaces fftw_plan_dft_c2r_2d + fftw_execute (ncs.c:433-435, 244):
 * complex inverse DFT along the slow axis, then a 1-D c2r along the fast
 * axis that reads only the real part of the l=0 and l=r/2 bins (the
 * Hermitian extension of the half spectrum). */
static void dft_c2r(ncsSubRegion *sr) {
  int r = sr->r_size, fs = sr->fft_size, h = r / 2;
  for (int x = 0; x < r; x++) {
    for (int l = 0; l < fs; l++) {
      double re = 0.0, im = 0.0;
      for (int k = 0; k < r; k++) {
        int m = (k * x) % r;
        double c = sr->cos_tab[m], s = sr->sin_tab[m];
        double a = sr->g_fft_re[k * fs + l], b = sr->g_fft_im[k * fs + l];
        /* (a + ib)(c + is) */
        re += a * c - b * s;
        im += a * s + b * c;
      }
      sr->w_re[x * fs + l] = re;
      sr->w_im[x * fs + l] = im;
    }
  }
  for (int x = 0; x < r; x++) {
    for (int y = 0; y < r; y++) {
      double v = sr->w_re[x * fs + 0];
      v += sr->w_re[x * fs + h] * ((y & 1) ? -1.0 : 1.0);
      for (int l = 1; l < h; l++) {
        int m = (l * y) % r;
        double a = sr->w_re[x * fs + l], b = sr->w_im[x * fs + l];
        v += 2.0 * (a * sr->cos_tab[m] - b * sr->sin_tab[m]);
      }
      sr->g[x * r + y] = v;
    }
  }
}

/* ------------------------------------------------------------------ */
/* Sub-region functions (ncs.c:150-561).                               */
/* ------------------------------------------------------------------ */

/* ncs.c:158-175: 1 - (d+gamma)/max(u+gamma, FITMIN).  Inside the guard this
 * is deliberately not the derivative of the guarded LL (as in the C). */
void ncsSRCalcLLGradient(ncsSubRegion *sr, double *gradient) {
  int n = sr->r_size * sr->r_size;
  for (int i = 0; i < n; i++) {
    double t1 = sr->data[i] + sr->gamma[i];
    double t2 = sr->u[i] + sr->gamma[i];
    if (t2 < FITMIN) t2 = FITMIN;
    if (sr->lin && sr->lin[i] == 0.0) { gradient[i] = 0.0; continue; }
    gradient[i] = 1.0 - t1 / t2;
  }
}

/* ncs.c:185-208: sum u - (d+gamma)*ln(max(u+gamma, FITMIN)). */
double ncsSRCalcLogLikelihood(ncsSubRegion *sr) {
  int n = sr->r_size * sr->r_size;
  double sum = 0.0;
  for (int i = 0; i < n; i++) {
    double t1 = sr->data[i] + sr->gamma[i];
    double t2 = sr->u[i] + sr->gamma[i];
    if (t2 < FITMIN) t2 = FITMIN;
    if (sr->lin && sr->lin[i] == 0.0) continue;
    sum += sr->u[i] - t1 * log(t2);
  }
  return sum;
}

/* ncs.c:220-250: c2r of 2*q*U over the half spectrum, times 1/r^2.
 * Uses the spectrum left by ncsSRCalcNoiseContribution (ncs.c:214-216). */
void ncsSRCalcNCGradient(ncsSubRegion *sr, double *gradient) {
  int r = sr->r_size, fs = sr->fft_size;
  for (int i = 0; i < r; i++) {
    for (int j = 0; j < fs; j++) {
      int k = i * fs + j;
      double t1 = 2.0 * sr->otf_mask_sqr[i * r + j];
      sr->g_fft_re[k] = sr->u_fft_re[k] * t1;
      sr->g_fft_im[k] = sr->u_fft_im[k] * t1;
    }
  }
  dft_c2r(sr);
  for (int i = 0; i < r * r; i++) gradient[i] = sr->g[i] * sr->normalization;
}

/* ncs.c:259-299: r2c of u, then sum_{i,j} q[i][j] |U[i][j]|^2 over the full
 * r x r grid, reading the bins j >= r/2+1 from their Hermitian partner
 * ((r-i) mod r, r-j) of the half spectrum (the index formula at
 * ncs.c:281-291), times 1/r^2. */
double ncsSRCalcNoiseContribution(ncsSubRegion *sr) {
  int r = sr->r_size, fs = sr->fft_size;
  dft_r2c(sr, sr->u);
  double sum = 0.0;
  for (int i = 0; i < r; i++) {
    for (int j = 0; j < r; j++) {
      int l;
      if (j >= fs)
        l = ((r - i) % r) * fs + (r - j);
      else
        l = i * fs + j;
      double p = sr->u_fft_re[l] * sr->u_fft_re[l] +
                 sr->u_fft_im[l] * sr->u_fft_im[l];
      sum += p * sr->otf_mask_sqr[i * r + j];
    }
  }
  return sum * sr->normalization;
}

void ncsSRCleanup(ncsSubRegion *sr) {
  if (!sr) return;
  free(sr->data);
  free(sr);
}

void ncsSRGetU(ncsSubRegion *sr, double *u) {
  memcpy(u, sr->u, sizeof(double) * sr->r_size * sr->r_size);
}

/* ncs.c:394-446.  One allocation holds every array. */
ncsSubRegion *ncsSRInitialize(int r_size) {
  if (r_size <= 0 || (r_size % 2) != 0) {
    printf("ROI size must be divisible by 2!\n");
    return NULL;
  }
  ncsSubRegion *sr = (ncsSubRegion *)calloc(1, sizeof(ncsSubRegion));
  int n = r_size * r_size, fs = r_size / 2 + 1;
  size_t total = (size_t)7 * n + (size_t)4 * r_size * fs + 2 * (size_t)r_size +
                 2 * (size_t)n;
  double *blk = (double *)calloc(total, sizeof(double));
  sr->r_size = r_size;
  sr->fft_size = fs;
  sr->alpha = 0.0;
  sr->normalization = 1.0 / ((double)n);
  sr->data = blk; blk += n;
  sr->gamma = blk; blk += n;
  sr->otf_mask_sqr = blk; blk += n;
  sr->t1 = blk; blk += n;
  sr->t2 = blk; blk += n;
  sr->u = blk; blk += n;
  sr->g = blk; blk += n;
  sr->u_fft_re = blk; blk += r_size * fs;
  sr->u_fft_im = blk; blk += r_size * fs;
  sr->g_fft_re = blk; blk += r_size * fs;
  sr->g_fft_im = blk; blk += r_size * fs;
  sr->cos_tab = blk; blk += r_size;
  sr->sin_tab = blk; blk += r_size;
  sr->w_re = blk; blk += n;
  sr->w_im = blk; blk += n;
  for (int m = 0; m < r_size; m++) {
    double a = 2.0 * M_PI * (double)m / (double)r_size;
    sr->cos_tab[m] = cos(a);
    sr->sin_tab[m] = sin(a);
  }
  /* exact values at the quarter points so that the DFT of a constant is
   * exactly zero away from DC */
  sr->sin_tab[0] = 0.0;
  sr->cos_tab[r_size / 2] = -1.0;
  sr->sin_tab[r_size / 2] = 0.0;
  if ((r_size % 4) == 0) {
    sr->cos_tab[r_size / 4] = 0.0;
    sr->sin_tab[r_size / 4] = 1.0;
    sr->cos_tab[3 * r_size / 4] = 0.0;
    sr->sin_tab[3 * r_size / 4] = -1.0;
  }
  ncso_param_init(&sr->param);
  return sr;
}

void ncsSRNewRegion(ncsSubRegion *sr, double *image, double *gamma) {
  int n = sr->r_size * sr->r_size;
  memcpy(sr->data, image, sizeof(double) * n);
  memcpy(sr->gamma, gamma, sizeof(double) * n);
}

/* ncs.c:502-510: the mask is squared, FFT order expected. */
void ncsSRSetOTFMask(ncsSubRegion *sr, double *otf_mask) {
  int n = sr->r_size * sr->r_size;
  for (int i = 0; i < n; i++) sr->otf_mask_sqr[i] = otf_mask[i] * otf_mask[i];
}

void ncsSRSetU(ncsSubRegion *sr, double *u) {
  memcpy(sr->u, u, sizeof(double) * sr->r_size * sr->r_size);
}

void ncso_sr_set_param(ncsSubRegion *sr, const ncso_param *p) { sr->param = *p; }

void ncso_sr_last_stats(ncsSubRegion *sr, int *iterations, int *evaluations,
                        double *final_cost) {
  if (iterations) *iterations = sr->last_iters;
  if (evaluations) *evaluations = sr->last_evals;
  if (final_cost) *final_cost = sr->last_cost;
}

/* ncs.c:335-364: f = LL + alpha*NC, g = gLL + alpha*gNC.  The callback
 * evaluates at sr->u, which is the solver's x (ncs.c:553-557 pass sr->u). */
static double sr_evaluate(ncsSubRegion *sr, double *g) {
  int n = sr->r_size * sr->r_size;
  double fx = ncsSRCalcLogLikelihood(sr);
  fx += sr->alpha * ncsSRCalcNoiseContribution(sr);
  ncsSRCalcLLGradient(sr, sr->t1);
  ncsSRCalcNCGradient(sr, sr->t2);
  for (int i = 0; i < n; i++) g[i] = sr->t1[i] + sr->alpha * sr->t2[i];
  return fx;
}

/* ------------------------------------------------------------------ */
/* L-BFGS with the More-Thuente line search.                            */
/* Restates the published algorithms that liblbfgs 1.10 implements:    */
/* outer loop, stop rule and codes per lbfgs.h:198-500; line search per */
/* More & Thuente (1994), sections 2-4 (the safeguarded step of their   */
/* Fortran subroutines cvsrch/cstep).                                   */
/* ------------------------------------------------------------------ */

static double dot(const double *a, const double *b, int n) {
  double s = 0.0;
  for (int i = 0; i < n; i++) s += a[i] * b[i];
  return s;
}

static double max3(double a, double b, double c) {
  double m = a > b ? a : b;
  return m > c ? m : c;
}

/* One safeguarded step of More & Thuente (1994), section 4: update the
 * interval of uncertainty [stx, sty] with the trial (stp, fp, dp) and choose
 * the next trial from cubic / quadratic interpolants.  Returns 0 on
 * success, or a liblbfgs error code when the inputs are inconsistent. */
static int mt_step(double *stx, double *fx, double *dx, double *sty,
                   double *fy, double *dy, double *stp, double fp, double dp,
                   int *brackt, double stmin, double stmax) {
  double t = *stp;
  if (*brackt) {
    double lo = *stx < *sty ? *stx : *sty, hi = *stx > *sty ? *stx : *sty;
    if (t <= lo || hi <= t) return NCSO_LBFGSERR_OUTOFINTERVAL;
    if (0.0 <= *dx * (t - *stx)) return NCSO_LBFGSERR_INCREASEGRADIENT;
    if (stmax < stmin) return NCSO_LBFGSERR_INCORRECT_TMINMAX;
  }
  double sgnd = dp * (*dx / fabs(*dx));
  double stpf, stpc, stpq, theta, s, gam, p, q;
  int bound;
  if (fp > *fx) {
    /* case 1: higher value, the minimum is bracketed */
    bound = 1;
    theta = 3.0 * (*fx - fp) / (t - *stx) + *dx + dp;
    s = max3(fabs(theta), fabs(*dx), fabs(dp));
    gam = s * sqrt((theta / s) * (theta / s) - (*dx / s) * (dp / s));
    if (t < *stx) gam = -gam;
    p = (gam - *dx) + theta;
    q = ((gam - *dx) + gam) + dp;
    stpc = *stx + (p / q) * (t - *stx);
    stpq = *stx + ((*dx / ((*fx - fp) / (t - *stx) + *dx)) / 2.0) * (t - *stx);
    if (fabs(stpc - *stx) < fabs(stpq - *stx))
      stpf = stpc;
    else
      stpf = stpc + (stpq - stpc) / 2.0;
    *brackt = 1;
  } else if (sgnd < 0.0) {
    /* case 2: lower value, derivatives of opposite sign */
    bound = 0;
    theta = 3.0 * (*fx - fp) / (t - *stx) + *dx + dp;
    s = max3(fabs(theta), fabs(*dx), fabs(dp));
    gam = s * sqrt((theta / s) * (theta / s) - (*dx / s) * (dp / s));
    if (t > *stx) gam = -gam;
    p = (gam - dp) + theta;
    q = ((gam - dp) + gam) + *dx;
    stpc = t + (p / q) * (*stx - t);
    stpq = t + (dp / (dp - *dx)) * (*stx - t);
    stpf = (fabs(stpc - t) > fabs(stpq - t)) ? stpc : stpq;
    *brackt = 1;
  } else if (fabs(dp) < fabs(*dx)) {
    /* case 3: lower value, same sign, derivative magnitude decreases */
    bound = 1;
    theta = 3.0 * (*fx - fp) / (t - *stx) + *dx + dp;
    s = max3(fabs(theta), fabs(*dx), fabs(dp));
    double disc = (theta / s) * (theta / s) - (*dx / s) * (dp / s);
    gam = s * sqrt(disc > 0.0 ? disc : 0.0);
    if (t > *stx) gam = -gam;
    p = (gam - dp) + theta;
    q = (gam + (*dx - dp)) + gam;
    double rr = p / q;
    if (rr < 0.0 && gam != 0.0)
      stpc = t + rr * (*stx - t);
    else if (t > *stx)
      stpc = stmax;
    else
      stpc = stmin;
    stpq = t + (dp / (dp - *dx)) * (*stx - t);
    if (*brackt)
      stpf = (fabs(t - stpc) < fabs(t - stpq)) ? stpc : stpq;
    else
      stpf = (fabs(t - stpc) > fabs(t - stpq)) ? stpc : stpq;
  } else {
    /* case 4: lower value, same sign, derivative does not decrease */
    bound = 0;
    if (*brackt) {
      theta = 3.0 * (fp - *fy) / (*sty - t) + *dy + dp;
      s = max3(fabs(theta), fabs(*dy), fabs(dp));
      gam = s * sqrt((theta / s) * (theta / s) - (*dy / s) * (dp / s));
      if (t > *sty) gam = -gam;
      p = (gam - dp) + theta;
      q = ((gam - dp) + gam) + *dy;
      stpc = t + (p / q) * (*sty - t);
      stpf = stpc;
    } else if (t > *stx) {
      stpf = stmax;
    } else {
      stpf = stmin;
    }
  }
  /* update the interval of uncertainty */
  if (fp > *fx) {
    *sty = t; *fy = fp; *dy = dp;
  } else {
    if (sgnd < 0.0) {
      *sty = *stx; *fy = *fx; *dy = *dx;
    }
    *stx = t; *fx = fp; *dx = dp;
  }
  if (stpf > stmax) stpf = stmax;
  if (stpf < stmin) stpf = stmin;
  if (*brackt && bound) {
    double mq = *stx + 0.66 * (*sty - *stx);
    if (*stx < *sty) {
      if (mq < stpf) stpf = mq;
    } else {
      if (stpf < mq) stpf = mq;
    }
  }
  *stp = stpf;
  return 0;
}

/* More-Thuente line search along s from xp (More & Thuente 1994, sec. 2-3:
 * sufficient decrease with ftol, curvature with gtol, modified function in
 * the first stage, bisection safeguard when the interval shrinks slowly).
 * x, f, g are updated to the accepted point.  Returns the number of
 * evaluations (>0) or a negative liblbfgs code. */
static int mt_search(ncsSubRegion *sr, int n, double *x, double *f, double *g,
                     const double *s, double *stp, const double *xp,
                     const ncso_param *pr, int *nevals) {
  int count = 0, brackt = 0, stage1 = 1, uinfo = 0;
  if (*stp <= 0.0) return NCSO_LBFGSERR_INVALIDPARAMETERS;
  double dginit = dot(g, s, n);
  if (0.0 < dginit) return NCSO_LBFGSERR_INCREASEGRADIENT;
  double finit = *f, dgtest = pr->ftol * dginit;
  double width = pr->max_step - pr->min_step, prev_width = 2.0 * width;
  double stx = 0.0, sty = 0.0, fx = finit, fy = finit, dgx = dginit,
         dgy = dginit;
  for (;;) {
    double stmin, stmax;
    if (brackt) {
      stmin = stx < sty ? stx : sty;
      stmax = stx > sty ? stx : sty;
    } else {
      stmin = stx;
      stmax = *stp + 4.0 * (*stp - stx);
    }
    if (*stp < pr->min_step) *stp = pr->min_step;
    if (pr->max_step < *stp) *stp = pr->max_step;
    /* fall back to the best step when further progress is impossible */
    if ((brackt && ((*stp <= stmin || stmax <= *stp) ||
                    pr->max_linesearch <= count + 1 || uinfo != 0)) ||
        (brackt && (stmax - stmin <= pr->xtol * stmax))) {
      *stp = stx;
    }
    for (int i = 0; i < n; i++) x[i] = xp[i] + (*stp) * s[i];
    *f = sr_evaluate(sr, g);
    (*nevals)++;
    double dg = dot(g, s, n);
    double ftest1 = finit + (*stp) * dgtest;
    ++count;
    if (brackt && ((*stp <= stmin || stmax <= *stp) || uinfo != 0))
      return NCSO_LBFGSERR_ROUNDING_ERROR;
    if (*stp == pr->max_step && *f <= ftest1 && dg <= dgtest)
      return NCSO_LBFGSERR_MAXIMUMSTEP;
    if (*stp == pr->min_step && (ftest1 < *f || dgtest <= dg))
      return NCSO_LBFGSERR_MINIMUMSTEP;
    if (brackt && (stmax - stmin) <= pr->xtol * stmax)
      return NCSO_LBFGSERR_WIDTHTOOSMALL;
    if (pr->max_linesearch <= count) return NCSO_LBFGSERR_MAXIMUMLINESEARCH;
    if (*f <= ftest1 && fabs(dg) <= pr->gtol * (-dginit)) return count;
    double ftol_gtol = pr->ftol < pr->gtol ? pr->ftol : pr->gtol;
    if (stage1 && *f <= ftest1 && ftol_gtol * dginit <= dg) stage1 = 0;
    if (stage1 && ftest1 < *f && *f <= fx) {
      /* modified function psi(a) = f(a) - f(0) - ftol*a*f'(0) */
      double fm = *f - *stp * dgtest, fxm = fx - stx * dgtest,
             fym = fy - sty * dgtest;
      double dgm = dg - dgtest, dgxm = dgx - dgtest, dgym = dgy - dgtest;
      uinfo = mt_step(&stx, &fxm, &dgxm, &sty, &fym, &dgym, stp, fm, dgm,
                      &brackt, stmin, stmax);
      fx = fxm + stx * dgtest;
      fy = fym + sty * dgtest;
      dgx = dgxm + dgtest;
      dgy = dgym + dgtest;
    } else {
      uinfo = mt_step(&stx, &fx, &dgx, &sty, &fy, &dgy, stp, *f, dg, &brackt,
                      stmin, stmax);
    }
    if (brackt) {
      if (0.66 * prev_width <= fabs(sty - stx)) *stp = stx + 0.5 * (sty - stx);
      prev_width = width;
      width = fabs(sty - stx);
    }
  }
}

/* L-BFGS outer loop (lbfgs.h:432-435 stop rule; two-loop recursion). */
static int lbfgs_solve(ncsSubRegion *sr, const ncso_param *pr) {
  int n = sr->r_size * sr->r_size, m = pr->m;
  double *x = sr->u;
  double *buf = (double *)malloc(sizeof(double) * (size_t)(4 + 2 * m) * n +
                                 sizeof(double) * 3 * (size_t)m);
  double *g = buf, *xp = g + n, *gp = xp + n, *d = gp + n;
  double *S = d + n, *Y = S + (size_t)m * n;
  double *ys_h = Y + (size_t)m * n, *alpha_h = ys_h + m;
  int ret, k, end = 0, evals = 0;
  double fx, xnorm, gnorm, step;

  sr->last_iters = 0;
  fx = sr_evaluate(sr, g);
  evals++;
  for (int i = 0; i < n; i++) d[i] = -g[i];
  xnorm = sqrt(dot(x, x, n));
  gnorm = sqrt(dot(g, g, n));
  if (xnorm < 1.0) xnorm = 1.0;
  if (gnorm / xnorm <= pr->epsilon) {
    ret = NCSO_LBFGS_ALREADY_MINIMIZED;
    goto done;
  }
  step = 1.0 / sqrt(dot(d, d, n));
  k = 1;
  for (;;) {
    memcpy(xp, x, sizeof(double) * n);
    memcpy(gp, g, sizeof(double) * n);
    int ls = mt_search(sr, n, x, &fx, g, d, &step, xp, pr, &evals);
    if (ls < 0) {
      memcpy(x, xp, sizeof(double) * n);
      memcpy(g, gp, sizeof(double) * n);
      /* the cost at the restored point */
      fx = sr_evaluate(sr, g);
      ret = ls;
      sr->last_iters = k;
      goto done;
    }
    sr->last_iters = k;
    xnorm = sqrt(dot(x, x, n));
    gnorm = sqrt(dot(g, g, n));
    if (xnorm < 1.0) xnorm = 1.0;
    if (gnorm / xnorm <= pr->epsilon) {
      ret = NCSO_LBFGS_SUCCESS;
      break;
    }
    if (pr->max_iterations != 0 && pr->max_iterations < k + 1) {
      ret = NCSO_LBFGSERR_MAXIMUMITERATION;
      break;
    }
    /* newest correction pair into the ring */
    double *sk = S + (size_t)end * n, *yk = Y + (size_t)end * n;
    for (int i = 0; i < n; i++) {
      sk[i] = x[i] - xp[i];
      yk[i] = g[i] - gp[i];
    }
    double ys = dot(yk, sk, n), yy = dot(yk, yk, n);
    ys_h[end] = ys;
    int bound = (m <= k) ? m : k;
    ++k;
    end = (end + 1) % m;
    for (int i = 0; i < n; i++) d[i] = -g[i];
    int j = end;
    for (int t = 0; t < bound; t++) {
      j = (j + m - 1) % m; /* newest to oldest */
      double a = dot(S + (size_t)j * n, d, n) / ys_h[j];
      alpha_h[j] = a;
      const double *yj = Y + (size_t)j * n;
      for (int i = 0; i < n; i++) d[i] -= a * yj[i];
    }
    for (int i = 0; i < n; i++) d[i] *= ys / yy;
    for (int t = 0; t < bound; t++) { /* oldest to newest */
      double b = dot(Y + (size_t)j * n, d, n) / ys_h[j];
      const double *sj = S + (size_t)j * n;
      for (int i = 0; i < n; i++) d[i] += (alpha_h[j] - b) * sj[i];
      j = (j + 1) % m;
    }
    step = 1.0;
  }
done:
  sr->last_evals = evals;
  sr->last_cost = fx;
  free(buf);
  return ret;
}

/* ncs.c:539-561: u <- data (plain start), then L-BFGS. */
int ncsSRSolve(ncsSubRegion *sr, double alpha, int verbose) {
  (void)verbose;
  sr->alpha = alpha;
  memcpy(sr->u, sr->data, sizeof(double) * sr->r_size * sr->r_size);
  return lbfgs_solve(sr, &sr->param);
}

int ncso_sr_solve_from_u(ncsSubRegion *sr, double alpha) {
  sr->alpha = alpha;
  return lbfgs_solve(sr, &sr->param);
}

double ncso_cost_grad(const double *u, const double *data, const double *gamma,
                      const double *mask_fft, double alpha, int r,
                      double *grad) {
  ncsSubRegion *sr = ncsSRInitialize(r);
  if (!sr) return NAN;
  ncsSRNewRegion(sr, (double *)data, (double *)gamma);
  ncsSRSetOTFMask(sr, (double *)mask_fft);
  ncsSRSetU(sr, (double *)u);
  sr->alpha = alpha;
  double f;
  if (grad)
    f = sr_evaluate(sr, grad);
  else
    f = ncsSRCalcLogLikelihood(sr) + alpha * ncsSRCalcNoiseContribution(sr);
  ncsSRCleanup(sr);
  return f;
}

/* ------------------------------------------------------------------ */
/* Batches.                                                            */
/* ------------------------------------------------------------------ */

static int resolve_threads(int nthreads) {
#ifdef _OPENMP
  return nthreads > 0 ? nthreads : omp_get_max_threads();
#else
  (void)nthreads;
  return 1;
#endif
}

int ncso_solve_tiles(const double *data, const double *gamma,
                     const double *u0, const double *mask_fft, double alpha,
                     int r, int ntiles, const ncso_param *p, int nthreads,
                     double *u_out, double *cost_out, int *status_out,
                     int *iters_out, int *evals_out) {
  if (r <= 0 || (r % 2) != 0 || ntiles < 0) return -1;
  int n = r * r;
  ncso_param prm;
  if (p) prm = *p; else ncso_param_init(&prm);
  int nt = resolve_threads(nthreads);
#pragma omp parallel num_threads(nt)
  {
    ncsSubRegion *sr = ncsSRInitialize(r);
    ncsSRSetOTFMask(sr, (double *)mask_fft);
    sr->param = prm;
#pragma omp for schedule(dynamic, 4)
    for (int t = 0; t < ntiles; t++) {
      ncsSRNewRegion(sr, (double *)(data + (size_t)t * n),
                     (double *)(gamma + (size_t)t * n));
      int st;
      if (u0) {
        ncsSRSetU(sr, (double *)(u0 + (size_t)t * n));
        st = ncso_sr_solve_from_u(sr, alpha);
      } else {
        st = ncsSRSolve(sr, alpha, 0);
      }
      if (u_out) memcpy(u_out + (size_t)t * n, sr->u, sizeof(double) * n);
      if (cost_out) cost_out[t] = sr->last_cost;
      if (status_out) status_out[t] = st;
      if (iters_out) iters_out[t] = sr->last_iters;
      if (evals_out) evals_out[t] = sr->last_evals;
    }
    ncsSRCleanup(sr);
  }
  return 0;
}

int ncso_solve_tiles_f32(const float *data, const float *gamma,
                         const float *u0, const float *mask_fft, double alpha,
                         int r, int ntiles, const ncso_param *p, int nthreads,
                         float *u_out, double *cost_out, int *status_out,
                         int *iters_out, int *evals_out) {
  if (r <= 0 || (r % 2) != 0 || ntiles < 0) return -1;
  int n = r * r;
  ncso_param prm;
  if (p) prm = *p; else ncso_param_init(&prm);
  double mask[1024];
  if (n > 1024) return -1;
  for (int i = 0; i < n; i++) mask[i] = mask_fft[i];
  int nt = resolve_threads(nthreads);
#pragma omp parallel num_threads(nt)
  {
    ncsSubRegion *sr = ncsSRInitialize(r);
    ncsSRSetOTFMask(sr, mask);
    sr->param = prm;
#pragma omp for schedule(dynamic, 4)
    for (int t = 0; t < ntiles; t++) {
      for (int i = 0; i < n; i++) {
        sr->data[i] = data[(size_t)t * n + i];
        sr->gamma[i] = gamma[(size_t)t * n + i];
      }
      int st;
      if (u0) {
        for (int i = 0; i < n; i++) sr->u[i] = u0[(size_t)t * n + i];
        st = ncso_sr_solve_from_u(sr, alpha);
      } else {
        st = ncsSRSolve(sr, alpha, 0);
      }
      if (u_out)
        for (int i = 0; i < n; i++) u_out[(size_t)t * n + i] = (float)sr->u[i];
      if (cost_out) cost_out[t] = sr->last_cost;
      if (status_out) status_out[t] = st;
      if (iters_out) iters_out[t] = sr->last_iters;
      if (evals_out) evals_out[t] = sr->last_evals;
    }
    ncsSRCleanup(sr);
  }
  return 0;
}

void ncso_eval_cost_f32(const float *u, const float *data, const float *gamma,
                        const float *mask_fft, double alpha, int r, int ntiles,
                        int nthreads, double *cost_out, double *gnorm_out,
                        double *unorm_out) {
  int n = r * r;
  double mask[1024];
  if (n > 1024 || (r % 2) != 0) return;
  for (int i = 0; i < n; i++) mask[i] = mask_fft[i];
  int nt = resolve_threads(nthreads);
#pragma omp parallel num_threads(nt)
  {
    ncsSubRegion *sr = ncsSRInitialize(r);
    ncsSRSetOTFMask(sr, mask);
    double *g = (double *)malloc(sizeof(double) * n);
#pragma omp for schedule(static)
    for (int t = 0; t < ntiles; t++) {
      for (int i = 0; i < n; i++) {
        sr->data[i] = data[(size_t)t * n + i];
        sr->gamma[i] = gamma[(size_t)t * n + i];
        sr->u[i] = u[(size_t)t * n + i];
      }
      sr->alpha = alpha;
      double f = sr_evaluate(sr, g);
      if (cost_out) cost_out[t] = f;
      if (gnorm_out) gnorm_out[t] = sqrt(dot(g, g, n));
      if (unorm_out) unorm_out[t] = sqrt(dot(sr->u, sr->u, n));
    }
    free(g);
    ncsSRCleanup(sr);
  }
}

/* Tiles whose pixels may be excluded from the likelihood (lin = 0): the
 * zero-padded edge blocks of the non-overlapping block tiling.  lin is
 * [ntiles][r*r] with 0/1 entries; padded pixels stay variables of the NC
 * term only.  Not a reference feature: the reference always replicates
 * edges (ncs.c:87-113). */
int ncso_solve_tiles_masked_f32(const float *data, const float *gamma,
                                const float *lin, const float *mask_fft,
                                double alpha, int r, int ntiles,
                                const ncso_param *p, int nthreads, float *u_out,
                                double *cost_out, int *status_out,
                                int *iters_out) {
  if (r <= 0 || (r % 2) != 0 || ntiles < 0 || r * r > 1024) return -1;
  int n = r * r;
  ncso_param prm;
  if (p) prm = *p; else ncso_param_init(&prm);
  double mask[1024];
  for (int i = 0; i < n; i++) mask[i] = mask_fft[i];
  int nt = resolve_threads(nthreads);
#pragma omp parallel num_threads(nt)
  {
    ncsSubRegion *sr = ncsSRInitialize(r);
    ncsSRSetOTFMask(sr, mask);
    sr->param = prm;
    double *w = (double *)malloc(sizeof(double) * n);
    sr->lin = w;
#pragma omp for schedule(dynamic, 4)
    for (int t = 0; t < ntiles; t++) {
      for (int i = 0; i < n; i++) {
        sr->data[i] = data[(size_t)t * n + i];
        sr->gamma[i] = gamma[(size_t)t * n + i];
        w[i] = lin[(size_t)t * n + i];
      }
      int st = ncsSRSolve(sr, alpha, 0);
      if (u_out)
        for (int i = 0; i < n; i++) u_out[(size_t)t * n + i] = (float)sr->u[i];
      if (cost_out) cost_out[t] = sr->last_cost;
      if (status_out) status_out[t] = st;
      if (iters_out) iters_out[t] = sr->last_iters;
    }
    sr->lin = NULL;
    free(w);
    ncsSRCleanup(sr);
  }
  return 0;
}

/* ------------------------------------------------------------------ */
/* Whole-image tiling (ncs.c:41-147).                                   */
/* ------------------------------------------------------------------ */

/* Window origins along one axis of size im: i = -1, -1+s, ...; a window
 * that would pass im+1 is clamped to im+1-r and ends the loop
 * (ncs.c:71-78, 136-142). */
static int axis_origins(int im, int r, int *out) {
  int s = r - 2, cnt = 0;
  for (int i = -1; i < im + 1; i += s) {
    int b = (i + r > im + 1) ? im + 1 - r : i;
    if (out) out[cnt] = b;
    cnt++;
    if (b == im + 1 - r) break;
  }
  return cnt;
}

int ncso_tiles_per_axis(int n, int r_size) {
  if (r_size <= 2 || (r_size % 2) != 0) return -1;
  return axis_origins(n, r_size, NULL);
}

/* Gather with edge clamp (ncs.c:87-113). */
static void gather(const double *image, const double *gamma, int im_x,
                   int im_y, int r, int bx, int by, double *d, double *gm) {
  for (int k = 0; k < r; k++) {
    int x = k + bx;
    if (x < 0) x = 0;
    if (x >= im_x) x = im_x - 1;
    for (int n = 0; n < r; n++) {
      int y = n + by;
      if (y < 0) y = 0;
      if (y >= im_y) y = im_y - 1;
      d[k * r + n] = image[(size_t)x * im_y + y];
      gm[k * r + n] = gamma[(size_t)x * im_y + y];
    }
  }
}

/* Scatter interior rows/cols 1..r-2 (ncs.c:122-130). */
static void scatter(double *ncs_image, int im_y, int r, int bx, int by,
                    const double *u) {
  for (int k = 1; k < r - 1; k++)
    for (int n = 1; n < r - 1; n++)
      ncs_image[(size_t)(k + bx) * im_y + (n + by)] = u[k * r + n];
}

void ncsReduceNoise(double *ncs_image, double *image, double *gamma,
                    double *otf_mask, double alpha, int im_x, int im_y,
                    int r_size) {
  if ((r_size % 2) != 0) {
    printf("OTF mask size of %d is not divisible by 2!", r_size);
    return;
  }
  ncso_reduce_noise(ncs_image, image, gamma, otf_mask, alpha, im_x, im_y,
                    r_size, NULL, 1, NULL, NULL, NULL);
}

int ncso_reduce_noise(double *ncs_image, const double *image,
                      const double *gamma, const double *mask_fft,
                      double alpha, int im_x, int im_y, int r_size,
                      const ncso_param *p, int nthreads, double *tile_cost,
                      int *tile_status, int *tile_iters) {
  if (r_size <= 2 || (r_size % 2) != 0 || im_x < r_size - 2 ||
      im_y < r_size - 2)
    return -1;
  int nx = axis_origins(im_x, r_size, NULL), ny = axis_origins(im_y, r_size, NULL);
  int *ox = (int *)malloc(sizeof(int) * (nx + ny));
  int *oy = ox + nx;
  axis_origins(im_x, r_size, ox);
  axis_origins(im_y, r_size, oy);
  int ntiles = nx * ny, n = r_size * r_size;
  double *d = (double *)malloc(sizeof(double) * (size_t)ntiles * n * 3);
  double *gm = d + (size_t)ntiles * n, *u = gm + (size_t)ntiles * n;
  for (int a = 0; a < nx; a++)
    for (int b = 0; b < ny; b++) {
      size_t t = (size_t)a * ny + b;
      gather(image, gamma, im_x, im_y, r_size, ox[a], oy[b], d + t * n,
             gm + t * n);
    }
  ncso_solve_tiles(d, gm, NULL, mask_fft, alpha, r_size, ntiles, p, nthreads,
                   u, tile_cost, tile_status, tile_iters, NULL);
  /* stitch in raster order: later windows overwrite (last writer wins) */
  for (int a = 0; a < nx; a++)
    for (int b = 0; b < ny; b++) {
      size_t t = (size_t)a * ny + b;
      scatter(ncs_image, im_y, r_size, ox[a], oy[b], u + t * n);
    }
  free(d);
  free(ox);
  return ntiles;
}
