// ncs_solver.hip -- persistent per-tile L-BFGS solver of the NCS cost on gfx950.
//
// One wave64 owns one 16x16 tile at a time and keeps the whole solver state
// (data, gamma, iterate, gradient, NC gradient, search direction and the
// L-BFGS (s, y) history) in VGPRs: lane l = 16*g + j holds pixels
// (4g + r, j), r = 0..3 -- the accumulator layout of
// v_mfma_f32_16x16x4_f32.  Waves pull tiles from a device-side atomic work
// queue, so tiles of every frame of a batch (any sizes, ROIs, per-frame
// alpha) share one launch and finish at their own iteration counts.
//
// Cost per tile (clib/pyCNCS/ncs.c:335-364, opencl/ncs.cl:875-935):
//   f(u) = sum_p [ lin_p u_p - c_p ln max(u_p + gamma_p, 1e-6) ] + alpha NC(u)
//   NC(u) = (1/N^2) sum_kl q_kl |DFT2(u)_kl|^2,   q = mask^2 (FFT order)
// with c = data + gamma and lin = 1 (lin = c = 0 on zero-padded pixels).
//
// NC on the matrix cores.  With R the orthonormal real DFT basis
// (rows: cos k, k = 0..8, then sin k, k = 1..7) and Z = R X R^T,
// NC = sum_ab qe(ab) Z_ab^2 + sum over (cc,ss)/(cs,sc) partner pairs
// (the odd-odd part of q, zero for masks even in each axis), so
//   grad NC = R^T W R,  W = 2 qe * Z + qo * partner(Z),
// four 16x16x16 FP32 GEMMs = 16 MFMA 16x16x4 per application, every operand
// taken straight from the register it sits in (no LDS, no shuffles; the
// general-mask partner is one lane permutation, lane ^ 40).  The evaluation
// kernel works this way.  The solver keeps its state in the R basis, where
// K is diagonal, and pays only the basis changes: 4 MFMA 16x16x32 f16 each
// on an FP16 hi/lo split of both operands, or 8 FP32 MFMA where FP16 is not
// accurate enough (a non-finite result, or a gradient near the stop rule).
//
// NC is quadratic, so along a search direction d:
//   NC(u + t d) = NC(u) + t (gnc . d) + t^2/2 (d . K d),  gnc(u + t d) = gnc + t K d.
// Each iteration applies K once (to d); line-search trials only re-evaluate
// the likelihood, and the decrease f(u + t d) - f(u) is formed from log1p of
// the relative change, which keeps the Armijo / Wolfe tests meaningful in
// FP32 down to the gradient tolerance.
#include "ncs_common.h"

namespace ncs {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef unsigned int u4 __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------
// Wave primitives
// ---------------------------------------------------------------------------

template <int CTRL, int RM = 0xF, int BM = 0xF>
__device__ __forceinline__ float dppmov(float v) {
  return __int_as_float(
      __builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, RM, BM, false));
}

// Sum of N values over the 64 lanes; every lane receives the totals.
// Four values share one register: two levels of gfx950 row swaps
// (permlane16_swap exchanges odd rows of one register with even rows of
// another, permlane32_swap the upper half of one with the lower half of the
// other) and adds leave row k holding value k's lane-wise sums over the four
// rows; four DPP butterflies within the row (quad_perm, half-mirror, mirror)
// finish every row at once, and row k's total is read from lane 16 k.
__device__ __forceinline__ void swap_add16(float a, float b, float &s) {
  const auto p = __builtin_amdgcn_permlane16_swap(__float_as_uint(a), __float_as_uint(b),
                                                  false, false);
  s = __uint_as_float(p[0]) + __uint_as_float(p[1]);
}

template <int N>
__device__ __forceinline__ void wave_sum(float (&x)[N]) {
#pragma unroll
  for (int g0 = 0; g0 < N; g0 += 4) {
    const int n = N - g0 < 4 ? N - g0 : 4;
    float s01, s23 = 0.0f;
    swap_add16(x[g0], n > 1 ? x[g0 + 1] : 0.0f, s01);  // rows [a01, b01, a23, b23]
    if (n > 2) swap_add16(x[g0 + 2], n > 3 ? x[g0 + 3] : 0.0f, s23);
    const auto q = __builtin_amdgcn_permlane32_swap(__float_as_uint(s01), __float_as_uint(s23),
                                                    false, false);
    float s = __uint_as_float(q[0]) + __uint_as_float(q[1]);  // row k: value k
    s += dppmov<0xB1>(s);
    s += dppmov<0x4E>(s);
    s += dppmov<0x141>(s);
    s += dppmov<0x140>(s);
#pragma unroll
    for (int k = 0; k < n; k++)
      x[g0 + k] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(s), 16 * k));
  }
}

// Largest of non-negative per-lane values over the wave (uniform): DPP
// butterflies within rows, then the two row swaps.
__device__ __forceinline__ float wave_max(float v) {
  v = fmaxf(v, dppmov<0xB1>(v));
  v = fmaxf(v, dppmov<0x4E>(v));
  v = fmaxf(v, dppmov<0x141>(v));
  v = fmaxf(v, dppmov<0x140>(v));
  const auto p = __builtin_amdgcn_permlane16_swap(__float_as_uint(v), __float_as_uint(v),
                                                  false, false);
  v = fmaxf(__uint_as_float(p[0]), __uint_as_float(p[1]));
  const auto q = __builtin_amdgcn_permlane32_swap(__float_as_uint(v), __float_as_uint(v),
                                                  false, false);
  v = fmaxf(__uint_as_float(q[0]), __uint_as_float(q[1]));
  return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}

// Exact power-of-two scale for the FP16 basis changes of a tile's
// data-like operands (iterates, directions, gamma, the learned
// correction): it brings the tile's largest |data| into [32, 64), which
// leaves FP16 a factor of about 200 of headroom (an overflow beyond it
// falls back to FP32), and values down to about 1/400 of the maximum keep
// a normal FP16 low half.  Zero, subnormal or non-finite maxima give 1.
__device__ __forceinline__ float tile_scale(const float (&x)[4]) {
  const float m = wave_max(
      fmaxf(fmaxf(fabsf(x[0]), fabsf(x[1])), fmaxf(fabsf(x[2]), fabsf(x[3]))));
  const int e = (int)(__float_as_uint(m) >> 23);  // biased exponent
  // 2^(132 - e) has biased exponent 259 - e
  return (e >= 5 && e <= 254) ? __uint_as_float((unsigned int)(259 - e) << 23) : 1.0f;
}

__device__ __forceinline__ float wave_dot(const float (&a)[4], const float (&b)[4]) {
  float s[1] = {a[0] * b[0] + a[1] * b[1] + (a[2] * b[2] + a[3] * b[3])};
  wave_sum<1>(s);
  return s[0];
}

__device__ __forceinline__ float rcpf(float x) { return __builtin_amdgcn_rcpf(x); }

// Natural log of a normal, positive float on the transcendental unit:
// v_log_f32 (log2) times ln 2, two instructions.  The libm-accurate logf
// adds denormal rescaling and an extended-precision ln 2 product (about a
// dozen instructions) that the solver's arguments (>= 1e-6) never need.
constexpr float kLn2 = 0.693147182f;
__device__ __forceinline__ float ln_hw(float x) {
  return __builtin_amdgcn_logf(x) * kLn2;
}

// log(1 + x) for x > -1, branch-free: the rounding error of u = 1 + x is
// added back to log(u), log1p(x) = log(u) + (x - (u - 1)) / u (the
// correction term of Goldberg's log1p), with 1/u taken as 2 - u: the
// correction is below half an ulp of 1, so the x^2 error of that factor is
// negligible.
__device__ __forceinline__ float log1p_acc(float x) {
  const float u = 1.0f + x;
  return fmaf(x - (u - 1.0f), 2.0f - u, ln_hw(u));
}

// ---------------------------------------------------------------------------
// NC gradient operator on the matrix cores
// ---------------------------------------------------------------------------

// Per-lane tables, 16 floats per lane: R[lane%16][4g + r], R[4g + r][lane%16],
// 2 * the even-even part of q at Z position (lane%16, 4g + r), and +-2 * the
// odd-odd part coupling to the partner lane ^ 40.
struct MaskTables {  // held in VGPRs
  f4 ka_, kb_, qe_, qo_;
  __device__ f4 ka() const { return ka_; }
  __device__ f4 kb() const { return kb_; }
  __device__ f4 qe() const { return qe_; }
  __device__ f4 qo() const { return qo_; }
};

// Lane 0's value (the DC coefficient of a spectral tile sits there).
__device__ __forceinline__ float lane0(float x) {
  return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(x)));
}

// The solver's tables, read from LDS at each use (one copy per workgroup,
// which saves 12-16 VGPRs per lane across the main loop).  The FP16 build
// (F16) holds the mask weights, the learned linear start's coefficients and
// the basis entries as FP16 hi/lo pairs, each duplicated into slots r and
// 4 + r of an 8-wide operand (table entries kSolveTableFirst..); the FP32
// build holds the FP32 basis entries, the mask weights and the linear
// start's coefficients (table entries 0..5).
template <bool F16>
struct LdsTables {
  static constexpr bool kF16 = F16;
  static constexpr int kFirst = F16 ? kSolveTableFirst : 0;  // first table entry held
  static constexpr int kCount = F16 ? kSolveTableF4 : 6;     // entries held
  static constexpr int o = F16 ? 0 : 2;
  const f4 *p;  // this lane's kCount entries
  __device__ f4 ka() const { return p[0]; }  // FP32 build only
  __device__ f4 kb() const { return p[1]; }  // FP32 build only
  __device__ f4 qe() const { return p[o + 0]; }
  __device__ f4 qo() const { return p[o + 1]; }
  __device__ f4 la() const { return p[o + 2]; }
  __device__ f4 lb() const { return p[o + 3]; }
  __device__ h8 kah() const { return __builtin_bit_cast(h8, p[4]); }  // FP16 build only
  __device__ h8 kal() const { return __builtin_bit_cast(h8, p[5]); }
  __device__ h8 kbh() const { return __builtin_bit_cast(h8, p[6]); }
  __device__ h8 kbl() const { return __builtin_bit_cast(h8, p[7]); }
};

// The FP32 basis entries, read from the global table when the FP32 basis
// changes run (past the precision switch, or as the FP16 path's fallback).
// The lane offset is formed where it is used, so no VGPRs stay reserved for
// it across the solver's loop.
struct GlobalBasis {
  const float *tables;  // uniform base of the per-lane table
  __device__ const f4 *lane_p() const {
    int ln = __lane_id();
    asm volatile("" : "+v"(ln));
    return reinterpret_cast<const f4 *>(tables + ln * kTableLaneFloats);
  }
  __device__ f4 ka() const { return lane_p()[0]; }
  __device__ f4 kb() const { return lane_p()[1]; }
};

// Real-DFT coefficients of a tile: Z^T = (R X R^T)^T in the MFMA
// accumulator layout (the transform is orthonormal, so dot products are
// the same in either basis).  Two GEMM stages, 8 MFMA 16x16x4.
template <class TB>
__device__ __forceinline__ void to_spec(const float (&x)[4], const TB &T, float (&z)[4]) {
  const f4 ka = T.ka();
  f4 z1 = {0.f, 0.f, 0.f, 0.f};
  // (R X)^T = X^T R^T : data as the A operand
  z1 = __builtin_amdgcn_mfma_f32_16x16x4f32(x[0], ka[0], z1, 0, 0, 0);
  z1 = __builtin_amdgcn_mfma_f32_16x16x4f32(x[1], ka[1], z1, 0, 0, 0);
  z1 = __builtin_amdgcn_mfma_f32_16x16x4f32(x[2], ka[2], z1, 0, 0, 0);
  z1 = __builtin_amdgcn_mfma_f32_16x16x4f32(x[3], ka[3], z1, 0, 0, 0);
  f4 zz = {0.f, 0.f, 0.f, 0.f};
  // Z^T = R (R X)^T : data as the B operand
  zz = __builtin_amdgcn_mfma_f32_16x16x4f32(ka[0], z1[0], zz, 0, 0, 0);
  zz = __builtin_amdgcn_mfma_f32_16x16x4f32(ka[1], z1[1], zz, 0, 0, 0);
  zz = __builtin_amdgcn_mfma_f32_16x16x4f32(ka[2], z1[2], zz, 0, 0, 0);
  zz = __builtin_amdgcn_mfma_f32_16x16x4f32(ka[3], z1[3], zz, 0, 0, 0);
#pragma unroll
  for (int r = 0; r < 4; r++) z[r] = zz[r];
}

// Inverse of to_spec: X = R^T W R back in the tile layout.  8 MFMA.
template <class TB>
__device__ __forceinline__ void to_pix(const float (&w)[4], const TB &T, float (&x)[4]) {
  const f4 kb = T.kb();
  f4 e = {0.f, 0.f, 0.f, 0.f};
  // W R : W^T held, so data as the A operand
  e = __builtin_amdgcn_mfma_f32_16x16x4f32(w[0], kb[0], e, 0, 0, 0);
  e = __builtin_amdgcn_mfma_f32_16x16x4f32(w[1], kb[1], e, 0, 0, 0);
  e = __builtin_amdgcn_mfma_f32_16x16x4f32(w[2], kb[2], e, 0, 0, 0);
  e = __builtin_amdgcn_mfma_f32_16x16x4f32(w[3], kb[3], e, 0, 0, 0);
  f4 g = {0.f, 0.f, 0.f, 0.f};
  // R^T (W R)
  g = __builtin_amdgcn_mfma_f32_16x16x4f32(kb[0], e[0], g, 0, 0, 0);
  g = __builtin_amdgcn_mfma_f32_16x16x4f32(kb[1], e[1], g, 0, 0, 0);
  g = __builtin_amdgcn_mfma_f32_16x16x4f32(kb[2], e[2], g, 0, 0, 0);
  g = __builtin_amdgcn_mfma_f32_16x16x4f32(kb[3], e[3], g, 0, 0, 0);
#pragma unroll
  for (int r = 0; r < 4; r++) x[r] = g[r];
}

// ---------------------------------------------------------------------------
// The same basis changes on the FP16 matrix cores.  FP32 MFMA shares the
// SIMD's vector datapath (it does not overlap VALU work), FP16 MFMA runs on
// the matrix cores at 8x the work per instruction and partly overlaps it.
// Each FP32 operand x becomes an FP16 pair hi = rn(x), lo = rn(x - hi)
// (x - hi is exact), so hi + lo carries 22 significant bits; with the basis
// split the same way and FP32 accumulation, one 16x16x32 MFMA takes
// [hi | lo] of 16 K values against one basis half, and two MFMA per GEMM
// stage give the product to about FP32 accuracy.  The K slot order (slots
// r and 4 + r of lane group g are tile row 4 g + r) is the same in both
// operands, which is all the contraction needs.  Operands must lie well
// inside FP16's range: callers scale by an exact power of two, and a
// non-finite result (overflow, or a non-finite input) is recomputed in FP32.
// ---------------------------------------------------------------------------

// Relative gradient norm below which the solver's basis changes switch from
// the FP16 split (about 2^-22 per operand) to FP32: the default tolerance
// (1e-5) stops at it, tighter tolerances finish in FP32.
constexpr float kF16Floor = 1.0e-5f;

// Bound on the error one incremental update adds to the spectral NC
// gradient (K z += t K d), relative to |K| max_j |u_j|: the fma's rounding
// (2^-24 of |K z| + |t K d|) and the mismatch between the spectral step t d
// and the pixel step its basis change produced (below 1e-6 |t d| in FP32
// and in the FP16 split), with |t d| = |u_new - u_old| <= 2 max_j |u_j|:
// 2.2e-6, taken with a factor of about 2 to spare.
constexpr float kNcDrift = 4.0e-6f;

// [rn(x0..3), rn(x0..3 - hi)] as one 8-wide FP16 operand: 8 VALU.  The
// packed conversions are compiler code on purpose: they are the first
// readers of x, which often is an MFMA result, and the compiler inserts the
// MFMA-to-VALU wait states only before instructions it can see.  Only the
// residual x - hi is inline asm (v_fma_mix_f32 takes the packed FP16 hi
// directly; the compiler would convert it back to FP32 first), and it
// depends on the conversion, so it always comes after those waits.
typedef _Float16 h2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ h8 split_f16(float x0, float x1, float x2, float x3) {
  const h2 h01 = {(_Float16)x0, (_Float16)x1};
  const h2 h23 = {(_Float16)x2, (_Float16)x3};
  const unsigned int u01 = __builtin_bit_cast(unsigned int, h01);
  const unsigned int u23 = __builtin_bit_cast(unsigned int, h23);
  float r0, r1, r2, r3;
  asm("v_fma_mix_f32 %0, -%1, 1.0, %2 op_sel_hi:[1,0,0]" : "=v"(r0) : "v"(u01), "v"(x0));
  asm("v_fma_mix_f32 %0, -%1, 1.0, %2 op_sel:[1,0,0] op_sel_hi:[1,0,0]"
      : "=v"(r1) : "v"(u01), "v"(x1));
  asm("v_fma_mix_f32 %0, -%1, 1.0, %2 op_sel_hi:[1,0,0]" : "=v"(r2) : "v"(u23), "v"(x2));
  asm("v_fma_mix_f32 %0, -%1, 1.0, %2 op_sel:[1,0,0] op_sel_hi:[1,0,0]"
      : "=v"(r3) : "v"(u23), "v"(x3));
  const h2 l01 = {(_Float16)r0, (_Float16)r1};
  const h2 l23 = {(_Float16)r2, (_Float16)r3};
  return (h8){h01[0], h01[1], h23[0], h23[1], l01[0], l01[1], l23[0], l23[1]};
}

// Every lane's four values finite (uniform).
__device__ __forceinline__ bool all_finite(const float (&z)[4]) {
  const bool ok = __builtin_isfinite(z[0]) & __builtin_isfinite(z[1]) &
                  __builtin_isfinite(z[2]) & __builtin_isfinite(z[3]);
  return __all(ok);
}

// to_spec on the FP16 matrix cores: z = to_spec(x * s) / s (s a power of
// two).  4 MFMA 16x16x32 f16; in FP32 (8 MFMA 16x16x4 f32, basis from G)
// when `exact` (uniform) is set, and always in the FP32 build.
template <bool CHECK = true, class TT, class GB>
__device__ __forceinline__ void to_spec16(const float (&x)[4], const TT &T, const GB &G,
                                          float s, float (&z)[4], bool exact) {
  if (!TT::kF16 || exact) {
    to_spec(x, G, z);
    return;
  }
  const h8 th = T.kah(), tl = T.kal();
  const h8 xs = split_f16(x[0] * s, x[1] * s, x[2] * s, x[3] * s);
  f4 y = {0.f, 0.f, 0.f, 0.f};
  // (R X)^T = X^T R^T : data as the A operand
  y = __builtin_amdgcn_mfma_f32_16x16x32_f16(xs, th, y, 0, 0, 0);
  y = __builtin_amdgcn_mfma_f32_16x16x32_f16(xs, tl, y, 0, 0, 0);
  const h8 ys = split_f16(y[0], y[1], y[2], y[3]);
  f4 zz = {0.f, 0.f, 0.f, 0.f};
  // Z^T = R (R X)^T : data as the B operand
  zz = __builtin_amdgcn_mfma_f32_16x16x32_f16(th, ys, zz, 0, 0, 0);
  zz = __builtin_amdgcn_mfma_f32_16x16x32_f16(tl, ys, zz, 0, 0, 0);
  const float is = 1.0f / s;
#pragma unroll
  for (int r = 0; r < 4; r++) z[r] = zz[r] * is;
  if (CHECK && !all_finite(z)) to_spec(x, G, z);
}

// to_pix on the FP16 matrix cores: x = to_pix(w * s) / s.  4 MFMA; in
// FP32 when `exact` is set, and always in the FP32 build.
template <bool CHECK = true, class TT, class GB>
__device__ __forceinline__ void to_pix16(const float (&w)[4], const TT &T, const GB &G,
                                         float s, float (&x)[4], bool exact) {
  if (!TT::kF16 || exact) {
    to_pix(w, G, x);
    return;
  }
  const h8 th = T.kbh(), tl = T.kbl();
  const h8 ws = split_f16(w[0] * s, w[1] * s, w[2] * s, w[3] * s);
  f4 e = {0.f, 0.f, 0.f, 0.f};
  // W R : W^T held, so data as the A operand
  e = __builtin_amdgcn_mfma_f32_16x16x32_f16(ws, th, e, 0, 0, 0);
  e = __builtin_amdgcn_mfma_f32_16x16x32_f16(ws, tl, e, 0, 0, 0);
  const h8 es = split_f16(e[0], e[1], e[2], e[3]);
  f4 g = {0.f, 0.f, 0.f, 0.f};
  // R^T (W R)
  g = __builtin_amdgcn_mfma_f32_16x16x32_f16(th, es, g, 0, 0, 0);
  g = __builtin_amdgcn_mfma_f32_16x16x32_f16(tl, es, g, 0, 0, 0);
  const float is = 1.0f / s;
#pragma unroll
  for (int r = 0; r < 4; r++) x[r] = g[r] * is;
  if (CHECK && !all_finite(x)) to_pix(w, G, x);
}

// The NC Hessian K in the spectral basis: the diagonal 2 q_ee plus, for a
// mask with an odd-odd part (GEN), the coupling to the partner lane ^ 40.
template <bool GEN, class TB>
__device__ __forceinline__ void nc_spec(const float (&z)[4], const TB &T, float (&out)[4]) {
  const f4 qe = T.qe();
  float w[4];
#pragma unroll
  for (int r = 0; r < 4; r++) w[r] = qe[r] * z[r];
  if constexpr (GEN) {
    const f4 qo = T.qo();
#pragma unroll
    for (int r = 0; r < 4; r++) w[r] = fmaf(qo[r], __shfl_xor(z[r], 40), w[r]);
  }
#pragma unroll
  for (int r = 0; r < 4; r++) out[r] = w[r];
}

// gnc = K x in the tile layout (the evaluation kernel's NC gradient).
template <bool GEN, class TB>
__device__ __forceinline__ void apply_k(const float (&x)[4], const TB &T, float (&out)[4]) {
  float z[4], w[4];
  to_spec(x, T, z);
  nc_spec<GEN>(z, T, w);
  to_pix(w, T, out);
}

// x <- H0 x in the spectral basis, H0 = (kappa I + alpha K)^-1: the inverse
// of the Hessian with the likelihood part replaced by its tile mean kappa.
// Without an odd-odd part K is the diagonal 2 qe; with one (GEN) it couples
// each coefficient with one partner (lane ^ 40, same slot) through the
// symmetric weight qo, and H0 inverts each 2 x 2 block
//   [kappa + alpha qe, alpha qo; alpha qo, kappa + alpha qe'].
template <bool GEN, class TB>
__device__ __forceinline__ void h0_spec(float (&x)[4], const TB &T, float kappa, float alpha) {
  const f4 q = T.qe();
  if constexpr (!GEN) {
#pragma unroll
    for (int r = 0; r < 4; r++) x[r] *= rcpf(fmaf(alpha, q[r], kappa));
  } else {
    const f4 qo = T.qo();
    float y[4];
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const float xp = __shfl_xor(x[r], 40), qp = __shfl_xor(q[r], 40);
      const float a = fmaf(alpha, q[r], kappa), d = fmaf(alpha, qp, kappa), b = alpha * qo[r];
      y[r] = fmaf(d, x[r], -b * xp) * rcpf(fmaf(a, d, -b * b));
    }
#pragma unroll
    for (int r = 0; r < 4; r++) x[r] = y[r];
  }
}

// Pixel curvatures above (1 + kH0Clamp) kappa count as that much in the
// curvature-corrected H0.
constexpr float kH0Clamp = 0.5f;

// The curvature correction applies while the NC curvature, at most
// alpha |K|, is within kH0CorrRatio of the mean likelihood curvature kappa.
// Beyond that the likelihood alone holds the low frequencies, where P is
// 1 / kappa and the curvature varies by more than kappa over the tile: the
// first Neumann term is then a poor step there, and the gradient rule,
// scaled by that small curvature, does not see the error.  Measured in
// FP64: photon tiles at alpha 0.1 have ratios 5-11 and save one iteration in
// three; at alpha 2 (ratios 110-215) the plain start took more iterations
// with it, and data scaled by 1e4 (ratios near 1e5) stopped 3e-3 above the
// minimiser's cost.  The spectral H0 is used beyond the ratio, decided per
// tile from kappa at its start.
constexpr float kH0CorrRatio = 30.0f;

// The linear start's slopes in ln(alpha) are extrapolated at most a factor 8
// from the alpha the coefficients were fitted at: ln 8.
constexpr float kLinSlopeSpan = 2.0794415f;

// x <- H0 x with the likelihood curvature's variation over the tile taken
// to first order (precond 2).  The Hessian is P^-1 + D, P = (kappa I +
// alpha K)^-1 (spectral, h0_spec) and D = diag(c_i - kappa) (pixels, c_i =
// c / max(u + gamma, FITMIN)^2); H0 = P - P D P is the first Neumann term of
// (P^-1 + D)^-1 = P (I + D P)^-1, applied with one basis change each way.
// D is clamped at kH0Clamp kappa from above, which keeps H0 positive
// definite: K is positive semidefinite, so P <= I / kappa and
// P^1/2 D P^1/2 <= kH0Clamp I.  Only the preconditioner depends on this
// approximation, not the minimiser.  iv is 1 / max(u + gamma, FITMIN) when
// CACHED, else recomputed from u; s scales the transform to pixels like the
// search direction's.
template <bool GEN, bool CACHED, class TT, class GB>
__device__ __forceinline__ void h0_corr(float (&x)[4], const TT &T, const GB &G,
                                        const float (&c)[4], const float (&iv)[4],
                                        const float (&u)[4], const float (&gm)[4], float kappa,
                                        float alpha, float s, bool exact, float4 *stash) {
  h0_spec<GEN>(x, T, kappa, alpha);
  float xp[4], w[4];
  // P x waits in LDS across the two basis changes (registers are the
  // scarcer resource there); the clobbers keep the compiler from
  // forwarding the stored value instead of reloading it
  *stash = make_float4(x[0], x[1], x[2], x[3]);
  asm volatile("" ::: "memory");
  to_pix16(x, T, G, s, xp, exact);
  const float hi = (1.0f + kH0Clamp) * kappa;
#pragma unroll
  for (int r = 0; r < 4; r++) {
    const float v = CACHED ? iv[r] : rcpf(fmaxf(u[r] + gm[r], kFitMin));
    xp[r] *= fminf(c[r] * v * v, hi) - kappa;
  }
  to_spec16(xp, T, G, 1.0f, w, exact);
  h0_spec<GEN>(w, T, kappa, alpha);
  asm volatile("" ::: "memory");
  const float4 z = *stash;
  x[0] = z.x - w[0];
  x[1] = z.y - w[1];
  x[2] = z.z - w[2];
  x[3] = z.w - w[3];
}

// ---------------------------------------------------------------------------
// More-Thuente line search on uniform scalars (More & Thuente 1994; the
// liblbfgs default, lbfgs.h:704-719).  phi(t) = f(u + t d) - f(u).
// ---------------------------------------------------------------------------

enum LsResult { LS_CONTINUE = 0, LS_ACCEPT = 1, LS_FAIL_PRECISION = -1,
                LS_FAIL_MINSTEP = -2, LS_FAIL_INCREASE = -3 };

struct MtState {
  float stx, fx, dx, sty, fy, dy;
  float stmin, stmax, width, prev_width;
  float dginit, dgtest;
  int count;
  bool brackt, stage1, uinfo;
};

__device__ __forceinline__ float fmax3(float a, float b, float c) {
  return fmaxf(fmaxf(a, b), c);
}

// Safeguarded step (section 4 of the paper); false on inconsistent input.
__device__ bool mt_cstep(float &stx, float &fx, float &dx, float &sty, float &fy,
                         float &dy, float &stp, float fp, float dp, bool &brackt,
                         float stmin, float stmax) {
  const float t = stp;
  if (brackt) {
    float lo = fminf(stx, sty), hi = fmaxf(stx, sty);
    if (t <= lo || hi <= t) return false;
    if (0.0f <= dx * (t - stx)) return false;
    if (stmax < stmin) return false;
  }
  const float sgnd = dp * (dx / fabsf(dx));
  float stpf, stpc, stpq, theta, s, gam, p, q;
  bool bound;
  if (fp > fx) {
    bound = true;
    theta = 3.0f * (fx - fp) / (t - stx) + dx + dp;
    s = fmax3(fabsf(theta), fabsf(dx), fabsf(dp));
    gam = s * sqrtf(fmaxf(0.0f, (theta / s) * (theta / s) - (dx / s) * (dp / s)));
    if (t < stx) gam = -gam;
    p = (gam - dx) + theta;
    q = ((gam - dx) + gam) + dp;
    stpc = stx + (p / q) * (t - stx);
    stpq = stx + ((dx / ((fx - fp) / (t - stx) + dx)) / 2.0f) * (t - stx);
    stpf = (fabsf(stpc - stx) < fabsf(stpq - stx)) ? stpc : stpc + (stpq - stpc) / 2.0f;
    brackt = true;
  } else if (sgnd < 0.0f) {
    bound = false;
    theta = 3.0f * (fx - fp) / (t - stx) + dx + dp;
    s = fmax3(fabsf(theta), fabsf(dx), fabsf(dp));
    gam = s * sqrtf(fmaxf(0.0f, (theta / s) * (theta / s) - (dx / s) * (dp / s)));
    if (t > stx) gam = -gam;
    p = (gam - dp) + theta;
    q = ((gam - dp) + gam) + dx;
    stpc = t + (p / q) * (stx - t);
    stpq = t + (dp / (dp - dx)) * (stx - t);
    stpf = (fabsf(stpc - t) > fabsf(stpq - t)) ? stpc : stpq;
    brackt = true;
  } else if (fabsf(dp) < fabsf(dx)) {
    bound = true;
    theta = 3.0f * (fx - fp) / (t - stx) + dx + dp;
    s = fmax3(fabsf(theta), fabsf(dx), fabsf(dp));
    gam = s * sqrtf(fmaxf(0.0f, (theta / s) * (theta / s) - (dx / s) * (dp / s)));
    if (t > stx) gam = -gam;
    p = (gam - dp) + theta;
    q = (gam + (dx - dp)) + gam;
    float rr = p / q;
    if (rr < 0.0f && gam != 0.0f)
      stpc = t + rr * (stx - t);
    else
      stpc = (t > stx) ? stmax : stmin;
    stpq = t + (dp / (dp - dx)) * (stx - t);
    if (brackt)
      stpf = (fabsf(t - stpc) < fabsf(t - stpq)) ? stpc : stpq;
    else
      stpf = (fabsf(t - stpc) > fabsf(t - stpq)) ? stpc : stpq;
  } else {
    bound = false;
    if (brackt) {
      theta = 3.0f * (fp - fy) / (sty - t) + dy + dp;
      s = fmax3(fabsf(theta), fabsf(dy), fabsf(dp));
      gam = s * sqrtf(fmaxf(0.0f, (theta / s) * (theta / s) - (dy / s) * (dp / s)));
      if (t > sty) gam = -gam;
      p = (gam - dp) + theta;
      q = ((gam - dp) + gam) + dy;
      stpc = t + (p / q) * (sty - t);
      stpf = stpc;
    } else {
      stpf = (t > stx) ? stmax : stmin;
    }
  }
  if (fp > fx) {
    sty = t; fy = fp; dy = dp;
  } else {
    if (sgnd < 0.0f) { sty = stx; fy = fx; dy = dx; }
    stx = t; fx = fp; dx = dp;
  }
  stpf = fminf(stpf, stmax);
  stpf = fmaxf(stpf, stmin);
  if (brackt && bound) {
    float mq = stx + 0.66f * (sty - stx);
    stpf = (stx < sty) ? fminf(mq, stpf) : fmaxf(mq, stpf);
  }
  stp = stpf;
  return true;
}

__device__ __forceinline__ void mt_init(MtState &S, float dginit, const SolveArgs &A) {
  S.stx = S.sty = 0.0f;
  S.fx = S.fy = 0.0f;
  S.dx = S.dy = dginit;
  S.dginit = dginit;
  S.dgtest = A.ftol * dginit;
  S.width = A.max_step - A.min_step;
  S.prev_width = 2.0f * S.width;
  S.count = 0;
  S.brackt = false;
  S.stage1 = true;
  S.uinfo = false;
}

// Clamp / fall back before a trial; returns the step to evaluate.
__device__ __forceinline__ float mt_pre(MtState &S, float stp, const SolveArgs &A) {
  if (S.brackt) {
    S.stmin = fminf(S.stx, S.sty);
    S.stmax = fmaxf(S.stx, S.sty);
  } else {
    S.stmin = S.stx;
    S.stmax = stp + 4.0f * (stp - S.stx);
  }
  stp = fmaxf(stp, A.min_step);
  stp = fminf(stp, A.max_step);
  if ((S.brackt && ((stp <= S.stmin || S.stmax <= stp) ||
                    A.max_ls <= S.count + 1 || S.uinfo)) ||
      (S.brackt && (S.stmax - S.stmin <= A.xtol * S.stmax)))
    stp = S.stx;
  return stp;
}

// Tests after a trial (f = phi(stp), dg = phi'(stp)); on LS_CONTINUE the
// next trial step is left in stp.
__device__ __forceinline__ int mt_post(MtState &S, float &stp, float f, float dg,
                                       const SolveArgs &A) {
  const float ftest1 = stp * S.dgtest;
  S.count++;
  if (S.brackt && ((stp <= S.stmin || S.stmax <= stp) || S.uinfo))
    return LS_FAIL_PRECISION;
  if (stp == A.max_step && f <= ftest1 && dg <= S.dgtest) return LS_FAIL_PRECISION;
  if (stp == A.min_step && (ftest1 < f || S.dgtest <= dg)) return LS_FAIL_MINSTEP;
  if (S.brackt && (S.stmax - S.stmin) <= A.xtol * S.stmax) return LS_FAIL_PRECISION;
  if (A.max_ls <= S.count) return LS_FAIL_PRECISION;
  if (f <= ftest1 && fabsf(dg) <= A.gtol * (-S.dginit)) return LS_ACCEPT;
  if (S.stage1 && f <= ftest1 && fminf(A.ftol, A.gtol) * S.dginit <= dg) S.stage1 = false;
  bool ok;
  if (S.stage1 && ftest1 < f && f <= S.fx) {
    float fm = f - stp * S.dgtest, fxm = S.fx - S.stx * S.dgtest,
          fym = S.fy - S.sty * S.dgtest;
    float dgm = dg - S.dgtest, dgxm = S.dx - S.dgtest, dgym = S.dy - S.dgtest;
    ok = mt_cstep(S.stx, fxm, dgxm, S.sty, fym, dgym, stp, fm, dgm, S.brackt,
                  S.stmin, S.stmax);
    S.fx = fxm + S.stx * S.dgtest;
    S.fy = fym + S.sty * S.dgtest;
    S.dx = dgxm + S.dgtest;
    S.dy = dgym + S.dgtest;
  } else {
    ok = mt_cstep(S.stx, S.fx, S.dx, S.sty, S.fy, S.dy, stp, f, dg, S.brackt,
                  S.stmin, S.stmax);
  }
  S.uinfo = !ok;
  if (S.brackt) {
    if (0.66f * S.prev_width <= fabsf(S.sty - S.stx)) stp = S.stx + 0.5f * (S.sty - S.stx);
    S.prev_width = S.width;
    S.width = fabsf(S.sty - S.stx);
  }
  // non-finite step (NaN data): give up rather than loop
  if (!(fabsf(stp) <= 3.0e38f)) return LS_FAIL_PRECISION;
  return LS_CONTINUE;
}

// ---------------------------------------------------------------------------
// The persistent solver kernel
// ---------------------------------------------------------------------------

// The frame table seen through the constant address space (scalar loads).
typedef __attribute__((address_space(4))) const FrameDev FrameConst;

struct TileGeom {
  int bx, by;           // window origin in the frame (may be -1)
  int r_lo, r_hi;       // owned local rows [r_lo, r_hi)
  int c_lo, c_hi;       // owned local cols
};

__device__ __forceinline__ TileGeom tile_geom(const FrameConst &F, long long local) {
  TileGeom G;
  // tile indices fit 32 bits (the host caps a batch below 2^32 tiles)
  const unsigned int l32 = (unsigned int)local, ty = (unsigned int)F.ty;
  const int a = (int)(l32 / ty), b = (int)(l32 - (unsigned int)a * ty);
  if (F.geom == NCS_GEOM_REFERENCE) {
    // windows of side r at -1, -1 + (r - 2), ...; the last one clamped to
    // end one pixel past the frame (ncs.c:71-85); each owns its interior up
    // to the next window's interior
    const int r = F.win, s = r - 2;
    const int lastx = F.H + 1 - r, lasty = F.W + 1 - r;
    G.bx = min(-1 + s * a, lastx);
    G.by = min(-1 + s * b, lasty);
    G.r_lo = 1;
    G.r_hi = (G.bx == lastx) ? r - 1 : min(r - 1, lastx + 1 - G.bx);
    G.c_lo = 1;
    G.c_hi = (G.by == lasty) ? r - 1 : min(r - 1, lasty + 1 - G.by);
  } else {
    // non-overlapping blocks of the window side r (16, or the side of an
    // r x r mask) in the top-left of the tile, zero-padded past the frame
    const int r = F.win;
    G.bx = r * a;
    G.by = r * b;
    G.r_lo = 0;
    G.r_hi = min(r, F.H - G.bx);
    G.c_lo = 0;
    G.c_hi = min(r, F.W - G.by);
  }
  return G;
}

// Per-lane likelihood value at u: log2 on the transcendental unit times
// ln 2 (a few ulp per term; the cost is a sum of terms of one sign-mixed
// magnitude, so its relative error stays near 1e-7).
__device__ __forceinline__ float ll_value_lane(const float (&u)[4], const float (&gm)[4],
                                               const float (&c)[4], const float (&lin)[4]) {
  float s = 0.0f;
#pragma unroll
  for (int r = 0; r < 4; r++) {
    const float v = fmaxf(u[r] + gm[r], kFitMin);
    s += lin[r] * u[r] - c[r] * ln_hw(v);
  }
  return s;
}

// Per-lane likelihood gradient lin - c / max(u + gamma, FITMIN)
// (ncs.c:158-175); iv receives 1 / max(u + gamma, FITMIN), which the line
// search and the curvature reuse until u moves.
__device__ __forceinline__ void ll_grad_lane(const float (&u)[4], const float (&gm)[4],
                                             const float (&c)[4], const float (&lin)[4],
                                             float (&gll)[4], float (&iv)[4]) {
#pragma unroll
  for (int r = 0; r < 4; r++) {
    iv[r] = rcpf(fmaxf(u[r] + gm[r], kFitMin));
    gll[r] = lin[r] - c[r] * iv[r];
  }
}

__device__ __forceinline__ float ll_lane(const float (&u)[4], const float (&gm)[4],
                                         const float (&c)[4], const float (&lin)[4],
                                         float (&gll)[4], float (&iv)[4]) {
  ll_grad_lane(u, gm, c, lin, gll, iv);
  return ll_value_lane(u, gm, c, lin);
}

// One trial of the line search at step t: the likelihood part of
// phi(t) = f(u + t d) - f(u) and of phi'(t), and the likelihood gradient at
// u + t d.  The NC part is the quadratic t a1 + t^2 a2 / 2 (added by the
// caller).
// Per-lane sum of the likelihood curvature c / max(u + gamma, FITMIN)^2,
// from iv = 1 / max(u + gamma, FITMIN).
__device__ __forceinline__ float curv_lane(const float (&c)[4], const float (&iv)[4]) {
  float s = 0.0f;
#pragma unroll
  for (int r = 0; r < 4; r++) s = fmaf(c[r] * iv[r], iv[r], s);
  return s;
}

// iv = 1 / max(u + gamma, FITMIN) at u (recomputed here unless CACHED:
// keeping it across iterations costs four VGPRs, which only short histories
// can spare); ivt receives it at u + t d.
template <bool PAD, bool CACHED>
__device__ __forceinline__ void ll_trial(float t, const float (&u)[4], const float (&dv)[4],
                                         const float (&gm)[4], const float (&c)[4],
                                         const float (&lin)[4], const float (&iv)[4],
                                         float (&gt)[4], float (&ivt)[4], float (&red)[2]) {
  red[0] = 0.0f;
  red[1] = 0.0f;
#pragma unroll
  for (int r = 0; r < 4; r++) {
    const float td = t * dv[r];
    const float w0 = u[r] + gm[r];
    const float w = w0 + td;
    const float vt = fmaxf(w, kFitMin);
    // relative change of max(u + gamma, FITMIN); exact step t d when no guard
    const float delta = (w0 >= kFitMin && w >= kFitMin) ? td : vt - fmaxf(w0, kFitMin);
    const float lg = log1p_acc(delta * (CACHED ? iv[r] : rcpf(fmaxf(w0, kFitMin))));
    const float l = PAD ? lin[r] : 1.0f;
    red[0] += l * td - c[r] * lg;
    ivt[r] = rcpf(vt);
    gt[r] = l - c[r] * ivt[r];
    red[1] += gt[r] * dv[r];
  }
  wave_sum<2>(red);
}

// One to kMaxWavesPerGroup waves per workgroup, each solving its own tiles;
// a wave's L-BFGS history lives in its slice of LDS as two rings of M vectors,
// followed by one vector of scratch for the curvature-corrected H0 (each
// lane keeps its own 4 values at float4 index slot * 64 + lane, so no lane
// reads another lane's LDS and no barrier is needed).  PAD enables
// zero-padded pixels (block geometry).
//
// Occupancy target: 5 waves per SIMD (at most 96 VGPRs) for histories up to
// 4, 4 waves (128 VGPRs) above; the latency of the reduction chains and of
// the MFMA stages is hidden by other waves, so the extra wave pays
// (measured: -10 % at history 1).
//
// F16 selects the basis changes: FP16 hi/lo split on the FP16 matrix cores
// (FP32 near the stop rule and as fallback), or FP32 MFMA throughout.
template <bool F16> struct Basis32 { typedef GlobalBasis type; };
template <> struct Basis32<false> { typedef LdsTables<false> type; };

template <bool F16>
__device__ __forceinline__ typename Basis32<F16>::type basis32(const LdsTables<F16> &T,
                                                               const float *tables) {
  if constexpr (F16) return GlobalBasis{tables};
  else return T;
}

template <int M, bool GEN, bool PAD, bool F16>
__global__ __launch_bounds__(64 * kMaxWavesPerGroup)
__attribute__((amdgpu_waves_per_eu(M <= 4 ? 5 : 4, 8)))
void solve_kernel(SolveArgs A) {
  extern __shared__ float4 lds_hist[];
  __shared__ MtState mt_state[kMaxWavesPerGroup];  // line-search state per wave
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // uniform
  constexpr bool kCacheIv = M <= 3;  // keep 1/(u+gamma) across iterations
  // The curvature-corrected H0 is built for histories up to 3; above that its
  // two basis changes per direction cost more than the iterations it saves
  // (measured at history 5), and precond 2 runs as precond 1.
  constexpr bool kCorr = M <= 3;
  MtState &mt_lds = mt_state[wave];
  float4 *const Sb = lds_hist + wave * ((2 * M + 1) * 64);  // s vectors, ring of M
  float4 *const Yb = Sb + M * 64;                           // y vectors, ring of M
  float4 *const Hz = Yb + M * 64 + (threadIdx.x & 63);      // this lane's H0 stash
  const int lane = threadIdx.x & 63;
  const int col = lane & 15, grp = lane >> 4;

  // OTF / DFT tables: each wave writes its lanes' entries (the same values
  // for every wave of the group) and each lane reads back only its own, so
  // no barrier is needed.
  typedef LdsTables<F16> TT;
  __shared__ f4 tab_lds[kWave * TT::kCount];
  {
    const f4 *tb = reinterpret_cast<const f4 *>(A.tables + lane * 4 * kTableF4) + TT::kFirst;
#pragma unroll
    for (int i = 0; i < TT::kCount; i++) tab_lds[lane * TT::kCount + i] = tb[i];
  }
  const TT T{tab_lds + lane * TT::kCount};
  const auto GB = basis32<F16>(T, A.tables);

  // Tile schedule.  Device-scope atomics are slow on this part (they are
  // resolved beyond the XCD's L2), so half the tiles are assigned statically:
  // wave w owns the contiguous block [w S, w S + S), S = floor(total / 2 /
  // waves).  The rest are split into kQueues ranges, one per XCD, that waves
  // take kChunk at a time with one atomic per chunk, home range first, then
  // the others.  Tiles differ in iterations, so the dynamic half is what
  // balances the waves: a static 7/8 (or 3/4) left the launch 7% (4%) longer,
  // and a larger dynamic share gained nothing more.  next_tile() gives the wave's next
  // tile (uniform) or -1 when every range is exhausted.
  // Tile indices fit 32 bits (the host caps a batch below 2^32 tiles).  The
  // static block is the wave's first chunk [next, end).
  const unsigned int total = (unsigned int)A.total_tiles;
  unsigned int next, end;
  {
    const unsigned int n_waves = gridDim.x * (blockDim.x >> 6);
    const unsigned int per_wave = (total - total / 2) / n_waves;
    next = (blockIdx.x * (blockDim.x >> 6) + wave) * per_wave;
    end = next + per_wave;
  }
  int qv = 0;  // dynamic ranges found exhausted
  auto next_tile = [&]() -> long long {
    if (next >= end) {
      const unsigned int n_waves = gridDim.x * (blockDim.x >> 6);
      const unsigned int dyn0 = (total - total / 2) / n_waves * n_waves;
      const unsigned int nd = total - dyn0;
      while (qv < kQueues) {
        const int q = (blockIdx.x + qv) % kQueues;  // home range first
        const unsigned int lo = dyn0 + (unsigned int)((unsigned long long)nd * q / kQueues);
        const unsigned int hi = dyn0 + (unsigned int)((unsigned long long)nd * (q + 1) / kQueues);
        unsigned int got = 0;
        if (lane == 0) got = atomicAdd(A.queue + q * kQueueStride, (unsigned int)kChunk);
        got = (unsigned int)__builtin_amdgcn_readfirstlane((int)got);
        if (got < hi - lo) {
          next = lo + got;
          end = min(hi, next + (unsigned int)kChunk);
          break;
        }
        qv++;
      }
      if (next >= end) return -1;
    }
    return (long long)(next++);
  };

  // The frame table is read through the constant address space (scalar
  // loads into SGPRs; the kernel never writes it).  Tiles of a chunk are
  // consecutive, so the previous frame is tried before a binary search.
  const FrameConst *frames = (const FrameConst *)(A.frames);
  int fi = 0;
  auto frame_of = [&](long long t) -> int {
    if (!(frames[fi].first <= t && (fi + 1 >= A.n_frames || frames[fi + 1].first > t))) {
      int lo = 0, hi = A.n_frames - 1;
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (frames[mid].first <= t) lo = mid; else hi = mid - 1;
      }
      fi = lo;
    }
    return fi;
  };

  // Raw window of a tile (edge clamp for the reference geometry, zero pad
  // for blocks) and, for warm starts, its start tile.
  auto load_tile = [&](const FrameConst &F, const TileGeom &G, long long t, float (&dd)[4],
                       float (&gg)[4], float (&w0)[4]) {
    // lane-derived offsets are recomputed per tile: hoisted out of the tile
    // loop they would be spilled, and reloading them costs more
    int ln = lane;
    asm volatile("" : "+v"(ln));
    const int col = ln & 15, grp = ln >> 4;
    // 32-bit element offsets from uniform base pointers (the host keeps
    // every frame below 2^31 elements)
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int row = 4 * grp + r;
      if (!PAD) {
        const unsigned int gr = (unsigned int)min(max(G.bx + row, 0), F.H - 1);
        const unsigned int gc = (unsigned int)min(max(G.by + col, 0), F.W - 1);
        dd[r] = F.data[gr * (unsigned int)F.ds + gc];
        gg[r] = F.gamma[gr * (unsigned int)F.gs + gc];
      } else if (F.geom == NCS_GEOM_REFERENCE) {
        // an r x r window in the top-left of the tile, zeros around it
        const bool in = (row < F.win) && (col < F.win);
        const unsigned int gr = (unsigned int)min(max(G.bx + row, 0), F.H - 1);
        const unsigned int gc = (unsigned int)min(max(G.by + col, 0), F.W - 1);
        dd[r] = in ? F.data[gr * (unsigned int)F.ds + gc] : 0.0f;
        gg[r] = in ? F.gamma[gr * (unsigned int)F.gs + gc] : 0.0f;
      } else {
        const bool ok = (row < G.r_hi) && (col < G.c_hi);
        const unsigned int gr = ok ? (unsigned int)(G.bx + row) : 0u;
        const unsigned int gc = ok ? (unsigned int)(G.by + col) : 0u;
        dd[r] = ok ? F.data[gr * (unsigned int)F.ds + gc] : 0.0f;
        gg[r] = ok ? F.gamma[gr * (unsigned int)F.gs + gc] : 0.0f;
      }
    }
    if (A.u0 != nullptr) {
      const float *w = A.u0 + (F.slot + (t - F.first)) * kTilePix;  // uniform
#pragma unroll
      for (int r = 0; r < 4; r++) w0[r] = w[(unsigned int)((4 * grp + r) * kTile + col)];
    }
  };

  long long t = next_tile();
  if (t < 0) return;
  int fc = frame_of(t);  // frame of the current tile
  TileGeom G = tile_geom(frames[fc], t - frames[fc].first);
  float raw_d[4], gm[4], w0[4] = {0.f, 0.f, 0.f, 0.f};
  load_tile(frames[fc], G, t, raw_d, gm, w0);

  for (;;) {
    const FrameConst &F = frames[fc];
    const float alpha = F.alpha;
    float c[4], lin[4], u[4];
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int row = 4 * grp + r;
      lin[r] = (!PAD || (F.geom == NCS_GEOM_REFERENCE ? (row < F.win && col < F.win)
                                                      : (row < G.r_hi && col < G.c_hi)))
                   ? 1.0f : 0.0f;
      c[r] = lin[r] * (raw_d[r] + gm[r]);
      u[r] = raw_d[r];
    }
    const float st = tile_scale(raw_d);

    // Start.  The L-BFGS state (gradient, NC gradient, direction, history)
    // lives in the orthonormal real-DFT basis, where K and the spectral H0
    // are diagonal: each iteration needs one transform to pixels (the
    // direction, for the likelihood trial) and one back (the likelihood
    // gradient).  Dot products are basis-independent.
    float gs[4], gncs[4], zs[4], iv[4];  // spectral g, K u; pixel 1/(u+gamma)
    float f = 0.0f;  // f(u) - f(u_start): only differences drive the search
    // Basis changes run on the FP16 matrix cores (22-bit operands) until the
    // gradient is within kF16Floor of the stop rule's scale, and in FP32
    // after that: a default-tolerance solve (eps 1e-5) stops before the
    // switch, a tighter one finishes with FP32 accuracy.
    bool exact = !F16;
    {
      float gll[4];
      to_spec16<false>(u, T, GB, st, zs, exact);
      nc_spec<GEN>(zs, T, gncs);
      const bool warm = A.u0 != nullptr || A.lin_start;
      if (!warm) {
        ll_grad_lane(u, gm, c, lin, gll, iv);
      } else {
        // warm start: keep it unless the plain start scores a lower cost
        float both[4];
        both[0] = ll_lane(u, gm, c, lin, gll, iv);
        both[1] = zs[0] * gncs[0] + zs[1] * gncs[1] + (zs[2] * gncs[2] + zs[3] * gncs[3]);
        float w[4], wz[4], wgs[4], wll[4], wiv[4];
        if (A.u0 != nullptr) {
#pragma unroll
          for (int r = 0; r < 4; r++) w[r] = w0[r];
          to_spec16(w, T, GB, st, wz, exact);
        } else {
          // Learned linear start in the real-DFT basis:
          //   Z_w = Z_d + sigma (la * Z_x0 + lb * Z_x1 + c e_DC),
          //   x0 = (d - mu) / sigma, x1 = gamma / sigma^2,
          //   mu = mean d, sigma = sqrt(max(mu, 1) + mean gamma),
          // with the tile means read off the DC coefficients (sum / 16).
          float zg[4], pz[4], wp[4];
          to_spec16(gm, T, GB, st, zg, exact);
          const float mu = lane0(zs[0]) * (1.0f / 16.0f);
          const float mgam = lane0(zg[0]) * (1.0f / 16.0f);
          const float sig = sqrtf(fmaxf(mu, 1.0f) + mgam);
          const float isig = rcpf(sig);
          f4 la = T.la(), lb = T.lb();
          float lc = A.lin_c;
          if (A.lin_slope && alpha != A.lin_aref && alpha > 0.0f) {
            // the coefficients at this frame's alpha: plus ln(alpha /
            // alpha_ref), clamped to +-kLinSlopeSpan, times their slopes
            // (global table, read once per tile); frames at alpha_ref keep
            // the coefficients exactly
            const float ell0 =
                (__builtin_amdgcn_logf(alpha) - __builtin_amdgcn_logf(A.lin_aref)) * kLn2;
            const float ell = __builtin_bit_cast(
                float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(
                           int, fminf(fmaxf(ell0, -kLinSlopeSpan), kLinSlopeSpan))));
            const f4 *sl = reinterpret_cast<const f4 *>(A.tables + lane * kTableLaneFloats) +
                           kTableSlope;
            const f4 la1 = sl[0], lb1 = sl[1];
#pragma unroll
            for (int r = 0; r < 4; r++) {
              la[r] = fmaf(ell, la1[r], la[r]);
              lb[r] = fmaf(ell, lb1[r], lb[r]);
            }
            lc = fmaf(ell, A.lin_c1, lc);
          }
#pragma unroll
          for (int r = 0; r < 4; r++)
            pz[r] = sig * fmaf(la[r], zs[r] * isig, lb[r] * zg[r] * (isig * isig));
          if (lane == 0)  // DC: x0 has none; the bias term lives here
            pz[0] = sig * fmaf(lb[0], zg[0] * (isig * isig), 16.0f * lc);
          to_pix16<false>(pz, T, GB, st, wp, exact);
#pragma unroll
          for (int r = 0; r < 4; r++) {
            w[r] = u[r] + wp[r];
            wz[r] = zs[r] + pz[r];
          }
        }
        nc_spec<GEN>(wz, T, wgs);
        both[2] = ll_lane(w, gm, c, lin, wll, wiv);
        both[3] = wz[0] * wgs[0] + wz[1] * wgs[1] + (wz[2] * wgs[2] + wz[3] * wgs[3]);
        wave_sum<4>(both);
        const float f_plain = both[0] + 0.5f * alpha * both[1];
        const float f_warm = both[2] + 0.5f * alpha * both[3];
        if (!A.fallback || f_warm < f_plain || !(f_plain == f_plain)) {
#pragma unroll
          for (int r = 0; r < 4; r++) {
            u[r] = w[r];
            gncs[r] = wgs[r];
            gll[r] = wll[r];
            iv[r] = wiv[r];
          }
        }
      }
      float gls[4];
      to_spec16(gll, T, GB, 1.0f, gls, exact);
#pragma unroll
      for (int r = 0; r < 4; r++) gs[r] = fmaf(alpha, gncs[r], gls[r]);
    }

    int status = NCS_REACHED_MAXITERS, iters = 0, evals = 1;
    bool gnc_exact = true;
    int n_inc = 0;  // incremental NC-gradient updates since the last exact one
    float gg, uu, kappa;
    {
      float red[3] = {gs[0] * gs[0] + gs[1] * gs[1] + (gs[2] * gs[2] + gs[3] * gs[3]),
                      u[0] * u[0] + u[1] * u[1] + (u[2] * u[2] + u[3] * u[3]),
                      curv_lane(c, iv)};
      wave_sum<3>(red);
      gg = red[0];
      uu = red[1];
      kappa = red[2] * (1.0f / 256.0f);
    }
    bool done = gg <= A.eps * A.eps * fmaxf(1.0f, uu);
    float uu_max = uu;  // largest |u|^2 of the iterates
    if (done) status = NCS_SUCCESS;
    else if (gg <= kF16Floor * kF16Floor * fmaxf(1.0f, uu)) exact = true;

    float rho[M], al[M];  // rho[j], al[j]: j-th newest pair
    int nh = 0, head = M - 1;
    float ds[4];  // search direction, spectral
#pragma unroll
    for (int r = 0; r < 4; r++) ds[r] = -gs[r];
    float step = rsqrtf(gg);
    // The spectral H0 (precond) models the likelihood curvature by its tile
    // mean.  From a start far from the solution that mean can change by
    // orders of magnitude along the way and the line search then fail; a
    // tile whose search fails under it drops its history and continues with
    // liblbfgs's scalar H0 from the same point.  A search that fails under
    // the curvature-corrected H0 restarts the tile with the spectral one.
    bool spec_h0 = A.precond != 0;
    // curvature-corrected H0 (uniform), decided per tile at its start
    bool corr_h0 = kCorr && A.precond == 2 && alpha * A.k_norm <= kH0CorrRatio * kappa;
    if (spec_h0 && !done) {
      if (kCorr && corr_h0)
        h0_corr<GEN, true>(ds, T, GB, c, iv, u, gm, kappa, alpha, st, exact, Hz);
      else
        h0_spec<GEN>(ds, T, kappa, alpha);
      step = 1.0f;
    }
    float ys_last = 1.0f, yy_last = 1.0f;

    for (int k = 1; !done && k <= A.max_it; k++) {
      // K d and the quadratic coefficients of NC along d.  A direction that
      // is not a descent direction (rounding at the FP32 noise floor can
      // spoil the history) is replaced once by the history-free one, -H0 g.
      float dp[4];  // d in pixels
      float dg0, a1, a2;
      for (int attempt = 0;; attempt++) {
        float kds[4];  // K d (spectral), recomputed at the accepted step
        nc_spec<GEN>(ds, T, kds);
        float red[3] = {
            gs[0] * ds[0] + gs[1] * ds[1] + (gs[2] * ds[2] + gs[3] * ds[3]),
            gncs[0] * ds[0] + gncs[1] * ds[1] + (gncs[2] * ds[2] + gncs[3] * ds[3]),
            kds[0] * ds[0] + kds[1] * ds[1] + (kds[2] * ds[2] + kds[3] * ds[3])};
        wave_sum<3>(red);
        dg0 = red[0];
        a1 = red[1];
        a2 = red[2];
        if (dg0 <= 0.0f || attempt == 1 || nh == 0) break;
        nh = 0;
#pragma unroll
        for (int r = 0; r < 4; r++) ds[r] = -gs[r];
        if (spec_h0) {
          h0_spec<GEN>(ds, T, kappa, alpha);
        } else {
          const float sc = rsqrtf(gg);
#pragma unroll
          for (int r = 0; r < 4; r++) ds[r] *= sc;
        }
        step = 1.0f;
      }
      iters = k;
      if (!(dg0 <= 0.0f)) {  // not a descent direction (or NaN)
        status = NCS_INCREASING_GRADIENT;
        break;
      }
      to_pix16(ds, T, GB, st, dp, exact);

      // line search; the first trial is taken at the clamped initial step
      float gt[4], ivt[4], phi, dphi_acc = 0.0f;  // dphi_acc: phi'(t) of the accepted trial
      float tstep = fminf(fmaxf(step, A.min_step), A.max_step);
      int ls;
      {
        float red[2];
        ll_trial<PAD, kCacheIv>(tstep, u, dp, gm, c, lin, iv, gt, ivt, red);
        evals++;
        phi = red[0] + alpha * tstep * (a1 + 0.5f * tstep * a2);
        const float dphi = red[1] + alpha * (a1 + tstep * a2);
        if (A.ls == NCS_LS_MORE_THUENTE) {
          if (A.max_ls > 1 && tstep != A.max_step && tstep != A.min_step &&
              phi <= A.ftol * dg0 * tstep && fabsf(dphi) <= A.gtol * (-dg0)) {
            ls = LS_ACCEPT;  // what the full search decides on this trial
            dphi_acc = dphi;
          } else {
            // Full More-Thuente search (rare: on the 10 x 2048^2 benchmark
            // batch every first trial is accepted).  Its state is kept in
            // LDS and the memory clobbers stop the compiler from holding it
            // in VGPRs across the vector trials.
            MtState &S_ = mt_lds;
            mt_init(S_, dg0, A);
            S_.stmin = 0.0f;
            S_.stmax = 5.0f * tstep;
            ls = mt_post(S_, tstep, phi, dphi, A);
            dphi_acc = dphi;
            while (ls == LS_CONTINUE) {
              tstep = mt_pre(S_, tstep, A);
              asm volatile("" ::: "memory");
              ll_trial<PAD, kCacheIv>(tstep, u, dp, gm, c, lin, iv, gt, ivt, red);
              asm volatile("" ::: "memory");
              evals++;
              phi = red[0] + alpha * tstep * (a1 + 0.5f * tstep * a2);
              const float dpt = red[1] + alpha * (a1 + tstep * a2);
              ls = mt_post(S_, tstep, phi, dpt, A);
              dphi_acc = dpt;
            }
          }
        } else {
          // Armijo backtracking (ncs.cl:979-1035)
          int trials = 1;
          for (;;) {
            if (phi <= A.ftol * dg0 * tstep) {
              ls = LS_ACCEPT;
              dphi_acc = red[1] + alpha * (a1 + tstep * a2);
              break;
            }
            tstep *= A.stepm;
            if (!(tstep >= A.min_step)) { ls = LS_FAIL_MINSTEP; break; }
            if (trials >= A.max_ls) { ls = LS_FAIL_PRECISION; break; }
            ll_trial<PAD, kCacheIv>(tstep, u, dp, gm, c, lin, iv, gt, ivt, red);
            evals++;
            trials++;
            phi = red[0] + alpha * tstep * (a1 + 0.5f * tstep * a2);
          }
        }
      }
      // A failed search that the tile can still recover from refreshes its
      // state exactly (FP32 basis changes, history dropped) and searches
      // again:
      //  * under the curvature-corrected H0, from the plain start (the data)
      //    with the spectral H0: the corrected steps may have led where no
      //    H0 recovers (on stiff masks, pixels at the FITMIN guard, where
      //    the gradient is not the cost's derivative); the cost reference
      //    is cleared and the iterations keep counting.  The fixed-iteration
      //    mode (eps 0) ends in failed searches by design and skips this;
      //  * in the FP16 split, from u in FP32: the split's error can stall
      //    the search on an ill-conditioned tile, where the likelihood and NC
      //    parts of phi'(t) nearly cancel.
      const bool restart = kCorr && corr_h0 && A.eps > 0.0f;
      if (ls != LS_ACCEPT && (restart || !exact)) {
        if (restart) {
          float nw[4] = {0.f, 0.f, 0.f, 0.f}, ng[4];
          load_tile(F, G, t, u, ng, nw);
          f = 0.0f;
        }
        corr_h0 = false;
        exact = true;
        nh = 0;
        float gp[4];
        to_spec16<false>(u, T, GB, st, zs, exact);
        nc_spec<GEN>(zs, T, gncs);
        ll_grad_lane(u, gm, c, lin, gp, iv);
        to_spec16(gp, T, GB, 1.0f, gs, exact);
#pragma unroll
        for (int r = 0; r < 4; r++) {
          gs[r] = fmaf(alpha, gncs[r], gs[r]);
          ds[r] = -gs[r];
        }
        float red[3] = {gs[0] * gs[0] + gs[1] * gs[1] + (gs[2] * gs[2] + gs[3] * gs[3]),
                        u[0] * u[0] + u[1] * u[1] + (u[2] * u[2] + u[3] * u[3]),
                        curv_lane(c, iv)};
        wave_sum<3>(red);
        gg = red[0];
        uu = red[1];
        uu_max = fmaxf(uu_max, uu);
        kappa = red[2] * (1.0f / 256.0f);
        gnc_exact = true;
        n_inc = 0;
        if (gg <= A.eps * A.eps * fmaxf(1.0f, uu)) {
          // the exact gradient already meets the stop rule at u (the point
          // the previous iteration accepted, or the plain start)
          iters = k - 1;
          status = NCS_SUCCESS;
          break;
        }
        if (spec_h0) {
          h0_spec<GEN>(ds, T, kappa, alpha);
          step = 1.0f;
        } else {
          step = rsqrtf(gg);
        }
        continue;
      }
      if (ls != LS_ACCEPT && spec_h0) {  // retry with the scalar H0
        spec_h0 = false;
        nh = 0;
#pragma unroll
        for (int r = 0; r < 4; r++) ds[r] = -gs[r];
        step = rsqrtf(gg);
        continue;
      }
      if (ls != LS_ACCEPT) {
        status = (ls == LS_FAIL_MINSTEP) ? NCS_MINIMUM_STEP
                 : (ls == LS_FAIL_INCREASE) ? NCS_INCREASING_GRADIENT
                                            : NCS_REACHED_MAXPRECISION;
        break;
      }

      // accept u + t d
      float ut[4], gnews[4], gncn[4], gts[4];
      nc_spec<GEN>(ds, T, gncn);
#pragma unroll
      for (int r = 0; r < 4; r++) {
        ut[r] = u[r] + tstep * dp[r];
        gncn[r] = fmaf(tstep, gncn[r], gncs[r]);
      }
      const bool moved = __any((ut[0] != u[0]) | (ut[1] != u[1]) | (ut[2] != u[2]) |
                               (ut[3] != u[3]));
      gnc_exact = false;
      n_inc++;
      if ((k & 15) == 0) {  // periodic exact refresh of the NC gradient
        to_spec16<false>(ut, T, GB, st, zs, exact);
        nc_spec<GEN>(zs, T, gncn);
        gnc_exact = true;
        n_inc = 0;
      }
      to_spec16(gt, T, GB, 1.0f, gts, exact);
      f += phi;
      // y.s = t (phi'(t) - phi'(0)) needs no reduction; the third slot is
      // y.y for the scalar initial Hessian, the mean curvature for the
      // spectral one
      const float ys = tstep * (dphi_acc - dg0);
      float yy = 0.0f;
      for (int pass = 0; pass < 2; pass++) {
#pragma unroll
        for (int r = 0; r < 4; r++) gnews[r] = fmaf(alpha, gncn[r], gts[r]);
        float red[3] = {
            gnews[0] * gnews[0] + gnews[1] * gnews[1] + (gnews[2] * gnews[2] + gnews[3] * gnews[3]),
            ut[0] * ut[0] + ut[1] * ut[1] + (ut[2] * ut[2] + ut[3] * ut[3]),
            spec_h0 ? curv_lane(c, ivt)
                      : ((gnews[0] - gs[0]) * (gnews[0] - gs[0]) +
                         (gnews[1] - gs[1]) * (gnews[1] - gs[1]) +
                         ((gnews[2] - gs[2]) * (gnews[2] - gs[2]) +
                          (gnews[3] - gs[3]) * (gnews[3] - gs[3])))};
        wave_sum<3>(red);
        gg = red[0];
        uu = red[1];
        uu_max = fmaxf(uu_max, uu);
        kappa = red[2] * (1.0f / 256.0f);
        yy = red[2];
        done = gg <= A.eps * A.eps * fmaxf(1.0f, uu);
        if (!done && gg <= kF16Floor * kF16Floor * fmaxf(1.0f, uu)) exact = true;
        if (!done || gnc_exact) break;
        // The incremental NC gradient differs from K u by rounding, at most
        // kNcDrift |K| max_j |u_j| per update (|K| <= k_norm).  When that
        // much drift, times alpha, cannot lift |g| past the stop rule's
        // bound, an exact NC gradient would confirm convergence too;
        // otherwise confirm with one.  (The FP16 build always confirms: the
        // bound's registers cost it more spills than the refresh.)
        if (!F16 && sqrtf(gg) + (float)n_inc * kNcDrift * alpha * A.k_norm * sqrtf(uu_max) <=
                        A.eps * fmaxf(1.0f, sqrtf(uu)))
          break;
        to_spec16<false>(ut, T, GB, st, zs, exact);
        nc_spec<GEN>(zs, T, gncn);
        gnc_exact = true;
        n_inc = 0;
      }
      // history pair: s = t d, y = g_new - g (spectral)
      const bool added = ys > 0.0f && !done && moved && k != A.max_it;
      if (added) {
        head = (head + 1 == M) ? 0 : head + 1;
        Sb[head * 64 + lane] = make_float4(tstep * ds[0], tstep * ds[1], tstep * ds[2],
                                           tstep * ds[3]);
        Yb[head * 64 + lane] = make_float4(gnews[0] - gs[0], gnews[1] - gs[1],
                                           gnews[2] - gs[2], gnews[3] - gs[3]);
      }
#pragma unroll
      for (int r = 0; r < 4; r++) {
        u[r] = ut[r];
        gs[r] = gnews[r];
        gncs[r] = gncn[r];
        iv[r] = ivt[r];
      }
      if (done) { status = NCS_SUCCESS; break; }
      if (!moved) { status = NCS_REACHED_MAXPRECISION; break; }
      if (k == A.max_it) { status = NCS_REACHED_MAXITERS; break; }

      if (added) {
#pragma unroll
        for (int j = M - 1; j > 0; j--) rho[j] = rho[j - 1];
        rho[0] = rcpf(ys);
        ys_last = ys;
        yy_last = yy;
        nh = min(nh + 1, M);
      }

      // two-loop recursion in the spectral basis, newest to oldest then back
#pragma unroll
      for (int r = 0; r < 4; r++) ds[r] = -gs[r];
#pragma unroll
      for (int j = 0; j < M; j++) {
        if (j < nh) {
          const int slot = (head - j < 0) ? head - j + M : head - j;
          const float4 sv = Sb[slot * 64 + lane];
          const float4 yv = Yb[slot * 64 + lane];
          const float sj[4] = {sv.x, sv.y, sv.z, sv.w};
          // the pair added in this step: s.(-g) = -t phi'(t), no reduction
          al[j] = rho[j] * ((j == 0 && added) ? -tstep * dphi_acc : wave_dot(sj, ds));
          ds[0] = fmaf(-al[j], yv.x, ds[0]);
          ds[1] = fmaf(-al[j], yv.y, ds[1]);
          ds[2] = fmaf(-al[j], yv.z, ds[2]);
          ds[3] = fmaf(-al[j], yv.w, ds[3]);
        }
      }
      if (kCorr && spec_h0 && corr_h0) {
        h0_corr<GEN, kCacheIv>(ds, T, GB, c, iv, u, gm, kappa, alpha, st, exact, Hz);
      } else if (spec_h0) {
        h0_spec<GEN>(ds, T, kappa, alpha);
      } else {  // liblbfgs: H0 = (s.y / y.y) I
        const float hscale = (nh > 0) ? ys_last * rcpf(yy_last) : rsqrtf(gg);
#pragma unroll
        for (int r = 0; r < 4; r++) ds[r] *= hscale;
      }
#pragma unroll
      for (int j = M - 1; j >= 0; j--) {
        if (j < nh) {
          const int slot = (head - j < 0) ? head - j + M : head - j;
          const float4 sv = Sb[slot * 64 + lane];
          const float4 yv = Yb[slot * 64 + lane];
          const float yj[4] = {yv.x, yv.y, yv.z, yv.w};
          const float b = al[j] - rho[j] * wave_dot(yj, ds);
          ds[0] = fmaf(b, sv.x, ds[0]);
          ds[1] = fmaf(b, sv.y, ds[1]);
          ds[2] = fmaf(b, sv.z, ds[2]);
          ds[3] = fmaf(b, sv.w, ds[3]);
        }
      }
      step = 1.0f;
    }

    // prefetch the next tile: its loads overlap the final evaluation and
    // stitch of this one
    const long long tn = next_tile();
    int fn = fc;
    TileGeom Gn = G;
    float nd[4] = {0.f, 0.f, 0.f, 0.f}, ng[4] = {0.f, 0.f, 0.f, 0.f};
    float nw[4] = {0.f, 0.f, 0.f, 0.f};
    if (tn >= 0) {
      fn = frame_of(tn);
      Gn = tile_geom(frames[fn], tn - frames[fn].first);
      load_tile(frames[fn], Gn, tn, nd, ng, nw);
    }

    // final cost at the returned point (when asked for), NC from an exact
    // transform of u
    if (A.cost != nullptr) {
      to_spec16<false>(u, T, GB, st, zs, exact);
      nc_spec<GEN>(zs, T, gncs);
      float red[2];
      red[0] = ll_value_lane(u, gm, c, lin);
      red[1] = zs[0] * gncs[0] + zs[1] * gncs[1] + (zs[2] * gncs[2] + zs[3] * gncs[3]);
      wave_sum<2>(red);
      f = red[0] + 0.5f * alpha * red[1];
    }

    // stitch the owned interior and per-tile results (per-tile arrays in the
    // caller's frame order)
    int ln = lane;
    asm volatile("" : "+v"(ln));
    const int col = ln & 15, grp = ln >> 4;
    const long long ts = F.slot + (t - F.first);
    float *const ut_base = A.u_tiles != nullptr ? A.u_tiles + ts * kTilePix : nullptr;
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int row = 4 * grp + r;
      if (F.out != nullptr && row >= G.r_lo && row < G.r_hi && col >= G.c_lo &&
          col < G.c_hi)
        F.out[(unsigned int)(G.bx + row) * (unsigned int)F.os + (unsigned int)(G.by + col)] = u[r];
      if (ut_base != nullptr) ut_base[(unsigned int)(row * kTile + col)] = u[r];
    }
    if (lane == 0) {
      if (A.iterations) A.iterations[ts] = iters;
      if (A.status) A.status[ts] = status;
      if (A.cost) A.cost[ts] = f;
      if (A.evaluations) A.evaluations[ts] = evals;
    }
    if (tn < 0) break;
    t = tn;
    fc = fn;
    G = Gn;
#pragma unroll
    for (int r = 0; r < 4; r++) {
      raw_d[r] = nd[r];
      gm[r] = ng[r];
      w0[r] = nw[r];
    }
  }
}

// Cost and gradient of given tiles (one wave per tile), the analogue of
// ncsSREvaluate (clib/pyCNCS/ncs.c:335-364) for parity tests.
template <bool GEN>
__global__ __launch_bounds__(256) void eval_kernel(const float *data, const float *gamma,
                                                   const float *u_in, const float *tables,
                                                   float alpha, int n, float *cost,
                                                   float *grad) {
  const int lane = threadIdx.x & 63;
  const int col = lane & 15, grp = lane >> 4;
  const long long t = (long long)blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
  if (t >= n) return;  // whole waves leave together
  MaskTables T;
  {
    const f4 *tb = reinterpret_cast<const f4 *>(tables + lane * kTableLaneFloats);
    T.ka_ = tb[0];
    T.kb_ = tb[1];
    T.qe_ = tb[2];
    T.qo_ = tb[3];
  }
  float u[4], gm[4], c[4], lin[4], gnc[4], gll[4];
  for (int r = 0; r < 4; r++) {
    const long long i = t * kTilePix + (4 * grp + r) * kTile + col;
    u[r] = u_in[i];
    gm[r] = gamma[i];
    c[r] = data[i] + gm[r];
    lin[r] = 1.0f;
  }
  apply_k<GEN>(u, T, gnc);
  float red[2];
  float iv[4];
  red[0] = ll_lane(u, gm, c, lin, gll, iv);
  red[1] = u[0] * gnc[0] + u[1] * gnc[1] + (u[2] * gnc[2] + u[3] * gnc[3]);
  wave_sum<2>(red);
  if (lane == 0 && cost) cost[t] = red[0] + 0.5f * alpha * red[1];
  if (grad)
    for (int r = 0; r < 4; r++)
      grad[t * kTilePix + (4 * grp + r) * kTile + col] = fmaf(alpha, gnc[r], gll[r]);
}

hipError_t launch_eval(bool general, const float *data, const float *gamma, const float *u,
                       const float *tables, float alpha, int n, float *cost, float *grad,
                       hipStream_t s) {
  if (n <= 0) return hipSuccess;
  unsigned int blocks = (unsigned int)((n + kWavesPerBlock - 1) / kWavesPerBlock);
  if (general)
    hipLaunchKernelGGL(eval_kernel<true>, dim3(blocks), dim3(kWave * kWavesPerBlock), 0, s,
                       data, gamma, u, tables, alpha, n, cost, grad);
  else
    hipLaunchKernelGGL(eval_kernel<false>, dim3(blocks), dim3(kWave * kWavesPerBlock), 0, s,
                       data, gamma, u, tables, alpha, n, cost, grad);
  return hipGetLastError();
}

// Resets the work-queue counters (a kernel, not a memset, so that captured
// graphs replay it correctly).
__global__ void reset_queue_kernel(unsigned int *q) {
  if (threadIdx.x < kQueues) q[threadIdx.x * kQueueStride] = 0u;
}

// Instantiations: history 1..8, with and without the odd-mask coupling and
// zero-padded pixels, FP32 or FP16-split basis changes.
typedef void (*SolveKernelFn)(SolveArgs);

template <int M, bool F16>
static SolveKernelFn pick2(bool general, bool pad) {
  if (general) return pad ? solve_kernel<M, true, true, F16> : solve_kernel<M, true, false, F16>;
  return pad ? solve_kernel<M, false, true, F16> : solve_kernel<M, false, false, F16>;
}

template <int M>
static SolveKernelFn pick(bool general, bool pad, bool f16) {
  return f16 ? pick2<M, true>(general, pad) : pick2<M, false>(general, pad);
}

static SolveKernelFn solve_kernel_for(int m, bool general, bool pad, bool f16) {
  switch (m) {
    case 1: return pick<1>(general, pad, f16);
    case 2: return pick<2>(general, pad, f16);
    case 3: return pick<3>(general, pad, f16);
    case 4: return pick<4>(general, pad, f16);
    case 5: return pick<5>(general, pad, f16);
    case 6: return pick<6>(general, pad, f16);
    case 7: return pick<7>(general, pad, f16);
    case 8: return pick<8>(general, pad, f16);
    default: return nullptr;
  }
}

static size_t solve_lds_bytes(int m, int wpg) {
  return (size_t)wpg * (2 * m + 1) * 64 * sizeof(float4);
}

// Waves of this instantiation one CU holds at once with workgroups of `wpg`
// waves (0 if unsupported).
int solve_waves_per_cu(int m, bool general, bool pad, bool f16, int wpg) {
  SolveKernelFn fn = solve_kernel_for(m, general, pad, f16);
  if (!fn || wpg < 1 || wpg > kMaxWavesPerGroup) return 0;
  int blocks = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
          &blocks, reinterpret_cast<const void *>(fn), kWave * wpg, solve_lds_bytes(m, wpg)) !=
      hipSuccess)
    return 0;
  return blocks * wpg;
}

// Resets the queue and launches `waves` persistent waves in workgroups of
// `wpg` waves (waves is rounded up to a multiple of wpg).
hipError_t launch_solve(int m, bool general, bool pad, bool f16, unsigned int waves, int wpg,
                        hipStream_t s, const SolveArgs &a) {
  SolveKernelFn fn = solve_kernel_for(m, general, pad, f16);
  if (!fn || wpg < 1 || wpg > kMaxWavesPerGroup || waves == 0) return hipErrorInvalidValue;
  const unsigned int groups = (waves + wpg - 1) / wpg;
  hipLaunchKernelGGL(reset_queue_kernel, dim3(1), dim3(64), 0, s, a.queue);
  hipLaunchKernelGGL(fn, dim3(groups), dim3(kWave * wpg), solve_lds_bytes(m, wpg), s, a);
  return hipGetLastError();
}

}  // namespace ncs
