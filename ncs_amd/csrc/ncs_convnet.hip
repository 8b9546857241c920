// ncs_convnet.hip -- the learned warm start: a 4-layer 3x3 conv net
// (2 -> 32 -> 32 -> 32 -> 1, ReLU after layers 1-3, 'same' zero padding
// inside each 16x16 tile) on BF16 MFMA with FP32 accumulation.
//
// Input channels per tile (normalisation shared with ncs_amd/warmstart.py):
//   mu = mean(d), sigma = sqrt(max(mu, 1) + mean(gamma)),
//   x0 = (d - mu) / sigma,  x1 = gamma / sigma^2;
// the net predicts y and the warm start is u0 = d + sigma * y.
//
// Convolution as GEMM on v_mfma_f32_16x16x32_bf16: for one output pixel row
// y and a block of 16 output channels,
//   D[cout][x] += sum_taps  W_tap[cout][cin] * In[cin][y + dy][x + dx],
// A = weights (pre-packed per tap and block: lane (g, i) holds cout i,
// K chunk 8g..8g+7), B = activations (lane (g, x) holds K chunk g of pixel
// x), D: lane (g, x) holds couts 4g..4g+3 of pixel x.  In the forward pass a
// tile's activations stay in registers (16 rows x 4 VGPRs per lane, see
// fwd_channel); the dx = -1 / +1 taps are DPP row_shr:1 / row_shl:1 of a
// row, whose zero fill at the row end is exactly the zero padding of the
// tile edge.  LDS holds the forward weight packs and biases, staged once
// per workgroup, and per wave the input grid / layer 4's partial sums.
//
// Backward: the data gradient is the same convolution with flipped,
// transposed weight packs; the weight gradient is a GEMM over pixels
// (MFMA with gathered operands), accumulated into FP32 with atomics.
#include <algorithm>

#include "ncs_common.h"

namespace ncs {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef int i4 __attribute__((ext_vector_type(4)));

constexpr int kCh = 32;
// FP32 parameter vector: per layer weights [cout][cin][3][3] then bias
constexpr int OFF_W1 = 0, OFF_B1 = 576, OFF_W2 = 608, OFF_B2 = 9824, OFF_W3 = 9856,
              OFF_B3 = 19072, OFF_W4 = 19104, OFF_B4 = 19392, N_PARAMS = 19393;
static_assert(N_PARAMS == NCS_NET_PARAMS, "parameter count");
// MFMA operand packs (64 lanes x 8 BF16 = 1 KB each)
constexpr int PK_L1 = 0;    // 2: layer 1 A, K = (tap, cin) padded to 32
constexpr int PK_L2 = 2;    // 18: layer 2 A, block * 9 + tap, K order fwd_channel
constexpr int PK_L3 = 20;   // 18
constexpr int PK_L4T = 38;  // 3: layer 4 B per dy, column n = dx (n < 3), K order fwd_channel
constexpr int PK_B4 = 41;   // 2: layer-4 data gradient, K = tap padded to 32
constexpr int PK_B3 = 43;   // 18: layer-3 data gradient (flipped, transposed)
constexpr int PK_B2 = 61;   // 18: layer-2 data gradient
constexpr int N_PACKS = 79;

// Forward activations live in registers.  With the K index k = 8g + e of a
// 32-channel layer's operand mapped to channel fwd_channel(k) (e < 4: 4g + e,
// else 16 + 4g + e - 4), the D fragments of a row's two output blocks at lane
// (g, x) -- couts 4g..4g+3, then 16 + 4g..16 + 4g + 3, of pixel x -- are in
// that order exactly the 8-channel fragment lane (g, x) supplies to the next
// layer.  A tile's activations are 16 rows x 4 VGPRs per lane and never cross
// lanes; the packs of layers 2-4 carry the permutation in their K order.
__host__ __device__ constexpr int fwd_channel(int k) {
  return (k & 7) < 4 ? 4 * (k >> 3) + (k & 7) : 12 + 4 * (k >> 3) + (k & 7);
}

__device__ __forceinline__ float w_at(const float *P, int off, int cin_n, int co, int ci,
                                      int tap) {
  return P[off + (co * cin_n + ci) * 9 + tap];
}

// Builds the BF16 A-operand packs from the FP32 parameters.
__global__ void net_pack_kernel(const float *__restrict__ P, __bf16 *__restrict__ packs) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= N_PACKS * 512) return;
  const int pack = idx >> 9, lane = (idx >> 3) & 63, e = idx & 7;
  const int i = lane & 15, g = lane >> 4;
  const int k = 8 * g + e;
  float v = 0.0f;
  if (pack < PK_L2) {  // layer 1: k = tap * 2 + cin
    const int co = 16 * (pack - PK_L1) + i;
    if (k < 18) v = w_at(P, OFF_W1, 2, co, k & 1, k >> 1);
  } else if (pack < PK_L3) {
    const int q = pack - PK_L2, co = 16 * (q / 9) + i;
    v = w_at(P, OFF_W2, kCh, co, fwd_channel(k), q % 9);
  } else if (pack < PK_L4T) {
    const int q = pack - PK_L3, co = 16 * (q / 9) + i;
    v = w_at(P, OFF_W3, kCh, co, fwd_channel(k), q % 9);
  } else if (pack < PK_B4) {  // B operand: lane (g, n) holds K = cin 8g..8g+7 of column n
    const int dy = pack - PK_L4T;
    if (i < 3) v = w_at(P, OFF_W4, kCh, 0, fwd_channel(k), dy * 3 + i);
  } else if (pack < PK_B3) {  // d(layer-3 act) from d(out): cout' = L4 cin, k = tap
    const int co = 16 * (pack - PK_B4) + i;
    if (k < 9) v = w_at(P, OFF_W4, kCh, 0, co, 8 - k);
  } else if (pack < PK_B2) {
    const int q = pack - PK_B3, co = 16 * (q / 9) + i;  // co = L3 cin, k = L3 cout
    v = w_at(P, OFF_W3, kCh, k, co, 8 - q % 9);
  } else {
    const int q = pack - PK_B2, co = 16 * (q / 9) + i;
    v = w_at(P, OFF_W2, kCh, k, co, 8 - q % 9);
  }
  packs[idx] = (__bf16)v;
}

__device__ __forceinline__ f4 mfma_bf16(i4 a, i4 b, f4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a),
                                                 __builtin_bit_cast(bf16x8, b), c, 0, 0, 0);
}

template <int CTRL>
__device__ __forceinline__ i4 dpp_i4(i4 v) {
  i4 r;
  r.x = __builtin_amdgcn_update_dpp(0, v.x, CTRL, 0xF, 0xF, true);
  r.y = __builtin_amdgcn_update_dpp(0, v.y, CTRL, 0xF, 0xF, true);
  r.z = __builtin_amdgcn_update_dpp(0, v.z, CTRL, 0xF, 0xF, true);
  r.w = __builtin_amdgcn_update_dpp(0, v.w, CTRL, 0xF, 0xF, true);
  return r;
}
// A zero the optimiser cannot see through.  A loop-invariant address offset
// by it keeps its loads inside the per-tile loop (the operand packs of all
// four layers would not fit in registers beside the working set), and the
// address keeps its address space.
__device__ __forceinline__ int opaque_zero() {
  int z = 0;
  asm volatile("" : "+s"(z));
  return z;
}

// Orders one wave's LDS accesses across its lanes.  The LDS performs a
// wave's instructions in issue order, so only the compiler has to be kept
// from moving accesses across this point (wavefront-scope fences emit no
// instructions).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Accumulator start of the 16-channel block b: the bias of the output
// channels lane (g, x) holds (16 b + 4 g + r, r = 0..3).
__device__ __forceinline__ f4 bias4(const float *bias, int b, int lane) {
  const int g = lane >> 4;
  return (f4){bias[16 * b + 4 * g], bias[16 * b + 4 * g + 1], bias[16 * b + 4 * g + 2],
              bias[16 * b + 4 * g + 3]};
}

constexpr int kRowShr1 = 0x111;  // lane x <- lane x-1 in its row (0 at x = 0)
constexpr int kRowShl1 = 0x101;  // lane x <- lane x+1 in its row (0 at x = 15)

// Row y' of a [256][32] BF16 LDS activation buffer as a B operand
// (lane (g, x): channels 8g..8g+7 of pixel (y', x)); zero outside the tile.
__device__ __forceinline__ i4 load_row(const __bf16 *buf, int yy, int lane) {
  i4 z = {0, 0, 0, 0};
  if (yy < 0 || yy > 15) return z;
  const int x = lane & 15, g = lane >> 4;
  return *reinterpret_cast<const i4 *>(buf + ((yy * 16 + x) * kCh + 8 * g));
}

__device__ __forceinline__ void store4(__bf16 *buf, int y, int lane, int blk, f4 v) {
  const int x = lane & 15, g = lane >> 4;
  bf16x4 h;
  h.x = (__bf16)v.x;
  h.y = (__bf16)v.y;
  h.z = (__bf16)v.z;
  h.w = (__bf16)v.w;
  *reinterpret_cast<bf16x4 *>(buf + ((y * 16 + x) * kCh + 16 * blk + 4 * g)) = h;
}

__device__ __forceinline__ f4 load4(const __bf16 *buf, int y, int lane, int blk) {
  const int x = lane & 15, g = lane >> 4;
  bf16x4 h = *reinterpret_cast<const bf16x4 *>(buf + ((y * 16 + x) * kCh + 16 * blk + 4 * g));
  f4 r = {(float)h.x, (float)h.y, (float)h.z, (float)h.w};
  return r;
}

// One 32 -> 16*NB channel layer over a tile: out rows via fn(y, blk, acc).
// A operands: packs + (blk * 9 + tap); B: rows of `in`.
template <int NB, class Epi>
__device__ __forceinline__ void conv32(const __bf16 *in, const i4 *__restrict__ packs,
                                       int lane, Epi epi) {
  i4 A[NB][9];
#pragma unroll
  for (int b = 0; b < NB; b++)
#pragma unroll
    for (int t = 0; t < 9; t++) A[b][t] = packs[(b * 9 + t) * 64 + lane];
  i4 rm = {0, 0, 0, 0};
  i4 r0 = load_row(in, 0, lane);
  for (int y = 0; y < 16; y++) {
    const i4 rp = load_row(in, y + 1, lane);
    f4 acc[NB];
#pragma unroll
    for (int b = 0; b < NB; b++) acc[b] = (f4){0.f, 0.f, 0.f, 0.f};
    const i4 rows[3] = {rm, r0, rp};
#pragma unroll
    for (int dy = 0; dy < 3; dy++) {
      const i4 c = rows[dy];
      const i4 l = dpp_i4<kRowShr1>(c);
      const i4 r = dpp_i4<kRowShl1>(c);
#pragma unroll
      for (int b = 0; b < NB; b++) {
        acc[b] = mfma_bf16(A[b][dy * 3 + 0], l, acc[b]);
        acc[b] = mfma_bf16(A[b][dy * 3 + 1], c, acc[b]);
        acc[b] = mfma_bf16(A[b][dy * 3 + 2], r, acc[b]);
      }
    }
#pragma unroll
    for (int b = 0; b < NB; b++) epi(y, b, acc[b]);
    rm = r0;
    r0 = rp;
  }
}

// Two floats as packed BF16 (round to nearest even), low half first.
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f2 __attribute__((ext_vector_type(2)));
typedef short s2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ int bf16_pair(float lo, float hi) {
  const f2 v = {lo, hi};
  return __builtin_bit_cast(int, __builtin_convertvector(v, bf16x2));
}

// ReLU of a packed BF16 pair, clamping the halves as signed 16-bit integers
// at 0: a BF16 with its sign bit set is a negative int16 and a positive one
// keeps its order, so this equals rounding after the ReLU (negative values,
// -0 included, become +0).
__device__ __forceinline__ int relu_pair(float lo, float hi) {
  const s2 z = {0, 0};
  const s2 h = __builtin_bit_cast(s2, bf16_pair(lo, hi));
  return __builtin_bit_cast(int, __builtin_elementwise_max(h, z));
}

// ReLU of one row's two block accumulators as a BF16 register row.
__device__ __forceinline__ i4 relu_row(f4 a0, f4 a1) {
  i4 r;
  r.x = relu_pair(a0.x, a0.y);
  r.y = relu_pair(a0.z, a0.w);
  r.z = relu_pair(a1.x, a1.y);
  r.w = relu_pair(a1.z, a1.w);
  return r;
}

// A register row written as row y of a [256][32] BF16 buffer in the
// standard channel order (the activations kept for the backward pass).
__device__ __forceinline__ void store_row_std(__bf16 *dst, int y, int lane, i4 v) {
  const int x = lane & 15, g = lane >> 4;
  __bf16 *px = dst + (y * 16 + x) * kCh;
  *reinterpret_cast<int2 *>(px + 4 * g) = make_int2(v.x, v.y);
  *reinterpret_cast<int2 *>(px + 16 + 4 * g) = make_int2(v.z, v.w);
}

// Layer 1 (2 -> 32 channels), K = tap * 2 + cin padded to 32.  The inputs sit
// in LDS as an 18 x 18 grid of (x0, x1) float pairs with a zero border, so
// lane (g, x) reads its K = 8g..8g+7 (taps 4g..4g+3) of row y as four float2
// at fixed offsets; taps past 8 read the zero border column of the row.
__device__ __forceinline__ int in_pad_offset(int tap, int x) {
  return tap < 9 ? ((tap / 3) * 18 + x + tap % 3) * 2 : 0;
}

template <class Keep>
__device__ __forceinline__ void layer_in2(const float *in_pad, const i4 *packs,
                                          const float *bias, int lane, i4 (&act)[16],
                                          Keep keep) {
  const i4 A0 = packs[lane], A1 = packs[64 + lane];
  const f4 c0 = bias4(bias, 0, lane), c1 = bias4(bias, 1, lane);
  const int x = lane & 15, g = lane >> 4;
  int off[4];
#pragma unroll
  for (int j = 0; j < 4; j++) off[j] = in_pad_offset(4 * g + j, x);
#pragma unroll
  for (int y = 0; y < 16; y++) {
    i4 B;
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const float2 v = *reinterpret_cast<const float2 *>(in_pad + y * 36 + off[j]);
      B[j] = bf16_pair(v.x, v.y);
    }
    act[y] = relu_row(mfma_bf16(A0, B, c0), mfma_bf16(A1, B, c1));
    keep(y, act[y]);
  }
}

// A 32 -> 32 channel layer in place over the register rows, two output rows
// per step (four independent accumulation chains: two rows x two blocks).
// Input rows y - 1 .. y + 2 and their dx = -1 / +1 shifts (DPP row_shr:1 /
// row_shl:1, whose zero fill at the row end is the tile's zero padding) are
// held as values when rows y and y + 1 are overwritten; rows y + 1 and y + 2
// carry over to the next step, so each row is shifted once per layer.
template <class Keep>
__device__ __forceinline__ void layer32(i4 (&act)[16], const i4 *packs_in, const float *bias_in,
                                        int lane, Keep keep) {
  const int z0 = opaque_zero();
  const i4 *packs = packs_in + z0;
  const float *bias = bias_in + z0;
  i4 A[2][9];
#pragma unroll
  for (int b = 0; b < 2; b++)
#pragma unroll
    for (int t = 0; t < 9; t++) A[b][t] = packs[(b * 9 + t) * 64 + lane];
  const f4 c0 = bias4(bias, 0, lane), c1 = bias4(bias, 1, lane);
  const i4 z = {0, 0, 0, 0};
  i4 C[4], L[4], R[4];  // rows y - 1 .. y + 2: centre, dx = -1, dx = +1
  C[0] = L[0] = R[0] = z;
  C[1] = act[0];
  L[1] = dpp_i4<kRowShr1>(C[1]);
  R[1] = dpp_i4<kRowShl1>(C[1]);
#pragma unroll
  for (int y = 0; y < 16; y += 2) {
    C[2] = act[y + 1];
    L[2] = dpp_i4<kRowShr1>(C[2]);
    R[2] = dpp_i4<kRowShl1>(C[2]);
    if (y + 2 < 16) {
      C[3] = act[y + 2];
      L[3] = dpp_i4<kRowShr1>(C[3]);
      R[3] = dpp_i4<kRowShl1>(C[3]);
    }
    f4 acc[2][2] = {{c0, c1}, {c0, c1}};
#pragma unroll
    for (int q = 0; q < 4; q++) {  // input row y - 1 + q
      if ((q == 0 && y == 0) || (q == 3 && y + 2 >= 16)) continue;  // zero rows
#pragma unroll
      for (int dx = 0; dx < 3; dx++) {
        const i4 in = dx == 0 ? L[q] : (dx == 1 ? C[q] : R[q]);
#pragma unroll
        for (int o = 0; o < 2; o++) {  // output row y + o: dy = q - o
          const int dy = q - o;
          if (dy < 0 || dy > 2) continue;
#pragma unroll
          for (int b = 0; b < 2; b++) acc[o][b] = mfma_bf16(A[b][dy * 3 + dx], in, acc[o][b]);
        }
      }
    }
    act[y] = relu_row(acc[0][0], acc[0][1]);
    keep(y, act[y]);
    act[y + 1] = relu_row(acc[1][0], acc[1][1]);
    keep(y + 1, act[y + 1]);
    C[0] = C[2];
    L[0] = L[2];
    R[0] = R[2];
    C[1] = C[3];
    L[1] = L[3];
    R[1] = R[3];
  }
}

// Layer 4 (32 -> 1 channel) with the register rows as the A operand: lane
// (g, x) holds A[x][8g..8g+7] (pixel x, K chunk g), and B_dy[k][n] =
// W4[fwd_channel(k)][dy][n] for n = dx < 3.  Three chained MFMA per row give
// D_y[x'][n] = sum_dy sum_ci a3[ci][y + dy - 1][x'] W4[ci][dy][n] (D lane
// (g, n) holds x' = 4g..4g+3); all 16 rows go to dbuf ([16][16][3] FP32 LDS)
// and, after one barrier,
//   out[y][x] = b4 + D_y[x - 1][0] + D_y[x][1] + D_y[x + 1][2]
// (zero past the tile) for the pixels (4g + r, x) of lane (g, x), r = 0..3.
__device__ __forceinline__ f4 layer_out1(const i4 (&act)[16], const i4 *packs,
                                         float *dbuf, float b4, int lane) {
  const i4 B0 = packs[lane], B1 = packs[64 + lane], B2 = packs[128 + lane];
  const int n = lane & 15, g = lane >> 4;
#pragma unroll
  for (int y = 0; y < 16; y++) {
    f4 d = {0.f, 0.f, 0.f, 0.f};
    if (y > 0) d = mfma_bf16(act[y - 1], B0, d);
    d = mfma_bf16(act[y], B1, d);
    if (y < 15) d = mfma_bf16(act[y + 1], B2, d);
    if (n < 3)
#pragma unroll
      for (int r = 0; r < 4; r++) dbuf[(y * 16 + 4 * g + r) * 3 + n] = d[r];
  }
  wave_sync();
  f4 out;
#pragma unroll
  for (int r = 0; r < 4; r++) {
    const float *D = dbuf + (4 * g + r) * 48;
    float v = b4 + D[n * 3 + 1];
    if (n > 0) v += D[(n - 1) * 3 + 0];
    if (n < 15) v += D[(n + 1) * 3 + 2];
    out[r] = v;
  }
  return out;
}

// Writes the zero border of the 18 x 18 input grid (68 cells).
__device__ __forceinline__ void in_pad_border(float *in_pad, int lane) {
  for (int c = lane; c < 68; c += 64) {
    int cell;
    if (c < 18) cell = c;                          // top row
    else if (c < 36) cell = 17 * 18 + (c - 18);    // bottom row
    else if (c < 52) cell = (c - 35) * 18;         // left column, rows 1..16
    else cell = (c - 51) * 18 + 17;                // right column, rows 1..16
    *reinterpret_cast<float2 *>(in_pad + 2 * cell) = make_float2(0.0f, 0.0f);
  }
}

__device__ __forceinline__ void in_pad_put(float *in_pad, int row, int col, float x0, float x1) {
  *reinterpret_cast<float2 *>(in_pad + 2 * ((row + 1) * 18 + col + 1)) = make_float2(x0, x1);
}

// A layer whose input has NC <= 3 channels (K = tap * NC + c padded to 32),
// input FP32 in LDS as [NC][256]; 2 output blocks, accumulators starting
// from `bias` (zero when null).
template <int NC, class Epi>
__device__ __forceinline__ void conv_small(const float *in, const i4 *__restrict__ packs,
                                           int lane, Epi epi, const float *bias = nullptr) {
  const i4 A0 = packs[lane], A1 = packs[64 + lane];
  f4 c0 = {0.f, 0.f, 0.f, 0.f}, c1 = c0;
  if (bias != nullptr) {
    c0 = bias4(bias, 0, lane);
    c1 = bias4(bias, 1, lane);
  }
  const int x = lane & 15, g = lane >> 4;
  for (int y = 0; y < 16; y++) {
    bf16x8 bv;
#pragma unroll
    for (int e = 0; e < 8; e++) {
      const int k = 8 * g + e;
      float v = 0.0f;
      if (k < 9 * NC) {
        const int tap = k / NC, ch = k % NC;
        const int yy = y + tap / 3 - 1, xx = x + tap % 3 - 1;
        if (yy >= 0 && yy < 16 && xx >= 0 && xx < 16) v = in[ch * 256 + yy * 16 + xx];
      }
      bv[e] = (__bf16)v;
    }
    const i4 B = __builtin_bit_cast(i4, bv);
    epi(y, 0, mfma_bf16(A0, B, c0));
    epi(y, 1, mfma_bf16(A1, B, c1));
  }
}

// Normalises one tile held in the solver's lane layout (lane (g, x) has
// rows 4g + r, column x) into the padded input grid; returns sigma.
__device__ __forceinline__ void tile_inputs(const float *d_tile_lane, const float *gm_lane,
                                            float *in_pad, int lane, float &sigma) {
  float red[2] = {d_tile_lane[0] + d_tile_lane[1] + (d_tile_lane[2] + d_tile_lane[3]),
                  gm_lane[0] + gm_lane[1] + (gm_lane[2] + gm_lane[3])};
  for (int off = 32; off >= 1; off >>= 1) {
    red[0] += __shfl_xor(red[0], off);
    red[1] += __shfl_xor(red[1], off);
  }
  const float mu = red[0] * (1.0f / 256.0f), gmean = red[1] * (1.0f / 256.0f);
  sigma = sqrtf(fmaxf(mu, 1.0f) + gmean);
  const float inv = 1.0f / sigma;
  const int x = lane & 15, g = lane >> 4;
  in_pad_border(in_pad, lane);
#pragma unroll
  for (int r = 0; r < 4; r++)
    in_pad_put(in_pad, 4 * g + r, x, (d_tile_lane[r] - mu) * inv, gm_lane[r] * inv * inv);
}

// The network body on one tile whose inputs are in the padded grid `io`
// (18 x 18 float pairs, then reused as layer 4's [16][16][3] partial sums);
// returns the output y at pixels (4g + r, x) of lane (g, x).  pk / bl: the
// forward packs and biases staged in LDS.  acts (may be NULL): post-ReLU
// layer 1-3 activations, BF16 [3][n][256][32] in the standard channel order,
// of tile t.  One wave runs it; io is the wave's own.
__device__ __forceinline__ f4 net_body(float *io, const i4 *pk, const float *bl, int lane,
                                       __bf16 *acts, long long n, long long t) {
  auto keep = [&](int layer) {
    return [=](int y, i4 v) {
      if (acts != nullptr)
        store_row_std(acts + (((long long)layer * n + t) * 256) * kCh, y, lane, v);
    };
  };
  const int z = opaque_zero();
  i4 act[16];
  wave_sync();  // the input grid is complete
  layer_in2(io, pk + (PK_L1 - PK_L1) * 64 + z, bl + z, lane, act, keep(0));
  layer32(act, pk + (PK_L2 - PK_L1) * 64, bl + 32, lane, keep(1));
  layer32(act, pk + (PK_L3 - PK_L1) * 64, bl + 64, lane, keep(2));
  wave_sync();  // every read of the grid precedes layer 4's writes over it
  return layer_out1(act, pk + (PK_L4T - PK_L1) * 64 + z, io, bl[96 + z], lane);
}

// Forward workgroups: kNetWaves waves, each running its own tiles, share the
// forward packs (layers 1-4, 41 KB) and biases staged once in LDS; each
// wave owns an input grid / layer-4 buffer and a copy of its tile's data.
constexpr int kNetWaves = 4;
constexpr int kFwdPacks = PK_B4 - PK_L1;
constexpr int kNetIoFloats = 768;  // max(18 * 18 * 2, 16 * 16 * 3)
constexpr int kNetWaveFloats = kNetIoFloats + 256;
constexpr int kNetFwdLdsBytes = kFwdPacks * 1024 + 128 * 4 + kNetWaves * kNetWaveFloats * 4;
constexpr int kNetLdsBytes = 2 * 256 * kCh * 2 + 3 * 256 * 4;  // backward: two buffers, x, dY

// Copies the forward packs and the biases [b1 | b2 | b3 | b4] into LDS.
__device__ __forceinline__ void stage_forward(const i4 *__restrict__ packs,
                                              const float *__restrict__ P, i4 *pk, float *bl) {
  for (int i = threadIdx.x; i < kFwdPacks * 64; i += blockDim.x) pk[i] = packs[PK_L1 * 64 + i];
  for (int i = threadIdx.x; i < 97; i += blockDim.x)
    bl[i] = i < 32 ? P[OFF_B1 + i]
                   : (i < 64 ? P[OFF_B2 + i - 32] : (i < 96 ? P[OFF_B3 + i - 64] : P[OFF_B4]));
  __syncthreads();
}

// Training / tensor mode: x [n][2][256] normalised inputs -> y [n][256].
// One wave per tile, its activations in registers.
__global__ __launch_bounds__(64 * kNetWaves) __attribute__((amdgpu_waves_per_eu(2)))
void net_fwd_kernel(const float *__restrict__ xin, float *__restrict__ yout,
                    __bf16 *__restrict__ acts, const i4 *__restrict__ packs,
                    const float *__restrict__ P, long long n) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  i4 *pk = reinterpret_cast<i4 *>(lds_raw);
  float *bl = reinterpret_cast<float *>(pk + kFwdPacks * 64);
  // the wave index is uniform: read it into a scalar register so the tile
  // loop's index arithmetic and frame-table reads stay on the scalar unit
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  float *io = bl + 128 + wave * kNetWaveFloats;
  const int col = lane & 15, grp = lane >> 4;
  stage_forward(packs, P, pk, bl);
  for (long long t = (long long)blockIdx.x * kNetWaves + wave; t < n;
       t += (long long)gridDim.x * kNetWaves) {
    in_pad_border(io, lane);
    for (int i = lane; i < 256; i += 64)
      in_pad_put(io, i >> 4, i & 15, xin[t * 512 + i], xin[t * 512 + 256 + i]);
    const f4 y = net_body(io, pk, bl, lane, acts, n, t);
#pragma unroll
    for (int r = 0; r < 4; r++) yout[t * 256 + (4 * grp + r) * 16 + col] = y[r];
    wave_sync();
  }
}

typedef __attribute__((address_space(4))) const FrameDev FrameConstN;

// Inference mode over frames: gather each window exactly like the solver,
// normalise, run the net and write u0 = d + sigma * y as [tiles][256].
__global__ __launch_bounds__(64 * kNetWaves) __attribute__((amdgpu_waves_per_eu(2)))
void net_warm_kernel(const FrameDev *__restrict__ frames, int n_frames, long long total,
                     float *__restrict__ u0, const i4 *__restrict__ packs,
                     const float *__restrict__ P) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  i4 *pk = reinterpret_cast<i4 *>(lds_raw);
  float *bl = reinterpret_cast<float *>(pk + kFwdPacks * 64);
  // the wave index is uniform: read it into a scalar register so the tile
  // loop's index arithmetic and frame-table reads stay on the scalar unit
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  // the frame table through the constant address space (scalar loads; the
  // kernel never writes it)
  const FrameConstN *fc = (const FrameConstN *)frames;
  float *io = bl + 128 + wave * kNetWaveFloats;
  float *dkeep = io + kNetIoFloats;
  const int col = lane & 15, grp = lane >> 4;
  stage_forward(packs, P, pk, bl);
  for (long long t = (long long)blockIdx.x * kNetWaves + wave; t < total;
       t += (long long)gridDim.x * kNetWaves) {
    int lo = 0, hi = n_frames - 1;
    while (lo < hi) {
      int mid = (lo + hi + 1) >> 1;
      if (fc[mid].first <= t) lo = mid; else hi = mid - 1;
    }
    const FrameConstN &F = fc[lo];
    const int local = (int)(t - F.first);  // a frame has fewer than 2^31 windows
    const int a = local / F.ty, b = local - a * F.ty;
    int bx, by, hv = 16, wv = 16;
    if (F.geom == NCS_GEOM_REFERENCE) {
      bx = min(-1 + 14 * a, F.H - 15);
      by = min(-1 + 14 * b, F.W - 15);
    } else {
      bx = 16 * a;
      by = 16 * b;
      hv = min(16, F.H - bx);
      wv = min(16, F.W - by);
    }
    float d[4], gm[4];
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int row = 4 * grp + r;
      // 32-bit element offsets from the uniform base pointers (the host keeps
      // every frame below 2^31 elements)
      if (F.geom == NCS_GEOM_REFERENCE) {
        const int gr = min(max(bx + row, 0), F.H - 1), gc = min(max(by + col, 0), F.W - 1);
        d[r] = F.data[(unsigned int)(gr * F.ds + gc)];
        gm[r] = F.gamma[(unsigned int)(gr * F.gs + gc)];
      } else {
        const bool ok = row < hv && col < wv;
        d[r] = ok ? F.data[(unsigned int)((bx + row) * F.ds + by + col)] : 0.0f;
        gm[r] = ok ? F.gamma[(unsigned int)((bx + row) * F.gs + by + col)] : 0.0f;
      }
    }
    float sigma;
    tile_inputs(d, gm, io, lane, sigma);
#pragma unroll
    for (int r = 0; r < 4; r++) dkeep[r * 64 + lane] = d[r];
    const f4 y = net_body(io, pk, bl, lane, nullptr, 0, 0);
    // starts are indexed in the caller's frame order
    float *const w = u0 + (F.slot + local) * 256;
#pragma unroll
    for (int r = 0; r < 4; r++)
      w[(4 * grp + r) * 16 + col] = dkeep[r * 64 + lane] + sigma * y[r];
    wave_sync();
  }
}

// ---------------------------------------------------------------------------
// backward
// ---------------------------------------------------------------------------

// dW[co][ci][tap] += sum_p dz[p][co] * act[p + tap][ci] for 32-channel dz / act
// ([256][32] BF16 LDS), MFMA over pixel chunks of 32 with gathered operands.
__device__ __forceinline__ void grad_w32(const __bf16 *dz, const __bf16 *act, float *gW,
                                         int lane) {
  const int i = lane & 15, g = lane >> 4;
  for (int tap = 0; tap < 9; tap++) {
    const int dy = tap / 3 - 1, dx = tap % 3 - 1;
#pragma unroll
    for (int bo = 0; bo < 2; bo++)
#pragma unroll
      for (int bi = 0; bi < 2; bi++) {
        f4 acc = {0.f, 0.f, 0.f, 0.f};
        for (int ch = 0; ch < 8; ch++) {  // pixel chunks of 32
          bf16x8 av, bv;
#pragma unroll
          for (int e = 0; e < 8; e++) {
            const int p = ch * 32 + 8 * g + e;
            av[e] = dz[p * kCh + 16 * bo + i];
            const int yy = (p >> 4) + dy, xx = (p & 15) + dx;
            bv[e] = (yy >= 0 && yy < 16 && xx >= 0 && xx < 16)
                        ? act[(yy * 16 + xx) * kCh + 16 * bi + i]
                        : (__bf16)0.0f;
          }
          acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc, 0, 0, 0);
        }
        // acc: lane (g, i) holds co = 16 bo + 4g + r, ci = 16 bi + i
#pragma unroll
        for (int r = 0; r < 4; r++)
          atomicAdd(&gW[((16 * bo + 4 * g + r) * kCh + 16 * bi + i) * 9 + tap], acc[r]);
      }
  }
}

// d(bias)[co] += sum_p dz[p][co]
__device__ __forceinline__ void grad_b32(const __bf16 *dz, float *gb, int lane) {
  if (lane < kCh) {
    float s = 0.0f;
    for (int p = 0; p < 256; p++) s += (float)dz[p * kCh + lane];
    atomicAdd(&gb[lane], s);
  }
}

// Data gradient of a 32-channel layer output given via packs (flipped,
// transposed weights), masked by the ReLU of `mask_act` and written in place
// of it (each lane rewrites exactly the entries it read).
__device__ __forceinline__ void grad_data32(const __bf16 *dz_next, __bf16 *act_inout,
                                            const i4 *__restrict__ packs, int lane) {
  conv32<2>(dz_next, packs, lane, [&](int y, int b, f4 acc) {
    f4 a = load4(act_inout, y, lane, b);
    f4 v;
    v.x = a.x > 0.0f ? acc.x : 0.0f;
    v.y = a.y > 0.0f ? acc.y : 0.0f;
    v.z = a.z > 0.0f ? acc.z : 0.0f;
    v.w = a.w > 0.0f ? acc.w : 0.0f;
    store4(act_inout, y, lane, b, v);
  });
}

__global__ __launch_bounds__(64) void net_bwd_kernel(const float *__restrict__ xin,
                                                     const __bf16 *__restrict__ acts,
                                                     const float *__restrict__ gout,
                                                     float *__restrict__ gP,
                                                     const i4 *__restrict__ packs,
                                                     long long n) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  __bf16 *buf0 = reinterpret_cast<__bf16 *>(lds_raw);
  __bf16 *buf1 = buf0 + 256 * kCh;
  float *in2 = reinterpret_cast<float *>(buf1 + 256 * kCh);  // x, then dY
  float *dy4 = in2 + 512;
  const int lane = threadIdx.x;
  const int i = lane & 15, g = lane >> 4;
  auto load_act = [&](__bf16 *dst, int layer, long long t) {
    const i4 *src = reinterpret_cast<const i4 *>(acts + (((long long)layer * n + t) * 256) * kCh);
    i4 *d = reinterpret_cast<i4 *>(dst);
    for (int k = lane; k < 256 * kCh / 8; k += 64) d[k] = src[k];
  };
  for (long long t = blockIdx.x; t < n; t += gridDim.x) {
    for (int k = lane; k < 256; k += 64) dy4[k] = gout[t * 256 + k];
    load_act(buf0, 2, t);  // layer-3 activations
    __syncthreads();
    // layer 4: dW4[0][ci][tap] = sum_p dy4[p] a3[p + tap][ci]; db4 = sum dy4
    for (int q = lane; q < 9 * kCh; q += 64) {
      const int ci = q % kCh, tap = q / kCh, dy = tap / 3 - 1, dx = tap % 3 - 1;
      float s = 0.0f;
      for (int p = 0; p < 256; p++) {
        const int yy = (p >> 4) + dy, xx = (p & 15) + dx;
        if (yy >= 0 && yy < 16 && xx >= 0 && xx < 16)
          s += dy4[p] * (float)buf0[(yy * 16 + xx) * kCh + ci];
      }
      atomicAdd(&gP[OFF_W4 + ci * 9 + tap], s);
    }
    if (lane == 0) {
      float s = 0.0f;
      for (int p = 0; p < 256; p++) s += dy4[p];
      atomicAdd(&gP[OFF_B4], s);
    }
    // dz3 = conv(dy4, flipped W4) masked by a3 > 0 -> buf1
    conv_small<1>(dy4, packs + PK_B4 * 64, lane, [&](int y, int b, f4 acc) {
      f4 a = load4(buf0, y, lane, b);
      f4 v;
      v.x = a.x > 0.0f ? acc.x : 0.0f;
      v.y = a.y > 0.0f ? acc.y : 0.0f;
      v.z = a.z > 0.0f ? acc.z : 0.0f;
      v.w = a.w > 0.0f ? acc.w : 0.0f;
      store4(buf1, y, lane, b, v);
    });
    __syncthreads();
    load_act(buf0, 1, t);  // layer-2 activations
    __syncthreads();
    grad_w32(buf1, buf0, gP + OFF_W3, lane);
    grad_b32(buf1, gP + OFF_B3, lane);
    __syncthreads();
    grad_data32(buf1, buf0, packs + PK_B3 * 64, lane);  // buf0 = dz2
    __syncthreads();
    load_act(buf1, 0, t);  // layer-1 activations
    __syncthreads();
    grad_w32(buf0, buf1, gP + OFF_W2, lane);
    grad_b32(buf0, gP + OFF_B2, lane);
    __syncthreads();
    grad_data32(buf0, buf1, packs + PK_B2 * 64, lane);  // buf1 = dz1
    // layer 1: dW1[co][ci][tap] = sum_p dz1[p][co] x[ci][p + tap]; K = pixels
    for (int k = lane; k < 512; k += 64) in2[k] = xin[t * 512 + k];
    __syncthreads();
    for (int tap = 0; tap < 9; tap++) {
      const int dy = tap / 3 - 1, dx = tap % 3 - 1;
#pragma unroll
      for (int bo = 0; bo < 2; bo++) {
        f4 acc = {0.f, 0.f, 0.f, 0.f};
        for (int ch = 0; ch < 8; ch++) {
          bf16x8 av, bv;
#pragma unroll
          for (int e = 0; e < 8; e++) {
            const int p = ch * 32 + 8 * g + e;
            av[e] = buf1[p * kCh + 16 * bo + i];
            const int yy = (p >> 4) + dy, xx = (p & 15) + dx;
            bv[e] = (i < 2 && yy >= 0 && yy < 16 && xx >= 0 && xx < 16)
                        ? (__bf16)in2[i * 256 + yy * 16 + xx]
                        : (__bf16)0.0f;
          }
          acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc, 0, 0, 0);
        }
        if (i < 2)
#pragma unroll
          for (int r = 0; r < 4; r++)
            atomicAdd(&gP[OFF_W1 + ((16 * bo + 4 * g + r) * 2 + i) * 9 + tap], acc[r]);
      }
    }
    grad_b32(buf1, gP + OFF_B1, lane);
    __syncthreads();
  }
}

__global__ void fill_kernel(float *p, long long n, float v) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    p[i] = v;
}

// ---------------------------------------------------------------------------
// launches
// ---------------------------------------------------------------------------

hipError_t net_pack(const float *P, void *packs, hipStream_t s) {
  const int total = N_PACKS * 512;
  hipLaunchKernelGGL(net_pack_kernel, dim3((total + 255) / 256), dim3(256), 0, s, P,
                     reinterpret_cast<__bf16 *>(packs));
  return hipGetLastError();
}

size_t net_packs_bytes() { return (size_t)N_PACKS * 512 * 2; }

// Forward workgroups (kNetWaves waves, one tile per wave at a time) that fill
// the device: the occupancy limit of the forward kernels times the CUs, at
// most one tile per wave.
int net_forward_grid(long long n, int cus) {
  int blocks = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
          &blocks, reinterpret_cast<const void *>(net_warm_kernel), 64 * kNetWaves,
          kNetFwdLdsBytes) != hipSuccess ||
      blocks < 1)
    blocks = 1;
  const long long need = (n + kNetWaves - 1) / kNetWaves;
  return (int)std::max<long long>(1, std::min<long long>(need, (long long)blocks * cus));
}

hipError_t net_forward(const float *x, float *y, void *acts, const void *packs, const float *P,
                       long long n, int grid, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(net_fwd_kernel, dim3(grid), dim3(64 * kNetWaves), kNetFwdLdsBytes, s, x, y,
                     reinterpret_cast<__bf16 *>(acts), reinterpret_cast<const i4 *>(packs), P, n);
  return hipGetLastError();
}

hipError_t net_backward(const float *x, const void *acts, const float *gout, float *gP,
                        const void *packs, long long n, int grid, hipStream_t s) {
  hipLaunchKernelGGL(fill_kernel, dim3(64), dim3(256), 0, s, gP, (long long)N_PARAMS, 0.0f);
  if (n <= 0) return hipGetLastError();
  hipLaunchKernelGGL(net_bwd_kernel, dim3(grid), dim3(64), kNetLdsBytes, s, x,
                     reinterpret_cast<const __bf16 *>(acts), gout, gP,
                     reinterpret_cast<const i4 *>(packs), n);
  return hipGetLastError();
}

hipError_t net_warm(const FrameDev *frames, int n_frames, long long total, float *u0,
                    const void *packs, const float *P, int grid, hipStream_t s) {
  if (total <= 0) return hipSuccess;
  hipLaunchKernelGGL(net_warm_kernel, dim3(grid), dim3(64 * kNetWaves), kNetFwdLdsBytes, s,
                     frames, n_frames,
                     total, u0, reinterpret_cast<const i4 *>(packs), P);
  return hipGetLastError();
}

}  // namespace ncs
