// ncs_common.h -- shared device/host definitions of libncs_hip.so (gfx950).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/ncs_hip.h"

namespace ncs {

constexpr int kTile = 16;          // tile edge (the reference's r_size = 16)
constexpr int kTilePix = 256;      // pixels per tile
constexpr int kWave = 64;          // gfx950 wavefront
constexpr int kWavesPerBlock = 4;  // 256-thread workgroups, one tile per wave
constexpr int kMaxWavesPerGroup = 4;  // persistent solver: waves per workgroup 1..4

// Work queue of the persistent solver: 7/8 of the tiles are assigned to the
// waves statically in contiguous blocks; the rest are split into kQueues
// contiguous ranges, one per XCD (workgroups are dispatched round-robin over
// the 8 XCDs), each with its own counter on its own 128-byte line, taken
// kChunk tiles per atomic.
constexpr int kQueues = 8;
constexpr int kQueueStride = 32;   // unsigned ints between counters
constexpr int kChunk = 4;
constexpr size_t kQueueBytes = kQueues * kQueueStride * sizeof(unsigned int);
constexpr float kFitMin = 1.0e-6f; // negative-value guard (ncs.cl:45, ncs.c:169)

// Per-lane tables, kTableF4 16-byte entries per lane: the real-DFT basis
// entries in FP32 (two), the mask's even-even and odd-odd weights, the
// learned linear start's two coefficient sets, the basis entries again as
// FP16 hi/lo pairs for the FP16 matrix-core basis changes (four; see
// ncs_solver.hip, LdsTables), and the linear start's two slopes in
// ln(alpha) (kTableSlope, kTableSlope + 1).  The solver keeps entries
// kSolveTableFirst .. kSolveTableFirst + kSolveTableF4 - 1 in LDS; the
// slopes, read once per tile, stay in the global table.
constexpr int kTableF4 = 12;
constexpr int kSolveTableFirst = 2;
constexpr int kSolveTableF4 = 8;
constexpr int kTableSlope = 10;
constexpr int kTableLaneFloats = 4 * kTableF4;
constexpr int kTableFloats = kWave * kTableLaneFloats;

// One frame of a batch, as the kernel sees it.  A batch's tiles are claimed
// in frame-table order (`first`, increasing along the table); per-tile arrays
// (warm starts, tile outputs, iterations, status, cost) are indexed in the
// caller's frame order (`slot`).  The two differ when the host schedules the
// frames in another order than the caller passed them.
struct FrameDev {
  const float *data;
  const float *gamma;
  float *out;
  int H, W;
  int ds, gs, os;      // row strides (elements)
  float alpha;
  int geom;            // NCS_GEOM_*
  int ty;              // tiles along the fast axis
  int win;             // window side r of the reference geometry (even, 4..16);
                       // a window fills the top-left r x r of the 16 x 16 tile
  long long first;     // schedule position of the frame's first tile
  long long slot;      // per-tile array index of the frame's first tile
};

// Solver parameters in the kernel's units.
struct SolveArgs {
  const FrameDev *frames;
  int n_frames;
  long long total_tiles;
  unsigned int *queue;         // kQueues work-queue counters (kQueueBytes)
  const float *tables;         // [64][16] per-lane OTF/DFT tables
  const float *u0;             // [tiles][256] warm start or nullptr
  float *u_tiles;              // [tiles][256] solved tiles or nullptr
  int *iterations;             // [tiles] or nullptr
  int *status;                 // [tiles] or nullptr
  float *cost;                 // [tiles] or nullptr
  int *evaluations;            // [tiles] or nullptr
  int max_it;
  float eps;
  int ls;
  int max_ls;
  float ftol, gtol, xtol, min_step, max_step, stepm;
  int fallback;
  int precond;                 // spectral initial inverse Hessian (else s.y / y.y)
  int lin_start;               // learned linear start (coefficients in the tables)
  float lin_c;                 // its DC bias
  float lin_c1;                // the bias's slope in ln(alpha)
  float lin_aref;              // alpha the coefficients were fitted at
  int lin_slope;               // nonzero: the coefficients follow alpha
  float k_norm;                // bound on the NC Hessian's norm (row sums of the mask tables)
};

}  // namespace ncs
