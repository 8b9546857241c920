// ncs_camera.hip -- sCMOS camera conversion on the device (gfx950).
//
// Raw camera frames (uint16 ADU) become the photo-electron images the solver
// takes, and the camera's variance map becomes gamma:
//   data  = (ADU - offset) / gain, values <= 0 set to 1e-6
//   gamma = variance / gain^2
// (python3-6/pyNCS/denoisetools.py:218-224 and 266; the same conversion in
// python3-6/NCSdemo_experiment.py:42-49).  Both are HBM-bound streams: 10 B
// read and 4 B written per pixel for the frame, so one 2304^2 frame moves
// 74 MB.  Eight pixels per lane (one 16-byte load of ADU, two of each map).
#include <algorithm>

#include "ncs_common.h"

namespace ncs {

__device__ __forceinline__ float adu_to_photons(unsigned int adu, float offset, float gain) {
  // IEEE division, as the host reference computes it in float32
  const float d = ((float)adu - offset) / gain;
  return d <= 0.0f ? 1.0e-6f : d;
}

__global__ __launch_bounds__(256) void adu_kernel(const uint16_t *__restrict__ adu,
                                                  const float *__restrict__ offset,
                                                  const float *__restrict__ gain,
                                                  float *__restrict__ out, long long n) {
  const long long stride = (long long)gridDim.x * blockDim.x * 8;
  for (long long i = ((long long)blockIdx.x * blockDim.x + threadIdx.x) * 8; i < n;
       i += stride) {
    if (i + 8 <= n) {
      const uint4 a = *reinterpret_cast<const uint4 *>(adu + i);
      const float4 o0 = *reinterpret_cast<const float4 *>(offset + i);
      const float4 o1 = *reinterpret_cast<const float4 *>(offset + i + 4);
      const float4 g0 = *reinterpret_cast<const float4 *>(gain + i);
      const float4 g1 = *reinterpret_cast<const float4 *>(gain + i + 4);
      float4 r0, r1;
      r0.x = adu_to_photons(a.x & 0xFFFFu, o0.x, g0.x);
      r0.y = adu_to_photons(a.x >> 16, o0.y, g0.y);
      r0.z = adu_to_photons(a.y & 0xFFFFu, o0.z, g0.z);
      r0.w = adu_to_photons(a.y >> 16, o0.w, g0.w);
      r1.x = adu_to_photons(a.z & 0xFFFFu, o1.x, g1.x);
      r1.y = adu_to_photons(a.z >> 16, o1.y, g1.y);
      r1.z = adu_to_photons(a.w & 0xFFFFu, o1.z, g1.z);
      r1.w = adu_to_photons(a.w >> 16, o1.w, g1.w);
      *reinterpret_cast<float4 *>(out + i) = r0;
      *reinterpret_cast<float4 *>(out + i + 4) = r1;
    } else {
      for (long long j = i; j < n; j++) out[j] = adu_to_photons(adu[j], offset[j], gain[j]);
    }
  }
}

__global__ __launch_bounds__(256) void gamma_kernel(const float *__restrict__ var,
                                                    const float *__restrict__ gain,
                                                    float *__restrict__ gamma, long long n) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const float g = gain[i];
    gamma[i] = var[i] / (g * g);
  }
}

// Several gain amplifiers (matlab/genmaps.m:29-62): every pixel of every
// frame picks its level by its own ADU count, L = the number of division
// points <= ADU (the points ascend; unused ones are 65536, above any ADU),
// and then its gain and gamma maps.  gains/gammas are [levels][stride]
// floats.  Photons as adu_kernel; gamma is written per frame.  Four pixels per
// lane: one 8-byte ADU load, 16-byte offset loads, and per pixel one gain and
// one gamma load from the level it picked.
__global__ __launch_bounds__(256) void adu_levels_kernel(
    const uint16_t *__restrict__ adu, const float *__restrict__ offset,
    const float *__restrict__ gains, const float *__restrict__ gammas, long long stride,
    int t0, int t1, int t2, float *__restrict__ out, float *__restrict__ gamma, long long n) {
  const long long step = (long long)gridDim.x * blockDim.x * 4;
  for (long long i = ((long long)blockIdx.x * blockDim.x + threadIdx.x) * 4; i < n; i += step) {
    const int m = (int)(n - i < 4 ? n - i : 4);
    unsigned int a[4];
    float o[4];
    if (m == 4) {
      const uint2 v = *reinterpret_cast<const uint2 *>(adu + i);
      a[0] = v.x & 0xFFFFu;
      a[1] = v.x >> 16;
      a[2] = v.y & 0xFFFFu;
      a[3] = v.y >> 16;
      const float4 ov = *reinterpret_cast<const float4 *>(offset + i);
      o[0] = ov.x;
      o[1] = ov.y;
      o[2] = ov.z;
      o[3] = ov.w;
    } else {
      for (int k = 0; k < 4; k++) {
        a[k] = k < m ? adu[i + k] : 0u;
        o[k] = k < m ? offset[i + k] : 0.0f;
      }
    }
    float d[4], g[4];
    for (int k = 0; k < 4; k++) {
      const int v = (int)a[k];
      const long long L = (long long)((v >= t0) + (v >= t1) + (v >= t2));
      const long long j = k < m ? L * stride + i + k : i;  // in range for every lane
      d[k] = adu_to_photons(a[k], o[k], gains[j]);
      g[k] = gammas[j];
    }
    if (m == 4) {
      *reinterpret_cast<float4 *>(out + i) = make_float4(d[0], d[1], d[2], d[3]);
      *reinterpret_cast<float4 *>(gamma + i) = make_float4(g[0], g[1], g[2], g[3]);
    } else {
      for (int k = 0; k < m; k++) {
        out[i + k] = d[k];
        gamma[i + k] = g[k];
      }
    }
  }
}

static unsigned int stream_grid(long long work, int per_thread) {
  long long blocks = (work + 256LL * per_thread - 1) / (256LL * per_thread);
  return (unsigned int)std::max<long long>(1, std::min<long long>(blocks, 8192));
}

// All pointers must be 16-byte aligned (every buffer the library allocates
// for a frame starts on a 16-pixel boundary).
hipError_t launch_adu_to_photons(const uint16_t *adu, const float *offset, const float *gain,
                                 float *out, long long n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(adu_kernel, dim3(stream_grid(n, 8)), dim3(256), 0, s, adu, offset, gain,
                     out, n);
  return hipGetLastError();
}

// adu, offset, out and gamma 16-byte aligned; gains/gammas hold the frame's
// level maps at offsets 0, stride, 2 stride, ... (stride >= n); t[] the
// division points as integers (ADU >= t), unused ones 65536.
hipError_t launch_adu_levels(const uint16_t *adu, const float *offset, const float *gains,
                             const float *gammas, long long stride, const int *t, float *out,
                             float *gamma, long long n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(adu_levels_kernel, dim3(stream_grid(n, 4)), dim3(256), 0, s, adu, offset,
                     gains, gammas, stride, t[0], t[1], t[2], out, gamma, n);
  return hipGetLastError();
}

hipError_t launch_gamma_from_camera(const float *var, const float *gain, float *gamma,
                                    long long n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(gamma_kernel, dim3(stream_grid(n, 1)), dim3(256), 0, s, var, gain, gamma,
                     n);
  return hipGetLastError();
}

}  // namespace ncs
