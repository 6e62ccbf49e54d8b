// Shared definitions for the simkit_amd HIP library (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/simkit_amd.h"

namespace sk {

// Last error message of the calling host thread (returned by sk_last_error).
void set_error(const std::string& msg);

#define SK_HIP_CHECK(expr)                                                            \
  do {                                                                                \
    hipError_t _e = (expr);                                                           \
    if (_e != hipSuccess) {                                                           \
      ::sk::set_error(std::string(#expr) + ": " + hipGetErrorString(_e));             \
      return SK_ERR_HIP;                                                              \
    }                                                                                 \
  } while (0)

#define SK_REQUIRE(cond, msg)                                                         \
  do {                                                                                \
    if (!(cond)) {                                                                    \
      ::sk::set_error(msg);                                                           \
      return SK_ERR_INVALID;                                                          \
    }                                                                                 \
  } while (0)

#define SK_TRY(expr)                                                                  \
  do {                                                                                \
    int _s = (expr);                                                                  \
    if (_s != SK_OK) return _s;                                                       \
  } while (0)

constexpr int kWave = 64;
constexpr int kBlock = 256;

inline int grid_for(long long n, int block = kBlock) {
  long long g = (n + block - 1) / block;
  return static_cast<int>(g < 1 ? 1 : g);
}

// ---------------------------------------------------------------------------
// Deterministic reductions. A block reduces in a fixed tree (wave shuffles,
// then LDS across the 4 waves of a 256-thread block); a grid reduction writes
// one partial per block and the last block to finish sums the partials in a
// fixed order, so results are bit-reproducible for a fixed launch shape.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_down(v, off, kWave);
  return v;
}

template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double* lds /* NV * kBlock/kWave */) {
  const int lane = threadIdx.x % kWave;
  const int wid = threadIdx.x / kWave;
  constexpr int nw = kBlock / kWave;
#pragma unroll
  for (int k = 0; k < NV; ++k) v[k] = wave_sum(v[k]);
  __syncthreads();
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) lds[k * nw + wid] = v[k];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      double s = 0.0;
#pragma unroll
      for (int w = 0; w < nw; ++w) s += lds[k * nw + w];
      v[k] = s;
    }
  }
}

// Grid-wide sum of NV values. Every block must call it (all threads). Returns
// true in the last block to finish; there, thread 0 holds the totals in out[].
// partials: NV * gridDim.x doubles; counter: one zero-initialised unsigned
// that the last block resets to zero.
template <int NV>
__device__ bool grid_sum(double (&v)[NV], double* partials, unsigned* counter, double (&out)[NV]) {
  __shared__ double lds[NV * kBlock / kWave];
  __shared__ bool is_last;
  block_sum<NV>(v, lds);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) partials[k * gridDim.x + blockIdx.x] = v[k];
    __threadfence();
    unsigned ticket = atomicAdd(counter, 1u);
    is_last = (ticket == gridDim.x - 1);
  }
  __syncthreads();
  if (!is_last) return false;
  __threadfence();
  double acc[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    double s = 0.0;
    for (int b = threadIdx.x; b < (int)gridDim.x; b += blockDim.x)
      s += partials[k * gridDim.x + b];
    acc[k] = s;
  }
  block_sum<NV>(acc, lds);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) out[k] = acc[k];
    *counter = 0u;
  }
  return true;
}

}  // namespace sk
