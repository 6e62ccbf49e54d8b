// Shared definitions for the simkit_amd HIP library (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/simkit_amd.h"

namespace sk {

// Newton's relative-gradient test ||g|| <= rtol ||g(x0)|| cannot be met once
// the accepted update no longer changes the positions: when
// max |t dx| <= kStepFloor max |x| over the free DOFs (4 machine epsilons, the
// spacing of doubles near max |x|), further iterations only re-round x, and
// the iterate is converged to working precision. This only binds when x0 is
// already an equilibrium to within that resolution (a damped mesh at rest),
// where ||g(x0)|| itself is tiny. newton.py:69 stops on the step length too.
constexpr double kStepFloor = 4.0 * 2.220446049250313e-16;

// Whether a material reads lam (every density except ARAP).
constexpr bool material_needs_lam(int material) { return material != SK_ARAP; }

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

// Grid reductions without fences or atomics: every block writes its block
// sum to partials[k * gridDim.x + blockIdx.x] with plain stores; a later
// kernel (ordered after this one on the stream) sums the partials in a fixed
// order, either in a dedicated one-block finalize launch or redundantly in
// every block of the consuming kernel. Results are bit-reproducible for a
// fixed launch shape, and no block waits on another block.
template <int NV>
__device__ __forceinline__ void store_block_partials(double (&v)[NV], double* partials) {
  __shared__ double lds[NV * kBlock / kWave];
  block_sum<NV>(v, lds);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) partials[k * gridDim.x + blockIdx.x] = v[k];
  }
}

// Whole-block sum of n partials per value (value k at partials[k * stride + i]);
// every thread of the block receives the totals.
template <int NV>
__device__ __forceinline__ void sum_partials(const double* __restrict__ partials, int n, int stride,
                                             double (&out)[NV]) {
  __shared__ double lds[NV * kBlock / kWave];
  __shared__ double res[NV];
  double acc[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) s += partials[k * stride + i];
    acc[k] = s;
  }
  block_sum<NV>(acc, lds);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) res[k] = acc[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; ++k) out[k] = res[k];
}

// One-block launch: result[k] = sum of the n partials of value k.
template <int NV>
__global__ __launch_bounds__(kBlock) void finalize_sum_kernel(const double* __restrict__ partials, int n,
                                                              double* __restrict__ result) {
  double out[NV];
  sum_partials<NV>(partials, n, n, out);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) result[k] = out[k];
  }
}

}  // namespace sk
