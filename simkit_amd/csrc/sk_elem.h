// Per-tet helpers on the mesh's SoA element layout, shared by the mesh-wide
// element passes and the batched per-problem kernels.
#pragma once

#include "sk_internal.h"

namespace sk {

// ---------------------------------------------------------------------------
// Element helpers (SoA mesh layout)
// ---------------------------------------------------------------------------
struct ElemView {
  int64_t nt;
  const int32_t* __restrict__ T;  // [4][nt]
  const double* __restrict__ D;   // [12][nt]
};

__device__ __forceinline__ void load_elem(const ElemView& m, int64_t e, int32_t (&v)[4], double (&D)[12]) {
#pragma unroll
  for (int k = 0; k < 4; ++k) v[k] = m.T[k * m.nt + e];
#pragma unroll
  for (int k = 0; k < 12; ++k) D[k] = m.D[k * m.nt + e];
}

__device__ __forceinline__ void gather_x(const double* __restrict__ x, const int32_t (&v)[4], double (&xe)[12]) {
#pragma unroll
  for (int k = 0; k < 4; ++k)
#pragma unroll
    for (int i = 0; i < 3; ++i) xe[3 * k + i] = x[3 * (int64_t)v[k] + i];
}

// F[3i+c] = sum_v xe[3v+i] D[4c+v]
__device__ __forceinline__ void make_F(const double (&xe)[12], const double (&D)[12], double (&F)[9]) {
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int c = 0; c < 3; ++c)
      F[3 * i + c] = xe[i] * D[4 * c] + xe[3 + i] * D[4 * c + 1] + xe[6 + i] * D[4 * c + 2] + xe[9 + i] * D[4 * c + 3];
}

__device__ __forceinline__ double det3(const double (&F)[9]) {
  return F[0] * (F[4] * F[8] - F[5] * F[7]) - F[1] * (F[3] * F[8] - F[5] * F[6]) + F[2] * (F[3] * F[7] - F[4] * F[6]);
}

__device__ __forceinline__ void add_F0(const double* __restrict__ F0, int64_t e, double (&F)[9]) {
  if (F0) {
#pragma unroll
    for (int k = 0; k < 9; ++k) F[k] += F0[9 * e + k];
  }
}

inline ElemView view(const sk_mesh* m) { return ElemView{m->nt, m->T.p, m->D.p}; }

}  // namespace sk
