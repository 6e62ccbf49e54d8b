// Two-level PCG preconditioner: block-Jacobi plus an additive coarse
// correction over piecewise-rigid aggregates,
//   M^-1 r = B^-1 r + W E^-1 W^T r,   E = W^T A0 W,
// where the vertices are grouped into spatial aggregates, W holds each
// aggregate's six rigid motions (three translations, three infinitesimal
// rotations about the aggregate centroid) restricted to the free DOFs and
// orthonormalised per aggregate, and A0 is the first system the simulator
// solves. Any symmetric positive definite E keeps M^-1 symmetric positive
// definite, so E is formed and inverted once and reused by every later solve.
//
// Layout (n_agg aggregates, nc coarse unknowns):
//   agg_ptr[n_agg + 1], agg_vert[nv]: the vertices of aggregate I are
//     agg_vert[agg_ptr[I] .. agg_ptr[I + 1]), ascending (a view's agg_vert
//     is null when that is the identity: vertex t at position t);
//   vagg[nv]: aggregate of each vertex;
//   dv[3 * v]: offset of vertex v from its aggregate's centroid;
//   dva[3 * t]: the same offset for vertex agg_vert[t] (aggregate order);
//   tm[36 I + 6 a + j]: column j of aggregate I's orthonormal basis in the
//     raw rigid-motion coordinates a (0..2 translation, 3..5 rotation);
//   agg_col[n_agg + 1]: the coarse unknowns of aggregate I are
//     agg_col[I] .. agg_col[I + 1) (at most 6; fewer where pinned DOFs or a
//     degenerate vertex set leave fewer independent motions);
//   dsc[ld]: D = diag(E^-1)^(1/2) (zero past nc);
//   einv[nc * ld]: S = D^-1 E^-1 D^-1 (unit diagonal, |S_ij| <= 1), symmetrised
//     and rounded to FP16, rows padded with zeros to ld = nc rounded up to a
//     multiple of 8; shift = max_i sum_j |S16_ij - S_ij|. The coarse solve
//     applies D (S16 + shift I) D: by Gershgorin the rounding moves no
//     eigenvalue of S by more than the shift, so the applied operator is
//     symmetric positive definite whenever E is, and M^-1 stays SPD;
//   c[ld]: the scaled restricted residual D W^T r of the current iteration
//     (zero past nc).
// A vertex's raw rigid motion (u, w) displaces it by u + w x d.
#pragma once

#include "sk_internal.h"

namespace sk {

constexpr int kMaxCoarse = 4096;  // coarse unknowns at most (the FP64 E^-1 of the build: 128 MiB)

template <typename T>
struct CoarseView {
  // vertex at aggregate-order position t
  __device__ __forceinline__ int64_t vert(int64_t t) const { return agg_vert ? (int64_t)agg_vert[t] : t; }
  int n_agg;
  int nc;
  int ld;
  const int32_t* agg_ptr;
  const int32_t* agg_vert;
  const int32_t* agg_col;
  const T* dv;
  const T* dva;
  const double* tm;
  const _Float16* einv;
  const double* dsc;
  double shift;
  double* c;
};

struct CoarseSpace {
  int n_agg = 0, nc = 0, ld = 0;
  bool built = false;
  bool ident = false;  // agg_vert[t] == t (aggregates are contiguous vertex ranges): views pass no list
  double build_ms = 0.0;
  DevBuf<int32_t> agg_ptr, agg_vert, agg_col, vagg, adj_ptr, adj;
  DevBuf<double> tm, dv64, dva64, E, pivot, c;
  DevBuf<double> gjP, gjC, gjR;  // blocked Gauss-Jordan: panel inverse, panel columns, new panel rows
  DevBuf<float> dv32, dva32;
  DevBuf<_Float16> einv;
  DevBuf<double> dsc;
  double shift = 0.0;
  template <typename T>
  CoarseView<T> view() const;
};

template <>
inline CoarseView<double> CoarseSpace::view<double>() const {
  return {n_agg, nc, ld, agg_ptr.p, ident ? nullptr : agg_vert.p, agg_col.p, dv64.p, dva64.p, tm.p, einv.p, dsc.p, shift, c.p};
}
template <>
inline CoarseView<float> CoarseSpace::view<float>() const {
  return {n_agg, nc, ld, agg_ptr.p, ident ? nullptr : agg_vert.p, agg_col.p, dv32.p, dva32.p, tm.p, einv.p, dsc.p, shift, c.p};
}

// Orthonormal basis of an aggregate's rigid motions restricted to its free
// DOFs (host): the Gram matrix of the 6 raw motions (offsets dv[3 v], v in
// verts[0..nverts) or vbase + 0..nverts when verts is null; fixed: per-DOF
// 0/1 or null) is eigen-decomposed, eigenvalues above 1e-9 of the largest
// are kept. tm[6 a + j]: column j in raw coordinates a; returns the count.
int rigid_basis(const double* dv, const uint8_t* fixed, const int32_t* verts, int nverts, int64_t vbase,
                double tm[36]);

// Spatial bins of the rest positions (host): vagg[v], and the vertices of
// aggregate I at agg_vert[agg_ptr[I] .. agg_ptr[I + 1]), ascending.
int coarse_bins(const double* X, int64_t nv, int per_agg, std::vector<int32_t>& vagg,
                std::vector<int32_t>& agg_ptr, std::vector<int32_t>& agg_vert);

// Aggregates from the rest positions X (host, nv x 3): each axis of the
// bounding box is cut evenly into bins of close to the edge s of a cube that
// holds per_agg vertices at the mesh's mean density; non-empty bins become
// aggregates. fixed_host: per-DOF 0/1 (may be null).
int coarse_setup(CoarseSpace& C, sk_mesh* m, const double* X, const uint8_t* fixed_host, int per_agg);

// E = W^T A W from the assembled SELL values (FP32 or FP64), then E^-1 in
// FP64 by Gauss-Jordan elimination (E is symmetric positive definite: no
// pivoting), symmetrised and stored in FP32. Synchronises the stream.
int coarse_build(CoarseSpace& C, sk_mesh* m, const void* vals, bool fp32, const uint8_t* fixed, hipStream_t st);

}  // namespace sk
