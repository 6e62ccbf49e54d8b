// Batched implicit steps: many independent meshes (problems) of uneven size
// advanced together, one launch per Newton phase for all of them.
//
// The problems are stored as one union mesh whose vertices and tets are
// grouped per problem (contiguous ranges). Mesh-wide phases (element
// gradient, element Hessian + PSD projection, block gather into FP32 values)
// run once over the union. The values are sliced per problem: chunks of 64
// consecutive rows (one wave's rows in the per-problem sweeps) stored
// component-major, the chunk's rows on its lanes in descending block count, so
// that a wave's load of one block component is one contiguous run of the lanes
// whose rows still have blocks; components 0-3 and 4-7 of a block are float4s
// (three loads per block). The phases with per-problem
// control flow are persistent kernels with one workgroup per problem (looping
// over problems when there are more problems than workgroups):
//   * PCG: block-Jacobi CG on the problem's free DOFs in FP32 with FP64
//     residual replacement (the mixed-precision scheme of the single-mesh
//     simulator), all scalars reduced inside the workgroup; optionally with a
//     two-level preconditioner (block-Jacobi + an additive coarse correction
//     over runs of consecutive vertices, rigid motions per run, one coarse
//     inverse per problem);
//   * line search: the Armijo backtracking of backtracking_line_search.py:8-67
//     on the problem's own incremental potential, in FP64 difference form;
//   * gradient norm and Newton stopping test (newton.py:46-69 with the
//     relative-gradient criterion), which clears the problem's active flag.
// Each problem keeps its own Newton iteration count, line-search step and
// PCG state, so the answer of every problem is the one it would get alone.
// One Newton iteration for the whole batch is a fixed kernel sequence that
// reads the active flags on the device; it is captured once in a hipGraph
// and replayed. Workgroup size and rows-per-lane layout of the per-problem
// kernels are chosen by sk_batch_autotune (timed on the batch itself).
#include "sk_coarse.h"
#include "sk_elem.h"
#include "sk_internal.h"
#include "sk_materials.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <numeric>
#include <vector>

namespace sk {
namespace {

struct BatchArgs {
  int P;
  const int32_t* vstart;   // [P+1]
  const int32_t* tstart;   // [P+1]
  const int32_t* pid;      // [nv] problem of each vertex
  // Per-problem sliced matrix: each problem's rows are cut into chunks of
  // kWave consecutive rows starting at its first vertex. Lane i of chunk c
  // holds the chunk's row wlane[c kWave + i] & 0xff (the rows in descending
  // block count, so the lanes still in a slot group are the first ones), the
  // row's fixed-DOF mask in bits 8-10 and its block count from bit 11; its
  // block s sits at slot q = woff[c] + s kWave + i,
  // in the slot_index<kQuadC, kQuadVW> layout (components 0-3 and 4-7 as
  // float4s, 8 alone). Slots past a row's blocks are never read.
  const int32_t* wchunk0;  // [P+1] first chunk of each problem
  const int32_t* woff;     // [nchunk+1] first slot of each chunk (a multiple of kWave)
  const int32_t* wcol;     // [nslot] column vertex of each slot
  const int32_t* wlane;    // [nchunk kWave] (block count << 11) | (fixed mask << 8) | row in the chunk
  const float* vals;
  const float* binv;       // chunk c's lane i: slot kWave c + i of the same layout
  const uint8_t* fixed;
  ElemView em;
  const double* mu;
  const double* lam;
  const double* vol;
  const double* mass;
  const double* fext;
  double inv_h2;
  double* x;
  const double* xt;
  double* g;
  double* dx;
  double* r64;
  float* r32;
  float* p32;
  float* Ap32;
  float* x32;
  // per problem
  int* active;
  int* status;
  int* newton_iters;
  long long* pcg_iters;
  int* ls_trials;
  int* refines;
  double* g0;
  double* gnorm;
  int* stalled;            // 1: the last accepted update was below position resolution
  double* tstep;
  int* n_active;
  // parameters
  double pcg_rtol2, inner_rtol2;
  int pcg_max_iter;
  double ls_alpha, ls_beta, ls_threshold;
  int ls_max_iter;
  double newton_rtol;
  int newton_max_iter;
  int count_pcg;           // 0: autotuning run, counters untouched
  int max_rows;            // largest problem (vertices)
  // two-level preconditioner (used by the KC kernel variants)
  const float* cdv;         // [3 nv] offset of each vertex from its aggregate's centroid
  const float* ctm;         // [36 n_agg] aggregate g's orthonormal basis: T[a][j] at 36 g + 6 a + j
  const int32_t* cagg0;     // [P + 1] first aggregate of each problem
  const int32_t* ccol;      // [n_agg + 1] first coarse unknown of each aggregate (batch numbering)
  const int64_t* einv_off;  // [P] offset of each problem's E^-1 (n_q x n_q, row-major, FP32)
  const float* einv;
  int agg_rows;             // vertices per aggregate (a multiple of 64)
  int max_agg, max_nc;      // per problem
  int vec_floats;           // dynamic-LDS floats before the coarse scratch
};

// Every thread of an NT-thread block receives the NV totals (fixed order).
// The leading barrier also orders the callers' earlier global stores.
template <int NT, int NV>
__device__ __forceinline__ void all_sum(double (&v)[NV], double* lds) {
  constexpr int nw = NT / kWave;
  const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
#pragma unroll
  for (int k = 0; k < NV; ++k) v[k] = wave_sum(v[k]);
  __syncthreads();
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) lds[k * nw + wid] = v[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < nw; ++w) s += lds[k * nw + w];
    v[k] = s;
  }
}

// Every thread receives the NV maxima (non-negative inputs).
template <int NT, int NV>
__device__ __forceinline__ void all_max(double (&v)[NV], double* lds) {
  constexpr int nw = NT / kWave;
  const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
#pragma unroll
  for (int k = 0; k < NV; ++k)
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) v[k] = fmax(v[k], __shfl_down(v[k], off, kWave));
  __syncthreads();
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) lds[k * nw + wid] = v[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    double m = 0.0;
#pragma unroll
    for (int w = 0; w < nw; ++w) m = fmax(m, lds[k * nw + w]);
    v[k] = m;
  }
}

__device__ __forceinline__ bool is_fixed(const uint8_t* fixed, int64_t k) { return fixed && fixed[k]; }

static_assert(kQuadC == kWave, "a slot group of the quad layout is one wave's rows");

// A thread's row in a sweep: problem-relative sweep position rel -> chunk c,
// lane i, vertex v and block count.
struct RowAt {
  int64_t v, c;
  int i, nb;
  int fx;   // fixed-DOF mask of the row (bit q: DOF 3 v + q)
  bool ok;  // rel is within the problem
};
// wl: the problem's lane words (wlane from its first chunk on).
__device__ __forceinline__ RowAt row_at(const int32_t* wl, int64_t v0, int64_t nrow, int64_t wc0, int64_t rel) {
  RowAt r;
  r.ok = rel < nrow;
  r.c = wc0 + rel / kWave;
  r.i = (int)(rel % kWave);
  const int32_t w = r.ok ? wl[rel] : 0;
  r.v = v0 + (rel - r.i) + (w & 0xff);
  r.fx = (w >> 8) & 7;
  r.nb = w >> 11;
  return r;
}

// y = A_row . src for the row of lane r.i of chunk r.c (when r.ok); LANES lanes
// share a row (lane j takes slot groups j, j + LANES, ...) and combine with
// xor shuffles. src is indexed by 3 * column - off; R: the accumulation type.
template <int LANES, typename R, typename V>
__device__ __forceinline__ void row_product(const BatchArgs& a, const RowAt& r, const V* __restrict__ src,
                                            int64_t off, int j, R (&y)[3]) {
  constexpr int step = LANES * kWave, kU = 4;
  const int64_t c = r.c;
  const int i = r.i;
  y[0] = y[1] = y[2] = R(0);
  if (r.ok) {
    const int32_t g1 = a.woff[c] + r.nb * kWave;
    int32_t g = a.woff[c] + j * kWave;
    // kU slot groups at a time: their loads are all issued before the first
    // product (one memory latency per kU groups); products in slot order
    for (; g + (kU - 1) * step < g1; g += kU * step) {
      int64_t w[kU];
      float A[kU][9];
#pragma unroll
      for (int u = 0; u < kU; ++u) {
        w[u] = 3 * (int64_t)a.wcol[g + u * step + i] - off;
        load_slot<float, kQuadC, kQuadVW>(a.vals, g + u * step, i, A[u]);
      }
#pragma unroll
      for (int u = 0; u < kU; ++u) {
        const R s0 = (R)src[w[u]], s1 = (R)src[w[u] + 1], s2 = (R)src[w[u] + 2];
#pragma unroll
        for (int p = 0; p < 3; ++p) y[p] += (R)A[u][3 * p] * s0 + (R)A[u][3 * p + 1] * s1 + (R)A[u][3 * p + 2] * s2;
      }
    }
    for (; g < g1; g += step) {
      const int64_t w = 3 * (int64_t)a.wcol[g + i] - off;
      float A[9];
      load_slot<float, kQuadC, kQuadVW>(a.vals, g, i, A);
      const R s0 = (R)src[w], s1 = (R)src[w + 1], s2 = (R)src[w + 2];
#pragma unroll
      for (int p = 0; p < 3; ++p) y[p] += (R)A[3 * p] * s0 + (R)A[3 * p + 1] * s1 + (R)A[3 * p + 2] * s2;
    }
  }
#pragma unroll
  for (int off = 1; off < LANES; off <<= 1)
#pragma unroll
    for (int p = 0; p < 3; ++p) y[p] += __shfl_xor(y[p], off, kWave);
}

// z = B^-1 r with the diagonal inverse of row i of chunk c.
__device__ __forceinline__ void apply_binv3(const BatchArgs& a, int64_t c, int i, float r0, float r1, float r2,
                                            float (&z)[3]) {
  float B[9];
  load_slot<float, kQuadC, kQuadVW>(a.binv, kWave * c, i, B);
#pragma unroll
  for (int p = 0; p < 3; ++p) z[p] = B[3 * p] * r0 + B[3 * p + 1] * r1 + B[3 * p + 2] * r2;
}

// The workgroup's shared flag read: one thread reads, all threads branch alike.
__device__ __forceinline__ int block_read(const int* flag) {
  __shared__ int s;
  __syncthreads();
  if (threadIdx.x == 0) s = *flag;
  __syncthreads();
  return s;
}

// ---------------------------------------------------------------------------
// Two-level preconditioner of the per-problem PCG:
//   z = B^-1 r + W_q E_q^-1 W_q^T r,
// W_q: the rigid motions (3 translations, 3 rotations about the centroid,
// restricted to the free DOFs and orthonormalised) of each run of agg_rows
// consecutive vertices of problem q; E_q = W_q^T A_q W_q from the first Newton
// system, inverted once. A wave's rows always lie in one aggregate (agg_rows
// is a multiple of the 64 / LANES rows of a wave), so the restriction is a
// wave sum per (aggregate, wave slot) followed by a fixed-order sum of the
// slots: deterministic and atomic-free.
// ---------------------------------------------------------------------------
struct CoarseLds {
  float* raw;    // [max_agg][slots][6] wave partials of (r, d x r)
  double* c;     // [max_nc] restricted residual
  double* y;     // [max_nc] E^-1 c
  double* uw;    // [max_agg][6] rigid motion (u, w) of each aggregate
};

__device__ __host__ inline int coarse_raw_floats(int max_agg, int agg_rows) {
  const int n = max_agg * (agg_rows / 16) * 6;  // slots per aggregate <= agg_rows / 16 (LANES = 4)
  return (n + 1) & ~1;                          // keep the doubles after it 8-byte aligned
}

__device__ inline CoarseLds coarse_lds(float* dyn, const BatchArgs& a) {
  CoarseLds c;
  c.raw = dyn + a.vec_floats;
  double* d = reinterpret_cast<double*>(c.raw + coarse_raw_floats(a.max_agg, a.agg_rows));
  c.c = d;
  c.y = d + a.max_nc;
  c.uw = d + 2 * a.max_nc;
  return c;
}

// Wave partial of the restriction: every lane passes its row's (r, d x r)
// (zeros for lanes without a row); lane 0 stores the wave sum in the slot of
// the wave's aggregate. row0: the wave's first row relative to the problem.
template <int LANES>
__device__ __forceinline__ void restrict_wave(const BatchArgs& a, const CoarseLds& cl, int64_t row0, int nagg,
                                              const float (&v)[6]) {
  constexpr int rpw = kWave / LANES;  // rows per wave
  float s[6];
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    float t = v[q];
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) t += __shfl_xor(t, off, kWave);
    s[q] = t;
  }
  const int l = (int)(row0 / a.agg_rows);
  if (threadIdx.x % kWave == 0 && l < nagg) {
    const int slots = a.agg_rows / rpw;
    const int ws = (int)((row0 % a.agg_rows) / rpw);
#pragma unroll
    for (int q = 0; q < 6; ++q) cl.raw[(l * slots + ws) * 6 + q] = s[q];
  }
}

// d: the vertex's offset from its aggregate's centroid (3 floats).
__device__ __forceinline__ void restrict_terms(const float* d, const float (&r)[3], float (&t)[6]) {
  const float d0 = d[0], d1 = d[1], d2 = d[2];
  t[0] = r[0];
  t[1] = r[1];
  t[2] = r[2];
  t[3] = d1 * r[2] - d2 * r[1];
  t[4] = d2 * r[0] - d0 * r[2];
  t[5] = d0 * r[1] - d1 * r[0];
}

// From the restriction slots (written before the preceding barrier): c = W^T r,
// y = E^-1 c, the aggregates' rigid motions; returns c . y in every thread.
template <int NT, int LANES>
__device__ double coarse_apply(const BatchArgs& a, const CoarseLds& cl, int p, double* lds) {
  constexpr int rpw = kWave / LANES;
  const int slots = a.agg_rows / rpw;
  const int g0 = a.cagg0[p], nagg = a.cagg0[p + 1] - g0;
  const int cb = a.ccol[g0], nc = a.ccol[g0 + nagg] - cb;
  for (int t = threadIdx.x; t < nagg * 6; t += NT) {
    const int l = t / 6, j = t % 6, g = g0 + l;
    const int c0 = a.ccol[g] - cb, k = a.ccol[g + 1] - a.ccol[g];
    if (j < k) {
      double s = 0.0;
#pragma unroll
      for (int q = 0; q < 6; ++q) {
        float rq = 0.f;
        for (int w = 0; w < slots; ++w) rq += cl.raw[(l * slots + w) * 6 + q];
        s += (double)a.ctm[36 * (int64_t)g + 6 * q + j] * (double)rq;
      }
      cl.c[c0 + j] = s;
    }
  }
  __syncthreads();
  // y = E^-1 c. E^-1 is stored symmetric, so y_i = sum_j E[j][i] c_j: the lanes
  // of a wave take 64 consecutive i (coalesced reads along row j), wave w the
  // j = w, w + nw, ... (kU rows' loads in flight at once); the nw partials of
  // each i are then summed in wave order (lds holds nw x 64 doubles).
  const float* E = a.einv + a.einv_off[p];
  const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
  constexpr int nw = NT / kWave, kU = 8;
  for (int i0 = 0; i0 < nc; i0 += kWave) {
    const int i = i0 + lane;
    double s = 0.0;
    if (i < nc) {
      int jr = wid;
      for (; jr + (kU - 1) * nw < nc; jr += kU * nw) {
        float e[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u) e[u] = E[(int64_t)(jr + u * nw) * nc + i];
#pragma unroll
        for (int u = 0; u < kU; ++u) s += (double)e[u] * cl.c[jr + u * nw];
      }
      for (; jr < nc; jr += nw) s += (double)E[(int64_t)jr * nc + i] * cl.c[jr];
    }
    lds[wid * kWave + lane] = s;
    __syncthreads();
    if (wid == 0 && i < nc) {
      double t = 0.0;
      for (int w = 0; w < nw; ++w) t += lds[w * kWave + lane];
      cl.y[i] = t;
    }
    __syncthreads();
  }
  for (int t = threadIdx.x; t < nagg * 6; t += NT) {
    const int l = t / 6, q = t % 6, g = g0 + l;
    const int c0 = a.ccol[g] - cb, k = a.ccol[g + 1] - a.ccol[g];
    double s = 0.0;
    for (int j = 0; j < k; ++j) s += (double)a.ctm[36 * (int64_t)g + 6 * q + j] * cl.y[c0 + j];
    cl.uw[t] = s;
  }
  double cy[1] = {0.0};
  for (int t = threadIdx.x; t < nc; t += NT) cy[0] += cl.c[t] * cl.y[t];
  all_sum<NT, 1>(cy, lds);  // its barriers also publish uw
  return cy[0];
}

// Coarse part of z at problem-relative row `row` with centroid offset d (3
// floats), zero on the DOFs of the fixed mask fx.
__device__ __forceinline__ void coarse_term(const BatchArgs& a, const CoarseLds& cl, const float* d, int64_t row,
                                            int fx, float (&zc)[3]) {
  const double* uw = cl.uw + 6 * (row / a.agg_rows);
  const double d0 = d[0], d1 = d[1], d2 = d[2];
  const double t[3] = {uw[0] + uw[4] * d2 - uw[5] * d1, uw[1] + uw[5] * d0 - uw[3] * d2,
                       uw[2] + uw[3] * d1 - uw[4] * d0};
#pragma unroll
  for (int q = 0; q < 3; ++q) zc[q] = ((fx >> q) & 1) ? 0.f : (float)t[q];
}

// Dynamic-LDS floats of the PCG kernel's LDS-resident per-problem data, before
// the coarse scratch: p, r, A p, x32 and z (3 max_rows floats each), the lane
// words (kWave per chunk) and, with the coarse space, the centroid offsets
// (3 max_rows floats).
__host__ __device__ inline int pcg_lane_words(int max_rows) { return kWave * ((max_rows + kWave - 1) / kWave); }
__host__ __device__ inline int pcg_lds_vec_floats(int max_rows, bool kc) {
  return 15 * max_rows + pcg_lane_words(max_rows) + (kc ? 3 * max_rows : 0);
}

// ---------------------------------------------------------------------------
// Per-problem PCG: H dx = -g on the free DOFs, FP32 CG with FP64 residual
// replacement whenever the FP32 recurrence residual has dropped by
// inner_rtol; stops at ||b - H dx|| <= pcg_rtol ||b|| (FP64).
// Thread rows: row slot i = threadIdx.x / LANES covers vertices v0 + i,
// v0 + i + NT / LANES, ...
// ---------------------------------------------------------------------------
template <int NT, int LANES, bool KLDS, bool KC>
__global__ __launch_bounds__(NT) void batch_pcg_kernel(BatchArgs a) {
  // reductions (4 doubles per wave); with KC also the coarse product's partials (NT doubles)
  __shared__ double lds[KC ? NT : 4 * (NT / kWave)];
  // KLDS: the problem's vectors, lane words and centroid offsets (pcg_lds_vec_floats); KC: the
  // coarse scratch after them (declared as double so that the base is 8-byte aligned for the
  // scratch's FP64 part)
  extern __shared__ double dyn_d[];
  float* dyn = reinterpret_cast<float*>(dyn_d);
  constexpr int RS = NT / LANES;  // rows per sweep
  constexpr int rpw = kWave / LANES;
  const int j = threadIdx.x % LANES, slot = threadIdx.x / LANES;
  const int64_t wrow = (int64_t)(threadIdx.x / kWave) * rpw;  // this wave's first row within a sweep
  CoarseLds cl{};
  if (KC) cl = coarse_lds(dyn, a);
  for (int p = blockIdx.x; p < a.P; p += gridDim.x) {
    int nagg = 0;
    if (KC) {
      nagg = a.cagg0[p + 1] - a.cagg0[p];
      const int nraw = nagg * (a.agg_rows / rpw) * 6;
      for (int t = threadIdx.x; t < nraw; t += NT) cl.raw[t] = 0.f;
    }
    if (!block_read(a.active + p)) continue;
    const int64_t v0 = a.vstart[p], v1 = a.vstart[p + 1], wc0 = a.wchunk0[p];
    const int64_t nrow = v1 - v0;
    const int64_t sweeps = (nrow + RS - 1) / RS;
    // p, r, A p and x32: in LDS (indexed from the problem's first DOF) or in
    // the global vectors
    const int64_t off = KLDS ? 3 * v0 : 0;
    float* P = KLDS ? dyn : a.p32;
    float* Rv = KLDS ? dyn + 3 * a.max_rows : a.r32;
    float* AP = KLDS ? dyn + 6 * a.max_rows : a.Ap32;
    float* X = KLDS ? dyn + 9 * a.max_rows : a.x32;
    // KLDS: z = B^-1 r of the current residual kept from the sweep that forms
    // it, and the problem's lane words and centroid offsets staged once
    float* Z = KLDS ? dyn + 12 * a.max_rows : nullptr;
    const int32_t* wl = a.wlane + wc0 * kWave;
    float* CD = (KLDS && KC) ? dyn + 15 * a.max_rows + pcg_lane_words(a.max_rows) : nullptr;
    if (KLDS) {
      int32_t* wls = reinterpret_cast<int32_t*>(dyn + 15 * a.max_rows);
      for (int64_t t = threadIdx.x; t < kWave * ((nrow + kWave - 1) / kWave); t += NT) wls[t] = wl[t];
      if (KC)
        for (int64_t t = threadIdx.x; t < 3 * nrow; t += NT) CD[t] = a.cdv[3 * v0 + t];
      wl = wls;
      __syncthreads();
    }
    auto cdv_at = [&](int64_t v) -> const float* { return (KLDS && KC) ? CD + 3 * (v - v0) : a.cdv + 3 * v; };
    // z of row v: kept (KLDS) or B^-1 r
    auto z_of = [&](const RowAt& R, int64_t v, float (&z)[3]) {
      if (KLDS) {
#pragma unroll
        for (int q = 0; q < 3; ++q) z[q] = Z[3 * v + q - off];
      } else {
        const float* rv = Rv + 3 * v - off;
        apply_binv3(a, R.c, R.i, rv[0], rv[1], rv[2], z);
      }
    };
    auto keep_z = [&](int64_t v, const float (&z)[3]) {
      if (KLDS) {
#pragma unroll
        for (int q = 0; q < 3; ++q) Z[3 * v + q - off] = z[q];
      }
    };
    // init: r64 = b = -g (free), dx = 0, r32 = b, x32 = 0, z = M^-1 r, p = z
    double acc[2] = {0.0, 0.0};
    for (int64_t s = 0; s < sweeps; ++s) {
      const RowAt R = row_at(wl, v0, nrow, wc0, s * RS + slot);
      const int64_t v = R.v;
      float tr[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
      if (R.ok && j == 0) {
        float r[3], z[3];
#pragma unroll
        for (int q = 0; q < 3; ++q) {
          const int64_t k = 3 * v + q;
          const double b = ((R.fx >> q) & 1) ? 0.0 : -a.g[k];
          a.r64[k] = b;
          a.dx[k] = 0.0;
          r[q] = (float)b;
          Rv[k - off] = r[q];
          X[k - off] = 0.f;
          acc[1] += b * b;
        }
        apply_binv3(a, R.c, R.i, r[0], r[1], r[2], z);
        keep_z(v, z);
#pragma unroll
        for (int q = 0; q < 3; ++q) {
          if (!KC) P[3 * v + q - off] = z[q];
          acc[0] += (double)r[q] * (double)z[q];
        }
        if (KC) restrict_terms(cdv_at(v), r, tr);
      }
      if (KC) restrict_wave<LANES>(a, cl, s * RS + wrow, nagg, tr);
    }
    all_sum<NT, 2>(acc, lds);
    double rz = acc[0];
    const double bb = acc[1];
    if (KC) {
      rz += coarse_apply<NT, LANES>(a, cl, p, lds);
      for (int64_t s = 0; s < sweeps; ++s) {
        const RowAt R = row_at(wl, v0, nrow, wc0, s * RS + slot);
        const int64_t v = R.v;
        if (R.ok && j == 0) {
          float z[3], zc[3];
          z_of(R, v, z);
          coarse_term(a, cl, cdv_at(v), v - v0, R.fx, zc);
#pragma unroll
          for (int q = 0; q < 3; ++q) P[3 * v + q - off] = z[q] + zc[q];
        }
      }
      __syncthreads();
    }
    const double target = a.pcg_rtol2 * bb;
    double seg_rr = bb;
    int it = 0, refines = 0, st = SK_OK;
    bool done = !(bb > 0.0);
    if (!isfinite(bb)) {
      st = SK_ERR_NONFINITE;
      done = true;
    }
    while (!done && it < a.pcg_max_iter) {
      // Ap = A p (fixed rows 0), p.Ap; the restriction slots are cleared for this iteration
      if (KC) {
        const int nraw = nagg * (a.agg_rows / rpw) * 6;
        for (int t = threadIdx.x; t < nraw; t += NT) cl.raw[t] = 0.f;
      }
      double d1[1] = {0.0};
      for (int64_t s = 0; s < sweeps; ++s) {
        const RowAt R = row_at(wl, v0, nrow, wc0, s * RS + slot);
        const int64_t v = R.v;
        float y[3];
        row_product<LANES, float>(a, R, P, off, j, y);
        if (R.ok && j == 0) {
#pragma unroll
          for (int q = 0; q < 3; ++q) {
            const int64_t k = 3 * v + q;
            const float ap = ((R.fx >> q) & 1) ? 0.f : y[q];
            AP[k - off] = ap;
            d1[0] += (double)P[k - off] * (double)ap;
          }
        }
      }
      all_sum<NT, 1>(d1, lds);
      const double pAp = d1[0];
      if (!(pAp > 0.0)) {
        st = isfinite(pAp) ? SK_ERR_PCG_BREAKDOWN : SK_ERR_NONFINITE;
        break;
      }
      const float alpha = (float)(rz / pAp);
      // x += alpha p, r -= alpha A p, r.r; with KC also r.(B^-1 r) and the restriction of r
      double d2[2] = {0.0, 0.0};
      for (int64_t s = 0; s < sweeps; ++s) {
        const RowAt R = row_at(wl, v0, nrow, wc0, s * RS + slot);
        const int64_t v = R.v;
        float tr[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        if (R.ok && j == 0) {
          float r[3];
#pragma unroll
          for (int q = 0; q < 3; ++q) {
            const int64_t k = 3 * v + q;
            X[k - off] += alpha * P[k - off];
            r[q] = Rv[k - off] - alpha * AP[k - off];
            Rv[k - off] = r[q];
            d2[0] += (double)r[q] * (double)r[q];
          }
          if (KC) {
            float z[3];
            apply_binv3(a, R.c, R.i, r[0], r[1], r[2], z);
            keep_z(v, z);
#pragma unroll
            for (int q = 0; q < 3; ++q) d2[1] += (double)r[q] * (double)z[q];
            restrict_terms(cdv_at(v), r, tr);
          }
        }
        if (KC) restrict_wave<LANES>(a, cl, s * RS + wrow, nagg, tr);
      }
      if (KC) {
        all_sum<NT, 2>(d2, lds);
      } else {
        double t1[1] = {d2[0]};
        all_sum<NT, 1>(t1, lds);
        d2[0] = t1[0];
      }
      ++it;
      if (!isfinite(d2[0])) {
        st = SK_ERR_NONFINITE;
        break;
      }
      if (d2[0] <= a.inner_rtol2 * seg_rr || d2[0] <= target) {
        // residual replacement: dx += x32 (FP64), r64 = b - A dx, r32 = r64
        for (int64_t s = 0; s < sweeps; ++s) {
          const int64_t v = v0 + s * RS + slot;
          if (v < v1 && j == 0) {
#pragma unroll
            for (int q = 0; q < 3; ++q) {
              const int64_t k = 3 * v + q;
              a.dx[k] += (double)X[k - off];
              X[k - off] = 0.f;
            }
          }
        }
        __syncthreads();
        double d3[2] = {0.0, 0.0};
        for (int64_t s = 0; s < sweeps; ++s) {
          const RowAt R = row_at(wl, v0, nrow, wc0, s * RS + slot);
          const int64_t v = R.v;
          double y[3];
          float tr[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
          row_product<LANES, double>(a, R, a.dx, 0, j, y);
          if (R.ok && j == 0) {
            float rf[3];
#pragma unroll
            for (int q = 0; q < 3; ++q) {
              const int64_t k = 3 * v + q;
              const double r = ((R.fx >> q) & 1) ? 0.0 : -a.g[k] - y[q];
              a.r64[k] = r;
              rf[q] = (float)r;
              Rv[k - off] = rf[q];
              d3[0] += r * r;
            }
            if (KC) {
              float z[3];
              apply_binv3(a, R.c, R.i, rf[0], rf[1], rf[2], z);
              keep_z(v, z);
#pragma unroll
              for (int q = 0; q < 3; ++q) d3[1] += (double)rf[q] * (double)z[q];
              restrict_terms(cdv_at(v), rf, tr);
            }
          }
          if (KC) restrict_wave<LANES>(a, cl, s * RS + wrow, nagg, tr);
        }
        if (KC) {
          all_sum<NT, 2>(d3, lds);
        } else {
          double t1[1] = {d3[0]};
          all_sum<NT, 1>(t1, lds);
          d3[0] = t1[0];
        }
        ++refines;
        if (d3[0] <= target) break;
        seg_rr = d3[0];
        d2[1] = d3[1];
      }
      if (KC) {
        // z = B^-1 r + W E^-1 W^T r; r.z = r.B^-1 r + c.y; p = z + beta p
        const double rz_new = d2[1] + coarse_apply<NT, LANES>(a, cl, p, lds);
        const float beta = (float)(rz_new / rz);
        rz = rz_new;
        for (int64_t s = 0; s < sweeps; ++s) {
          const RowAt R = row_at(wl, v0, nrow, wc0, s * RS + slot);
          const int64_t v = R.v;
          if (R.ok && j == 0) {
            float z[3], zc[3];
            z_of(R, v, z);
            coarse_term(a, cl, cdv_at(v), v - v0, R.fx, zc);
#pragma unroll
            for (int q = 0; q < 3; ++q) P[3 * v + q - off] = z[q] + zc[q] + beta * P[3 * v + q - off];
          }
        }
        __syncthreads();
        continue;
      }
      // z = B^-1 r, r.z, then p = z + beta p
      double d4[1] = {0.0};
      for (int64_t s = 0; s < sweeps; ++s) {
        const RowAt R = row_at(wl, v0, nrow, wc0, s * RS + slot);
        const int64_t v = R.v;
        if (R.ok && j == 0) {
          float z[3];
          const float* rv = Rv + 3 * v - off;
          apply_binv3(a, R.c, R.i, rv[0], rv[1], rv[2], z);
          keep_z(v, z);
#pragma unroll
          for (int q = 0; q < 3; ++q) d4[0] += (double)rv[q] * (double)z[q];
        }
      }
      all_sum<NT, 1>(d4, lds);
      const float beta = (float)(d4[0] / rz);
      rz = d4[0];
      for (int64_t s = 0; s < sweeps; ++s) {
        const RowAt R = row_at(wl, v0, nrow, wc0, s * RS + slot);
        const int64_t v = R.v;
        if (R.ok && j == 0) {
          float z[3];
          z_of(R, v, z);
#pragma unroll
          for (int q = 0; q < 3; ++q) P[3 * v + q - off] = z[q] + beta * P[3 * v + q - off];
        }
      }
      __syncthreads();
    }
    if (it >= a.pcg_max_iter && st == SK_OK) {
      // iteration limit: the FP64 correction so far is used as the direction
      for (int64_t s = 0; s < sweeps; ++s) {
        const int64_t v = v0 + s * RS + slot;
        if (v < v1 && j == 0)
#pragma unroll
          for (int q = 0; q < 3; ++q) {
            a.dx[3 * v + q] += (double)X[3 * v + q - off];
            X[3 * v + q - off] = 0.f;
          }
      }
    }
    __syncthreads();
    if (threadIdx.x == 0 && a.count_pcg) {
      a.pcg_iters[p] += it;
      a.refines[p] += refines;
      if (st != SK_OK) {
        a.status[p] = st;
        a.active[p] = 0;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Per-problem Armijo line search in difference form:
//   dE(t) = sum_e vol (psi(x + t dx) - psi(x)) + sum_free [c/(2h^2) m ((x1-xt)^2 - (x0-xt)^2) - f (x1 - x0)]
// accepted at the first t in {1, beta, beta^2, ...} with dE <= alpha t g.dx + threshold.
// ---------------------------------------------------------------------------
template <class M, int NT>
__global__ __launch_bounds__(NT) void batch_ls_kernel(BatchArgs a) {
  __shared__ double lds[2 * (NT / kWave)];  // all_max<NT, 2> uses 2 * NT / 64
  for (int p = blockIdx.x; p < a.P; p += gridDim.x) {
    if (!block_read(a.active + p)) continue;
    const int64_t v0 = a.vstart[p], v1 = a.vstart[p + 1];
    const int64_t t0 = a.tstart[p], t1 = a.tstart[p + 1];
    double sl[1] = {0.0};
    for (int64_t k = 3 * v0 + threadIdx.x; k < 3 * v1; k += NT)
      if (!is_fixed(a.fixed, k)) sl[0] += a.g[k] * a.dx[k];
    all_sum<NT, 1>(sl, lds);
    const double slope = sl[0];
    double t = 1.0;
    int trials = 0;
    bool accepted = false;
    for (int ls = 0; ls < a.ls_max_iter; ++ls) {
      double de[1] = {0.0};
      for (int64_t e = t0 + threadIdx.x; e < t1; e += NT) {
        int32_t v[4];
        double D[12], xe[12], F[9], F1[9], de12[12];
        load_elem(a.em, e, v, D);
        gather_x(a.x, v, xe);
        gather_x(a.dx, v, de12);
        make_F(xe, D, F);
#pragma unroll
        for (int k = 0; k < 12; ++k) de12[k] = xe[k] + t * de12[k];
        make_F(de12, D, F1);
        const double mu = a.mu[e], lam = a.lam ? a.lam[e] : 0.0;
        de[0] += a.vol[e] * M::energy(F1, mu, lam) - a.vol[e] * M::energy(F, mu, lam);
      }
      for (int64_t k = 3 * v0 + threadIdx.x; k < 3 * v1; k += NT) {
        if (is_fixed(a.fixed, k)) continue;
        const double x0 = a.x[k];
        const double x1 = x0 + t * a.dx[k];
        const double d0 = x0 - a.xt[k], d1 = x1 - a.xt[k];
        de[0] += 0.5 * a.inv_h2 * a.mass[k] * (d1 - d0) * (d1 + d0);
        if (a.fext) de[0] -= a.fext[k] * (x1 - x0);
      }
      all_sum<NT, 1>(de, lds);
      ++trials;
      if (de[0] <= a.ls_alpha * t * slope + a.ls_threshold) {
        accepted = true;
        break;
      }
      t *= a.ls_beta;
    }
    // does the accepted update still move the positions?
    double mx[2] = {0.0, 0.0};
    if (accepted)
      for (int64_t k = 3 * v0 + threadIdx.x; k < 3 * v1; k += NT)
        if (!is_fixed(a.fixed, k)) {
          const double step = t * a.dx[k];
          mx[0] = fmax(mx[0], fabs(step));
          mx[1] = fmax(mx[1], fabs(a.x[k] + step));
        }
    all_max<NT, 2>(mx, lds);
    __syncthreads();
    if (threadIdx.x == 0) {
      a.stalled[p] = (accepted && mx[0] <= kStepFloor * mx[1]) ? 1 : 0;
      a.ls_trials[p] += trials;
      a.newton_iters[p] += 1;
      a.tstep[p] = accepted ? t : 0.0;
      if (!accepted) {
        a.status[p] = SK_ERR_LS_FAIL;
        a.active[p] = 0;
      }
    }
  }
}

// x += t_p dx on the free DOFs of problems still active.
__global__ void batch_update_kernel(int64_t n3, BatchArgs a) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n3) return;
  const int p = a.pid[k / 3];
  if (a.active[p] && !is_fixed(a.fixed, k)) a.x[k] += a.tstep[p] * a.dx[k];
}

// Per-problem ||g_free||. mode 0: start of a step (g0, reset counters and
// flags); mode 1: Newton stopping test of an active problem.
template <int NT>
__global__ __launch_bounds__(NT) void batch_gnorm_kernel(BatchArgs a, int mode) {
  __shared__ double lds[NT / kWave];
  for (int p = blockIdx.x; p < a.P; p += gridDim.x) {
    if (mode == 1 && !block_read(a.active + p)) continue;
    const int64_t v0 = a.vstart[p], v1 = a.vstart[p + 1];
    double s[1] = {0.0};
    for (int64_t k = 3 * v0 + threadIdx.x; k < 3 * v1; k += NT)
      if (!is_fixed(a.fixed, k)) s[0] += a.g[k] * a.g[k];
    all_sum<NT, 1>(s, lds);
    __syncthreads();
    if (threadIdx.x == 0) {
      const double gn = sqrt(s[0]);
      a.gnorm[p] = gn;
      if (mode == 0) {
        a.g0[p] = gn;
        a.newton_iters[p] = 0;
        a.pcg_iters[p] = 0;
        a.ls_trials[p] = 0;
        a.refines[p] = 0;
        a.status[p] = isfinite(gn) ? SK_OK : SK_ERR_NONFINITE;
        a.stalled[p] = 0;
        a.active[p] = (isfinite(gn) && gn > 0.0) ? 1 : 0;
      } else if (!isfinite(gn)) {
        a.status[p] = SK_ERR_NONFINITE;
        a.active[p] = 0;
      } else if (gn <= a.newton_rtol * a.g0[p] || a.stalled[p]) {
        a.active[p] = 0;
      } else if (a.newton_iters[p] >= a.newton_max_iter) {
        a.status[p] = SK_ERR_NEWTON_MAXIT;
        a.active[p] = 0;
      }
    }
  }
}

__global__ __launch_bounds__(kBlock) void count_active_kernel(const int* active, int P, int* n_active) {
  __shared__ double lds[kBlock / kWave];
  double s[1] = {0.0};
  for (int p = threadIdx.x; p < P; p += kBlock) s[0] += active[p];
  all_sum<kBlock, 1>(s, lds);
  if (threadIdx.x == 0) *n_active = (int)s[0];
}

__global__ void batch_init_kernel(int64_t n, const double* xc, const double* xp, const double* xbc,
                                  const uint8_t* fixed, double h, double* x, double* xt) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const double c = xc[k];
  const double t = c + h * ((c - xp[k]) / h);
  xt[k] = t;
  x[k] = (fixed && fixed[k]) ? (xbc ? xbc[k] : c) : t;
}

hipStream_t as_stream(sk_stream_t s) { return reinterpret_cast<hipStream_t>(s); }

}  // namespace
}  // namespace sk

using namespace sk;

struct sk_batch {
  sk_mesh* mesh = nullptr;
  int P = 0;
  sk_batch_params prm{};
  DevBuf<int32_t> vstart, tstart, pid;
  DevBuf<double> mu, lam, vol, mass, fext;
  DevBuf<uint8_t> fixed;
  DevBuf<double> x, xt, g, dx, r64;
  DevBuf<float> vals32, binv32, r32, p32, Ap32, x32;
  DevBuf<int32_t> wchunk0, woff, wcol, wlane, wpos, wrow;  // the sliced per-problem matrix (BatchArgs)
  DevBuf<int> active, status, newton_iters, ls_trials, refines, stalled, n_active;
  DevBuf<long long> pcg_iters;
  DevBuf<double> g0, gnorm, tstep;
  int* n_active_host = nullptr;
  hipStream_t st = nullptr;
  hipEvent_t ev_in = nullptr, ev_out = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int graph_block = 0, graph_lanes = 0, graph_lds = 0;
  int grid = 0;
  int max_rows = 0;
  // two-level preconditioner (agg_rows > 0 when enabled)
  int agg_rows = 0, n_agg = 0, max_agg = 0, max_nc = 0;
  bool coarse_built = false;
  int coarse_period = 0, steps_since_coarse = 0;  // re-form every coarse_period steps (0: once)
  DevBuf<float> cdv, ctm, einv;
  DevBuf<int32_t> cagg0, ccol, cagg_prob, cadj_ptr, cadj, cnq;
  DevBuf<int64_t> einv_off;
  DevBuf<double> cE;
  DevBuf<int> cbad;
  bool coarse() const { return agg_rows > 0; }
  BatchArgs args() const;
};

BatchArgs sk_batch::args() const {
  BatchArgs a{};
  a.P = P;
  a.vstart = vstart.p;
  a.tstart = tstart.p;
  a.pid = pid.p;
  a.wchunk0 = wchunk0.p;
  a.woff = woff.p;
  a.wcol = wcol.p;
  a.wlane = wlane.p;
  a.vals = vals32.p;
  a.binv = binv32.p;
  a.fixed = fixed.p;
  a.em = view(mesh);
  a.mu = mu.p;
  a.lam = lam.p;
  a.vol = vol.p;
  a.mass = mass.p;
  a.fext = fext.p;
  a.inv_h2 = 1.0 / (prm.h * prm.h);
  a.x = x.p;
  a.xt = xt.p;
  a.g = g.p;
  a.dx = dx.p;
  a.r64 = r64.p;
  a.r32 = r32.p;
  a.p32 = p32.p;
  a.Ap32 = Ap32.p;
  a.x32 = x32.p;
  a.active = active.p;
  a.status = status.p;
  a.newton_iters = newton_iters.p;
  a.pcg_iters = pcg_iters.p;
  a.ls_trials = ls_trials.p;
  a.refines = refines.p;
  a.g0 = g0.p;
  a.gnorm = gnorm.p;
  a.stalled = stalled.p;
  a.tstep = tstep.p;
  a.n_active = n_active.p;
  a.pcg_rtol2 = prm.pcg_rtol * prm.pcg_rtol;
  a.inner_rtol2 = prm.inner_rtol * prm.inner_rtol;
  a.pcg_max_iter = prm.pcg_max_iter;
  a.ls_alpha = prm.ls_alpha;
  a.ls_beta = prm.ls_beta;
  a.ls_threshold = prm.ls_threshold;
  a.ls_max_iter = prm.ls_max_iter;
  a.newton_rtol = prm.newton_rtol;
  a.newton_max_iter = prm.newton_max_iter;
  a.count_pcg = 1;
  a.max_rows = max_rows;
  a.cdv = cdv.p;
  a.ctm = ctm.p;
  a.cagg0 = cagg0.p;
  a.ccol = ccol.p;
  a.einv_off = einv_off.p;
  a.einv = einv.p;
  a.agg_rows = agg_rows;
  a.max_agg = max_agg;
  a.max_nc = max_nc;
  return a;
}

namespace {

// Dynamic LDS of the PCG kernel: the LDS-resident per-problem data
// (pcg_lds_vec_floats) followed by the coarse scratch.
int pcg_vec_floats(const sk_batch* B, int lds) { return lds ? pcg_lds_vec_floats(B->max_rows, B->coarse()) : 0; }
size_t pcg_shm_bytes(const sk_batch* B, int lds) {
  size_t b = (size_t)pcg_vec_floats(B, lds) * sizeof(float);
  if (B->coarse())
    b += (size_t)coarse_raw_floats(B->max_agg, B->agg_rows) * sizeof(float) +
         (2 * (size_t)B->max_nc + 6 * (size_t)B->max_agg) * sizeof(double);
  return b;
}
size_t pcg_lds_bytes(const sk_batch* B) { return pcg_shm_bytes(B, 1); }
// 150 KiB keeps the total, with the kernel's static scratch (up to 8.5 KiB at
// 1024 threads), within the 160 KiB a workgroup may use.
bool pcg_lds_fits(const sk_batch* B) { return B->max_rows > 0 && pcg_lds_bytes(B) <= 150 * 1024; }

int launch_pcg(const sk_batch* B, const BatchArgs& a, int block, int lanes, int lds) {
  const int g = B->grid;
  const size_t shm = pcg_shm_bytes(B, lds);
  SK_REQUIRE(!lds || pcg_lds_fits(B), "sk_batch: problems too large for the LDS-resident PCG");
  SK_REQUIRE(!B->coarse() || B->coarse_built, "sk_batch: coarse operator not built");
  BatchArgs a2 = a;
  a2.vec_floats = pcg_vec_floats(B, lds);
  const bool kc = B->coarse();
#define SK_PCG_CASE(NT, L)                                                         \
  if (block == NT && lanes == L) {                                                 \
    if (kc) {                                                                      \
      if (lds) batch_pcg_kernel<NT, L, true, true><<<g, NT, shm, B->st>>>(a2);     \
      else batch_pcg_kernel<NT, L, false, true><<<g, NT, shm, B->st>>>(a2);        \
    } else {                                                                       \
      if (lds) batch_pcg_kernel<NT, L, true, false><<<g, NT, shm, B->st>>>(a2);    \
      else batch_pcg_kernel<NT, L, false, false><<<g, NT, 0, B->st>>>(a2);         \
    }                                                                              \
    return SK_OK;                                                                  \
  }
  SK_PCG_CASE(128, 1) SK_PCG_CASE(128, 4) SK_PCG_CASE(256, 1) SK_PCG_CASE(256, 4)
  SK_PCG_CASE(512, 1) SK_PCG_CASE(512, 4) SK_PCG_CASE(1024, 1) SK_PCG_CASE(1024, 4)
#undef SK_PCG_CASE
  set_error("sk_batch: unsupported block/lanes");
  return SK_ERR_INVALID;
}

template <int NT>
int launch_ls_nt(const sk_batch* B, const BatchArgs& a) {
  const bool ok = dispatch_material(B->prm.material, [&](auto tag) {
    batch_ls_kernel<typename decltype(tag)::template type<double>, NT><<<B->grid, NT, 0, B->st>>>(a);
  });
  SK_REQUIRE(ok, "sk_batch: unknown material");
  return SK_OK;
}

int launch_ls(const sk_batch* B, const BatchArgs& a, int block) {
  if (block == 128) return launch_ls_nt<128>(B, a);
  if (block == 256) return launch_ls_nt<256>(B, a);
  if (block == 512) return launch_ls_nt<512>(B, a);
  if (block == 1024) return launch_ls_nt<1024>(B, a);
  set_error("sk_batch: unsupported block");
  return SK_ERR_INVALID;
}

// ---------------------------------------------------------------------------
// Set-up of the per-problem coarse operators (once, from an assembled system).
// ---------------------------------------------------------------------------
struct GalerkinBatchArgs {
  const int32_t* vstart;
  const int32_t* bptr;
  const int32_t* bcol;
  const int32_t* wpos;     // slot of each block in the sliced FP32 values
  const float* vals;
  const uint8_t* fixed;
  const float* cdv;
  const float* ctm;
  const int32_t* cagg0;
  const int32_t* ccol;
  const int32_t* cagg_prob;
  const int32_t* cadj_ptr;
  const int32_t* cadj;
  const int32_t* cnq;      // [P] coarse unknowns per problem
  const int64_t* einv_off;
  double* E;               // per problem, n_q x n_q, at einv_off[q]
  int agg_rows;
};

// Row p of R_v = [I | e_0 x d, e_1 x d, e_2 x d] (3 x 6).
__device__ __forceinline__ void rigid_row_f(const float* d, int p, double (&out)[6]) {
  const double dx = d[0], dy = d[1], dz = d[2];
  out[0] = p == 0;
  out[1] = p == 1;
  out[2] = p == 2;
  const double c0[3] = {0.0, -dz, dy}, c1[3] = {dz, 0.0, -dx}, c2[3] = {-dy, dx, 0.0};
  out[3] = c0[p];
  out[4] = c1[p];
  out[5] = c2[p];
}

// Workgroup g (aggregate): E_q blocks (g, J) = T_g^T (sum over blocks (v, w),
// v in g, w in J, of R_v^T M_v A_vw M_w R_w) T_J for every adjacent J, with
// fixed-order block sums.
__global__ __launch_bounds__(kBlock) void batch_galerkin_kernel(GalerkinBatchArgs a) {
  __shared__ double lds[36 * kBlock / kWave];
  const int g = blockIdx.x, q = a.cagg_prob[g];
  const int64_t v0 = a.vstart[q], v1 = a.vstart[q + 1];
  const int g0 = a.cagg0[q], l = g - g0;
  const int64_t r0 = v0 + (int64_t)l * a.agg_rows, r1 = std::min<int64_t>(v1, r0 + a.agg_rows);
  const int cb = a.ccol[g0], nq = a.cnq[q];
  const int cI = a.ccol[g] - cb, kI = a.ccol[g + 1] - a.ccol[g];
  for (int jj = a.cadj_ptr[g]; jj < a.cadj_ptr[g + 1]; ++jj) {
    const int J = a.cadj[jj];
    const int cJ = a.ccol[J] - cb, kJ = a.ccol[J + 1] - a.ccol[J];
    double acc[36];
#pragma unroll
    for (int k = 0; k < 36; ++k) acc[k] = 0.0;
    for (int64_t v = r0 + threadIdx.x; v < r1; v += blockDim.x) {
      for (int32_t b = a.bptr[v]; b < a.bptr[v + 1]; ++b) {
        const int64_t w = a.bcol[b];
        if (g0 + (int)((w - v0) / a.agg_rows) != J) continue;
        double A[9];
        const int64_t q = a.wpos[b];
#pragma unroll
        for (int m = 0; m < 9; ++m) {
          const bool fr = !(a.fixed && (a.fixed[3 * v + m / 3] || a.fixed[3 * w + m % 3]));
          A[m] = fr ? (double)a.vals[slot_index<kQuadC, kQuadVW>(q, m)] : 0.0;
        }
        double Rw[3][6], AR[3][6];
#pragma unroll
        for (int p = 0; p < 3; ++p) rigid_row_f(a.cdv + 3 * w, p, Rw[p]);
#pragma unroll
        for (int p = 0; p < 3; ++p)
#pragma unroll
          for (int c = 0; c < 6; ++c) AR[p][c] = A[3 * p] * Rw[0][c] + A[3 * p + 1] * Rw[1][c] + A[3 * p + 2] * Rw[2][c];
#pragma unroll
        for (int p = 0; p < 3; ++p) {
          double Rv[6];
          rigid_row_f(a.cdv + 3 * v, p, Rv);
#pragma unroll
          for (int r = 0; r < 6; ++r)
#pragma unroll
            for (int c = 0; c < 6; ++c) acc[6 * r + c] += Rv[r] * AR[p][c];
        }
      }
    }
    block_sum<36>(acc, lds);
    if (threadIdx.x == 0) {
      const float* TI = a.ctm + 36 * (int64_t)g;
      const float* TJ = a.ctm + 36 * (int64_t)J;
      double* E = a.E + a.einv_off[q];
      for (int j = 0; j < kI; ++j)
        for (int m = 0; m < kJ; ++m) {
          double s = 0.0;
          for (int r = 0; r < 6; ++r)
            for (int c = 0; c < 6; ++c) s += (double)TI[6 * r + j] * acc[6 * r + c] * (double)TJ[6 * c + m];
          E[(int64_t)(cI + j) * nq + cJ + m] = s;
        }
    }
    __syncthreads();
  }
}

// Workgroup q: Gauss-Jordan inversion of E_q in place (symmetric positive
// definite, no pivoting), then E_q^-1 symmetrised into FP32. A non-positive
// or non-finite pivot sets bad[q] and stops that problem's elimination.
constexpr int kGjBlock = 1024;

__global__ __launch_bounds__(kGjBlock) void batch_gj_kernel(double* E, const int64_t* off, const int32_t* nq_arr,
                                                            float* einv, int* bad) {
  __shared__ double s_piv;
  const int q = blockIdx.x, n = nq_arr[q];
  double* A = E + off[q];
  for (int k = 0; k < n; ++k) {
    if (threadIdx.x == 0) s_piv = A[(int64_t)k * n + k];
    __syncthreads();
    const double piv = s_piv;
    if (!(piv > 0.0) || !isfinite(piv)) {
      if (threadIdx.x == 0) bad[q] = 1;
      return;  // uniform: every thread read the same pivot
    }
    for (int64_t idx = threadIdx.x; idx < (int64_t)n * n; idx += blockDim.x) {
      const int i = (int)(idx / n), j = (int)(idx % n);
      if (i != k && j != k) A[idx] -= A[(int64_t)i * n + k] * A[(int64_t)k * n + j] / piv;
    }
    __syncthreads();
    for (int idx = threadIdx.x; idx < 2 * n; idx += blockDim.x) {
      if (idx < n) {
        A[(int64_t)k * n + idx] = (idx == k) ? 1.0 / piv : A[(int64_t)k * n + idx] / piv;
      } else if (idx - n != k) {
        A[(int64_t)(idx - n) * n + k] = -A[(int64_t)(idx - n) * n + k] / piv;
      }
    }
    __syncthreads();
  }
  for (int64_t idx = threadIdx.x; idx < (int64_t)n * n; idx += blockDim.x) {
    const int i = (int)(idx / n), j = (int)(idx % n);
    einv[off[q] + idx] = (float)(0.5 * (A[idx] + A[(int64_t)j * n + i]));
  }
}

int hessian_phase(sk_batch* B, bool only_active);

// Forms and inverts every problem's coarse operator from the system assembled
// at the batch's current positions (synchronises the batch stream).
int build_coarse(sk_batch* B) {
  SK_TRY(hessian_phase(B, false));
  GalerkinBatchArgs g{B->vstart.p, B->mesh->bptr.p, B->mesh->bcol.p, B->wpos.p, B->vals32.p, B->fixed.p,
                      B->cdv.p, B->ctm.p, B->cagg0.p, B->ccol.p, B->cagg_prob.p, B->cadj_ptr.p, B->cadj.p,
                      B->cnq.p, B->einv_off.p, B->cE.p, B->agg_rows};
  SK_HIP_CHECK(hipMemsetAsync(B->cE.p, 0, B->cE.n * sizeof(double), B->st));
  SK_HIP_CHECK(hipMemsetAsync(B->cbad.p, 0, B->P * sizeof(int), B->st));
  batch_galerkin_kernel<<<B->n_agg, kBlock, 0, B->st>>>(g);
  batch_gj_kernel<<<B->P, kGjBlock, 0, B->st>>>(B->cE.p, B->einv_off.p, B->cnq.p, B->einv.p, B->cbad.p);
  SK_HIP_CHECK(hipGetLastError());
  std::vector<int> bad(B->P);
  SK_HIP_CHECK(hipMemcpyAsync(bad.data(), B->cbad.p, B->P * sizeof(int), hipMemcpyDeviceToHost, B->st));
  SK_HIP_CHECK(hipStreamSynchronize(B->st));
  for (int q = 0; q < B->P; ++q)
    if (bad[q]) {
      set_error("sk_batch: coarse operator of problem " + std::to_string(q) + " is not positive definite");
      return SK_ERR_PCG_BREAKDOWN;
    }
  B->coarse_built = true;
  return SK_OK;
}

int gradient_phase(sk_batch* B) {
  sk_mesh* m = B->mesh;
  SK_TRY(launch_elem_grad(m, B->prm.material, B->x.p, nullptr, B->mu.p, B->lam.p, B->vol.p, B->st));
  return launch_gather_grad(m, B->g.p, B->mass.p, B->x.p, B->xt.p, B->fext.p, B->fixed.p,
                            1.0 / (B->prm.h * B->prm.h), B->st);
}

// Element Hessian + FP32 gather into the sliced layout. only_active: skip the elements and
// blocks of problems whose active flag (read on the device) is 0 (their
// values are not read until they are assembled again); else every problem.
int hessian_phase(sk_batch* B, bool only_active) {
  sk_mesh* m = B->mesh;
  const int* act = only_active ? B->active.p : nullptr;
  const int32_t* pid = only_active ? B->pid.p : nullptr;
  SK_TRY(launch_elem_hess(m, B->prm.material, B->x.p, nullptr, B->mu.p, B->lam.p, B->vol.p, B->prm.psd, true,
                          B->st, act, pid, nullptr, true));
  return launch_gather_hess_quad(m, B->vals32.p, B->mass.p, 1.0 / (B->prm.h * B->prm.h), B->fixed.p,
                                 B->binv32.p, B->wpos.p, B->wrow.p, true, B->st, act, pid);
}

// One Newton iteration of every active problem (no host synchronization).
int newton_iteration(sk_batch* B, int block, int lanes, int lds) {
  const BatchArgs a = B->args();
  const int64_t n3 = 3 * B->mesh->nv;
  SK_TRY(hessian_phase(B, true));
  SK_TRY(launch_pcg(B, a, block, lanes, lds));
  SK_TRY(launch_ls(B, a, block));
  batch_update_kernel<<<grid_for(n3), kBlock, 0, B->st>>>(n3, a);
  SK_TRY(gradient_phase(B));
  batch_gnorm_kernel<256><<<B->grid, 256, 0, B->st>>>(a, 1);
  count_active_kernel<<<1, kBlock, 0, B->st>>>(B->active.p, B->P, B->n_active.p);
  SK_HIP_CHECK(hipGetLastError());
  return SK_OK;
}

int start_step(sk_batch* B, const double* x_curr, const double* x_prev, const double* x_bc) {
  const int64_t n3 = 3 * B->mesh->nv;
  batch_init_kernel<<<grid_for(n3), kBlock, 0, B->st>>>(n3, x_curr, x_prev, x_bc, B->fixed.p, B->prm.h, B->x.p,
                                                        B->xt.p);
  SK_TRY(gradient_phase(B));
  const BatchArgs a = B->args();
  batch_gnorm_kernel<256><<<B->grid, 256, 0, B->st>>>(a, 0);
  count_active_kernel<<<1, kBlock, 0, B->st>>>(B->active.p, B->P, B->n_active.p);
  SK_HIP_CHECK(hipGetLastError());
  return SK_OK;
}

int read_active(sk_batch* B, int* n) {
  SK_HIP_CHECK(hipMemcpyAsync(B->n_active_host, B->n_active.p, sizeof(int), hipMemcpyDeviceToHost, B->st));
  SK_HIP_CHECK(hipStreamSynchronize(B->st));
  *n = *B->n_active_host;
  return SK_OK;
}

int ensure_graph(sk_batch* B, int block, int lanes, int lds) {
  if (B->exec && B->graph_block == block && B->graph_lanes == lanes && B->graph_lds == lds) return SK_OK;
  if (B->exec) (void)hipGraphExecDestroy(B->exec);
  if (B->graph) (void)hipGraphDestroy(B->graph);
  B->exec = nullptr;
  B->graph = nullptr;
  SK_HIP_CHECK(hipStreamBeginCapture(B->st, hipStreamCaptureModeThreadLocal));
  const int s = newton_iteration(B, block, lanes, lds);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(B->st, &g);
  if (s != SK_OK) {
    if (g) (void)hipGraphDestroy(g);
    return s;
  }
  SK_HIP_CHECK(e);
  B->graph = g;
  SK_HIP_CHECK(hipGraphInstantiate(&B->exec, g, nullptr, nullptr, 0));
  B->graph_block = block;
  B->graph_lanes = lanes;
  B->graph_lds = lds;
  return SK_OK;
}

int batch_block(const sk_batch* B) { return B->prm.block > 0 ? B->prm.block : 256; }
int batch_lanes(const sk_batch* B) { return B->prm.lanes > 0 ? B->prm.lanes : 1; }
int batch_lds(const sk_batch* B) { return B->prm.lds && pcg_lds_fits(B) ? 1 : 0; }

}  // namespace

extern "C" {

int sk_batch_create(sk_mesh* mesh, int nprob, const int64_t* vstart, const int64_t* tstart, const int32_t* T_host,
                    const sk_batch_params* p, const double* mu, const double* lam, const double* vol,
                    const double* mass, const double* f_ext, const uint8_t* fixed, sk_batch** out) {
  SK_REQUIRE(out, "sk_batch_create: null out");
  *out = nullptr;
  SK_REQUIRE(mesh && p && vstart && tstart && T_host && mu && vol && mass, "sk_batch_create: null argument");
  SK_REQUIRE(nprob >= 1, "sk_batch_create: need at least one problem");
  SK_REQUIRE(p->material >= SK_SNH && p->material <= SK_FCR, "sk_batch_create: unknown material");
  SK_REQUIRE(!material_needs_lam(p->material) || lam, "sk_batch_create: this material needs lam");
  SK_REQUIRE(p->h > 0.0 && p->ls_alpha > 0.0 && p->ls_alpha <= 0.5 && p->ls_beta > 0.0 && p->ls_beta < 1.0,
             "sk_batch_create: invalid h or line-search constants");
  SK_REQUIRE(p->inner_rtol > 0.0 && p->inner_rtol < 1.0, "sk_batch_create: inner_rtol must be in (0, 1)");
  SK_REQUIRE(p->psd == SK_PSD_NONE || p->psd == SK_PSD_PROJ || p->psd == SK_PSD_ABS, "sk_batch_create: bad psd");
  SK_REQUIRE(p->block == 0 || p->block == 128 || p->block == 256 || p->block == 512 || p->block == 1024,
             "sk_batch_create: block must be 0, 128, 256, 512 or 1024");
  SK_REQUIRE(p->lanes == 0 || p->lanes == 1 || p->lanes == 4, "sk_batch_create: lanes must be 0, 1 or 4");
  SK_REQUIRE(p->lds == 0 || p->lds == 1, "sk_batch_create: lds must be 0 or 1");
  SK_REQUIRE(vstart[0] == 0 && vstart[nprob] == mesh->nv && tstart[0] == 0 && tstart[nprob] == mesh->nt,
             "sk_batch_create: problem ranges must cover the mesh");
  std::vector<int32_t> vs(nprob + 1), ts(nprob + 1), pid(mesh->nv);
  for (int q = 0; q <= nprob; ++q) {
    vs[q] = (int32_t)vstart[q];
    ts[q] = (int32_t)tstart[q];
    if (q > 0) SK_REQUIRE(vstart[q] > vstart[q - 1] && tstart[q] >= tstart[q - 1], "sk_batch_create: bad ranges");
  }
  for (int q = 0; q < nprob; ++q) {
    for (int64_t v = vstart[q]; v < vstart[q + 1]; ++v) pid[v] = q;
    for (int64_t e = tstart[q]; e < tstart[q + 1]; ++e)
      for (int c = 0; c < 4; ++c) {
        const int64_t v = T_host[4 * e + c];
        SK_REQUIRE(v >= vstart[q] && v < vstart[q + 1], "sk_batch_create: a tet references another problem");
      }
  }
  // The per-problem sliced matrix: chunks of kWave rows from each problem's
  // first vertex, the rows of a chunk on its lanes in descending block count
  // (ties in row order), slot groups as wide as the chunk's longest row.
  std::vector<int32_t> wchunk0(nprob + 1), woff(1, 0), wcol, wlane, wpos(mesh->nblk), wrow(mesh->nv);
  {
    // The per-problem SpMV reads only the problem's own rows: every block of
    // the mesh's pattern must couple two vertices of one problem.
    std::vector<int32_t> bp(mesh->nv + 1), bc(mesh->nblk);
    SK_HIP_CHECK(hipMemcpy(bp.data(), mesh->bptr.p, bp.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
    if (mesh->nblk) SK_HIP_CHECK(hipMemcpy(bc.data(), mesh->bcol.p, bc.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
    for (int64_t v = 0; v < mesh->nv; ++v)
      for (int32_t b = bp[v]; b < bp[v + 1]; ++b)
        SK_REQUIRE(bc[b] >= 0 && bc[b] < mesh->nv && pid[bc[b]] == pid[v],
                   "sk_batch_create: the mesh couples vertices of different problems");
    std::vector<uint8_t> fx(3 * mesh->nv, 0);  // the fixed-DOF flags, packed into wlane
    if (fixed) SK_HIP_CHECK(hipMemcpy(fx.data(), fixed, fx.size(), hipMemcpyDeviceToHost));
    int64_t nslot = 0;
    for (int q = 0; q < nprob; ++q) {
      wchunk0[q] = (int32_t)(woff.size() - 1);
      for (int64_t r0 = vstart[q]; r0 < vstart[q + 1]; r0 += kWave) {
        const int m = (int)(std::min<int64_t>(vstart[q + 1], r0 + kWave) - r0);
        const int64_t c = (int64_t)woff.size() - 1;
        auto nb = [&](int o) { return bp[r0 + o + 1] - bp[r0 + o]; };
        std::vector<int> lane(m);
        std::iota(lane.begin(), lane.end(), 0);
        std::stable_sort(lane.begin(), lane.end(), [&](int x, int y) { return nb(x) > nb(y); });
        const int32_t width = m ? nb(lane[0]) : 0;
        SK_REQUIRE(width < (1 << 20), "sk_batch_create: row with too many blocks");
        const int64_t base = nslot;
        nslot += (int64_t)width * kWave;
        SK_REQUIRE(nslot <= INT32_MAX, "sk_batch_create: matrix too large for 32-bit slot indices");
        woff.push_back((int32_t)nslot);
        wcol.resize(nslot, (int32_t)r0);  // slots past a row's blocks are never read
        wlane.resize((c + 1) * kWave, 0);
        for (int i = 0; i < m; ++i) {
          const int o = lane[i];
          const int64_t v = r0 + o;
          const int mask = (fx[3 * v] ? 1 : 0) | (fx[3 * v + 1] ? 2 : 0) | (fx[3 * v + 2] ? 4 : 0);
          wlane[c * kWave + i] = (nb(o) << 11) | (mask << 8) | o;
          wrow[v] = (int32_t)(c * kWave + i);
          for (int32_t k = 0; k < nb(o); ++k) {
            const int64_t sl = base + (int64_t)k * kWave + i;
            wcol[sl] = bc[bp[v] + k];
            wpos[bp[v] + k] = (int32_t)sl;
          }
        }
      }
    }
    wchunk0[nprob] = (int32_t)(woff.size() - 1);
    if (wcol.empty()) wcol.push_back(0);
  }
  sk_batch* B = new sk_batch();
  B->mesh = mesh;
  B->P = nprob;
  B->prm = *p;
  B->grid = nprob;
  for (int q = 0; q < nprob; ++q) B->max_rows = std::max<int>(B->max_rows, (int)(vstart[q + 1] - vstart[q]));
  const size_t nt = mesh->nt, n3 = 3 * mesh->nv;
  auto copy = [](DevBuf<double>& d, const double* s, size_t n) -> int {
    SK_TRY(d.alloc(n));
    SK_HIP_CHECK(hipMemcpy(d.p, s, n * sizeof(double), hipMemcpyDeviceToDevice));
    return SK_OK;
  };
  auto run = [&]() -> int {
    SK_TRY(B->vstart.upload(vs));
    SK_TRY(B->tstart.upload(ts));
    SK_TRY(B->pid.upload(pid));
    SK_TRY(copy(B->mu, mu, nt));
    if (lam) SK_TRY(copy(B->lam, lam, nt));
    SK_TRY(copy(B->vol, vol, nt));
    SK_TRY(copy(B->mass, mass, n3));
    if (f_ext) SK_TRY(copy(B->fext, f_ext, n3));
    if (fixed) {
      SK_TRY(B->fixed.alloc(n3));
      SK_HIP_CHECK(hipMemcpy(B->fixed.p, fixed, n3, hipMemcpyDeviceToDevice));
    }
    for (DevBuf<double>* b : {&B->x, &B->xt, &B->g, &B->dx, &B->r64}) SK_TRY(b->alloc(n3));
    for (DevBuf<float>* b : {&B->r32, &B->p32, &B->Ap32, &B->x32}) SK_TRY(b->alloc(n3));
    const size_t nchunk = woff.size() - 1, nslot = (size_t)woff.back();
    SK_TRY(B->vals32.alloc(9 * std::max<size_t>(nslot, 1)));
    SK_HIP_CHECK(hipMemset(B->vals32.p, 0, B->vals32.n * sizeof(float)));  // padding blocks stay zero
    SK_TRY(B->binv32.alloc(9 * (size_t)kWave * std::max<size_t>(nchunk, 1)));
    SK_TRY(B->wchunk0.upload(wchunk0));
    SK_TRY(B->woff.upload(woff));
    SK_TRY(B->wcol.upload(wcol));
    SK_TRY(B->wlane.upload(wlane));
    SK_TRY(B->wpos.upload(wpos));
    SK_TRY(B->wrow.upload(wrow));
    for (DevBuf<int>* b : {&B->active, &B->status, &B->newton_iters, &B->ls_trials, &B->refines, &B->stalled}) {
      SK_TRY(b->alloc(nprob));
      SK_HIP_CHECK(hipMemset(b->p, 0, nprob * sizeof(int)));
    }
    SK_TRY(B->pcg_iters.alloc(nprob));
    SK_HIP_CHECK(hipMemset(B->pcg_iters.p, 0, nprob * sizeof(long long)));
    for (DevBuf<double>* b : {&B->g0, &B->gnorm, &B->tstep}) SK_TRY(b->alloc(nprob));
    SK_TRY(B->n_active.alloc(1));
    SK_HIP_CHECK(hipHostMalloc((void**)&B->n_active_host, sizeof(int), hipHostMallocDefault));
    SK_HIP_CHECK(hipStreamCreate(&B->st));
    SK_HIP_CHECK(hipEventCreateWithFlags(&B->ev_in, hipEventDisableTiming));
    SK_HIP_CHECK(hipEventCreateWithFlags(&B->ev_out, hipEventDisableTiming));
    return SK_OK;
  };
  const int s = run();
  if (s != SK_OK) {
    sk_batch_destroy(B);
    return s;
  }
  *out = B;
  return SK_OK;
}

int sk_batch_destroy(sk_batch* B) {
  if (!B) return SK_OK;
  if (B->st) (void)hipStreamSynchronize(B->st);
  if (B->exec) (void)hipGraphExecDestroy(B->exec);
  if (B->graph) (void)hipGraphDestroy(B->graph);
  if (B->ev_in) (void)hipEventDestroy(B->ev_in);
  if (B->ev_out) (void)hipEventDestroy(B->ev_out);
  if (B->st) (void)hipStreamDestroy(B->st);
  if (B->n_active_host) (void)hipHostFree(B->n_active_host);
  delete B;
  return SK_OK;
}

int sk_batch_step(sk_batch* B, const double* x_curr, const double* x_prev, const double* x_bc, double* x_next,
                  sk_batch_info* info, int* rounds, sk_stream_t s) {
  SK_REQUIRE(B && x_curr && x_prev && x_next, "sk_batch_step: null argument");
  hipStream_t caller = as_stream(s);
  SK_HIP_CHECK(hipEventRecord(B->ev_in, caller));
  SK_HIP_CHECK(hipStreamWaitEvent(B->st, B->ev_in, 0));
  const int block = batch_block(B), lanes = batch_lanes(B), lds = batch_lds(B);
  SK_TRY(start_step(B, x_curr, x_prev, x_bc));
  if (B->coarse() && (!B->coarse_built || (B->coarse_period > 0 && B->steps_since_coarse >= B->coarse_period))) {
    SK_TRY(build_coarse(B));
    B->steps_since_coarse = 0;
  }
  ++B->steps_since_coarse;
  int n_active = 0, r = 0;
  SK_TRY(read_active(B, &n_active));
  if (B->prm.use_graph) SK_TRY(ensure_graph(B, block, lanes, lds));
  while (n_active > 0 && r < B->prm.newton_max_iter) {
    if (B->prm.use_graph) SK_HIP_CHECK(hipGraphLaunch(B->exec, B->st));
    else SK_TRY(newton_iteration(B, block, lanes, lds));
    ++r;
    SK_TRY(read_active(B, &n_active));
  }
  const int64_t n3 = 3 * B->mesh->nv;
  SK_HIP_CHECK(hipMemcpyAsync(x_next, B->x.p, n3 * sizeof(double), hipMemcpyDeviceToDevice, B->st));
  if (info) {
    std::vector<int> st(B->P), ni(B->P), ls(B->P), rf(B->P);
    std::vector<long long> pc(B->P);
    std::vector<double> g0(B->P), gn(B->P);
    std::vector<int> sl(B->P);
    SK_HIP_CHECK(hipMemcpyAsync(st.data(), B->status.p, B->P * sizeof(int), hipMemcpyDeviceToHost, B->st));
    SK_HIP_CHECK(hipMemcpyAsync(ni.data(), B->newton_iters.p, B->P * sizeof(int), hipMemcpyDeviceToHost, B->st));
    SK_HIP_CHECK(hipMemcpyAsync(ls.data(), B->ls_trials.p, B->P * sizeof(int), hipMemcpyDeviceToHost, B->st));
    SK_HIP_CHECK(hipMemcpyAsync(rf.data(), B->refines.p, B->P * sizeof(int), hipMemcpyDeviceToHost, B->st));
    SK_HIP_CHECK(hipMemcpyAsync(pc.data(), B->pcg_iters.p, B->P * sizeof(long long), hipMemcpyDeviceToHost, B->st));
    SK_HIP_CHECK(hipMemcpyAsync(g0.data(), B->g0.p, B->P * sizeof(double), hipMemcpyDeviceToHost, B->st));
    SK_HIP_CHECK(hipMemcpyAsync(gn.data(), B->gnorm.p, B->P * sizeof(double), hipMemcpyDeviceToHost, B->st));
    SK_HIP_CHECK(hipMemcpyAsync(sl.data(), B->stalled.p, B->P * sizeof(int), hipMemcpyDeviceToHost, B->st));
    SK_HIP_CHECK(hipStreamSynchronize(B->st));
    for (int q = 0; q < B->P; ++q) {
      const bool rel = gn[q] <= B->prm.newton_rtol * g0[q] || g0[q] == 0.0;
      const bool conv = rel || sl[q];
      info[q].status = (st[q] == SK_OK && !conv) ? SK_ERR_NEWTON_MAXIT : st[q];
      info[q].newton_iters = ni[q];
      info[q].pcg_iters = pc[q];
      info[q].ls_trials = ls[q];
      info[q].refine_steps = rf[q];
      info[q].g0_norm = g0[q];
      info[q].g_norm = gn[q];
      info[q].stalled = (st[q] == SK_OK && !rel && sl[q]) ? 1 : 0;
    }
  }
  if (rounds) *rounds = r;
  SK_HIP_CHECK(hipEventRecord(B->ev_out, B->st));
  SK_HIP_CHECK(hipStreamWaitEvent(caller, B->ev_out, 0));
  return SK_OK;
}

int sk_batch_set_coarse(sk_batch* B, const double* X, int agg_rows, int period) {
  SK_REQUIRE(B, "sk_batch_set_coarse: null batch");
  SK_REQUIRE(period >= 0, "sk_batch_set_coarse: period must be >= 0");
  if (B->exec) (void)hipGraphExecDestroy(B->exec);
  if (B->graph) (void)hipGraphDestroy(B->graph);
  B->exec = nullptr;
  B->graph = nullptr;
  B->agg_rows = B->n_agg = B->max_agg = B->max_nc = 0;
  B->coarse_built = false;
  for (DevBuf<float>* b : {&B->cdv, &B->ctm, &B->einv}) b->release();
  for (DevBuf<int32_t>* b : {&B->cagg0, &B->ccol, &B->cagg_prob, &B->cadj_ptr, &B->cadj, &B->cnq}) b->release();
  B->einv_off.release();
  B->cE.release();
  B->cbad.release();
  if (agg_rows <= 0) return SK_OK;
  SK_REQUIRE(X, "sk_batch_set_coarse: null positions");
  SK_REQUIRE(agg_rows % kWave == 0 && agg_rows <= 512, "sk_batch_set_coarse: agg_rows must be 64, 128, ..., 512");
  sk_mesh* m = B->mesh;
  const int64_t nv = m->nv;
  const int P = B->P;
  std::vector<int32_t> vs(P + 1), bptr(nv + 1), bcol(m->nblk);
  SK_HIP_CHECK(hipMemcpy(vs.data(), B->vstart.p, (P + 1) * sizeof(int32_t), hipMemcpyDeviceToHost));
  SK_HIP_CHECK(hipMemcpy(bptr.data(), m->bptr.p, (nv + 1) * sizeof(int32_t), hipMemcpyDeviceToHost));
  if (m->nblk) SK_HIP_CHECK(hipMemcpy(bcol.data(), m->bcol.p, m->nblk * sizeof(int32_t), hipMemcpyDeviceToHost));
  std::vector<uint8_t> fixed;
  if (B->fixed.p) {
    fixed.resize(3 * nv);
    SK_HIP_CHECK(hipMemcpy(fixed.data(), B->fixed.p, 3 * nv, hipMemcpyDeviceToHost));
  }
  // aggregates: runs of agg_rows consecutive vertices of each problem
  std::vector<int32_t> cagg0(P + 1, 0), ccol(1, 0), cagg_prob, cnq(P), vagg(nv);
  std::vector<int64_t> einv_off(P);
  std::vector<double> dv(3 * nv), tm;
  int64_t etot = 0;
  int max_agg = 0, max_nc = 0;
  for (int q = 0; q < P; ++q) {
    const int64_t v0 = vs[q], v1 = vs[q + 1];
    const int nagg = (int)((v1 - v0 + agg_rows - 1) / agg_rows);
    cagg0[q + 1] = cagg0[q] + nagg;
    const int c_start = ccol.back();
    for (int l = 0; l < nagg; ++l) {
      const int64_t r0 = v0 + (int64_t)l * agg_rows, r1 = std::min<int64_t>(v1, r0 + agg_rows);
      double cen[3] = {0.0, 0.0, 0.0};
      for (int64_t v = r0; v < r1; ++v)
        for (int a = 0; a < 3; ++a) cen[a] += X[3 * v + a];
      for (int a = 0; a < 3; ++a) cen[a] /= (double)(r1 - r0);
      for (int64_t v = r0; v < r1; ++v) {
        vagg[v] = cagg0[q] + l;
        for (int a = 0; a < 3; ++a) dv[3 * v + a] = X[3 * v + a] - cen[a];
      }
      double t[36];
      const int k = rigid_basis(dv.data(), fixed.empty() ? nullptr : fixed.data(), nullptr, (int)(r1 - r0), r0, t);
      tm.insert(tm.end(), t, t + 36);
      ccol.push_back(ccol.back() + k);
      cagg_prob.push_back(q);
    }
    cnq[q] = ccol.back() - c_start;
    einv_off[q] = etot;
    etot += (int64_t)cnq[q] * cnq[q];
    max_agg = std::max(max_agg, nagg);
    max_nc = std::max(max_nc, cnq[q]);
  }
  const int n_agg = cagg0[P];
  SK_REQUIRE(max_nc <= 1024, "sk_batch_set_coarse: a problem has more than 1024 coarse unknowns");
  std::vector<std::vector<int32_t>> nbr(n_agg);
  for (int64_t v = 0; v < nv; ++v)
    for (int32_t b = bptr[v]; b < bptr[v + 1]; ++b) nbr[vagg[v]].push_back(vagg[bcol[b]]);
  std::vector<int32_t> adj_ptr(n_agg + 1, 0), adj;
  for (int g = 0; g < n_agg; ++g) {
    auto& l = nbr[g];
    std::sort(l.begin(), l.end());
    l.erase(std::unique(l.begin(), l.end()), l.end());
    adj.insert(adj.end(), l.begin(), l.end());
    adj_ptr[g + 1] = (int32_t)adj.size();
  }
  B->agg_rows = agg_rows;
  B->coarse_period = period;
  B->steps_since_coarse = 0;
  B->n_agg = n_agg;
  B->max_agg = max_agg;
  B->max_nc = max_nc;
  auto fail = [&](int st) {
    B->agg_rows = 0;
    return st;
  };
  if (pcg_shm_bytes(B, 0) > 60 * 1024) {
    set_error("sk_batch_set_coarse: coarse scratch exceeds the workgroup's LDS");
    return fail(SK_ERR_INVALID);
  }
  std::vector<float> dv32(dv.begin(), dv.end()), tm32(tm.begin(), tm.end());
  int s = SK_OK;
  if ((s = B->cdv.upload(dv32)) || (s = B->ctm.upload(tm32)) || (s = B->cagg0.upload(cagg0)) ||
      (s = B->ccol.upload(ccol)) || (s = B->cagg_prob.upload(cagg_prob)) || (s = B->cadj_ptr.upload(adj_ptr)) ||
      (s = B->cadj.upload(adj)) || (s = B->cnq.upload(cnq)) || (s = B->einv_off.upload(einv_off)) ||
      (s = B->cE.alloc(etot)) || (s = B->einv.alloc(etot)) || (s = B->cbad.alloc(P)))
    return fail(s);
  return SK_OK;
}

int sk_batch_set_shape(sk_batch* B, int block, int lanes, int lds) {
  SK_REQUIRE(B, "sk_batch_set_shape: null batch");
  SK_REQUIRE(block == 128 || block == 256 || block == 512 || block == 1024, "sk_batch_set_shape: bad block");
  SK_REQUIRE(lanes == 1 || lanes == 4, "sk_batch_set_shape: lanes must be 1 or 4");
  SK_REQUIRE(lds == 0 || (lds == 1 && pcg_lds_fits(B)), "sk_batch_set_shape: lds must be 0, or 1 when it fits");
  B->prm.block = block;
  B->prm.lanes = lanes;
  B->prm.lds = lds;
  return SK_OK;
}

int sk_batch_autotune(sk_batch* B, const double* x_curr, const double* x_prev, const double* x_bc, int reps,
                      int* block_out, int* lanes_out, int* lds_out, double* ms_out, sk_stream_t s) {
  SK_REQUIRE(B && x_curr && x_prev && block_out && lanes_out, "sk_batch_autotune: null argument");
  SK_REQUIRE(reps >= 1, "sk_batch_autotune: reps must be >= 1");
  hipStream_t caller = as_stream(s);
  SK_HIP_CHECK(hipEventRecord(B->ev_in, caller));
  SK_HIP_CHECK(hipStreamWaitEvent(B->st, B->ev_in, 0));
  // The first Newton system of a step: gradient, Hessian, gather.
  SK_TRY(start_step(B, x_curr, x_prev, x_bc));
  if (B->coarse() && !B->coarse_built) SK_TRY(build_coarse(B));
  SK_TRY(hessian_phase(B, false));
  BatchArgs a = B->args();
  a.count_pcg = 0;
  hipEvent_t e0, e1;
  SK_HIP_CHECK(hipEventCreate(&e0));
  SK_HIP_CHECK(hipEventCreate(&e1));
  const int blocks[4] = {128, 256, 512, 1024}, lanes[2] = {1, 4};
  double best = 1e300;
  int bb = 256, bl = 1, bd = 0, c = 0;
  for (int ld = 0; ld < 2; ++ld)
    for (int bi = 0; bi < 4; ++bi)
      for (int li = 0; li < 2; ++li, ++c) {
        if (ld && !pcg_lds_fits(B)) {
          if (ms_out) ms_out[c] = -1.0;
          continue;
        }
        SK_TRY(launch_pcg(B, a, blocks[bi], lanes[li], ld));  // warm-up
        SK_HIP_CHECK(hipEventRecord(e0, B->st));
        for (int k = 0; k < reps; ++k) SK_TRY(launch_pcg(B, a, blocks[bi], lanes[li], ld));
        SK_HIP_CHECK(hipEventRecord(e1, B->st));
        SK_HIP_CHECK(hipEventSynchronize(e1));
        float ms = 0.f;
        SK_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
        const double t = ms / reps;
        if (ms_out) ms_out[c] = t;
        if (t < best) {
          best = t;
          bb = blocks[bi];
          bl = lanes[li];
          bd = ld;
        }
      }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  B->prm.block = bb;
  B->prm.lanes = bl;
  B->prm.lds = bd;
  *block_out = bb;
  *lanes_out = bl;
  if (lds_out) *lds_out = bd;
  SK_HIP_CHECK(hipEventRecord(B->ev_out, B->st));
  SK_HIP_CHECK(hipStreamWaitEvent(caller, B->ev_out, 0));
  return SK_OK;
}

}  // extern "C"
