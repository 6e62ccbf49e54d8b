// Preconditioned conjugate gradients and the device backward-Euler driver.
//
// PCG iteration k = 2 launches, no fences, no atomics:
//   spmv(k):   every block first sums the previous launch's per-block
//              partials (r.z, r.r) in a fixed order, derives beta and the
//              convergence test itself (all blocks get identical values),
//              then forms p = z + beta p_prev on the fly for every column it
//              touches (ping-pong p buffers), writes the new p and A p for its
//              rows and one partial of p.Ap.
//   update(k): every block sums the spmv partials (p.Ap), alpha = r.z / p.Ap,
//              updates x, r, z = B^-1 r for its vertices and writes partials
//              of the new r.z and r.r.
// Block 0 of each launch records the scalars (r.z by iteration parity, r.r,
// iteration count, stop flag) in a device struct that only later launches
// read, so every launch returns at once after convergence and the host reads
// the struct once per chunk of enqueued iterations.
#include "sk_coarse.h"

#include <cmath>

namespace sk {

struct PcgState {
  double rz[2];  // r.z of iteration k in slot k & 1
  double tol2, rr, bb, pAp;
  int iter, done;  // done: 0 running, 1 converged, 2 breakdown (p.Ap <= 0), 3 non-finite
};

constexpr int kRowLanes = 8;      // lanes cooperating on one (block) row in SpMV
constexpr int kSpmvBlocks = 1024; // SpMV grid cap; blocks walk contiguous row chunks

// Shared prologue of the SpMV launches: reduce the update partials, decide
// convergence, compute beta. Returns false if the iteration must not run.
struct PcgCtl {
  double beta;
  bool run;
};

// Block 0 of a launch may set st->done while other blocks of the same launch
// are starting, so the flag is read once per block (thread 0, through LDS):
// every thread of a block takes the same branch around the barriers below.
__device__ __forceinline__ bool block_reads_done(const PcgState* st) {
  __shared__ int s_done;
  if (threadIdx.x == 0) s_done = st->done;
  __syncthreads();
  return s_done != 0;
}

// restart: first launch after a residual replacement (the stopping threshold
// restarts from the replaced residual; beta continues from r.z of k - 1).
__device__ __forceinline__ PcgCtl spmv_prologue(PcgState* st, const double* upart, int gu, int k, double rtol2,
                                                int restart = 0) {
  PcgCtl c{0.0, false};
  if (block_reads_done(st)) return c;
  double red[2];
  sum_partials<2>(upart, gu, gu, red);
  const double rz = red[0], rr = red[1];
  const bool fresh = (k == 0) || restart;
  const double tol2 = fresh ? rtol2 * rr : st->tol2;
  const bool finite = isfinite(rr) && isfinite(rz);
  const bool conv = finite && !(rr > tol2);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->rz[k & 1] = rz;
    st->rr = rr;
    if (fresh) st->tol2 = tol2;
    if (k == 0) st->bb = rr;
    if (!finite) st->done = 3;
    else if (conv) st->done = 1;
  }
  if (!finite || conv) return c;
  c.beta = (k == 0) ? 0.0 : rz / st->rz[(k - 1) & 1];
  c.run = true;
  return c;
}

// The same prologue in two halves, so that its loads (the flag, this thread's
// update partials, r.z of k - 1) are in flight while the caller issues its own
// first loads: spmv_prologue_issue only loads (every address valid, values
// used in the second half); spmv_prologue_finish publishes the flag, reduces
// and decides exactly as spmv_prologue. Partials past kPreParts per thread are
// read in the second half.
constexpr int kPreParts = 4;
struct PrologueLoads {
  int done;
  double part[2][kPreParts];  // upart[v gu + threadIdx.x + q kBlock], valid where that index < gu
  double rzprev;              // st->rz[(k - 1) & 1] (unused for k = 0)
};

__device__ __forceinline__ PrologueLoads spmv_prologue_issue(const PcgState* st, const double* upart, int gu,
                                                             int k) {
  PrologueLoads L;
  L.done = st->done;
#pragma unroll
  for (int v = 0; v < 2; ++v)
#pragma unroll
    for (int q = 0; q < kPreParts; ++q) {
      const int i = (int)threadIdx.x + q * kBlock;
      L.part[v][q] = upart[v * gu + (i < gu ? i : gu - 1)];
    }
  L.rzprev = st->rz[(k + 1) & 1];
  return L;
}

// hst (optional): the host's pinned mirror of iter and done, written with st.
__device__ __forceinline__ PcgCtl spmv_prologue_finish(const PrologueLoads& L, PcgState* st, const double* upart,
                                                       int gu, int k, double rtol2, int restart,
                                                       PcgState* hst = nullptr) {
  __shared__ int s_done;
  __shared__ double lds[2 * kBlock / kWave];
  __shared__ double res[2];
  PcgCtl c{0.0, false};
  if (threadIdx.x == 0) s_done = L.done;  // thread 0's read decides for the block
  __syncthreads();
  if (s_done != 0) return c;
  double acc[2];
#pragma unroll
  for (int v = 0; v < 2; ++v) {
    double sum = 0.0;
#pragma unroll
    for (int q = 0; q < kPreParts; ++q)
      if ((int)threadIdx.x + q * kBlock < gu) sum += L.part[v][q];
    for (int i = (int)threadIdx.x + kPreParts * kBlock; i < gu; i += kBlock) sum += upart[v * gu + i];
    acc[v] = sum;
  }
  block_sum<2>(acc, lds);
  if (threadIdx.x == 0) {
    res[0] = acc[0];
    res[1] = acc[1];
  }
  __syncthreads();
  const double rz = res[0], rr = res[1];
  const bool fresh = (k == 0) || restart;
  const double tol2 = fresh ? rtol2 * rr : st->tol2;
  const bool finite = isfinite(rr) && isfinite(rz);
  const bool conv = finite && !(rr > tol2);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->rz[k & 1] = rz;
    st->rr = rr;
    if (fresh) st->tol2 = tol2;
    if (k == 0) st->bb = rr;
    const int d = !finite ? 3 : conv ? 1 : 0;
    if (d) {
      st->done = d;
      if (hst) hst->done = d;
    }
  }
  if (!finite || conv) return c;
  c.beta = (k == 0) ? 0.0 : rz / L.rzprev;
  c.run = true;
  return c;
}

// The update prologue in two halves, as for the SpMV: update_prologue_issue
// loads the flag, this thread's SpMV partials of p.Ap (kPreSParts, clamped
// indices) and r.z of k; update_prologue_finish publishes the flag, reduces in
// the order of sum_partials and decides exactly as update_prologue.
constexpr int kPreSParts = 8;
struct UpdatePrologueLoads {
  int done;
  double part[kPreSParts];  // spart[threadIdx.x + q kBlock], valid where that index < gs
  double rz;                // st->rz[k & 1]
};

__device__ __forceinline__ UpdatePrologueLoads update_prologue_issue(const PcgState* st, const double* spart, int gs,
                                                                     int k) {
  UpdatePrologueLoads L;
  L.done = st->done;
#pragma unroll
  for (int q = 0; q < kPreSParts; ++q) {
    const int i = (int)threadIdx.x + q * kBlock;
    L.part[q] = spart[i < gs ? i : gs - 1];
  }
  L.rz = st->rz[k & 1];
  return L;
}

__device__ __forceinline__ bool update_prologue_finish(const UpdatePrologueLoads& L, PcgState* st,
                                                       const double* spart, int gs, int k, double& alpha,
                                                       PcgState* hst = nullptr) {
  __shared__ int s_done;
  __shared__ double lds[kBlock / kWave];
  __shared__ double res;
  if (threadIdx.x == 0) s_done = L.done;  // thread 0's read decides for the block
  __syncthreads();
  if (s_done != 0) return false;
  double acc[1];
  {
    double sum = 0.0;
#pragma unroll
    for (int q = 0; q < kPreSParts; ++q)
      if ((int)threadIdx.x + q * kBlock < gs) sum += L.part[q];
    for (int i = (int)threadIdx.x + kPreSParts * kBlock; i < gs; i += kBlock) sum += spart[i];
    acc[0] = sum;
  }
  block_sum<1>(acc, lds);
  if (threadIdx.x == 0) res = acc[0];
  __syncthreads();
  const double pAp = res;
  if (!(pAp > 0.0)) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      st->pAp = pAp;
      st->done = isfinite(pAp) ? 2 : 3;
      if (hst) hst->done = st->done;
    }
    return false;
  }
  alpha = L.rz / pAp;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->pAp = pAp;
    st->iter = k + 1;
    if (hst) hst->iter = k + 1;
  }
  return true;
}

// Shared prologue of the update launches: reduce p.Ap, detect breakdown.
__device__ __forceinline__ bool update_prologue(PcgState* st, const double* spart, int gs, int k, double& alpha) {
  if (block_reads_done(st)) return false;
  double red[1];
  sum_partials<1>(spart, gs, gs, red);
  const double pAp = red[0];
  if (!(pAp > 0.0)) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      st->pAp = pAp;
      st->done = isfinite(pAp) ? 2 : 3;
    }
    return false;
  }
  alpha = st->rz[k & 1] / pAp;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->pAp = pAp;
    st->iter = k + 1;
  }
  return true;
}

// ---------------------------------------------------------------------------
// FEM operator on the simulator's BSR values (block b's 3x3 at
// vals[9 b .. 9 b + 9)), templated on the precision T of values and vectors
// (double: plain path; float: inner solve of the mixed-precision path). Dot
// products are always accumulated in FP64.
// ---------------------------------------------------------------------------
// SELL-C-sigma description of the matrix (see sk_mesh).
struct SellView {
  int64_t nchunk;
  const int32_t* sell_row;
  const int32_t* chunk_deg;
  const int32_t* chunk_off;
  const int32_t* scol;
  const int32_t* corder;  // XCD-balanced chunk order, 8 lists of lpad (-1: none)
  int64_t lpad;
  const int4* cmeta;      // corder's entries with their chunks' slot count and offset
};

template <typename T>
struct BlockPcgArgs {
  int64_t nv;
  SellView sv;
  const T* vals;
  const T* binv;
  const uint8_t* fixed;
  const T* b;
  T* x;
  T* r;
  T* z;
  const T* pin;
  T* pout;
  T* Ap;
  PcgState* st;
  double rtol2;
  double* upart;  // [2][gu] partials of r.z and r.r
  int gu;
  double* spart;  // [gs] partials of p.Ap
  int gs;
  int k;          // iteration index of this launch
  int restart;    // 1 on the first launch after a residual replacement
  CoarseView<T> cv;  // two-level preconditioner (cv.n_agg = 0: block-Jacobi only)
  PcgState* hst;     // pinned host mirror of st's iter and done (block 0 writes both)
};

template <typename T>
__device__ __forceinline__ void apply_binv(const T* __restrict__ binv, int64_t v, const T (&r)[3], T (&z)[3]) {
  T B[9];
  binv_load(binv, v, B);
#pragma unroll
  for (int p = 0; p < 3; ++p) z[p] = B[3 * p] * r[0] + B[3 * p + 1] * r[1] + B[3 * p + 2] * r[2];
}

// Vertex loop of the vertex-wise kernels: grid-stride in index order, or
// (kAgg, two-level preconditioner) workgroup I over the vertices of
// aggregate I, which also restricts the residual to that aggregate.
template <bool kAgg, typename T, typename F>
__device__ __forceinline__ void for_vertices(const BlockPcgArgs<T>& a, F f) {
  if (kAgg) {
    const int32_t t1 = a.cv.agg_ptr[blockIdx.x + 1];
    for (int32_t t = a.cv.agg_ptr[blockIdx.x] + threadIdx.x; t < t1; t += blockDim.x) f(a.cv.vert(t));
  } else {
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < a.nv; v += (int64_t)gridDim.x * blockDim.x)
      f(v);
  }
}

// raw[0..2] += r, raw[3..5] += d x r: the vertex's term of R_v^T r (d: its
// offset from the aggregate's centroid).
template <typename T>
__device__ __forceinline__ void restrict_add(const T (&d)[3], const T (&r)[3], double (&raw)[6]) {
  const double d0 = (double)d[0], d1 = (double)d[1], d2 = (double)d[2];
  const double r0 = (double)r[0], r1 = (double)r[1], r2 = (double)r[2];
  raw[0] += r0;
  raw[1] += r1;
  raw[2] += r2;
  raw[3] += d1 * r2 - d2 * r1;
  raw[4] += d2 * r0 - d0 * r2;
  raw[5] += d0 * r1 - d1 * r0;
}

// The aggregate's coarse columns and, in threads 0..5, column threadIdx.x of
// its basis T_I: loaded at kernel entry so that the restriction at the end
// does not wait for them.
struct AggPre {
  double col[6];
  double dsc;  // thread j < k: the scale of coarse unknown c0 + j
  int c0, k;
};

template <bool kAgg, typename T>
__device__ __forceinline__ AggPre agg_prefetch(const BlockPcgArgs<T>& a) {
  AggPre pre;
#pragma unroll
  for (int q = 0; q < 6; ++q) pre.col[q] = 0.0;
  pre.c0 = pre.k = 0;
  pre.dsc = 0.0;
  if (kAgg) {
    pre.c0 = a.cv.agg_col[blockIdx.x];
    pre.k = a.cv.agg_col[blockIdx.x + 1] - pre.c0;
    if ((int)threadIdx.x < pre.k) pre.dsc = a.cv.dsc[pre.c0 + threadIdx.x];
    if (threadIdx.x < 6) {
      const double* tm = a.cv.tm + 36 * (int64_t)blockIdx.x + threadIdx.x;
#pragma unroll
      for (int q = 0; q < 6; ++q) pre.col[q] = tm[6 * q];
    }
  }
  return pre;
}

// Per-workgroup partials of r.z and r.r; with kAgg also the aggregate's
// scaled restricted residual c_I = D_I T_I^T raw (thread j < k writes c_I[j]).
template <bool kAgg, typename T>
__device__ __forceinline__ void finish_vertex_pass(const BlockPcgArgs<T>& a, double rz, double rr,
                                                   const double (&raw)[6], const AggPre& pre) {
  if (!kAgg) {
    double vals[2] = {rz, rr};
    store_block_partials<2>(vals, a.upart);
    return;
  }
  __shared__ double lds[8 * kBlock / kWave];
  __shared__ double s_raw[6];
  double v[8] = {rz, rr, raw[0], raw[1], raw[2], raw[3], raw[4], raw[5]};
  block_sum<8>(v, lds);
  if (threadIdx.x == 0) {
    a.upart[blockIdx.x] = v[0];
    a.upart[gridDim.x + blockIdx.x] = v[1];
#pragma unroll
    for (int q = 0; q < 6; ++q) s_raw[q] = v[2 + q];
  }
  __syncthreads();
  if ((int)threadIdx.x < pre.k) {
    double c = 0.0;
#pragma unroll
    for (int q = 0; q < 6; ++q) c += pre.col[q] * s_raw[q];
    a.cv.c[pre.c0 + threadIdx.x] = c * pre.dsc;
  }
}

template <typename T, bool kAgg>
__global__ __launch_bounds__(kBlock) void bpcg_init_kernel(BlockPcgArgs<T> a) {
  const AggPre pre = agg_prefetch<kAgg>(a);
  double rz = 0.0, rr = 0.0, raw[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for_vertices<kAgg>(a, [&](int64_t v) {
    T r[3], z[3];
#pragma unroll
    for (int p = 0; p < 3; ++p) {
      const int64_t k = 3 * v + p;
      r[p] = (a.fixed && a.fixed[k]) ? T(0) : a.b[k];
    }
    apply_binv(a.binv, v, r, z);
#pragma unroll
    for (int p = 0; p < 3; ++p) {
      const int64_t k = 3 * v + p;
      a.r[k] = r[p];
      a.z[k] = z[p];
      a.x[k] = T(0);
      a.pout[k] = T(0);
      rz += (double)r[p] * (double)z[p];
      rr += (double)r[p] * (double)r[p];
    }
    if (kAgg) {
      const T d[3] = {a.cv.dv[3 * v], a.cv.dv[3 * v + 1], a.cv.dv[3 * v + 2]};
      restrict_add(d, r, raw);
    }
  });
  finish_vertex_pass<kAgg>(a, rz, rr, raw, pre);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    a.st->done = 0;
    a.st->iter = 0;
    if (a.hst) {
      a.hst->done = 0;
      a.hst->iter = 0;
    }
  }
}

// Residual replacement: r already holds the replaced residual; z = B^-1 r,
// the correction accumulator x restarts from 0, the search direction is kept.
template <typename T, bool kAgg>
__global__ __launch_bounds__(kBlock) void bpcg_restart_kernel(BlockPcgArgs<T> a) {
  const AggPre pre = agg_prefetch<kAgg>(a);
  double rz = 0.0, rr = 0.0, raw[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for_vertices<kAgg>(a, [&](int64_t v) {
    T r[3], z[3];
#pragma unroll
    for (int p = 0; p < 3; ++p) r[p] = a.r[3 * v + p];
    apply_binv(a.binv, v, r, z);
#pragma unroll
    for (int p = 0; p < 3; ++p) {
      const int64_t k = 3 * v + p;
      a.z[k] = z[p];
      a.x[k] = T(0);
      rz += (double)r[p] * (double)z[p];
      rr += (double)r[p] * (double)r[p];
    }
    if (kAgg) {
      const T d[3] = {a.cv.dv[3 * v], a.cv.dv[3 * v + 1], a.cv.dv[3 * v + 2]};
      restrict_add(d, r, raw);
    }
  });
  finish_vertex_pass<kAgg>(a, rz, rr, raw, pre);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    a.st->done = 0;
    if (a.hst) a.hst->done = 0;
  }
}

// The SpMV kernels take their chunks from the XCD-balanced order: workgroup
// b runs on XCD b % 8 and is that XCD's (b / 8)-th workgroup (the grid is a
// multiple of 8). One chunk of kSellC rows per wave, kSellLanes lanes per row
// (lane = j * kSellC + i: row i of the chunk, slots s = j, j + kSellLanes, ...).
//
// Range form: each workgroup walks `per` consecutive groups of kWaves chunks
// of its XCD's list.
__device__ __forceinline__ int64_t chunks_per_block(const SellView& sv) {
  constexpr int kWaves = kBlock / kWave;
  const int64_t per_xcd = gridDim.x / 8;
  return (sv.lpad + per_xcd * kWaves - 1) / (per_xcd * kWaves);
}

// Position q of this wave in its XCD's list; -2 past the end, -1 padding.
__device__ __forceinline__ int64_t block_chunk(const SellView& sv, int wid, int64_t q, int64_t per) {
  constexpr int kWaves = kBlock / kWave;
  const int64_t pos = ((int64_t)(blockIdx.x / 8) * per + q) * kWaves + wid;
  if (pos >= sv.lpad) return -2;
  return sv.corder[(blockIdx.x % 8) * sv.lpad + pos];
}

// The entry (chunk, slot count, offset) of position q of this wave in its XCD's
// list, as block_chunk / sell_chunk (range = false: q = 0 of the one-chunk-
// per-wave form) but with one 16-byte load; chunk -2 past the end, -1 padding.
template <bool kRange>
__device__ __forceinline__ int4 chunk_entry(const SellView& sv, int wid, int64_t q, int64_t per) {
  constexpr int kWaves = kBlock / kWave;
  const int64_t pos = kRange ? ((int64_t)(blockIdx.x / 8) * per + q) * kWaves + wid
                             : (int64_t)(blockIdx.x / 8) * kWaves + wid;
  int4 e = sv.cmeta[(blockIdx.x % 8) * sv.lpad + (pos < sv.lpad ? pos : sv.lpad - 1)];
  if (pos >= sv.lpad) e.x = kRange ? -2 : -1;
  return e;
}

// One-chunk-per-wave form (the grid covers 8 * lpad waves); -1: no chunk.
__device__ __forceinline__ int64_t sell_chunk(const SellView& sv, int wid) {
  constexpr int kWaves = kBlock / kWave;
  const int64_t pos = (int64_t)(blockIdx.x / 8) * kWaves + wid;
  if (pos >= sv.lpad) return -1;
  return sv.corder[(blockIdx.x % 8) * sv.lpad + pos];
}

// y[p] = sum over the row's slots of A_s[p][:] . xv(col_s), where xv(w) gives
// the 3-vector multiplied at column vertex w; summed over the row's
// kSellLanes lanes (fixed xor order) and returned to every lane of the row.
template <typename T, typename V, typename XV>
__device__ __forceinline__ void sell_row_product(const SellView& sv, const T* __restrict__ vals, int64_t c, int i,
                                                 int j, XV xv, V (&y)[3]) {
  const int32_t deg = sv.chunk_deg[c];
  const int64_t off = sv.chunk_off[c];
  for (int32_t s = j; s < deg; s += kSellLanes) {
    const int64_t g = off + (int64_t)s * kSellC;  // slot group of the chunk's rows
    const int64_t w = 3 * (int64_t)sv.scol[g + i];
    V pv[3];
    xv(w, pv);
    T A[9];
    load_slot<T, kSellC, 1>(vals, g, i, A);
#pragma unroll
    for (int p = 0; p < 3; ++p) y[p] += (V)A[3 * p] * pv[0] + (V)A[3 * p + 1] * pv[1] + (V)A[3 * p + 2] * pv[2];
  }
#pragma unroll
  for (int off2 = kSellC; off2 < kWave; off2 <<= 1) {
#pragma unroll
    for (int p = 0; p < 3; ++p) y[p] += __shfl_xor(y[p], off2, kWave);
  }
}

// SpMV of PCG iteration k on the SELL matrix: forms p = z + beta p_prev on
// the fly for every column touched, writes the new p and A p of its rows and
// a partial of p.Ap.
// One SELL chunk's rows of the PCG SpMV (see bpcg_spmv_kernel).
template <typename T>
__device__ __forceinline__ void spmv_chunk(const BlockPcgArgs<T>& a, T beta, int64_t ch, int lane, double& dot) {
  const int i = lane % kSellC, j = lane / kSellC;
  T y[3] = {T(0), T(0), T(0)};
  const T* __restrict__ z = a.z;
  const T* __restrict__ pin = a.pin;
  sell_row_product<T, T>(a.sv, a.vals, ch, i, j, [&](int64_t w, T (&pv)[3]) {
#pragma unroll
    for (int q = 0; q < 3; ++q) pv[q] = z[w + q] + beta * pin[w + q];
  }, y);
  const int32_t row = a.sv.sell_row[ch * kSellC + i];
  if (row >= 0 && j == 0) {
#pragma unroll
    for (int p = 0; p < 3; ++p) {
      const int64_t k = 3 * (int64_t)row + p;
      const T pn = z[k] + beta * pin[k];
      const T ap = (a.fixed && a.fixed[k]) ? T(0) : y[p];
      a.pout[k] = pn;
      a.Ap[k] = ap;
      dot += (double)pn * (double)ap;
    }
  }
}

// A lane's first slot (s = j) of a chunk: its block and column, in FP32 also
// the z and previous-p entries at that column, loaded before beta is known,
// and the column of its next slot (s = j + kSellLanes). Every load is
// unconditional (slots clamped to the chunk's last; a padding entry carries
// chunk 0's slot count and offset); has says whether the lane has the slot.
// In FP64 the six vector entries would take the kernel past 64 VGPRs, below
// full residency of its grid (measured: 22.35 us without them against 22.63
// us with; FP32: 13.85 us with them against 14.3 us without).
template <typename T>
constexpr bool kPrefetchVec = sizeof(T) == 4;

template <typename T>
struct FirstSlot {
  T A[9], z[3], pin[3];
  int64_t w;
  int32_t deg, off, ncol;
  bool has;
};

// e: the chunk's entry (chunk_entry).
template <typename T>
__device__ __forceinline__ FirstSlot<T> first_slot(const BlockPcgArgs<T>& a, int4 e, int i, int j) {
  FirstSlot<T> f;
  f.deg = e.y;
  f.off = e.z;
  f.has = e.x >= 0 && j < f.deg;
  const int64_t g = f.off + (int64_t)(j < f.deg ? j : f.deg - 1) * kSellC;
  const int64_t w = 3 * (int64_t)a.sv.scol[g + i];
  const int32_t jn = j + kSellLanes;
  f.ncol = a.sv.scol[f.off + (int64_t)(jn < f.deg ? jn : f.deg - 1) * kSellC + i];
  const T* __restrict__ A = a.vals + g * 9 + i;
#pragma unroll
  for (int m = 0; m < 9; ++m) f.A[m] = A[m * kSellC];
  f.w = w;
  if constexpr (kPrefetchVec<T>) {
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      f.z[q] = a.z[w + q];
      f.pin[q] = a.pin[w + q];
    }
  }
  return f;
}

// spmv_chunk with the lane's first slot already loaded: the same products in
// the same order (slot j first, then j + kSellLanes, ...). The loop is
// software-pipelined: each pass issues its slot's vector and block loads
// together with the next slot's column, so a slot costs one memory round trip
// instead of two (column, then the entries at that column).
template <typename T>
__device__ __forceinline__ void spmv_chunk_first(const BlockPcgArgs<T>& a, T beta, int64_t ch, int lane,
                                                 const FirstSlot<T>& f, double& dot) {
  const int i = lane % kSellC, j = lane / kSellC;
  T y[3] = {T(0), T(0), T(0)};
  const T* __restrict__ z = a.z;
  const T* __restrict__ pin = a.pin;
  if (f.has) {
    T pv[3];
#pragma unroll
    for (int q = 0; q < 3; ++q)
      pv[q] = kPrefetchVec<T> ? f.z[q] + beta * f.pin[q] : z[f.w + q] + beta * pin[f.w + q];
#pragma unroll
    for (int p = 0; p < 3; ++p)
      y[p] += (T)f.A[3 * p] * pv[0] + (T)f.A[3 * p + 1] * pv[1] + (T)f.A[3 * p + 2] * pv[2];
  }
  const int64_t off = f.off;
  int32_t col = f.ncol;
  for (int32_t s = j + kSellLanes; s < f.deg; s += kSellLanes) {
    const int64_t g = off + (int64_t)s * kSellC;
    const int64_t w = 3 * (int64_t)col;
    const int32_t sn = s + kSellLanes;
    col = a.sv.scol[off + (int64_t)(sn < f.deg ? sn : s) * kSellC + i];
    T pv[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) pv[q] = z[w + q] + beta * pin[w + q];
    const T* __restrict__ A = a.vals + g * 9 + i;
#pragma unroll
    for (int p = 0; p < 3; ++p)
      y[p] += (T)A[(3 * p) * kSellC] * pv[0] + (T)A[(3 * p + 1) * kSellC] * pv[1] +
              (T)A[(3 * p + 2) * kSellC] * pv[2];
    // keeps the next column's load in this pass (issued with the others, waited for last)
    asm volatile("" : "+v"(col));
  }
#pragma unroll
  for (int off2 = kSellC; off2 < kWave; off2 <<= 1) {
#pragma unroll
    for (int p = 0; p < 3; ++p) y[p] += __shfl_xor(y[p], off2, kWave);
  }
  const int32_t row = a.sv.sell_row[ch * kSellC + i];
  if (row >= 0 && j == 0) {
#pragma unroll
    for (int p = 0; p < 3; ++p) {
      const int64_t k = 3 * (int64_t)row + p;
      const T pn = z[k] + beta * pin[k];
      const T ap = (a.fixed && a.fixed[k]) ? T(0) : y[p];
      a.pout[k] = pn;
      a.Ap[k] = ap;
      dot += (double)pn * (double)ap;
    }
  }
}

// SpMV of PCG iteration k on the SELL matrix: forms p = z + beta p_prev on
// the fly for every column touched, writes the new p and A p of its rows and
// a partial of p.Ap. kRange = false: one chunk per wave (the grid covers every
// chunk); true: each workgroup walks a contiguous range of chunk groups. The
// first chunk's first slots are loaded while the prologue's loads are in
// flight.
template <typename T, bool kRange>
__global__ __launch_bounds__(kBlock) void bpcg_spmv_kernel(BlockPcgArgs<T> a) {
  const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
  double dot = 0.0;
  const PrologueLoads L = spmv_prologue_issue(a.st, a.upart, a.gu, a.k);
  const int64_t per = kRange ? chunks_per_block(a.sv) : 1;
  const int4 e0 = chunk_entry<kRange>(a.sv, wid, 0, per);
  const int64_t ch0 = e0.x;
  const FirstSlot<T> f = first_slot(a, e0, lane % kSellC, lane / kSellC);
  const PcgCtl c = spmv_prologue_finish(L, a.st, a.upart, a.gu, a.k, a.rtol2, a.restart, a.hst);
  if (!c.run) return;
  const T beta = (T)c.beta;
  if (ch0 >= 0) spmv_chunk_first(a, beta, ch0, lane, f, dot);
  if (kRange && ch0 != -2) {
    for (int64_t q = 1; q < per; ++q) {
      const int64_t ch = block_chunk(a.sv, wid, q, per);
      if (ch == -2) break;
      if (ch >= 0) spmv_chunk(a, beta, ch, lane, dot);
    }
  }
  double vals[1] = {dot};
  store_block_partials<1>(vals, a.spart);
}

// The per-vertex data of the update (x, p, r, A p, B^-1), loaded before the
// prologue's reductions so that the loads overlap them.
template <typename T>
struct UpdateIn {
  T x[3], p[3], r[3], ap[3], B[9], d[3];  // d: centroid offset (two-level restriction)
};

template <typename T, bool kAgg>
__device__ __forceinline__ void load_update(const BlockPcgArgs<T>& a, int64_t v, UpdateIn<T>& u) {
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const int64_t k = 3 * v + q;
    u.x[q] = a.x[k];
    u.p[q] = a.pout[k];
    u.r[q] = a.r[k];
    u.ap[q] = a.Ap[k];
    u.d[q] = kAgg ? a.cv.dv[k] : T(0);
  }
  binv_load(a.binv, v, u.B);
}

template <typename T, bool kAgg>
__device__ __forceinline__ void apply_update(const BlockPcgArgs<T>& a, int64_t v, const UpdateIn<T>& u, T alpha,
                                             double& rz, double& rr, double (&raw)[6]) {
  T r[3], z[3];
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    a.x[3 * v + q] = u.x[q] + alpha * u.p[q];
    r[q] = u.r[q] - alpha * u.ap[q];
  }
#pragma unroll
  for (int q = 0; q < 3; ++q) z[q] = u.B[3 * q] * r[0] + u.B[3 * q + 1] * r[1] + u.B[3 * q + 2] * r[2];
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const int64_t k = 3 * v + q;
    a.r[k] = r[q];
    a.z[k] = z[q];
    rz += (double)r[q] * (double)z[q];
    rr += (double)r[q] * (double)r[q];
  }
  if (kAgg) restrict_add(u.d, r, raw);
}

template <typename T, bool kAgg>
__global__ __launch_bounds__(kBlock) void bpcg_update_kernel(BlockPcgArgs<T> a) {
  const UpdatePrologueLoads L = update_prologue_issue(a.st, a.spart, a.gs, a.k);
  const AggPre pre = agg_prefetch<kAgg>(a);
  // this thread's first vertex and the stride to its next ones (for_vertices' order)
  int64_t v = -1, t = 0, t_end = 0, stride = 0;
  if (kAgg) {
    t = a.cv.agg_ptr[blockIdx.x] + threadIdx.x;
    t_end = a.cv.agg_ptr[blockIdx.x + 1];
    stride = blockDim.x;
    if (t < t_end) v = a.cv.vert(t);
  } else {
    t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    t_end = a.nv;
    stride = (int64_t)gridDim.x * blockDim.x;
    if (t < t_end) v = t;
  }
  UpdateIn<T> u;
  if (v >= 0) load_update<T, kAgg>(a, v, u);
  // aggregates larger than the workgroup: the thread's second vertex is loaded
  // with the first, so the largest aggregates do not add a dependent round
  int64_t v2 = -1;
  UpdateIn<T> u2;
  if (kAgg && t + stride < t_end) {
    v2 = a.cv.vert(t + stride);
    load_update<T, kAgg>(a, v2, u2);
  }
  double alpha_d;
  if (!update_prologue_finish(L, a.st, a.spart, a.gs, a.k, alpha_d, a.hst)) return;
  const T alpha = (T)alpha_d;
  double rz = 0.0, rr = 0.0, raw[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  if (v >= 0) apply_update<T, kAgg>(a, v, u, alpha, rz, rr, raw);
  if (kAgg) {
    if (v2 >= 0) apply_update<T, kAgg>(a, v2, u2, alpha, rz, rr, raw);
    t += stride;
  }
  for (t += stride; t < t_end; t += stride) {
    const int64_t w = kAgg ? a.cv.vert(t) : t;
    UpdateIn<T> uw;
    load_update<T, kAgg>(a, w, uw);
    apply_update<T, kAgg>(a, w, uw, alpha, rz, rr, raw);
  }
  finish_vertex_pass<kAgg>(a, rz, rr, raw, pre);
}

// The FP16 values in the low and high halves of a 32-bit word.
__device__ __forceinline__ float half_lo(uint32_t w) { return (float)__builtin_bit_cast(_Float16, (uint16_t)(w & 0xFFFFu)); }
__device__ __forceinline__ float half_hi(uint32_t w) { return (float)__builtin_bit_cast(_Float16, (uint16_t)(w >> 16)); }

// A vertex of the coarse kernel's prolongation: aggregate-order position t
// (none when t >= t1), its vertex, centroid offset, z and free mask.
template <typename T>
struct ProlongIn {
  bool has;
  int64_t v;
  double d[3];
  T z[3];
  bool free[3];
};

template <typename T>
__device__ __forceinline__ void load_prolong(const BlockPcgArgs<T>& a, int32_t t, int32_t t1, ProlongIn<T>& p) {
  p.has = t < t1;
  p.v = 0;
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    p.d[q] = 0.0;
    p.z[q] = T(0);
    p.free[q] = false;
  }
  if (!p.has) return;
  p.v = a.cv.vert(t);
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    p.d[q] = (double)a.cv.dva[3 * (int64_t)t + q];
    p.z[q] = a.z[3 * p.v + q];
    p.free[q] = !(a.fixed && a.fixed[3 * p.v + q]);
  }
}

// z += u + w x d on the vertex's free DOFs (uw = (u, w)).
template <typename T>
__device__ __forceinline__ void apply_prolong(const BlockPcgArgs<T>& a, const ProlongIn<T>& p, const double (&uw)[6]) {
  if (!p.has) return;
  const double corr[3] = {uw[0] + uw[4] * p.d[2] - uw[5] * p.d[1], uw[1] + uw[5] * p.d[0] - uw[3] * p.d[2],
                          uw[2] + uw[3] * p.d[1] - uw[4] * p.d[0]};
#pragma unroll
  for (int q = 0; q < 3; ++q)
    if (p.free[q]) a.z[3 * p.v + q] = (T)((double)p.z[q] + corr[q]);
}

// Coarse correction of the two-level preconditioner, one workgroup per
// aggregate I (launched after a kAgg vertex pass of the same grid): with the
// scaled residual c = D W^T r, y'_I = ((S16 + shift I) c)_I and y_I = D_I y'_I
// (so y = D (S16 + shift I) D W^T r, the coarse solve), the aggregate's rigid
// motion (u, w) = T_I y_I is added to z on its free DOFs, and c_I . y'_I is
// added to the workgroup's r.z partial, so the next launch's r.z is
// r . (B^-1 r + W E^-1 W^T r) with the FP16 coarse inverse. No block of this
// launch writes the stop flag (only the SpMV and update launches before it
// do), so every thread reads the same final value at entry: a launch past
// convergence returns before streaming S16 and writes nothing.
template <typename T>
__global__ __launch_bounds__(kBlock) void bpcg_coarse_kernel(BlockPcgArgs<T> a) {
  __shared__ double lds[6 * kBlock / kWave];
  __shared__ double s_y[6], s_uw[6];
  const CoarseView<T>& cv = a.cv;
  const int I = blockIdx.x;
  if (a.st->done) return;
  const int c0 = cv.agg_col[I], k = cv.agg_col[I + 1] - c0;
  const int32_t t0 = cv.agg_ptr[I], t1 = cv.agg_ptr[I + 1];
  double trow[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};  // thread q < 6: row q of T_I
  if (threadIdx.x < 6) {
#pragma unroll
    for (int j = 0; j < 6; ++j) trow[j] = cv.tm[36 * (int64_t)I + 6 * threadIdx.x + j];
  }
  double cs[6], ds[6];  // the aggregate's scaled residual and scales (the same for every thread)
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    cs[q] = q < k ? cv.c[c0 + q] : 0.0;
    ds[q] = q < k ? cv.dsc[c0 + q] : 0.0;
  }
  const double part0 = threadIdx.x == 0 ? a.upart[I] : 0.0;
  // The first two vertices this thread prolongs into (the second in aggregates
  // larger than the workgroup), loaded now: z is final (the previous launch
  // wrote it) and only this thread writes them below.
  const int32_t tp = t0 + (int32_t)threadIdx.x;
  ProlongIn<T> p0, p1;
  load_prolong(a, tp, t1, p0);
  load_prolong(a, tp + kBlock, t1, p1);
  // y'_q = sum_j S16[c0 + q][j] c_j: eight FP16 columns per 16-byte load, up
  // to kU loads per row in flight per thread, FP64 accumulation
  constexpr int kU = 2;
  const int ld8 = cv.ld >> 3;
  const uint4* __restrict__ E8 = reinterpret_cast<const uint4*>(cv.einv);
  const double2* __restrict__ c2 = reinterpret_cast<const double2*>(cv.c);
  double y[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (int base = threadIdx.x; base < ld8; base += kU * kBlock) {
    double2 cc[kU][4];
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      const int j8 = base + u * kBlock;
#pragma unroll
      for (int h = 0; h < 4; ++h) cc[u][h] = j8 < ld8 ? c2[4 * j8 + h] : make_double2(0.0, 0.0);
    }
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      if (q < k) {
        const uint4* row = E8 + (int64_t)(c0 + q) * ld8;
        uint4 e[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u) {
          const int j8 = base + u * kBlock;
          e[u] = j8 < ld8 ? row[j8] : make_uint4(0u, 0u, 0u, 0u);
        }
#pragma unroll
        for (int u = 0; u < kU; ++u) {
          const uint32_t wd[4] = {e[u].x, e[u].y, e[u].z, e[u].w};  // two FP16 columns per word, low first
#pragma unroll
          for (int h = 0; h < 4; ++h)
            y[q] += (double)half_lo(wd[h]) * cc[u][h].x + (double)half_hi(wd[h]) * cc[u][h].y;
        }
      }
    }
  }
  block_sum<6>(y, lds);
  if (threadIdx.x == 0) {
    double cy = 0.0;
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      const double yq = y[q] + cv.shift * cs[q];
      s_y[q] = ds[q] * yq;
      cy += cs[q] * yq;
    }
    a.upart[I] = part0 + cy;
  }
  __syncthreads();
  if (threadIdx.x < 6) {  // (u, w) = T_I y_I
    double u = 0.0;
#pragma unroll
    for (int j = 0; j < 6; ++j) u += trow[j] * s_y[j];
    s_uw[threadIdx.x] = u;
  }
  __syncthreads();
  const double uw[6] = {s_uw[0], s_uw[1], s_uw[2], s_uw[3], s_uw[4], s_uw[5]};
  apply_prolong(a, p0, uw);
  apply_prolong(a, p1, uw);
  for (int32_t t = tp + 2 * kBlock; t < t1; t += kBlock) {  // aggregates above twice the workgroup
    ProlongIn<T> pt;
    load_prolong(a, t, t1, pt);
    apply_prolong(a, pt, uw);
  }
}

// ---------------------------------------------------------------------------
// Iterative refinement of the mixed-precision path: FP64 residual of the
// FP32-stored matrix, r = b - A x (masked to the free DOFs), its FP32 copy for
// the next inner solve, and per-block partials of r.r.
// ---------------------------------------------------------------------------
struct ResidualArgs {
  int64_t nv;
  SellView sv;
  const float* vals;
  const uint8_t* fixed;
  const double* b;
  const double* x;  // nullptr: x = 0
  double* r;
  float* r32;
  double* partials;
};

__global__ __launch_bounds__(kBlock) void ir_residual_kernel(ResidualArgs a) {
  const int lane = threadIdx.x % kWave;
  const int64_t ch = sell_chunk(a.sv, threadIdx.x / kWave);
  double rr = 0.0;
  if (ch >= 0) {
    const int i = lane % kSellC, j = lane / kSellC;
    double y[3] = {0.0, 0.0, 0.0};
    if (a.x) {
      const double* __restrict__ x = a.x;
      sell_row_product<float, double>(a.sv, a.vals, ch, i, j, [&](int64_t w, double (&pv)[3]) {
#pragma unroll
        for (int q = 0; q < 3; ++q) pv[q] = x[w + q];
      }, y);
    }
    const int32_t row = a.sv.sell_row[ch * kSellC + i];
    if (row >= 0 && j == 0) {
#pragma unroll
      for (int p = 0; p < 3; ++p) {
        const int64_t k = 3 * (int64_t)row + p;
        const double rk = (a.fixed && a.fixed[k]) ? 0.0 : a.b[k] - y[p];
        a.r[k] = rk;
        a.r32[k] = (float)rk;
        rr += rk * rk;
      }
    }
  }
  double vals[1] = {rr};
  store_block_partials<1>(vals, a.partials);
}

__global__ void ir_accumulate_kernel(int64_t n, const float* d, double* x, int first) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) x[k] = first ? (double)d[k] : x[k] + (double)d[k];
}

// ---------------------------------------------------------------------------
// Pipelined block-Jacobi PCG: one launch per iteration.
//
// CG needs two grid-wide reductions per iteration (p.Ap, then r.z), so the
// classic form above runs two dependent launches. The pipelined recurrences
// of Ghysels & Vanroose (Parallel Computing 40, 2014), with the block-Jacobi
// inverse B^-1 folded into the matrix as Ah = A B^-1 (block columns scaled
// once per Newton system), need one reduction phase of three dot products:
//   u = B^-1 r (never stored), w = Ah r = A u,
//   gamma = r.u, delta = w.u, beta = gamma / gamma_prev,
//   alpha = gamma / (delta - beta gamma / alpha_prev)   (= gamma / p.Ap),
//   t = Ah w + beta t   (= A B^-1 s),  s = w + beta s   (= A p),
//   p = u + beta p,  x += alpha p,  r -= alpha s,  w -= alpha t.
// Launch k reduces (gamma_k, delta_k, r.r) written by launch k - 1, forms
// alpha_k and beta_k, multiplies Ah w_k (w is double-buffered: the product
// gathers w_k at neighbouring rows that other workgroups update in the same
// launch), updates its own rows and writes the partials of iteration k + 1.
// Exact arithmetic gives the classic iterates; the stopping test is on the
// recurrence residual as in the classic form.
// ---------------------------------------------------------------------------
constexpr int kPipeBlock = 512;                   // 8 waves: fewer partials for every prologue to sum
constexpr int kPipeWaves = kPipeBlock / kWave;

template <typename T>
struct PipeArgs {
  int64_t nv;
  SellView sv;
  const T* vhat;       // Ah = A B^-1, the SELL layout of the matrix
  const T* binv;       // fixed-DOF-masked B^-1, 9 per vertex
  const uint8_t* fixed;
  const T* b;          // right-hand side (init)
  T *x, *r, *s, *t, *p;
  const T* win;        // w_k
  T* wout;             // w_k+1
  PcgState* st;
  double* ab;          // alpha of launch k at ab[k & 1]
  double rtol2;
  const double* pin;   // [3][np] gamma, delta, r.r partials of the previous launch
  double* pout;        // [3][gridDim.x] this launch's
  int np;
  int k;
  int restart;         // first iteration after a residual replacement (tolerance from this residual)
  PcgState* hst;
};

// Sum over a kPipeBlock-thread block of NV values (fixed order: wave trees,
// then a tree over the wave sums in wave 0), result in every thread. The second
// stage keeps one value per lane rather than all 16 x NV in one thread.
template <int NV>
__device__ __forceinline__ void pipe_block_sum(double (&v)[NV]) {
  __shared__ double lds[NV * kPipeWaves];
  __shared__ double res[NV];
  const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
#pragma unroll
  for (int k = 0; k < NV; ++k) v[k] = wave_sum(v[k]);
  __syncthreads();
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) lds[k * kPipeWaves + wid] = v[k];
  }
  __syncthreads();
  if (wid == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      const double t = wave_sum(lane < kPipeWaves ? lds[k * kPipeWaves + lane] : 0.0);
      if (lane == 0) res[k] = t;
    }
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; ++k) v[k] = res[k];
}

// The wave's chunk entry (chunk, slots, offset) in the XCD-balanced order; chunk -1: none.
__device__ __forceinline__ int4 pipe_chunk(const SellView& sv) {
  const int wid = threadIdx.x / kWave;
  const int64_t pos = (int64_t)(blockIdx.x / 8) * kPipeWaves + wid;
  int4 e = sv.cmeta[(blockIdx.x % 8) * sv.lpad + (pos < sv.lpad ? pos : sv.lpad - 1)];
  if (pos >= sv.lpad) e.x = -1;
  return e;
}

// y = (Ah v)_row for the lane's row, v(w) read through `get`; summed over the
// row's kSellLanes lanes and returned to all of them.
template <typename T, typename G>
__device__ __forceinline__ void pipe_row_product(const SellView& sv, const T* __restrict__ vals, int4 e, int i, int j,
                                                 G get, T (&y)[3]) {
  const int32_t deg = e.y;
  const int64_t off = e.z;
#pragma unroll 1
  for (int32_t sl = j; sl < deg; sl += kSellLanes) {
    const int64_t g = off + (int64_t)sl * kSellC;
    const int64_t w = 3 * (int64_t)sv.scol[g + i];
    T v[3];
    get(w, v);
    T A[9];
    load_slot<T, kSellC, 1>(vals, g, i, A);
#pragma unroll
    for (int q = 0; q < 3; ++q) y[q] += A[3 * q] * v[0] + A[3 * q + 1] * v[1] + A[3 * q + 2] * v[2];
  }
#pragma unroll
  for (int o = kSellC; o < kWave; o <<= 1) {
#pragma unroll
    for (int q = 0; q < 3; ++q) y[q] += __shfl_xor(y[q], o, kWave);
  }
}

template <typename T>
__device__ __forceinline__ void pipe_apply_binv(const T* __restrict__ binv, int64_t v, const T (&r)[3], T (&u)[3]) {
  T B[9];
  binv_load(binv, v, B);
#pragma unroll
  for (int q = 0; q < 3; ++q) u[q] = B[3 * q] * r[0] + B[3 * q + 1] * r[1] + B[3 * q + 2] * r[2];
}

template <typename T>
__device__ __forceinline__ void pipe_store_partials(const PipeArgs<T>& a, double (&v)[3]) {
  pipe_block_sum<3>(v);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < 3; ++q) a.pout[q * gridDim.x + blockIdx.x] = v[q];
  }
}

// Start (or restart after a residual replacement) of a pipelined run: r = b
// masked (init) or the replaced residual already in r (restart), w = Ah r,
// partials of gamma = r.B^-1 r, delta = w.B^-1 r and r.r. The init also
// zeroes x, s, t, p; a restart keeps s, t, p (the search direction) and
// zeroes x (the correction accumulator).
template <typename T>
__global__ __launch_bounds__(kPipeBlock) void ppcg_start_kernel(PipeArgs<T> a) {
  const int lane = threadIdx.x % kWave, i = lane % kSellC, j = lane / kSellC;
  const int4 e = pipe_chunk(a.sv);
  double acc[3] = {0.0, 0.0, 0.0};
  if (e.x >= 0) {
    const T* __restrict__ src = a.restart ? a.r : a.b;
    const uint8_t* __restrict__ fx = a.fixed;
    T y[3] = {T(0), T(0), T(0)};
    pipe_row_product<T>(a.sv, a.vhat, e, i, j, [&](int64_t w, T (&v)[3]) {
#pragma unroll
      for (int q = 0; q < 3; ++q) v[q] = (fx && fx[w + q]) ? T(0) : src[w + q];
    }, y);
    const int32_t row = a.sv.sell_row[(int64_t)e.x * kSellC + i];
    if (row >= 0 && j == 0) {
      const int64_t v0 = 3 * (int64_t)row;
      T r[3], u[3];
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        const bool f = fx && fx[v0 + q];
        r[q] = f ? T(0) : src[v0 + q];
        y[q] = f ? T(0) : y[q];
      }
      pipe_apply_binv(a.binv, row, r, u);
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        const int64_t k = v0 + q;
        a.r[k] = r[q];
        a.wout[k] = y[q];
        a.x[k] = T(0);
        if (!a.restart) {
          a.s[k] = T(0);
          a.t[k] = T(0);
          a.p[k] = T(0);
        }
        acc[0] += (double)r[q] * (double)u[q];
        acc[1] += (double)y[q] * (double)u[q];
        acc[2] += (double)r[q] * (double)r[q];
      }
    }
  }
  pipe_store_partials(a, acc);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    a.st->done = 0;
    if (!a.restart) a.st->iter = 0;
    if (a.hst) {
      a.hst->done = 0;
      if (!a.restart) a.hst->iter = 0;
    }
  }
}

// Iteration k of a pipelined run (see above).
template <typename T>
__global__ __launch_bounds__(kPipeBlock, 8) void ppcg_iter_kernel(PipeArgs<T> a) {
  __shared__ int s_done;
  __shared__ double s_red[3 * kPipeWaves];
  __shared__ double s_ab[3];
  const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave, i = lane % kSellC, j = lane / kSellC;
  // prologue loads: the stop flag, this thread's partials, alpha and gamma of launch k - 1
  const int done0 = a.st->done;
  double part[3] = {0.0, 0.0, 0.0};
  for (int q = (int)threadIdx.x; q < a.np; q += kPipeBlock) {
    part[0] += a.pin[q];
    part[1] += a.pin[a.np + q];
    part[2] += a.pin[2 * a.np + q];
  }
  const double alpha_prev = a.k > 0 ? a.ab[(a.k - 1) & 1] : 1.0;
  const double gamma_prev = a.k > 0 ? a.st->rz[(a.k - 1) & 1] : 1.0;
  const int4 e = pipe_chunk(a.sv);
  if (threadIdx.x == 0) s_done = done0;  // thread 0's read decides for the block
  // fixed-order block reduction of the partials (as pipe_block_sum)
#pragma unroll
  for (int q = 0; q < 3; ++q) part[q] = wave_sum(part[q]);
  if (lane == 0) {
#pragma unroll
    for (int q = 0; q < 3; ++q) s_red[q * kPipeWaves + wid] = part[q];
  }
  __syncthreads();
  if (s_done != 0) return;
  if (wid == 0) {
    double tot[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) tot[q] = wave_sum(lane < kPipeWaves ? s_red[q * kPipeWaves + lane] : 0.0);
    if (lane == 0) {
      const double g = tot[0], d = tot[1], rr = tot[2];
      const bool fresh = a.k == 0 || a.restart;
      const double tol2 = fresh ? a.rtol2 * rr : a.st->tol2;
      int dn = 0;
      double alpha = 0.0, beta = 0.0;
      if (!(isfinite(g) && isfinite(d) && isfinite(rr))) {
        dn = 3;
      } else if (!(rr > tol2)) {
        dn = 1;
      } else {
        beta = a.k > 0 ? g / gamma_prev : 0.0;
        const double den = a.k > 0 ? d - beta * g / alpha_prev : d;  // p.Ap
        if (!(den > 0.0)) dn = isfinite(den) ? 2 : 3;
        else alpha = g / den;
      }
      s_ab[0] = alpha;
      s_ab[1] = beta;
      s_ab[2] = dn;
      if (blockIdx.x == 0) {
        a.st->rz[a.k & 1] = g;
        a.ab[a.k & 1] = alpha;
        a.st->rr = rr;
        a.st->pAp = alpha > 0.0 ? g / alpha : d;
        if (fresh) a.st->tol2 = tol2;
        if (a.k == 0) a.st->bb = rr;
        if (dn) {
          a.st->done = dn;
          if (a.hst) a.hst->done = dn;
        } else {
          a.st->iter = a.k + 1;
          if (a.hst) a.hst->iter = a.k + 1;
        }
      }
    }
  }
  __syncthreads();
  if (s_ab[2] != 0.0) return;
  const T alpha = (T)s_ab[0], beta = (T)s_ab[1];
  double acc[3] = {0.0, 0.0, 0.0};
  if (e.x >= 0) {
    const T* __restrict__ win = a.win;
    T y[3] = {T(0), T(0), T(0)};
    pipe_row_product<T>(a.sv, a.vhat, e, i, j, [&](int64_t w, T (&v)[3]) {
#pragma unroll
      for (int q = 0; q < 3; ++q) v[q] = win[w + q];
    }, y);
    const int32_t row = a.sv.sell_row[(int64_t)e.x * kSellC + i];
    if (row >= 0 && j == 0) {
      const int64_t v0 = 3 * (int64_t)row;
      T r[3], u[3];
#pragma unroll
      for (int q = 0; q < 3; ++q) r[q] = a.r[v0 + q];
      pipe_apply_binv(a.binv, row, r, u);
      T rn[3], wn[3];
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        const int64_t k = v0 + q;
        const bool f = a.fixed && a.fixed[k];
        const T tq = (f ? T(0) : y[q]) + beta * a.t[k];
        const T wq = win[k];
        const T sq = wq + beta * a.s[k];
        const T pq = u[q] + beta * a.p[k];
        a.t[k] = tq;
        a.s[k] = sq;
        a.p[k] = pq;
        a.x[k] += alpha * pq;
        rn[q] = r[q] - alpha * sq;
        wn[q] = wq - alpha * tq;
        a.r[k] = rn[q];
        a.wout[k] = wn[q];
      }
      T un[3];
      pipe_apply_binv(a.binv, row, rn, un);
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        acc[0] += (double)rn[q] * (double)un[q];
        acc[1] += (double)wn[q] * (double)un[q];
        acc[2] += (double)rn[q] * (double)rn[q];
      }
    }
  }
  pipe_store_partials(a, acc);
}

// Ah = A B^-1 per slot: block (row v, column w) times the masked inverse of
// w's diagonal block, in the SELL layout of A (padding slots stay zero).
template <typename T>
__global__ __launch_bounds__(kBlock) void precond_scale_kernel(int64_t nslot, const int32_t* __restrict__ scol,
                                                              const T* __restrict__ A, const T* __restrict__ binv,
                                                              T* __restrict__ Ah) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nslot) return;
  const int64_t g = q & ~(int64_t)(kSellC - 1);
  const int i = (int)(q - g);
  const int32_t w = scol[q];
  T a[9];
  load_slot<T, kSellC, 1>(A, g, i, a);
  T B[9];
  binv_load(binv, w >= 0 ? (int64_t)w : 0, B);
  if (w < 0) {
#pragma unroll
    for (int m = 0; m < 9; ++m) B[m] = T(0);
  }
  T* o = Ah + 9 * g + i;
#pragma unroll
  for (int pq = 0; pq < 3; ++pq)
#pragma unroll
    for (int qq = 0; qq < 3; ++qq)
      o[(3 * pq + qq) * kSellC] = a[3 * pq] * B[qq] + a[3 * pq + 1] * B[3 + qq] + a[3 * pq + 2] * B[6 + qq];
}

// ---------------------------------------------------------------------------
// Generic CSR operator with scalar Jacobi preconditioner.
// ---------------------------------------------------------------------------
struct CsrPcgArgs {
  int64_t n;
  const int32_t* indptr;
  const int32_t* indices;
  const double* vals;
  double* dinv;
  const double* b;
  double* x;
  double* r;
  double* z;
  const double* pin;
  double* pout;
  double* Ap;
  PcgState* st;
  double rtol2;
  double* upart;
  int gu;
  double* spart;
  int gs;
  int k;
};

__global__ __launch_bounds__(kBlock) void cpcg_init_kernel(CsrPcgArgs a) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double rz = 0.0, rr = 0.0;
  if (i < a.n) {
    double d = 0.0;
    for (int32_t k = a.indptr[i]; k < a.indptr[i + 1]; ++k)
      if (a.indices[k] == i) d += a.vals[k];
    const double di = (d != 0.0) ? 1.0 / d : 1.0;
    a.dinv[i] = di;
    const double r = a.b[i];
    const double z = di * r;
    a.r[i] = r;
    a.z[i] = z;
    a.x[i] = 0.0;
    a.pout[i] = 0.0;
    rz = r * z;
    rr = r * r;
  }
  double vals[2] = {rz, rr};
  store_block_partials<2>(vals, a.upart);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    a.st->done = 0;
    a.st->iter = 0;
  }
}

__global__ __launch_bounds__(kBlock) void cpcg_spmv_kernel(CsrPcgArgs a) {
  const PcgCtl c = spmv_prologue(a.st, a.upart, a.gu, a.k, a.rtol2);
  if (!c.run) return;
  const double beta = c.beta;
  constexpr int rows_per_pass = kBlock / kRowLanes;
  const int64_t chunk = (a.n + gridDim.x - 1) / gridDim.x;
  const int64_t i_begin = blockIdx.x * chunk;
  const int64_t i_end = i_begin + chunk < a.n ? i_begin + chunk : a.n;
  const int lane = threadIdx.x % kRowLanes;
  double dot = 0.0;
  for (int64_t i0 = i_begin; i0 < i_end; i0 += rows_per_pass) {
    const int64_t i = i0 + threadIdx.x / kRowLanes;
    double acc = 0.0;
    if (i < i_end) {
      for (int32_t k = a.indptr[i] + lane; k < a.indptr[i + 1]; k += kRowLanes) {
        const int64_t j = a.indices[k];
        acc += a.vals[k] * (a.z[j] + beta * a.pin[j]);
      }
    }
#pragma unroll
    for (int off = kRowLanes / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, kRowLanes);
    if (i < i_end && lane == 0) {
      const double pn = a.z[i] + beta * a.pin[i];
      a.pout[i] = pn;
      a.Ap[i] = acc;
      dot += pn * acc;
    }
  }
  double vals[1] = {dot};
  store_block_partials<1>(vals, a.spart);
}

__global__ __launch_bounds__(kBlock) void cpcg_update_kernel(CsrPcgArgs a) {
  double alpha;
  if (!update_prologue(a.st, a.spart, a.gs, a.k, alpha)) return;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double rz = 0.0, rr = 0.0;
  if (i < a.n) {
    a.x[i] += alpha * a.pout[i];
    const double r = a.r[i] - alpha * a.Ap[i];
    const double z = a.dinv[i] * r;
    a.r[i] = r;
    a.z[i] = z;
    rz = r * z;
    rr = r * r;
  }
  double vals[2] = {rz, rr};
  store_block_partials<2>(vals, a.upart);
}

// ---------------------------------------------------------------------------
// Backward-Euler helpers
// ---------------------------------------------------------------------------
// Caller's index of internal DOF k: with a vertex order (internal vertex t is
// the caller's vertex order[t]) states cross the boundary permuted here, in
// the kernels that read and write them.
__device__ __forceinline__ int64_t caller_dof(const int32_t* order, int64_t k) {
  return order ? 3 * (int64_t)order[k / 3] + k % 3 : k;
}

// x~ = x_c + h ((x_c - x_p) / h) (kinetic.py:87-90); x0 = x~ on free DOFs,
// prescribed (or current) values on fixed DOFs.
__global__ void be_init_kernel(int64_t n, const double* xc, const double* xp, const double* xbc,
                               const uint8_t* fixed, double h, const int32_t* order, double* x, double* xt) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int64_t q = caller_dof(order, k);
  const double c = xc[q];
  const double t = c + h * ((c - xp[q]) / h);
  xt[k] = t;
  x[k] = (fixed && fixed[k]) ? (xbc ? xbc[q] : c) : t;
}

// BDF2 inertial target (kinetic.py bdf2_target / velocity_bdf2, in the
// reference's order of operations):
//   v_c = (3 x_c - 4 x_p + x_p2) / (2h),  v_p = (3 x_p - 4 x_p2 + x_p3) / (2h)
//   x~ = (4/3) x_c - (1/3) x_p + (8h/9) v_c - (2h/9) v_p.
__global__ void bdf2_init_kernel(int64_t n, const double* xc, const double* xp, const double* xp2, const double* xp3,
                                 const double* xbc, const uint8_t* fixed, double h, const int32_t* order, double* x,
                                 double* xt) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int64_t q = caller_dof(order, k);
  const double c = xc[q], p = xp[q], p2 = xp2[q], p3 = xp3[q];
  const double vc = (3.0 * c - 4.0 * p + p2) / (2.0 * h);
  const double vp = (3.0 * p - 4.0 * p2 + p3) / (2.0 * h);
  const double t = (4.0 / 3.0) * c - (1.0 / 3.0) * p + (8.0 * h / 9.0) * vc - (2.0 * h / 9.0) * vp;
  xt[k] = t;
  x[k] = (fixed && fixed[k]) ? (xbc ? xbc[q] : c) : t;
}

// dx = guess - x on free DOFs, 0 on fixed DOFs (which keep their prescribed
// values whichever start point is taken).
__global__ void guess_dx_kernel(int64_t n, const double* guess, const double* x, const uint8_t* fixed,
                                const int32_t* order, double* dx) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  dx[k] = (fixed && fixed[k]) ? 0.0 : guess[caller_dof(order, k)] - x[k];
}

// out (caller's numbering) = x (internal numbering).
__global__ void store_state_kernel(int64_t n, const double* __restrict__ x, const int32_t* order,
                                   double* __restrict__ out) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  out[caller_dof(order, k)] = x[k];
}

// Largest of each of NV values over a block (fixed tree), result in thread 0.
template <int NV>
__device__ __forceinline__ void block_max(double (&v)[NV]) {
  __shared__ double lds[NV * kBlock / kWave];
  const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
#pragma unroll
  for (int k = 0; k < NV; ++k)
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) v[k] = fmax(v[k], __shfl_down(v[k], off, kWave));
  __syncthreads();
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < NV; ++k) lds[k * (kBlock / kWave) + wid] = v[k];
  __syncthreads();
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      double m = 0.0;
      for (int w = 0; w < kBlock / kWave; ++w) m = fmax(m, lds[k * (kBlock / kWave) + w]);
      v[k] = m;
    }
}

// x += t dx; per-block maxima of |t dx| and |x| (after the update) over the
// free DOFs into partials[k * gridDim.x + block].
__global__ __launch_bounds__(kBlock) void update_absmax_kernel(int64_t n, double t, const double* dx,
                                                               const uint8_t* fixed, double* x, double* partials) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double v[2] = {0.0, 0.0};
  if (k < n) {
    const double step = t * dx[k];
    const double xn = x[k] + step;
    x[k] = xn;
    if (!(fixed && fixed[k])) {
      v[0] = fabs(step);
      v[1] = fabs(xn);
    }
  }
  block_max<2>(v);
  if (threadIdx.x == 0) {
    partials[blockIdx.x] = v[0];
    partials[gridDim.x + blockIdx.x] = v[1];
  }
}

__global__ __launch_bounds__(kBlock) void finalize_max2_kernel(const double* partials, int n, double* out) {
  double v[2] = {0.0, 0.0};
  for (int i = threadIdx.x; i < n; i += kBlock) {
    v[0] = fmax(v[0], partials[i]);
    v[1] = fmax(v[1], partials[n + i]);
  }
  block_max<2>(v);
  if (threadIdx.x == 0) {
    out[0] = v[0];
    out[1] = v[1];
  }
}

__global__ void axpy_kernel(int64_t n, double t, const double* dx, double* x) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) x[k] = x[k] + t * dx[k];
}

__global__ __launch_bounds__(kBlock) void dot_kernel(int64_t n, const double* a, const double* b, double* partials) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double v[1] = {k < n ? a[k] * b[k] : 0.0};
  store_block_partials<1>(v, partials);
}

// ---------------------------------------------------------------------------
// Generic CSR PCG on caller-provided scratch
// ---------------------------------------------------------------------------
namespace {
constexpr int kPcgChunk = 16;  // iterations enqueued between the generic PCG's convergence checks

int spmv_grid(int64_t rows) {
  const int64_t g = (rows + (kBlock / kRowLanes) - 1) / (kBlock / kRowLanes);
  return (int)(g < kSpmvBlocks ? (g < 1 ? 1 : g) : kSpmvBlocks);
}

constexpr size_t kStateBytes = 64;  // PcgState, padded so the doubles after it stay 64-byte aligned
static_assert(sizeof(PcgState) <= kStateBytes, "PcgState outgrew its scratch slot");
}  // namespace

// Scratch layout: the PcgState, then dinv, r, z, two search directions and
// A p (n doubles each), then the reduction partials (2 gu + gs doubles).
size_t csr_pcg_scratch_bytes(int64_t n) {
  return kStateBytes + sizeof(double) * (6 * (size_t)n + 2 * (size_t)grid_for(n) + spmv_grid(n));
}

// Every scratch entry the kernels read is written first by the init launch,
// so the scratch need not be cleared. At most maxit iterations are launched.
int run_csr_pcg(int64_t n, const int32_t* indptr, const int32_t* indices, const double* values, const double* b,
                double* x, double rtol, int64_t maxit, int64_t* iters, double* relres, void* work, hipStream_t st) {
  const int gu = grid_for(n);
  const int gs = spmv_grid(n);
  PcgState* state = reinterpret_cast<PcgState*>(work);
  double* dinv = reinterpret_cast<double*>(reinterpret_cast<char*>(work) + kStateBytes);
  double* r = dinv + n;
  double* z = r + n;
  double* p0 = z + n;
  double* p1 = p0 + n;
  double* Ap = p1 + n;
  double* partials = Ap + n;
  CsrPcgArgs a{n, indptr, indices, values, dinv, b, x, r, z, p1, p0, Ap, state, rtol * rtol,
               partials, gu, partials + 2 * gu, gs, 0};
  cpcg_init_kernel<<<gu, kBlock, 0, st>>>(a);
  SK_HIP_CHECK(hipGetLastError());
  PcgState hs{};
  int64_t launched = 0;
  for (;;) {
    for (int k = 0; k < kPcgChunk && launched < maxit; ++k, ++launched) {
      a.k = (int)launched;
      a.pin = (launched & 1) ? p1 : p0;
      a.pout = (launched & 1) ? p0 : p1;
      cpcg_spmv_kernel<<<gs, kBlock, 0, st>>>(a);
      cpcg_update_kernel<<<gu, kBlock, 0, st>>>(a);
    }
    SK_HIP_CHECK(hipGetLastError());
    SK_HIP_CHECK(hipMemcpyAsync(&hs, state, sizeof(PcgState), hipMemcpyDeviceToHost, st));
    SK_HIP_CHECK(hipStreamSynchronize(st));
    if (hs.done || launched >= maxit) break;
  }
  if (iters) *iters = hs.iter;
  if (relres) *relres = hs.bb > 0.0 ? std::sqrt(hs.rr / hs.bb) : 0.0;
  if (hs.done == 2) {
    set_error("sk_pcg_csr: p^T A p <= 0 (matrix not positive definite)");
    return SK_ERR_PCG_BREAKDOWN;
  }
  if (hs.done == 3) {
    set_error("sk_pcg_csr: non-finite residual");
    return SK_ERR_NONFINITE;
  }
  if (hs.done == 0) {
    set_error("sk_pcg_csr: iteration limit reached before rtol");
    return SK_ERR_PCG_MAXIT;
  }
  return SK_OK;
}

}  // namespace sk

using namespace sk;

struct sk_sim {
  sk_mesh* mesh = nullptr;
  sk_sim_params prm{};
  int spmv_blocks = 0;
  int update_blocks = 0;
  DevBuf<double> mu, lam, vol, mass, fext;
  DevBuf<uint8_t> fixed;
  DevBuf<double> x, xt, g, dx, r, z, p0, p1, Ap, vals, binv;
  // mixed precision: FP32 matrix, preconditioner and inner-PCG vectors
  DevBuf<float> vals32, binv32, x32, r32, z32, p0_32, p1_32, Ap32;
  DevBuf<PcgState> st;
  PcgState* st_host = nullptr;  // pinned, coherent: the PCG kernels write iter and done through to it
  DevBuf<unsigned long long> fp64_count;  // element Hessians that took the FP64 polar branch this step
  DevBuf<int32_t> order;  // optional: internal vertex t is the caller's vertex order[t] (sk_sim_set_order)
  // pipelined PCG (pcg_variant 1): Ah = A B^-1 in the matrix's layout, the p vector, alpha of the last two launches
  DevBuf<double> vhat, pp;
  DevBuf<float> vhat32, pp32;
  DevBuf<double> ab;
  hipEvent_t chunk_ev[2] = {};
  hipEvent_t ev[6] = {};
  bool have_events = false;
  sk::CoarseSpace coarse;  // two-level preconditioner (coarse.n_agg > 0 when enabled)
  bool mixed() const { return prm.precision == SK_PREC_MIXED; }
  bool two_level() const { return coarse.n_agg > 0; }
  // The mixed path keeps the classic form: its FP64 residual replacement breaks the pipelined
  // recurrences' identity p.Ap = delta - beta gamma / alpha_prev in FP32 (the replaced residual is
  // no longer r_prev - alpha s_prev to the recurrences' accuracy; a numpy emulation of the FP32 run
  // breaks down at the second replacement).
  bool pipelined() const { return prm.pcg_variant == 1 && !two_level() && !mixed(); }
};

namespace {

hipStream_t as_stream(sk_stream_t s) { return reinterpret_cast<hipStream_t>(s); }

int copy_dev(DevBuf<double>& dst, const double* src, size_t n) {
  SK_TRY(dst.alloc(n));
  SK_HIP_CHECK(hipMemcpy(dst.p, src, n * sizeof(double), hipMemcpyDeviceToDevice));
  return SK_OK;
}

// PCG iterations enqueued per convergence check: past convergence up to about
// 1.5 chunks of launches return in their prologue, so short solves (the
// two-level preconditioner's) use shorter chunks.
// Defaults measured at 48^3 (PCG time per iteration against 2..16): the
// two-level launches past convergence cost a full coarse product, so its
// chunks are shorter, and shortest with the mixed path's short refinement runs.
int pcg_chunk(const sk_sim* S) {
  if (S->prm.pcg_chunk > 0) return S->prm.pcg_chunk;
  if (S->two_level()) return S->mixed() ? 2 : 4;
  return 8;
}

// One SELL chunk per wave; the grid is rounded up to a multiple of 8 so the
// XCD-aware block order is a permutation (surplus waves do no work).
int sell_grid(const sk_mesh* m) { return (int)(8 * (m->lpad / (kBlock / kWave))); }

SellView sell_view(sk_mesh* m) {
  return SellView{m->nchunk, m->sell_row.p, m->chunk_deg.p, m->chunk_off.p, m->scol.p, m->corder.p, m->lpad,
                  m->cmeta.p};
}

// Buffers of one precision's PCG.
template <typename T>
struct PcgBufs {
  const T* vals;
  const T* binv;
  T *x, *r, *z, *p0, *p1, *Ap;
};

PcgBufs<double> bufs64(sk_sim* S) {
  return {S->vals.p, S->binv.p, S->dx.p, S->r.p, S->z.p, S->p0.p, S->p1.p, S->Ap.p};
}
PcgBufs<float> bufs32(sk_sim* S) {
  return {S->vals32.p, S->binv32.p, S->x32.p, S->r32.p, S->z32.p, S->p0_32.p, S->p1_32.p, S->Ap32.p};
}

template <typename T>
BlockPcgArgs<T> block_pcg_args(sk_sim* S, const PcgBufs<T>& B, const T* b, double rtol2) {
  sk_mesh* m = S->mesh;
  // Launch widths (measured on the 48^3 cube): the vertex-wise kernels run one
  // lane per vertex unless update_blocks caps them (they stride over the
  // vertices); the SpMV runs one chunk per wave (FP32 14.4 us against 16.1 us
  // with 1024 workgroups walking chunk ranges), unless spmv_blocks caps it.
  // With the two-level preconditioner the vertex-wise kernels run one
  // workgroup per aggregate.
  const int gu = S->two_level() ? S->coarse.n_agg
                 : S->update_blocks > 0 ? std::min(grid_for(m->nv), S->update_blocks)
                                        : grid_for(m->nv);
  const int cap = S->spmv_blocks;
  const int gs = cap > 0 ? std::min(sell_grid(m), (cap + 7) / 8 * 8) : sell_grid(m);
  double* part = m->red.partials.p;
  CoarseView<T> cv{};
  if (S->two_level()) cv = S->coarse.view<T>();
  return BlockPcgArgs<T>{m->nv, sell_view(m), B.vals, B.binv, S->fixed.p, b, B.x, B.r, B.z, B.p1, B.p0, B.Ap,
                         S->st.p, rtol2, part, gu, part + 2 * gu, gs, 0, 0, cv, S->st_host};
}

template <typename T>
void launch_spmv(const BlockPcgArgs<T>& a, const sk_mesh* m, hipStream_t st) {
  if (a.gs >= sell_grid(m)) bpcg_spmv_kernel<T, false><<<a.gs, kBlock, 0, st>>>(a);
  else bpcg_spmv_kernel<T, true><<<a.gs, kBlock, 0, st>>>(a);
}

// Block-Jacobi (or two-level) PCG on the assembled system with right-hand side b (x0 = 0).
// With k0 > 0 it instead continues an earlier run after a residual
// replacement: B.r holds the replaced residual, iterations continue from
// index k0 with the kept search direction, and x accumulates the correction
// from 0. Stops at ||r|| <= rtol ||r(start of this run)||; *iters is the
// total iteration index reached. iters = nullptr: returns once the last
// launches are enqueued, without waiting for them; the caller reads the
// result with pcg_result after its next synchronization of the stream.
template <typename T>
int run_block_pcg(sk_sim* S, const PcgBufs<T>& B, const T* b, double rtol, int64_t max_iter, hipStream_t st,
                  int64_t* iters, int* done, int64_t k0 = 0) {
  // The coarse operator is formed from the first system solved and kept.
  if (S->two_level() && !S->coarse.built)
    SK_TRY(coarse_build(S->coarse, S->mesh, B.vals, sizeof(T) == 4, S->fixed.p, st));
  BlockPcgArgs<T> a = block_pcg_args<T>(S, B, b, rtol * rtol);
  SK_REQUIRE(2 * a.gu + a.gs <= 8 * Reducer::kMaxBlocks, "mesh too large for the PCG reductions");
  const bool tl = S->two_level();
  if (k0 == 0) {
    a.pout = B.p0;  // init zeroes p0, the first "previous p"
    if (tl) bpcg_init_kernel<T, true><<<a.gu, kBlock, 0, st>>>(a);
    else bpcg_init_kernel<T, false><<<a.gu, kBlock, 0, st>>>(a);
  } else {
    if (tl) bpcg_restart_kernel<T, true><<<a.gu, kBlock, 0, st>>>(a);
    else bpcg_restart_kernel<T, false><<<a.gu, kBlock, 0, st>>>(a);
  }
  if (tl) bpcg_coarse_kernel<T><<<a.gu, kBlock, 0, st>>>(a);
  SK_HIP_CHECK(hipGetLastError());
  const int chunk = pcg_chunk(S);
  // Chunks of `chunk` iterations; the state after chunk c is checked only once
  // chunk c + 1 has been enqueued, so the GPU is not idle while the host reads
  // it. The kernels write iter and done through to the pinned mirror, so the
  // check needs no copy in the stream; it may see a later state than chunk
  // c's end, which only stops sooner. Launches past convergence exit in their
  // prologue.
  const volatile PcgState* h = S->st_host;
  int64_t launched = k0;
  int c = 0;
  for (bool stop = false; !stop; ++c) {
    for (int k = 0; k < chunk && launched < max_iter; ++k, ++launched) {
      a.k = (int)launched;
      a.restart = (k0 > 0 && launched == k0) ? 1 : 0;
      a.pin = (launched & 1) ? B.p1 : B.p0;
      a.pout = (launched & 1) ? B.p0 : B.p1;
      launch_spmv<T>(a, S->mesh, st);
      if (tl) {
        bpcg_update_kernel<T, true><<<a.gu, kBlock, 0, st>>>(a);
        bpcg_coarse_kernel<T><<<a.gu, kBlock, 0, st>>>(a);
      } else {
        bpcg_update_kernel<T, false><<<a.gu, kBlock, 0, st>>>(a);
      }
    }
    SK_HIP_CHECK(hipGetLastError());
    SK_HIP_CHECK(hipEventRecord(S->chunk_ev[c & 1], st));
    if (c > 0) {
      SK_HIP_CHECK(hipEventSynchronize(S->chunk_ev[(c - 1) & 1]));
      const int hd = h->done, hi = h->iter;
      if (hd || hi >= max_iter) stop = true;
      if (k0 > 0 && hi < k0) stop = true;  // continuation did not start
    }
    if (launched >= max_iter) stop = true;  // no launches past the iteration limit
  }
  if (!iters) return SK_OK;
  SK_HIP_CHECK(hipEventSynchronize(S->chunk_ev[(c - 1) & 1]));
  *iters = h->iter;
  *done = h->done;
  return SK_OK;
}

// Pipelined block-Jacobi PCG (ppcg_*_kernel) on the system whose Ah = A B^-1
// the caller has formed (precond_scale_kernel), with the same contract as
// run_block_pcg: x0 = 0 (k0 = 0) or the continuation after a residual
// replacement (k0 > 0: r holds the replaced residual, the search direction is
// kept, x restarts from 0); iters = nullptr returns without waiting.
template <typename T>
int run_pipe_pcg(sk_sim* S, const PcgBufs<T>& B, const T* vhat, T* pvec, const T* b, double rtol,
                 int64_t max_iter, hipStream_t st, int64_t* iters, int* done, int64_t k0 = 0) {
  sk_mesh* m = S->mesh;
  const int grid = (int)(8 * ((m->lpad + kPipeWaves - 1) / kPipeWaves));
  SK_REQUIRE(6 * (int64_t)grid <= 8 * (int64_t)Reducer::kMaxBlocks, "mesh too large for the PCG reductions");
  double* part = m->red.partials.p;
  // vectors: s in z, t in Ap, w double-buffered in p0 / p1
  PipeArgs<T> a{m->nv, sell_view(m), vhat, B.binv, S->fixed.p, b, B.x, B.r, B.z, B.Ap, pvec, nullptr, nullptr,
                S->st.p, S->ab.p, rtol * rtol, nullptr, nullptr, grid, 0, 0, S->st_host};
  auto wbuf = [&](int64_t k) { return (k & 1) ? B.p1 : B.p0; };
  auto pbuf = [&](int64_t k) { return part + 3 * (size_t)grid * (k & 1); };
  a.restart = k0 > 0 ? 1 : 0;
  a.wout = wbuf(k0);
  a.pout = pbuf(k0);
  ppcg_start_kernel<T><<<grid, kPipeBlock, 0, st>>>(a);
  SK_HIP_CHECK(hipGetLastError());
  const int chunk = pcg_chunk(S);
  const volatile PcgState* h = S->st_host;
  int64_t launched = k0;
  int c = 0;
  for (bool stop = false; !stop; ++c) {
    for (int k = 0; k < chunk && launched < max_iter; ++k, ++launched) {
      a.k = (int)launched;
      a.restart = (k0 > 0 && launched == k0) ? 1 : 0;
      a.win = wbuf(launched);
      a.wout = wbuf(launched + 1);
      a.pin = pbuf(launched);
      a.pout = pbuf(launched + 1);
      ppcg_iter_kernel<T><<<grid, kPipeBlock, 0, st>>>(a);
    }
    SK_HIP_CHECK(hipGetLastError());
    SK_HIP_CHECK(hipEventRecord(S->chunk_ev[c & 1], st));
    if (c > 0) {
      SK_HIP_CHECK(hipEventSynchronize(S->chunk_ev[(c - 1) & 1]));
      const int hd = h->done, hi = h->iter;
      if (hd || hi >= max_iter) stop = true;
      if (k0 > 0 && hi < k0) stop = true;  // continuation did not start
    }
    if (launched >= max_iter) stop = true;  // no launches past the iteration limit
  }
  if (!iters) return SK_OK;
  SK_HIP_CHECK(hipEventSynchronize(S->chunk_ev[(c - 1) & 1]));
  *iters = h->iter;
  *done = h->done;
  return SK_OK;
}

// Ah = A B^-1 of the assembled system (pipelined PCG), once per Newton system.
template <typename T>
int form_precond_matrix(sk_sim* S, const T* A, const T* binv, T* Ah, hipStream_t st) {
  sk_mesh* m = S->mesh;
  precond_scale_kernel<T><<<grid_for(m->nslot), kBlock, 0, st>>>(m->nslot, m->scol.p, A, binv, Ah);
  SK_HIP_CHECK(hipGetLastError());
  return SK_OK;
}

// The result of a run_block_pcg launched without waiting, once the stream
// has been synchronized since.
void pcg_result(const sk_sim* S, int64_t* iters, int* done) {
  const volatile PcgState* h = S->st_host;
  *iters = h->iter;
  *done = h->done;
}

// Mixed-precision solve of A dx = b (b FP64 in S->g, dx FP64 in S->dx). One
// FP32 block-Jacobi PCG run on the FP32 matrix, refined in FP64: whenever the
// FP32 recurrence residual has dropped by inner_rtol, the FP32 correction is
// accumulated into the FP64 dx, the true residual r = b - A32 dx is recomputed
// in FP64 and replaces the FP32 one, and CG continues with its current search
// direction (residual replacement; the Krylov space is kept). Stops when the
// FP64 residual satisfies ||r|| <= pcg_rtol ||b||.
int run_mixed_solve(sk_sim* S, hipStream_t st, int64_t* iters, int* refines, int* done) {
  sk_mesh* m = S->mesh;
  const int64_t n3 = 3 * m->nv;
  ResidualArgs ra{m->nv, sell_view(m), S->vals32.p, S->fixed.p, S->g.p, nullptr, S->r.p, S->r32.p,
                  m->red.partials.p + 4 * (size_t)Reducer::kMaxBlocks};
  const int gr = sell_grid(m);
  double* rr_dev = m->red.result.p + 1;
  ir_residual_kernel<<<gr, kBlock, 0, st>>>(ra);
  finalize_sum_kernel<1><<<1, kBlock, 0, st>>>(ra.partials, gr, rr_dev);
  SK_TRY(read_scalars(m, rr_dev, 1, st));
  const double bb = m->pinned[0];
  const double target = S->prm.pcg_rtol * S->prm.pcg_rtol * bb;
  double rr = bb;
  *iters = 0;
  *refines = 0;
  *done = 1;
  const PcgBufs<float> B = bufs32(S);
  int64_t k = 0;
  for (int seg = 0; seg < S->prm.max_refine && rr > target && bb > 0.0; ++seg) {
    // The right-hand side may alias the residual r32: the init launch reads
    // each entry before the same thread overwrites it.
    // The run is not waited for: the refinement launches follow it in the
    // stream, and its result is read after the residual's readback (on a
    // breakdown the step fails, so the correction accumulated meanwhile is moot).
    int d = 0;
    if (S->pipelined())
      SK_TRY(run_pipe_pcg<float>(S, B, S->vhat32.p, S->pp32.p, B.r, S->prm.inner_rtol, S->prm.pcg_max_iter, st,
                                 nullptr, nullptr, seg == 0 ? 0 : k));
    else
      SK_TRY(run_block_pcg<float>(S, B, B.r, S->prm.inner_rtol, S->prm.pcg_max_iter, st, nullptr, nullptr,
                                  seg == 0 ? 0 : k));
    ++*refines;
    ir_accumulate_kernel<<<grid_for(n3), kBlock, 0, st>>>(n3, B.x, S->dx.p, seg == 0 ? 1 : 0);
    ra.x = S->dx.p;
    ir_residual_kernel<<<gr, kBlock, 0, st>>>(ra);
    finalize_sum_kernel<1><<<1, kBlock, 0, st>>>(ra.partials, gr, rr_dev);
    SK_TRY(read_scalars(m, rr_dev, 1, st));
    pcg_result(S, &k, &d);
    if (d == 2 || d == 3) {
      *iters = k;
      *done = d;
      return SK_OK;
    }
    rr = m->pinned[0];
    if (!std::isfinite(rr)) {
      *iters = k;
      *done = 3;
      return SK_OK;
    }
    if (k >= S->prm.pcg_max_iter || d != 1) break;
  }
  *iters = k;
  if (bb == 0.0) SK_HIP_CHECK(hipMemsetAsync(S->dx.p, 0, n3 * sizeof(double), st));
  *done = (rr <= target) ? 1 : 0;
  return SK_OK;
}

float elapsed(hipEvent_t a, hipEvent_t b) {
  float ms = 0.f;
  if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0.f;
  return ms;
}

}  // namespace

extern "C" {

int sk_pcg_csr(int64_t n, const int32_t* indptr, const int32_t* indices, const double* values, const double* b,
               double* x, double rtol, int64_t maxit, int64_t* iters, double* relres, sk_stream_t s) {
  SK_REQUIRE(n >= 0, "sk_pcg_csr: negative size");
  SK_REQUIRE(rtol >= 0.0 && maxit >= 0, "sk_pcg_csr: invalid rtol or maxit");
  if (iters) *iters = 0;
  if (relres) *relres = 0.0;
  if (n == 0) return SK_OK;
  SK_REQUIRE(indptr && indices && values && b && x, "sk_pcg_csr: null pointer");
  SK_REQUIRE(grid_for(n) <= Reducer::kMaxBlocks, "sk_pcg_csr: system too large");
  DevBuf<char> work;
  SK_TRY(work.alloc(csr_pcg_scratch_bytes(n)));
  return run_csr_pcg(n, indptr, indices, values, b, x, rtol, maxit, iters, relres, work.p, as_stream(s));
}

int sk_sim_create(sk_mesh* mesh, const sk_sim_params* p, const double* mu, const double* lam, const double* vol,
                  const double* mass, const double* f_ext, const uint8_t* fixed, sk_sim** out) {
  SK_REQUIRE(out, "sk_sim_create: null out");
  *out = nullptr;
  SK_REQUIRE(mesh && p && mu && vol && mass, "sk_sim_create: null argument");
  SK_REQUIRE(p->material >= SK_SNH && p->material <= SK_FCR, "sk_sim_create: unknown material");
  SK_REQUIRE(!material_needs_lam(p->material) || lam, "sk_sim_create: this material needs lam");
  SK_REQUIRE(p->h > 0.0, "sk_sim_create: h must be positive");
  SK_REQUIRE(p->ls_alpha > 0.0 && p->ls_alpha <= 0.5, "sk_sim_create: ls_alpha must be in (0, 0.5]");
  SK_REQUIRE(p->ls_beta > 0.0 && p->ls_beta < 1.0, "sk_sim_create: ls_beta must be in (0, 1)");
  SK_REQUIRE(p->psd == SK_PSD_NONE || p->psd == SK_PSD_PROJ || p->psd == SK_PSD_ABS, "sk_sim_create: bad psd");
  SK_REQUIRE(p->precision == SK_PREC_FP64 || p->precision == SK_PREC_MIXED, "sk_sim_create: bad precision");
  SK_REQUIRE(p->precision == SK_PREC_FP64 || (p->inner_rtol > 0.0 && p->inner_rtol < 1.0 && p->max_refine > 0),
             "sk_sim_create: mixed precision needs 0 < inner_rtol < 1 and max_refine > 0");
  SK_REQUIRE(p->spmv_blocks >= 0 && p->update_blocks >= 0, "sk_sim_create: launch widths must be >= 0");
  SK_REQUIRE(p->pcg_chunk >= 0 && p->pcg_chunk <= 1024, "sk_sim_create: pcg_chunk must be in [0, 1024]");
  SK_REQUIRE(p->pcg_variant == 0 || p->pcg_variant == 1, "sk_sim_create: pcg_variant must be 0 or 1");
  SK_REQUIRE(p->psd_jacobi == 0 || p->psd_jacobi == 1, "sk_sim_create: psd_jacobi must be 0 or 1");
  sk_sim* S = new sk_sim();
  S->mesh = mesh;
  S->prm = *p;
  S->spmv_blocks = p->spmv_blocks;
  S->update_blocks = p->update_blocks;
  const size_t nt = mesh->nt, n3 = 3 * mesh->nv;
  auto run = [&]() -> int {
    SK_TRY(copy_dev(S->mu, mu, nt));
    if (lam) SK_TRY(copy_dev(S->lam, lam, nt));
    SK_TRY(copy_dev(S->vol, vol, nt));
    SK_TRY(copy_dev(S->mass, mass, n3));
    if (f_ext) SK_TRY(copy_dev(S->fext, f_ext, n3));
    if (fixed) {
      SK_TRY(S->fixed.alloc(n3));
      SK_HIP_CHECK(hipMemcpy(S->fixed.p, fixed, n3, hipMemcpyDeviceToDevice));
    }
    for (DevBuf<double>* b : {&S->x, &S->xt, &S->g, &S->dx, &S->r})
      SK_TRY(b->alloc(n3));
    if (S->mixed()) {
      for (DevBuf<float>* b : {&S->x32, &S->r32, &S->z32, &S->p0_32, &S->p1_32, &S->Ap32})
        SK_TRY(b->alloc(n3));
      SK_TRY(S->vals32.alloc(9 * (size_t)mesh->nslot));
      SK_HIP_CHECK(hipMemset(S->vals32.p, 0, 9 * (size_t)mesh->nslot * sizeof(float)));
      SK_TRY(S->binv32.alloc(kBinv * (size_t)mesh->nv));
    } else {
      for (DevBuf<double>* b : {&S->z, &S->p0, &S->p1, &S->Ap})
        SK_TRY(b->alloc(n3));
      SK_TRY(S->vals.alloc(9 * (size_t)mesh->nslot));
      SK_HIP_CHECK(hipMemset(S->vals.p, 0, 9 * (size_t)mesh->nslot * sizeof(double)));
      SK_TRY(S->binv.alloc(kBinv * (size_t)mesh->nv));
    }
    if (p->pcg_variant == 1) {
      if (S->mixed()) {
        SK_TRY(S->vhat32.alloc(9 * (size_t)mesh->nslot));
        SK_TRY(S->pp32.alloc(n3));
      } else {
        SK_TRY(S->vhat.alloc(9 * (size_t)mesh->nslot));
        SK_TRY(S->pp.alloc(n3));
      }
      SK_TRY(S->ab.alloc(4));
    }
    SK_TRY(S->st.alloc(1));
    SK_TRY(S->fp64_count.alloc(1));
    SK_HIP_CHECK(hipHostMalloc((void**)&S->st_host, sizeof(PcgState), hipHostMallocCoherent));
    *S->st_host = PcgState{};
    for (auto& e : S->ev) SK_HIP_CHECK(hipEventCreate(&e));
    for (auto& e : S->chunk_ev) SK_HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    S->have_events = true;
    return SK_OK;
  };
  int s = run();
  if (s != SK_OK) {
    sk_sim_destroy(S);
    return s;
  }
  *out = S;
  return SK_OK;
}

int sk_sim_destroy(sk_sim* S) {
  if (!S) return SK_OK;
  if (S->st_host) (void)hipHostFree(S->st_host);
  if (S->have_events) {
    for (auto& e : S->ev) (void)hipEventDestroy(e);
    for (auto& e : S->chunk_ev)
      if (e) (void)hipEventDestroy(e);
  }
  delete S;
  return SK_OK;
}

int sk_sim_set_coarse(sk_sim* S, const double* X, int per_agg) {
  SK_REQUIRE(S, "sk_sim_set_coarse: null simulator");
  S->coarse = CoarseSpace();
  if (per_agg <= 0) return SK_OK;
  SK_REQUIRE(X, "sk_sim_set_coarse: null positions");
  const int64_t n3 = 3 * S->mesh->nv;
  std::vector<uint8_t> fixed;
  if (S->fixed.p) {
    fixed.resize(n3);
    SK_HIP_CHECK(hipMemcpy(fixed.data(), S->fixed.p, n3, hipMemcpyDeviceToHost));
  }
  const int s = coarse_setup(S->coarse, S->mesh, X, S->fixed.p ? fixed.data() : nullptr, per_agg);
  if (s != SK_OK) S->coarse = CoarseSpace();
  return s;
}

int sk_sim_set_order(sk_sim* S, const int32_t* order_host) {
  SK_REQUIRE(S, "sk_sim_set_order: null simulator");
  if (!order_host) {
    S->order.release();
    return SK_OK;
  }
  const int64_t nv = S->mesh->nv;
  std::vector<uint8_t> seen(nv, 0);
  for (int64_t t = 0; t < nv; ++t) {
    const int32_t v = order_host[t];
    SK_REQUIRE(v >= 0 && v < nv && !seen[v], "sk_sim_set_order: order is not a permutation of the vertices");
    seen[v] = 1;
  }
  std::vector<int32_t> h(order_host, order_host + nv);
  return S->order.upload(h);
}

int sk_sim_coarse_info(const sk_sim* S, int* n_agg, int* nc, int* built, double* build_ms) {
  SK_REQUIRE(S, "sk_sim_coarse_info: null simulator");
  if (n_agg) *n_agg = S->coarse.n_agg;
  if (nc) *nc = S->coarse.nc;
  if (built) *built = S->coarse.built ? 1 : 0;
  if (build_ms) *build_ms = S->coarse.build_ms;
  return SK_OK;
}

}  // extern "C"

namespace {

// Newton on E(x) = V(x) + c/(2h^2) |x - x~|_M^2 - f_ext.x for one implicit
// step; S->x and S->xt hold the start point and x~ (written by the caller's
// init kernel). With n_guess candidate start points (device arrays in the
// caller's numbering), Newton starts from the one with the lowest incremental
// potential if that is below E(x~) (E(guess) - E(x~), FP64, difference form;
// ties go to the earlier candidate); the stopping test stays relative to the
// gradient norm at x~, so every start converges to the same absolute tolerance.
int step_core(sk_sim* S, double c_mass, const double* const* guesses, int n_guess, double* x_next,
              sk_step_info* info, hipStream_t st) {
  sk_mesh* m = S->mesh;
  const sk_sim_params& P = S->prm;
  const bool mixed = S->mixed();
  const int64_t n3 = 3 * m->nv;
  const double inv_h2 = c_mass / (P.h * P.h);
  const double* lam = S->lam.p;
  const double* fext = S->fext.p;
  const uint8_t* fixed = S->fixed.p;
  sk_step_info inf{};
  inf.status = SK_OK;
  SK_HIP_CHECK(hipMemsetAsync(S->fp64_count.p, 0, sizeof(unsigned long long), st));

  SK_TRY(launch_elem_grad(m, P.material, S->x.p, nullptr, S->mu.p, lam, S->vol.p, st));
  SK_TRY(launch_gather_grad(m, S->g.p, S->mass.p, S->x.p, S->xt.p, fext, fixed, inv_h2, st));
  SK_TRY(read_scalars(m, m->red.result.p, 1, st));
  const double g0 = std::sqrt(m->pinned[0]);
  double gn = g0;
  bool stalled = false;
  inf.g0_norm = g0;
  if (!std::isfinite(g0)) inf.status = SK_ERR_NONFINITE;
  if (n_guess > 0 && inf.status == SK_OK) {
    int best = -1;
    double best_dE = 0.0;
    // every candidate's potential difference into its own result slot, read back together (the
    // launches reuse dx and the partials in stream order)
    double* cres = m->red.result.p + Reducer::kCandidateSlot;
    static_assert(Reducer::kResults - Reducer::kCandidateSlot >= 8, "room for 8 candidates");
    for (int j = 0; j < n_guess; ++j) {
      guess_dx_kernel<<<grid_for(n3), kBlock, 0, st>>>(n3, guesses[j], S->x.p, fixed, S->order.p, S->dx.p);
      SK_TRY(launch_elem_energy(m, P.material, S->x.p, nullptr, S->dx.p, 1.0, S->mu.p, lam, S->vol.p, nullptr,
                                S->mass.p, S->xt.p, fext, fixed, inv_h2, st, cres + j));
    }
    SK_TRY(read_scalars(m, cres, n_guess, st));
    for (int j = 0; j < n_guess; ++j) {
      const double dE = m->pinned[j];
      if (j == 0) inf.guess_dE = dE;
      if (dE < best_dE) {
        best = j;
        best_dE = dE;
      }
    }
    if (best >= 0) {
      inf.guess_used = best + 1;
      inf.start_dE = best_dE;
      if (best != n_guess - 1)  // S->dx holds the last candidate's offset
        guess_dx_kernel<<<grid_for(n3), kBlock, 0, st>>>(n3, guesses[best], S->x.p, fixed, S->order.p, S->dx.p);
      axpy_kernel<<<grid_for(n3), kBlock, 0, st>>>(n3, 1.0, S->dx.p, S->x.p);
      SK_TRY(launch_elem_grad(m, P.material, S->x.p, nullptr, S->mu.p, lam, S->vol.p, st));
      SK_TRY(launch_gather_grad(m, S->g.p, S->mass.p, S->x.p, S->xt.p, fext, fixed, inv_h2, st));
      SK_TRY(read_scalars(m, m->red.result.p, 1, st));
      gn = std::sqrt(m->pinned[0]);
      inf.guess_g_norm = gn;
      if (!std::isfinite(gn)) inf.status = SK_ERR_NONFINITE;
    }
  }

  bool converged = false;
  while (inf.status == SK_OK) {
    if (gn <= P.newton_rtol * g0 || gn == 0.0 || stalled) {
      converged = true;
      break;
    }
    if (inf.newton_iters >= P.newton_max_iter) break;
    if (P.timing) SK_HIP_CHECK(hipEventRecord(S->ev[0], st));
    SK_TRY(launch_elem_hess(m, P.material, S->x.p, nullptr, S->mu.p, lam, S->vol.p, P.psd, mixed, st, nullptr,
                            nullptr, S->fp64_count.p, P.psd_jacobi == 0));
    if (mixed)
      SK_TRY(launch_gather_hess(m, S->vals32.p, S->mass.p, inv_h2, fixed, S->binv32.p, kLayoutSELL, true, st));
    else
      SK_TRY(launch_gather_hess(m, S->vals.p, S->mass.p, inv_h2, fixed, S->binv.p, kLayoutSELL, false, st));
    if (S->pipelined()) {
      if (mixed) SK_TRY(form_precond_matrix<float>(S, S->vals32.p, S->binv32.p, S->vhat32.p, st));
      else SK_TRY(form_precond_matrix<double>(S, S->vals.p, S->binv.p, S->vhat.p, st));
    }
    if (P.timing) SK_HIP_CHECK(hipEventRecord(S->ev[1], st));
    // Newton direction: H dx = -g on the free DOFs. The RHS is negated in place.
    axpy_kernel<<<grid_for(n3), kBlock, 0, st>>>(n3, -2.0, S->g.p, S->g.p);
    int64_t it = 0;
    int done = 0, refines = 0;
    // The FP64 solve is not waited for: the slope's launches follow it in the
    // stream and its result is read after the slope's readback.
    if (mixed) SK_TRY(run_mixed_solve(S, st, &it, &refines, &done));
    else if (S->pipelined())
      SK_TRY(run_pipe_pcg<double>(S, bufs64(S), S->vhat.p, S->pp.p, S->g.p, P.pcg_rtol, P.pcg_max_iter, st,
                                  nullptr, nullptr));
    else SK_TRY(run_block_pcg<double>(S, bufs64(S), S->g.p, P.pcg_rtol, P.pcg_max_iter, st, nullptr, nullptr));
    axpy_kernel<<<grid_for(n3), kBlock, 0, st>>>(n3, -2.0, S->g.p, S->g.p);
    if (P.timing) SK_HIP_CHECK(hipEventRecord(S->ev[2], st));
    dot_kernel<<<grid_for(n3), kBlock, 0, st>>>(n3, S->g.p, S->dx.p, m->red.partials.p);
    finalize_sum_kernel<1><<<1, kBlock, 0, st>>>(m->red.partials.p, grid_for(n3), m->red.result.p);
    SK_TRY(read_scalars(m, m->red.result.p, 1, st));
    if (!mixed) pcg_result(S, &it, &done);
    inf.pcg_iters += it;
    inf.refine_steps += refines;
    if (done == 2) inf.status = SK_ERR_PCG_BREAKDOWN;
    if (done == 3) inf.status = SK_ERR_NONFINITE;
    if (inf.status != SK_OK) break;
    const double slope = m->pinned[0];
    double t = 1.0;
    bool accepted = false;
    for (int ls = 0; ls < P.ls_max_iter; ++ls) {
      SK_TRY(launch_elem_energy(m, P.material, S->x.p, nullptr, S->dx.p, t, S->mu.p, lam, S->vol.p, nullptr,
                                S->mass.p, S->xt.p, fext, fixed, inv_h2, st));
      SK_TRY(read_scalars(m, m->red.result.p, 1, st));
      inf.ls_trials++;
      const double dE = m->pinned[0];
      if (dE <= P.ls_alpha * t * slope + P.ls_threshold) {
        accepted = true;
        break;
      }
      t *= P.ls_beta;
    }
    inf.newton_iters++;
    if (!accepted) {
      inf.status = SK_ERR_LS_FAIL;
      break;
    }
    // x += t dx, with the largest |t dx| and |x| over the free DOFs into result[2..3]
    update_absmax_kernel<<<grid_for(n3), kBlock, 0, st>>>(n3, t, S->dx.p, fixed, S->x.p, m->red.partials.p);
    finalize_max2_kernel<<<1, kBlock, 0, st>>>(m->red.partials.p, grid_for(n3), m->red.result.p + 2);
    if (P.timing) SK_HIP_CHECK(hipEventRecord(S->ev[3], st));
    SK_TRY(launch_elem_grad(m, P.material, S->x.p, nullptr, S->mu.p, lam, S->vol.p, st));
    SK_TRY(launch_gather_grad(m, S->g.p, S->mass.p, S->x.p, S->xt.p, fext, fixed, inv_h2, st));
    if (P.timing) SK_HIP_CHECK(hipEventRecord(S->ev[4], st));
    SK_TRY(read_scalars(m, m->red.result.p, 4, st));
    gn = std::sqrt(m->pinned[0]);
    stalled = m->pinned[2] <= kStepFloor * m->pinned[3];
    if (!std::isfinite(gn)) inf.status = SK_ERR_NONFINITE;
    if (P.timing) {
      inf.ms_hessian += elapsed(S->ev[0], S->ev[1]);
      inf.ms_pcg += elapsed(S->ev[1], S->ev[2]);
      inf.ms_line_search += elapsed(S->ev[2], S->ev[3]);
      inf.ms_gradient += elapsed(S->ev[3], S->ev[4]);
    }
  }
  if (inf.status == SK_OK && !converged) inf.status = SK_ERR_NEWTON_MAXIT;
  inf.g_norm = gn;
  inf.stalled = (converged && stalled && !(gn <= P.newton_rtol * g0)) ? 1 : 0;
  // element counts: det F <= 0 at x_next (a double count in result[4]) and the FP64 polar branch
  SK_TRY(launch_count_inverted(m, S->x.p, m->red.result.p + 4, st));
  SK_HIP_CHECK(hipMemcpyAsync(m->pinned + 1, S->fp64_count.p, sizeof(unsigned long long), hipMemcpyDeviceToHost,
                              st));
  SK_TRY(read_scalars(m, m->red.result.p + 4, 1, st));
  inf.inverted_elems = (int)m->pinned[0];
  inf.polar_fp64_elems = (int64_t)*reinterpret_cast<const unsigned long long*>(m->pinned + 1);
  if (inf.status == SK_OK && P.max_inverted >= 0 && inf.inverted_elems > P.max_inverted)
    inf.status = SK_ERR_INVERTED;
  store_state_kernel<<<grid_for(n3), kBlock, 0, st>>>(n3, S->x.p, S->order.p, x_next);
  SK_HIP_CHECK(hipGetLastError());
  if (info) *info = inf;
  return SK_OK;
}

}  // namespace

extern "C" {

int sk_sim_step(sk_sim* S, const double* x_curr, const double* x_prev, const double* x_bc, double* x_next,
                sk_step_info* info, sk_stream_t s) {
  SK_REQUIRE(S && x_curr && x_prev && x_next, "sk_sim_step: null argument");
  hipStream_t st = as_stream(s);
  const int64_t n3 = 3 * S->mesh->nv;
  be_init_kernel<<<grid_for(n3), kBlock, 0, st>>>(n3, x_curr, x_prev, x_bc, S->fixed.p, S->prm.h, S->order.p,
                                                   S->x.p, S->xt.p);
  SK_HIP_CHECK(hipGetLastError());
  return step_core(S, 1.0, nullptr, 0, x_next, info, st);
}

int sk_sim_step_guess(sk_sim* S, const double* x_curr, const double* x_prev, const double* x_bc,
                      const double* x_guess, double* x_next, sk_step_info* info, sk_stream_t s) {
  SK_REQUIRE(S && x_curr && x_prev && x_next, "sk_sim_step_guess: null argument");
  hipStream_t st = as_stream(s);
  const int64_t n3 = 3 * S->mesh->nv;
  be_init_kernel<<<grid_for(n3), kBlock, 0, st>>>(n3, x_curr, x_prev, x_bc, S->fixed.p, S->prm.h, S->order.p,
                                                   S->x.p, S->xt.p);
  SK_HIP_CHECK(hipGetLastError());
  return step_core(S, 1.0, &x_guess, x_guess ? 1 : 0, x_next, info, st);
}

int sk_sim_step_guesses(sk_sim* S, const double* x_curr, const double* x_prev, const double* x_bc, int n_guess,
                        const double* const* guesses, double* x_next, sk_step_info* info, sk_stream_t s) {
  SK_REQUIRE(S && x_curr && x_prev && x_next, "sk_sim_step_guesses: null argument");
  SK_REQUIRE(n_guess >= 0 && n_guess <= 8 && (n_guess == 0 || guesses), "sk_sim_step_guesses: 0..8 guesses");
  for (int j = 0; j < n_guess; ++j) SK_REQUIRE(guesses[j], "sk_sim_step_guesses: null guess");
  hipStream_t st = as_stream(s);
  const int64_t n3 = 3 * S->mesh->nv;
  be_init_kernel<<<grid_for(n3), kBlock, 0, st>>>(n3, x_curr, x_prev, x_bc, S->fixed.p, S->prm.h, S->order.p,
                                                   S->x.p, S->xt.p);
  SK_HIP_CHECK(hipGetLastError());
  return step_core(S, 1.0, guesses, n_guess, x_next, info, st);
}

int sk_sim_step_bdf2(sk_sim* S, const double* x_curr, const double* x_prev, const double* x_prev2,
                     const double* x_prev3, const double* x_bc, double* x_next, sk_step_info* info, sk_stream_t s) {
  SK_REQUIRE(S && x_curr && x_prev && x_prev2 && x_prev3 && x_next, "sk_sim_step_bdf2: null argument");
  hipStream_t st = as_stream(s);
  const int64_t n3 = 3 * S->mesh->nv;
  bdf2_init_kernel<<<grid_for(n3), kBlock, 0, st>>>(n3, x_curr, x_prev, x_prev2, x_prev3, x_bc, S->fixed.p,
                                                     S->prm.h, S->order.p, S->x.p, S->xt.p);
  SK_HIP_CHECK(hipGetLastError());
  return step_core(S, 9.0 / 4.0, nullptr, 0, x_next, info, st);
}

int sk_sim_spmv_reps(sk_sim* S, int reps, sk_stream_t s) {
  SK_REQUIRE(S && reps >= 0, "sk_sim_spmv_reps: invalid argument");
  hipStream_t st = as_stream(s);
  // Per-step scratch vectors only; rtol 0 so every launch runs iteration 0 in
  // full. Right-hand side: the last gradient (FP64) or the last FP32 residual.
  // The pipelined form runs its iteration kernel: one start launch, then reps
  // consecutive iterations (rtol 0: none stops early on an SPD system).
  if (S->pipelined()) {
    sk_mesh* m = S->mesh;
    const int grid = (int)(8 * ((m->lpad + kPipeWaves - 1) / kPipeWaves));
    double* part = m->red.partials.p;
    auto run = [&](auto B, auto* vhat, auto* pvec, const auto* b) {
      using T = std::remove_const_t<std::remove_pointer_t<decltype(vhat)>>;
      PipeArgs<T> a{m->nv, sell_view(m), vhat, B.binv, S->fixed.p, b, B.x, B.r, B.z, B.Ap, pvec, nullptr,
                    B.p0, S->st.p, S->ab.p, 0.0, nullptr, part, grid, 0, 0, nullptr};
      ppcg_start_kernel<T><<<grid, kPipeBlock, 0, st>>>(a);
      for (int k = 0; k < reps; ++k) {
        a.k = k;
        a.win = (k & 1) ? B.p1 : B.p0;
        a.wout = (k & 1) ? B.p0 : B.p1;
        a.pin = part + 3 * (size_t)grid * (k & 1);
        a.pout = part + 3 * (size_t)grid * ((k + 1) & 1);
        ppcg_iter_kernel<T><<<grid, kPipeBlock, 0, st>>>(a);
      }
    };
    if (S->mixed()) {
      const PcgBufs<float> B = bufs32(S);
      run(B, S->vhat32.p, S->pp32.p, (const float*)B.r);
    } else {
      run(bufs64(S), S->vhat.p, S->pp.p, (const double*)S->g.p);
    }
    SK_HIP_CHECK(hipGetLastError());
    return SK_OK;
  }
  if (S->mixed()) {
    const PcgBufs<float> B = bufs32(S);
    BlockPcgArgs<float> a = block_pcg_args<float>(S, B, B.r, 0.0);
    a.pout = B.p0;
    bpcg_init_kernel<float, false><<<a.gu, kBlock, 0, st>>>(a);
    a.pin = B.p0;
    a.pout = B.p1;
    for (int k = 0; k < reps; ++k) launch_spmv<float>(a, S->mesh, st);
  } else {
    const PcgBufs<double> B = bufs64(S);
    BlockPcgArgs<double> a = block_pcg_args<double>(S, B, S->g.p, 0.0);
    a.pout = B.p0;
    bpcg_init_kernel<double, false><<<a.gu, kBlock, 0, st>>>(a);
    a.pin = B.p0;
    a.pout = B.p1;
    for (int k = 0; k < reps; ++k) launch_spmv<double>(a, S->mesh, st);
  }
  SK_HIP_CHECK(hipGetLastError());
  return SK_OK;
}

int sk_sim_hess_reps(sk_sim* S, int reps, sk_stream_t s) {
  SK_REQUIRE(S && reps >= 0, "sk_sim_hess_reps: invalid argument");
  hipStream_t st = as_stream(s);
  const double* lam = S->lam.p;
  for (int k = 0; k < reps; ++k)
    SK_TRY(launch_elem_hess(S->mesh, S->prm.material, S->x.p, nullptr, S->mu.p, lam, S->vol.p, S->prm.psd,
                            S->mixed(), st, nullptr, nullptr, nullptr, S->prm.psd_jacobi == 0));
  return SK_OK;
}

}  // extern "C"
