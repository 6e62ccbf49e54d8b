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
#include "sk_internal.h"

#include <cmath>

namespace sk {

struct PcgState {
  double rz[2];  // r.z of iteration k in slot k & 1
  double tol2, rr, bb, pAp;
  int iter, done;  // done: 0 running, 1 converged, 2 breakdown (p.Ap <= 0), 3 non-finite
};

constexpr int kRowLanes = 8;      // lanes cooperating on one (block) row in SpMV
constexpr int kSpmvBlocks = 1024; // SpMV grid cap; blocks walk contiguous row chunks

// Logical block index with consecutive logical blocks on the same XCD: the
// hardware deals workgroups round-robin over the 8 XCDs, so XCD j receives
// hardware blocks j, j+8, ... and is given the contiguous logical range
// [j*G/8, (j+1)*G/8). Row chunks processed by one XCD then share its L2.
__device__ __forceinline__ int xcd_block(int b, int G) {
  if (G % 8) return b;
  return (b % 8) * (G / 8) + b / 8;
}

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

__device__ __forceinline__ PcgCtl spmv_prologue(PcgState* st, const double* upart, int gu, int k, double rtol2) {
  PcgCtl c{0.0, false};
  if (block_reads_done(st)) return c;
  double red[2];
  sum_partials<2>(upart, gu, gu, red);
  const double rz = red[0], rr = red[1];
  const double tol2 = (k == 0) ? rtol2 * rr : st->tol2;
  const bool finite = isfinite(rr) && isfinite(rz);
  const bool conv = finite && !(rr > tol2);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->rz[k & 1] = rz;
    st->rr = rr;
    if (k == 0) {
      st->tol2 = tol2;
      st->bb = rr;
    }
    if (!finite) st->done = 3;
    else if (conv) st->done = 1;
  }
  if (!finite || conv) return c;
  c.beta = (k == 0) ? 0.0 : rz / st->rz[(k - 1) & 1];
  c.run = true;
  return c;
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
// FEM operator: sorted block-CSR pattern read from the scalar CSR values.
// Block row v occupies values[9*bptr[v] .. 9*bptr[v+1]) as three scalar rows
// of 3*deg entries each; block r of the row holds columns 3*bcol[bptr[v]+r]+q.
// ---------------------------------------------------------------------------
struct BlockPcgArgs {
  int64_t nv;
  const int32_t* bptr;
  const int32_t* bcol;
  const double* vals;
  const double* binv;
  const uint8_t* fixed;
  const double* b;
  double* x;
  double* r;
  double* z;
  const double* pin;
  double* pout;
  double* Ap;
  PcgState* st;
  double rtol2;
  double* upart;  // [2][gu] partials of r.z and r.r
  int gu;
  double* spart;  // [gs] partials of p.Ap
  int gs;
  int k;          // iteration index of this launch
};

__device__ __forceinline__ void apply_binv(const double* __restrict__ binv, int64_t v, const double (&r)[3],
                                           double (&z)[3]) {
  const double* B = binv + 9 * v;
#pragma unroll
  for (int p = 0; p < 3; ++p) z[p] = B[3 * p] * r[0] + B[3 * p + 1] * r[1] + B[3 * p + 2] * r[2];
}

__global__ __launch_bounds__(kBlock) void bpcg_init_kernel(BlockPcgArgs a) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double rz = 0.0, rr = 0.0;
  if (v < a.nv) {
    double r[3], z[3];
#pragma unroll
    for (int p = 0; p < 3; ++p) {
      const int64_t k = 3 * v + p;
      r[p] = (a.fixed && a.fixed[k]) ? 0.0 : a.b[k];
    }
    apply_binv(a.binv, v, r, z);
#pragma unroll
    for (int p = 0; p < 3; ++p) {
      const int64_t k = 3 * v + p;
      a.r[k] = r[p];
      a.z[k] = z[p];
      a.x[k] = 0.0;
      a.pout[k] = 0.0;
      rz += r[p] * z[p];
      rr += r[p] * r[p];
    }
  }
  double vals[2] = {rz, rr};
  store_block_partials<2>(vals, a.upart);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    a.st->done = 0;
    a.st->iter = 0;
  }
}

__global__ __launch_bounds__(kBlock) void bpcg_spmv_kernel(BlockPcgArgs a) {
  const PcgCtl c = spmv_prologue(a.st, a.upart, a.gu, a.k, a.rtol2);
  if (!c.run) return;
  const double beta = c.beta;
  constexpr int rows_per_pass = kBlock / kRowLanes;
  const int lb = xcd_block(blockIdx.x, gridDim.x);
  const int64_t chunk = (a.nv + gridDim.x - 1) / gridDim.x;
  const int64_t v_begin = lb * chunk;
  const int64_t v_end = v_begin + chunk < a.nv ? v_begin + chunk : a.nv;
  const int lane = threadIdx.x % kRowLanes;
  double dot = 0.0;
  for (int64_t v0 = v_begin; v0 < v_end; v0 += rows_per_pass) {
    const int64_t v = v0 + threadIdx.x / kRowLanes;
    double acc[3] = {0.0, 0.0, 0.0};
    if (v < v_end) {
      const int32_t b0 = a.bptr[v];
      const int32_t len = 3 * (a.bptr[v + 1] - b0);
      const double* __restrict__ row = a.vals + 9 * (int64_t)b0;
      for (int32_t m = lane; m < len; m += kRowLanes) {
        const int32_t rb = m / 3;
        const int64_t col = 3 * (int64_t)a.bcol[b0 + rb] + (m - 3 * rb);
        const double pv = a.z[col] + beta * a.pin[col];
        acc[0] += row[m] * pv;
        acc[1] += row[len + m] * pv;
        acc[2] += row[2 * len + m] * pv;
      }
    }
#pragma unroll
    for (int off = kRowLanes / 2; off > 0; off >>= 1) {
#pragma unroll
      for (int p = 0; p < 3; ++p) acc[p] += __shfl_xor(acc[p], off, kRowLanes);
    }
    if (v < v_end && lane < 3) {
      const int64_t k = 3 * v + lane;
      const double pn = a.z[k] + beta * a.pin[k];
      const double ap = (a.fixed && a.fixed[k]) ? 0.0 : (lane == 0 ? acc[0] : (lane == 1 ? acc[1] : acc[2]));
      a.pout[k] = pn;
      a.Ap[k] = ap;
      dot += pn * ap;
    }
  }
  double vals[1] = {dot};
  store_block_partials<1>(vals, a.spart);
}

__global__ __launch_bounds__(kBlock) void bpcg_update_kernel(BlockPcgArgs a) {
  double alpha;
  if (!update_prologue(a.st, a.spart, a.gs, a.k, alpha)) return;
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double rz = 0.0, rr = 0.0;
  if (v < a.nv) {
    double r[3], z[3];
#pragma unroll
    for (int p = 0; p < 3; ++p) {
      const int64_t k = 3 * v + p;
      a.x[k] += alpha * a.pout[k];
      r[p] = a.r[k] - alpha * a.Ap[k];
    }
    apply_binv(a.binv, v, r, z);
#pragma unroll
    for (int p = 0; p < 3; ++p) {
      const int64_t k = 3 * v + p;
      a.r[k] = r[p];
      a.z[k] = z[p];
      rz += r[p] * z[p];
      rr += r[p] * r[p];
    }
  }
  double vals[2] = {rz, rr};
  store_block_partials<2>(vals, a.upart);
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
// x~ = x_c + h ((x_c - x_p) / h) (kinetic.py:87-90); x0 = x~ on free DOFs,
// prescribed (or current) values on fixed DOFs.
__global__ void be_init_kernel(int64_t n, const double* xc, const double* xp, const double* xbc,
                               const uint8_t* fixed, double h, double* x, double* xt) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const double c = xc[k];
  const double t = c + h * ((c - xp[k]) / h);
  xt[k] = t;
  x[k] = (fixed && fixed[k]) ? (xbc ? xbc[k] : c) : t;
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

}  // namespace sk

using namespace sk;

struct sk_sim {
  sk_mesh* mesh = nullptr;
  sk_sim_params prm{};
  DevBuf<double> mu, lam, vol, mass, fext;
  DevBuf<uint8_t> fixed;
  DevBuf<double> x, xt, g, dx, r, z, p0, p1, Ap, vals, binv;
  DevBuf<PcgState> st;
  PcgState* st_host = nullptr;
  hipEvent_t ev[6] = {};
  bool have_events = false;
};

namespace {

hipStream_t as_stream(sk_stream_t s) { return reinterpret_cast<hipStream_t>(s); }

int copy_dev(DevBuf<double>& dst, const double* src, size_t n) {
  SK_TRY(dst.alloc(n));
  SK_HIP_CHECK(hipMemcpy(dst.p, src, n * sizeof(double), hipMemcpyDeviceToDevice));
  return SK_OK;
}

constexpr int kPcgChunk = 16;

int spmv_grid(int64_t rows) {
  const int64_t g = (rows + (kBlock / kRowLanes) - 1) / (kBlock / kRowLanes);
  return (int)(g < kSpmvBlocks ? (g < 1 ? 1 : g) : kSpmvBlocks);
}

BlockPcgArgs block_pcg_args(sk_sim* S, const double* b, double rtol2) {
  sk_mesh* m = S->mesh;
  const int gu = grid_for(m->nv);
  const int gs = spmv_grid(m->nv);
  double* part = m->red.partials.p;
  return BlockPcgArgs{m->nv, m->bptr.p, m->bcol.p, S->vals.p, S->binv.p, S->fixed.p, b, S->dx.p, S->r.p, S->z.p,
                      S->p1.p, S->p0.p, S->Ap.p, S->st.p, rtol2, part, gu, part + 2 * gu, gs, 0};
}

// Runs block-Jacobi PCG on the sim's assembled system with right-hand side b.
int run_block_pcg(sk_sim* S, const double* b, hipStream_t st, int64_t* iters, int* done) {
  sk_mesh* m = S->mesh;
  BlockPcgArgs a = block_pcg_args(S, b, S->prm.pcg_rtol * S->prm.pcg_rtol);
  SK_REQUIRE(2 * a.gu + a.gs <= 8 * Reducer::kMaxBlocks, "mesh too large for the PCG reductions");
  a.pout = S->p0.p;  // init zeroes p0, the first "previous p"
  bpcg_init_kernel<<<a.gu, kBlock, 0, st>>>(a);
  SK_HIP_CHECK(hipGetLastError());
  int64_t launched = 0;
  for (;;) {
    for (int k = 0; k < kPcgChunk; ++k, ++launched) {
      a.k = (int)launched;
      a.pin = (launched & 1) ? S->p1.p : S->p0.p;
      a.pout = (launched & 1) ? S->p0.p : S->p1.p;
      bpcg_spmv_kernel<<<a.gs, kBlock, 0, st>>>(a);
      bpcg_update_kernel<<<a.gu, kBlock, 0, st>>>(a);
    }
    SK_HIP_CHECK(hipGetLastError());
    SK_HIP_CHECK(hipMemcpyAsync(S->st_host, S->st.p, sizeof(PcgState), hipMemcpyDeviceToHost, st));
    SK_HIP_CHECK(hipStreamSynchronize(st));
    if (S->st_host->done || S->st_host->iter >= S->prm.pcg_max_iter) break;
    if (launched > (int64_t)S->prm.pcg_max_iter + kPcgChunk) break;  // device state not advancing
  }
  (void)m;
  *iters = S->st_host->iter;
  *done = S->st_host->done;
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
  hipStream_t st = as_stream(s);
  DevBuf<double> dinv, r, z, p0, p1, Ap, partials;
  DevBuf<PcgState> state;
  SK_TRY(dinv.alloc(n));
  SK_TRY(r.alloc(n));
  SK_TRY(z.alloc(n));
  SK_TRY(p0.alloc(n));
  SK_TRY(p1.alloc(n));
  SK_TRY(Ap.alloc(n));
  const int gu = grid_for(n);
  const int gs = spmv_grid(n);
  SK_REQUIRE(gu <= Reducer::kMaxBlocks, "sk_pcg_csr: system too large");
  SK_TRY(partials.alloc(2 * (size_t)gu + gs));
  SK_TRY(state.alloc(1));
  CsrPcgArgs a{n, indptr, indices, values, dinv.p, b, x, r.p, z.p, p1.p, p0.p, Ap.p, state.p, rtol * rtol,
               partials.p, gu, partials.p + 2 * gu, gs, 0};
  cpcg_init_kernel<<<gu, kBlock, 0, st>>>(a);
  SK_HIP_CHECK(hipGetLastError());
  PcgState hs{};
  int64_t launched = 0;
  for (;;) {
    for (int k = 0; k < kPcgChunk; ++k, ++launched) {
      a.k = (int)launched;
      a.pin = (launched & 1) ? p1.p : p0.p;
      a.pout = (launched & 1) ? p0.p : p1.p;
      cpcg_spmv_kernel<<<gs, kBlock, 0, st>>>(a);
      cpcg_update_kernel<<<gu, kBlock, 0, st>>>(a);
    }
    SK_HIP_CHECK(hipGetLastError());
    SK_HIP_CHECK(hipMemcpyAsync(&hs, state.p, sizeof(PcgState), hipMemcpyDeviceToHost, st));
    SK_HIP_CHECK(hipStreamSynchronize(st));
    if (hs.done || hs.iter >= maxit) break;
    if (launched > maxit + kPcgChunk) break;  // device state not advancing
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

int sk_sim_create(sk_mesh* mesh, const sk_sim_params* p, const double* mu, const double* lam, const double* vol,
                  const double* mass, const double* f_ext, const uint8_t* fixed, sk_sim** out) {
  SK_REQUIRE(out, "sk_sim_create: null out");
  *out = nullptr;
  SK_REQUIRE(mesh && p && mu && vol && mass, "sk_sim_create: null argument");
  SK_REQUIRE(p->material == SK_ARAP || lam, "sk_sim_create: stable neo-Hookean needs lam");
  SK_REQUIRE(p->material == SK_SNH || p->material == SK_ARAP, "sk_sim_create: unknown material");
  SK_REQUIRE(p->h > 0.0, "sk_sim_create: h must be positive");
  SK_REQUIRE(p->ls_alpha > 0.0 && p->ls_alpha <= 0.5, "sk_sim_create: ls_alpha must be in (0, 0.5]");
  SK_REQUIRE(p->ls_beta > 0.0 && p->ls_beta < 1.0, "sk_sim_create: ls_beta must be in (0, 1)");
  SK_REQUIRE(p->psd == SK_PSD_NONE || p->psd == SK_PSD_PROJ || p->psd == SK_PSD_ABS, "sk_sim_create: bad psd");
  sk_sim* S = new sk_sim();
  S->mesh = mesh;
  S->prm = *p;
  const size_t nt = mesh->nt, n3 = 3 * mesh->nv, nnz = 9 * mesh->nblk;
  int s = SK_OK;
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
    for (DevBuf<double>* b : {&S->x, &S->xt, &S->g, &S->dx, &S->r, &S->z, &S->p0, &S->p1, &S->Ap})
      SK_TRY(b->alloc(n3));
    SK_TRY(S->vals.alloc(nnz));
    SK_TRY(S->binv.alloc(9 * (size_t)mesh->nv));
    SK_TRY(S->st.alloc(1));
    SK_HIP_CHECK(hipHostMalloc((void**)&S->st_host, sizeof(PcgState), hipHostMallocDefault));
    for (auto& e : S->ev) SK_HIP_CHECK(hipEventCreate(&e));
    S->have_events = true;
    return SK_OK;
  };
  s = run();
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
  if (S->have_events)
    for (auto& e : S->ev) (void)hipEventDestroy(e);
  delete S;
  return SK_OK;
}

int sk_sim_step(sk_sim* S, const double* x_curr, const double* x_prev, const double* x_bc, double* x_next,
                sk_step_info* info, sk_stream_t s) {
  SK_REQUIRE(S && x_curr && x_prev && x_next, "sk_sim_step: null argument");
  sk_mesh* m = S->mesh;
  hipStream_t st = as_stream(s);
  const sk_sim_params& P = S->prm;
  const int64_t n3 = 3 * m->nv;
  const double inv_h2 = 1.0 / (P.h * P.h);
  const double* lam = S->lam.p;
  const double* fext = S->fext.p;
  const uint8_t* fixed = S->fixed.p;
  sk_step_info inf{};
  inf.status = SK_OK;

  be_init_kernel<<<grid_for(n3), kBlock, 0, st>>>(n3, x_curr, x_prev, x_bc, fixed, P.h, S->x.p, S->xt.p);
  SK_HIP_CHECK(hipGetLastError());
  SK_TRY(launch_elem_grad(m, P.material, S->x.p, nullptr, S->mu.p, lam, S->vol.p, st));
  SK_TRY(launch_gather_grad(m, S->g.p, S->mass.p, S->x.p, S->xt.p, fext, fixed, inv_h2, st));
  SK_TRY(read_scalars(m, m->red.result.p, 1, st));
  const double g0 = std::sqrt(m->pinned[0]);
  double gn = g0;
  inf.g0_norm = g0;
  if (!std::isfinite(g0)) inf.status = SK_ERR_NONFINITE;

  bool converged = false;
  while (inf.status == SK_OK) {
    if (gn <= P.newton_rtol * g0 || gn == 0.0) {
      converged = true;
      break;
    }
    if (inf.newton_iters >= P.newton_max_iter) break;
    if (P.timing) SK_HIP_CHECK(hipEventRecord(S->ev[0], st));
    SK_TRY(launch_elem_hess(m, P.material, S->x.p, nullptr, S->mu.p, lam, S->vol.p, P.psd, st));
    SK_TRY(launch_gather_hess(m, S->vals.p, S->mass.p, inv_h2, fixed, S->binv.p, st));
    if (P.timing) SK_HIP_CHECK(hipEventRecord(S->ev[1], st));
    // Newton direction: H dx = -g on the free DOFs. The RHS is negated in place.
    axpy_kernel<<<grid_for(n3), kBlock, 0, st>>>(n3, -2.0, S->g.p, S->g.p);
    int64_t it = 0;
    int done = 0;
    SK_TRY(run_block_pcg(S, S->g.p, st, &it, &done));
    axpy_kernel<<<grid_for(n3), kBlock, 0, st>>>(n3, -2.0, S->g.p, S->g.p);
    inf.pcg_iters += it;
    if (done == 2) inf.status = SK_ERR_PCG_BREAKDOWN;
    if (done == 3) inf.status = SK_ERR_NONFINITE;
    if (inf.status != SK_OK) break;
    if (P.timing) SK_HIP_CHECK(hipEventRecord(S->ev[2], st));
    dot_kernel<<<grid_for(n3), kBlock, 0, st>>>(n3, S->g.p, S->dx.p, m->red.partials.p);
    finalize_sum_kernel<1><<<1, kBlock, 0, st>>>(m->red.partials.p, grid_for(n3), m->red.result.p);
    SK_TRY(read_scalars(m, m->red.result.p, 1, st));
    const double slope = m->pinned[0];
    double t = 1.0;
    bool accepted = false;
    for (int ls = 0; ls < P.ls_max_iter; ++ls) {
      SK_TRY(launch_elem_energy(m, P.material, S->x.p, nullptr, S->dx.p, t, S->mu.p, lam, S->vol.p, nullptr, S->mass.p,
                                S->xt.p, fext, fixed, inv_h2, st));
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
    axpy_kernel<<<grid_for(n3), kBlock, 0, st>>>(n3, t, S->dx.p, S->x.p);
    if (P.timing) SK_HIP_CHECK(hipEventRecord(S->ev[3], st));
    SK_TRY(launch_elem_grad(m, P.material, S->x.p, nullptr, S->mu.p, lam, S->vol.p, st));
    SK_TRY(launch_gather_grad(m, S->g.p, S->mass.p, S->x.p, S->xt.p, fext, fixed, inv_h2, st));
    if (P.timing) SK_HIP_CHECK(hipEventRecord(S->ev[4], st));
    SK_TRY(read_scalars(m, m->red.result.p, 1, st));
    gn = std::sqrt(m->pinned[0]);
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
  SK_HIP_CHECK(hipMemcpyAsync(x_next, S->x.p, n3 * sizeof(double), hipMemcpyDeviceToDevice, st));
  if (info) *info = inf;
  return SK_OK;
}


int sk_sim_spmv_reps(sk_sim* S, int reps, sk_stream_t s) {
  SK_REQUIRE(S && reps >= 0, "sk_sim_spmv_reps: invalid argument");
  hipStream_t st = as_stream(s);
  // Per-step scratch vectors only (dx, r, z, p0, p1, Ap); RHS = the last
  // gradient; rtol 0 so every launch runs iteration 0 in full.
  BlockPcgArgs a = block_pcg_args(S, S->g.p, 0.0);
  a.pout = S->p0.p;
  bpcg_init_kernel<<<a.gu, kBlock, 0, st>>>(a);
  a.pin = S->p0.p;
  a.pout = S->p1.p;
  for (int k = 0; k < reps; ++k) bpcg_spmv_kernel<<<a.gs, kBlock, 0, st>>>(a);
  SK_HIP_CHECK(hipGetLastError());
  return SK_OK;
}

}  // extern "C"
