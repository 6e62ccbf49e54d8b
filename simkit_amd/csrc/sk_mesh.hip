// Mesh handle, sparsity pattern and the global x tier.
//
// Assembly is atomic-free and sort-free at run time. Set-up (host, once per
// mesh) builds the sorted vertex-pair block pattern, the element -> block map
// and, per block, the list of (element, local pair) contributions in
// ascending element order. Each Hessian evaluation is then two launches:
//   1. element pass: one lane per tet computes F, the 9x9 Hessian, its
//      vol-weighted PSD projection and the 10 upper 3x3 blocks of
//      K_e = B_e^T H+ B_e, staged per element (kKstR values, blocks padded
//      to 16 bytes);
//   2. block gather: one lane per CSR block sums its contributions in list
//      order and writes the 3 row segments of the block. No two lanes write the
//      same value and the summation order is fixed, so values are
//      bit-reproducible.
#include "sk_elem.h"
#include "sk_internal.h"
#include "sk_materials.h"

#include <algorithm>
#include <numeric>

namespace sk {

// ---------------------------------------------------------------------------
// Element passes
// ---------------------------------------------------------------------------
struct EnergyArgs {
  ElemView m;
  int64_t nv;
  const double* x;
  const double* F0;  // optional per-element offset added to F (u tier: F = J u + J x_bar)
  const double* dx;  // non-null: evaluate E(x + t dx) - E(x)
  double t;
  const double* mu;
  const double* lam;
  const double* vol;
  double* elem_energy;
  const double* mass;  // per DOF; non-null adds the kinetic and load difference
  const double* xt;
  const double* fext;
  const uint8_t* fixed;
  double inv_h2;
  double* partials;
};

template <class M>
__global__ __launch_bounds__(kBlock) void elem_energy_kernel(EnergyArgs a) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double acc = 0.0;
  if (e < a.m.nt) {
    int32_t v[4];
    double D[12], xe[12], F[9];
    load_elem(a.m, e, v, D);
    gather_x(a.x, v, xe);
    make_F(xe, D, F);
    add_F0(a.F0, e, F);
    const double mu = a.mu[e], lam = a.lam ? a.lam[e] : 0.0, vol = a.vol[e];
    double val = vol * M::energy(F, mu, lam);
    if (a.dx) {
      double de[12], F1[9];
      gather_x(a.dx, v, de);
#pragma unroll
      for (int k = 0; k < 12; ++k) de[k] = xe[k] + a.t * de[k];
      make_F(de, D, F1);
      add_F0(a.F0, e, F1);
      val = vol * M::energy(F1, mu, lam) - val;
    }
    if (a.elem_energy) a.elem_energy[e] = val;
    acc = val;
  }
  if (a.mass && e < a.nv) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const int64_t k = 3 * e + c;
      if (a.fixed && a.fixed[k]) continue;
      const double x0 = a.x[k];
      const double x1 = x0 + a.t * a.dx[k];
      const double d0 = x0 - a.xt[k], d1 = x1 - a.xt[k];
      acc += 0.5 * a.inv_h2 * a.mass[k] * (d1 - d0) * (d1 + d0);
      if (a.fext) acc -= a.fext[k] * (x1 - x0);
    }
  }
  double v1[1] = {acc};
  store_block_partials<1>(v1, a.partials);
}

struct ElemArgs {
  ElemView m;
  const double* x;
  const double* F0;
  const double* mu;
  const double* lam;
  const double* vol;
  double* fst;
  double* Kst;
  int psd;
  const int* active;    // optional (batched problems): skip elements whose problem is inactive
  const int32_t* pid;   // problem of each vertex, with `active`
  unsigned long long* fp64_count;  // optional: += elements whose FP32 Hessian took the FP64 polar branch
  uint8_t* flags;       // closed-form eigensystem passes: 1 where the element falls back to the Jacobi pass
};

template <class M>
__global__ __launch_bounds__(kBlock) void elem_grad_kernel(ElemArgs a) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= a.m.nt) return;
  int32_t v[4];
  double D[12], xe[12], F[9], P[9];
  load_elem(a.m, e, v, D);
  gather_x(a.x, v, xe);
  make_F(xe, D, F);
  add_F0(a.F0, e, F);
  M::gradient(F, a.mu[e], a.lam ? a.lam[e] : 0.0, P);
  const double vol = a.vol[e];
  double* out = a.fst + kFst * e;
#pragma unroll
  for (int k = 0; k < 4; ++k)
#pragma unroll
    for (int i = 0; i < 3; ++i)
      out[3 * k + i] = vol * (P[3 * i] * D[k] + P[3 * i + 1] * D[4 + k] + P[3 * i + 2] * D[8 + k]);
}

// Element Hessian in precision R (double, or float for the mixed-precision
// path). F is formed in FP64 from the FP64 positions and rounded to R; in FP32
// the densities that use the polar decomposition (ARAP, FCR) still evaluate
// elements with det F < 1e-3 in FP64, where the FP32 SVD of a near-singular F
// loses the rotation. Returns whether that FP64 branch was taken.
template <template <typename> class M, typename R>
__device__ __forceinline__ bool elem_hessian_eval(const double (&F)[9], double mu, double lam, R (&H)[45]) {
  if (sizeof(R) == 4 && M<R>::kPolar) {
    if (det3(F) < 1e-3) {
      double Hd[45];
      M<double>::hessian(F, mu, lam, Hd);
#pragma unroll
      for (int k = 0; k < 45; ++k) H[k] = (R)Hd[k];
      return true;
    }
  }
  R Fr[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) Fr[k] = (R)F[k];
  M<R>::hessian(Fr, (R)mu, (R)lam, H);
  return false;
}

// Adds the number of lanes of the wave with `flag` set to *count: one atomic
// per wave, by the lowest flagged lane (integer sums do not depend on order).
__device__ __forceinline__ void wave_count(bool flag, unsigned long long* count) {
  const unsigned long long m = __ballot(flag);
  if (flag && (int)__lane_id() == __ffsll((long long)m) - 1) atomicAdd(count, (unsigned long long)__popcll(m));
}

// kMode 0: Hessian, vol weight and 9x9 Jacobi PSD projection for every
// element. kMode 1 (densities with a closed-form eigensystem, PSD on): the
// eigenvalue modification from that eigensystem; elements it declines are
// flagged and left to kMode 2, which runs the kMode 0 path on flagged
// elements only. Splitting the two paths over two kernels keeps the
// closed-form kernel's registers (and occupancy) free of the 9x9 Jacobi's.
template <template <typename> class M, int PSD, typename R, int kMode = 0>
__global__ __launch_bounds__(kBlock) void elem_hess_kernel(ElemArgs a) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= a.m.nt) return;
  if (a.active && !a.active[a.pid[a.m.T[e]]]) return;
  if (kMode == 2 && !a.flags[e]) return;
  int32_t v[4];
  double Dd[12], xe[12], F[9];
  load_elem(a.m, e, v, Dd);
  gather_x(a.x, v, xe);
  make_F(xe, Dd, F);
  add_F0(a.F0, e, F);
  R H[45];
  if constexpr (kMode == 1) {
    R Fr[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) Fr[k] = (R)F[k];
    const bool ok = M<R>::hessian_psd_analytic(Fr, (R)a.mu[e], a.lam ? (R)a.lam[e] : R(0), (R)a.vol[e], PSD,
                                               R(1e-6), H);
    a.flags[e] = ok ? 0 : 1;
    if (!ok) return;
  } else {
    const bool fp64_polar = elem_hessian_eval<M, R>(F, a.mu[e], a.lam ? a.lam[e] : 0.0, H);
    if (sizeof(R) == 4 && M<R>::kPolar && a.fp64_count) wave_count(fp64_polar, a.fp64_count);
    const R vol = (R)a.vol[e];
#pragma unroll
    for (int k = 0; k < 45; ++k) H[k] *= vol;
    if (PSD != SK_PSD_NONE) psd_project_sym<9, R>(H, PSD, R(1e-6));
  }
  R D[12];
#pragma unroll
  for (int k = 0; k < 12; ++k) D[k] = (R)Dd[k];
  // K_e[3i+p][3j+q] = sum_{c,c'} D[c][i] H[3p+c][3q+c'] D[c'][j], upper blocks i <= j,
  // written one local row i at a time: row i's blocks are consecutive in the
  // record, so its cache lines are complete before the next row starts
  R* out = reinterpret_cast<R*>(a.Kst) + (int64_t)kKstR<R> * e;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int p = 0; p < 3; ++p) {
      R w[9];  // w[3q+c'] = sum_c D[c][i] H[3p+c][3q+c']
#pragma unroll
      for (int m = 0; m < 9; ++m)
        w[m] = D[i] * H[Sym<9>::idx(3 * p, m)] + D[4 + i] * H[Sym<9>::idx(3 * p + 1, m)] +
               D[8 + i] * H[Sym<9>::idx(3 * p + 2, m)];
#pragma unroll
      for (int j = i; j < 4; ++j)
#pragma unroll
        for (int q = 0; q < 3; ++q)
          out[ublock(i, j) * kBlkStride<R> + 3 * p + q] = w[3 * q] * D[j] + w[3 * q + 1] * D[4 + j] + w[3 * q + 2] * D[8 + j];
    }
}

// ---------------------------------------------------------------------------
// Gathers
// ---------------------------------------------------------------------------
struct GatherGArgs {
  int64_t nv;
  const int32_t* vptr;
  const int32_t* vlist;
  const double* fst;
  double* g;
  const double* mass;
  const double* x;
  const double* xt;
  const double* fext;
  const uint8_t* fixed;
  double inv_h2;
  double* partials;
};

__global__ __launch_bounds__(kBlock) void gather_grad_kernel(GatherGArgs a) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double nrm = 0.0;
  if (v < a.nv) {
    double acc[3] = {0.0, 0.0, 0.0};
    const int32_t k0 = a.vptr[v], k1 = a.vptr[v + 1];
    for (int32_t k = k0; k < k1; ++k) {
      const int32_t pk = a.vlist[k];
      const double* f = a.fst + (int64_t)kFst * (pk >> 2) + 3 * (pk & 3);
      acc[0] += f[0];
      acc[1] += f[1];
      acc[2] += f[2];
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const int64_t k = 3 * v + c;
      double gk = acc[c];
      if (a.mass) {
        gk += a.mass[k] * (a.x[k] - a.xt[k]) * a.inv_h2;
        if (a.fext) gk -= a.fext[k];
        if (a.fixed && a.fixed[k]) gk = 0.0;
        nrm += gk * gk;
      }
      a.g[k] = gk;
    }
  }
  if (a.mass) {
    double v1[1] = {nrm};
    store_block_partials<1>(v1, a.partials);
  }
}

struct GatherHArgs {
  int64_t nblk;
  const int32_t* bptr;
  const int32_t* bcol;
  const int32_t* brow;
  const int32_t* cptr;
  const int32_t* clist;
  const void* Kst;
  void* values;
  const double* mass;
  double inv_h2;
  const uint8_t* fixed;
  void* binv;
  const int32_t* sell_pos;  // non-null: sliced layout (slot_index; slot of each block): the
  int quad;                 //   SELL <kSellC, 1>, or with quad = 1 <kQuadC, kQuadVW> (FP32)
  const int32_t* binv_pos;  // non-null: the diagonal inverses in the quad layout (slot of each vertex)
  int bsr;                  // 1: BSR layout (block b at values[9 b + 3 p + q])
  const int* active;        // optional (batched problems): skip blocks whose problem is inactive
  const int32_t* pid;
};

// A staged block: FP64 in five 16-byte loads (16-byte aligned, padded to 10),
// FP32 in nine 4-byte loads.
template <typename R>
__device__ __forceinline__ void load_staged_block(const R* __restrict__ s, R (&B)[kBlkStride<R>]) {
  if constexpr (sizeof(R) == 8) {
    const double2* v = reinterpret_cast<const double2*>(s);
#pragma unroll
    for (int t = 0; t < 5; ++t) {
      const double2 d = v[t];
      B[2 * t] = d.x;
      B[2 * t + 1] = d.y;
    }
  } else {
#pragma unroll
    for (int m = 0; m < 9; ++m) B[m] = s[m];
  }
}

// Contributions loaded per pass of the gather loop. Measured at 48^3 as the
// FP64 Hessian phase per Newton iteration (element pass + gather): 1 per pass
// 0.780 ms, 2 per pass 0.710 ms, 4 per pass 0.725 ms.
constexpr int kGatherU = 2;

// One lane per block: sums the staged contributions in list order (precision
// R), adds the lumped mass / h^2 on diagonal blocks, writes the block in CSR,
// BSR or a sliced order and, for diagonal blocks, the inverse of the
// fixed-DOF-masked block (inverse of diag(f) A diag(f) + diag(1 - f), masked to
// the free DOFs).
template <typename R>
__global__ __launch_bounds__(kBlock) void gather_hess_kernel(GatherHArgs a) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= a.nblk) return;
  if (a.active && !a.active[a.pid[a.brow[b]]]) return;
  const R* __restrict__ Kst = reinterpret_cast<const R*>(a.Kst);
  R acc[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) acc[k] = R(0);
  // the staged upper block (min(i,j), max(i,j)) of contribution pk, used
  // transposed when i > j: one load path for every lane, the transpose by selects
  auto source = [&](int32_t pk, bool& tr) {
    const int i = (pk >> 2) & 3, j = pk & 3;
    tr = i > j;
    return Kst + kKstR<R> * (int64_t)(pk >> 4) + kBlkStride<R> * ublock(tr ? j : i, tr ? i : j);
  };
  auto accumulate = [&](const R (&B)[kBlkStride<R>], bool tr) {
    acc[0] += B[0];
    acc[4] += B[4];
    acc[8] += B[8];
    acc[1] += tr ? B[3] : B[1];
    acc[3] += tr ? B[1] : B[3];
    acc[2] += tr ? B[6] : B[2];
    acc[6] += tr ? B[2] : B[6];
    acc[5] += tr ? B[7] : B[5];
    acc[7] += tr ? B[5] : B[7];
  };
  const int32_t k0 = a.cptr[b], k1 = a.cptr[b + 1];
  int32_t k = k0;
  // kGatherU contributions per pass: their blocks' loads in flight together, summed in list order
  constexpr int U = kGatherU;
  for (; k + U <= k1; k += U) {
    bool t[U];
    R B[U][kBlkStride<R>];
#pragma unroll
    for (int u = 0; u < U; ++u) load_staged_block<R>(source(a.clist[k + u], t[u]), B[u]);
#pragma unroll
    for (int u = 0; u < U; ++u) accumulate(B[u], t[u]);
  }
  for (; k < k1; ++k) {
    bool t0;
    R B0[kBlkStride<R>];
    load_staged_block<R>(source(a.clist[k], t0), B0);
    accumulate(B0, t0);
  }
  const int32_t v = a.brow[b], w = a.bcol[b];
  const int32_t b0 = a.bptr[v], deg = a.bptr[v + 1] - b0, r = (int32_t)(b - b0);
  if (v == w && a.mass) {
#pragma unroll
    for (int p = 0; p < 3; ++p) acc[4 * p] += (R)(a.mass[3 * (int64_t)v + p] * a.inv_h2);
  }
  R* values = reinterpret_cast<R*>(a.values);
  if (a.sell_pos) {
    const int64_t q = a.sell_pos[b];
#pragma unroll
    for (int m = 0; m < 9; ++m)
      values[a.quad ? slot_index<kQuadC, kQuadVW>(q, m) : slot_index<kSellC, 1>(q, m)] = acc[m];
  } else if (a.bsr) {
    R* out = values + 9 * b;
#pragma unroll
    for (int m = 0; m < 9; ++m) out[m] = acc[m];
  } else {
    R* out = values + 9 * (int64_t)b0 + 3 * r;
#pragma unroll
    for (int p = 0; p < 3; ++p)
#pragma unroll
      for (int q = 0; q < 3; ++q) out[3 * p * deg + q] = acc[3 * p + q];
  }
  if (v == w && a.binv) {
    R f[3];
#pragma unroll
    for (int p = 0; p < 3; ++p) f[p] = (a.fixed && a.fixed[3 * (int64_t)v + p]) ? R(0) : R(1);
    R A[9];
#pragma unroll
    for (int p = 0; p < 3; ++p)
#pragma unroll
      for (int q = 0; q < 3; ++q) A[3 * p + q] = f[p] * f[q] * acc[3 * p + q] + (p == q ? R(1) - f[p] : R(0));
    R c[9];
    cofactor3(A, c);
    const R det = A[0] * c[0] + A[1] * c[1] + A[2] * c[2];
    const R inv = R(1) / det;
    R* binv = reinterpret_cast<R*>(a.binv);
    if (a.binv_pos) {  // quad layout: all 9 values
#pragma unroll
      for (int p = 0; p < 3; ++p)
#pragma unroll
        for (int q = 0; q < 3; ++q)
          binv[slot_index<kQuadC, kQuadVW>(a.binv_pos[v], 3 * p + q)] = f[p] * f[q] * c[3 * q + p] * inv;
    } else {  // the upper triangle (the inverse of a symmetric block)
#pragma unroll
      for (int u = 0; u < kBinv; ++u) {
        const int p = kBinvPQ[u][0], q = kBinvPQ[u][1];
        binv[kBinv * (int64_t)v + u] = f[p] * f[q] * c[3 * q + p] * inv;
      }
    }
  }
}

// Per-block partial counts of elements with det F(x) <= 0 (inverted or
// collapsed), summed by a finalize launch.
__global__ __launch_bounds__(kBlock) void count_inverted_kernel(ElemView m, const double* __restrict__ x,
                                                                double* partials) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double c[1] = {0.0};
  if (e < m.nt) {
    int32_t v[4];
    double D[12], xe[12], F[9];
    load_elem(m, e, v, D);
    gather_x(x, v, xe);
    make_F(xe, D, F);
    c[0] = det3(F) <= 0.0 ? 1.0 : 0.0;
  }
  store_block_partials<1>(c, partials);
}

__global__ void csr_pattern_kernel(int64_t nv, int64_t nblk, const int32_t* bptr, const int32_t* bcol,
                                   const int32_t* brow, int32_t* indptr, int32_t* indices) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (indptr && b <= nv) {
    if (b == nv) {
      indptr[3 * nv] = (int32_t)(9 * nblk);
    } else {
      const int32_t b0 = bptr[b], deg = bptr[b + 1] - b0;
#pragma unroll
      for (int p = 0; p < 3; ++p) indptr[3 * b + p] = 9 * b0 + 3 * p * deg;
    }
  }
  if (indices && b < nblk) {
    const int32_t v = brow[b], w = bcol[b];
    const int32_t b0 = bptr[v], deg = bptr[v + 1] - b0, r = (int32_t)(b - b0);
#pragma unroll
    for (int p = 0; p < 3; ++p)
#pragma unroll
      for (int q = 0; q < 3; ++q) indices[9 * (int64_t)b0 + 3 * p * deg + 3 * r + q] = 3 * w + q;
  }
}

__global__ void element_nnz_kernel(int64_t nt, const int32_t* T, const int32_t* bptr, const int32_t* emap,
                                   int32_t* out) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nt) return;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int32_t v = T[i * nt + e];
    const int32_t b0 = bptr[v], deg = bptr[v + 1] - b0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int32_t r = emap[(4 * i + j) * nt + e] - b0;
#pragma unroll
      for (int p = 0; p < 3; ++p)
#pragma unroll
        for (int q = 0; q < 3; ++q) out[144 * e + 12 * (3 * i + p) + 3 * j + q] = 9 * b0 + 3 * p * deg + 3 * r + q;
    }
  }
}

__global__ void aos_to_soa_kernel(const double* __restrict__ D, int64_t nt, double* __restrict__ Ds) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nt) return;
#pragma unroll
  for (int k = 0; k < 12; ++k) Ds[k * nt + e] = D[12 * e + k];
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------

int launch_elem_energy(sk_mesh* m, int material, const double* x, const double* F0, const double* dx, double t,
                       const double* mu,
                       const double* lam, const double* vol, double* elem_energy, const double* mass,
                       const double* xt, const double* fext, const uint8_t* fixed, double inv_h2,
                       hipStream_t st, double* result) {
  EnergyArgs a{view(m), m->nv, x, F0, dx, t, mu, lam, vol, elem_energy, mass, xt, fext, fixed, inv_h2,
               m->red.partials.p};
  const int64_t n = std::max<int64_t>(m->nt, mass ? m->nv : 0);
  const int g = grid_for(n);
  SK_REQUIRE(g <= Reducer::kMaxBlocks, "mesh too large for the energy reduction");
  SK_REQUIRE(dispatch_material(material, [&](auto tag) {
               elem_energy_kernel<typename decltype(tag)::template type<double>><<<g, kBlock, 0, st>>>(a);
             }),
             "unknown material");
  finalize_sum_kernel<1><<<1, kBlock, 0, st>>>(m->red.partials.p, g, result ? result : m->red.result.p);
  SK_HIP_CHECK(hipGetLastError());
  return SK_OK;
}

int launch_elem_grad(sk_mesh* m, int material, const double* x, const double* F0, const double* mu,
                     const double* lam, const double* vol, hipStream_t st) {
  ElemArgs a{view(m), x, F0, mu, lam, vol, m->fst.p, m->Kst.p, 0, nullptr, nullptr, nullptr};
  const int g = grid_for(m->nt);
  SK_REQUIRE(dispatch_material(material, [&](auto tag) {
               elem_grad_kernel<typename decltype(tag)::template type<double>><<<g, kBlock, 0, st>>>(a);
             }),
             "unknown material");
  SK_HIP_CHECK(hipGetLastError());
  return SK_OK;
}

template <template <typename> class M, typename R>
static void hess_dispatch(int psd, int g, hipStream_t st, const ElemArgs& a) {
  if constexpr (M<R>::kAnalyticEig) {
    if (a.flags && psd != SK_PSD_NONE) {
      if (psd == SK_PSD_PROJ) {
        elem_hess_kernel<M, SK_PSD_PROJ, R, 1><<<g, kBlock, 0, st>>>(a);
        elem_hess_kernel<M, SK_PSD_PROJ, R, 2><<<g, kBlock, 0, st>>>(a);
      } else {
        elem_hess_kernel<M, SK_PSD_ABS, R, 1><<<g, kBlock, 0, st>>>(a);
        elem_hess_kernel<M, SK_PSD_ABS, R, 2><<<g, kBlock, 0, st>>>(a);
      }
      return;
    }
  }
  if (psd == SK_PSD_PROJ) elem_hess_kernel<M, SK_PSD_PROJ, R><<<g, kBlock, 0, st>>>(a);
  else if (psd == SK_PSD_ABS) elem_hess_kernel<M, SK_PSD_ABS, R><<<g, kBlock, 0, st>>>(a);
  else elem_hess_kernel<M, SK_PSD_NONE, R><<<g, kBlock, 0, st>>>(a);
}

template <template <typename> class M>
static void hess_dispatch_prec(bool fp32, int psd, int g, hipStream_t st, const ElemArgs& a) {
  if (fp32) hess_dispatch<M, float>(psd, g, st, a);
  else hess_dispatch<M, double>(psd, g, st, a);
}

int launch_elem_hess(sk_mesh* m, int material, const double* x, const double* F0, const double* mu,
                     const double* lam, const double* vol, int psd, bool fp32, hipStream_t st, const int* active,
                     const int32_t* pid, unsigned long long* fp64_count, bool closed_form) {
  SK_REQUIRE(psd == SK_PSD_NONE || psd == SK_PSD_PROJ || psd == SK_PSD_ABS, "invalid psd method");
  SK_REQUIRE(!active == !pid, "launch_elem_hess: active and pid go together");
  ElemArgs a{view(m), x, F0, mu, lam, vol, m->fst.p, m->Kst.p, psd, active, pid, fp64_count,
             closed_form ? m->eflag.p : nullptr};
  const int g = grid_for(m->nt);
  bool known = true;
  switch (material) {
    case SK_SNH: hess_dispatch_prec<SNH>(fp32, psd, g, st, a); break;
    case SK_ARAP: hess_dispatch_prec<ARAP>(fp32, psd, g, st, a); break;
    case SK_LINEAR: hess_dispatch_prec<Linear>(fp32, psd, g, st, a); break;
    case SK_MMNH: hess_dispatch_prec<MMNH>(fp32, psd, g, st, a); break;
    case SK_FCR: hess_dispatch_prec<FCR>(fp32, psd, g, st, a); break;
    default: known = false;
  }
  SK_REQUIRE(known, "unknown material");
  SK_HIP_CHECK(hipGetLastError());
  return SK_OK;
}

int launch_count_inverted(sk_mesh* m, const double* x, double* result, hipStream_t st) {
  const int g = grid_for(m->nt);
  SK_REQUIRE(g <= Reducer::kMaxBlocks, "mesh too large for the inverted-element count");
  count_inverted_kernel<<<g, kBlock, 0, st>>>(view(m), x, m->red.partials.p);
  finalize_sum_kernel<1><<<1, kBlock, 0, st>>>(m->red.partials.p, g, result);
  SK_HIP_CHECK(hipGetLastError());
  return SK_OK;
}

int launch_gather_grad(sk_mesh* m, double* g, const double* mass, const double* x, const double* xt,
                       const double* fext, const uint8_t* fixed, double inv_h2, hipStream_t st) {
  GatherGArgs a{m->nv, m->vptr.p, m->vlist.p, m->fst.p, g, mass, x, xt, fext, fixed, inv_h2, m->red.partials.p};
  const int gr = grid_for(m->nv);
  SK_REQUIRE(gr <= Reducer::kMaxBlocks, "mesh too large for the gradient reduction");
  gather_grad_kernel<<<gr, kBlock, 0, st>>>(a);
  if (mass) finalize_sum_kernel<1><<<1, kBlock, 0, st>>>(m->red.partials.p, gr, m->red.result.p);
  SK_HIP_CHECK(hipGetLastError());
  return SK_OK;
}

int launch_gather_hess(sk_mesh* m, void* values, const double* mass, double inv_h2, const uint8_t* fixed,
                       void* binv, int layout, bool fp32, hipStream_t st, const int* active, const int32_t* pid) {
  SK_REQUIRE(layout == kLayoutCSR || layout == kLayoutSELL || layout == kLayoutBSR, "invalid value layout");
  SK_REQUIRE(!active == !pid, "launch_gather_hess: active and pid go together");
  GatherHArgs a{m->nblk, m->bptr.p, m->bcol.p, m->brow.p, m->cptr.p, m->clist.p, m->Kst.p,
                values, mass, inv_h2, fixed, binv, layout == kLayoutSELL ? m->sell_pos.p : nullptr, 0,
                nullptr, layout == kLayoutBSR ? 1 : 0, active, pid};
  if (fp32) gather_hess_kernel<float><<<grid_for(m->nblk), kBlock, 0, st>>>(a);
  else gather_hess_kernel<double><<<grid_for(m->nblk), kBlock, 0, st>>>(a);
  SK_HIP_CHECK(hipGetLastError());
  return SK_OK;
}

int launch_gather_hess_quad(sk_mesh* m, void* values, const double* mass, double inv_h2, const uint8_t* fixed,
                            void* binv, const int32_t* pos, const int32_t* row_pos, bool fp32, hipStream_t st,
                            const int* active, const int32_t* pid) {
  SK_REQUIRE(pos && (!binv || row_pos), "launch_gather_hess_quad: missing slot map");
  SK_REQUIRE(fp32, "launch_gather_hess_quad: the quad layout holds FP32 values");
  SK_REQUIRE(!active == !pid, "launch_gather_hess_quad: active and pid go together");
  GatherHArgs a{m->nblk, m->bptr.p, m->bcol.p, m->brow.p, m->cptr.p, m->clist.p, m->Kst.p,
                values, mass, inv_h2, fixed, binv, pos, 1, row_pos, 0, active, pid};
  if (fp32) gather_hess_kernel<float><<<grid_for(m->nblk), kBlock, 0, st>>>(a);
  else gather_hess_kernel<double><<<grid_for(m->nblk), kBlock, 0, st>>>(a);
  SK_HIP_CHECK(hipGetLastError());
  return SK_OK;
}

int read_scalars(sk_mesh* m, const double* dev, int n, hipStream_t st) {
  SK_HIP_CHECK(hipMemcpyAsync(m->pinned, dev, n * sizeof(double), hipMemcpyDeviceToHost, st));
  SK_HIP_CHECK(hipStreamSynchronize(st));
  return SK_OK;
}

// ---------------------------------------------------------------------------
// Host pattern build (once per mesh)
// ---------------------------------------------------------------------------
static int build_mesh(const int32_t* T, int64_t nv, int64_t nt, sk_mesh* m) {
  SK_REQUIRE(nt < (int64_t(1) << 27), "sk_mesh_create: at most 2^27 tets");
  SK_REQUIRE(nv < (int64_t(1) << 29), "sk_mesh_create: too many vertices");
  for (int64_t k = 0; k < 4 * nt; ++k)
    SK_REQUIRE(T[k] >= 0 && T[k] < nv, "sk_mesh_create: tet index out of range");

  // vertex -> (element, local corner), ascending element order
  std::vector<int32_t> vptr(nv + 1, 0), vlist(4 * nt);
  for (int64_t k = 0; k < 4 * nt; ++k) vptr[T[k] + 1]++;
  std::partial_sum(vptr.begin(), vptr.end(), vptr.begin());
  {
    std::vector<int32_t> fill(vptr.begin(), vptr.end() - 1);
    for (int64_t e = 0; e < nt; ++e)
      for (int i = 0; i < 4; ++i) vlist[fill[T[4 * e + i]]++] = (int32_t)((e << 2) | i);
  }
  // sorted neighbour blocks per vertex (including the vertex itself)
  std::vector<int32_t> bptr(nv + 1, 0), bcol;
  bcol.reserve(16 * nv);
  std::vector<int32_t> nb;
  for (int64_t v = 0; v < nv; ++v) {
    nb.clear();
    nb.push_back((int32_t)v);
    for (int32_t k = vptr[v]; k < vptr[v + 1]; ++k) {
      const int64_t e = vlist[k] >> 2;
      for (int j = 0; j < 4; ++j) nb.push_back(T[4 * e + j]);
    }
    std::sort(nb.begin(), nb.end());
    nb.erase(std::unique(nb.begin(), nb.end()), nb.end());
    bcol.insert(bcol.end(), nb.begin(), nb.end());
    bptr[v + 1] = (int32_t)bcol.size();
  }
  const int64_t nblk = (int64_t)bcol.size();
  SK_REQUIRE(9 * nblk < (int64_t(1) << 31), "sk_mesh_create: nnz exceeds int32 indexing");
  std::vector<int32_t> brow(nblk);
  for (int64_t v = 0; v < nv; ++v)
    for (int32_t b = bptr[v]; b < bptr[v + 1]; ++b) brow[b] = (int32_t)v;
  // element -> block map and per-block contribution lists
  std::vector<int32_t> emap(16 * nt), cptr(nblk + 1, 0), clist(16 * nt);
  for (int64_t e = 0; e < nt; ++e)
    for (int i = 0; i < 4; ++i) {
      const int32_t v = T[4 * e + i];
      const int32_t* lo = bcol.data() + bptr[v];
      const int32_t* hi = bcol.data() + bptr[v + 1];
      for (int j = 0; j < 4; ++j) {
        const int32_t b = (int32_t)(std::lower_bound(lo, hi, T[4 * e + j]) - bcol.data());
        emap[(4 * i + j) * nt + e] = b;
        cptr[b + 1]++;
      }
    }
  std::partial_sum(cptr.begin(), cptr.end(), cptr.begin());
  {
    std::vector<int32_t> fill(cptr.begin(), cptr.end() - 1);
    for (int64_t e = 0; e < nt; ++e)
      for (int ij = 0; ij < 16; ++ij) clist[fill[emap[ij * nt + e]]++] = (int32_t)((e << 4) | ij);
  }
  // SELL-C-sigma layout: rows sorted by block count (descending, stable, so
  // rows of equal degree keep their spatial order), chunks of kSellC rows.
  std::vector<int32_t> order(nv);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) {
    return bptr[a + 1] - bptr[a] > bptr[b + 1] - bptr[b];
  });
  const int64_t nchunk = (nv + kSellC - 1) / kSellC;
  std::vector<int32_t> sell_row(nchunk * kSellC, -1), chunk_deg(nchunk), chunk_off(nchunk + 1, 0);
  for (int64_t i = 0; i < nv; ++i) sell_row[i] = order[i];
  for (int64_t c = 0; c < nchunk; ++c) {
    int32_t d = 0;
    for (int i = 0; i < kSellC; ++i) {
      const int32_t r = sell_row[c * kSellC + i];
      if (r >= 0) d = std::max(d, bptr[r + 1] - bptr[r]);
    }
    chunk_deg[c] = d;
    SK_REQUIRE((int64_t)chunk_off[c] + (int64_t)kSellC * d < (int64_t(1) << 31) / 9, "sk_mesh_create: SELL too large");
    chunk_off[c + 1] = chunk_off[c] + kSellC * d;
  }
  const int64_t nslot = chunk_off[nchunk];
  // XCD-balanced chunk order (see sk_mesh::corder). Within an XCD's list the
  // runs are merged by relative position (the k-th of n chunks of a run's
  // share sorts at k / n): on a spatially ordered mesh, chunks of different
  // degree from the same region share columns, and now run at about the same
  // time on the same XCD, whose L2 then serves their common vector entries.
  std::vector<std::vector<int32_t>> xl(8);
  std::vector<std::vector<std::pair<double, int32_t>>> xk(8);
  for (int64_t c0 = 0; c0 < nchunk;) {
    int64_t c1 = c0;
    while (c1 < nchunk && chunk_deg[c1] == chunk_deg[c0]) ++c1;
    for (int j = 0; j < 8; ++j) {
      const int64_t a0 = c0 + (c1 - c0) * j / 8, a1 = c0 + (c1 - c0) * (j + 1) / 8;
      for (int64_t c = a0; c < a1; ++c) xk[j].emplace_back((double)(c - a0) / (double)(a1 - a0), (int32_t)c);
    }
    c0 = c1;
  }
  for (int j = 0; j < 8; ++j) {
    std::stable_sort(xk[j].begin(), xk[j].end(),
                     [](const std::pair<double, int32_t>& a, const std::pair<double, int32_t>& b) {
                       return a.first < b.first;
                     });
    for (const auto& e : xk[j]) xl[j].push_back(e.second);
  }
  size_t lmax = 1;
  for (auto& l : xl) lmax = std::max(lmax, l.size());
  const int64_t lpad = (int64_t)((lmax + 3) / 4 * 4);
  std::vector<int32_t> corder(8 * lpad, -1);
  for (int j = 0; j < 8; ++j) std::copy(xl[j].begin(), xl[j].end(), corder.begin() + j * lpad);
  std::vector<int4> cmeta(8 * lpad);
  for (size_t e = 0; e < cmeta.size(); ++e) {
    const int32_t c = corder[e];
    cmeta[e] = make_int4(c, chunk_deg[c >= 0 ? c : 0], chunk_off[c >= 0 ? c : 0], 0);
  }
  std::vector<int32_t> scol(nslot, 0), sell_pos(nblk);
  for (int64_t c = 0; c < nchunk; ++c)
    for (int i = 0; i < kSellC; ++i) {
      const int32_t r = sell_row[c * kSellC + i];
      for (int32_t s = 0; s < chunk_deg[c]; ++s) {
        const int32_t q = chunk_off[c] + s * kSellC + i;
        if (r >= 0 && s < bptr[r + 1] - bptr[r]) {
          scol[q] = bcol[bptr[r] + s];
          sell_pos[bptr[r] + s] = q;
        } else {
          scol[q] = r >= 0 ? r : 0;  // padding slot: value 0, any valid column
        }
      }
    }
  std::vector<int32_t> Tsoa(4 * nt);
  for (int64_t e = 0; e < nt; ++e)
    for (int k = 0; k < 4; ++k) Tsoa[k * nt + e] = T[4 * e + k];

  m->nv = nv;
  m->nt = nt;
  m->nblk = nblk;
  SK_TRY(m->T.upload(Tsoa));
  SK_TRY(m->bptr.upload(bptr));
  SK_TRY(m->bcol.upload(bcol));
  SK_TRY(m->brow.upload(brow));
  SK_TRY(m->emap.upload(emap));
  SK_TRY(m->cptr.upload(cptr));
  SK_TRY(m->clist.upload(clist));
  SK_TRY(m->vptr.upload(vptr));
  SK_TRY(m->vlist.upload(vlist));
  SK_TRY(m->sell_row.upload(sell_row));
  SK_TRY(m->chunk_deg.upload(chunk_deg));
  SK_TRY(m->chunk_off.upload(chunk_off));
  SK_TRY(m->scol.upload(scol));
  SK_TRY(m->sell_pos.upload(sell_pos));
  SK_TRY(m->corder.upload(corder));
  SK_TRY(m->cmeta.upload(cmeta));
  m->lpad = lpad;
  m->nchunk = nchunk;
  m->nslot = nslot;
  SK_TRY(m->D.alloc(12 * (size_t)nt));
  SK_TRY(m->Kst.alloc((size_t)kKst * nt));
  SK_TRY(m->fst.alloc((size_t)kFst * nt));
  SK_TRY(m->eflag.alloc((size_t)nt));
  SK_TRY(m->red.init());
  SK_HIP_CHECK(hipHostMalloc((void**)&m->pinned, 16 * sizeof(double), hipHostMallocDefault));
  return SK_OK;
}

}  // namespace sk

using namespace sk;

static hipStream_t as_stream(sk_stream_t s) { return reinterpret_cast<hipStream_t>(s); }

extern "C" {

int sk_mesh_create(const int32_t* T_host, const double* D, int64_t nv, int64_t nt, sk_mesh** out) {
  SK_REQUIRE(out, "sk_mesh_create: null out");
  *out = nullptr;
  SK_REQUIRE(nv > 0 && nt > 0, "sk_mesh_create: empty mesh");
  SK_REQUIRE(T_host && D, "sk_mesh_create: null T or D");
  sk_mesh* m = new sk_mesh();
  int s = build_mesh(T_host, nv, nt, m);
  if (s == SK_OK) {
    aos_to_soa_kernel<<<grid_for(nt), kBlock>>>(D, nt, m->D.p);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) {
      set_error(std::string("sk_mesh_create: ") + hipGetErrorString(e));
      s = SK_ERR_HIP;
    }
  }
  if (s != SK_OK) {
    sk_mesh_destroy(m);
    return s;
  }
  *out = m;
  return SK_OK;
}

int sk_sell_order(const int32_t* T, int64_t nv, int64_t nt, int32_t* order) {
  SK_REQUIRE(T && order, "sk_sell_order: null argument");
  SK_REQUIRE(nv > 0 && nt > 0, "sk_sell_order: empty mesh");
  for (int64_t k = 0; k < 4 * nt; ++k) SK_REQUIRE(T[k] >= 0 && T[k] < nv, "sk_sell_order: tet index out of range");
  std::vector<int32_t> vptr(nv + 1, 0), vlist(4 * nt);
  for (int64_t k = 0; k < 4 * nt; ++k) vptr[T[k] + 1]++;
  std::partial_sum(vptr.begin(), vptr.end(), vptr.begin());
  {
    std::vector<int32_t> fill(vptr.begin(), vptr.end() - 1);
    for (int64_t e = 0; e < nt; ++e)
      for (int i = 0; i < 4; ++i) vlist[fill[T[4 * e + i]]++] = (int32_t)e;
  }
  std::vector<int32_t> deg(nv), nb;
  for (int64_t v = 0; v < nv; ++v) {
    nb.clear();
    nb.push_back((int32_t)v);
    for (int32_t k = vptr[v]; k < vptr[v + 1]; ++k)
      for (int j = 0; j < 4; ++j) nb.push_back(T[4 * (int64_t)vlist[k] + j]);
    std::sort(nb.begin(), nb.end());
    deg[v] = (int32_t)(std::unique(nb.begin(), nb.end()) - nb.begin());
  }
  std::iota(order, order + nv, 0);
  std::stable_sort(order, order + nv, [&](int32_t a, int32_t b) { return deg[a] > deg[b]; });
  return SK_OK;
}

int sk_mesh_destroy(sk_mesh* m) {
  if (!m) return SK_OK;
  if (m->pinned) (void)hipHostFree(m->pinned);
  delete m;
  return SK_OK;
}

int sk_mesh_info(const sk_mesh* m, int64_t* nv, int64_t* nt, int64_t* nnz, int64_t* nblocks) {
  SK_REQUIRE(m, "sk_mesh_info: null mesh");
  if (nv) *nv = m->nv;
  if (nt) *nt = m->nt;
  if (nnz) *nnz = 9 * m->nblk;
  if (nblocks) *nblocks = m->nblk;
  return SK_OK;
}

int sk_mesh_csr_pattern(const sk_mesh* m, int32_t* indptr, int32_t* indices, sk_stream_t s) {
  SK_REQUIRE(m, "sk_mesh_csr_pattern: null mesh");
  const int64_t n = std::max<int64_t>(m->nv + 1, m->nblk);
  csr_pattern_kernel<<<grid_for(n), kBlock, 0, as_stream(s)>>>(m->nv, m->nblk, m->bptr.p, m->bcol.p, m->brow.p,
                                                                indptr, indices);
  SK_HIP_CHECK(hipGetLastError());
  return SK_OK;
}

int sk_mesh_element_nnz(const sk_mesh* m, int32_t* elem_nnz, sk_stream_t s) {
  SK_REQUIRE(m && elem_nnz, "sk_mesh_element_nnz: null argument");
  element_nnz_kernel<<<grid_for(m->nt), kBlock, 0, as_stream(s)>>>(m->nt, m->T.p, m->bptr.p, m->emap.p, elem_nnz);
  SK_HIP_CHECK(hipGetLastError());
  return SK_OK;
}

int sk_mesh_energy(sk_mesh* m, int material, const double* x, const double* F0, const double* mu,
                   const double* lam, const double* vol, double* energy_host, double* elem_energy, sk_stream_t s) {
  SK_REQUIRE(m && x && mu && vol, "sk_mesh_energy: null argument");
  SK_REQUIRE(!material_needs_lam(material) || lam, "sk_mesh_energy: this material needs lam");
  hipStream_t st = as_stream(s);
  SK_TRY(launch_elem_energy(m, material, x, F0, nullptr, 0.0, mu, lam, vol, elem_energy, nullptr, nullptr, nullptr,
                            nullptr, 0.0, st));
  if (energy_host) {
    SK_TRY(read_scalars(m, m->red.result.p, 1, st));
    *energy_host = m->pinned[0];
  }
  return SK_OK;
}

int sk_mesh_gradient(sk_mesh* m, int material, const double* x, const double* F0, const double* mu,
                     const double* lam, const double* vol, double* g, sk_stream_t s) {
  SK_REQUIRE(m && x && mu && vol && g, "sk_mesh_gradient: null argument");
  SK_REQUIRE(!material_needs_lam(material) || lam, "sk_mesh_gradient: this material needs lam");
  hipStream_t st = as_stream(s);
  SK_TRY(launch_elem_grad(m, material, x, F0, mu, lam, vol, st));
  SK_TRY(launch_gather_grad(m, g, nullptr, nullptr, nullptr, nullptr, nullptr, 0.0, st));
  return SK_OK;
}

int sk_mesh_hessian(sk_mesh* m, int material, const double* x, const double* F0, const double* mu,
                    const double* lam, const double* vol, int psd, double* values, sk_stream_t s) {
  SK_REQUIRE(m && x && mu && vol && values, "sk_mesh_hessian: null argument");
  SK_REQUIRE(!material_needs_lam(material) || lam, "sk_mesh_hessian: this material needs lam");
  hipStream_t st = as_stream(s);
  SK_TRY(launch_elem_hess(m, material, x, F0, mu, lam, vol, psd, false, st, nullptr, nullptr, nullptr, true));
  SK_TRY(launch_gather_hess(m, values, nullptr, 0.0, nullptr, nullptr, kLayoutCSR, false, st));
  return SK_OK;
}

}  // extern "C"
