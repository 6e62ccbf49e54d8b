// Set-up of the two-level preconditioner's coarse space (see sk_coarse.h):
// aggregation and per-aggregate orthonormal rigid-motion bases on the host,
// the Galerkin product E = W^T A W and its inverse on the device.
#include "sk_coarse.h"

#include <algorithm>
#include <chrono>
#include <cmath>

namespace sk {

namespace {

// Row p of R_v = [I_3 | e_0 x d, e_1 x d, e_2 x d] (3 x 6): the displacement
// of a vertex at offset d under the raw rigid motion a is sum_a R[p][a] a_a.
__host__ __device__ inline void rigid_row(const double d[3], int p, double out[6]) {
  out[0] = p == 0;
  out[1] = p == 1;
  out[2] = p == 2;
  // e_0 x d = (0, -dz, dy), e_1 x d = (dz, 0, -dx), e_2 x d = (-dy, dx, 0)
  const double c0[3] = {0.0, -d[2], d[1]}, c1[3] = {d[2], 0.0, -d[0]}, c2[3] = {-d[1], d[0], 0.0};
  out[3] = c0[p];
  out[4] = c1[p];
  out[5] = c2[p];
}

// Cyclic Jacobi eigen-decomposition of a symmetric 6x6 matrix (host):
// A = V diag(w) V^T, V column-major in v[6 * j + i] = V[i][j].
void jacobi6(double A[6][6], double w[6], double v[36]) {
  for (int i = 0; i < 36; ++i) v[i] = (i % 7 == 0) ? 1.0 : 0.0;
  for (int sweep = 0; sweep < 60; ++sweep) {
    double off = 0.0, diag = 0.0;
    for (int i = 0; i < 6; ++i)
      for (int j = 0; j < 6; ++j) (i == j ? diag : off) += A[i][j] * A[i][j];
    if (off <= 1e-30 * diag || off == 0.0) break;
    for (int p = 0; p < 5; ++p)
      for (int q = p + 1; q < 6; ++q) {
        if (A[p][q] == 0.0) continue;
        const double theta = (A[q][q] - A[p][p]) / (2.0 * A[p][q]);
        const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < 6; ++k) {  // A <- A G (columns p, q)
          const double akp = A[k][p], akq = A[k][q];
          A[k][p] = c * akp - s * akq;
          A[k][q] = s * akp + c * akq;
        }
        for (int k = 0; k < 6; ++k) {  // A <- G^T A (rows p, q)
          const double apk = A[p][k], aqk = A[q][k];
          A[p][k] = c * apk - s * aqk;
          A[q][k] = s * apk + c * aqk;
        }
        for (int k = 0; k < 6; ++k) {
          const double vkp = v[6 * p + k], vkq = v[6 * q + k];
          v[6 * p + k] = c * vkp - s * vkq;
          v[6 * q + k] = s * vkp + c * vkq;
        }
      }
  }
  for (int i = 0; i < 6; ++i) w[i] = A[i][i];
}

template <typename R>
struct GalerkinArgs {
  const int32_t* agg_ptr;
  const int32_t* agg_vert;
  const int32_t* agg_col;
  const int32_t* vagg;
  const int32_t* adj_ptr;
  const int32_t* adj;
  const int32_t* bptr;
  const int32_t* bcol;
  const int32_t* sell_pos;
  const R* vals;
  const uint8_t* fixed;
  const double* dv;
  const double* tm;
  double* E;
  int nc;
};

// Workgroup I forms the blocks E_IJ = T_I^T (sum over blocks (v, w), v in I,
// w in J, of R_v^T M_v A_vw M_w R_w) T_J for every aggregate J adjacent to I
// (M: the free-DOF mask), one J at a time, with a fixed-order block sum.
template <typename R>
__global__ __launch_bounds__(kBlock) void coarse_galerkin_kernel(GalerkinArgs<R> a) {
  __shared__ double lds[36 * kBlock / kWave];
  const int I = blockIdx.x;
  const int t0 = a.agg_ptr[I], t1 = a.agg_ptr[I + 1];
  const int cI = a.agg_col[I], kI = a.agg_col[I + 1] - cI;
  for (int jj = a.adj_ptr[I]; jj < a.adj_ptr[I + 1]; ++jj) {
    const int J = a.adj[jj];
    const int cJ = a.agg_col[J], kJ = a.agg_col[J + 1] - cJ;
    double acc[36];
#pragma unroll
    for (int k = 0; k < 36; ++k) acc[k] = 0.0;
    for (int t = t0 + threadIdx.x; t < t1; t += blockDim.x) {
      const int v = a.agg_vert[t];
      const double dvv[3] = {a.dv[3 * v], a.dv[3 * v + 1], a.dv[3 * v + 2]};
      for (int32_t b = a.bptr[v]; b < a.bptr[v + 1]; ++b) {
        const int w = a.bcol[b];
        if (a.vagg[w] != J) continue;
        const int64_t q = a.sell_pos[b];
        double A[9];
#pragma unroll
        for (int m = 0; m < 9; ++m) {
          const bool fr = !(a.fixed && (a.fixed[3 * (int64_t)v + m / 3] || a.fixed[3 * (int64_t)w + m % 3]));
          A[m] = fr ? (double)a.vals[slot_index<kSellC, 1>(q, m)] : 0.0;
        }
        const double dw[3] = {a.dv[3 * w], a.dv[3 * w + 1], a.dv[3 * w + 2]};
        double Rw[3][6], AR[3][6];
#pragma unroll
        for (int p = 0; p < 3; ++p) rigid_row(dw, p, Rw[p]);
#pragma unroll
        for (int p = 0; p < 3; ++p)
#pragma unroll
          for (int c = 0; c < 6; ++c) AR[p][c] = A[3 * p] * Rw[0][c] + A[3 * p + 1] * Rw[1][c] + A[3 * p + 2] * Rw[2][c];
#pragma unroll
        for (int p = 0; p < 3; ++p) {
          double Rv[6];
          rigid_row(dvv, p, Rv);
#pragma unroll
          for (int r = 0; r < 6; ++r)
#pragma unroll
            for (int c = 0; c < 6; ++c) acc[6 * r + c] += Rv[r] * AR[p][c];
        }
      }
    }
    block_sum<36>(acc, lds);
    if (threadIdx.x == 0) {
      const double* TI = a.tm + 36 * (int64_t)I;
      const double* TJ = a.tm + 36 * (int64_t)J;
      for (int j = 0; j < kI; ++j)
        for (int l = 0; l < kJ; ++l) {
          double s = 0.0;
          for (int r = 0; r < 6; ++r)
            for (int c = 0; c < 6; ++c) s += TI[6 * r + j] * acc[6 * r + c] * TJ[6 * c + l];
          a.E[(int64_t)(cI + j) * a.nc + cJ + l] = s;
        }
    }
    __syncthreads();
  }
}

// Blocked Gauss-Jordan inversion in place (E symmetric positive definite, no
// pivoting), kGJ pivots per panel P. With A11 = A[P][P], A12 = A[P][rest],
// A21 = A[rest][P], sweeping the panel's pivots turns A into
//   A11^-1,  A11^-1 A12,  -A21 A11^-1,  A22 - A21 A11^-1 A12
// (the same matrix as sweeping the pivots one at a time), so after every
// panel A holds E^-1. Per panel: the panel's inverse in LDS (its pivots, the
// Schur complements' diagonals, recorded for the host's definiteness check),
// the panel's columns copied out, its new rows, the rank-kGJ update of every
// entry outside the panel, then the panel's rows, columns and block.
constexpr int kGJ = 48;    // pivots per panel
constexpr int kGJT = 64;   // rank-update output tile (rows and columns)

__global__ __launch_bounds__(kBlock) void gj_panel_inverse_kernel(const double* A, int n, int k0, int kb,
                                                                  double* Pinv, double* piv) {
  __shared__ double M[kGJ][kGJ + 1];
  for (int e = threadIdx.x; e < kb * kb; e += blockDim.x) M[e / kb][e % kb] = A[(int64_t)(k0 + e / kb) * n + k0 + e % kb];
  __syncthreads();
  for (int j = 0; j < kb; ++j) {
    const double p = M[j][j];
    if (threadIdx.x == 0) piv[k0 + j] = p;
    for (int e = threadIdx.x; e < kb * kb; e += blockDim.x) {
      const int r = e / kb, c = e % kb;
      if (r != j && c != j) M[r][c] -= M[r][j] * M[j][c] / p;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < 2 * kb; e += blockDim.x) {
      if (e < kb) M[j][e] = (e == j) ? 1.0 / p : M[j][e] / p;
      else if (e - kb != j) M[e - kb][j] = -M[e - kb][j] / p;
    }
    __syncthreads();
  }
  for (int e = threadIdx.x; e < kb * kb; e += blockDim.x) Pinv[(e / kb) * kGJ + e % kb] = M[e / kb][e % kb];
}

// Cb[i][t] = A[i][k0 + t] (the panel's columns, all rows); Rb[t][j] =
// (A11^-1 A[P][j])_t for the columns outside the panel.
__global__ __launch_bounds__(kBlock) void gj_panel_copy_kernel(const double* A, int n, int k0, int kb,
                                                               const double* Pinv, double* Cb, double* Rb) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t nc = (int64_t)n * kb;
  if (idx < nc) {
    const int i = (int)(idx / kb), t = (int)(idx % kb);
    Cb[(int64_t)i * kGJ + t] = A[(int64_t)i * n + k0 + t];
  } else if (idx < 2 * nc) {
    const int t = (int)((idx - nc) / n), j = (int)((idx - nc) % n);
    if (j >= k0 && j < k0 + kb) return;
    double acc = 0.0;
    for (int q = 0; q < kb; ++q) acc += Pinv[t * kGJ + q] * A[(int64_t)(k0 + q) * n + j];
    Rb[(int64_t)t * n + j] = acc;
  }
}

// A[i][j] -= sum_t Cb[i][t] Rb[t][j] for i, j outside the panel: one
// kGJT x kGJT tile per workgroup, each thread 4 x 4 outputs, operands in LDS.
__global__ __launch_bounds__(kBlock) void gj_update_kernel(double* A, int n, int k0, int kb, const double* Cb,
                                                           const double* Rb) {
  __shared__ double Cs[kGJT][kGJ + 1];
  __shared__ double Rs[kGJ][kGJT];
  const int tiles = (n + kGJT - 1) / kGJT;
  const int i0 = (blockIdx.x / tiles) * kGJT, j0 = (blockIdx.x % tiles) * kGJT;
  for (int e = threadIdx.x; e < kGJT * kGJ; e += blockDim.x) {
    const int r = e / kGJ, t = e % kGJ;
    Cs[r][t] = (i0 + r < n && t < kb) ? Cb[(int64_t)(i0 + r) * kGJ + t] : 0.0;
  }
  for (int e = threadIdx.x; e < kGJ * kGJT; e += blockDim.x) {
    const int t = e / kGJT, c = e % kGJT;
    const int j = j0 + c;
    Rs[t][c] = (t < kb && j < n && !(j >= k0 && j < k0 + kb)) ? Rb[(int64_t)t * n + j] : 0.0;
  }
  __syncthreads();
  const int ty = threadIdx.x / 16, tx = threadIdx.x % 16;
  double acc[4][4] = {};
  for (int t = 0; t < kb; ++t) {
    double c[4], r[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) c[a] = Cs[ty * 4 + a][t];
#pragma unroll
    for (int b = 0; b < 4; ++b) r[b] = Rs[t][tx * 4 + b];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) acc[a][b] += c[a] * r[b];
  }
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    const int i = i0 + ty * 4 + a;
    if (i >= n || (i >= k0 && i < k0 + kb)) continue;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int j = j0 + tx * 4 + b;
      if (j < n && !(j >= k0 && j < k0 + kb)) A[(int64_t)i * n + j] -= acc[a][b];
    }
  }
}

// The panel's rows (Rb), columns (-Cb A11^-1) and block (A11^-1).
__global__ __launch_bounds__(kBlock) void gj_panel_finish_kernel(double* A, int n, int k0, int kb, const double* Pinv,
                                                                 const double* Cb, const double* Rb) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t nr = (int64_t)n * kb;
  if (idx < nr) {
    const int t = (int)(idx / n), j = (int)(idx % n);
    if (j >= k0 && j < k0 + kb) return;
    A[(int64_t)(k0 + t) * n + j] = Rb[(int64_t)t * n + j];
  } else if (idx < 2 * nr) {
    const int i = (int)((idx - nr) / kb), t = (int)((idx - nr) % kb);
    if (i >= k0 && i < k0 + kb) return;
    double acc = 0.0;
    for (int q = 0; q < kb; ++q) acc += Cb[(int64_t)i * kGJ + q] * Pinv[q * kGJ + t];
    A[(int64_t)i * n + k0 + t] = -acc;
  } else if (idx < 2 * nr + (int64_t)kb * kb) {
    const int e = (int)(idx - 2 * nr), s = e / kb, t = e % kb;
    A[(int64_t)(k0 + s) * n + k0 + t] = Pinv[s * kGJ + t];
  }
}

// dsc[i] = sqrt(E^-1_ii).
__global__ __launch_bounds__(kBlock) void coarse_scale_kernel(const double* A, int n, double* dsc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dsc[i] = sqrt(A[(int64_t)i * n + i]);
}

// out[i * ld + j] = FP16 of S_ij = (A_ij + A_ji) / 2 / (dsc_i dsc_j), one workgroup
// per row i; dev[i] = sum_j |S16_ij - S_ij| in a fixed order (per-thread
// strided sums, then the workgroup sum). The padding columns were zeroed.
__global__ __launch_bounds__(kBlock) void coarse_f16_kernel(const double* A, int n, int ld, const double* dsc,
                                                           _Float16* out, double* dev) {
  __shared__ double lds[kBlock / kWave];
  const int i = blockIdx.x;
  const double di = dsc[i];
  double acc[1] = {0.0};
  for (int j = threadIdx.x; j < n; j += blockDim.x) {
    const double sij = 0.5 * (A[(int64_t)i * n + j] + A[(int64_t)j * n + i]) / (di * dsc[j]);
    const _Float16 h = (_Float16)sij;
    out[(int64_t)i * ld + j] = h;
    acc[0] += fabs((double)h - sij);
  }
  block_sum<1>(acc, lds);
  if (threadIdx.x == 0) dev[i] = acc[0];
}

}  // namespace

int rigid_basis(const double* dv, const uint8_t* fixed, const int32_t* verts, int nverts, int64_t vbase,
                double tm[36]) {
  double G[6][6] = {};
  for (int t = 0; t < nverts; ++t) {
    const int64_t v = verts ? verts[t] : vbase + t;
    for (int p = 0; p < 3; ++p) {
      if (fixed && fixed[3 * v + p]) continue;
      double R[6];
      rigid_row(&dv[3 * v], p, R);
      for (int r = 0; r < 6; ++r)
        for (int c = 0; c < 6; ++c) G[r][c] += R[r] * R[c];
    }
  }
  double w[6], V[36];
  jacobi6(G, w, V);
  const double wmax = *std::max_element(w, w + 6);
  for (int i = 0; i < 36; ++i) tm[i] = 0.0;
  int k = 0;
  for (int j = 0; j < 6; ++j) {
    if (!(wmax > 0.0) || !(w[j] > 1e-9 * wmax)) continue;
    const double sc = 1.0 / std::sqrt(w[j]);
    for (int a = 0; a < 6; ++a) tm[6 * a + k] = V[6 * j + a] * sc;
    ++k;
  }
  return k;
}

int coarse_bins(const double* X, int64_t nv, int per_agg, std::vector<int32_t>& vagg,
                std::vector<int32_t>& agg_ptr, std::vector<int32_t>& agg_vert) {
  SK_REQUIRE(X, "coarse set-up: null positions");
  SK_REQUIRE(per_agg >= 1, "coarse set-up: vertices per aggregate must be >= 1");
  SK_REQUIRE(nv > 0, "coarse set-up: empty mesh");
  double lo[3], hi[3];
  for (int a = 0; a < 3; ++a) lo[a] = hi[a] = X[a];
  for (int64_t v = 0; v < nv; ++v)
    for (int a = 0; a < 3; ++a) {
      SK_REQUIRE(std::isfinite(X[3 * v + a]), "coarse set-up: non-finite position");
      lo[a] = std::min(lo[a], X[3 * v + a]);
      hi[a] = std::max(hi[a], X[3 * v + a]);
    }
  const double emax = std::max({hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2], 1e-300});
  double box = 1.0;
  for (int a = 0; a < 3; ++a) box *= std::max(hi[a] - lo[a], 1e-6 * emax);
  const double nbins = std::max(1.0, (double)nv / per_agg);
  const double s = std::cbrt(box / nbins);
  // bins per axis: the extent over s, rounded; each axis is then cut evenly
  int64_t nb[3];
  double sa[3];
  for (int a = 0; a < 3; ++a) {
    const double ext = hi[a] - lo[a];
    nb[a] = std::min<int64_t>(1 << 20, std::max<int64_t>(1, std::llround(ext / s)));
    sa[a] = ext > 0.0 ? ext / nb[a] : 1.0;
  }
  std::vector<int64_t> key(nv);
  for (int64_t v = 0; v < nv; ++v) {
    int64_t ix[3];
    for (int a = 0; a < 3; ++a)
      ix[a] = std::min<int64_t>(nb[a] - 1, std::max<int64_t>(0, (int64_t)((X[3 * v + a] - lo[a]) / sa[a])));
    key[v] = (ix[0] * nb[1] + ix[1]) * nb[2] + ix[2];
  }
  std::vector<int64_t> bins(key);
  std::sort(bins.begin(), bins.end());
  bins.erase(std::unique(bins.begin(), bins.end()), bins.end());
  const int n_agg = (int)bins.size();
  vagg.assign(nv, 0);
  agg_ptr.assign(n_agg + 1, 0);
  agg_vert.assign(nv, 0);
  for (int64_t v = 0; v < nv; ++v) {
    vagg[v] = (int32_t)(std::lower_bound(bins.begin(), bins.end(), key[v]) - bins.begin());
    ++agg_ptr[vagg[v] + 1];
  }
  for (int I = 0; I < n_agg; ++I) agg_ptr[I + 1] += agg_ptr[I];
  std::vector<int32_t> fill(agg_ptr.begin(), agg_ptr.end() - 1);
  for (int64_t v = 0; v < nv; ++v) agg_vert[fill[vagg[v]]++] = (int32_t)v;
  return SK_OK;
}

int coarse_setup(CoarseSpace& C, sk_mesh* m, const double* X, const uint8_t* fixed_host, int per_agg) {
  const int64_t nv = m->nv;
  std::vector<int32_t> vagg, agg_ptr, agg_vert;
  SK_TRY(coarse_bins(X, nv, per_agg, vagg, agg_ptr, agg_vert));
  const int n_agg = (int)agg_ptr.size() - 1;
  // centroids, offsets, per-aggregate Gram matrices of the masked rigid motions
  std::vector<double> cen(3 * (size_t)n_agg, 0.0), dv(3 * nv), tm(36 * (size_t)n_agg, 0.0);
  for (int64_t v = 0; v < nv; ++v)
    for (int a = 0; a < 3; ++a) cen[3 * vagg[v] + a] += X[3 * v + a];
  for (int I = 0; I < n_agg; ++I)
    for (int a = 0; a < 3; ++a) cen[3 * I + a] /= (agg_ptr[I + 1] - agg_ptr[I]);
  for (int64_t v = 0; v < nv; ++v)
    for (int a = 0; a < 3; ++a) dv[3 * v + a] = X[3 * v + a] - cen[3 * vagg[v] + a];
  std::vector<int32_t> agg_col(n_agg + 1, 0);
  for (int I = 0; I < n_agg; ++I) {
    const int k = rigid_basis(dv.data(), fixed_host, agg_vert.data() + agg_ptr[I], agg_ptr[I + 1] - agg_ptr[I], 0,
                              &tm[36 * (size_t)I]);
    agg_col[I + 1] = agg_col[I] + k;
  }
  const int nc = agg_col[n_agg];
  if (nc > kMaxCoarse) {
    set_error("coarse set-up: " + std::to_string(nc) + " coarse unknowns exceed " + std::to_string(kMaxCoarse) +
              "; use more vertices per aggregate");
    return SK_ERR_INVALID;
  }
  SK_REQUIRE(nc > 0, "coarse set-up: every DOF is fixed");
  // aggregate adjacency from the block pattern
  std::vector<int32_t> bptr(nv + 1), bcol(m->nblk);
  SK_HIP_CHECK(hipMemcpy(bptr.data(), m->bptr.p, bptr.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
  if (m->nblk) SK_HIP_CHECK(hipMemcpy(bcol.data(), m->bcol.p, bcol.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
  std::vector<std::vector<int32_t>> nbr(n_agg);
  for (int64_t v = 0; v < nv; ++v)
    for (int32_t b = bptr[v]; b < bptr[v + 1]; ++b) nbr[vagg[v]].push_back(vagg[bcol[b]]);
  std::vector<int32_t> adj_ptr(n_agg + 1, 0), adj;
  for (int I = 0; I < n_agg; ++I) {
    auto& l = nbr[I];
    std::sort(l.begin(), l.end());
    l.erase(std::unique(l.begin(), l.end()), l.end());
    adj.insert(adj.end(), l.begin(), l.end());
    adj_ptr[I + 1] = (int32_t)adj.size();
  }
  std::vector<float> dv32(dv.begin(), dv.end());
  std::vector<double> dva(3 * nv);
  for (int64_t t = 0; t < nv; ++t)
    for (int a = 0; a < 3; ++a) dva[3 * t + a] = dv[3 * (int64_t)agg_vert[t] + a];
  std::vector<float> dva32(dva.begin(), dva.end());
  const int ld = (nc + 7) / 8 * 8;
  C.n_agg = n_agg;
  C.nc = nc;
  C.ld = ld;
  C.built = false;
  C.ident = true;
  for (int64_t t = 0; t < nv && C.ident; ++t) C.ident = agg_vert[t] == (int32_t)t;
  SK_TRY(C.agg_ptr.upload(agg_ptr));
  SK_TRY(C.agg_vert.upload(agg_vert));
  SK_TRY(C.agg_col.upload(agg_col));
  SK_TRY(C.vagg.upload(vagg));
  SK_TRY(C.adj_ptr.upload(adj_ptr));
  SK_TRY(C.adj.upload(adj));
  SK_TRY(C.tm.upload(tm));
  SK_TRY(C.dv64.upload(dv));
  SK_TRY(C.dv32.upload(dv32));
  SK_TRY(C.dva64.upload(dva));
  SK_TRY(C.dva32.upload(dva32));
  SK_TRY(C.c.alloc(ld));
  SK_HIP_CHECK(hipMemset(C.c.p, 0, ld * sizeof(double)));
  SK_TRY(C.pivot.alloc(nc));
  SK_TRY(C.E.alloc((size_t)nc * nc));
  SK_TRY(C.gjP.alloc((size_t)kGJ * kGJ));
  SK_TRY(C.gjC.alloc((size_t)nc * kGJ));
  SK_TRY(C.gjR.alloc((size_t)kGJ * nc));
  SK_TRY(C.einv.alloc((size_t)nc * ld));
  SK_TRY(C.dsc.alloc(ld));
  SK_HIP_CHECK(hipMemset(C.dsc.p, 0, ld * sizeof(double)));
  return SK_OK;
}

int coarse_build(CoarseSpace& C, sk_mesh* m, const void* vals, bool fp32, const uint8_t* fixed, hipStream_t st) {
  SK_REQUIRE(C.nc > 0, "coarse build: no coarse space");
  const auto t0 = std::chrono::steady_clock::now();
  const int nc = C.nc;
  const int64_t nn = (int64_t)nc * nc;
  SK_HIP_CHECK(hipMemsetAsync(C.E.p, 0, nn * sizeof(double), st));
  if (fp32) {
    GalerkinArgs<float> a{C.agg_ptr.p, C.agg_vert.p, C.agg_col.p, C.vagg.p, C.adj_ptr.p, C.adj.p, m->bptr.p,
                          m->bcol.p, m->sell_pos.p, (const float*)vals, fixed, C.dv64.p, C.tm.p, C.E.p, nc};
    coarse_galerkin_kernel<float><<<C.n_agg, kBlock, 0, st>>>(a);
  } else {
    GalerkinArgs<double> a{C.agg_ptr.p, C.agg_vert.p, C.agg_col.p, C.vagg.p, C.adj_ptr.p, C.adj.p, m->bptr.p,
                           m->bcol.p, m->sell_pos.p, (const double*)vals, fixed, C.dv64.p, C.tm.p, C.E.p, nc};
    coarse_galerkin_kernel<double><<<C.n_agg, kBlock, 0, st>>>(a);
  }
  SK_HIP_CHECK(hipGetLastError());
  const int tiles = (nc + kGJT - 1) / kGJT;
  for (int k0 = 0; k0 < nc; k0 += kGJ) {
    const int kb = std::min(kGJ, nc - k0);
    const int64_t nk = (int64_t)nc * kb;
    gj_panel_inverse_kernel<<<1, kBlock, 0, st>>>(C.E.p, nc, k0, kb, C.gjP.p, C.pivot.p);
    gj_panel_copy_kernel<<<(int)((2 * nk + kBlock - 1) / kBlock), kBlock, 0, st>>>(C.E.p, nc, k0, kb, C.gjP.p,
                                                                                   C.gjC.p, C.gjR.p);
    gj_update_kernel<<<tiles * tiles, kBlock, 0, st>>>(C.E.p, nc, k0, kb, C.gjC.p, C.gjR.p);
    gj_panel_finish_kernel<<<(int)((2 * nk + (int64_t)kb * kb + kBlock - 1) / kBlock), kBlock, 0, st>>>(
        C.E.p, nc, k0, kb, C.gjP.p, C.gjC.p, C.gjR.p);
  }
  SK_HIP_CHECK(hipGetLastError());
  std::vector<double> piv(nc), dev(nc);
  SK_HIP_CHECK(hipMemcpyAsync(piv.data(), C.pivot.p, nc * sizeof(double), hipMemcpyDeviceToHost, st));
  SK_HIP_CHECK(hipStreamSynchronize(st));
  for (int k = 0; k < nc; ++k)
    if (!(piv[k] > 0.0) || !std::isfinite(piv[k])) {
      set_error("coarse build: the coarse matrix is not positive definite (pivot " + std::to_string(k) + ")");
      return SK_ERR_PCG_BREAKDOWN;
    }
  SK_HIP_CHECK(hipMemsetAsync(C.einv.p, 0, (size_t)nc * C.ld * sizeof(_Float16), st));
  coarse_scale_kernel<<<(nc + kBlock - 1) / kBlock, kBlock, 0, st>>>(C.E.p, nc, C.dsc.p);
  // the pivots have been checked: their buffer takes the rows' rounding deviations
  coarse_f16_kernel<<<nc, kBlock, 0, st>>>(C.E.p, nc, C.ld, C.dsc.p, C.einv.p, C.pivot.p);
  SK_HIP_CHECK(hipGetLastError());
  SK_HIP_CHECK(hipMemcpyAsync(dev.data(), C.pivot.p, nc * sizeof(double), hipMemcpyDeviceToHost, st));
  SK_HIP_CHECK(hipStreamSynchronize(st));
  C.shift = 0.0;
  for (int k = 0; k < nc; ++k) {
    SK_REQUIRE(std::isfinite(dev[k]), "coarse build: non-finite coarse inverse");
    C.shift = std::max(C.shift, dev[k]);
  }
  C.built = true;
  C.build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return SK_OK;
}

}  // namespace sk

extern "C" int sk_coarse_order(const double* X, int64_t nv, int per_agg, int32_t* order, int* n_agg,
                               int32_t* agg_ptr_out) {
  SK_REQUIRE(order, "sk_coarse_order: null output");
  std::vector<int32_t> vagg, agg_ptr, agg_vert;
  SK_TRY(sk::coarse_bins(X, nv, per_agg, vagg, agg_ptr, agg_vert));
  std::copy(agg_vert.begin(), agg_vert.end(), order);
  if (n_agg) *n_agg = (int)agg_ptr.size() - 1;
  if (agg_ptr_out) std::copy(agg_ptr.begin(), agg_ptr.end(), agg_ptr_out);
  return SK_OK;
}
