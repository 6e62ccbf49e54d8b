// Linear modal analysis on the GPU: the lowest k generalized eigenpairs of
// (K + shift M) phi = lambda M phi with the fixed DOFs eliminated, K the
// assembled element Hessian at the given positions and M the lumped mass.
// Replaces linear_modal_analysis.py:18-73 -> eigs.py:66-110 (shift-invert
// Lanczos on a UMFPACK factorization) with a Chebyshev-filtered subspace
// iteration that needs only block sparse products:
//
//   * with D = M^{-1/2} on free DOFs (0 on fixed ones) the problem is the
//     standard one A psi = lambda psi, A = D (K + shift M) D, phi = D psi;
//   * a block Y of kb >= k vectors (rows = DOFs, kb contiguous per row) is
//     filtered by a degree-m Chebyshev polynomial that damps [a, b] (a = the
//     largest Ritz value of the block, b = a Gershgorin bound of A) using the
//     scaled three-term recurrence of Zhou and Saad (2007);
//   * then orthonormalized (SVQB, Stathopoulos and Wu 2002, applied twice)
//     and rotated by a Rayleigh-Ritz step, until the k lowest Ritz pairs have
//     ||A y - theta y|| <= tol max(|theta|, 1e-6 b).
//
// Device work: one block SpMV per filter degree (one wave per vertex row,
// lanes over the kb columns), tall-skinny Gram products Y^T Z and rotations
// Y C staged through LDS, with per-workgroup partials summed in a fixed
// order (deterministic). The kb x kb symmetric eigenproblems run on the host
// (cyclic Jacobi).
#include "sk_internal.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <vector>

namespace sk {
namespace {

constexpr int kGramRows = 16;   // rows per LDS stage in the Gram kernel
constexpr int kRotRows = 32;    // rows per workgroup in the rotation kernel
constexpr int kRotK = 32;       // k-chunk of C staged per step

// A = D (K + shift M) D per 3x3 block, stored 9 contiguous per block.
__global__ void scale_blocks_kernel(int64_t nblk, const int32_t* bptr, const int32_t* brow, const int32_t* bcol,
                                    const double* K, const double* dscale, double shift, double* A) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nblk) return;
  const int32_t v = brow[b], w = bcol[b];
  const int32_t b0 = bptr[v], deg = bptr[v + 1] - b0, r = (int32_t)(b - b0);
  const double* in = K + 9 * (int64_t)b0 + 3 * r;
#pragma unroll
  for (int p = 0; p < 3; ++p) {
    const double dp = dscale[3 * (int64_t)v + p];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const double dq = dscale[3 * (int64_t)w + q];
      double a = dp * dq * in[3 * p * deg + q];
      if (v == w && p == q && dp != 0.0) a += shift;
      A[9 * b + 3 * p + q] = a;
    }
  }
}

// Per vertex row: max over its 3 DOF rows of sum_j |A_ij|; per-block maxima.
__global__ __launch_bounds__(kBlock) void gershgorin_kernel(int64_t nv, const int32_t* bptr, const double* A,
                                                            double* partials) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double m = 0.0;
  if (v < nv) {
    double s[3] = {0.0, 0.0, 0.0};
    for (int32_t b = bptr[v]; b < bptr[v + 1]; ++b)
#pragma unroll
      for (int p = 0; p < 3; ++p)
        s[p] += fabs(A[9 * (int64_t)b + 3 * p]) + fabs(A[9 * (int64_t)b + 3 * p + 1]) +
                fabs(A[9 * (int64_t)b + 3 * p + 2]);
    m = fmax(s[0], fmax(s[1], s[2]));
  }
  __shared__ double red[kBlock];
  red[threadIdx.x] = m;
  __syncthreads();
  for (int o = kBlock / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + o]);
    __syncthreads();
  }
  if (threadIdx.x == 0) partials[blockIdx.x] = red[0];
}

__global__ void max_finalize_kernel(const double* partials, int n, double* out) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  double m = 0.0;
  for (int i = 0; i < n; ++i) m = fmax(m, partials[i]);
  out[0] = m;
}

// Y[r][j] uniform in [-1, 1) from a counter-based hash of (seed, r, j);
// 0 on rows where dscale is 0 (fixed DOFs).
__global__ void random_block_kernel(int64_t rows, int kb, uint64_t seed, const double* dscale, double* Y) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= rows * kb) return;
  uint64_t z = seed + 0x9E3779B97F4A7C15ull * (uint64_t)(i + 1);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  const double u = (double)(z >> 11) * (1.0 / 9007199254740992.0);
  Y[i] = dscale[i / kb] != 0.0 ? 2.0 * u - 1.0 : 0.0;
}

// Z = alpha (A X - c X) + beta W, one wave per vertex row, lane l handles
// columns l, l + 64, ... (CPL = kb / 64 per lane). W may alias Z (each entry is
// read before the same lane writes it); X must not alias Z.
template <int CPL>
__global__ __launch_bounds__(kBlock) void cheb_spmm_kernel(int64_t nv, const int32_t* __restrict__ bptr,
                                                           const int32_t* __restrict__ bcol,
                                                           const double* __restrict__ A,
                                                           const double* __restrict__ X, const double* W, double* Z,
                                                           double alpha, double beta, double c) {
  constexpr int kb = CPL * kWave;
  const int64_t v = (int64_t)blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave;
  const int lane = threadIdx.x % kWave;
  if (v >= nv) return;
  double acc[3][CPL];
#pragma unroll
  for (int p = 0; p < 3; ++p)
#pragma unroll
    for (int j = 0; j < CPL; ++j) acc[p][j] = 0.0;
  const int32_t b1 = bptr[v + 1];
  for (int32_t b = bptr[v]; b < b1; ++b) {
    const int64_t w = bcol[b];
    double a[9];
#pragma unroll
    for (int m = 0; m < 9; ++m) a[m] = A[9 * (int64_t)b + m];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const double* xr = X + (3 * w + q) * kb + lane;
#pragma unroll
      for (int j = 0; j < CPL; ++j) {
        const double x = xr[j * kWave];
#pragma unroll
        for (int p = 0; p < 3; ++p) acc[p][j] += a[3 * p + q] * x;
      }
    }
  }
#pragma unroll
  for (int p = 0; p < 3; ++p)
#pragma unroll
    for (int j = 0; j < CPL; ++j) {
      const int64_t idx = (3 * v + p) * kb + j * kWave + lane;
      double z = alpha * (acc[p][j] - c * X[idx]);
      if (beta != 0.0) z += beta * W[idx];
      Z[idx] = z;
    }
}

// Per-workgroup partial of G = X^T Y over a contiguous row range; thread
// (ti, tj) = (t / 16, t % 16) owns the TILE x TILE sub-block at (TILE ti, TILE tj).
template <int KB>
__global__ __launch_bounds__(kBlock) void gram_partial_kernel(int64_t rows, const double* __restrict__ X,
                                                              const double* __restrict__ Y, int64_t rows_per_block,
                                                              double* partials) {
  constexpr int TILE = KB / 16;
  __shared__ double xs[kGramRows][KB];
  __shared__ double ys[kGramRows][KB];
  const int t = threadIdx.x;
  const int ti = t / 16, tj = t % 16;
  double acc[TILE][TILE];
#pragma unroll
  for (int i = 0; i < TILE; ++i)
#pragma unroll
    for (int j = 0; j < TILE; ++j) acc[i][j] = 0.0;
  const int64_t r0 = (int64_t)blockIdx.x * rows_per_block;
  const int64_t r1 = r0 + rows_per_block < rows ? r0 + rows_per_block : rows;
  for (int64_t rs = r0; rs < r1; rs += kGramRows) {
    __syncthreads();
    for (int k = t; k < kGramRows * KB; k += kBlock) {
      const int rr = k / KB, cc = k % KB;
      const int64_t r = rs + rr;
      xs[rr][cc] = r < r1 ? X[r * KB + cc] : 0.0;
      ys[rr][cc] = r < r1 ? Y[r * KB + cc] : 0.0;
    }
    __syncthreads();
#pragma unroll 4
    for (int rr = 0; rr < kGramRows; ++rr) {
      double xa[TILE], yb[TILE];
#pragma unroll
      for (int i = 0; i < TILE; ++i) xa[i] = xs[rr][ti * TILE + i];
#pragma unroll
      for (int j = 0; j < TILE; ++j) yb[j] = ys[rr][tj * TILE + j];
#pragma unroll
      for (int i = 0; i < TILE; ++i)
#pragma unroll
        for (int j = 0; j < TILE; ++j) acc[i][j] += xa[i] * yb[j];
    }
  }
  double* out = partials + (int64_t)blockIdx.x * KB * KB;
#pragma unroll
  for (int i = 0; i < TILE; ++i)
#pragma unroll
    for (int j = 0; j < TILE; ++j) out[(ti * TILE + i) * KB + tj * TILE + j] = acc[i][j];
}

// G[e] = sum over workgroups of partials[g][e], in workgroup order.
__global__ void sum_partials_kernel(const double* partials, int nparts, int64_t n, double* G) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  double s = 0.0;
  for (int g = 0; g < nparts; ++g) s += partials[(int64_t)g * n + e];
  G[e] = s;
}

// Z = X C (X: rows x KB, C: KB x KB row-major), 32 rows per workgroup; thread
// (r, cg) = (t / 8, t % 8) computes row r, columns cg * KB/8 ... + KB/8.
template <int KB>
__global__ __launch_bounds__(kBlock) void rotate_kernel(int64_t rows, const double* __restrict__ X,
                                                        const double* __restrict__ C, double* __restrict__ Z) {
  constexpr int NC = KB / 8;
  __shared__ double xs[kRotRows][kRotK];
  __shared__ double cs[kRotK][KB];
  const int t = threadIdx.x;
  const int r = t / 8, cg = t % 8;
  const int64_t row0 = (int64_t)blockIdx.x * kRotRows;
  double acc[NC];
#pragma unroll
  for (int j = 0; j < NC; ++j) acc[j] = 0.0;
  for (int k0 = 0; k0 < KB; k0 += kRotK) {
    __syncthreads();
    for (int k = t; k < kRotRows * kRotK; k += kBlock) {
      const int rr = k / kRotK, kk = k % kRotK;
      const int64_t row = row0 + rr;
      xs[rr][kk] = row < rows ? X[row * KB + k0 + kk] : 0.0;
    }
    for (int k = t; k < kRotK * KB; k += kBlock) cs[k / KB][k % KB] = C[(int64_t)(k0 + k / KB) * KB + k % KB];
    __syncthreads();
#pragma unroll 4
    for (int kk = 0; kk < kRotK; ++kk) {
      const double xv = xs[r][kk];
#pragma unroll
      for (int j = 0; j < NC; ++j) acc[j] += xv * cs[kk][cg * NC + j];
    }
  }
  const int64_t row = row0 + r;
  if (row < rows) {
#pragma unroll
    for (int j = 0; j < NC; ++j) Z[row * KB + cg * NC + j] = acc[j];
  }
}

// Per-workgroup column partials of sum_r (Z[r][j] - theta_j Y[r][j])^2.
template <int KB>
__global__ __launch_bounds__(kBlock) void resid_partial_kernel(int64_t rows, const double* __restrict__ Y,
                                                               const double* __restrict__ Z,
                                                               const double* __restrict__ theta,
                                                               int64_t rows_per_block, double* partials) {
  constexpr int RL = kBlock / KB;  // row lanes per column
  __shared__ double red[kBlock];
  const int t = threadIdx.x;
  const int j = t % KB, rl = t / KB;
  const int64_t r0 = (int64_t)blockIdx.x * rows_per_block;
  const int64_t r1 = r0 + rows_per_block < rows ? r0 + rows_per_block : rows;
  const double th = theta[j];
  double s = 0.0;
  for (int64_t r = r0 + rl; r < r1; r += RL) {
    const double d = Z[r * KB + j] - th * Y[r * KB + j];
    s += d * d;
  }
  red[t] = s;
  __syncthreads();
  if (rl == 0) {
    double tot = 0.0;
    for (int q = 0; q < RL; ++q) tot += red[q * KB + j];
    partials[(int64_t)blockIdx.x * KB + j] = tot;
  }
}

// phi[r][j] = dscale[r] * Y[r][j] for j < k (drops the guard columns).
__global__ void extract_basis_kernel(int64_t rows, int kb, int k, const double* Y, const double* dscale,
                                     double* out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= rows * k) return;
  const int64_t r = i / k;
  const int j = (int)(i % k);
  out[i] = dscale[r] * Y[r * kb + j];
}

__global__ void dscale_kernel(int64_t n, const double* mass, const uint8_t* fixed, double* d) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  d[i] = (fixed && fixed[i]) || !(mass[i] > 0.0) ? 0.0 : 1.0 / sqrt(mass[i]);
}

// Per-workgroup partials of q_j = sum_r basis[r][j] w[r] (x[r] - base[r]),
// j < k (k <= 256): thread t handles column t % k on rows t / k, t / k + 256/k, ...
__global__ __launch_bounds__(kBlock) void project_partial_kernel(int64_t rows, int k, const double* __restrict__ B,
                                                                 const double* __restrict__ w,
                                                                 const double* __restrict__ x,
                                                                 const double* __restrict__ base,
                                                                 int64_t rows_per_block, double* partials) {
  __shared__ double red[kBlock];
  const int t = threadIdx.x;
  const int lanes = kBlock / k;  // row lanes per column
  const int j = t % k, rl = t / k;
  const int64_t r0 = (int64_t)blockIdx.x * rows_per_block;
  const int64_t r1 = r0 + rows_per_block < rows ? r0 + rows_per_block : rows;
  double s = 0.0;
  if (rl < lanes) {
    for (int64_t r = r0 + rl; r < r1; r += lanes) {
      const double d = (x[r] - (base ? base[r] : 0.0)) * (w ? w[r] : 1.0);
      s += B[r * k + j] * d;
    }
  }
  red[t] = s;
  __syncthreads();
  if (t < k) {
    double tot = 0.0;
    for (int q = 0; q < lanes; ++q) tot += red[q * k + t];
    partials[(int64_t)blockIdx.x * k + t] = tot;
  }
}

// out[r] = base[r] + sum_j basis[r][j] c[j].
__global__ void expand_kernel(int64_t rows, int k, const double* __restrict__ B, const double* __restrict__ c,
                              const double* base, double* out) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= rows) return;
  double s = base ? base[r] : 0.0;
  for (int j = 0; j < k; ++j) s += B[r * k + j] * c[j];
  out[r] = s;
}

__global__ void axpby_kernel(int64_t n, double a, const double* x, double b, const double* y, double* out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = a * x[i] + b * y[i];
}

// ---------------------------------------------------------------------------
// Host: symmetric eigen-decomposition (cyclic Jacobi) of a small dense matrix.
// A (n x n, row-major) is overwritten; eigenvalues ascending in w, vectors as
// columns of V (row-major).
// ---------------------------------------------------------------------------
void host_eigh(int n, std::vector<double>& A, std::vector<double>& w, std::vector<double>& V) {
  V.assign((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) V[(size_t)i * n + i] = 1.0;
  for (int sweep = 0; sweep < 60; ++sweep) {
    double off = 0.0, dia = 0.0;
    for (int p = 0; p < n; ++p) {
      dia += A[(size_t)p * n + p] * A[(size_t)p * n + p];
      for (int q = p + 1; q < n; ++q) off += A[(size_t)p * n + q] * A[(size_t)p * n + q];
    }
    if (!(off > 1e-30 * dia)) break;
    for (int p = 0; p < n - 1; ++p) {
      for (int q = p + 1; q < n; ++q) {
        const double apq = A[(size_t)p * n + q];
        if (apq == 0.0) continue;
        const double app = A[(size_t)p * n + p], aqq = A[(size_t)q * n + q];
        const double d = aqq - app;
        const double rt = std::sqrt(d * d + 4.0 * apq * apq);
        if (!(rt > 0.0)) continue;
        double t = 2.0 * apq / (std::fabs(d) + rt);
        if (d < 0.0) t = -t;
        const double c = 1.0 / std::sqrt(1.0 + t * t), s = t * c;
        for (int r = 0; r < n; ++r) {
          const double g = A[(size_t)r * n + p], h = A[(size_t)r * n + q];
          A[(size_t)r * n + p] = c * g - s * h;
          A[(size_t)r * n + q] = s * g + c * h;
        }
        for (int r = 0; r < n; ++r) {
          const double g = A[(size_t)p * n + r], h = A[(size_t)q * n + r];
          A[(size_t)p * n + r] = c * g - s * h;
          A[(size_t)q * n + r] = s * g + c * h;
        }
        for (int r = 0; r < n; ++r) {
          const double g = V[(size_t)r * n + p], h = V[(size_t)r * n + q];
          V[(size_t)r * n + p] = c * g - s * h;
          V[(size_t)r * n + q] = s * g + c * h;
        }
      }
    }
  }
  std::vector<int> idx(n);
  for (int i = 0; i < n; ++i) idx[i] = i;
  std::stable_sort(idx.begin(), idx.end(),
                   [&](int a, int b) { return A[(size_t)a * n + a] < A[(size_t)b * n + b]; });
  w.resize(n);
  std::vector<double> Vs((size_t)n * n);
  for (int j = 0; j < n; ++j) {
    w[j] = A[(size_t)idx[j] * n + idx[j]];
    for (int r = 0; r < n; ++r) Vs[(size_t)r * n + j] = V[(size_t)r * n + idx[j]];
  }
  V.swap(Vs);
}

struct Modal {
  sk_mesh* m = nullptr;
  hipStream_t st = nullptr;
  int kb = 0;
  int64_t rows = 0;
  int gram_blocks = 0;
  int64_t gram_rows = 0;
  DevBuf<double> A, dscale, Y, Z, W, G, C, theta, partials;
  int64_t spmvs = 0;

  template <int KB>
  int gram_t(const double* X, const double* Yv, std::vector<double>& out) {
    gram_partial_kernel<KB><<<gram_blocks, kBlock, 0, st>>>(rows, X, Yv, gram_rows, partials.p);
    const int64_t n = (int64_t)KB * KB;
    sum_partials_kernel<<<grid_for(n), kBlock, 0, st>>>(partials.p, gram_blocks, n, G.p);
    SK_HIP_CHECK(hipGetLastError());
    out.resize(n);
    SK_HIP_CHECK(hipMemcpyAsync(out.data(), G.p, n * sizeof(double), hipMemcpyDeviceToHost, st));
    SK_HIP_CHECK(hipStreamSynchronize(st));
    return SK_OK;
  }
  int gram(const double* X, const double* Yv, std::vector<double>& out) {
    return kb == 64 ? gram_t<64>(X, Yv, out) : gram_t<128>(X, Yv, out);
  }
  // dst = src C (C host, kb x kb); dst must not alias src.
  int rotate(const double* src, const std::vector<double>& Ch, double* dst) {
    SK_HIP_CHECK(hipMemcpyAsync(C.p, Ch.data(), Ch.size() * sizeof(double), hipMemcpyHostToDevice, st));
    const int g = (int)((rows + kRotRows - 1) / kRotRows);
    if (kb == 64) rotate_kernel<64><<<g, kBlock, 0, st>>>(rows, src, C.p, dst);
    else rotate_kernel<128><<<g, kBlock, 0, st>>>(rows, src, C.p, dst);
    SK_HIP_CHECK(hipGetLastError());
    return SK_OK;
  }
  int spmm(const double* X, const double* Wv, double* Zv, double alpha, double beta, double c) {
    const int g = (int)((m->nv + kBlock / kWave - 1) / (kBlock / kWave));
    if (kb == 64)
      cheb_spmm_kernel<1><<<g, kBlock, 0, st>>>(m->nv, m->bptr.p, m->bcol.p, A.p, X, Wv, Zv, alpha, beta, c);
    else
      cheb_spmm_kernel<2><<<g, kBlock, 0, st>>>(m->nv, m->bptr.p, m->bcol.p, A.p, X, Wv, Zv, alpha, beta, c);
    SK_HIP_CHECK(hipGetLastError());
    ++spmvs;
    return SK_OK;
  }
  // SVQB: Y <- Y D V Lambda^{-1/2} with G = Y^T Y, D = diag(G)^{-1/2},
  // D G D = V Lambda V^T (eigenvalues clamped at 1e-14 of the largest).
  int orthonormalize() {
    for (int pass = 0; pass < 2; ++pass) {
      std::vector<double> Gh, w, V;
      SK_TRY(gram(Y.p, Y.p, Gh));
      std::vector<double> d(kb);
      for (int i = 0; i < kb; ++i) {
        const double gii = Gh[(size_t)i * kb + i];
        SK_REQUIRE(std::isfinite(gii), "sk_modal_basis: non-finite block");
        d[i] = gii > 0.0 ? 1.0 / std::sqrt(gii) : 1.0;
      }
      for (int i = 0; i < kb; ++i)
        for (int j = 0; j < kb; ++j) Gh[(size_t)i * kb + j] *= d[i] * d[j];
      host_eigh(kb, Gh, w, V);
      const double floor_ = 1e-14 * std::max(w[kb - 1], 1e-300);
      std::vector<double> Ch((size_t)kb * kb);
      for (int i = 0; i < kb; ++i)
        for (int j = 0; j < kb; ++j) Ch[(size_t)i * kb + j] = d[i] * V[(size_t)i * kb + j] / std::sqrt(std::max(w[j], floor_));
      SK_TRY(rotate(Y.p, Ch, W.p));
      std::swap(Y.p, W.p);
    }
    return SK_OK;
  }
  // Rayleigh-Ritz: Z = A Y, H = Y^T Z, H = V Theta V^T; Y <- Y V, Z <- Z V;
  // residual norms of the Ritz pairs into res.
  int rayleigh_ritz(std::vector<double>& th, std::vector<double>& res) {
    SK_TRY(spmm(Y.p, nullptr, Z.p, 1.0, 0.0, 0.0));
    std::vector<double> Hh, V;
    SK_TRY(gram(Y.p, Z.p, Hh));
    for (int i = 0; i < kb; ++i)
      for (int j = i + 1; j < kb; ++j) {
        const double s = 0.5 * (Hh[(size_t)i * kb + j] + Hh[(size_t)j * kb + i]);
        Hh[(size_t)i * kb + j] = Hh[(size_t)j * kb + i] = s;
      }
    host_eigh(kb, Hh, th, V);
    SK_TRY(rotate(Y.p, V, W.p));
    std::swap(Y.p, W.p);
    SK_TRY(rotate(Z.p, V, W.p));
    std::swap(Z.p, W.p);
    SK_HIP_CHECK(hipMemcpyAsync(theta.p, th.data(), kb * sizeof(double), hipMemcpyHostToDevice, st));
    if (kb == 64) resid_partial_kernel<64><<<gram_blocks, kBlock, 0, st>>>(rows, Y.p, Z.p, theta.p, gram_rows, partials.p);
    else resid_partial_kernel<128><<<gram_blocks, kBlock, 0, st>>>(rows, Y.p, Z.p, theta.p, gram_rows, partials.p);
    sum_partials_kernel<<<grid_for(kb), kBlock, 0, st>>>(partials.p, gram_blocks, kb, G.p);
    SK_HIP_CHECK(hipGetLastError());
    res.resize(kb);
    SK_HIP_CHECK(hipMemcpyAsync(res.data(), G.p, kb * sizeof(double), hipMemcpyDeviceToHost, st));
    SK_HIP_CHECK(hipStreamSynchronize(st));
    for (auto& r : res) r = std::sqrt(r);
    return SK_OK;
  }
};

}  // namespace
}  // namespace sk

using namespace sk;

extern "C" {

int sk_modal_basis(sk_mesh* m, int material, const double* x, const double* mu, const double* lam,
                   const double* vol, const double* mass, const uint8_t* fixed, int k, double shift, double tol,
                   int degree, int max_sweeps, uint64_t seed, double* basis, double* eigenvalues,
                   sk_modal_info* info, sk_stream_t s) {
  SK_REQUIRE(m && x && mu && vol && mass && basis && eigenvalues, "sk_modal_basis: null argument");
  SK_REQUIRE(!material_needs_lam(material) || lam, "sk_modal_basis: this material needs lam");
  SK_REQUIRE(k >= 1 && k <= 96, "sk_modal_basis: k must be in [1, 96]");
  SK_REQUIRE(tol > 0.0 && degree >= 1 && max_sweeps >= 1, "sk_modal_basis: tol, degree and max_sweeps must be > 0");
  SK_REQUIRE(shift >= 0.0, "sk_modal_basis: negative shift");
  const auto t0 = std::chrono::steady_clock::now();
  hipStream_t st = reinterpret_cast<hipStream_t>(s);
  Modal md;
  md.m = m;
  md.st = st;
  md.kb = (k + std::max(k / 2, 16) <= 64) ? 64 : 128;
  md.rows = 3 * m->nv;
  SK_REQUIRE(md.rows >= md.kb, "sk_modal_basis: mesh has fewer DOFs than the subspace block");
  const int kb = md.kb;
  // Gram / residual partials: about 2048 rows per workgroup, at most 1024 workgroups.
  int64_t gr = std::max<int64_t>(2048, (md.rows + 1023) / 1024);
  gr = (gr + kGramRows - 1) / kGramRows * kGramRows;
  md.gram_rows = gr;
  md.gram_blocks = (int)((md.rows + gr - 1) / gr);

  DevBuf<double> K;
  SK_TRY(K.alloc(9 * (size_t)m->nblk));
  SK_TRY(md.A.alloc(9 * (size_t)m->nblk));
  SK_TRY(md.dscale.alloc(md.rows));
  for (DevBuf<double>* b : {&md.Y, &md.Z, &md.W}) SK_TRY(b->alloc((size_t)md.rows * kb));
  SK_TRY(md.G.alloc((size_t)kb * kb));
  SK_TRY(md.C.alloc((size_t)kb * kb));
  SK_TRY(md.theta.alloc(kb));
  SK_TRY(md.partials.alloc((size_t)md.gram_blocks * kb * kb));

  // K at x (no PSD projection unless the density needs it: psd = NONE keeps
  // the reference's linear-elasticity stiffness exact).
  SK_TRY(launch_elem_hess(m, material, x, nullptr, mu, lam, vol, SK_PSD_NONE, false, st));
  SK_TRY(launch_gather_hess(m, K.p, nullptr, 0.0, nullptr, nullptr, kLayoutCSR, false, st));
  dscale_kernel<<<grid_for(md.rows), kBlock, 0, st>>>(md.rows, mass, fixed, md.dscale.p);
  scale_blocks_kernel<<<grid_for(m->nblk), kBlock, 0, st>>>(m->nblk, m->bptr.p, m->brow.p, m->bcol.p, K.p,
                                                            md.dscale.p, shift, md.A.p);
  const int gg = grid_for(m->nv);
  DevBuf<double> gpart;
  SK_TRY(gpart.alloc(gg + 1));
  gershgorin_kernel<<<gg, kBlock, 0, st>>>(m->nv, m->bptr.p, md.A.p, gpart.p);
  max_finalize_kernel<<<1, 1, 0, st>>>(gpart.p, gg, gpart.p + gg);
  random_block_kernel<<<grid_for(md.rows * kb), kBlock, 0, st>>>(md.rows, kb, seed, md.dscale.p, md.Y.p);
  SK_HIP_CHECK(hipGetLastError());
  double bmax = 0.0;
  SK_HIP_CHECK(hipMemcpyAsync(&bmax, gpart.p + gg, sizeof(double), hipMemcpyDeviceToHost, st));
  SK_HIP_CHECK(hipStreamSynchronize(st));
  SK_REQUIRE(std::isfinite(bmax) && bmax > 0.0, "sk_modal_basis: empty or non-finite operator");

  std::vector<double> th, res;
  SK_TRY(md.orthonormalize());
  SK_TRY(md.rayleigh_ritz(th, res));
  int sweeps = 0;
  double worst = 0.0;
  for (;;) {
    // Relative residual; eigenvalues below 1e-6 of the spectrum bound (rigid
    // motions of an unconstrained mesh) are measured against that floor.
    worst = 0.0;
    for (int j = 0; j < k; ++j) worst = std::max(worst, res[j] / std::max(std::fabs(th[j]), 1e-6 * bmax));
    if (worst <= tol || sweeps >= max_sweeps) break;
    // Chebyshev filter damping [a, b]; lower = smallest Ritz value.
    const double a = th[kb - 1], lower = th[0];
    const double b = bmax * 1.01;
    SK_REQUIRE(a < b, "sk_modal_basis: spectrum bound below the Ritz values");
    const double e = 0.5 * (b - a), c = 0.5 * (b + a);
    double sigma = e / (lower - c);
    const double tau = 2.0 / sigma;
    // Y1 = (A Y - c Y) sigma / e into Z; then Y_{i+1} = 2 s'/e (A Y_i - c Y_i) - s s' Y_{i-1}.
    SK_TRY(md.spmm(md.Y.p, nullptr, md.Z.p, sigma / e, 0.0, c));
    double* prev = md.Y.p;
    double* cur = md.Z.p;
    for (int i = 1; i < degree; ++i) {
      const double s2 = 1.0 / (tau - sigma);
      SK_TRY(md.spmm(cur, prev, prev, 2.0 * s2 / e, -sigma * s2, c));
      std::swap(prev, cur);
      sigma = s2;
    }
    if (cur != md.Y.p) std::swap(md.Y.p, md.Z.p);
    SK_TRY(md.orthonormalize());
    SK_TRY(md.rayleigh_ritz(th, res));
    ++sweeps;
  }
  const int64_t nk = md.rows * k;
  extract_basis_kernel<<<grid_for(nk), kBlock, 0, st>>>(md.rows, kb, k, md.Y.p, md.dscale.p, basis);
  SK_HIP_CHECK(hipGetLastError());
  SK_HIP_CHECK(hipStreamSynchronize(st));
  for (int j = 0; j < k; ++j) eigenvalues[j] = th[j];
  if (info) {
    info->sweeps = sweeps;
    info->spmvs = md.spmvs;
    info->block = kb;
    info->max_rel_residual = worst;
    info->spectrum_bound = bmax;
    info->ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }
  if (worst > tol) {
    set_error("sk_modal_basis: residual tolerance not reached within max_sweeps");
    return SK_ERR_EIG_MAXIT;
  }
  return SK_OK;
}

int sk_dense_project(int64_t rows, int k, const double* basis, const double* weight, const double* x,
                     const double* base, double* q, sk_stream_t s) {
  SK_REQUIRE(rows >= 0 && k >= 1 && k <= kBlock, "sk_dense_project: need rows >= 0 and 1 <= k <= 256");
  SK_REQUIRE(basis && x && q, "sk_dense_project: null pointer");
  hipStream_t st = reinterpret_cast<hipStream_t>(s);
  // about 1024 workgroups (each row lane then walks a few dozen rows), at least 64 rows each
  const int64_t rpb = std::max<int64_t>(64, (rows + 1023) / 1024);
  const int nb = (int)std::max<int64_t>(1, (rows + rpb - 1) / rpb);
  DevBuf<double> partials;
  SK_TRY(partials.alloc((size_t)nb * k));
  project_partial_kernel<<<nb, kBlock, 0, st>>>(rows, k, basis, weight, x, base, rpb, partials.p);
  sum_partials_kernel<<<grid_for(k), kBlock, 0, st>>>(partials.p, nb, k, q);
  SK_HIP_CHECK(hipGetLastError());
  SK_HIP_CHECK(hipStreamSynchronize(st));  // partials are released on return
  return SK_OK;
}

int sk_dense_expand(int64_t rows, int k, const double* basis, const double* coef, const double* base, double* out,
                    sk_stream_t s) {
  SK_REQUIRE(rows >= 0 && k >= 1, "sk_dense_expand: need rows >= 0 and k >= 1");
  SK_REQUIRE(basis && coef && out, "sk_dense_expand: null pointer");
  if (rows == 0) return SK_OK;
  expand_kernel<<<grid_for(rows), kBlock, 0, reinterpret_cast<hipStream_t>(s)>>>(rows, k, basis, coef, base, out);
  SK_HIP_CHECK(hipGetLastError());
  return SK_OK;
}

int sk_axpby(int64_t n, double a, const double* x, double b, const double* y, double* out, sk_stream_t s) {
  SK_REQUIRE(n >= 0, "sk_axpby: negative size");
  if (n == 0) return SK_OK;
  SK_REQUIRE(x && y && out, "sk_axpby: null pointer");
  axpby_kernel<<<grid_for(n), kBlock, 0, reinterpret_cast<hipStream_t>(s)>>>(n, a, x, b, y, out);
  SK_HIP_CHECK(hipGetLastError());
  return SK_OK;
}

}  // extern "C"
