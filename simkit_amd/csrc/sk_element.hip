// Element-tier kernels: per-tet gradient operator and rest volume, deformation
// gradients, SNH/ARAP densities on user F arrays, batched PSD projection and
// the polar SVD / rotation gradient. One lane owns one element (or matrix);
// all per-element linear algebra stays in registers.
#include "sk_internal.h"
#include "sk_materials.h"

#include <mutex>
#include <thread>

namespace sk {

static thread_local std::string g_last_error;
void set_error(const std::string& msg) { g_last_error = msg; }

// D_e = (H Dm^-1)^T and vol_e = det(Dm)/6 with Dm = [x1-x0, x2-x0, x3-x0]
// (deformation_jacobian.py:58-61, tetrahedron_volumes.py:26-27).
// D is stored [e][c][v] (t x 3 x 4, row-major).
__global__ void tet_operator_kernel(const double* __restrict__ X, const int32_t* __restrict__ T, int64_t nt,
                                    double* __restrict__ D, double* __restrict__ vol) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nt) return;
  const int32_t v0 = T[4 * e], v1 = T[4 * e + 1], v2 = T[4 * e + 2], v3 = T[4 * e + 3];
  double Dm[9];  // Dm[i][k] = x_{k+1,i} - x_{0,i}
  const int32_t vv[3] = {v1, v2, v3};
#pragma unroll
  for (int k = 0; k < 3; ++k)
#pragma unroll
    for (int i = 0; i < 3; ++i) Dm[i * 3 + k] = X[3 * (int64_t)vv[k] + i] - X[3 * (int64_t)v0 + i];
  double c[9];
  cofactor3(Dm, c);
  const double det = Dm[0] * c[0] + Dm[1] * c[1] + Dm[2] * c[2];
  const double inv = 1.0 / det;
  // Dm^-1 = cof(Dm)^T / det, so (Dm^-1)[k][c] = c[c*3+k] / det.
  double* De = D + 12 * e;
#pragma unroll
  for (int cc = 0; cc < 3; ++cc) {
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const double m = c[cc * 3 + k] * inv;
      De[cc * 4 + 1 + k] = m;
      s += m;
    }
    De[cc * 4 + 0] = -s;
  }
  if (vol) vol[e] = det / 6.0;
}

// F[e][i][c] = sum_v U[T[e][v]][i] D[e][c][v] (deformation_gradient.py:65-67).
__global__ void deformation_gradient_kernel(const double* __restrict__ D, const int32_t* __restrict__ T,
                                            const double* __restrict__ U, int64_t nt, double* __restrict__ F) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nt) return;
  double x[12];
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int64_t vi = T[4 * e + v];
#pragma unroll
    for (int i = 0; i < 3; ++i) x[3 * v + i] = U[3 * vi + i];
  }
  const double* De = D + 12 * e;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int cc = 0; cc < 3; ++cc) {
      double s = 0.0;
#pragma unroll
      for (int v = 0; v < 4; ++v) s += x[3 * v + i] * De[cc * 4 + v];
      F[9 * e + 3 * i + cc] = s;
    }
}

template <class M>
__global__ void element_F_kernel(const double* __restrict__ Fin, const double* __restrict__ mu,
                                 const double* __restrict__ lam, int64_t nt, double* __restrict__ psi,
                                 double* __restrict__ P, double* __restrict__ H) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nt) return;
  double F[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) F[k] = Fin[9 * e + k];
  const double m = mu[e], l = lam ? lam[e] : 0.0;
  if (psi) psi[e] = M::energy(F, m, l);
  if (P) {
    double p[9];
    M::gradient(F, m, l, p);
#pragma unroll
    for (int k = 0; k < 9; ++k) P[9 * e + k] = p[k];
  }
  if (H) {
    double h[45];
    M::hessian(F, m, l, h);
#pragma unroll
    for (int a = 0; a < 9; ++a)
#pragma unroll
      for (int b = 0; b < 9; ++b) H[81 * e + 9 * a + b] = h[Sym<9>::idx(a, b)];
  }
}

// Batched PSD projection of d x d matrices (psd_project.py:28-45). Like
// numpy.linalg.eigh, only the lower triangle of each input is read.
template <int N>
__global__ void psd_project_kernel(const double* __restrict__ Hin, int64_t n, int method, double* __restrict__ Hout) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  double a[Sym<N>::size];
  const double* h = Hin + (int64_t)N * N * e;
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j <= i; ++j) a[Sym<N>::idx(i, j)] = h[i * N + j];
  psd_project_sym<N, double>(a, method, 1e-6);
  double* o = Hout + (int64_t)N * N * e;
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) o[i * N + j] = a[Sym<N>::idx(i, j)];
}

// Any other size: the same cyclic Jacobi (rotation rule and stopping test of
// jacobi_eigh) with the matrix and its eigenvectors in global memory, one lane
// per matrix: A and V are n d^2 scratch each, the result is V diag(s') V^T.
__global__ void psd_project_general_kernel(const double* __restrict__ Hin, int64_t n, int d, int method,
                                           double* __restrict__ A, double* __restrict__ V,
                                           double* __restrict__ Hout) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const int64_t o = (int64_t)d * d * e;
  const double* h = Hin + o;
  double* a = A + o;
  double* v = V + o;
  for (int i = 0; i < d; ++i)
    for (int j = 0; j < d; ++j) {
      a[i * d + j] = i >= j ? h[i * d + j] : h[j * d + i];  // lower triangle, as numpy.linalg.eigh
      v[i * d + j] = i == j ? 1.0 : 0.0;
    }
  for (int sweep = 0; sweep < 30; ++sweep) {
    double off = 0.0, dia = 0.0;
    for (int p = 0; p < d; ++p) {
      dia += a[p * d + p] * a[p * d + p];
      for (int q = p + 1; q < d; ++q) off += a[p * d + q] * a[p * d + q];
    }
    if (!(off > 1e-32 * dia)) break;
    for (int p = 0; p < d - 1; ++p)
      for (int q = p + 1; q < d; ++q) {
        const double apq = a[p * d + q], app = a[p * d + p], aqq = a[q * d + q];
        const double dd = aqq - app;
        const double rt = sqrt(dd * dd + 4.0 * apq * apq);
        double t = 2.0 * apq / (fabs(dd) + rt);
        t = dd < 0.0 ? -t : t;
        t = (apq == 0.0 || !(rt > 0.0)) ? 0.0 : t;
        const double c = 1.0 / sqrt(1.0 + t * t), sn = t * c, hh = t * apq;
        a[p * d + p] = app - hh;
        a[q * d + q] = aqq + hh;
        a[p * d + q] = a[q * d + p] = 0.0;
        for (int r = 0; r < d; ++r) {
          if (r == p || r == q) continue;
          const double g = a[r * d + p], k = a[r * d + q];
          a[r * d + p] = a[p * d + r] = c * g - sn * k;
          a[r * d + q] = a[q * d + r] = sn * g + c * k;
        }
        for (int r = 0; r < d; ++r) {
          const double g = v[r * d + p], k = v[r * d + q];
          v[r * d + p] = c * g - sn * k;
          v[r * d + q] = sn * g + c * k;
        }
      }
  }
  for (int k = 0; k < d; ++k) {
    const double ev = a[k * d + k];
    a[k * d + k] = method == SK_PSD_ABS ? fabs(ev) : method == SK_PSD_PROJ ? (ev < 1e-6 ? 1e-6 : ev) : ev;
  }
  double* out = Hout + o;
  for (int i = 0; i < d; ++i)
    for (int j = 0; j < d; ++j) {
      double acc = 0.0;
      for (int k = 0; k < d; ++k) acc += v[i * d + k] * a[k * d + k] * v[j * d + k];
      out[i * d + j] = acc;
    }
}

// Signed polar SVD per 3x3: U, s (3), V, R = U V^T (svd_rv.py:33-51, polar_svd.py:209-221).
__global__ void polar_svd_kernel(const double* __restrict__ Fin, int64_t n, double* __restrict__ U,
                                 double* __restrict__ S, double* __restrict__ V, double* __restrict__ Rout) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  double F[9], u[9], s[3], v[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) F[k] = Fin[9 * e + k];
  svd3_signed(F, u, s, v);
  if (U)
#pragma unroll
    for (int k = 0; k < 9; ++k) U[9 * e + k] = u[k];
  if (V)
#pragma unroll
    for (int k = 0; k < 9; ++k) V[9 * e + k] = v[k];
  if (S)
#pragma unroll
    for (int k = 0; k < 3; ++k) S[3 * e + k] = s[k];
  if (Rout) {
    double r[9];
    ARAP<double>::rotation(u, v, r);
#pragma unroll
    for (int k = 0; k < 9; ++k) Rout[9 * e + k] = r[k];
  }
}

// dR/dF = sum_k 2/max(s_i+s_j, 1e-8) t_k t_k^T (rotation_gradient.py:42-72).
__global__ void rotation_gradient_kernel(const double* __restrict__ Fin, int64_t n, double* __restrict__ K) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  double F[9], t[3][9], coef[3];
#pragma unroll
  for (int k = 0; k < 9; ++k) F[k] = Fin[9 * e + k];
  ARAP<double>::twists(F, t, coef);
#pragma unroll
  for (int a = 0; a < 9; ++a)
#pragma unroll
    for (int b = 0; b < 9; ++b)
      K[81 * e + 9 * a + b] = coef[0] * t[0][a] * t[0][b] + coef[1] * t[1][a] * t[1][b] + coef[2] * t[2][a] * t[2][b];
}

}  // namespace sk

using namespace sk;

static hipStream_t as_stream(sk_stream_t s) { return reinterpret_cast<hipStream_t>(s); }

extern "C" {

const char* sk_last_error(void) { return g_last_error.c_str(); }

const char* sk_status_string(int status) {
  switch (status) {
    case SK_OK: return "ok";
    case SK_ERR_INVALID: return "invalid argument";
    case SK_ERR_HIP: return "HIP runtime error";
    case SK_ERR_LS_FAIL: return "line search failed";
    case SK_ERR_PCG_BREAKDOWN: return "PCG breakdown (matrix not positive definite)";
    case SK_ERR_PCG_MAXIT: return "PCG reached its iteration limit";
    case SK_ERR_NONFINITE: return "non-finite value";
    case SK_ERR_NEWTON_MAXIT: return "Newton reached its iteration limit";
    case SK_ERR_UNSUPPORTED: return "unsupported";
    case SK_ERR_EIG_MAXIT: return "eigen-solver sweep limit reached before tol";
    default: return "unknown status";
  }
}

int sk_version(void) { return SK_VERSION; }

int sk_tet_operator(const double* X, const int32_t* T, int64_t nv, int64_t nt, double* D, double* vol,
                    sk_stream_t s) {
  SK_REQUIRE(nv >= 0 && nt >= 0, "sk_tet_operator: negative size");
  if (nt == 0) return SK_OK;
  SK_REQUIRE(X && T && D, "sk_tet_operator: null pointer");
  tet_operator_kernel<<<grid_for(nt), kBlock, 0, as_stream(s)>>>(X, T, nt, D, vol);
  SK_HIP_CHECK(hipGetLastError());
  return SK_OK;
}

int sk_deformation_gradient(const double* D, const int32_t* T, const double* U, int64_t nt, double* F,
                            sk_stream_t s) {
  SK_REQUIRE(nt >= 0, "sk_deformation_gradient: negative size");
  if (nt == 0) return SK_OK;
  SK_REQUIRE(D && T && U && F, "sk_deformation_gradient: null pointer");
  deformation_gradient_kernel<<<grid_for(nt), kBlock, 0, as_stream(s)>>>(D, T, U, nt, F);
  SK_HIP_CHECK(hipGetLastError());
  return SK_OK;
}

int sk_element_F(int material, const double* F, const double* mu, const double* lam, int64_t nt, double* psi,
                 double* P, double* H, sk_stream_t s) {
  SK_REQUIRE(nt >= 0, "sk_element_F: negative size");
  if (nt == 0) return SK_OK;
  SK_REQUIRE(F && mu, "sk_element_F: null F or mu");
  const int g = grid_for(nt);
  SK_REQUIRE(!material_needs_lam(material) || lam, "sk_element_F: this material needs lam");
  SK_REQUIRE(dispatch_material(material, [&](auto tag) {
               element_F_kernel<typename decltype(tag)::template type<double>>
                   <<<g, kBlock, 0, as_stream(s)>>>(F, mu, lam, nt, psi, P, H);
             }),
             "sk_element_F: unknown material");
  SK_HIP_CHECK(hipGetLastError());
  return SK_OK;
}

int sk_psd_project(const double* H, int64_t n, int d, int method, double* out, sk_stream_t s) {
  SK_REQUIRE(n >= 0, "sk_psd_project: negative batch");
  SK_REQUIRE(method == SK_PSD_PROJ || method == SK_PSD_ABS || method == SK_PSD_NONE,
             "sk_psd_project: method must be proj or abs");
  if (n == 0) return SK_OK;
  SK_REQUIRE(H && out, "sk_psd_project: null pointer");
  const int g = grid_for(n, 64);
  hipStream_t st = as_stream(s);
  switch (d) {
    case 1: psd_project_kernel<1><<<g, 64, 0, st>>>(H, n, method, out); break;
    case 2: psd_project_kernel<2><<<g, 64, 0, st>>>(H, n, method, out); break;
    case 3: psd_project_kernel<3><<<g, 64, 0, st>>>(H, n, method, out); break;
    case 4: psd_project_kernel<4><<<g, 64, 0, st>>>(H, n, method, out); break;
    case 5: psd_project_kernel<5><<<g, 64, 0, st>>>(H, n, method, out); break;
    case 6: psd_project_kernel<6><<<g, 64, 0, st>>>(H, n, method, out); break;
    case 7: psd_project_kernel<7><<<g, 64, 0, st>>>(H, n, method, out); break;
    case 8: psd_project_kernel<8><<<g, 64, 0, st>>>(H, n, method, out); break;
    case 9: psd_project_kernel<9><<<g, 64, 0, st>>>(H, n, method, out); break;
    case 10: psd_project_kernel<10><<<g, 64, 0, st>>>(H, n, method, out); break;
    case 11: psd_project_kernel<11><<<g, 64, 0, st>>>(H, n, method, out); break;
    case 12: psd_project_kernel<12><<<g, 64, 0, st>>>(H, n, method, out); break;
    default: {
      SK_REQUIRE(d >= 1, "sk_psd_project: matrix size must be positive");
      DevBuf<double> A, V;
      SK_TRY(A.alloc((size_t)n * d * d));
      SK_TRY(V.alloc((size_t)n * d * d));
      psd_project_general_kernel<<<g, 64, 0, st>>>(H, n, d, method, A.p, V.p, out);
      SK_HIP_CHECK(hipGetLastError());
      SK_HIP_CHECK(hipStreamSynchronize(st));  // the scratch is freed on return
    }
  }
  SK_HIP_CHECK(hipGetLastError());
  return SK_OK;
}

int sk_polar_svd(const double* F, int64_t n, double* U, double* S, double* V, double* R, sk_stream_t s) {
  SK_REQUIRE(n >= 0, "sk_polar_svd: negative batch");
  if (n == 0) return SK_OK;
  SK_REQUIRE(F, "sk_polar_svd: null F");
  polar_svd_kernel<<<grid_for(n), kBlock, 0, as_stream(s)>>>(F, n, U, S, V, R);
  SK_HIP_CHECK(hipGetLastError());
  return SK_OK;
}

int sk_rotation_gradient(const double* F, int64_t n, double* dRdF, sk_stream_t s) {
  SK_REQUIRE(n >= 0, "sk_rotation_gradient: negative batch");
  if (n == 0) return SK_OK;
  SK_REQUIRE(F && dRdF, "sk_rotation_gradient: null pointer");
  rotation_gradient_kernel<<<grid_for(n), kBlock, 0, as_stream(s)>>>(F, n, dRdF);
  SK_HIP_CHECK(hipGetLastError());
  return SK_OK;
}

}  // extern "C"
