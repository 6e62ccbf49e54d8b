// Rest-measure and kinetic helpers: lumped mass, CSR SpMV and the
// backward-Euler kinetic terms on user arrays.
#include "sk_internal.h"

namespace sk {

// M_vv = (sum_{e ni v} vol_e rho_e) / 4 (massmatrix.py:40-49), gathered per
// vertex in ascending element order.
__global__ void lumped_mass_kernel(int64_t nv, const int32_t* vptr, const int32_t* vlist, const double* vol,
                                   const double* rho, double* mass) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nv) return;
  double s = 0.0;
  for (int32_t k = vptr[v]; k < vptr[v + 1]; ++k) {
    const int64_t e = vlist[k] >> 2;
    s += vol[e] * (rho ? rho[e] : 1.0);
  }
  mass[v] = s / 4.0;
}

constexpr int kLanes = 8;

__global__ void csr_spmv_kernel(int64_t n, const int32_t* indptr, const int32_t* indices, const double* vals,
                                const double* x, double* y) {
  const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t i = gid / kLanes;
  const int lane = (int)(gid % kLanes);
  double acc = 0.0;
  if (i < n)
    for (int32_t k = indptr[i] + lane; k < indptr[i + 1]; k += kLanes) acc += vals[k] * x[indices[k]];
#pragma unroll
  for (int off = kLanes / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, kLanes);
  if (i < n && lane == 0) y[i] = acc;
}

// d = x - x~ with x~ = x_c + h ((x_c - x_p) / h) (kinetic.py:87-90), or d = x
// when x_c is null; also v = (x_c - x_p)/h and x~ when requested.
__global__ void kinetic_prep_kernel(int64_t n, const double* x, const double* xc, const double* xp, double h,
                                    double* d, double* xt_out, double* v_out) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  if (!xc) {
    d[k] = x[k];
    return;
  }
  const double v = (xc[k] - xp[k]) / h;
  const double t = xc[k] + h * v;
  if (v_out) v_out[k] = v;
  if (xt_out) xt_out[k] = t;
  if (x) d[k] = x[k] - t;
}

// BDF2 target (kinetic.py bdf2_target / velocity_bdf2): d = x - x~ with
//   v_c = (3 x_c - 4 x_p + x_p2)/(2h), v_p = (3 x_p - 4 x_p2 + x_p3)/(2h),
//   x~ = (4/3) x_c - (1/3) x_p + (8h/9) v_c - (2h/9) v_p;
// also v_c and x~ when requested.
__global__ void kinetic_bdf2_prep_kernel(int64_t n, const double* x, const double* xc, const double* xp,
                                         const double* xp2, const double* xp3, double h, double* d, double* xt_out,
                                         double* v_out) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const double c = xc[k], p = xp[k], p2 = xp2[k], p3 = xp3[k];
  const double vc = (3.0 * c - 4.0 * p + p2) / (2.0 * h);
  const double vp = (3.0 * p - 4.0 * p2 + p3) / (2.0 * h);
  const double t = (4.0 / 3.0) * c - (1.0 / 3.0) * p + (8.0 * h / 9.0) * vc - (2.0 * h / 9.0) * vp;
  if (v_out) v_out[k] = vc;
  if (xt_out) xt_out[k] = t;
  if (x) d[k] = x[k] - t;
}

// grad = c (M d) (1/h^2); energy partials of d.(M d).
__global__ __launch_bounds__(kBlock) void kinetic_finish_kernel(int64_t n, const double* d, const double* Md,
                                                                double c, double inv_h2, double* grad,
                                                                double* partials) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double s = 0.0;
  if (k < n) {
    s = d[k] * Md[k];
    if (grad) grad[k] = c * Md[k] * inv_h2;
  }
  double v[1] = {s};
  store_block_partials<1>(v, partials);
}

}  // namespace sk

using namespace sk;

static hipStream_t as_stream(sk_stream_t s) { return reinterpret_cast<hipStream_t>(s); }

extern "C" {

int sk_mesh_lumped_mass(sk_mesh* m, const double* vol, const double* rho, double* mass, sk_stream_t s) {
  SK_REQUIRE(m && vol && mass, "sk_mesh_lumped_mass: null argument");
  lumped_mass_kernel<<<grid_for(m->nv), kBlock, 0, as_stream(s)>>>(m->nv, m->vptr.p, m->vlist.p, vol, rho, mass);
  SK_HIP_CHECK(hipGetLastError());
  return SK_OK;
}

int sk_csr_spmv(int64_t n, const int32_t* indptr, const int32_t* indices, const double* values, const double* x,
                double* y, sk_stream_t s) {
  SK_REQUIRE(n >= 0, "sk_csr_spmv: negative size");
  if (n == 0) return SK_OK;
  SK_REQUIRE(indptr && indices && values && x && y, "sk_csr_spmv: null pointer");
  csr_spmv_kernel<<<grid_for(n * kLanes), kBlock, 0, as_stream(s)>>>(n, indptr, indices, values, x, y);
  SK_HIP_CHECK(hipGetLastError());
  return SK_OK;
}

}  // extern "C"

// energy = 0.5 c d.(M d) / h^2 and grad = c (M d) / h^2 for a prepared d.
static int kinetic_finish(int64_t n, const double* d, double h, double c, const int32_t* M_indptr,
                          const int32_t* M_indices, const double* M_values, double* energy_host, double* grad,
                          sk_stream_t s) {
  hipStream_t st = as_stream(s);
  DevBuf<double> Md, partials, result;
  SK_TRY(Md.alloc(n));
  SK_TRY(sk_csr_spmv(n, M_indptr, M_indices, M_values, d, Md.p, s));
  const int g = grid_for(n);
  SK_REQUIRE(g <= Reducer::kMaxBlocks, "kinetic: vector too long");
  SK_TRY(partials.alloc(g));
  SK_TRY(result.alloc(1));
  const double inv_h2 = 1.0 / (h * h);
  kinetic_finish_kernel<<<g, kBlock, 0, st>>>(n, d, Md.p, c, inv_h2, grad, partials.p);
  finalize_sum_kernel<1><<<1, kBlock, 0, st>>>(partials.p, g, result.p);
  SK_HIP_CHECK(hipGetLastError());
  double e = 0.0;
  SK_HIP_CHECK(hipMemcpyAsync(&e, result.p, sizeof(double), hipMemcpyDeviceToHost, st));
  SK_HIP_CHECK(hipStreamSynchronize(st));
  if (energy_host) *energy_host = 0.5 * c * e * inv_h2;
  return SK_OK;
}

extern "C" {

int sk_kinetic(int64_t n, const double* x, const double* x_curr, const double* x_prev, double h, double c,
               const int32_t* M_indptr, const int32_t* M_indices, const double* M_values, double* energy_host,
               double* grad, double* x_tilde, double* velocity, sk_stream_t s) {
  SK_REQUIRE(n >= 0, "sk_kinetic: negative size");
  SK_REQUIRE(h > 0.0 || !x_curr, "sk_kinetic: h must be positive");
  SK_REQUIRE(!x_curr || x_prev, "sk_kinetic: x_prev required with x_curr");
  if (n == 0) {
    if (energy_host) *energy_host = 0.0;
    return SK_OK;
  }
  hipStream_t st = as_stream(s);
  const bool want_d = energy_host || grad;
  SK_REQUIRE(!want_d || (x && M_indptr && M_indices && M_values), "sk_kinetic: energy/gradient need x and M");
  DevBuf<double> d;
  SK_TRY(d.alloc(n));
  if (x_curr || x) {
    kinetic_prep_kernel<<<grid_for(n), kBlock, 0, st>>>(n, x, x_curr, x_prev, h, d.p, x_tilde, velocity);
    SK_HIP_CHECK(hipGetLastError());
  }
  if (!want_d) return SK_OK;
  return kinetic_finish(n, d.p, h, c, M_indptr, M_indices, M_values, energy_host, grad, s);
}

int sk_kinetic_bdf2(int64_t n, const double* x, const double* x_curr, const double* x_prev, const double* x_prev2,
                    const double* x_prev3, double h, const int32_t* M_indptr, const int32_t* M_indices,
                    const double* M_values, double* energy_host, double* grad, double* x_tilde, double* velocity,
                    sk_stream_t s) {
  SK_REQUIRE(n >= 0, "sk_kinetic_bdf2: negative size");
  SK_REQUIRE(h > 0.0, "sk_kinetic_bdf2: h must be positive");
  if (n == 0) {
    if (energy_host) *energy_host = 0.0;
    return SK_OK;
  }
  SK_REQUIRE(x_curr && x_prev && x_prev2 && x_prev3, "sk_kinetic_bdf2: the four history levels are required");
  const bool want_d = energy_host || grad;
  SK_REQUIRE(!want_d || (x && M_indptr && M_indices && M_values), "sk_kinetic_bdf2: energy/gradient need x and M");
  hipStream_t st = as_stream(s);
  DevBuf<double> d;
  SK_TRY(d.alloc(n));
  kinetic_bdf2_prep_kernel<<<grid_for(n), kBlock, 0, st>>>(n, x, x_curr, x_prev, x_prev2, x_prev3, h, d.p, x_tilde,
                                                            velocity);
  SK_HIP_CHECK(hipGetLastError());
  if (!want_d) return SK_OK;
  return kinetic_finish(n, d.p, h, 9.0 / 4.0, M_indptr, M_indices, M_values, energy_host, grad, s);
}

}  // extern "C"
