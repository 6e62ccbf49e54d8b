// Multilayer perceptron with BF16 weights and activations and FP32
// accumulation on CDNA4 matrix cores: forward, MSE backward and Adam, all in
// hand-written kernels. Used by the learned Newton warm start (the network maps
// linear-modal coordinates of the previous two states and the boundary-motion
// parameters to the modal coordinates of the next state); no counterpart in
// the reference.
//
// Numerics (what the CPU oracle restates):
//   forward, layer l:  Z_l = bf16(A_l) bf16(W_l)^T + b_l   (FP32 accumulate)
//                      A_{l+1} = GELU(Z_l) (exact erf form), rounded to BF16;
//                      last layer: Y = Z_{L-1} (FP32, no activation);
//   loss:              mean over batch x outputs of (Y - T)^2;
//   backward:          dZ_{L-1} = 2 (Y - T) / (batch dL);
//                      dW_l = bf16(dZ_l)^T bf16(A_l), db_l = sum_b dZ_l (FP32);
//                      dA_l = bf16(dZ_l) bf16(W_l);  dZ_{l-1} = dA_l * GELU'(Z_{l-1});
//   Adam:              FP32 master weights and moments; the BF16 copy is
//                      refreshed after every update.
// Rounding to BF16 is round-to-nearest-even.
//
// Every GEMM runs 64 x 64 output tiles per workgroup (4 waves, 32 x 32 each:
// 2 x 2 tiles of v_mfma_f32_16x16x32_bf16), stepping K by 32: the workgroup
// stages the A and B tiles in LDS with 16-byte loads along whichever operand
// dimension is contiguous in memory (so transposed operands cost the same),
// storing both k-contiguous; lane l then reads its fragments with one 16-byte
// LDS load each: A[l % 16][8 (l / 16) + 0..7] and B[8 (l / 16) + 0..7][l % 16],
// and receives D[4 (l / 16) + 0..3][l % 16]. Feature and batch dimensions are
// padded to multiples of 16 with zeros; a K tail of 16 is zero-filled in LDS.
#include "sk_internal.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

typedef float sk_v4f __attribute__((ext_vector_type(4)));

namespace sk {
namespace {

__host__ __device__ inline uint16_t f2bf(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);  // quiet NaN
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

__host__ __device__ inline float bf2f(uint16_t b) {
  const uint32_t u = (uint32_t)b << 16;
  float f;
  memcpy(&f, &u, 4);
  return f;
}

__device__ inline float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }

__device__ inline float gelu_grad(float z) {
  return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) + z * 0.39894228040143268f * expf(-0.5f * z * z);
}

constexpr int kTile = 16;

struct GemmArgs {
  int M, N, K;                      // multiples of 16
  const uint16_t* A;                // A(m, k) = A[m * sam + k * sak]
  int64_t sam, sak;
  const uint16_t* B;                // B(k, n) = B[k * sbk + n * sbn]
  int64_t sbk, sbn;
  float* C;                         // C(m, n) = acc + bias[n], row stride ldc (may be null)
  int64_t ldc;
  const float* bias;                // may be null
  uint16_t* Cbf;                    // bf16(act(C)), row stride ldcb (may be null)
  int64_t ldcb;
  int gelu_act;                     // 1: Cbf holds GELU(C)
};

constexpr int kGT = 64;           // workgroup output tile, M and N
constexpr int kGK = 32;           // K step (one 16x16x32 MFMA)
constexpr int kGP = kGK + 8;      // LDS row pitch in BF16 (80 bytes: 16-byte aligned, rows on different banks)

typedef __bf16 sk_bf16x8 __attribute__((ext_vector_type(8)));
typedef uint32_t sk_u4 __attribute__((ext_vector_type(4)));

// S[r][c] = P[(r0 + r) srow + (k0 + c) sk] for r < 64, c < 32 (zero outside
// R x K). 256 threads, 8 elements each: 16-byte loads along k (sk == 1) or
// along r (srow == 1) when the addresses allow, else element by element.
// K is a multiple of 16, so an aligned run of 8 along k is all inside or all
// outside.
__device__ __forceinline__ void stage_tile(const uint16_t* __restrict__ P, int64_t srow, int64_t sk, int r0, int R,
                                           int k0, int K, bool vec_k, bool vec_r, uint16_t (*S)[kGP]) {
  const int t = threadIdx.x;
  if (vec_k) {
    const int r = t / 4, c = (t % 4) * 8;
    sk_u4 v = {0u, 0u, 0u, 0u};
    if (r0 + r < R && k0 + c < K) v = *reinterpret_cast<const sk_u4*>(P + (int64_t)(r0 + r) * srow + (k0 + c));
    *reinterpret_cast<sk_u4*>(&S[r][c]) = v;
  } else if (vec_r) {
    const int c = t / 8, r = (t % 8) * 8;
    sk_u4 v = {0u, 0u, 0u, 0u};
    if (r0 + r < R && k0 + c < K) v = *reinterpret_cast<const sk_u4*>(P + (int64_t)(k0 + c) * sk + (r0 + r));
    const uint16_t* h = reinterpret_cast<const uint16_t*>(&v);
#pragma unroll
    for (int q = 0; q < 8; ++q) S[r + q][c] = h[q];
  } else {
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int e = t * 8 + q, r = e / kGK, c = e % kGK;
      S[r][c] = (r0 + r < R && k0 + c < K) ? P[(int64_t)(r0 + r) * srow + (int64_t)(k0 + c) * sk] : (uint16_t)0;
    }
  }
}

__global__ __launch_bounds__(kBlock) void gemm_bf16_kernel(GemmArgs g) {
  __shared__ __align__(16) uint16_t As[kGT][kGP];  // [m][k]
  __shared__ __align__(16) uint16_t Bs[kGT][kGP];  // [n][k]
  const int tiles_n = (g.N + kGT - 1) / kGT;
  const int m0 = (blockIdx.x / tiles_n) * kGT, n0 = (blockIdx.x % tiles_n) * kGT;
  const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
  const int wm = (wid / 2) * 32, wn = (wid % 2) * 32;
  const int i = lane % 16, kq = lane / 16;
  // 16-byte staging needs 8-element runs at 16-byte aligned addresses
  auto aligned = [](const uint16_t* p, int64_t s) { return ((uintptr_t)p % 16 == 0) && (s % 8 == 0); };
  const bool a_k = g.sak == 1 && aligned(g.A, g.sam), a_r = !a_k && g.sam == 1 && aligned(g.A, g.sak);
  const bool b_k = g.sbk == 1 && aligned(g.B, g.sbn), b_r = !b_k && g.sbn == 1 && aligned(g.B, g.sbk);
  sk_v4f acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = sk_v4f{0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < g.K; k0 += kGK) {
    stage_tile(g.A, g.sam, g.sak, m0, g.M, k0, g.K, a_k, a_r, As);
    stage_tile(g.B, g.sbn, g.sbk, n0, g.N, k0, g.K, b_k, b_r, Bs);
    __syncthreads();
    sk_bf16x8 af[2], bfr[2];
#pragma unroll
    for (int a = 0; a < 2; ++a) af[a] = *reinterpret_cast<const sk_bf16x8*>(&As[wm + 16 * a + i][8 * kq]);
#pragma unroll
    for (int b = 0; b < 2; ++b) bfr[b] = *reinterpret_cast<const sk_bf16x8*>(&Bs[wn + 16 * b + i][8 * kq]);
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[a], bfr[b], acc[a][b], 0, 0, 0);
    __syncthreads();
  }
#pragma unroll
  for (int b = 0; b < 2; ++b) {
    const int n = n0 + wn + 16 * b + i;
    if (n >= g.N) continue;
    const float bn = g.bias ? g.bias[n] : 0.f;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t m = m0 + wm + 16 * a + 4 * kq + r;
        if (m >= g.M) continue;
        const float z = acc[a][b][r] + bn;
        if (g.C) g.C[m * g.ldc + n] = z;
        if (g.Cbf) g.Cbf[m * g.ldcb + n] = f2bf(g.gelu_act ? gelu(z) : z);
      }
  }
}

int launch_gemm(const GemmArgs& g, hipStream_t st) {
  SK_REQUIRE(g.M % kTile == 0 && g.N % kTile == 0 && g.K % kTile == 0 && g.M > 0 && g.N > 0 && g.K > 0,
             "bf16 GEMM: dimensions must be positive multiples of 16");
  const int64_t blocks = (int64_t)((g.M + kGT - 1) / kGT) * ((g.N + kGT - 1) / kGT);
  gemm_bf16_kernel<<<(unsigned)blocks, kBlock, 0, st>>>(g);
  SK_HIP_CHECK(hipGetLastError());
  return SK_OK;
}

// dst[b][k] = bf16(src[b][k]) for b < rows, k < cols; zero in the padding
// (b < prow, k < pcol).
__global__ void pack_bf16_kernel(int rows, int cols, int prow, int pcol, const float* src, uint16_t* dst) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)prow * pcol) return;
  const int b = (int)(e / pcol), k = (int)(e % pcol);
  dst[e] = (b < rows && k < cols) ? f2bf(src[(int64_t)b * cols + k]) : (uint16_t)0;
}

// dst = BF16 of the FP32 rounding of (X - mu) / sd (FP64 input, per-column
// mu and sd), zero in the padding: the input normalization of a predictor.
__global__ void pack_affine_bf16_kernel(int rows, int cols, int prow, int pcol, const double* X, const double* mu,
                                        const double* sd, uint16_t* dst) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)prow * pcol) return;
  const int b = (int)(e / pcol), k = (int)(e % pcol);
  dst[e] = (b < rows && k < cols) ? f2bf((float)((X[(int64_t)b * cols + k] - mu[k]) / sd[k])) : (uint16_t)0;
}

// out[b][n] = Y[b][n] sd[n] + mu[n] in FP64 for the unpadded block (the output
// de-normalization of a predictor).
__global__ void unpack_affine_kernel(int rows, int cols, int pcol, const float* Y, const double* sd,
                                     const double* mu, double* out) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)rows * cols) return;
  const int b = (int)(e / cols), n = (int)(e % cols);
  out[e] = (double)Y[(int64_t)b * pcol + n] * sd[n] + mu[n];
}

// out[b][n] = Y[b][n] for the unpadded block.
__global__ void unpack_kernel(int rows, int cols, int pcol, const float* Y, float* out) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)rows * cols) return;
  const int b = (int)(e / cols), n = (int)(e % cols);
  out[e] = Y[(int64_t)b * pcol + n];
}

// dZ = 2 (Y - T) / (rows cols) (zero in the padding), its BF16 copy, and per
// block partial sums of (Y - T)^2.
__global__ __launch_bounds__(kBlock) void mse_grad_kernel(int rows, int cols, int prow, int pcol, const float* Y,
                                                          const float* T, float* dZ, uint16_t* dZbf,
                                                          double* partials) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double sq = 0.0;
  if (e < (int64_t)prow * pcol) {
    const int b = (int)(e / pcol), n = (int)(e % pcol);
    float d = 0.f;
    if (b < rows && n < cols) {
      const float diff = Y[e] - T[(int64_t)b * cols + n];
      sq = (double)diff * (double)diff;
      d = 2.0f * diff / (float)((double)rows * cols);
    }
    dZ[e] = d;
    dZbf[e] = f2bf(d);
  }
  double v[1] = {sq};
  store_block_partials<1>(v, partials);
}

// dZ = dA * GELU'(Z) and its BF16 copy.
__global__ void gelu_backward_kernel(int64_t n, const float* dA, const float* Z, float* dZ, uint16_t* dZbf) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const float d = dA[e] * gelu_grad(Z[e]);
  dZ[e] = d;
  dZbf[e] = f2bf(d);
}

// db[n] = sum_b dZ[b][n] in batch order.
__global__ void colsum_kernel(int prow, int pcol, const float* dZ, float* db) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= pcol) return;
  float s = 0.f;
  for (int b = 0; b < prow; ++b) s += dZ[(int64_t)b * pcol + n];
  db[n] = s;
}

__global__ void adam_kernel(int64_t n, float* w, const float* g, float* m, float* v, uint16_t* wbf, float lr,
                            float b1, float b2, float eps, float c1, float c2) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const float gi = g[e];
  const float mi = b1 * m[e] + (1.f - b1) * gi;
  const float vi = b2 * v[e] + (1.f - b2) * gi * gi;
  m[e] = mi;
  v[e] = vi;
  const float wi = w[e] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  w[e] = wi;
  if (wbf) wbf[e] = f2bf(wi);
}

__global__ void to_bf16_kernel(int64_t n, const float* w, uint16_t* wbf) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) wbf[e] = f2bf(w[e]);
}

int pad16(int x) { return (x + kTile - 1) / kTile * kTile; }

}  // namespace
}  // namespace sk

using namespace sk;

struct sk_mlp {
  int L = 0;
  std::vector<int> dims, pdims;
  int max_batch = 0, pbatch = 0;
  // per layer: [pdims[l+1]][pdims[l]] weights, [pdims[l+1]] biases
  std::vector<DevBuf<float>> W, b, dW, db, mW, vW, mb, vb;
  std::vector<DevBuf<uint16_t>> Wbf;
  std::vector<DevBuf<uint16_t>> Abf;  // layer inputs [pbatch][pdims[l]], l = 0..L-1
  std::vector<DevBuf<float>> Z;       // pre-activations [pbatch][pdims[l+1]] (Z[L-1] = output)
  DevBuf<float> dA, dZ[2];
  DevBuf<uint16_t> dZbf[2];
  DevBuf<double> partials, loss;
  int64_t step = 0;
  bool have_grad = false;
};

namespace {

hipStream_t as_stream(sk_stream_t s) { return reinterpret_cast<hipStream_t>(s); }

// The layers on the packed input Abf[0] (batch rows, padded to pad16(batch)).
int mlp_layers(sk_mlp* n, int batch, hipStream_t st) {
  const int pb = pad16(batch);
  for (int l = 0; l < n->L; ++l) {
    GemmArgs g{};
    g.M = pb;
    g.N = n->pdims[l + 1];
    g.K = n->pdims[l];
    g.A = n->Abf[l].p;
    g.sam = n->pdims[l];
    g.sak = 1;
    g.B = n->Wbf[l].p;  // B(k, n) = W[n][k]
    g.sbk = 1;
    g.sbn = n->pdims[l];
    g.C = n->Z[l].p;
    g.ldc = n->pdims[l + 1];
    g.bias = n->b[l].p;
    if (l + 1 < n->L) {
      g.Cbf = n->Abf[l + 1].p;
      g.ldcb = n->pdims[l + 1];
      g.gelu_act = 1;
    }
    SK_TRY(launch_gemm(g, st));
  }
  return SK_OK;
}

int mlp_forward(sk_mlp* n, const float* X, int batch, hipStream_t st) {
  const int pb = pad16(batch);
  const int64_t e0 = (int64_t)pb * n->pdims[0];
  pack_bf16_kernel<<<grid_for(e0), kBlock, 0, st>>>(batch, n->dims[0], pb, n->pdims[0], X, n->Abf[0].p);
  SK_HIP_CHECK(hipGetLastError());
  return mlp_layers(n, batch, st);
}

}  // namespace

extern "C" {

int sk_gemm_bf16(int M, int N, int K, const uint16_t* A, int64_t sam, int64_t sak, const uint16_t* B, int64_t sbk,
                 int64_t sbn, float* C, int64_t ldc, sk_stream_t s) {
  SK_REQUIRE(A && B && C, "sk_gemm_bf16: null pointer");
  GemmArgs g{};
  g.M = M;
  g.N = N;
  g.K = K;
  g.A = A;
  g.sam = sam;
  g.sak = sak;
  g.B = B;
  g.sbk = sbk;
  g.sbn = sbn;
  g.C = C;
  g.ldc = ldc;
  return launch_gemm(g, as_stream(s));
}

int sk_mlp_create(int n_layers, const int* dims, int max_batch, sk_mlp** out) {
  SK_REQUIRE(out, "sk_mlp_create: null out");
  *out = nullptr;
  SK_REQUIRE(n_layers >= 1 && n_layers <= 16 && dims && max_batch >= 1, "sk_mlp_create: invalid shape");
  for (int l = 0; l <= n_layers; ++l) SK_REQUIRE(dims[l] >= 1 && dims[l] <= 4096, "sk_mlp_create: invalid width");
  sk_mlp* n = new sk_mlp();
  n->L = n_layers;
  n->dims.assign(dims, dims + n_layers + 1);
  for (int d : n->dims) n->pdims.push_back(pad16(d));
  n->max_batch = max_batch;
  n->pbatch = pad16(max_batch);
  auto run = [&]() -> int {
    const int L = n->L;
    for (auto* v : {&n->W, &n->b, &n->dW, &n->db, &n->mW, &n->vW, &n->mb, &n->vb, &n->Z}) v->resize(L);
    n->Wbf.resize(L);
    n->Abf.resize(L);
    int maxd = 0;
    for (int d : n->pdims) maxd = std::max(maxd, d);
    for (int l = 0; l < L; ++l) {
      const size_t nw = (size_t)n->pdims[l + 1] * n->pdims[l], nb = n->pdims[l + 1];
      for (auto* v : {&n->W[l], &n->dW[l], &n->mW[l], &n->vW[l]}) {
        SK_TRY(v->alloc(nw));
        SK_HIP_CHECK(hipMemset(v->p, 0, nw * sizeof(float)));
      }
      for (auto* v : {&n->b[l], &n->db[l], &n->mb[l], &n->vb[l]}) {
        SK_TRY(v->alloc(nb));
        SK_HIP_CHECK(hipMemset(v->p, 0, nb * sizeof(float)));
      }
      SK_TRY(n->Wbf[l].alloc(nw));
      SK_HIP_CHECK(hipMemset(n->Wbf[l].p, 0, nw * sizeof(uint16_t)));
      SK_TRY(n->Abf[l].alloc((size_t)n->pbatch * n->pdims[l]));
      SK_TRY(n->Z[l].alloc((size_t)n->pbatch * n->pdims[l + 1]));
    }
    SK_TRY(n->dA.alloc((size_t)n->pbatch * maxd));
    for (int i = 0; i < 2; ++i) {
      SK_TRY(n->dZ[i].alloc((size_t)n->pbatch * maxd));
      SK_TRY(n->dZbf[i].alloc((size_t)n->pbatch * maxd));
    }
    SK_TRY(n->partials.alloc(grid_for((int64_t)n->pbatch * n->pdims[L]) + 1));
    SK_TRY(n->loss.alloc(1));
    return SK_OK;
  };
  const int s = run();
  if (s != SK_OK) {
    delete n;
    return s;
  }
  *out = n;
  return SK_OK;
}

int sk_mlp_destroy(sk_mlp* n) {
  delete n;
  return SK_OK;
}

int64_t sk_mlp_num_params(const sk_mlp* n) {
  if (!n) return -1;
  int64_t c = 0;
  for (int l = 0; l < n->L; ++l) c += (int64_t)n->dims[l + 1] * n->dims[l] + n->dims[l + 1];
  return c;
}

int sk_mlp_set_params(sk_mlp* n, const float* params, sk_stream_t s) {
  SK_REQUIRE(n && params, "sk_mlp_set_params: null argument");
  hipStream_t st = as_stream(s);
  const float* p = params;
  for (int l = 0; l < n->L; ++l) {
    const int di = n->dims[l], dn = n->dims[l + 1], pi = n->pdims[l];
    std::vector<float> w((size_t)n->pdims[l + 1] * pi, 0.f), bb(n->pdims[l + 1], 0.f);
    for (int r = 0; r < dn; ++r)
      for (int c = 0; c < di; ++c) w[(size_t)r * pi + c] = *p++;
    for (int r = 0; r < dn; ++r) bb[r] = *p++;
    SK_HIP_CHECK(hipMemcpyAsync(n->W[l].p, w.data(), w.size() * sizeof(float), hipMemcpyHostToDevice, st));
    SK_HIP_CHECK(hipMemcpyAsync(n->b[l].p, bb.data(), bb.size() * sizeof(float), hipMemcpyHostToDevice, st));
    to_bf16_kernel<<<grid_for((int64_t)w.size()), kBlock, 0, st>>>((int64_t)w.size(), n->W[l].p, n->Wbf[l].p);
    SK_HIP_CHECK(hipGetLastError());
    for (auto* v : {&n->mW[l], &n->vW[l]}) SK_HIP_CHECK(hipMemsetAsync(v->p, 0, w.size() * sizeof(float), st));
    for (auto* v : {&n->mb[l], &n->vb[l]}) SK_HIP_CHECK(hipMemsetAsync(v->p, 0, bb.size() * sizeof(float), st));
    SK_HIP_CHECK(hipStreamSynchronize(st));
  }
  n->step = 0;
  n->have_grad = false;
  return SK_OK;
}

// which = 0: FP32 master parameters; 1: the last computed gradient.
static int mlp_export(sk_mlp* n, int which, float* out, hipStream_t st) {
  float* p = out;
  for (int l = 0; l < n->L; ++l) {
    const int di = n->dims[l], dn = n->dims[l + 1], pi = n->pdims[l];
    std::vector<float> w((size_t)n->pdims[l + 1] * pi), bb(n->pdims[l + 1]);
    const float* sw = which ? n->dW[l].p : n->W[l].p;
    const float* sb = which ? n->db[l].p : n->b[l].p;
    SK_HIP_CHECK(hipMemcpyAsync(w.data(), sw, w.size() * sizeof(float), hipMemcpyDeviceToHost, st));
    SK_HIP_CHECK(hipMemcpyAsync(bb.data(), sb, bb.size() * sizeof(float), hipMemcpyDeviceToHost, st));
    SK_HIP_CHECK(hipStreamSynchronize(st));
    for (int r = 0; r < dn; ++r)
      for (int c = 0; c < di; ++c) *p++ = w[(size_t)r * pi + c];
    for (int r = 0; r < dn; ++r) *p++ = bb[r];
  }
  return SK_OK;
}

int sk_mlp_get_params(sk_mlp* n, float* params, sk_stream_t s) {
  SK_REQUIRE(n && params, "sk_mlp_get_params: null argument");
  return mlp_export(n, 0, params, as_stream(s));
}

int sk_mlp_get_grad(sk_mlp* n, float* grad, sk_stream_t s) {
  SK_REQUIRE(n && grad, "sk_mlp_get_grad: null argument");
  SK_REQUIRE(n->have_grad, "sk_mlp_get_grad: no gradient computed yet");
  return mlp_export(n, 1, grad, as_stream(s));
}

int sk_mlp_predict(sk_mlp* n, const double* X, int batch, const double* in_mu, const double* in_sd,
                   const double* out_mu, const double* out_sd, double* Y, sk_stream_t s) {
  SK_REQUIRE(n && X && Y && in_mu && in_sd && out_mu && out_sd, "sk_mlp_predict: null argument");
  SK_REQUIRE(batch >= 1 && batch <= n->max_batch, "sk_mlp_predict: batch outside [1, max_batch]");
  hipStream_t st = as_stream(s);
  const int pb = pad16(batch);
  const int64_t e0 = (int64_t)pb * n->pdims[0];
  pack_affine_bf16_kernel<<<grid_for(e0), kBlock, 0, st>>>(batch, n->dims[0], pb, n->pdims[0], X, in_mu, in_sd,
                                                            n->Abf[0].p);
  SK_HIP_CHECK(hipGetLastError());
  SK_TRY(mlp_layers(n, batch, st));
  const int64_t e = (int64_t)batch * n->dims[n->L];
  unpack_affine_kernel<<<grid_for(e), kBlock, 0, st>>>(batch, n->dims[n->L], n->pdims[n->L], n->Z[n->L - 1].p,
                                                        out_sd, out_mu, Y);
  SK_HIP_CHECK(hipGetLastError());
  return SK_OK;
}

int sk_mlp_forward(sk_mlp* n, const float* X, int batch, float* Y, sk_stream_t s) {
  SK_REQUIRE(n && X && Y, "sk_mlp_forward: null argument");
  SK_REQUIRE(batch >= 1 && batch <= n->max_batch, "sk_mlp_forward: batch outside [1, max_batch]");
  hipStream_t st = as_stream(s);
  SK_TRY(mlp_forward(n, X, batch, st));
  const int64_t e = (int64_t)batch * n->dims[n->L];
  unpack_kernel<<<grid_for(e), kBlock, 0, st>>>(batch, n->dims[n->L], n->pdims[n->L], n->Z[n->L - 1].p, Y);
  SK_HIP_CHECK(hipGetLastError());
  return SK_OK;
}

int sk_mlp_backward(sk_mlp* n, const float* X, const float* T, int batch, double* loss_host, sk_stream_t s) {
  SK_REQUIRE(n && X && T, "sk_mlp_backward: null argument");
  SK_REQUIRE(batch >= 1 && batch <= n->max_batch, "sk_mlp_backward: batch outside [1, max_batch]");
  hipStream_t st = as_stream(s);
  SK_TRY(mlp_forward(n, X, batch, st));
  const int L = n->L, pb = pad16(batch);
  const int64_t eo = (int64_t)pb * n->pdims[L];
  const int go = grid_for(eo);
  int cur = 0;
  mse_grad_kernel<<<go, kBlock, 0, st>>>(batch, n->dims[L], pb, n->pdims[L], n->Z[L - 1].p, T, n->dZ[cur].p,
                                         n->dZbf[cur].p, n->partials.p);
  finalize_sum_kernel<1><<<1, kBlock, 0, st>>>(n->partials.p, go, n->loss.p);
  SK_HIP_CHECK(hipGetLastError());
  for (int l = L - 1; l >= 0; --l) {
    const int dn = n->pdims[l + 1], di = n->pdims[l];
    // dW[l] (dn x di) = dZ^T A_l: A(m, k) = dZ[k][m], B(k, n) = A_l[k][n]
    GemmArgs g{};
    g.M = dn;
    g.N = di;
    g.K = pb;
    g.A = n->dZbf[cur].p;
    g.sam = 1;
    g.sak = dn;
    g.B = n->Abf[l].p;
    g.sbk = di;
    g.sbn = 1;
    g.C = n->dW[l].p;
    g.ldc = di;
    SK_TRY(launch_gemm(g, st));
    colsum_kernel<<<(dn + kBlock - 1) / kBlock, kBlock, 0, st>>>(pb, dn, n->dZ[cur].p, n->db[l].p);
    SK_HIP_CHECK(hipGetLastError());
    if (l == 0) break;
    // dA (pb x di) = dZ W_l: A(m, k) = dZ[m][k], B(k, n) = W_l[k][n]
    GemmArgs h{};
    h.M = pb;
    h.N = di;
    h.K = dn;
    h.A = n->dZbf[cur].p;
    h.sam = dn;
    h.sak = 1;
    h.B = n->Wbf[l].p;
    h.sbk = di;
    h.sbn = 1;
    h.C = n->dA.p;
    h.ldc = di;
    SK_TRY(launch_gemm(h, st));
    const int64_t ez = (int64_t)pb * di;
    gelu_backward_kernel<<<grid_for(ez), kBlock, 0, st>>>(ez, n->dA.p, n->Z[l - 1].p, n->dZ[1 - cur].p,
                                                          n->dZbf[1 - cur].p);
    SK_HIP_CHECK(hipGetLastError());
    cur = 1 - cur;
  }
  n->have_grad = true;
  if (loss_host) {
    double sq = 0.0;
    SK_HIP_CHECK(hipMemcpyAsync(&sq, n->loss.p, sizeof(double), hipMemcpyDeviceToHost, st));
    SK_HIP_CHECK(hipStreamSynchronize(st));
    *loss_host = sq / ((double)batch * n->dims[L]);
  }
  return SK_OK;
}

int sk_mlp_adam(sk_mlp* n, float lr, float beta1, float beta2, float eps, sk_stream_t s) {
  SK_REQUIRE(n && n->have_grad, "sk_mlp_adam: no gradient computed");
  SK_REQUIRE(lr > 0.f && beta1 >= 0.f && beta1 < 1.f && beta2 >= 0.f && beta2 < 1.f && eps > 0.f,
             "sk_mlp_adam: invalid hyper-parameters");
  hipStream_t st = as_stream(s);
  n->step += 1;
  const float c1 = 1.f - std::pow(beta1, (float)n->step), c2 = 1.f - std::pow(beta2, (float)n->step);
  for (int l = 0; l < n->L; ++l) {
    const int64_t nw = (int64_t)n->pdims[l + 1] * n->pdims[l], nb = n->pdims[l + 1];
    adam_kernel<<<grid_for(nw), kBlock, 0, st>>>(nw, n->W[l].p, n->dW[l].p, n->mW[l].p, n->vW[l].p, n->Wbf[l].p, lr,
                                                 beta1, beta2, eps, c1, c2);
    adam_kernel<<<grid_for(nb), kBlock, 0, st>>>(nb, n->b[l].p, n->db[l].p, n->mb[l].p, n->vb[l].p, nullptr, lr,
                                                 beta1, beta2, eps, c1, c2);
  }
  SK_HIP_CHECK(hipGetLastError());
  return SK_OK;
}

}  // extern "C"
