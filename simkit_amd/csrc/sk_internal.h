// Internal structures shared by the mesh / x-tier and solver translation units.
#pragma once

#include "sk_common.h"

#include <type_traits>
#include <vector>

namespace sk {

// Owning device buffer.
template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n) {
    o.p = nullptr;
    o.n = 0;
  }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) {
      release();
      p = o.p;
      n = o.n;
      o.p = nullptr;
      o.n = 0;
    }
    return *this;
  }
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  int alloc(size_t count) {
    release();
    if (count == 0) return SK_OK;
    hipError_t e = hipMalloc(&p, count * sizeof(T));
    if (e != hipSuccess) {
      p = nullptr;
      set_error(std::string("hipMalloc failed: ") + hipGetErrorString(e));
      return SK_ERR_HIP;
    }
    n = count;
    return SK_OK;
  }
  int upload(const std::vector<T>& h) {
    SK_TRY(alloc(h.size()));
    if (!h.empty()) SK_HIP_CHECK(hipMemcpy(p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
    return SK_OK;
  }
};

// Upper-triangle 3x3-block index of local pair (i <= j) of a tet: 0..9. The
// blocks of one local row i are consecutive.
__host__ __device__ constexpr int ublock(int i, int j) { return i * (7 - i) / 2 + j; }

// The inverse of a vertex's masked diagonal block (the block-Jacobi
// preconditioner) is symmetric: the simulator stores its upper triangle
// (00, 01, 02, 11, 12, 22) at binv[kBinv v], read as three 2-wide loads.
constexpr int kBinv = 6;
constexpr int kBinvPQ[6][2] = {{0, 0}, {0, 1}, {0, 2}, {1, 1}, {1, 2}, {2, 2}};

template <typename T>
__device__ __forceinline__ void binv_load(const T* __restrict__ binv, int64_t v, T (&B)[9]) {
  using T2 = typename std::conditional<sizeof(T) == 8, double2, float2>::type;
  const T2* b2 = reinterpret_cast<const T2*>(binv + kBinv * v);
  const T2 u0 = b2[0], u1 = b2[1], u2 = b2[2];
  B[0] = u0.x;
  B[1] = B[3] = u0.y;
  B[2] = B[6] = u1.x;
  B[4] = u1.y;
  B[5] = B[7] = u2.x;
  B[8] = u2.y;
}

// Staged element stiffness: the 10 upper 3x3 blocks of K_e. Block u (ublock)
// of element e sits at R-unit offset kKstR<R> e + kBlkStride<R> u. FP64
// blocks are padded to 10 doubles (the padding is never written) so that the
// gather reads them with five 16-byte loads; FP32 blocks stay at 9 floats
// (padded to 12 the extra bytes cost more than the wider loads saved).
// Records are not padded to whole cache lines: at 48^3 that wrote 5% more
// bytes and took 6% longer in FP64.
template <typename R>
constexpr int kBlkStride = sizeof(R) == 8 ? 10 : 9;
template <typename R>
constexpr int kKstR = 10 * kBlkStride<R>;
constexpr int kKst = 100;  // staging doubles per element (the larger of the two layouts)
constexpr int kSellC = 16;       // rows per SELL chunk
constexpr int kSellLanes = 4;    // lanes per row in the SELL SpMV (kSellC * kSellLanes = wave size)
constexpr int kFst = 12;  // staged doubles per element: f_e = vol P D

// Grid-reduction scratch: per-block partial sums and result slots.
struct Reducer {
  DevBuf<double> partials;  // up to 8 values x kMaxBlocks
  DevBuf<double> result;    // kResults doubles; [kCandidateSlot, kResults): a step's candidate starts
  static constexpr int kMaxBlocks = 65536;
  static constexpr int kResults = 16, kCandidateSlot = 8;
  int init() {
    SK_TRY(partials.alloc(8 * (size_t)kMaxBlocks));
    SK_TRY(result.alloc(kResults));
    SK_HIP_CHECK(hipMemset(result.p, 0, kResults * sizeof(double)));
    return SK_OK;
  }
};

}  // namespace sk

struct sk_mesh {
  int64_t nv = 0, nt = 0, nblk = 0;
  sk::DevBuf<int32_t> T;      // SoA [4][nt]
  sk::DevBuf<double> D;       // SoA [12][nt], (4*c + v)*nt + e
  sk::DevBuf<int32_t> bptr;   // [nv+1] block-row pointer
  sk::DevBuf<int32_t> bcol;   // [nblk] sorted neighbour vertex per block
  sk::DevBuf<int32_t> brow;   // [nblk] row vertex per block
  sk::DevBuf<int32_t> emap;   // [16][nt] block of local pair (i, j)
  sk::DevBuf<int32_t> cptr;   // [nblk+1] contribution list per block
  sk::DevBuf<int32_t> clist;  // [16 nt] (e << 4) | (4 i + j), e ascending per block
  sk::DevBuf<int32_t> vptr;   // [nv+1] element incidences per vertex
  sk::DevBuf<int32_t> vlist;  // [4 nt] (e << 2) | i, e ascending per vertex
  // Sliced-ELL (SELL-C-sigma) layout of the simulator's matrix: rows sorted by
  // block count (descending), cut into chunks of kSellC rows; chunk c has
  // chunk_deg[c] slots per row and starts at slot chunk_off[c]. Slot
  // q = chunk_off[c] + s*kSellC + i holds block s of the chunk's row i; its
  // column vertex is scol[q] and its value m sits at
  // values[(q - i) * 9 + m * kSellC + i] (slot_index<kSellC, 1>:
  // component-major within a slot group), so each component load of a chunk
  // is one contiguous run. Padding slots have value 0.
  sk::DevBuf<int32_t> sell_row;   // [nchunk * kSellC] original row, -1 for padding
  sk::DevBuf<int32_t> chunk_deg;  // [nchunk]
  sk::DevBuf<int32_t> chunk_off;  // [nchunk + 1]
  sk::DevBuf<int32_t> scol;       // [nslot]
  sk::DevBuf<int32_t> sell_pos;   // [nblk] slot of each CSR block
  int64_t nchunk = 0, nslot = 0;
  // Chunk order of the SpMV kernels, balanced over the 8 XCDs: XCD j (the
  // hardware deals workgroup b to XCD b % 8) processes corder[j*lpad ..
  // (j+1)*lpad), the j-th eighth of every run of equal-degree chunks, so each
  // XCD gets the same share of heavy and light rows from one spatial region,
  // the runs merged by relative position within the list; -1 pads the lists
  // to the common length lpad (a multiple of 4).
  sk::DevBuf<int32_t> corder;
  // The same order with each entry's chunk metadata, one 16-byte load per
  // wave: (chunk, chunk_deg, chunk_off, 0); padding entries (-1, and chunk 0's
  // slot count and offset, so that addresses formed from them stay valid).
  sk::DevBuf<int4> cmeta;
  int64_t lpad = 0;
  sk::DevBuf<double> Kst;     // [nt][kKst] staged element stiffness blocks
  sk::DevBuf<double> fst;     // [nt][12] staged element forces
  sk::DevBuf<uint8_t> eflag;  // [nt] element pass: 1 where the closed-form eigensystem was not used
  sk::Reducer red;
  double* pinned = nullptr;   // 16 host-pinned doubles for scalar readback
};

namespace sk {

// Element-pass launchers shared by the x tier and the simulator.
int launch_elem_energy(sk_mesh* m, int material, const double* x, const double* F0, const double* dx, double t,
                       const double* mu,
                       const double* lam, const double* vol, double* elem_energy, const double* mass,
                       const double* xt, const double* fext, const uint8_t* fixed, double inv_h2,
                       hipStream_t st, double* result = nullptr);  // result: default red.result[0]
int launch_elem_grad(sk_mesh* m, int material, const double* x, const double* F0, const double* mu,
                     const double* lam, const double* vol, hipStream_t st);
// fp32: element Hessian, PSD projection and staged blocks in FP32 (staging
// buffer reinterpreted), otherwise FP64.
// active / pid (optional, together): skip elements whose problem (pid of the
// element's first vertex) has active == 0 (batched problems).
// fp64_count (optional): += elements whose FP32 Hessian of a polar density
// (ARAP, FCR) was evaluated in FP64 because det F < 1e-3.
// closed_form: densities with a closed-form Hessian eigensystem (stable
// neo-Hookean) modify eigenvalues from it instead of the 9x9 Jacobi when a
// PSD method is given (two launches; elements whose F is too close to
// singular for the SVD fall back to the Jacobi in the second).
int launch_elem_hess(sk_mesh* m, int material, const double* x, const double* F0, const double* mu,
                     const double* lam, const double* vol, int psd, bool fp32, hipStream_t st,
                     const int* active = nullptr, const int32_t* pid = nullptr,
                     unsigned long long* fp64_count = nullptr, bool closed_form = false);
// *result (device) = number of elements with det F(x) <= 0.
int launch_count_inverted(sk_mesh* m, const double* x, double* result, hipStream_t st);
// Gathers staged element forces into g; with mass != nullptr adds
// m (x - xt) * inv_h2 - fext, zeroes fixed DOFs and reduces ||g||^2 into red.result[0].
int launch_gather_grad(sk_mesh* m, double* g, const double* mass, const double* x, const double* xt,
                       const double* fext, const uint8_t* fixed, double inv_h2, hipStream_t st);
// Gathers staged element blocks into CSR values; with mass != nullptr adds
// mass * inv_h2 on the diagonal and writes the fixed-DOF-masked inverse of each
// diagonal 3x3 block into binv (its upper triangle, kBinv per vertex).
// layout: kLayoutCSR (scalar CSR of the sorted pattern), kLayoutSELL (the
// simulator's private SELL matrix) or kLayoutBSR (block b's 3x3 row-major at
// values[9 b]). fp32: values / binv are float arrays summed in FP32, otherwise
// double.
enum ValueLayout : int { kLayoutCSR = 0, kLayoutSELL = 1, kLayoutBSR = 2 };
int launch_gather_hess(sk_mesh* m, void* values, const double* mass, double inv_h2, const uint8_t* fixed,
                       void* binv, int layout, bool fp32, hipStream_t st, const int* active = nullptr,
                       const int32_t* pid = nullptr);
// Sliced layout of 9-component blocks in slot groups of C lanes (a power of
// two): slot q = G + i (G a multiple of C) keeps lane i's components 0-7 in
// vectors of VW components (VW = 1, or 16 bytes: 4 floats / 2 doubles) and
// component 8 alone,
//   component m < 8 at 9 G + (m / VW) C VW + VW i + m % VW,
//   component 8     at 9 G + 8 C + i,
// so that each load of a wave covers one contiguous run per slot group
// (VW = 1: component-major; VW = 4 for float: three loads per block).
template <int C, int VW>
__host__ __device__ inline int64_t slot_index(int64_t q, int m) {
  const int64_t i = q & (C - 1), b = 9 * (q - i);
  return m < 8 ? b + (m / VW) * C * VW + VW * i + m % VW : b + 8 * C + i;
}
// Lane i's block of the slot group starting at slot g (a multiple of C).
template <typename T, int C, int VW>
__device__ __forceinline__ void load_slot(const T* __restrict__ base, int64_t g, int i, T (&A)[9]) {
  const T* G = base + 9 * g;
  if constexpr (VW == 4 && sizeof(T) == 4) {
    const float4 q0 = *reinterpret_cast<const float4*>(G + 4 * i);
    const float4 q1 = *reinterpret_cast<const float4*>(G + 4 * C + 4 * i);
    A[0] = q0.x;
    A[1] = q0.y;
    A[2] = q0.z;
    A[3] = q0.w;
    A[4] = q1.x;
    A[5] = q1.y;
    A[6] = q1.z;
    A[7] = q1.w;
  } else {
    static_assert(VW == 1, "load_slot: VW is 1, or 4 for float");
#pragma unroll
    for (int m = 0; m < 8; ++m) A[m] = G[m * C + i];
  }
  A[8] = G[8 * C + i];
}
// Slot groups of the batch's per-problem matrix: one wave's rows, float4
// vectors of components.
constexpr int kQuadC = 64, kQuadVW = 4;
// The same gather into a caller's sliced layout of slot groups of kQuadC
// lanes: block b goes to slot pos[b], the diagonal inverse of vertex v to slot
// row_pos[v] of binv.
int launch_gather_hess_quad(sk_mesh* m, void* values, const double* mass, double inv_h2, const uint8_t* fixed,
                            void* binv, const int32_t* pos, const int32_t* row_pos, bool fp32, hipStream_t st,
                            const int* active = nullptr, const int32_t* pid = nullptr);
// Jacobi-preconditioned CG on a general CSR matrix (sk_pcg_csr) with
// caller-provided scratch: work holds csr_pcg_scratch_bytes(n) bytes,
// 64-byte aligned; nothing in it needs clearing. Launches at most maxit
// iterations; returns the sk_pcg_csr statuses.
size_t csr_pcg_scratch_bytes(int64_t n);
int run_csr_pcg(int64_t n, const int32_t* indptr, const int32_t* indices, const double* values, const double* b,
                double* x, double rtol, int64_t maxit, int64_t* iters, double* relres, void* work, hipStream_t st);
// Copies n doubles from device to the mesh's pinned buffer and waits.
int read_scalars(sk_mesh* m, const double* dev, int n, hipStream_t st);

}  // namespace sk
