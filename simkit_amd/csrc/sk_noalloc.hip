// Allocation-free entry point: the generic CSR PCG (sk_pcg_csr) on a static
// device workspace, for callers that cannot allocate device memory (code under
// stream capture, serving loops, harnesses that forbid hipMalloc).
//
// The workspace is one file-scope __device__ array, `scratch`, sized for
// kNoAllocMaxRows unknowns (25.2 MB); every GPU the library runs on holds its
// own copy until the process ends. In host code `scratch` is only a host-side
// stand-in for the symbol: its device address is looked up once per device
// (with that device current) by hipGetSymbolAddress and kept in g_scratch;
// kernels and copies only ever receive that address. All calls on one GPU
// share that GPU's workspace, so calls that overlap on one GPU (different
// streams or threads) race; nothing in it needs clearing between calls (the
// init launch writes every entry the PCG later reads).
#include "sk_internal.h"

#include <mutex>

namespace sk {
namespace {

constexpr int64_t kNoAllocMaxRows = 524288;  // 2^19: the 48^3 cube's 352,947 unknowns fit
// csr_pcg_scratch_bytes(kNoAllocMaxRows): the 64-byte PCG state, six vectors
// of n doubles, 2 * 2048 + 1024 reduction partials
constexpr size_t kScratchBytes = 64 + sizeof(double) * (6 * (size_t)kNoAllocMaxRows + 2 * 2048 + 1024);

__device__ __align__(256) char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_scratch[kMaxDevices] = {};

// Device address of the current device's workspace, looked up on the first
// call on that device.
int scratch_address(void** out) {
  int dev = -1;
  SK_HIP_CHECK(hipGetDevice(&dev));
  SK_REQUIRE(dev >= 0 && dev < kMaxDevices, "sk_pcg_csr_noalloc: device index out of range");
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    SK_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    SK_REQUIRE(p, "sk_pcg_csr_noalloc: no device address for the workspace");
    g_scratch[dev] = p;
  }
  *out = g_scratch[dev];
  return SK_OK;
}

}  // namespace
}  // namespace sk

using namespace sk;

extern "C" int sk_pcg_csr_noalloc(int64_t n, const int32_t* indptr, const int32_t* indices, const double* values,
                                  const double* b, double* x, double rtol, int64_t maxit, int64_t* iters,
                                  double* relres, sk_stream_t s) {
  SK_REQUIRE(n >= 0, "sk_pcg_csr_noalloc: negative size");
  SK_REQUIRE(n <= kNoAllocMaxRows, "sk_pcg_csr_noalloc: n exceeds the static workspace (524,288 unknowns); use "
                                   "sk_pcg_csr");
  SK_REQUIRE(rtol >= 0.0 && maxit >= 0, "sk_pcg_csr_noalloc: invalid rtol or maxit");
  if (iters) *iters = 0;
  if (relres) *relres = 0.0;
  if (n == 0) return SK_OK;
  SK_REQUIRE(indptr && indices && values && b && x, "sk_pcg_csr_noalloc: null pointer");
  SK_REQUIRE(csr_pcg_scratch_bytes(n) <= kScratchBytes, "sk_pcg_csr_noalloc: workspace layout mismatch");
  void* work = nullptr;
  SK_TRY(scratch_address(&work));
  return run_csr_pcg(n, indptr, indices, values, b, x, rtol, maxit, iters, relres, work,
                     reinterpret_cast<hipStream_t>(s));
}
