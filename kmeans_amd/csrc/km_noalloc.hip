// kmUpdateNoAlloc: the update of kmUpdate (counts, canonical FP64 means,
// transposed copy, norms) with its workspace in a fixed device array instead
// of a caller-supplied or library-allocated buffer, for callers that cannot
// allocate device memory (code under HIP graph capture, serving loops,
// harnesses that forbid hipMalloc).  The call allocates nothing and clears
// what must start at zero (the FP64 sums, counts and empty counter) with a
// kernel, not a memset.
//
// The array lives in this file's code object; each device that loads the
// library holds its own copy until the process ends.  Host code never uses
// the array's name as an address: the device address comes from
// hipGetSymbolAddress, taken once per device (with that device current) on
// the first call there, and is what the kernels receive.  All calls on one
// device share that device's array, so calls that overlap on one device
// (different streams, or threads) race; calls on one stream are ordered.
#include <mutex>

#include "km_common.hpp"

namespace km {

static __device__ __align__(256) char scratch[KM_NOALLOC_SCRATCH_BYTES];

#define KM_NOALLOC_MAX_DEVICES 64
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[KM_NOALLOC_MAX_DEVICES];

// the device address of `scratch` on the current device
static hipError_t scratch_address(void **out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= KM_NOALLOC_MAX_DEVICES) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!p) return hipErrorInvalidSymbol;
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return hipSuccess;
}

}  // namespace km

extern "C" kmeansCudaErrorStatus kmUpdateNoAlloc(void *stream, long n, int d, int k,
                                                 const float *X, const int *labels,
                                                 const float *C_old, float *C_new, float *Ct_new,
                                                 float *nc_new, long long *counts, int *empty) {
  if (n < 0 || d <= 0 || k <= 0 || !X || !labels || !C_old || !C_new)
    return ERROR_INVALID_ARGUMENT;
  const size_t need = km::update_ws_layout(n, d, k, nullptr, nullptr);
  if (need > (size_t)KM_NOALLOC_SCRATCH_BYTES) {
    fprintf(stderr,
            "kmeans_amd: kmUpdateNoAlloc(n=%ld, d=%d, k=%d) needs %zu bytes of workspace, more "
            "than its %ld-byte scratch array; use kmUpdate with a workspace\n",
            n, d, k, need, (long)KM_NOALLOC_SCRATCH_BYTES);
    return ERROR_INVALID_ARGUMENT;
  }
  void *base = nullptr;
  hipError_t e = km::scratch_address(&base);
  if (e != hipSuccess) {
    fprintf(stderr, "kmeans_amd: kmUpdateNoAlloc: hipGetSymbolAddress: %s\n", hipGetErrorString(e));
    return ERROR_UNSUPPORTED;
  }
  km::UpdateWs ws;
  km::update_ws_layout(n, d, k, (char *)base, &ws);
  e = km::launch_update((hipStream_t)stream, n, d, k, X, labels, C_old, C_new, Ct_new,
                        km::centers_ld(k), nc_new, counts, empty, ws);
  if (e != hipSuccess) {
    fprintf(stderr, "kmeans_amd: kmUpdateNoAlloc: %s\n", hipGetErrorString(e));
    return ERROR_CLUSTERCENTERS;
  }
  return NO_ERROR;
}
