// apc_morton_sort_noalloc: Morton keys and their stable sort on caller device
// arrays with a workspace that is a file-scope device array of fixed size, so
// that the call performs no device allocation (for code under HIP graph
// capture and other callers that cannot allocate).
//
// The array `scratch` below is a host-side stand-in in host code: its device
// address differs per GPU and is obtained once per device with
// hipGetSymbolAddress (with that device current) and kept in noalloc_base[].
// Host code uses only that address (kernel arguments), never the array's name.
//
// Every call on one GPU shares that GPU's array and finds the previous call's
// contents in it; nothing in it needs to start at zero (each histogram column
// and scan partial is written before it is read).  Calls that overlap on one
// GPU (two host threads, or two streams in flight) race on it.
#include <mutex>

#include "apc_internal.h"

namespace apc {

// Layout (bytes, each part 256-byte aligned): k2 [N_max u64], v2 [N_max int],
// hist and histx [256 * N_max/4096 int each], the scan slab.
constexpr long NA_MAXN = APC_NOALLOC_MAX_N;
constexpr long NA_TILES = (NA_MAXN + 4095) / 4096;
constexpr long na_align(long b) { return (b + 255) & ~255L; }
constexpr long NA_K2 = 0;
constexpr long NA_V2 = NA_K2 + na_align(8 * NA_MAXN);
constexpr long NA_HIST = NA_V2 + na_align(4 * NA_MAXN);
constexpr long NA_HISTX = NA_HIST + na_align(4 * 256 * NA_TILES);
constexpr long NA_SLAB = NA_HISTX + na_align(4 * 256 * NA_TILES);
constexpr long NA_SLAB_INTS = 16384;
constexpr long NA_BYTES = NA_SLAB + na_align(4 * NA_SLAB_INTS);

static __device__ __align__(256) unsigned char scratch[NA_BYTES];

constexpr int NA_MAX_DEV = 64;
static std::mutex noalloc_mu;
static unsigned char *noalloc_base[NA_MAX_DEV];  // device address of `scratch` per GPU

// The device address of this GPU's scratch; looked up once per device.
static hipError_t noalloc_scratch(int dev, unsigned char **out) {
  std::lock_guard<std::mutex> g(noalloc_mu);
  if (!noalloc_base[dev]) {
    void *ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!ptr) return hipErrorInvalidSymbol;
    noalloc_base[dev] = (unsigned char *)ptr;
  }
  *out = noalloc_base[dev];
  return hipSuccess;
}

}  // namespace apc

using namespace apc;

extern "C" int apc_morton_sort_noalloc(int n, const double *x, const double *y, const double *z,
                                       const double *bbox, int depth, uint64_t *keys, int *index,
                                       void *stream) {
  if (n < 0 || !bbox || (n > 0 && (!x || !y || !z || !keys || !index)))
    return fail_noctx(APC_ERR_ARG, "apc_morton_sort_noalloc: bad argument");
  if (n > APC_NOALLOC_MAX_N)
    return fail_noctx(APC_ERR_ARG,
                      "apc_morton_sort_noalloc: n = %d exceeds its fixed workspace (at most %d "
                      "particles)", n, (int)APC_NOALLOC_MAX_N);
  if (depth < 1 || depth > 21)
    return fail_noctx(APC_ERR_ARG, "apc_morton_sort_noalloc: depth %d outside 1..21", depth);
  if (scan_slab_len(256 * ((n + 4095) / 4096)) > (size_t)NA_SLAB_INTS)
    return fail_noctx(APC_ERR_ARG, "apc_morton_sort_noalloc: scan workspace too small");
  if (n == 0) return APC_OK;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return fail_noctx(APC_ERR_HIP, "hipGetDevice: %s", hipGetErrorString(e));
  if (dev < 0 || dev >= NA_MAX_DEV)
    return fail_noctx(APC_ERR_UNSUPPORTED, "apc_morton_sort_noalloc: device %d", dev);
  unsigned char *base = nullptr;
  e = noalloc_scratch(dev, &base);
  if (e != hipSuccess)
    return fail_noctx(APC_ERR_HIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
  e = morton_sort_ws((hipStream_t)stream, n, x, y, z, bbox, depth, keys, index,
                     (uint64_t *)(base + NA_K2), (int *)(base + NA_V2), (int *)(base + NA_HIST),
                     (int *)(base + NA_HISTX), (int *)(base + NA_SLAB));
  if (e != hipSuccess)
    return fail_noctx(APC_ERR_HIP, "apc_morton_sort_noalloc: %s", hipGetErrorString(e));
  return APC_OK;
}
