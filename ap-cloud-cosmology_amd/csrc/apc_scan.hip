// Block/grid scan and reduction primitives (wave64, 256-thread blocks).
#include "apc_internal.h"
#include "apc_prims.h"

namespace apc {

// Exclusive scan of a 2048-element tile per block (8 per thread, blocked).
__global__ void __launch_bounds__(256) k_scan_tile(const int *__restrict__ in, int *__restrict__ out,
                                                   int n, int *__restrict__ partial) {
  __shared__ int tile[2048];
  __shared__ int ws[4];
  const int t = threadIdx.x;
  const long base = (long)blockIdx.x * 2048;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    long i = base + k * 256 + t;
    tile[k * 256 + t] = (i < n) ? in[i] : 0;
  }
  __syncthreads();
  int v[8];
  int sum = 0;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    v[k] = tile[t * 8 + k];
    sum += v[k];
  }
  int tot;
  int off = block_excl_scan256(sum, ws, &tot);
  int run = off;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    tile[t * 8 + k] = run;
    run += v[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    long i = base + k * 256 + t;
    if (i < n) out[i] = tile[k * 256 + t];
  }
  if (t == 0 && partial) partial[blockIdx.x] = tot;
}

__global__ void k_scan_add(int *__restrict__ out, int n, const int *__restrict__ poff) {
  long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i < n) out[i] += poff[i / 2048];
}

__global__ void k_scan_total(const int *in, const int *out, int n, int *total) {
  if (threadIdx.x == 0) *total = (n > 0) ? out[n - 1] + in[n - 1] : 0;
}

// Recursive three-phase scan; `slab` holds this level's partial sums and
// their scan, the next level's slab follows.
static void scan_rec(hipStream_t s, const int *in, int *out, int n, int *slab) {
  if (n <= 0) return;
  int nb = (n + 2047) / 2048;
  int *part = slab;
  int *poff = slab + nb + 8;
  hipLaunchKernelGGL(k_scan_tile, dim3(nb), dim3(256), 0, s, in, out, n, nb > 1 ? part : nullptr);
  if (nb > 1) {
    scan_rec(s, part, poff, nb, slab + 2 * (nb + 16));
    hipLaunchKernelGGL(k_scan_add, dim3(nblocks(n, 256)), dim3(256), 0, s, out, n, poff);
  }
}

size_t scan_slab_len(long n) {
  size_t need = 64;
  for (long m = n; m > 1;) {
    m = (m + 2047) / 2048;
    need += 2 * (m + 16);
  }
  return need;
}

hipError_t scan_exclusive_ws(hipStream_t s, const int *in, int *out, int n, int *slab) {
  scan_rec(s, in, out, n, slab);
  return hipGetLastError();
}

// `in` and `out` must not alias.
int scan_exclusive(apc_ctx *c, const int *in, int *out, int n, int *total_dev) {
  if (n <= 0) {
    if (total_dev) APC_HIP(c, hipMemsetAsync(total_dev, 0, sizeof(int), c->s));
    return 0;
  }
  APC_ALLOC(c, c->scratch_i, scan_slab_len(n));
  APC_HIP(c, scan_exclusive_ws(c->s, in, out, n, c->scratch_i.p));
  if (total_dev) hipLaunchKernelGGL(k_scan_total, dim3(1), dim3(64), 0, c->s, in, out, n, total_dev);
  APC_HIP(c, hipGetLastError());
  return 0;
}

// Max of non-negative doubles: the IEEE bit pattern of a non-negative double
// orders like the value, so an unsigned 64-bit atomic max is exact.
__global__ void __launch_bounds__(256) k_max_nonneg(const double *__restrict__ in, long n,
                                                    unsigned long long *out) {
  __shared__ double ws[4];
  double m = 0.0;
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256)
    m = fmax(m, in[i]);
  m = block_max256(m, ws);
  if (threadIdx.x == 0) atomicMax(out, (unsigned long long)__double_as_longlong(m));
}

int reduce_max_nonneg(apc_ctx *c, const double *in, int n, double *out_host) {
  APC_ALLOC(c, c->red_u64, 4);
  APC_HIP(c, hipMemsetAsync(c->red_u64.p, 0, sizeof(unsigned long long), c->s));
  int nb = (int)std::min<long>(1024, nblocks(n, 256));
  if (nb < 1) nb = 1;
  hipLaunchKernelGGL(k_max_nonneg, dim3(nb), dim3(256), 0, c->s, in, (long)n, c->red_u64.p);
  unsigned long long bits = 0;
  APC_HIP(c, hipMemcpyAsync(&bits, c->red_u64.p, sizeof(bits), hipMemcpyDeviceToHost, c->s));
  APC_HIP(c, hipStreamSynchronize(c->s));
  double v;
  memcpy(&v, &bits, sizeof(v));
  *out_host = v;
  return 0;
}

}  // namespace apc
