// Morton keys, stable LSD radix sort of (key, index) pairs, the move of the
// particle arrays into key order, and per-level octree node extraction.
// Built with -ffp-contract=off: keys must match the reference's FP64
// bisection bit for bit.
#include "apc_internal.h"
#include "apc_prims.h"

namespace apc {

// octree.cpp:676-852 — FP64 bisection, 3 bits per level (x=4, y=2, z=1),
// ">= midpoint" takes the upper half; a NaN coordinate matches none of the
// reference's eight cases and leaves the key unshifted.
__global__ void __launch_bounds__(256) k_keys(int n, const double *__restrict__ x,
                                              const double *__restrict__ y,
                                              const double *__restrict__ z, double bx0,
                                              double bx1, double by0, double by1, double bz0,
                                              double bz1, int depth, uint64_t *__restrict__ key,
                                              int *__restrict__ idx, int nbox = 1,
                                              const int *__restrict__ pstart = nullptr) {
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  double px = x[i], py = y[i], pz = z[i];
  double lx = bx0, rx = bx1, ly = by0, ry = by1, lz = bz0, rz = bz1;
  uint64_t k = 0;
  const bool finite = (px == px) && (py == py) && (pz == pz);
  for (int l = 1; l <= depth; ++l) {
    double mx = (lx + rx) / 2.0, my = (ly + ry) / 2.0, mz = (lz + rz) / 2.0;
    int bxb = px >= mx, byb = py >= my, bzb = pz >= mz;
    if (finite) k = (k << 3) + (uint64_t)((bxb << 2) | (byb << 1) | bzb);
    if (bxb) lx = mx; else rx = mx;
    if (byb) ly = my; else ry = my;
    if (bzb) lz = mz; else rz = mz;
  }
  if (nbox > 1) {
    // packed boxes: the box owning position i (pstart ascending) above the
    // Morton bits, so the sort groups the boxes and keeps each box's order
    int lo = 0, hi = nbox - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (pstart[mid] <= i) lo = mid;
      else hi = mid - 1;
    }
    k |= (uint64_t)lo << (3 * depth);
  }
  key[i] = k;
  idx[i] = i;
}

// ------------------------------------------------------------ radix sort --
// Tile = 256 threads x 16 keys, striped: key t + 256*it of the tile.
constexpr int RX_TILE = 4096;

__global__ void __launch_bounds__(256) k_rx_hist(const uint64_t *__restrict__ k, int n, int sh,
                                                 int *__restrict__ hist, int nblk) {
  __shared__ int h[256];
  const int t = threadIdx.x;
  h[t] = 0;
  __syncthreads();
  const long base = (long)blockIdx.x * RX_TILE;
#pragma unroll 4
  for (int it = 0; it < 16; ++it) {
    long i = base + it * 256 + t;
    if (i < n) atomicAdd(&h[(int)((k[i] >> sh) & 255)], 1);
  }
  __syncthreads();
  hist[t * nblk + blockIdx.x] = h[t];
}

// Stable scatter: a key's output position is the block offset of its digit
// plus the keys of the same digit earlier in this tile (earlier iterations,
// lower waves, lower lanes).  The tile is first placed in LDS in that order
// (the digit's tile-local offset from this tile's histogram plus the same
// rank) and then written out in LDS order, so consecutive threads store
// consecutive positions of a digit's run instead of one scattered word each.
__global__ void __launch_bounds__(256) k_rx_scatter(const uint64_t *__restrict__ kin,
                                                    const int *__restrict__ vin,
                                                    uint64_t *__restrict__ kout,
                                                    int *__restrict__ vout, int n, int sh,
                                                    const int *__restrict__ offs,
                                                    const int *__restrict__ hist, int nblk) {
  __shared__ uint64_t sk[RX_TILE];
  __shared__ int sv[RX_TILE];
  __shared__ int gbase[256];  // output offset of digit t for this tile
  __shared__ int lpre[256];   // tile-local offset of digit t
  __shared__ int run[256];    // keys of digit t placed by earlier iterations
  __shared__ int wcnt[4][256];
  const int t = threadIdx.x, w = t >> 6, lane = t & 63;
  gbase[t] = offs[t * nblk + blockIdx.x];
  {  // exclusive scan of this tile's digit counts over the block
    const int c0 = hist[t * nblk + blockIdx.x];
    int inc = c0;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int u = __shfl_up(inc, o, 64);
      if (lane >= o) inc += u;
    }
    if (lane == 63) wcnt[0][w] = inc;
    __syncthreads();
    int before = 0;
    for (int q = 0; q < w; ++q) before += wcnt[0][q];
    lpre[t] = before + inc - c0;
    run[t] = 0;
    __syncthreads();
  }
  const long tile = (long)blockIdx.x * RX_TILE;
  const int tn = (int)((n - tile) < RX_TILE ? (n - tile) : RX_TILE);
  for (int it = 0; it < 16; ++it) {
    long i = tile + it * 256 + t;
    bool valid = i < n;
    uint64_t kk = valid ? kin[i] : 0;
    int vv = valid ? vin[i] : 0;
    int d = (int)((kk >> sh) & 255);
#pragma unroll
    for (int q = 0; q < 4; ++q) wcnt[q][t] = 0;
    __syncthreads();
    unsigned long long peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      unsigned long long bal = __ballot((d >> b) & 1);
      peers &= ((d >> b) & 1) ? bal : ~bal;
    }
    int rank = below(peers);
    int cnt = __popcll(peers);
    if (valid && rank == cnt - 1) wcnt[w][d] = cnt;
    __syncthreads();
    if (valid) {
      int lp = lpre[d] + run[d] + rank;
      for (int q = 0; q < w; ++q) lp += wcnt[q][d];
      sk[lp] = kk;
      sv[lp] = vv;
    }
    __syncthreads();
    run[t] += wcnt[0][t] + wcnt[1][t] + wcnt[2][t] + wcnt[3][t];
    __syncthreads();
  }
  for (int e = t; e < tn; e += 256) {
    const uint64_t kk = sk[e];
    const int d = (int)((kk >> sh) & 255);
    const long g = (long)gbase[d] + (e - lpre[d]);
    kout[g] = kk;
    vout[g] = sv[e];
  }
}

// Sorts (k, v) by the low `bits` bits of k, stably.  Result ends in (k, v);
// k2/v2 are same-sized scratch.
int radix_sort_pairs(apc_ctx *c, uint64_t *k, int *v, uint64_t *k2, int *v2, int n, int bits) {
  if (n <= 1) return 0;
  int nblk = (n + RX_TILE - 1) / RX_TILE;
  APC_ALLOC(c, c->hist, (size_t)256 * nblk);
  APC_ALLOC(c, c->histx, (size_t)256 * nblk);
  uint64_t *ka = k, *kb = k2;
  int *va = v, *vb = v2;
  int passes = (bits + 7) / 8;
  for (int p = 0; p < passes; ++p) {
    int sh = 8 * p;
    {
      KScope ks(c, "radix_hist", 8.0 * n);
      hipLaunchKernelGGL(k_rx_hist, dim3(nblk), dim3(256), 0, c->s, ka, n, sh, c->hist.p, nblk);
    }
    int r = scan_exclusive(c, c->hist.p, c->histx.p, 256 * nblk, nullptr);
    if (r) return r;
    {
      KScope ks(c, "radix_scatter", 24.0 * n);
      hipLaunchKernelGGL(k_rx_scatter, dim3(nblk), dim3(256), 0, c->s, ka, va, kb, vb, n, sh,
                         c->histx.p, c->hist.p, nblk);
    }
    std::swap(ka, kb);
    std::swap(va, vb);
  }
  if (ka != k) {
    APC_HIP(c, hipMemcpyAsync(k, ka, sizeof(uint64_t) * n, hipMemcpyDeviceToDevice, c->s));
    APC_HIP(c, hipMemcpyAsync(v, va, sizeof(int) * n, hipMemcpyDeviceToDevice, c->s));
  }
  APC_HIP(c, hipGetLastError());
  return 0;
}

__global__ void k_copy_pairs(int n, const uint64_t *__restrict__ ka, const int *__restrict__ va,
                             uint64_t *__restrict__ kb, int *__restrict__ vb) {
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) {
    kb[i] = ka[i];
    vb[i] = va[i];
  }
}

// The keys and sort of sort_particles with every buffer supplied: kernels
// only, on stream s (a caller may capture it into a graph).  The passes
// alternate between (keys, idx) and (k2, v2); an odd pass count ends with a
// copy kernel back into (keys, idx).
hipError_t morton_sort_ws(hipStream_t s, int n, const double *x, const double *y, const double *z,
                          const double *bbox, int depth, uint64_t *keys, int *idx, uint64_t *k2,
                          int *v2, int *hist, int *histx, int *slab) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_keys, dim3(nblocks(n, 256)), dim3(256), 0, s, n, x, y, z, bbox[0], bbox[1],
                     bbox[2], bbox[3], bbox[4], bbox[5], depth, keys, idx, 1, (const int *)nullptr);
  if (n > 1) {
    const int nblk = (n + RX_TILE - 1) / RX_TILE;
    uint64_t *ka = keys, *kb = k2;
    int *va = idx, *vb = v2;
    const int passes = (3 * depth + 7) / 8;
    for (int p = 0; p < passes; ++p) {
      const int sh = 8 * p;
      hipLaunchKernelGGL(k_rx_hist, dim3(nblk), dim3(256), 0, s, ka, n, sh, hist, nblk);
      hipError_t e = scan_exclusive_ws(s, hist, histx, 256 * nblk, slab);
      if (e != hipSuccess) return e;
      hipLaunchKernelGGL(k_rx_scatter, dim3(nblk), dim3(256), 0, s, ka, va, kb, vb, n, sh, histx,
                         hist, nblk);
      std::swap(ka, kb);
      std::swap(va, vb);
    }
    if (ka != keys)
      hipLaunchKernelGGL(k_copy_pairs, dim3(nblocks(n, 256)), dim3(256), 0, s, n, ka, va, keys, idx);
  }
  return hipGetLastError();
}

// Gather the particle state into sorted order.
__global__ void __launch_bounds__(256) k_permute(int n, const int *__restrict__ p,
                                                 const double *__restrict__ a0,
                                                 const double *__restrict__ a1,
                                                 const double *__restrict__ a2,
                                                 const double *__restrict__ a3,
                                                 const double *__restrict__ a4,
                                                 const double *__restrict__ a5,
                                                 const int *__restrict__ ai, double *__restrict__ b0,
                                                 double *__restrict__ b1, double *__restrict__ b2,
                                                 double *__restrict__ b3, double *__restrict__ b4,
                                                 double *__restrict__ b5, int *__restrict__ bi) {
  int s = blockIdx.x * 256 + threadIdx.x;
  if (s >= n) return;
  int q = p[s];
  b0[s] = a0[q]; b1[s] = a1[q]; b2[s] = a2[q];
  b3[s] = a3[q]; b4[s] = a4[q]; b5[s] = a5[q];
  bi[s] = ai[q];
}

int sort_particles(apc_ctx *c) {
  const int n = c->N;
  if (n < 1) return fail(c, APC_ERR_ARG, "no particles");
  const apc_params &p = c->prm;
  APC_ALLOC(c, c->key, n);
  APC_ALLOC(c, c->key2, n);
  APC_ALLOC(c, c->perm, n);
  APC_ALLOC(c, c->perm2, n);
  {
    KScope ks(c, "keys", 36.0 * n);
    hipLaunchKernelGGL(k_keys, dim3(nblocks(n, 256)), dim3(256), 0, c->s, n, c->x.p, c->y.p,
                       c->z.p, p.bbox[0], p.bbox[1], p.bbox[2], p.bbox[3], p.bbox[4], p.bbox[5],
                       p.depth, c->key.p, c->perm.p, c->nbox, c->pstart.p);
  }
  int r = radix_sort_pairs(c, c->key.p, c->perm.p, c->key2.p, c->perm2.p, n,
                           3 * p.depth + c->boxbits);
  if (r) return r;
  APC_ALLOC(c, c->t0, n); APC_ALLOC(c, c->t1, n); APC_ALLOC(c, c->t2, n);
  APC_ALLOC(c, c->t3, n); APC_ALLOC(c, c->t4, n); APC_ALLOC(c, c->t5, n);
  APC_ALLOC(c, c->tid, n);
  {
    KScope ks(c, "permute", 108.0 * n);
    hipLaunchKernelGGL(k_permute, dim3(nblocks(n, 256)), dim3(256), 0, c->s, n, c->perm.p, c->x.p,
                       c->y.p, c->z.p, c->vx.p, c->vy.p, c->vz.p, c->id.p, c->t0.p, c->t1.p,
                       c->t2.p, c->t3.p, c->t4.p, c->t5.p, c->tid.p);
  }
  std::swap(c->x, c->t0); std::swap(c->y, c->t1); std::swap(c->z, c->t2);
  std::swap(c->vx, c->t3); std::swap(c->vy, c->t4); std::swap(c->vz, c->t5);
  std::swap(c->id, c->tid);
  APC_HIP(c, hipGetLastError());
  c->sorted = true;
  c->have_comp = false;
  return 0;
}

// ---------------------------------------------------- octree level nodes --
__global__ void k_level_flags(const uint64_t *__restrict__ k, int n, int sh, int *__restrict__ f) {
  int s = blockIdx.x * 256 + threadIdx.x;
  if (s >= n) return;
  f[s] = (s == 0) || ((k[s] >> sh) != (k[s - 1] >> sh));
}

__global__ void k_level_emit(const uint64_t *__restrict__ k, int n, int sh,
                             const int *__restrict__ f, const int *__restrict__ pos,
                             uint64_t *__restrict__ okey, int *__restrict__ ofirst) {
  int s = blockIdx.x * 256 + threadIdx.x;
  if (s >= n || !f[s]) return;
  int q = pos[s];
  okey[q] = (sh >= 64) ? 0 : ((k[s] >> sh) << sh);
  ofirst[q] = s;
}

__global__ void k_level_count(const int *__restrict__ first, int m, int n, int *__restrict__ np) {
  int q = blockIdx.x * 256 + threadIdx.x;
  if (q >= m) return;
  np[q] = ((q + 1 < m) ? first[q + 1] : n) - first[q];
}

}  // namespace apc

using namespace apc;

extern "C" int apc_octree_level(apc_ctx *c, int level, int *count, uint64_t *key, int *first,
                                int *np) {
  if (!c || !count) return APC_ERR_ARG;
  if (!c->sorted) return fail(c, APC_ERR_ARG, "apc_octree_level: octree not built");
  if (level < 0 || level > c->prm.depth) return fail(c, APC_ERR_ARG, "level out of range");
  const int n = c->N;
  int sh = 3 * (c->prm.depth - level);
  APC_ALLOC(c, c->nflag, n);
  APC_ALLOC(c, c->nrank, n);
  APC_ALLOC(c, c->dcount, 8);
  if (sh >= 64) {
    // level 0: only the root
    *count = 1;
    if (key) key[0] = 0;
    if (first) first[0] = 0;
    if (np) np[0] = n;
    return 0;
  }
  hipLaunchKernelGGL(k_level_flags, dim3(nblocks(n, 256)), dim3(256), 0, c->s, c->key.p, n, sh,
                     c->nflag.p);
  int r = scan_exclusive(c, c->nflag.p, c->nrank.p, n, c->dcount.p);
  if (r) return r;
  int m = 0;
  APC_HIP(c, hipMemcpyAsync(&m, c->dcount.p, sizeof(int), hipMemcpyDeviceToHost, c->s));
  APC_HIP(c, hipStreamSynchronize(c->s));
  *count = m;
  if (!key && !first && !np) return 0;
  DBuf<uint64_t> ok;
  DBuf<int> of, on;
  if (ok.ensure(m, c->s) != hipSuccess || of.ensure(m, c->s) != hipSuccess || on.ensure(m, c->s) != hipSuccess) {
    ok.release(c->s); of.release(c->s); on.release(c->s);
    return fail(c, APC_ERR_NOMEM, "apc_octree_level: allocation");
  }
  hipLaunchKernelGGL(k_level_emit, dim3(nblocks(n, 256)), dim3(256), 0, c->s, c->key.p, n, sh,
                     c->nflag.p, c->nrank.p, ok.p, of.p);
  hipLaunchKernelGGL(k_level_count, dim3(nblocks(m, 256)), dim3(256), 0, c->s, of.p, m, n, on.p);
  hipError_t e = hipSuccess;
  if (key && e == hipSuccess) e = hipMemcpyAsync(key, ok.p, 8ull * m, hipMemcpyDeviceToHost, c->s);
  if (first && e == hipSuccess) e = hipMemcpyAsync(first, of.p, 4ull * m, hipMemcpyDeviceToHost, c->s);
  if (np && e == hipSuccess) e = hipMemcpyAsync(np, on.p, 4ull * m, hipMemcpyDeviceToHost, c->s);
  if (e == hipSuccess) e = hipStreamSynchronize(c->s);
  ok.release(c->s); of.release(c->s); on.release(c->s);
  if (e != hipSuccess) return hip_fail(c, e, "apc_octree_level copies");
  return 0;
}

extern "C" int apc_octree_build(apc_ctx *c) {
  if (!c) return APC_ERR_ARG;
  if (c->frozen) return fail(c, APC_ERR_ARG, "octree_build: the topology is frozen");
  int r = sort_particles(c);
  if (r) return r;
  APC_HIP(c, hipGetLastError());
  return 0;
}

extern "C" int apc_get_sorted(apc_ctx *c, uint64_t *keys, int *index) {
  if (!c) return APC_ERR_ARG;
  if (!c->sorted) return fail(c, APC_ERR_ARG, "apc_get_sorted: octree not built");
  if (keys) APC_HIP(c, hipMemcpyAsync(keys, c->key.p, 8ull * c->N, hipMemcpyDeviceToHost, c->s));
  if (index) APC_HIP(c, hipMemcpyAsync(index, c->id.p, 4ull * c->N, hipMemcpyDeviceToHost, c->s));
  APC_HIP(c, hipStreamSynchronize(c->s));
  return 0;
}
