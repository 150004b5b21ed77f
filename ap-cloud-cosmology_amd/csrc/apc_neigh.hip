// GFD neighbour selection, solver.cpp:456-570 (searchneighboroctant) and its
// quicksort, solver.cpp:600-632.  One thread per computational/vacuum node;
// the candidate lists live in a chunked device scratch laid out element-major
// (element e of the chunk's thread t at e*rows + t) so a wave's accesses to
// the same element coalesce.  Built with -ffp-contract=off: the octant test
// and the squared distances must match the reference bit for bit.
//
// Against the reference's loop structure (same output):
//  * membership of a candidate is tested on 16-bit fingerprints of the list,
//    eight per 16-byte load, and confirmed on the stored index;
//  * octant and distance are computed only for candidates that join the list;
//  * the octant-ordered snapshot and the penalised sort keys are built in one
//    pass each from per-octant offsets held in registers;
//  * a first pass runs with a short candidate list (FAST_CAP); the few nodes
//    whose expansion outgrows it are queued and searched again with the full
//    bound;
//  * one launch of as many threads as the GPU keeps resident, each walking
//    the rows with a stride and reusing its own scratch slot (no launch
//    tails, scratch sized to residency).  The expansion lists are
//    element-major (the scans run in lockstep and coalesce); the sort keys
//    are thread-major, so a thread's partition scans stay in its own lines.
#include <cstdlib>

#include "apc_internal.h"
#include "apc_prims.h"

namespace apc {

struct Frame {
  double A[9];
  double len[3];
  int bc;
};

__device__ inline double pwrap(double x, double len, int bc) {
  if (bc == 1 && isfinite(x)) {  // a non-finite value would never leave the loop
    while (x >= (0.5 * len)) x -= len;
    while (x < -(0.5 * len)) x += len;
  }
  return x;
}

struct Cand {
  // element-major views
  uint32_t *mo;  // neighbour index | octant << 29
  double *d2;    // squared distance
  double *wd;    // sort keys
  int *wm;       // sort payload
  uint4 *fp;     // word k: 16-bit fingerprints of entries 8k..8k+7
  int rows, t, cap;  // rows: scratch slots (stride of the element-major arrays)
  __device__ uint32_t &MO(int e) const { return mo[(size_t)e * rows + t]; }
  __device__ double &D2(int e) const { return d2[(size_t)e * rows + t]; }
  __device__ double &WD(int e) const { return wd[(size_t)t * cap + e]; }
  __device__ int &WM(int e) const { return wm[(size_t)t * cap + e]; }
  __device__ uint4 FPW(int k) const { return fp[(size_t)k * rows + t]; }
  __device__ void set_fp(int e, uint32_t f) const {
    reinterpret_cast<uint16_t *>(fp + (size_t)(e >> 3) * rows + t)[e & 7] = (uint16_t)f;
  }
};

__device__ inline uint32_t fprint(uint32_t m) { return (m * 0x9E3779B1u) >> 16; }

// Is m among the first l list entries?
__device__ bool listed(const Cand &c, int l, uint32_t m) {
  const uint32_t f = fprint(m);
  const uint32_t ff = f | (f << 16);
  const int nw = (l + 7) >> 3;
  for (int k = 0; k < nw; ++k) {
    const uint4 w = c.FPW(k);
    const uint32_t h[4] = {w.x ^ ff, w.y ^ ff, w.z ^ ff, w.w ^ ff};
#pragma unroll
    for (int p = 0; p < 4; ++p) {
      const int e = 8 * k + 2 * p;
      if ((h[p] & 0xffffu) == 0 && e < l && (c.MO(e) & 0x1fffffffu) == m) return true;
      if ((h[p] >> 16) == 0 && e + 1 < l && (c.MO(e + 1) & 0x1fffffffu) == m) return true;
    }
  }
  return false;
}

// per-octant registers (constant indices only, so they stay in VGPRs)
__device__ inline int sel8(const int (&a)[8], int o) {
  int r = a[0];
#pragma unroll
  for (int q = 1; q < 8; ++q) r = (q == o) ? a[q] : r;
  return r;
}
__device__ inline void inc8(int (&a)[8], int o) {
#pragma unroll
  for (int q = 0; q < 8; ++q) a[q] += (q == o) ? 1 : 0;
}

// apartition / aquickSort on the W arrays, ranges [l, r].  Sub-ranges are
// disjoint, so processing them from an explicit stack (smaller side first)
// gives the same permutation as the reference's recursion.
__device__ int qpart(const Cand &c, int l, int r) {
  double pivot = c.WD(l);
  int i = l, j = r + 1;
  for (;;) {
    do ++i; while (i <= r && c.WD(i) <= pivot);
    do --j; while (c.WD(j) > pivot);
    if (i >= j) break;
    double t = c.WD(i); c.WD(i) = c.WD(j); c.WD(j) = t;
    int ti = c.WM(i); c.WM(i) = c.WM(j); c.WM(j) = ti;
  }
  double t = c.WD(l); c.WD(l) = c.WD(j); c.WD(j) = t;
  int ti = c.WM(l); c.WM(l) = c.WM(j); c.WM(j) = ti;
  return j;
}

// Each pushed range is at least as large as the side the loop continues
// with, so d stacked ranges need n >= 2^d elements: at most log2(n) + 1
// entries (the arrays are larger than that needs; measured, a 16-entry stack
// made the whole search 20% slower).
__device__ void qsort_ref(const Cand &c, int l, int r) {
  int sl[40], sr[40];
  int top = 0;
  sl[top] = l; sr[top] = r; ++top;
  while (top > 0) {
    --top;
    int lo = sl[top], hi = sr[top];
    while (lo < hi) {
      int j = qpart(c, lo, hi);
      if (j - lo < hi - j) {
        sl[top] = j + 1; sr[top] = hi; ++top;
        hi = j - 1;
      } else {
        sl[top] = lo; sr[top] = j - 1; ++top;
        lo = j + 1;
      }
    }
  }
}

constexpr int FAST_CAP = 256;  // multiple of 8 (fingerprint words)

// Neighbours of node i0 with the candidate lists in scratch slot c.t.
// Returns 0, 1 if the list would exceed cap and ovf is given (the node is
// appended to ovf, count in status[2]), or -1 after setting status.
// xyz: node centres packed (x, y, z, unused), one 32-byte load per node.
__device__ int search_node(int i0, const Cand &c, int M, int S, int lm, int cap, const Frame &F,
                           const double4 *__restrict__ xyz, const int *__restrict__ n24, const int *__restrict__ n24cnt,
                           int *__restrict__ nbr, double *__restrict__ dist,
                           int *__restrict__ status, int *__restrict__ ovf) {
  const int nq = 10 * lm;
  const double4 p0 = xyz[i0];
  const double x0 = p0.x, y0 = p0.y, z0 = p0.z;
  // wrapped offset to node m: octant in the rotated frame and squared
  // distance (solver.cpp:476-479, 595-598)
  auto place = [&](int m, double *dd) {
    const double4 pm = xyz[m];
    double ex = pwrap(pm.x - x0, F.len[0], F.bc);
    double ey = pwrap(pm.y - y0, F.len[1], F.bc);
    double ez = pwrap(pm.z - z0, F.len[2], F.bc);
    double xt = F.A[0] * ex + F.A[1] * ey + F.A[2] * ez;
    double yt = F.A[3] * ex + F.A[4] * ey + F.A[5] * ez;
    double zt = F.A[6] * ex + F.A[7] * ey + F.A[8] * ez;
    *dd = ex * ex + ey * ey + ez * ez;
    return (xt > 0) + 2 * (yt > 0) + 4 * (zt > 0);
  };
  int nn[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  int l = 0;
  // bucket the node's own Neigh24 (no de-duplication, solver.cpp:481-492)
  {
    const int cnt = n24cnt[i0];
    for (int q = 0; q < cnt; ++q) {
      int m = n24[(size_t)i0 * 24 + q];
      double dd;
      int o = place(m, &dd);
      c.MO(l) = (uint32_t)m | ((uint32_t)o << 29);
      c.D2(l) = dd;
      c.set_fp(l, fprint((uint32_t)m));
      inc8(nn, o);
      l++;
    }
  }
  // expand through the neighbours' Neigh24 until >= lm (solver.cpp:495-551)
  int ite = 0;
  while (l < lm) {
    // snapshot of the current list in octant order (stable within an octant)
    {
      int off[8];
      off[0] = 0;
#pragma unroll
      for (int q = 1; q < 8; ++q) off[q] = off[q - 1] + nn[q - 1];
      for (int e = 0; e < l; ++e) {
        const uint32_t v = c.MO(e);
        const int o = (int)(v >> 29);
        c.WM(sel8(off, o)) = (int)(v & 0x1fffffffu);
        inc8(off, o);
      }
    }
    const int lsnap = l;
    for (int a = 0; a < lsnap; ++a) {
      int src = c.WM(a);
      int cnt = (src < M) ? n24cnt[src] : 0;
      for (int q = 0; q < cnt; ++q) {
        int m = n24[(size_t)src * 24 + q];
        if (m >= S) { status[0] = APC_ERR_NEIGH_INVALID; return -1; }
        if (m == i0 || listed(c, l, (uint32_t)m)) continue;
        double dd;
        int o = place(m, &dd);
        if (sel8(nn, o) >= nq) { status[0] = APC_ERR_NEIGH_MANY; return -1; }
        if (l >= cap) {
          if (ovf) {
            ovf[atomicAdd(&status[2], 1)] = i0;
            return 1;
          }
          status[0] = APC_ERR_UNSUPPORTED;
          status[1] = i0;
          return -1;
        }
        c.MO(l) = (uint32_t)m | ((uint32_t)o << 29);
        c.D2(l) = dd;
        c.set_fp(l, fprint((uint32_t)m));
        inc8(nn, o);
        l++;
      }
    }
    ite++;
    if (ite > 5) { status[0] = APC_ERR_NEIGH_FEW; return -1; }
  }
  // The reference sorts td[k..k+nn[o]-1] but reads octant o's own slots
  // (solver.cpp:557-561), so only octant 0 is sorted; for o >= 1 the range
  // touches only octant 0's zero-initialised tail unless it reaches slot nq.
  {
    int ko = nn[0];
#pragma unroll
    for (int o = 1; o < 8; ++o) {
      if (nn[o] >= 2 && ko + nn[o] - 1 >= nq) {
        status[0] = APC_ERR_UNSUPPORTED;
        status[1] = i0;
        return -1;
      }
      ko += nn[o];
    }
  }
  // keys: octant 0 first (raw distances, insertion order), then octants 1..7
  // in insertion order with the rank penalty 1e4^k applied
  {
    int off[8];
    off[0] = 0;
#pragma unroll
    for (int q = 1; q < 8; ++q) off[q] = off[q - 1] + nn[q - 1];
    double pen[8] = {1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0};
    for (int e = 0; e < l; ++e) {
      const uint32_t v = c.MO(e);
      const int o = (int)(v >> 29);
      const int pos = sel8(off, o);
      inc8(off, o);
      double key = c.D2(e);
      if (o != 0) {
        double p = pen[1];
#pragma unroll
        for (int q = 2; q < 8; ++q) p = (q == o) ? pen[q] : p;
        key = p * key;
#pragma unroll
        for (int q = 1; q < 8; ++q)
          if (q == o) pen[q] = pen[q] * 10000.0;
      }
      c.WD(pos) = key;
      c.WM(pos) = (int)(v & 0x1fffffffu);
    }
  }
  // octant 0 quicksorted on the squared distance, then penalised
  if (nn[0] > 1) qsort_ref(c, 0, nn[0] - 1);
  {
    double pen = 1.0;
    for (int q = 0; q < nn[0]; ++q) { c.WD(q) = pen * c.WD(q); pen = pen * 10000.0; }
  }
  qsort_ref(c, 0, l - 1);
  for (int q = 0; q < lm; ++q) {
    const int m = c.WM(q);
    double dd;
    (void)place(m, &dd);
    nbr[(size_t)q * M + i0] = m;
    dist[(size_t)q * M + i0] = sqrt(dd);
  }
  return 0;
}

__global__ void k_pack_xyz(int n, const double *__restrict__ x, const double *__restrict__ y,
                           const double *__restrict__ z, double4 *__restrict__ xyz) {
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) xyz[i] = make_double4(x[i], y[i], z[i], 0.0);
}

// Rows q = 0..nrows-1 (node q, or list[q] when a list is given), strided over
// nslots threads; thread t owns scratch slot t.
__global__ void __launch_bounds__(256) k_search(int M, int S, int nrows, const int *__restrict__ list,
                                                int lm, int cap, int nslots, Frame F,
                                                const double4 *__restrict__ xyz,
                                                const int *__restrict__ n24,
                                                const int *__restrict__ n24cnt, uint32_t *mo,
                                                double *d2, double *wd, int *wm, uint4 *fpw,
                                                int *__restrict__ nbr, double *__restrict__ dist,
                                                int *__restrict__ status, int *__restrict__ ovf) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= nslots) return;
  const Cand c{mo, d2, wd, wm, fpw, nslots, t, cap};
  for (int q = t; q < nrows; q += nslots) {
    const int i0 = list ? list[q] : q;
    if (search_node(i0, c, M, S, lm, cap, F, xyz, n24, n24cnt, nbr, dist, status, ovf) < 0)
      return;
  }
}

int search_neighbors(apc_ctx *c) {
  if (!c->have_comp) return fail(c, APC_ERR_ARG, "neighbours: no computational particles");
  const int M = c->nc + c->nv, S = M + c->nb, lm = c->prm.lm;
  if (lm < 9) return fail(c, APC_ERR_ARG, "NumberOfNeighbors must be >= 9");
  if (lm > 1000) return fail(c, APC_ERR_UNSUPPORTED, "NumberOfNeighbors above 1000 not supported");
  APC_ALLOC(c, c->nbr, (size_t)lm * M + 1);
  APC_ALLOC(c, c->dist, (size_t)lm * M + 1);
  APC_ALLOC(c, c->status_dev, 4);
  APC_HIP(c, hipMemsetAsync(c->status_dev.p, 0, 4 * sizeof(int), c->s));
  Frame F;
  {  // solver.cpp:464-475, computed on the host
    double alpha = 0.01, beta = 0.001, gamma = 0.0001;
    F.A[0] = cos(alpha) * cos(gamma) - sin(alpha) * sin(beta) * sin(gamma);
    F.A[1] = -sin(alpha) * cos(beta);
    F.A[2] = -cos(alpha) * sin(gamma) - sin(alpha) * sin(beta) * cos(gamma);
    F.A[3] = cos(alpha) * sin(beta) * sin(gamma) + sin(alpha) * cos(gamma);
    F.A[4] = cos(alpha) * cos(beta);
    F.A[5] = cos(alpha) * sin(beta) * cos(gamma) - sin(alpha) * sin(gamma);
    F.A[6] = cos(beta) * sin(gamma);
    F.A[7] = -sin(beta);
    F.A[8] = cos(beta) * cos(gamma);
    const double *b = c->prm.bbox;
    F.len[0] = b[1] - b[0];
    F.len[1] = b[3] - b[2];
    F.len[2] = b[5] - b[4];
    F.bc = c->prm.bc;
  }
  // full candidate bound: an expansion round starts with fewer than lm
  // candidates, each adding at most 24 (the 2:1 bound), so 25*lm suffices
  const int cap_full = (25 * lm + 32 + 7) & ~7;
  // APC_NEIGH_FAST_CAP shortens the first-pass list (tests drive the second
  // pass with it); clamped to [32, FAST_CAP], multiple of 8
  int fast_cap = FAST_CAP;
  if (const char *v = getenv("APC_NEIGH_FAST_CAP"))
    fast_cap = std::min(FAST_CAP, std::max(32, atoi(v))) & ~7;
  if (c->search_slots == 0) {  // resident threads of k_search on this device
    int nb = 0, ncu = 0;
    APC_HIP(c, hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_search, 256, 0));
    APC_HIP(c, hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, c->dev));
    c->search_slots = std::max(1, nb) * std::max(1, ncu) * 256;
  }
  int slot_cap = c->search_slots;
  if (const char *v = getenv("APC_NEIGH_SLOTS")) slot_cap = std::max(256, atoi(v));
  const int slots = std::max(1, std::min(M, std::min(slot_cap, c->search_slots)));
  const int slots_full = std::max(1, std::min(M, 1 << 16));
  const size_t elems = std::max((size_t)fast_cap * slots, (size_t)cap_full * slots_full);
  APC_ALLOC(c, c->nscratch, 2 * elems);
  APC_ALLOC(c, c->nscratch_i, 2 * elems);
  APC_ALLOC(c, c->nscratch_f, elems / 2 + 4);  // 16-bit fingerprints as int
  APC_ALLOC(c, c->novf, (size_t)M + 1);
  uint32_t *mo = (uint32_t *)c->nscratch_i.p;
  int *wm = c->nscratch_i.p + elems;
  double *d2 = c->nscratch.p;
  double *wd = c->nscratch.p + elems;
  uint4 *fpw = (uint4 *)c->nscratch_f.p;
  APC_ALLOC(c, c->nxyz, 4 * (size_t)S + 4);
  double4 *xyz = (double4 *)c->nxyz.p;
  hipLaunchKernelGGL(k_pack_xyz, dim3(nblocks(S, 256)), dim3(256), 0, c->s, S, c->cx.p, c->cy.p,
                     c->cz.p, xyz);
  {
    KScope ks(c, "neighbor_search");
    hipLaunchKernelGGL(k_search, dim3(nblocks(slots, 256)), dim3(256), 0, c->s, M, S, M,
                       (const int *)nullptr, lm, fast_cap, slots, F, (const double4 *)xyz, c->n24.p, c->n24cnt.p, mo, d2, wd, wm, fpw, c->nbr.p, c->dist.p,
                       c->status_dev.p, c->novf.p);
  }
  int st[3] = {0, 0, 0};
  APC_HIP(c, hipMemcpyAsync(st, c->status_dev.p, sizeof(st), hipMemcpyDeviceToHost, c->s));
  APC_HIP(c, hipStreamSynchronize(c->s));
  // nodes whose expansion outgrew the short list, with the full bound
  if (st[0] == 0 && st[2] > 0) {
    const int rows = st[2];
    const int ns = std::min(rows, slots_full);
    KScope ks(c, "neighbor_search_long");
    hipLaunchKernelGGL(k_search, dim3(nblocks(ns, 256)), dim3(256), 0, c->s, M, S, rows,
                       (const int *)c->novf.p, lm, cap_full, ns, F, (const double4 *)xyz, c->n24.p, c->n24cnt.p, mo, d2, wd, wm, fpw, c->nbr.p, c->dist.p,
                       c->status_dev.p, (int *)nullptr);
  }
  if (st[2] > 0 && st[0] == 0) {
    APC_HIP(c, hipMemcpyAsync(st, c->status_dev.p, 2 * sizeof(int), hipMemcpyDeviceToHost, c->s));
    APC_HIP(c, hipStreamSynchronize(c->s));
  }
  c->nsearch_long = st[2];
  if (st[0]) {
    if (st[0] == APC_ERR_UNSUPPORTED)
      return fail(c, st[0], "neighbours: node %d needs more than the supported candidate list", st[1]);
    return fail(c, st[0], "neighbour search failed (%d)", st[0]);
  }
  c->have_nbr = 1;
  return 0;
}

}  // namespace apc

using namespace apc;

extern "C" int apc_neighbors(apc_ctx *c, int *nbr, double *dist) {
  if (!c) return APC_ERR_ARG;
  if (!c->have_nbr) {
    int r = search_neighbors(c);
    if (r) return r;
  }
  const int M = c->nc + c->nv, lm = c->prm.lm;
  std::vector<int> tn((size_t)M * lm);
  std::vector<double> td((size_t)M * lm);
  APC_HIP(c, hipMemcpyAsync(tn.data(), c->nbr.p, 4ull * M * lm, hipMemcpyDeviceToHost, c->s));
  APC_HIP(c, hipMemcpyAsync(td.data(), c->dist.p, 8ull * M * lm, hipMemcpyDeviceToHost, c->s));
  APC_HIP(c, hipStreamSynchronize(c->s));
  for (int i = 0; i < M; ++i)
    for (int q = 0; q < lm; ++q) {
      if (nbr) nbr[(size_t)i * lm + q] = tn[(size_t)q * M + i];
      if (dist) dist[(size_t)i * lm + q] = td[(size_t)q * M + i];
    }
  return 0;
}
