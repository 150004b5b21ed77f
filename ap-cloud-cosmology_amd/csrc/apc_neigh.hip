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
#include <array>
#include <climits>
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
      if ((unsigned)m >= (unsigned)S) { status[0] = APC_ERR_NEIGH_INVALID; return -1; }
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
        if (m < 0 || m >= S) { status[0] = APC_ERR_NEIGH_INVALID; return -1; }
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

// ---------------------------------------------------------------- one wave per node
// The same search with the 64 lanes of a wave working on one node, its lists
// in LDS:
//  * expansion: a round's candidates (snapshot entry, Neigh24 slot) are
//    flattened in the reference's order and taken 64 at a time; the lanes
//    decide together which join the list (not the node itself, not listed —
//    an LDS hash set — and not equal to an earlier lane's), appending in lane
//    order, so the list order is the reference's; errors are raised at the
//    first candidate (in that order) that hits one;
//  * sorts: the reference's quicksort, with every partition of a range longer
//    than WSMALL computed by ranks instead of the two scans; the short ranges
//    left over are disjoint and each is sorted by one lane with the scans.  For a range [lo, hi] with pivot p = WD[lo], let
//    g_1 < g_2 < ... be the positions in (lo, hi] holding a value > p and
//    l_1 > l_2 > ... the positions holding a value <= p, lo last.  The scans
//    swap g_k with l_k for every k with g_k < l_k (a prefix, K swaps), and the
//    pivot then lands at max(l_{K+1}, g_K) (l_1 if K = 0).  Disjoint ranges are
//    independent, so they are taken from a stack in any order.
//  Nodes whose list would exceed WCAP entries are queued for the per-thread
//  search with the full bound.
constexpr int WCAP = 256;   // list entries held in LDS (< 65536: ranges pack in 16 bits)
constexpr int SCAP = 128;   // lists up to this long are sorted by k_search_sort (codes < 256)
constexpr int WHASH = 512;  // membership hash slots (power of two, >= 2 WCAP)

struct WaveLds {
  int mo[WCAP];       // list: index | octant << 29 (insertion order); then WM (sort payload)
  double d2[WCAP];    // list: squared distance; then WD (sort keys)
  int snap[WCAP];     // snapshot in octant order; then the partition's g positions
  int lasc[WCAP];     // candidate offsets of a round; then the partition's <= positions
  int smallr[WCAP / 2];  // short sort ranges left to single lanes, lo | hi << 16
  union {
    int ht[WHASH];    // membership hash of the list's indices (-1 = empty)
    int stk[WHASH];   // pending sort ranges, lo | hi << 16; then per-lane stacks
  };
  int own[WHASH];     // per hash slot: the tag (chunk << 6 | lane) of the earliest
                      // candidate holding that index; 0 for the node's Neigh24
};

// 1e4^k by repeated multiplication from 1 (the reference's running penalty),
// k = 0..WCAP-1; written by the host before the search
__constant__ double c_pen[WCAP];


// exclusive prefix sum over the 64 lanes; *total gets the sum
__device__ inline int wave_excl_scan_i(int v, int *total) {
  const int lane = threadIdx.x & 63;
  int inc = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int t = __shfl_up(inc, o, 64);
    if (lane >= o) inc += t;
  }
  *total = __shfl(inc, 63, 64);
  return inc - v;
}

__device__ inline int hslot(uint32_t m) { return (int)((m * 0x9E3779B1u) >> 23) & (WHASH - 1); }

// the slot holding m, claimed (CAS from -1) if m is absent; lanes of one call
// with equal m meet in one slot.  The table holds the list (at most WCAP
// keys) and one chunk's claims (at most 64) in 2 WCAP slots, so a free slot
// always exists.
__device__ inline int ht_claim(int *ht, int m) {
  for (int h = hslot((uint32_t)m);; h = (h + 1) & (WHASH - 1)) {
    const int old = atomicCAS(&ht[h], -1, m);
    if (old < 0 || old == m) return h;
  }
}

// wrapped offset to a node at pm: octant in the rotated frame and squared
// distance
__device__ inline int place_p(const Frame &F, const double4 pm, double x0, double y0, double z0,
                              double *dd) {
  double ex = pwrap(pm.x - x0, F.len[0], F.bc);
  double ey = pwrap(pm.y - y0, F.len[1], F.bc);
  double ez = pwrap(pm.z - z0, F.len[2], F.bc);
  double xt = F.A[0] * ex + F.A[1] * ey + F.A[2] * ez;
  double yt = F.A[3] * ex + F.A[4] * ey + F.A[5] * ez;
  double zt = F.A[6] * ex + F.A[7] * ey + F.A[8] * ez;
  *dd = ex * ex + ey * ey + ez * ez;
  return (xt > 0) + 2 * (yt > 0) + 4 * (zt > 0);
}

// x held in a vector register from here on (an empty asm with a VGPR
// operand): the wave search's frame constants live in VGPRs, which its LDS
// budget leaves spare, instead of SGPRs, which it spills
__device__ inline double in_vgpr(double x) {
  asm volatile("" : "+v"(x));
  return x;
}

// the same for node m
__device__ inline int place_w(const Frame &F, const double4 *__restrict__ xyz, double x0, double y0,
                              double z0, int m, double *dd) {
  return place_p(F, xyz[m], x0, y0, z0, dd);
}

// The reference's quicksort on one short range by one lane (apartition /
// aquickSort, solver.cpp:600-632, as qpart / qsort_ref above), on the LDS
// arrays; stk: this lane's stack of WLSTK entries.
constexpr int WSMALL = 8;  // ranges up to this length go to single lanes (sweep: 6-8 fastest)
constexpr int WLSTK = 8;    // > log2(WSMALL) + 1 pending ranges per lane
static_assert(64 * WLSTK <= WHASH, "per-lane stacks fit the stack storage");

__device__ int lane_qpart(double *WD, int *WM, int l, int r) {
  const double pivot = WD[l];
  int i = l, j = r + 1;
  for (;;) {
    do ++i; while (i <= r && WD[i] <= pivot);
    do --j; while (WD[j] > pivot);
    if (i >= j) break;
    const double t = WD[i]; WD[i] = WD[j]; WD[j] = t;
    const int ti = WM[i]; WM[i] = WM[j]; WM[j] = ti;
  }
  const double t = WD[l]; WD[l] = WD[j]; WD[j] = t;
  const int ti = WM[l]; WM[l] = WM[j]; WM[j] = ti;
  return j;
}

__device__ void lane_qsort(double *WD, int *WM, int l, int r, int need, int *stk) {
  int top = 0;
  stk[top++] = l | (r << 16);
  while (top > 0) {
    const int pk = stk[--top];
    int lo = pk & 0xffff, hi = pk >> 16;
    while (lo < hi && lo < need) {
      const int j = lane_qpart(WD, WM, lo, hi);
      if (j - lo < hi - j) {
        stk[top++] = (j + 1) | (hi << 16);
        hi = j - 1;
      } else {
        stk[top++] = lo | ((j - 1) << 16);
        lo = j + 1;
      }
    }
  }
}

// the reference's quicksort of (WD, WM)[lo0..hi0], partitions by ranks; only
// positions below `need` are wanted: a range lying wholly at or above it is
// left unsorted (a range's elements stay within it, so nothing below changes)
__device__ void wave_qsort(WaveLds &L, int lo0, int hi0, int need) {
  const int lane = threadIdx.x & 63;
  double *WD = L.d2;
  int *WM = L.mo;
  int *gp = L.snap, *la = L.lasc;
  if (hi0 <= lo0) return;
  if (lane == 0) L.stk[0] = lo0 | (hi0 << 16);
  int top = 1, nsmall = 0;
  __syncthreads();
  while (top > 0) {
    --top;
    const int packed = L.stk[top];
    const int lo = packed & 0xffff, hi = packed >> 16;
    if (hi - lo + 1 <= WSMALL) {  // disjoint from every other range: any lane, any time
      if (lane == 0) L.smallr[nsmall] = packed;
      ++nsmall;
      __syncthreads();
      continue;
    }
    const double pivot = WD[lo];
    int jstar;
    {
    int nG = 0, nL = 0;
    for (int base = lo + 1; base <= hi; base += 64) {
      const int p = base + lane;
      const bool valid = p <= hi;
      const bool gt = valid && (WD[valid ? p : lo] > pivot);
      const bool le = valid && !gt;
      const uint64_t mG = __ballot(gt), mL = __ballot(le);
      if (gt) gp[nG + below(mG)] = p;
      if (le) la[nL + below(mL)] = p;
      nG += __popcll(mG);
      nL += __popcll(mL);
    }
    __syncthreads();
    const int kmax = min(nG, nL);
    int K = 0;
    for (int base = 0; base < kmax; base += 64) {
      const int k = base + lane + 1;
      const bool c = (k <= kmax) && (gp[k <= kmax ? k - 1 : 0] < la[k <= kmax ? nL - k : 0]);
      K += __popcll(__ballot(c));
    }
    const int lK1 = (K + 1 <= nL) ? la[nL - K - 1] : lo;
    jstar = (K > 0) ? max(lK1, gp[K - 1]) : lK1;
    for (int base = 0; base < K; base += 64) {
      const int k = base + lane + 1;
      if (k <= K) {
        const int a = gp[k - 1], b = la[nL - k];
        const double wa = WD[a], wb = WD[b];
        const int ma = WM[a], mb = WM[b];
        WD[a] = wb;
        WD[b] = wa;
        WM[a] = mb;
        WM[b] = ma;
      }
    }
    __syncthreads();
    if (lane == 0 && jstar != lo) {
      const double t = WD[lo];
      WD[lo] = WD[jstar];
      WD[jstar] = t;
      const int tm = WM[lo];
      WM[lo] = WM[jstar];
      WM[jstar] = tm;
    }
    }
    const bool left = jstar - 1 > lo, right = hi > jstar + 1 && jstar + 1 < need;
    if (lane == 0) {
      int t = top;
      if (left) L.stk[t++] = lo | ((jstar - 1) << 16);
      if (right) L.stk[t++] = (jstar + 1) | (hi << 16);
    }
    top += (left ? 1 : 0) + (right ? 1 : 0);
    __syncthreads();
  }
  for (int k = lane; k < nsmall; k += 64) {
    const int pk = L.smallr[k];
    lane_qsort(WD, WM, pk & 0xffff, pk >> 16, need, &L.stk[lane * WLSTK]);
  }
  __syncthreads();
}

// per-octant rank of every list entry and its position in octant order
// (octants ascending, insertion order within an octant); calls
// f(r, e, m, o, rank, pos) for entry e = 64 r + lane
//
// The lanes of one octant are found from three ballots of the octant's bits:
// lane j shares lane i's octant iff bit j of every bit-mask b_k equals bit k
// of i's octant, so i's mask is valid & ~((b0 ^ x0) | (b1 ^ x1) | (b2 ^ x2))
// with x_k all ones where i's bit k is set.  Lane q < 8 keeps octant q's
// offset (off[q] plus the octant's entries in earlier rows), which every
// lane reads at its own octant.
__device__ inline uint64_t same_octant(uint64_t bv, uint64_t b0, uint64_t b1, uint64_t b2,
                                       int o) {
  const uint32_t x0 = -(uint32_t)(o & 1), x1 = -(uint32_t)((o >> 1) & 1),
                 x2 = -(uint32_t)((o >> 2) & 1);
  const uint32_t lo = ((uint32_t)b0 ^ x0) | ((uint32_t)b1 ^ x1) | ((uint32_t)b2 ^ x2);
  const uint32_t hi = ((uint32_t)(b0 >> 32) ^ x0) | ((uint32_t)(b1 >> 32) ^ x1) |
                      ((uint32_t)(b2 >> 32) ^ x2);
  return bv & ~(((uint64_t)hi << 32) | lo);
}

template <class F>
__device__ inline void octant_order(const WaveLds &L, int l, const int (&off)[8], F f) {
  const int lane = threadIdx.x & 63;
  int offq = 0;  // lane q < 8: off[q]
#pragma unroll
  for (int q = 0; q < 8; ++q) offq = (lane == q) ? off[q] : offq;
  int runq = 0;  // lane q < 8: octant q's entries in the rows before
#pragma unroll
  for (int r = 0; r < WCAP / 64; ++r) {
    if (64 * r >= l) break;
    const int e = 64 * r + lane;
    const bool valid = e < l;
    const uint32_t v = valid ? (uint32_t)L.mo[e] : 0u;
    const int o = (int)(v >> 29);  // 0 on lanes past the list
    const uint64_t bv = __ballot(valid);
    const uint64_t b0 = __ballot(o & 1), b1 = __ballot(o & 2), b2 = __ballot(o & 4);
    const int base = __shfl((runq << 16) | offq, o, 64);  // run | off of octant o
    const int rank = (base >> 16) + below(same_octant(bv, b0, b1, b2, o));
    const int pos = (base & 0xffff) + rank;
    if (64 * (r + 1) < l) runq += __popcll(same_octant(bv, b0, b1, b2, lane & 7));
    if (valid) f(r, e, (int)(v & 0x1fffffffu), o, rank, pos);
  }
}

__global__ void __launch_bounds__(64) k_search_wave(int M, int S, int lm, int cap, int xchunk,
                                                    Frame Fk,
                                                    const double4 *__restrict__ xyz,
                                                    const int *__restrict__ n24,
                                                    const int *__restrict__ n24cnt,
                                                    int *__restrict__ nbr, double *__restrict__ dist,
                                                    int *__restrict__ status, int *__restrict__ ovf,
                                                    unsigned long long *__restrict__ visits,
                                                    uint8_t *__restrict__ scodes,
                                                    int *__restrict__ sidx,
                                                    uint8_t *__restrict__ slen) {
  __shared__ WaveLds L;
  Frame F;
#pragma unroll
  for (int k = 0; k < 9; ++k) F.A[k] = in_vgpr(Fk.A[k]);
#pragma unroll
  for (int k = 0; k < 3; ++k) F.len[k] = in_vgpr(Fk.len[k]);
  F.bc = Fk.bc;
  // Neigh24 entries this wave visited (the node's own and every expansion
  // round's candidates): the search's algorithmic traffic is 28 bytes per
  // visit (index + centre) plus its output
  unsigned long long nvis = 0;
  unsigned long long nlist = 0;  // list entries handed to k_search_sort
  unsigned long long nrank = 0;  // octant-0 rank codes handed to it
  const int lane = threadIdx.x;
  const int nq = 10 * lm;
  // Workgroups are dispatched round-robin over the 8 XCDs, each with its own
  // L2.  Nodes are dealt to the XCDs in chunks of xchunk consecutive nodes
  // (chunk k to XCD k mod 8) and the blocks of one XCD walk its chunks in
  // order, so the nodes in flight on an XCD are consecutive: their candidate
  // positions and Neigh24 lists overlap in its L2, and the output lines they
  // write (slot-major, consecutive nodes adjacent) fill within one L2.
  // Interleaving the chunks keeps the XCDs' work balanced over the box (one
  // contiguous eighth per XCD was slower: the dense regions are not spread
  // evenly).  A grid that is not a multiple of 8 strides over the nodes.
  const int nx = (gridDim.x % 8 == 0) ? 8 : 1;
  const int xcd = (int)(blockIdx.x % nx), per = (int)(gridDim.x / nx);
  const long nchunk = ((long)M + xchunk - 1) / xchunk;
  const long mine = (nchunk - xcd + nx - 1) / nx * xchunk;  // node slots of this XCD
  for (long k = blockIdx.x / nx; k < mine; k += per) {
    const int i0 = (int)(((k / xchunk) * nx + xcd) * xchunk + k % xchunk);
    if (i0 >= M) continue;  // the last chunk is short
    for (int h = lane; h < WHASH; h += 64) {
      L.ht[h] = -1;
      L.own[h] = INT_MAX;
    }
    const double4 p0 = xyz[i0];
    const double x0 = p0.x, y0 = p0.y, z0 = p0.z;
    int nn[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    __syncthreads();
    // the node's own Neigh24, all of it (solver.cpp:481-492)
    int l;
    {
      const int cnt = n24cnt[i0];
      nvis += (unsigned)cnt;
      int m = -1, o = 8;
      // an index outside the nodes (possible only in a caller-filled set) is
      // an error here; the reference would read past its arrays
      const bool bad = lane < cnt && (unsigned)n24[(size_t)i0 * 24 + lane] >= (unsigned)S;
      if (__ballot(bad)) {
        if (lane == 0) status[0] = APC_ERR_NEIGH_INVALID;
        return;
      }
      if (lane < cnt) {
        m = n24[(size_t)i0 * 24 + lane];
        double dd;
        o = place_w(F, xyz, x0, y0, z0, m, &dd);
        L.mo[lane] = m | (o << 29);
        L.d2[lane] = dd;
      }
      // in the list (duplicates, if any, share a slot): owner 0, before every chunk
      if (lane < cnt) L.own[ht_claim(L.ht, m)] = 0;
#pragma unroll
      for (int q = 0; q < 8; ++q) nn[q] += __popcll(__ballot(o == q));
      l = cnt;
    }
    __syncthreads();
    // expansion through the neighbours' Neigh24 until >= lm (solver.cpp:495-551)
    int ite = 0, err = 0;
    int ctag = 1;  // chunk counter of the node (owner tags; 0 is the Neigh24)
    bool overflow = false;
    while (l < lm) {
      int off[8];
      off[0] = 0;
#pragma unroll
      for (int q = 1; q < 8; ++q) off[q] = off[q - 1] + nn[q - 1];
      octant_order(L, l, off, [&](int, int, int m, int, int, int pos) { L.snap[pos] = m; });
      __syncthreads();
      // the round's candidates, (snapshot entry a, Neigh24 slot q) in the
      // reference's order, flattened: coff[a] = first candidate of entry a
      const int lsnap = l;
      int *coff = L.lasc;
      int T = 0;
      for (int base = 0; base < lsnap; base += 64) {
        const int a = base + lane;
        int cnt = 0;
        if (a < lsnap) {
          const int src = L.snap[a];
          cnt = (src < M) ? n24cnt[src] : 0;
        }
        int tot = 0;
        const int ex = wave_excl_scan_i(cnt, &tot);
        if (a < lsnap) coff[a] = T + ex;
        T += tot;
      }
      if (lane == 0) coff[lsnap] = T;
      nvis += (unsigned)T;
      __syncthreads();
      // candidate t's index (-1 past the round's end): Neigh24 slot t - coff[a]
      // of the last snapshot entry a with coff[a] <= t
      auto cand_of = [&](int t) {
        if (t >= T) return -1;
        int lo = 0, hi = lsnap;
        while (lo < hi) {
          const int mid = (lo + hi + 1) >> 1;
          if (coff[mid] <= t) lo = mid;
          else hi = mid - 1;
        }
        return n24[(size_t)L.snap[lo] * 24 + (t - coff[lo])];
      };
      auto pos_of = [&](int m) {
        return (m >= 0 && m < S) ? xyz[m] : make_double4(0.0, 0.0, 0.0, 0.0);
      };
      // The chunk's loads run ahead: chunk c+2's index and chunk c+1's
      // position are in flight while chunk c claims its slots, so a wave
      // does not wait on the chain (offset search, Neigh24 gather, position
      // gather) once per chunk.  Positions are read for every valid
      // candidate, fresh or not.
      int mC = cand_of(lane), mB = cand_of(64 + lane);
      double4 pC = pos_of(mC);
      for (int tb = 0; tb < T; tb += 64) {
        const int mA = cand_of(tb + 128 + lane);
        const double4 pB = pos_of(mB);
        const int t = tb + lane;
        const bool live = t < T;
        const int m = mC;
        const double4 pm = pC;
        mC = mB;
        pC = pB;
        mB = mA;
        const bool invalid = live && (m < 0 || m >= S);
        // new to the list: the index is in no earlier chunk (nor the node's
        // Neigh24) and no earlier lane of this chunk holds it.  Every
        // candidate claims its index's slot and offers its tag to the slot's
        // owner word; the smallest tag is the earliest candidate, and a slot
        // owned before this chunk holds a list member.
        const bool cand = live && !invalid && m != i0;
        const int tag = (ctag << 6) | lane;
        int slot = 0;
        if (cand) {
          slot = ht_claim(L.ht, m);
          atomicMin(&L.own[slot], tag);
        }
        bool fresh = cand && L.own[slot] == tag;
        ++ctag;
        int o = 8;
        double dd = 0.0;
        if (fresh) o = place_p(F, pm, x0, y0, z0, &dd);
        const uint64_t mf = __ballot(fresh);
        const int rank = below(mf);  // position among this chunk's new entries
        int add[8];                  // new entries per octant (wave-uniform; o is 8
                                     // for every lane that adds none)
#pragma unroll
        for (int q = 0; q < 8; ++q) add[q] = __popcll(__ballot(o == q));
        // the reference's checks in its order: invalid index, octant full, list
        // full.  An entry's same-octant predecessors are at most l + rank, so
        // with nq >= cap the list fills first: the node then takes the long
        // pass, which makes the checks with the full bound.
        int before = 0;  // same-octant entries before this one
        if (nq < cap) {
#pragma unroll
          for (int q = 0; q < 8; ++q) {
            const uint64_t mk = __ballot(o == q);
            if (o == q) before = nn[q] + below(mk);
          }
        }
        int code = 0;
        if (invalid) code = APC_ERR_NEIGH_INVALID;
        else if (fresh && nq < cap && before >= nq) code = APC_ERR_NEIGH_MANY;
        else if (fresh && l + rank >= cap) code = 1;  // overflow: longer list needed
        const uint64_t me = __ballot(code != 0);
        if (me) {
          const int fl = __ffsll((unsigned long long)me) - 1;
          const int c0 = __shfl(code, fl, 64);
          if (c0 == 1) overflow = true;
          else err = c0;
          break;
        }
        if (fresh) {
          L.mo[l + rank] = m | (o << 29);
          L.d2[l + rank] = dd;
        }
#pragma unroll
        for (int q = 0; q < 8; ++q) nn[q] += add[q];
        l += __popcll(mf);
        __syncthreads();
      }
      if (err || overflow) break;
      ite++;
      if (ite > 5) {
        err = APC_ERR_NEIGH_FEW;
        break;
      }
    }
    if (err) {
      if (lane == 0) status[0] = err;
      return;
    }
    if (overflow) {
      if (lane == 0) {
        ovf[atomicAdd(&status[2], 1)] = i0;
        slen[i0] = 0;  // searched again by k_search
      }
      __syncthreads();
      continue;
    }
    // the reference sorts octant o's range but reads octant 0's slots
    // (solver.cpp:557-561): ranges that reach slot nq are not supported
    {
      int ko = nn[0];
      bool bad = false;
#pragma unroll
      for (int o = 1; o < 8; ++o) {
        if (nn[o] >= 2 && ko + nn[o] - 1 >= nq) bad = true;
        ko += nn[o];
      }
      if (bad) {
        if (lane == 0) {
          status[0] = APC_ERR_UNSUPPORTED;
          status[1] = i0;
        }
        return;
      }
    }
    // keys: octant 0 first (raw distances), then octants 1..7, each in
    // insertion order with 1e4^rank; built in registers, then written over
    // the list storage
    {
      int off[8];
      off[0] = 0;
#pragma unroll
      for (int q = 1; q < 8; ++q) off[q] = off[q - 1] + nn[q - 1];
      double key[WCAP / 64];
      int km[WCAP / 64], kp[WCAP / 64];
#pragma unroll
      for (int r = 0; r < WCAP / 64; ++r) kp[r] = -1;
      octant_order(L, l, off, [&](int r, int e, int m, int o, int rank, int pos) {
        const double d = L.d2[e];
        key[r] = (o != 0) ? c_pen[rank] * d : d;
        km[r] = m;
        kp[r] = pos;
      });
      __syncthreads();
#pragma unroll
      for (int r = 0; r < WCAP / 64; ++r)
        if (kp[r] >= 0) {
          L.d2[kp[r]] = key[r];
          L.mo[kp[r]] = km[r];
        }
      __syncthreads();
    }
    // octant 0 quicksorted on the squared distance, then penalised.  The key
    // at sorted position p is 1e4^p times the p-th smallest distance,
    // whatever the order among equal distances, so each lane places its key
    // at its rank (the keys below it, then the equal keys before it in list
    // order).  The indices: with distinct keys (57% of the nodes with two or
    // more at 256^3) the quicksort's order is the sorted order and each lane
    // places its index with its key.  With equal keys their order is the
    // quicksort's own: the indices stay in list order and k_search_sort
    // replays the quicksort (one lane per node) on the rank codes written
    // after the node's list codes, from byte r0 of its code row.  An octant
    // over 64 entries, or a list sorted here (over SCAP) or with no room for
    // the codes, takes the wave sort's exact replay.  (The final lists, where
    // each octant's first entry keeps its raw distance, nearly always hold
    // equal keys; their replay is k_search_sort's.)
    const int r0 = (l + 7) & ~7;
    int defer = 0;  // octant-0 entries whose order k_search_sort replays
    int rc = 0;     // rank code of this lane's octant-0 entry (lane < defer)
    {
      const int n0 = nn[0];
      bool placed = false;
      if (n0 > 1 && n0 <= 64) {
        double ke = 0.0;
        int me = 0, lt = 0, eq = 0, eb = 0;
        if (lane < n0) {
          ke = L.d2[lane];
          me = L.mo[lane];
          for (int j = 0; j < n0; ++j) {
            const double kj = L.d2[j];
            lt += (kj < ke) ? 1 : 0;
            eq += (kj == ke) ? 1 : 0;
            eb += (kj == ke && j < lane) ? 1 : 0;
          }
        }
        const bool ties = __ballot(lane < n0 && eq > 1) != 0ull;
        if (!ties || (l <= SCAP && r0 + n0 <= SCAP)) {
          __syncthreads();
          if (lane < n0) {
            L.d2[lt + eb] = c_pen[lt + eb] * ke;
            if (!ties) L.mo[lt] = me;
          }
          rc = lt;
          defer = ties ? n0 : 0;
          placed = true;
        }
      }
      if (!placed) {
        if (n0 > 1) wave_qsort(L, 0, n0 - 1, n0);
        for (int q = lane; q < n0; q += 64) L.d2[q] = c_pen[q] * L.d2[q];
      }
    }
    __syncthreads();
    if (l <= SCAP) {
      // The final quicksort goes to k_search_sort, one lane per node.  It
      // compares only the order of the keys, so each key is replaced by the
      // number of keys strictly below it: equal keys get equal codes, a
      // larger key a larger code, and the quicksort's permutation of the
      // codes is its permutation of the keys.
      const double k0 = (lane < l) ? L.d2[lane] : 0.0;
      const double k1 = (64 + lane < l) ? L.d2[64 + lane] : 0.0;
      int c0 = 0, c1 = 0;
      if (l <= 64) {
        for (int j = 0; j < l; ++j) c0 += (L.d2[j] < k0) ? 1 : 0;
      } else {
        for (int j = 0; j < l; ++j) {
          const double kj = L.d2[j];
          c0 += (kj < k0) ? 1 : 0;
          c1 += (kj < k1) ? 1 : 0;
        }
      }
      uint8_t *oc = scodes + (size_t)i0 * SCAP;
      int *oi = sidx + (size_t)i0 * SCAP;
      if (lane < l) {
        oc[lane] = (uint8_t)c0;
        oi[lane] = L.mo[lane];
      }
      if (64 + lane < l) {
        oc[64 + lane] = (uint8_t)c1;
        oi[64 + lane] = L.mo[64 + lane];
      }
      if (lane < defer) oc[r0 + lane] = (uint8_t)rc;
      if (lane == 0) {
        slen[i0] = (uint8_t)l;
        slen[(size_t)M + i0] = (uint8_t)defer;
      }
      nlist += (unsigned)l;
      nrank += (unsigned)defer;
    } else {
      wave_qsort(L, 0, l - 1, lm);  // positions 0..lm-1 are the output
      for (int q = lane; q < lm; q += 64) {
        const int m = L.mo[q];
        double dd;
        (void)place_w(F, xyz, x0, y0, z0, m, &dd);
        nbr[(size_t)q * M + i0] = m;
        dist[(size_t)q * M + i0] = sqrt(dd);
      }
      if (lane == 0) slen[i0] = 0;  // done here
    }
    __syncthreads();
  }
  if (lane == 0 && nvis) atomicAdd(visits, nvis);  // one atomic per wave
  if (lane == 0 && nlist) atomicAdd(visits + 1, nlist);
  if (lane == 0 && nrank) atomicAdd(visits + 2, nrank);
}

// The final quicksort of k_search_wave's lists of at most SCAP entries, one
// lane per node.  Each lane keeps its node's codes (order-equivalent keys,
// all below 128) and the list positions they carry in LDS rows of its own,
// and runs the reference's partition scans (apartition, solver.cpp:600-632)
// eight codes per 64-bit read: for codes b, p < 128, b + (127 - p) sets bit 7
// of its byte exactly when b > p and never carries into the next byte.
// Ranges lying wholly at or above lm are left unsorted (nothing below them
// changes).  The output rows are written slot-major by consecutive lanes,
// i.e. coalesced.
//
// The rows are packed: a lane's row holds its list's length rounded up to 8
// bytes, at the exclusive prefix sum of those sizes over the wave, and a
// workgroup sorts the lanes whose rows fit in SBUD bytes per array in one
// pass, the rest in further passes.  Lists average 47 entries at 256^3
// (78% below 64), so nearly every wave takes one pass with a quarter of the
// LDS that rows of SCAP entries would hold, and twice the waves per CU.
constexpr int SBUD = 4096;  // bytes per array (codes, positions) of one pass
constexpr int SSTK = 12;    // > log2(SCAP) + 1 pending ranges per lane
static_assert(SBUD >= SCAP, "every row fits a pass");

// eight codes per read: rows start at multiples of 8 bytes and hold their
// length rounded up to 8, so a scan's words stay in its row
__device__ inline uint64_t row_u64(const uint8_t *row, int d) {
  return *reinterpret_cast<const uint64_t *>(row + 8 * d);
}
constexpr uint64_t kHigh8 = 0x8080808080808080ull;

// first x in [x0, r] with a[x] > p, or r + 1 (the reference's i-scan)
__device__ inline int scan_gt(const uint8_t *a, int x0, int r, int p) {
  const uint64_t add = 0x0101010101010101ull * (uint64_t)(127 - p);
  int d = x0 >> 3;
  uint64_t m = ((row_u64(a, d) + add) & kHigh8) & (~0ull << (8 * (x0 & 7)));
  while (m == 0) {
    ++d;
    if (8 * d > r) return r + 1;
    m = (row_u64(a, d) + add) & kHigh8;
  }
  const int x = 8 * d + (__builtin_ctzll(m) >> 3);
  return x <= r ? x : r + 1;
}

// last y <= y0 with a[y] <= p; a[lo] == p stops it at lo (the j-scan)
__device__ inline int scan_le(const uint8_t *a, int y0, int p) {
  const uint64_t add = 0x0101010101010101ull * (uint64_t)(127 - p);
  int d = y0 >> 3;
  uint64_t m = (~(row_u64(a, d) + add) & kHigh8) & (~0ull >> (8 * (7 - (y0 & 7))));
  while (m == 0) {
    --d;
    m = ~(row_u64(a, d) + add) & kHigh8;
  }
  return 8 * d + ((63 - __builtin_clzll(m)) >> 3);
}

__device__ inline int code_part(uint8_t *a, uint8_t *b, int lo, int hi) {
  const int p = a[lo];
  int i = lo, j = hi + 1;
  for (;;) {
    i = scan_gt(a, i + 1, hi, p);
    j = scan_le(a, j - 1, p);
    if (i >= j) break;
    const uint8_t ta = a[i]; a[i] = a[j]; a[j] = ta;
    const uint8_t tb = b[i]; b[i] = b[j]; b[j] = tb;
  }
  const uint8_t ta = a[lo]; a[lo] = a[j]; a[j] = ta;
  const uint8_t tb = b[lo]; b[lo] = b[j]; b[j] = tb;
  return j;
}

// codes row[0..n-1] into a; bytes past n in its last word zeroed (the scans
// read whole words)
__device__ inline void load_codes(uint8_t *a, const uint2 *__restrict__ src, int n) {
  for (int w = 0; 8 * w < n; ++w) {
    uint2 v = src[w];
    const int left = n - 8 * w;  // codes in this 8-byte word
    if (left < 8) {
      const uint64_t keep = ~0ull >> (64 - 8 * left);
      uint64_t u = ((uint64_t)v.y << 32) | v.x;
      u &= keep;
      v = make_uint2((uint32_t)u, (uint32_t)(u >> 32));
    }
    *reinterpret_cast<uint2 *>(a + 8 * w) = v;
  }
}

// the reference's quicksort (aquickSort) of (a, b)[0..hi] by one lane;
// ranges lying wholly at or above `need` are left unsorted
__device__ inline void lane_code_sort(uint8_t *a, uint8_t *b, int hi0, int need, uint16_t *stk,
                                      int t) {
  int top = 0;
  stk[t] = (uint16_t)(0 | (hi0 << 8));
  top = 1;
  while (top > 0) {
    const int pk = stk[(--top) * 64 + t];
    int lo = pk & 0xff, hi = pk >> 8;
    while (lo < hi && lo < need) {
      const int j = code_part(a, b, lo, hi);
      if (j - lo < hi - j) {
        stk[(top++) * 64 + t] = (uint16_t)((j + 1) | (hi << 8));
        hi = j - 1;
      } else {
        stk[(top++) * 64 + t] = (uint16_t)(lo | ((j - 1) << 8));
        lo = j + 1;
      }
    }
  }
}

// one node's final sort (lane t's row at a, b) and its output rows
__device__ inline void search_sort_row(int M, int lm, const Frame &F,
                                       const double4 *__restrict__ xyz,
                                       const uint8_t *__restrict__ scodes,
                                       const int *__restrict__ sidx,
                                       const uint8_t *__restrict__ slen, int *__restrict__ nbr,
                                       double *__restrict__ dist, uint8_t *a, uint8_t *b,
                                       uint16_t *stk, int i0, int l) {
  const int t = threadIdx.x;
  const int n0 = slen[(size_t)M + i0];
  const uint2 *src = reinterpret_cast<const uint2 *>(scodes + (size_t)i0 * SCAP);
  // the positions 0..l-1 as payload
  for (int w = 0; 8 * w < l; ++w) {
    const uint32_t base = 0x03020100u + 0x08080808u * (uint32_t)w;
    *reinterpret_cast<uint2 *>(b + 8 * w) = make_uint2(base, base + 0x04040404u);
  }
  // octant 0 with equal keys: its quicksort (on the squared distance,
  // solver.cpp:557-561) replayed on the rank codes first; the permutation of
  // positions 0..n0-1 it leaves in b is the payload of those slots, whose
  // list indices are in list order
  if (n0 > 1) {
    load_codes(a, src + ((l + 7) >> 3), n0);
    lane_code_sort(a, b, n0 - 1, n0, stk, t);
  }
  // the list's codes, bytes past l zeroed
  load_codes(a, src, l);
  lane_code_sort(a, b, l - 1, lm, stk, t);
  // output, eight slots at a time: their indices, then their centres, are
  // loaded together so the gathers overlap
  const double4 p0 = xyz[i0];
  const int *li = sidx + (size_t)i0 * SCAP;
  for (int q0 = 0; q0 < lm; q0 += 8) {
    int mm[8];
    double4 pp[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) mm[k] = (q0 + k < lm) ? li[b[q0 + k]] : i0;
#pragma unroll
    for (int k = 0; k < 8; ++k) pp[k] = xyz[mm[k]];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      if (q0 + k < lm) {
        double dd;
        (void)place_p(F, pp[k], p0.x, p0.y, p0.z, &dd);
        nbr[(size_t)(q0 + k) * M + i0] = mm[k];
        dist[(size_t)(q0 + k) * M + i0] = sqrt(dd);
      }
    }
  }
}

__global__ void __launch_bounds__(64) k_search_sort(int M, int lm, Frame F,
                                                    const double4 *__restrict__ xyz,
                                                    const uint8_t *__restrict__ scodes,
                                                    const int *__restrict__ sidx,
                                                    const uint8_t *__restrict__ slen,
                                                    int *__restrict__ nbr,
                                                    double *__restrict__ dist) {
  __shared__ __align__(16) uint8_t sa[SBUD];
  __shared__ __align__(16) uint8_t sb[SBUD];
  __shared__ uint16_t stk[SSTK * 64];
  const int t = threadIdx.x;
  const int i0 = blockIdx.x * 64 + t;
  // l = 0: no node, or sorted by k_search_wave or by the long-list pass
  const int l = (i0 < M) ? slen[i0] : 0;
  const int rsz = (l + 7) & ~7;  // the row's bytes
  int tot = 0;
  const int off = wave_excl_scan_i(rsz, &tot);
  for (int base = 0; base < tot;) {
    // this pass: the rows from base that end within SBUD bytes of it
    const bool mine = rsz > 0 && off >= base && off + rsz <= base + SBUD;
    const uint64_t later = __ballot(rsz > 0 && off + rsz > base + SBUD);
    const int next = later ? __shfl(off, __ffsll((unsigned long long)later) - 1, 64) : tot;
    if (mine) search_sort_row(M, lm, F, xyz, scodes, sidx, slen, nbr, dist, sa + (off - base),
                              sb + (off - base), stk, i0, l);
    base = next;
    __syncthreads();
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
  APC_ALLOC(c, c->visits, 3);
  APC_HIP(c, hipMemsetAsync(c->visits.p, 0, 3 * sizeof(unsigned long long), c->s));
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
  // first pass: one wave per node (k_search_wave) when the list bound fits
  // its LDS lists, else one thread per node with lists in device scratch
  const char *wenv = getenv("APC_NEIGH_WAVE");
  const bool wave = lm <= WCAP / 2 && !(wenv && atoi(wenv) == 0);
  int ncu = 0;
  APC_HIP(c, hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, c->dev));
  if (c->search_slots == 0) {  // resident threads of k_search on this device
    int nb = 0;
    APC_HIP(c, hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_search, 256, 0));
    c->search_slots = std::max(1, nb) * std::max(1, ncu) * 256;
  }
  if (c->search_waves == 0) {  // resident waves of k_search_wave
    int nb = 0;
    APC_HIP(c, hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_search_wave, 64, 0));
    c->search_waves = std::max(1, nb) * std::max(1, ncu);
  }
  int slot_cap = c->search_slots;
  if (const char *v = getenv("APC_NEIGH_SLOTS")) slot_cap = std::max(256, atoi(v));
  const int slots = std::max(1, std::min(M, std::min(slot_cap, c->search_slots)));
  const int slots_full = std::max(1, std::min(M, 1 << 16));
  const size_t elems = wave ? (size_t)cap_full * slots_full
                            : std::max((size_t)fast_cap * slots, (size_t)cap_full * slots_full);
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
  int search_scope = -1;  // timer of the wave search; its bytes are known after the launch
  if (wave) {
    APC_ALLOC(c, c->scodes, (size_t)M * SCAP + 16);
    APC_ALLOC(c, c->sidx, (size_t)M * SCAP + 4);
    APC_ALLOC(c, c->slen, 2 * (size_t)M + 16);
    if (!c->search_pen) {  // per context: the constant table lives on its device
      static const std::array<double, WCAP> pen = [] {
        std::array<double, WCAP> t{};
        double p = 1.0;
        for (int k = 0; k < WCAP; ++k) {
          t[k] = p;
          p = p * 10000.0;
        }
        return t;
      }();
      APC_HIP(c, hipMemcpyToSymbolAsync(HIP_SYMBOL(c_pen), pen.data(), sizeof(double) * WCAP, 0,
                                        hipMemcpyHostToDevice, c->s));
      c->search_pen = 1;
    }
    KScope ks(c, "neighbor_search");
    search_scope = ks.idx;
    int grid = std::min(M, c->search_waves);
    if (grid >= 8) grid -= grid % 8;  // a multiple of 8: node chunks dealt to the XCDs
    grid = std::max(1, grid);
    // node chunks of the knob's size (default 256; 16 to 4096 measured alike
    // at 256^3, 58.3 -> 53.9 ms against a plain grid stride)
    const int xchunk = c->knob[APC_KNOB_SEARCH_CHUNK];
    hipLaunchKernelGGL(k_search_wave, dim3(grid), dim3(64), 0, c->s, M, S, lm,
                       std::min(fast_cap, WCAP), xchunk, F, (const double4 *)xyz, c->n24.p, c->n24cnt.p,
                       c->nbr.p, c->dist.p, c->status_dev.p, c->novf.p, c->visits.p,
                       c->scodes.p, c->sidx.p, c->slen.p);
  } else {
    KScope ks(c, "neighbor_search");
    hipLaunchKernelGGL(k_search, dim3(nblocks(slots, 256)), dim3(256), 0, c->s, M, S, M,
                       (const int *)nullptr, lm, fast_cap, slots, F, (const double4 *)xyz, c->n24.p, c->n24cnt.p, mo, d2, wd, wm, fpw, c->nbr.p, c->dist.p,
                       c->status_dev.p, c->novf.p);
  }
  int st[3] = {0, 0, 0};
  unsigned long long cnt2[3] = {0, 0, 0};
  APC_HIP(c, hipMemcpyAsync(st, c->status_dev.p, sizeof(st), hipMemcpyDeviceToHost, c->s));
  APC_HIP(c, hipMemcpyAsync(cnt2, c->visits.p, sizeof(cnt2), hipMemcpyDeviceToHost, c->s));
  APC_HIP(c, hipStreamSynchronize(c->s));
  const double nvis = (double)cnt2[0], nlist = (double)cnt2[1], nrank = (double)cnt2[2];
  c->search_visits = cnt2[0];
  // algorithmic bytes of the wave search: 28 per Neigh24 entry visited (its
  // index and the centre of the node it names), the node's own centre and
  // count, per list entry handed on its code (1) and index (4), per node
  // its list length and octant-0 replay count (2) and per octant-0 entry
  // handed on its rank code (1)
  if (search_scope >= 0)
    ktime_bytes(c, search_scope, 28.0 * nvis + 30.0 * M + 5.0 * nlist + nrank);
  // the final sorts (every list in the wave search's LDS was complete: no
  // error was raised, so every node's length entry is this search's)
  if (wave && st[0] == 0) {
    // per node: length and replay count 2, codes (read whole) and octant-0
    // rank codes, per output slot the index (4), the neighbour's centre
    // (24), index + distance written (12), own centre 24
    KScope ks(c, "neighbor_sort", nlist + nrank + (26.0 + 40.0 * lm) * M);
    hipLaunchKernelGGL(k_search_sort, dim3(nblocks(M, 64)), dim3(64), 0, c->s, M, lm, F,
                       (const double4 *)xyz, c->scodes.p, c->sidx.p, c->slen.p, c->nbr.p,
                       c->dist.p);
    APC_HIP(c, hipGetLastError());
  }
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
