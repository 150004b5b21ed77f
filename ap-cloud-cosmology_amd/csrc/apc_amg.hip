// Aggregation multigrid preconditioner for the GFD systems.
//
// The reference configures its solves with KSPSetFromOptions after choosing
// GMRES + Jacobi (solver.cpp:197-202, 269-274), so its users can select
// another preconditioner on the command line (`-pc_type gamg`).  This file is
// that option for the GPU path: an algebraic multigrid built on the device
// every solve, applied as a right preconditioner so GMRES converges on the
// true residual.
//
//  * Aggregation: maximal independent set on the matrix graph (priority = a
//    bijective hash of the row index, so the set is deterministic); every
//    other row joins the root it is most strongly coupled to.
//  * Coarse operators: Galerkin with the piecewise-constant prolongation,
//    A_c[I,J] = sum of a_ij over i in I, j in J, formed by a stable radix
//    sort of (I,J) keys and a segmented sum (deterministic).
//  * Smoother: damped Jacobi; correction scaled by alpha; V- or W-cycle.
//  * Coarsest level: dense inverse by Gauss-Jordan with partial pivoting.  On
//    the periodic operator (one identity row pins the constant mode) the
//    constant mode is nearly singular; the coarsest matrix gets the rank-one
//    shift sigma*11^T/n before inversion so that the preconditioner does not
//    amplify it (GMRES resolves that one direction itself).
#include <algorithm>
#include <cmath>
#include <vector>

#include "apc_solve.h"

namespace apc {

template <class T>
struct Csr {
  const int *rp, *col;
  const T *val;
  int n;
};

struct AmgLevel {
  int n = 0, nnz = 0;
  DBuf<int> rp, col;
  DBuf<double> val, dinv;
  DBuf<float> val32, dinv32;
  // sliced ELL copy for the cycle: slices of 64 rows, slot-major within a
  // slice, each padded to its longest row (padding: own column, value 0)
  long long sell_len = 0;
  DBuf<int> soff, swid, scol;  // slice offsets (elements), widths, columns
  DBuf<double> sval;
  DBuf<float> sval32;
  DBuf<int> agg;      // coarse row of each row (all but the coarsest level)
  DBuf<int> mp, ml;   // rows of each coarse row, grouped (CSR over next level)
  DBuf<int> box;      // packed boxes: the box of each row (levels below the finest)
  // packed boxes: 1 on the rows of a box whose coarsening has stopped at or
  // above this level (no smoothing here: dinv 0; the coarse correction taken
  // once, unscaled: the box's coarsest solve passes through unchanged)
  DBuf<unsigned char> pass;
  DBuf<double> x, b, r, t;
  DBuf<float> x32, b32, r32, t32;
  // with AmgOptions::ap: the product A P (rows of A, columns the
  // aggregates).  Coarse levels: sliced ELL like their A.  Level 0:
  // compacted slices (k_ap_build): slot k of a 64-row slice holds only the
  // rows with more than k entries, in row order, after slots 0..k-1; apn
  // counts each row's entries, apss / apmask give each slot's start in its
  // slice and its rows (bit = row within the slice).  ap_nnz counts the
  // entries, ap_len the slots x 64 of padded slices.
  bool ap = false;
  long long ap_nnz = 0, ap_len = 0;
  DBuf<int> apoff, apwid, apcol;
  DBuf<double> apval;
  DBuf<float> apval32;
  DBuf<unsigned char> apn;
  DBuf<int> apss;
  DBuf<unsigned long long> apmask;
  // packed boxes, coarse levels: the pass flag of each aggregate's rows (an
  // aggregate never joins rows of two boxes, so its rows share the flag)
  DBuf<unsigned char> apass;
  void release() {
    ap = false;
    apoff.release(); apwid.release(); apcol.release(); apval.release(); apval32.release();
    apn.release(); apss.release(); apmask.release(); apass.release();
    box.release(); pass.release();
    soff.release(); swid.release(); scol.release(); sval.release(); sval32.release();
    rp.release(); col.release(); val.release(); dinv.release(); val32.release();
    dinv32.release(); agg.release(); mp.release(); ml.release(); x.release(); b.release();
    r.release(); t.release(); x32.release(); b32.release(); r32.release(); t32.release();
  }
};

// A captured cycle: one per (input, output, status, precision) while the
// hierarchy is unchanged.
struct CycleGraph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  const void *in = nullptr, *st = nullptr, *a0 = nullptr;
  void *out = nullptr;
  int f32 = 0, gen = -1;
  // s: the stream the graph was launched on; it is drained first, so that no
  // launch of this exec is still queued or running when it is destroyed
  void release(hipStream_t s) {
    if (exec && s) (void)hipStreamSynchronize(s);
    if (exec) hipGraphExecDestroy(exec);
    if (graph) hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
  }
};

struct AmgHier {
  std::vector<AmgLevel *> L;
  std::vector<CycleGraph> graphs;
  int gen = 0;  // bumped by every setup; graphs of older generations are stale
  int nlev = 0;
  int nc = 0;
  bool dense = false;
  DBuf<double> gj, cinv, colk;
  // packed boxes: per-box coarsening stops (frozen), the coarsest level's box
  // blocks (first rows, Gauss-Jordan and inverse offsets, shifts)
  int nbox = 1;
  bool blocks = false;
  DBuf<int> frozen, bstart;
  DBuf<long long> goff, ioff;
  DBuf<double> bshift;
  DBuf<float> cinv32;
  AmgOptions o;
  bool fp32 = false;
  DBuf<uint64_t> key, key2;
  DBuf<int> idx, idx2, sa, sb, flag, pos, cnt;
  DBuf<double> ev, smax, diag;
  DBuf<int> gcol;     // Galerkin entries grouped by coarse row: coarse column
  DBuf<double> gval;  // and value
  DBuf<int> rcnt;     // undecided rows after each independent-set round of a batch
  DBuf<int> apsn;     // level 0's A P: real entries per slice
  DBuf<long long> aptot;
  DBuf<int> apiota, aprp, apcnt, apc;  // coarse levels' A P: rows as one-row groups, CSR
  DBuf<double> apv, apdiag;
  // geometric aggregation: the octree cell (x, y, z, level) of every row of
  // the level being aggregated, and of the next level's rows
  DBuf<int4> gcell, gcell2;
  bool geo_prev = false;  // the last level built was aggregated geometrically
  hipStream_t s = nullptr;  // the owning context's stream (graphs are launched on it)
  ~AmgHier() {
    for (auto &g : graphs) g.release(s);
    for (auto *l : L) { l->release(); delete l; }
    gj.release(); cinv.release(); colk.release(); cinv32.release(); key.release();
    frozen.release(); bstart.release(); goff.release(); ioff.release(); bshift.release();
    key2.release(); idx.release(); idx2.release(); sa.release(); sb.release(); flag.release();
    pos.release(); cnt.release(); ev.release(); smax.release(); diag.release();
    gcol.release(); gval.release(); apsn.release(); aptot.release(); rcnt.release();
    apiota.release(); aprp.release(); apcnt.release(); apc.release(); apv.release();
    apdiag.release(); gcell.release(); gcell2.release();
  }
};

// ---------------------------------------------------------------- views
// Off-diagonal entries of row i (duplicates included) and the diagonal.
struct EllView {
  Op<double> A;
  __device__ int n() const { return A.M; }
  // the diagonal, then every slot with an interior column, in slot order
  __device__ int count(int i) const {
    int c = 1;
    for (int q = 0; q < A.lm; ++q) c += (A.nbr[(size_t)q * A.M + i] < A.M) ? 1 : 0;
    return c;
  }
  template <class F>
  __device__ void entries(int i, F f) const {
    f(i, A.d[i]);
    const double fi = A.f[i];
    for (int q = 0; q < A.lm; ++q) {
      int m = A.nbr[(size_t)q * A.M + i];
      if (m < A.M) f(m, fi * A.s[(size_t)q * A.M + i]);
    }
  }
  __device__ double diag(int i) const { return A.d[i]; }
  template <class F>
  __device__ void offdiag(int i, F f) const {
    const double fi = A.f[i];
    for (int q = 0; q < A.lm; ++q) {
      int m = A.nbr[(size_t)q * A.M + i];
      if (m < A.M) f(m, fi * A.s[(size_t)q * A.M + i]);
    }
  }
  // every slot in slot order as f(j, a, ok): ok for an interior column j;
  // otherwise j = i (a valid index to gather from) and the slot is to be
  // ignored.  Slots are loaded nine at a time so that f's gathers at j do
  // not wait on one another.  VAL false: the values are not read (a = 0),
  // for callers that use only the pattern.
  template <bool VAL = true, class F>
  __device__ void offdiag_all(int i, F f) const {
    const double fi = VAL ? A.f[i] : 0.0;
    constexpr int RG = 9;
    int q = 0;
    for (; q + RG <= A.lm; q += RG) {
      int m[RG];
      double sv[RG];
#pragma unroll
      for (int u = 0; u < RG; ++u) {
        m[u] = A.nbr[(size_t)(q + u) * A.M + i];
        sv[u] = VAL ? A.s[(size_t)(q + u) * A.M + i] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < RG; ++u) {
        const bool ok = m[u] < A.M;
        f(ok ? m[u] : i, fi * sv[u], ok);
      }
    }
    for (; q < A.lm; ++q) {
      const int m = A.nbr[(size_t)q * A.M + i];
      const bool ok = m < A.M;
      f(ok ? m : i, VAL ? fi * A.s[(size_t)q * A.M + i] : 0.0, ok);
    }
  }
};

struct CsrView {
  Csr<double> A;
  __device__ int n() const { return A.n; }
  // every stored entry of row i in storage order (the Galerkin generation order)
  __device__ int count(int i) const { return A.rp[i + 1] - A.rp[i]; }
  template <class F>
  __device__ void entries(int i, F f) const {
    for (int e = A.rp[i]; e < A.rp[i + 1]; ++e) f(A.col[e], A.val[e]);
  }
  __device__ double diag(int i) const {
    double d = 0.0;
    for (int e = A.rp[i]; e < A.rp[i + 1]; ++e)
      if (A.col[e] == i) d += A.val[e];
    return d;
  }
  template <class F>
  __device__ void offdiag(int i, F f) const {
    for (int e = A.rp[i]; e < A.rp[i + 1]; ++e)
      if (A.col[e] != i) f(A.col[e], A.val[e]);
  }
  template <bool VAL = true, class F>
  __device__ void offdiag_all(int i, F f) const {
    for (int e = A.rp[i]; e < A.rp[i + 1]; ++e) f(A.col[e], VAL ? A.val[e] : 0.0, A.col[e] != i);
  }
};

__device__ inline uint32_t prio_of(uint32_t i) {  // bijective on 32 bits
  uint32_t x = i * 0x9E3779B1u;
  x ^= x >> 16;
  x *= 0x85EBCA6Bu;
  x ^= x >> 13;
  return x;
}

// ---------------------------------------------------------------- setup kernels
template <class V>
__global__ void __launch_bounds__(256) k_amg_smax(V v, double *__restrict__ smax) {
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= v.n()) return;
  double m = 0.0;
  v.offdiag_all(i, [&](int, double a, bool ok) {
    if (ok) m = fmax(m, fabs(a));
  });
  smax[i] = m;
}

// MIS round, part 1: an undecided row whose priority beats every undecided
// strong neighbour becomes a root.  a: states in (0 undecided, 1 root,
// 2 covered), b: states out.
template <class V>
__global__ void __launch_bounds__(256) k_amg_pick(V v, double theta, const double *__restrict__ smax,
                                                  const int *__restrict__ a, int *__restrict__ b,
                                                  const int *__restrict__ prev_und) {
  if (prev_und && *prev_und == 0) return;  // the set settled in an earlier round
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= v.n()) return;
  int s = a[i];
  if (s != 0) { b[i] = s; return; }
  const uint32_t pi = prio_of((uint32_t)i);
  const double lim = theta * smax[i];
  bool best = true;
  auto f = [&](int j, double x, bool ok) {
    const int aj = a[j];
    if (ok && fabs(x) >= lim && aj == 0 && prio_of((uint32_t)j) > pi) best = false;
  };
  if (theta > 0) v.template offdiag_all<true>(i, f);
  else v.template offdiag_all<false>(i, f);  // every coupling is strong: pattern only
  b[i] = best ? 1 : 0;
}

// MIS round, part 2: an undecided row with a strong root neighbour is
// covered; the rest are counted.  The grid is bounded (eight blocks per CU,
// striding over the rows) and each block adds its count with one atomic:
// atomics on one word serialise, and one per wave of a 7.5M-row level (117k)
// cost more than the round's loads.
template <class V>
__global__ void __launch_bounds__(256) k_amg_cover(V v, double theta, const double *__restrict__ smax,
                                                   int *__restrict__ b, int *__restrict__ undecided,
                                                   const int *__restrict__ prev_und) {
  if (prev_und && *prev_und == 0) return;
  __shared__ int ws[4];
  int und = 0;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < v.n(); i += gridDim.x * 256) {
    if (b[i] != 0) continue;
    const double lim = theta * smax[i];
    bool cov = false;
    auto f = [&](int j, double x, bool ok) {
      const int bj = b[j];
      if (ok && fabs(x) >= lim && bj == 1) cov = true;
    };
    if (theta > 0) v.template offdiag_all<true>(i, f);
    else v.template offdiag_all<false>(i, f);  // every coupling is strong: pattern only
    if (cov) b[i] = 2;
    else ++und;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) und += __shfl_xor(und, o, 64);
  if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = und;
  __syncthreads();
  if (threadIdx.x == 0) {
    const int t = ws[0] + ws[1] + ws[2] + ws[3];
    if (t) atomicAdd(undecided, t);
  }
}

__global__ void k_amg_rootflag(int n, const int *__restrict__ st, int *__restrict__ flag) {
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) flag[i] = (st[i] == 1) ? 1 : 0;
}

// Aggregate of every row: a root's rank, or the strongest root neighbour's
// (largest |a_ij|, then smallest j).
template <class V>
__global__ void __launch_bounds__(256) k_amg_agg(V v, double theta, const double *__restrict__ smax,
                                                 const int *__restrict__ st,
                                                 const int *__restrict__ rank, int *__restrict__ agg,
                                                 int *__restrict__ bad) {
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= v.n()) return;
  if (st[i] == 1) { agg[i] = rank[i]; return; }
  const double lim = theta * smax[i];
  double bv = -1.0;
  int bj = -1;
  v.offdiag_all(i, [&](int j, double x, bool ok) {
    const double ax = fabs(x);
    const int sj = st[j];
    if (ok && ax >= lim && sj == 1 && (ax > bv || (ax == bv && j < bj))) { bv = ax; bj = j; }
  });
  if (bj < 0) { atomicAdd(bad, 1); agg[i] = 0; return; }
  agg[i] = rank[bj];
}

__global__ void k_amg_keys_rows(int n, const int *__restrict__ agg, uint64_t *__restrict__ key,
                                int *__restrict__ idx) {
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  key[i] = (uint64_t)agg[i];
  idx[i] = i;
}

// first sorted position of every key (keys 0..nc-1 all present)
__global__ void k_amg_bounds(int n, int nc, const uint64_t *__restrict__ key, int *__restrict__ mp) {
  int k = blockIdx.x * 256 + threadIdx.x;
  if (k >= n) return;
  if (k == 0 || key[k] != key[k - 1]) mp[key[k]] = k;
  if (k == n - 1) mp[nc] = n;
}

constexpr uint64_t SENT = ~0ull;

// Galerkin entries from the ELL level: entry q*M + i, q = 0 the diagonal.
__global__ void __launch_bounds__(256) k_amg_entries_ell(Op<double> A, const int *__restrict__ agg,
                                                         int nc, uint64_t *__restrict__ key,
                                                         double *__restrict__ ev,
                                                         int *__restrict__ idx) {
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= A.M) return;
  const uint64_t I = (uint64_t)agg[i];
  const double fi = A.f[i];
  key[i] = I * nc + I;
  ev[i] = A.d[i];
  idx[i] = i;
  for (int q = 0; q < A.lm; ++q) {
    const size_t e = (size_t)(q + 1) * A.M + i;
    int m = A.nbr[(size_t)q * A.M + i];
    if (m < A.M) {
      key[e] = I * nc + (uint64_t)agg[m];
      ev[e] = fi * A.s[(size_t)q * A.M + i];
    } else {
      key[e] = SENT;
      ev[e] = 0.0;
    }
    idx[e] = (int)e;
  }
}

__global__ void __launch_bounds__(256) k_amg_entries_csr(Csr<double> A, const int *__restrict__ agg,
                                                         int nc, uint64_t *__restrict__ key,
                                                         double *__restrict__ ev,
                                                         int *__restrict__ idx) {
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= A.n) return;
  const uint64_t I = (uint64_t)agg[i];
  for (int e = A.rp[i]; e < A.rp[i + 1]; ++e) {
    key[e] = I * nc + (uint64_t)agg[A.col[e]];
    ev[e] = A.val[e];
    idx[e] = e;
  }
}

__global__ void k_amg_heads(int n, uint64_t lim, const uint64_t *__restrict__ key,
                            int *__restrict__ flag) {
  int k = blockIdx.x * 256 + threadIdx.x;
  if (k >= n) return;
  flag[k] = (key[k] < lim && (k == 0 || key[k] != key[k - 1])) ? 1 : 0;
}

// Segmented sum of the sorted entries into the coarse CSR.
__global__ void __launch_bounds__(256) k_amg_reduce(int n, int nc, const uint64_t *__restrict__ key,
                                                    const int *__restrict__ idx,
                                                    const double *__restrict__ ev,
                                                    const int *__restrict__ flag,
                                                    const int *__restrict__ pos,
                                                    int *__restrict__ rp, int *__restrict__ col,
                                                    double *__restrict__ val,
                                                    double *__restrict__ diag) {
  int k = blockIdx.x * 256 + threadIdx.x;
  if (k >= n || !flag[k]) return;
  const uint64_t kk = key[k];
  double s = 0.0;
  for (int q = k; q < n && key[q] == kk; ++q) s += ev[idx[q]];
  const int p = pos[k];
  const int I = (int)(kk / (uint64_t)nc), J = (int)(kk % (uint64_t)nc);
  col[p] = J;
  val[p] = s;
  if (J == I) diag[I] = s;
  if (k == 0 || (int)(key[k - 1] / (uint64_t)nc) != I) rp[I] = p;
}

// Galerkin product by coarse row, in three row-parallel steps and no global
// sort:
//  1. every fine row's entry count; an exclusive scan of the counts taken in
//     member order (the aggregates' rows grouped, ascending within each)
//     gives each fine row an offset such that the entries of a coarse row's
//     members are contiguous, members in row order;
//  2. each fine row writes its entries there as (coarse column, value), in
//     storage order (coalesced reads of the fine matrix);
//  3. one 64-lane block per coarse row reads its span and takes the distinct
//     coarse columns in increasing order by repeated wave minima, summing
//     each over its entries with a fixed reduction tree (deterministic; DPP row
//     operations, no LDS crossbar).  A
//     span of up to GAL_CAP entries is held in registers (as many per lane as
//     the span needs); a longer one is re-read from memory for every column.
//  Pass 0 of step 3 counts each coarse row's distinct columns; a span held in
//  registers also gets its result (column, sum) written over its own start,
//  the whole span having been read first.  Pass 1 copies those results to
//  rp[I], and runs the streamed reduction again for the longer spans, writing
//  at rp[I].
//  On the ELL level steps 1 and 2 are skipped: k_gal_rows_ell forms each
//  entry from the matrix where step 3 uses it (member k's slots in order),
//  which saves writing and re-reading every fine entry (8.6 -> 7.1 ms of
//  setup per step at 256^3).
constexpr int GAL_CAP = 1024;
constexpr int GAL_R = GAL_CAP / 64;
// Entries per row of level 0's A P: the diagonal and up to kApMax - 1 slots.
constexpr int kApMax = 32;

__device__ inline int wave_min_int(int v) { return wave_min_int_dpp(v); }

// entry counts in row order (coalesced reads of the fine matrix) ...
template <class V>
__global__ void k_gal_count(V v, int n, int *__restrict__ cnt_row) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) cnt_row[i] = v.count(i);
}

// ... and gathered into member order
__global__ void k_gal_gather(int n, const int *__restrict__ ml, const int *__restrict__ cnt_row,
                             int *__restrict__ cnt_ml) {
  const int k = blockIdx.x * 256 + threadIdx.x;
  if (k < n) cnt_ml[k] = cnt_row[ml[k]];
}

// off_row[ml[k]] = off_ml[k]
__global__ void k_gal_rowoff(int n, const int *__restrict__ ml, const int *__restrict__ off_ml,
                             int *__restrict__ off_row) {
  const int k = blockIdx.x * 256 + threadIdx.x;
  if (k < n) off_row[ml[k]] = off_ml[k];
}

template <class V>
__global__ void __launch_bounds__(256) k_gal_emit(V v, int n, const int *__restrict__ agg,
                                                  const int *__restrict__ off_row,
                                                  int *__restrict__ ej, double *__restrict__ ea) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  int e = off_row[i];
  v.entries(i, [&](int j, double a) {
    ej[e] = agg ? agg[j] : j;  // no map: the entries' columns are coarse already (A P)
    ea[e] = a;
    ++e;
  });
}

// One coarse row's E entries (E <= 64 R), src(e, J, A) giving entry e's
// coarse column and value (J = INT_MAX: no entry), held in R registers per
// lane: distinct columns ascending, each summed; the results (column, sum)
// go to oj/oa and their count is returned.  oj/oa may be the entries' own
// storage: every lane's loads have completed before the first write (the
// first minimum needs them).
template <int R, class SRC>
__device__ inline int gal_row_regs(int I, int E, SRC src, int *oj, double *oa, double *diag) {
  const int lane = threadIdx.x;
  int J[R];
  double A[R];
  int mn = INT_MAX;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int e = lane + 64 * r;
    J[r] = INT_MAX;
    A[r] = 0.0;
    if (e < E) src(e, J[r], A[r]);
    mn = min(mn, J[r]);
  }
  int cur = wave_min_int(mn);
  int u = 0;
  while (cur != INT_MAX) {
    double sum = 0.0;
    int nxt = INT_MAX;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      if (J[r] == cur) sum += A[r];
      else if (J[r] > cur) nxt = min(nxt, J[r]);
    }
    sum = wave_sum_dpp(sum);
    if (lane == 0) {
      oj[u] = cur;
      oa[u] = sum;
      if (cur == I) diag[I] = sum;
    }
    ++u;
    cur = wave_min_int(nxt);
  }
  return u;
}

template <class SRC>
__device__ inline int gal_row_regs_any(int I, int E, SRC src, int *oj, double *oa, double *diag) {
  if (E <= 128) return gal_row_regs<2>(I, E, src, oj, oa, diag);
  if (E <= 256) return gal_row_regs<4>(I, E, src, oj, oa, diag);
  if (E <= 384) return gal_row_regs<6>(I, E, src, oj, oa, diag);
  if (E <= 512) return gal_row_regs<8>(I, E, src, oj, oa, diag);
  return gal_row_regs<GAL_R>(I, E, src, oj, oa, diag);
}

// E > GAL_CAP entries, streamed once per distinct column; write: (column,
// sum) at o
template <class SRC>
__device__ inline int gal_row_stream(int I, int E, SRC src, bool write, int o, int *col,
                                     double *val, double *diag) {
  const int lane = threadIdx.x;
  int mn = INT_MAX;
  for (int e = lane; e < E; e += 64) {
    int j = INT_MAX;
    double a = 0.0;
    src(e, j, a);
    mn = min(mn, j);
  }
  int cur = wave_min_int(mn);
  int u = 0;
  while (cur != INT_MAX) {
    double sum = 0.0;
    int nxt = INT_MAX;
    for (int e = lane; e < E; e += 64) {
      int j = INT_MAX;
      double a = 0.0;
      src(e, j, a);
      if (j == cur) sum += a;
      else if (j > cur) nxt = min(nxt, j);
    }
    sum = wave_sum_dpp(sum);
    if (write && lane == 0) {
      col[o + u] = cur;
      val[o + u] = sum;
      if (cur == I) diag[I] = sum;
    }
    ++u;
    cur = wave_min_int(nxt);
  }
  return u;
}

// ej, ea: the emitted entries, overwritten by pass 0 (not __restrict__: read
// and written through the same pointers)
__global__ void __launch_bounds__(64) k_gal_rows(int nc, const int *__restrict__ mp,
                                                 const int *__restrict__ off_ml, int *ej,
                                                 double *ea, int pass, int *__restrict__ ucnt,
                                                 const int *__restrict__ rp,
                                                 int *__restrict__ col, double *__restrict__ val,
                                                 double *__restrict__ diag) {
  const int lane = threadIdx.x;
  for (int I = blockIdx.x; I < nc; I += gridDim.x) {
    const int e0 = off_ml[mp[I]], e1 = off_ml[mp[I + 1]];
    const int E = e1 - e0;
    auto src = [&](int e, int &J, double &A) {
      J = ej[e0 + e];
      A = ea[e0 + e];
    };
    if (pass == 0) {
      const int u = (E <= GAL_CAP) ? gal_row_regs_any(I, E, src, ej + e0, ea + e0, diag)
                                   : gal_row_stream(I, E, src, false, 0, nullptr, nullptr, nullptr);
      if (lane == 0) ucnt[I] = u;
    } else {
      const int o = rp[I];
      if (E <= GAL_CAP) {
        const int u = rp[I + 1] - o;
        for (int q = lane; q < u; q += 64) {
          col[o + q] = ej[e0 + q];
          val[o + q] = ea[e0 + q];
        }
      } else {
        (void)gal_row_stream(I, E, src, true, o, col, val, diag);
      }
    }
  }
}

// The same product on the ELL level without the emission step: coarse row
// I's entries are formed where they are used, entry e = k (lm + 1) + q of
// member k: q = 0 its diagonal, q >= 1 ELL slot q - 1 (no entry if that slot
// is empty), in the emission's order.  Pass 0 writes each row's results at
// (lm + 1) mp[I] in oj/oa (the members' upper bound) and counts them; pass
// 1 copies them to rp[I] (or streams a row longer than GAL_CAP again).
__global__ void __launch_bounds__(64) k_gal_rows_ell(int nc, Op<double> Am,
                                                     const int *__restrict__ agg,
                                                     const int *__restrict__ mp,
                                                     const int *__restrict__ ml, int pass,
                                                     int *__restrict__ ucnt,
                                                     const int *__restrict__ rp,
                                                     int *__restrict__ col,
                                                     double *__restrict__ val,
                                                     int *__restrict__ oj, double *__restrict__ oa,
                                                     double *__restrict__ diag) {
  const int lane = threadIdx.x;
  const int W = Am.lm + 1;
  for (int I = blockIdx.x; I < nc; I += gridDim.x) {
    const int k0 = mp[I];
    const int E = (mp[I + 1] - k0) * W;
    const long base = (long)k0 * W;
    auto src = [&](int e, int &J, double &A) {
      const int k = e / W, q = e - k * W;
      const int i = ml[k0 + k];
      if (q == 0) {
        J = agg[i];
        A = Am.d[i];
      } else {
        const size_t p = (size_t)(q - 1) * Am.M + i;
        const int m = Am.nbr[p];
        if (m < Am.M) {
          J = agg[m];
          A = Am.f[i] * Am.s[p];
        }
      }
    };
    if (pass == 0) {
      const int u = (E <= GAL_CAP) ? gal_row_regs_any(I, E, src, oj + base, oa + base, diag)
                                   : gal_row_stream(I, E, src, false, 0, nullptr, nullptr, nullptr);
      if (lane == 0) ucnt[I] = u;
    } else {
      const int o = rp[I];
      if (E <= GAL_CAP) {
        const int u = rp[I + 1] - o;
        for (int q = lane; q < u; q += 64) {
          col[o + q] = oj[base + q];
          val[o + q] = oa[base + q];
        }
      } else {
        (void)gal_row_stream(I, E, src, true, o, col, val, diag);
      }
    }
  }
}

// The level-0 product P^T (A P) from A P's compacted slices without an
// emission step: coarse row I's entries are its members' A P rows, member
// k's entry q (q below the members' widest row, W) at e = k W + q, taken in
// member order; A P is kApMax wide at most, so pass 0 writes row I's
// results at kApMax mp[I].  Row i's entry q: slot q's start in i's slice
// plus the slot's rows before i.
__global__ void __launch_bounds__(64) k_gal_rows_ap(int nc, const int *__restrict__ soff,
                                                    const int *__restrict__ pcol,
                                                    const double *__restrict__ pval,
                                                    const unsigned char *__restrict__ pcnt,
                                                    const int *__restrict__ sstart,
                                                    const unsigned long long *__restrict__ smask,
                                                    const int *__restrict__ mp,
                                                    const int *__restrict__ ml, int pass,
                                                    int *__restrict__ ucnt,
                                                    const int *__restrict__ rp,
                                                    int *__restrict__ col,
                                                    double *__restrict__ val,
                                                    int *__restrict__ oj, double *__restrict__ oa,
                                                    double *__restrict__ diag) {
  const int lane = threadIdx.x;
  for (int I = blockIdx.x; I < nc; I += gridDim.x) {
    const int k0 = mp[I], nmem = mp[I + 1] - k0;
    int W = 0;
    for (int k = lane; k < nmem; k += 64) W = max(W, pcnt[ml[k0 + k]]);
    for (int o = 32; o > 0; o >>= 1) W = max(W, __shfl_xor(W, o, 64));
    const int E = nmem * W;
    const long base = (long)k0 * kApMax;
    auto src = [&](int e, int &J, double &A) {
      const int k = e / W, q = e - k * W;
      const int i = ml[k0 + k];
      if (q < pcnt[i]) {
        const int sq = (i >> 6) * kApMax + q;
        const long long p = (long long)soff[i >> 6] * 64 + sstart[sq] +
                            __popcll(smask[sq] & ((1ull << (i & 63)) - 1));
        J = pcol[p];
        A = pval[p];
      }
    };
    if (pass == 0) {
      const int u = (E <= GAL_CAP) ? gal_row_regs_any(I, E, src, oj + base, oa + base, diag)
                                   : gal_row_stream(I, E, src, false, 0, nullptr, nullptr, nullptr);
      if (lane == 0) ucnt[I] = u;
    } else {
      const int o = rp[I];
      if (E <= GAL_CAP) {
        const int u = rp[I + 1] - o;
        for (int q = lane; q < u; q += 64) {
          col[o + q] = oj[base + q];
          val[o + q] = oa[base + q];
        }
      } else {
        (void)gal_row_stream(I, E, src, true, o, col, val, diag);
      }
    }
  }
}

__global__ void k_amg_dinv(int n, const double *__restrict__ diag, double *__restrict__ dinv) {
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) dinv[i] = (diag[i] != 0.0) ? 1.0 / diag[i] : 0.0;
}

// width of each 64-row slice
__global__ void k_sell_width(int n, const int *__restrict__ rp, int *__restrict__ wid) {
  const int s = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int r = s * 64 + lane;
  int w = (r < n) ? rp[r + 1] - rp[r] : 0;
  for (int o = 32; o > 0; o >>= 1) w = max(w, __shfl_xor(w, o, 64));
  if (lane == 0 && s * 64 < n) wid[s] = w;
}

// padding entries: value 0 and column padc[r] (the row itself without padc;
// a rectangular matrix passes a valid column of each row)
__global__ void k_sell_fill(int n, const int *__restrict__ rp, const int *__restrict__ col,
                            const double *__restrict__ val, const int *__restrict__ soff,
                            const int *__restrict__ wid, int *__restrict__ scol,
                            double *__restrict__ sval, const int *__restrict__ padc) {
  const int r = blockIdx.x * 256 + threadIdx.x;
  if (r >= n) return;
  const int s = r >> 6, lane = r & 63;
  const long long base = (long long)soff[s] * 64 + lane;
  const int b = rp[r], e = rp[r + 1], w = wid[s];
  const int pc = padc ? padc[r] : r;
  for (int k = 0; k < w; ++k) {
    const bool real = b + k < e;
    scol[base + (long long)k * 64] = real ? col[b + k] : pc;
    sval[base + (long long)k * 64] = real ? val[b + k] : 0.0;
  }
}

template <class S, class D>
__global__ void k_amg_convert(size_t n, const S *__restrict__ a, D *__restrict__ b) {
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256)
    b[i] = (D)a[i];
}

// Level 0's product A P with the aggregation's piecewise-constant
// prolongation: for every aggregate J that row i's entries meet (its
// diagonal and its slots), the sum of a_ij over the columns j in J, added in
// slot order, the aggregates in increasing order (repeated minima over the
// row's entries held in registers).  Stored in compacted slices: 64-row
// slices (one wave each), slice s at s * KMAX * 64; slot k holds the k-th
// entry of every row with more than k entries, in row order, right after
// slot k - 1 (no padding: a slot's rows are its bit mask, a row's place the
// mask's bits below it).  Per slice: its offset (in units of 64), its
// width (largest count) and number of entries, and per slot its start and
// mask; per row its count.  val32 (optional) gets the FP32 copy.
template <int KMAX>
__global__ void __launch_bounds__(256) k_ap_build(Op<double> A, const int *__restrict__ agg,
                                                  int *__restrict__ soff, int *__restrict__ wid,
                                                  int *__restrict__ snnz, int *__restrict__ col,
                                                  double *__restrict__ val,
                                                  float *__restrict__ val32,
                                                  unsigned char *__restrict__ ncnt,
                                                  int *__restrict__ sstart,
                                                  unsigned long long *__restrict__ smask) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  const int s = i >> 6, lane = threadIdx.x & 63;
  if (s >= (A.M + 63) >> 6) return;  // s is uniform over a wave: whole waves leave
  const bool live = i < A.M;
  int key[KMAX];
  double v[KMAX];
  int m[KMAX - 1];
  const int own = live ? agg[i] : INT_MAX;
  key[0] = own;
  v[0] = live ? A.d[i] : 0.0;
  const double fi = live ? A.f[i] : 0.0;
#pragma unroll
  for (int q = 0; q < KMAX - 1; ++q)
    m[q] = (live && q < A.lm) ? A.nbr[(size_t)q * A.M + i] : A.M;
#pragma unroll
  for (int q = 0; q < KMAX - 1; ++q) {
    const bool in = m[q] < A.M;
    key[q + 1] = in ? agg[m[q]] : INT_MAX;
    v[q + 1] = in ? fi * A.s[(size_t)q * A.M + i] : 0.0;
  }
  const long long base = (long long)s * KMAX * 64;
  int cur = INT_MAX;
#pragma unroll
  for (int k = 0; k < KMAX; ++k) cur = min(cur, key[k]);
  // slot k: every row with a k-th aggregate left; a row has at most KMAX
  // aggregates, so k < KMAX
  int cnt = 0, pos = 0, k = 0;
  for (;; ++k) {
    const unsigned long long mk = __ballot(cur != INT_MAX);
    if (mk == 0) break;
    if (lane == 0) {
      sstart[s * KMAX + k] = pos;
      smask[s * KMAX + k] = mk;
    }
    if (cur != INT_MAX) {
      double sum = 0.0;
      int nxt = INT_MAX;
#pragma unroll
      for (int q = 0; q < KMAX; ++q) {
        if (key[q] == cur) sum += v[q];
        else if (key[q] > cur) nxt = min(nxt, key[q]);
      }
      const long long e = base + pos + below(mk);
      col[e] = cur;
      val[e] = sum;
      if (val32) val32[e] = (float)sum;
      ++cnt;
      cur = nxt;
    }
    pos += __popcll(mk);
  }
  if (live) ncnt[i] = (unsigned char)cnt;
  if (lane == 0) {
    soff[s] = s * KMAX;
    wid[s] = k;
    snnz[s] = pos;
  }
}

// out[0] = padded entries (64 per slot of every slice), out[1] = real ones
__global__ void __launch_bounds__(1024) k_ap_total(int ns, const int *__restrict__ wid,
                                                   const int *__restrict__ snnz,
                                                   long long *__restrict__ out) {
  __shared__ long long a[1024], b[1024];
  long long sw = 0, sn = 0;
  for (int s = threadIdx.x; s < ns; s += 1024) {
    sw += wid[s];
    sn += snnz[s];
  }
  a[threadIdx.x] = sw;
  b[threadIdx.x] = sn;
  __syncthreads();
  for (int h = 512; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) {
      a[threadIdx.x] += a[threadIdx.x + h];
      b[threadIdx.x] += b[threadIdx.x + h];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = a[0] * 64;
    out[1] = b[0];
  }
}

// [A | I] for Gauss-Jordan, A from CSR plus shift/n everywhere.
__global__ void k_gj_fill(Csr<double> A, double shift, double *__restrict__ g) {
  const int n = A.n;
  int i = blockIdx.x;
  for (int j = threadIdx.x; j < 2 * n; j += blockDim.x)
    g[(size_t)i * 2 * n + j] = (j < n) ? shift / n : (j - n == i ? 1.0 : 0.0);
  __syncthreads();
  if (threadIdx.x == 0)
    for (int e = A.rp[i]; e < A.rp[i + 1]; ++e) g[(size_t)i * 2 * n + A.col[e]] += A.val[e];
}

// pivot step k: pick the largest |g[i][k]|, i >= k, swap it into row k,
// scale row k, save column k for the elimination
__global__ void __launch_bounds__(256) k_gj_pivot(int n, int k, double *__restrict__ g,
                                                  double *__restrict__ colk, int *__restrict__ bad) {
  __shared__ double bv[256];
  __shared__ int bi[256];
  const int t = threadIdx.x, w = 2 * n;
  double v = -1.0;
  int vi = k;
  for (int i = k + t; i < n; i += 256) {
    double a = fabs(g[(size_t)i * w + k]);
    if (a > v) { v = a; vi = i; }
  }
  bv[t] = v;
  bi[t] = vi;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (t < s && (bv[t + s] > bv[t] || (bv[t + s] == bv[t] && bi[t + s] < bi[t]))) {
      bv[t] = bv[t + s];
      bi[t] = bi[t + s];
    }
    __syncthreads();
  }
  const int p = bi[0];
  if (bv[0] <= 0.0) {
    if (t == 0) *bad = 1;
    return;
  }
  if (p != k)
    for (int j = t; j < w; j += 256) {
      double a = g[(size_t)k * w + j];
      g[(size_t)k * w + j] = g[(size_t)p * w + j];
      g[(size_t)p * w + j] = a;
    }
  __syncthreads();
  // every wave reads the pivot before any thread overwrites it with 1
  const double inv = 1.0 / g[(size_t)k * w + k];
  __syncthreads();
  for (int j = t; j < w; j += 256) g[(size_t)k * w + j] *= inv;
  __syncthreads();
  for (int i = t; i < n; i += 256) colk[i] = g[(size_t)i * w + k];
}

__global__ void __launch_bounds__(256) k_gj_elim(int n, int k, double *__restrict__ g,
                                                 const double *__restrict__ colk) {
  const int w = 2 * n;
  const size_t tot = (size_t)n * w;
  for (size_t e = (size_t)blockIdx.x * 256 + threadIdx.x; e < tot; e += (size_t)gridDim.x * 256) {
    const int i = (int)(e / w), j = (int)(e % w);
    if (i == k) continue;
    g[e] -= colk[i] * g[(size_t)k * w + j];
  }
}

__global__ void k_gj_extract(int n, const double *__restrict__ g, double *__restrict__ inv) {
  int i = blockIdx.x;
  for (int j = threadIdx.x; j < n; j += blockDim.x) inv[(size_t)i * n + j] = g[(size_t)i * 2 * n + n + j];
}

// ---------------------------------------------------------------- cycle kernels
template <class T>
__device__ inline double ell_apply(const Op<T> &A, const T *__restrict__ x, int i) {
  return row_apply(A, x, i);
}
template <class T>
struct Sell {
  const int *soff, *wid, *col;
  const T *val;
  int n;
};

template <class T>
__device__ inline double csr_apply(const Sell<T> &A, const T *__restrict__ x, int i) {
  const int s = i >> 6;
  const long long base = (long long)A.soff[s] * 64 + (i & 63);
  const int w = A.wid[s];
  double acc = 0.0;
  for (int k = 0; k < w; ++k) {
    const long long e = base + (long long)k * 64;
    acc += (double)A.val[e] * (double)x[A.col[e]];
  }
  return acc;
}

template <class T>
__global__ void k_cy_first(int n, const T *__restrict__ dinv, const T *__restrict__ b, T *__restrict__ x,
                           double omega, const GState *st) {
  if (st && halted(st)) return;
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) x[i] = (T)(omega * (double)dinv[i] * (double)b[i]);
}

template <class T, bool ELL>
__global__ void __launch_bounds__(256) k_cy_sweep(Op<T> E, Sell<T> C, const T *__restrict__ dinv,
                                                  const T *__restrict__ b, const T *__restrict__ xin,
                                                  T *__restrict__ xout, double omega,
                                                  const GState *st) {
  if (st && halted(st)) return;
  int i = blockIdx.x * 256 + threadIdx.x;
  const int n = ELL ? E.M : C.n;
  if (i >= n) return;
  const double ax = ELL ? ell_apply(E, xin, i) : csr_apply(C, xin, i);
  xout[i] = (T)((double)xin[i] + omega * (double)dinv[i] * ((double)b[i] - ax));
}

template <class T, bool ELL>
__global__ void __launch_bounds__(256) k_cy_resid(Op<T> E, Sell<T> C, const T *__restrict__ b,
                                                  const T *__restrict__ x, T *__restrict__ r,
                                                  const GState *st) {
  if (st && halted(st)) return;
  int i = blockIdx.x * 256 + threadIdx.x;
  const int n = ELL ? E.M : C.n;
  if (i >= n) return;
  const double ax = ELL ? ell_apply(E, x, i) : csr_apply(C, x, i);
  r[i] = (T)((double)b[i] - ax);
}

template <class T>
__global__ void k_cy_restrict(int nc, const int *__restrict__ mp, const int *__restrict__ ml,
                              const T *__restrict__ r, T *__restrict__ bc, const GState *st) {
  if (st && halted(st)) return;
  int I = blockIdx.x * 256 + threadIdx.x;
  if (I >= nc) return;
  double s = 0.0;
  for (int e = mp[I]; e < mp[I + 1]; ++e) s += (double)r[ml[e]];
  bc[I] = (T)s;
}

template <class T>
__global__ void k_cy_copy(int n, const T *__restrict__ a, T *__restrict__ b, const GState *st) {
  if (st && halted(st)) return;
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) b[i] = a[i];
}

// x += alpha xc[agg]; rows flagged in pass (packed boxes) take alpha_pass
template <class T>
__global__ void k_cy_prolong(int n, const int *__restrict__ agg, const T *__restrict__ xc,
                             T *__restrict__ x, double alpha, const unsigned char *__restrict__ pass,
                             double alpha_pass, const GState *st) {
  if (st && halted(st)) return;
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const double a = (pass && pass[i]) ? alpha_pass : alpha;
  x[i] = (T)((double)x[i] + a * (double)xc[agg[i]]);
}

// Level 0's A P in compacted slices (k_ap_build): per slice its offset (in
// units of 64 entries) and width, per row its count
template <class T>
struct CSell {
  const int *soff, *wid;
  const unsigned char *cnt;
  const int *col;
  const T *val;
  int n;
};

// (A P e)_i on level 0, called by every live row of the wave (row i of the
// wave's slice): slot k's rows are the lanes whose count exceeds k, a row's
// entry the slot's start plus its rank among them.  Sixteen index and
// weight loads are issued before their gathers; the additions run in slot
// order.
template <class T>
__device__ inline double ap_row(const CSell<T> &P, const T *__restrict__ e, int i) {
  const int s = i >> 6;
  long long base = (long long)P.soff[s] * 64;
  const int w = P.wid[s];
  const int c = P.cnt[i];
  double acc = 0.0;
  for (int k0 = 0; k0 < w; k0 += 16) {
    int cc[16];
    T vv[16], ee[16];
#pragma unroll
    for (int u = 0; u < 16; ++u)
      if (k0 + u < w) {
        const bool on = k0 + u < c;
        const uint64_t mk = __ballot(on);
        if (on) {
          const long long q = base + below(mk);
          cc[u] = P.col[q];
          vv[u] = P.val[q];
        }
        base += __popcll(mk);
      }
#pragma unroll
    for (int u = 0; u < 16; ++u)
      if (k0 + u < c) ee[u] = e[cc[u]];
#pragma unroll
    for (int u = 0; u < 16; ++u)
      if (k0 + u < c) acc += (double)vv[u] * (double)ee[u];
  }
  return acc;
}

// Level 0 after a coarse correction e that another follows: x += alpha e[agg]
// and r -= alpha (A P e), i.e. r = b - A x for the corrected x without
// applying A again.
template <class T>
__global__ void __launch_bounds__(256) k_cy_ap_resid(int n, const int *__restrict__ agg, CSell<T> P,
                                                     const T *__restrict__ e, T *__restrict__ x,
                                                     T *__restrict__ r, double alpha,
                                                     const GState *st) {
  if (st && halted(st)) return;
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  x[i] = (T)((double)x[i] + alpha * (double)e[agg[i]]);
  r[i] = (T)((double)r[i] - alpha * ap_row(P, e, i));
}

// Level 0 without pre-smoothing, after the first of two coarse corrections:
// x = alpha e[agg] (x was 0) and r = b - alpha (A P) e (the residual of x = 0
// is b itself).
template <class T>
__global__ void __launch_bounds__(256) k_cy_ap_resid_first(int n, const int *__restrict__ agg,
                                                           CSell<T> P, const T *__restrict__ e,
                                                           const T *__restrict__ b,
                                                           T *__restrict__ x, T *__restrict__ r,
                                                           double alpha, const GState *st) {
  if (st && halted(st)) return;
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  x[i] = (T)(alpha * (double)e[agg[i]]);
  r[i] = (T)((double)b[i] - alpha * ap_row(P, e, i));
}

// Level 0 after the last coarse correction: the correction and the first
// post-smoothing step, x2 = x + alpha e[agg], xout = x2 + omega D^-1 (b - A x2)
// with b - A x2 = r - alpha (A P e) (r = b - A x on entry).
template <class T>
__global__ void __launch_bounds__(256) k_cy_ap_sweep(int n, const int *__restrict__ agg, CSell<T> P,
                                                     const T *__restrict__ e,
                                                     const T *__restrict__ dinv,
                                                     const T *__restrict__ xin,
                                                     const T *__restrict__ r, T *__restrict__ xout,
                                                     double alpha, double omega, const GState *st) {
  if (st && halted(st)) return;
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const T x2 = (T)((double)xin[i] + alpha * (double)e[agg[i]]);
  const double res = (double)r[i] - alpha * ap_row(P, e, i);
  xout[i] = (T)((double)x2 + omega * (double)dinv[i] * res);
}

// packed boxes, coarsest level: x_i = (inverse of row i's box block) b, one
// wave per row
template <class T>
__global__ void __launch_bounds__(64) k_cy_dense_blocks(const int *__restrict__ box,
                                                        const int *__restrict__ bstart,
                                                        const long long *__restrict__ ioff,
                                                        const T *__restrict__ inv,
                                                        const T *__restrict__ b,
                                                        T *__restrict__ x, const GState *st) {
  if (st && halted(st)) return;
  const int i = blockIdx.x, lane = threadIdx.x;
  const int bx = box[i], r0 = bstart[bx], n = bstart[bx + 1] - r0;
  const T *row = inv + ioff[bx] + (size_t)(i - r0) * n;
  double s = 0.0;
  for (int j = lane; j < n; j += 64) s += (double)row[j] * (double)b[r0 + j];
  s = wave_sum(s);
  if (lane == 0) x[i] = (T)s;
}

template <class T>
__global__ void __launch_bounds__(64) k_cy_dense(int n, const T *__restrict__ inv, const T *__restrict__ b,
                                                 T *__restrict__ x, const GState *st) {
  if (st && halted(st)) return;
  const int i = blockIdx.x, lane = threadIdx.x;
  double s = 0.0;
  for (int j = lane; j < n; j += 64) s += (double)inv[(size_t)i * n + j] * (double)b[j];
  s = wave_sum(s);
  if (lane == 0) x[i] = (T)s;
}

// ---- small coarse levels: G lanes per row (per coarse row for the
// restriction), so each lane's chain of dependent gathers is G times shorter;
// the lanes' partial sums meet in a fixed shuffle tree (deterministic)
// (A x)_i with G lanes per row; xf(j) gives x_j (a load, or a value formed
// from loads at j)
template <class T, int G, class XF>
__device__ inline double sell_apply_gx(const Sell<T> &A, XF xf, int i, int sub) {
  const int sl = i >> 6;
  const long long base = (long long)A.soff[sl] * 64 + (i & 63);
  const int w = A.wid[sl];
  double acc = 0.0;
  // a lane's slots sub, sub+G, ... in groups of four: the group's column and
  // value loads, then its gathers, then the additions in slot order
  int k = sub;
  for (; k + 3 * G < w; k += 4 * G) {
    int cj[4];
    T av[4], xv[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const long long e = base + (long long)(k + u * G) * 64;
      cj[u] = A.col[e];
      av[u] = A.val[e];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) xv[u] = xf(cj[u]);
#pragma unroll
    for (int u = 0; u < 4; ++u) acc += (double)av[u] * (double)xv[u];
  }
  for (; k < w; k += G) {
    const long long e = base + (long long)k * 64;
    acc += (double)A.val[e] * (double)xf(A.col[e]);
  }
#pragma unroll
  for (int o = G / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o, G);
  return acc;
}

template <class T, int G>
__device__ inline double sell_apply_g(const Sell<T> &A, const T *__restrict__ x, int i, int sub) {
  return sell_apply_gx<T, G>(A, [&](int j) { return x[j]; }, i, sub);
}

// The first smoothing step and the residual after it, in one launch: x =
// omega D^-1 b and r = b - A x, each neighbour's x_j formed from b_j and
// dinv_j exactly as the first step stores it (so r is the same as from the
// two launches), the row's own x_i stored for the correction that follows.
template <class T, int G>
__global__ void __launch_bounds__(256) k_cy_first_resid_g(Sell<T> C, const T *__restrict__ dinv,
                                                          const T *__restrict__ b,
                                                          T *__restrict__ x, T *__restrict__ r,
                                                          double omega, const GState *st) {
  if (st && halted(st)) return;
  const int t = blockIdx.x * 256 + threadIdx.x;
  const int i = t / G, sub = t % G;
  if (i >= C.n) return;  // whole groups (G divides 64)
  auto xf = [&](int j) { return (T)(omega * (double)dinv[j] * (double)b[j]); };
  const double ax = sell_apply_gx<T, G>(C, xf, i, sub);
  if (sub == 0) {
    x[i] = xf(i);
    r[i] = (T)((double)b[i] - ax);
  }
}

template <class T, int G>
__global__ void __launch_bounds__(256) k_cy_sweep_g(Sell<T> C, const T *__restrict__ dinv,
                                                    const T *__restrict__ b,
                                                    const T *__restrict__ xin,
                                                    T *__restrict__ xout, double omega,
                                                    const GState *st) {
  if (st && halted(st)) return;
  const int t = blockIdx.x * 256 + threadIdx.x;
  const int i = t / G, sub = t % G;
  if (i >= C.n) return;  // whole groups (G divides 64)
  const double ax = sell_apply_g<T, G>(C, xin, i, sub);
  if (sub == 0) xout[i] = (T)((double)xin[i] + omega * (double)dinv[i] * ((double)b[i] - ax));
}

template <class T, int G>
__global__ void __launch_bounds__(256) k_cy_resid_g(Sell<T> C, const T *__restrict__ b,
                                                    const T *__restrict__ x, T *__restrict__ r,
                                                    const GState *st) {
  if (st && halted(st)) return;
  const int t = blockIdx.x * 256 + threadIdx.x;
  const int i = t / G, sub = t % G;
  if (i >= C.n) return;
  const double ax = sell_apply_g<T, G>(C, x, i, sub);
  if (sub == 0) r[i] = (T)((double)b[i] - ax);
}

template <class T, int G>
__global__ void __launch_bounds__(256) k_cy_restrict_g(int nc, const int *__restrict__ mp,
                                                       const int *__restrict__ ml,
                                                       const T *__restrict__ r, T *__restrict__ bc,
                                                       const GState *st) {
  if (st && halted(st)) return;
  const int t = blockIdx.x * 256 + threadIdx.x;
  const int I = t / G, sub = t % G;
  if (I >= nc) return;
  double s = 0.0;
  for (int e = mp[I] + sub; e < mp[I + 1]; e += G) s += (double)r[ml[e]];
#pragma unroll
  for (int o = G / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, G);
  if (sub == 0) bc[I] = (T)s;
}

// sum over J of (A P)_iJ a_J e_J on a coarse level (G lanes per row): a_J =
// alpha, or alpha_pass where apass flags aggregate J (packed boxes)
template <class T, int G>
__device__ inline double ap_apply_g(const Sell<T> &P, const T *__restrict__ e,
                                    const unsigned char *__restrict__ apass, double alpha,
                                    double alpha_pass, int i, int sub) {
  if (!apass) return alpha * sell_apply_g<T, G>(P, e, i, sub);
  return sell_apply_gx<T, G>(
      P, [&](int J) { return (T)((apass[J] ? alpha_pass : alpha) * (double)e[J]); }, i, sub);
}

// Coarse level after a coarse correction that another follows: x += a_i
// e[agg_i] and r -= (A P) (a e) (k_cy_ap_resid with G lanes per row and the
// packed boxes' pass rows)
// FIRST: the level's visit has no pre-smoothing, x was 0 and r is b.
template <class T, int G, bool FIRST>
__global__ void __launch_bounds__(256) k_cy_ap_resid_g(Sell<T> P, const int *__restrict__ agg,
                                                       const T *__restrict__ e,
                                                       const T *__restrict__ b, T *__restrict__ x,
                                                       T *__restrict__ r, double alpha,
                                                       const unsigned char *__restrict__ pass,
                                                       const unsigned char *__restrict__ apass,
                                                       double alpha_pass, const GState *st) {
  if (st && halted(st)) return;
  const int t = blockIdx.x * 256 + threadIdx.x;
  const int i = t / G, sub = t % G;
  if (i >= P.n) return;  // whole groups (G divides 64)
  const double ape = ap_apply_g<T, G>(P, e, apass, alpha, alpha_pass, i, sub);
  if (sub == 0) {
    const double a = (pass && pass[i]) ? alpha_pass : alpha;
    x[i] = (T)((FIRST ? 0.0 : (double)x[i]) + a * (double)e[agg[i]]);
    r[i] = (T)((double)(FIRST ? b[i] : r[i]) - ape);
  }
}

// Coarse level after the last coarse correction: the correction and the
// first post-smoothing step (k_cy_ap_sweep with G lanes per row)
template <class T, int G, bool FIRST>
__global__ void __launch_bounds__(256) k_cy_ap_sweep_g(Sell<T> P, const int *__restrict__ agg,
                                                       const T *__restrict__ e,
                                                       const T *__restrict__ dinv,
                                                       const T *__restrict__ xin,
                                                       const T *__restrict__ r,
                                                       T *__restrict__ xout, double alpha,
                                                       double omega,
                                                       const unsigned char *__restrict__ pass,
                                                       const unsigned char *__restrict__ apass,
                                                       double alpha_pass, const GState *st) {
  if (st && halted(st)) return;
  const int t = blockIdx.x * 256 + threadIdx.x;
  const int i = t / G, sub = t % G;
  if (i >= P.n) return;
  const double ape = ap_apply_g<T, G>(P, e, apass, alpha, alpha_pass, i, sub);
  if (sub == 0) {
    const double a = (pass && pass[i]) ? alpha_pass : alpha;
    const T x2 = (T)((FIRST ? 0.0 : (double)xin[i]) + a * (double)e[agg[i]]);
    xout[i] = (T)((double)x2 + omega * (double)dinv[i] * ((double)r[i] - ape));
  }
}

// ---------------------------------------------------------------- host side
static int ilog2_ceil(uint64_t v) {
  int b = 0;
  while (b < 64 && (v >> b) != 0) ++b;
  return std::max(b, 1);
}

// Packed boxes: the MIS of a level starts with every row of a stopped box a
// root, so its rows become singleton aggregates and its block is carried
// unchanged.
__global__ void k_amg_freeze_roots(int n, const int *__restrict__ box,
                                   const int *__restrict__ frozen, int *__restrict__ a) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) a[i] = frozen[box[i]] ? 1 : 0;
}

template <class V>
static int aggregate_level(apc_ctx *c, AmgHier *h, const V &v, int n, AmgLevel *Lf, int *nc_out,
                           const int *box = nullptr) {
  const AmgOptions &o = h->o;
  const unsigned g = nblocks(n, 256);
  APC_ALLOC(c, h->smax, n + 1);
  APC_ALLOC(c, h->sa, n + 1);
  APC_ALLOC(c, h->sb, n + 1);
  APC_ALLOC(c, h->flag, n + 1);
  APC_ALLOC(c, h->pos, n + 1);
  APC_ALLOC(c, h->cnt, 4);
  if (o.theta > 0) hipLaunchKernelGGL((k_amg_smax<V>), dim3(g), dim3(256), 0, c->s, v, h->smax.p);
  else APC_HIP(c, hipMemsetAsync(h->smax.p, 0, sizeof(double) * n, c->s));
  if (box)  // packed boxes: the rows of stopped boxes start as roots
    hipLaunchKernelGGL(k_amg_freeze_roots, dim3(g), dim3(256), 0, c->s, n, box, h->frozen.p,
                       h->sa.p);
  else
    APC_HIP(c, hipMemsetAsync(h->sa.p, 0, sizeof(int) * n, c->s));
  // rounds go out in batches of kRounds with one host read per batch: a
  // round whose predecessor left no undecided row returns at once, and the
  // set is read from the buffer of the first round that settled it
  constexpr int kRounds = 8;
  APC_ALLOC(c, h->rcnt, kRounds + 1);
  int *a = h->sa.p, *b = h->sb.p;
  for (int batch = 0;; ++batch) {
    if (batch * kRounds > 200)
      return fail(c, APC_ERR_UNSUPPORTED, "amg: independent set did not settle");
    APC_HIP(c, hipMemsetAsync(h->rcnt.p, 0, sizeof(int) * kRounds, c->s));
    int *after[kRounds];
    for (int k = 0; k < kRounds; ++k) {
      const int *prev = k ? h->rcnt.p + k - 1 : nullptr;
      hipLaunchKernelGGL((k_amg_pick<V>), dim3(g), dim3(256), 0, c->s, v, o.theta, h->smax.p, a, b,
                         prev);
      hipLaunchKernelGGL((k_amg_cover<V>), dim3(std::min(g, 2048u)), dim3(256), 0, c->s, v,
                         o.theta, h->smax.p, b, h->rcnt.p + k, prev);
      after[k] = b;
      std::swap(a, b);
    }
    int und[kRounds];
    APC_HIP(c, hipMemcpyAsync(und, h->rcnt.p, sizeof(und), hipMemcpyDeviceToHost, c->s));
    APC_HIP(c, hipStreamSynchronize(c->s));
    int settled = -1;
    for (int k = 0; k < kRounds && settled < 0; ++k)
      if (und[k] == 0) settled = k;
    if (settled >= 0) {
      a = after[settled];
      break;
    }
  }
  hipLaunchKernelGGL(k_amg_rootflag, dim3(g), dim3(256), 0, c->s, n, a, h->flag.p);
  int r = scan_exclusive(c, h->flag.p, h->pos.p, n, h->cnt.p + 1);
  if (r) return r;
  APC_ALLOC(c, Lf->agg, n + 1);
  APC_HIP(c, hipMemsetAsync(h->cnt.p + 2, 0, sizeof(int), c->s));
  hipLaunchKernelGGL((k_amg_agg<V>), dim3(g), dim3(256), 0, c->s, v, o.theta, h->smax.p, a, h->pos.p,
                     Lf->agg.p, h->cnt.p + 2);
  int hc[3] = {0, 0, 0};
  APC_HIP(c, hipMemcpyAsync(hc, h->cnt.p, sizeof(hc), hipMemcpyDeviceToHost, c->s));
  APC_HIP(c, hipStreamSynchronize(c->s));
  if (hc[2]) return fail(c, APC_ERR_UNSUPPORTED, "amg: %d rows without a root neighbour", hc[2]);
  *nc_out = hc[1];
  return 0;
}

// ---- geometric aggregation: an aggregate is every row whose octree cell
// lies in one ancestor cell `up` levels above it (the rows are the octree's
// computational nodes, so a level-0 aggregate is a compact block of up to
// 8^up leaves, and a coarse row is its ancestor cell).

// level 0: each row's cell from its centre and width (width = extent / 2^L)
__global__ void k_geo_cells0(int n, const double *__restrict__ cx, const double *__restrict__ cy,
                             const double *__restrict__ cz, const double *__restrict__ ch,
                             double x0, double ex, double y0, double ey, double z0, double ez,
                             int4 *__restrict__ cell) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const int L = min(max((int)rint(log2(ex / ch[i])), 0), 21);
  const double s = ldexp(1.0, L);
  const int top = (1 << L) - 1;
  const int ix = min(max((int)floor((cx[i] - x0) / ex * s), 0), top);
  const int iy = min(max((int)floor((cy[i] - y0) / ey * s), 0), top);
  const int iz = min(max((int)floor((cz[i] - z0) / ez * s), 0), top);
  cell[i] = make_int4(ix, iy, iz, L);
}

// key of row i's ancestor cell: the Morton code of its corner at R bits per
// axis (R = the deepest ancestor level present, at most kGeoRes), then its
// level (5 bits); the sort groups an ancestor's rows and orders the
// ancestors along the Morton curve.  bad counts cells whose ancestor lies
// deeper than R.
constexpr int kGeoRes = 19;
__device__ inline uint64_t geo_key(int4 c, int up, int R) {
  const int la = max(c.w - up, 0), sh = c.w - la;
  const int up2 = R - min(la, R);
  const uint64_t m = spread3((uint32_t)(c.x >> sh) << up2) |
                     (spread3((uint32_t)(c.y >> sh) << up2) << 1) |
                     (spread3((uint32_t)(c.z >> sh) << up2) << 2);
  return (m << 5) | (uint64_t)la;
}

// the deepest ancestor level over the rows (*lev starts at 0), so that the
// keys need only 3 bits per level of it
__global__ void k_geo_maxlev(int n, const int4 *__restrict__ cell, int up, int *__restrict__ lev) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  int la = (i < n) ? max(cell[i].w - up, 0) : 0;
  for (int o = 32; o > 0; o >>= 1) la = max(la, __shfl_xor(la, o, 64));
  if ((threadIdx.x & 63) == 0) atomicMax(lev, la);
}

__global__ void k_geo_keys(int n, const int4 *__restrict__ cell, int up, int R,
                           uint64_t *__restrict__ key, int *__restrict__ idx,
                           int *__restrict__ bad) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const int4 c = cell[i];
  if (c.w - up > R) atomicAdd(bad, 1);
  key[i] = geo_key(c, up, R);
  idx[i] = i;
}

// packed boxes: after the sort by cell, the box of each sorted row as the
// key of a last stable pass (box-major aggregates, never two boxes in one)
__global__ void k_geo_boxkeys(int n, const int *__restrict__ idx, const int *__restrict__ box,
                              uint64_t *__restrict__ key) {
  const int s = blockIdx.x * 256 + threadIdx.x;
  if (s < n) key[s] = (uint64_t)box[idx[s]];
}

__global__ void k_geo_flags_box(int n, const int *__restrict__ idx, const int *__restrict__ box,
                                const int4 *__restrict__ cell, int up, int R,
                                int *__restrict__ f) {
  const int s = blockIdx.x * 256 + threadIdx.x;
  if (s >= n) return;
  if (s == 0) { f[0] = 1; return; }
  const int i = idx[s], j = idx[s - 1];
  f[s] = (box[i] != box[j] || geo_key(cell[i], up, R) != geo_key(cell[j], up, R)) ? 1 : 0;
}

__global__ void k_geo_flags(int n, const uint64_t *__restrict__ key, int *__restrict__ f) {
  const int s = blockIdx.x * 256 + threadIdx.x;
  if (s < n) f[s] = (s == 0 || key[s] != key[s - 1]) ? 1 : 0;
}

// sorted position s: its row joins aggregate a (the runs of equal keys in
// order); the members list is the sorted rows (row order within a run: the
// sort is stable), a run's start its offset; the run's first row writes
// the aggregate's cell
__global__ void k_geo_assign(int n, int nc, const int *__restrict__ idx,
                             const int *__restrict__ f, const int *__restrict__ pos,
                             const int4 *__restrict__ cell, int up, int *__restrict__ agg,
                             int *__restrict__ mp, int *__restrict__ ml,
                             int4 *__restrict__ ccell) {
  const int s = blockIdx.x * 256 + threadIdx.x;
  if (s >= n) return;
  const int i = idx[s];
  const int a = pos[s] + f[s] - 1;
  agg[i] = a;
  ml[s] = i;
  if (f[s]) {
    mp[a] = s;
    const int4 c = cell[i];
    const int la = max(c.w - up, 0), sh = c.w - la;
    ccell[a] = make_int4(c.x >> sh, c.y >> sh, c.z >> sh, la);
  }
  if (s == n - 1) mp[nc] = n;
}


// Level l's aggregates by ancestor cells `up` levels above its rows' cells
// (level 0: the computational nodes' cells; below: the cells the previous
// call left in h->gcell).  Writes agg, the members (mp, ml) and the next
// level's cells; *nc_out = the number of aggregates, or -1 (nothing
// written) where a cell lies too deep for the keys.  box (packed boxes):
// each row's box; the aggregates are then box-major.
static int aggregate_geo(apc_ctx *c, AmgHier *h, int l, int n, int up, AmgLevel *Lf,
                         int *nc_out, const int *box) {
  const unsigned g = nblocks(n, 256);
  APC_ALLOC(c, h->key, n + 1);
  APC_ALLOC(c, h->key2, n + 1);
  APC_ALLOC(c, h->idx, n + 1);
  APC_ALLOC(c, h->idx2, n + 1);
  APC_ALLOC(c, h->flag, n + 1);
  APC_ALLOC(c, h->pos, n + 1);
  APC_ALLOC(c, h->cnt, 4);
  APC_ALLOC(c, h->gcell2, n + 1);
  if (l == 0) {
    APC_ALLOC(c, h->gcell, n + 1);
    const double *bb = c->prm.bbox;
    hipLaunchKernelGGL(k_geo_cells0, dim3(g), dim3(256), 0, c->s, n, c->cx.p, c->cy.p, c->cz.p,
                       c->ch.p, bb[0], bb[1] - bb[0], bb[2], bb[3] - bb[2], bb[4], bb[5] - bb[4],
                       h->gcell.p);
  }
  APC_HIP(c, hipMemsetAsync(h->cnt.p, 0, 4 * sizeof(int), c->s));
  hipLaunchKernelGGL(k_geo_maxlev, dim3(g), dim3(256), 0, c->s, n, (const int4 *)h->gcell.p, up,
                     h->cnt.p + 3);
  int R = 0;
  APC_HIP(c, hipMemcpyAsync(&R, h->cnt.p + 3, sizeof(int), hipMemcpyDeviceToHost, c->s));
  APC_HIP(c, hipStreamSynchronize(c->s));
  if (R > kGeoRes) {  // cells deeper than the keys hold: the caller aggregates by sets
    *nc_out = -1;
    return 0;
  }
  hipLaunchKernelGGL(k_geo_keys, dim3(g), dim3(256), 0, c->s, n, (const int4 *)h->gcell.p, up, R,
                     h->key.p, h->idx.p, h->cnt.p + 2);
  int r = radix_sort_pairs(c, h->key.p, h->idx.p, h->key2.p, h->idx2.p, n, 3 * R + 5);
  if (r) return r;
  if (box) {
    hipLaunchKernelGGL(k_geo_boxkeys, dim3(g), dim3(256), 0, c->s, n, (const int *)h->idx.p, box,
                       h->key.p);
    r = radix_sort_pairs(c, h->key.p, h->idx.p, h->key2.p, h->idx2.p, n, std::max(1, c->boxbits));
    if (r) return r;
    hipLaunchKernelGGL(k_geo_flags_box, dim3(g), dim3(256), 0, c->s, n, (const int *)h->idx.p,
                       box, (const int4 *)h->gcell.p, up, R, h->flag.p);
  } else {
    hipLaunchKernelGGL(k_geo_flags, dim3(g), dim3(256), 0, c->s, n, (const uint64_t *)h->key.p,
                       h->flag.p);
  }
  r = scan_exclusive(c, h->flag.p, h->pos.p, n, h->cnt.p + 1);
  if (r) return r;
  int hc[3] = {0, 0, 0};
  APC_HIP(c, hipMemcpyAsync(hc, h->cnt.p, sizeof(hc), hipMemcpyDeviceToHost, c->s));
  APC_HIP(c, hipStreamSynchronize(c->s));
  if (hc[2]) return fail(c, APC_ERR_UNSUPPORTED, "amg: %d cells too deep to aggregate", hc[2]);
  const int nc = hc[1];
  APC_ALLOC(c, Lf->agg, n + 1);
  APC_ALLOC(c, Lf->mp, nc + 1);
  APC_ALLOC(c, Lf->ml, n + 1);
  hipLaunchKernelGGL(k_geo_assign, dim3(g), dim3(256), 0, c->s, n, nc, (const int *)h->idx.p,
                     (const int *)h->flag.p, (const int *)h->pos.p, (const int4 *)h->gcell.p, up,
                     Lf->agg.p, Lf->mp.p, Lf->ml.p, h->gcell2.p);
  APC_HIP(c, hipGetLastError());
  std::swap(h->gcell, h->gcell2);
  *nc_out = nc;
  return 0;
}

// rows of each aggregate, grouped in row order (restriction order)
static int group_members(apc_ctx *c, AmgHier *h, AmgLevel *Lf, int n, int nc) {
  APC_ALLOC(c, h->key, n + 1);
  APC_ALLOC(c, h->key2, n + 1);
  APC_ALLOC(c, h->idx, n + 1);
  APC_ALLOC(c, h->idx2, n + 1);
  APC_ALLOC(c, Lf->mp, nc + 1);
  APC_ALLOC(c, Lf->ml, n + 1);
  const unsigned g = nblocks(n, 256);
  hipLaunchKernelGGL(k_amg_keys_rows, dim3(g), dim3(256), 0, c->s, n, Lf->agg.p, h->key.p, h->idx.p);
  int r = radix_sort_pairs(c, h->key.p, h->idx.p, h->key2.p, h->idx2.p, n, ilog2_ceil((uint64_t)nc));
  if (r) return r;
  hipLaunchKernelGGL(k_amg_bounds, dim3(g), dim3(256), 0, c->s, n, nc, h->key.p, Lf->mp.p);
  APC_HIP(c, hipMemcpyAsync(Lf->ml.p, h->idx.p, sizeof(int) * n, hipMemcpyDeviceToDevice, c->s));
  return 0;
}

// Galerkin product into level Lc (CSR) by coarse rows (k_gal_count,
// k_gal_rowoff, k_gal_emit, k_gal_rows).
template <class V>
static int galerkin_rows(apc_ctx *c, AmgHier *h, const V &v, AmgLevel *Lf, int n, int nc,
                         AmgLevel *Lc, bool coarse_cols = false) {
  if ((size_t)Lf->nnz >= (size_t)INT32_MAX)  // entry offsets are int
    return fail(c, APC_ERR_UNSUPPORTED, "amg: %d stored entries on one level", Lf->nnz);
  APC_ALLOC(c, h->flag, (size_t)n + 1);  // entry counts, then coarse-row counts
  APC_ALLOC(c, h->pos, (size_t)n + 1);   // entry offsets in member order
  APC_ALLOC(c, h->sa, (size_t)n + 1);    // entry offsets by fine row
  APC_ALLOC(c, h->cnt, 4);
  const unsigned g = nblocks(n, 256);
  hipLaunchKernelGGL((k_gal_count<V>), dim3(g), dim3(256), 0, c->s, v, n, h->sa.p);
  hipLaunchKernelGGL(k_gal_gather, dim3(g), dim3(256), 0, c->s, n, Lf->ml.p, h->sa.p, h->flag.p);
  int r = scan_exclusive(c, h->flag.p, h->pos.p, n, h->cnt.p);
  if (r) return r;
  int ne = 0;
  APC_HIP(c, hipMemcpyAsync(&ne, h->cnt.p, sizeof(int), hipMemcpyDeviceToHost, c->s));
  APC_HIP(c, hipStreamSynchronize(c->s));
  APC_HIP(c, hipMemcpyAsync(h->pos.p + n, h->cnt.p, sizeof(int), hipMemcpyDeviceToDevice, c->s));
  APC_ALLOC(c, h->gcol, (size_t)ne + 1);
  APC_ALLOC(c, h->gval, (size_t)ne + 1);
  hipLaunchKernelGGL(k_gal_rowoff, dim3(g), dim3(256), 0, c->s, n, Lf->ml.p, h->pos.p, h->sa.p);
  hipLaunchKernelGGL((k_gal_emit<V>), dim3(g), dim3(256), 0, c->s, v, n,
                     coarse_cols ? (const int *)nullptr : (const int *)Lf->agg.p, h->sa.p,
                     h->gcol.p, h->gval.p);
  const unsigned grid = (unsigned)std::min(nc, 1 << 16);
  APC_ALLOC(c, h->diag, nc + 1);
  APC_HIP(c, hipMemsetAsync(h->diag.p, 0, sizeof(double) * nc, c->s));  // both passes write it
  hipLaunchKernelGGL(k_gal_rows, dim3(grid), dim3(64), 0, c->s, nc, Lf->mp.p, h->pos.p, h->gcol.p,
                     h->gval.p, 0, h->flag.p, (const int *)nullptr, (int *)nullptr,
                     (double *)nullptr, h->diag.p);
  APC_ALLOC(c, h->sb, (size_t)nc + 1);
  r = scan_exclusive(c, h->flag.p, h->sb.p, nc, h->cnt.p + 1);
  if (r) return r;
  int nnz = 0;
  APC_HIP(c, hipMemcpyAsync(&nnz, h->cnt.p + 1, sizeof(int), hipMemcpyDeviceToHost, c->s));
  APC_HIP(c, hipStreamSynchronize(c->s));
  Lc->n = nc;
  Lc->nnz = nnz;
  APC_ALLOC(c, Lc->rp, nc + 1);
  APC_ALLOC(c, Lc->col, nnz + 1);
  APC_ALLOC(c, Lc->val, nnz + 1);
  APC_ALLOC(c, Lc->dinv, nc + 1);
  APC_HIP(c, hipMemcpyAsync(Lc->rp.p, h->sb.p, sizeof(int) * nc, hipMemcpyDeviceToDevice, c->s));
  APC_HIP(c, hipMemcpyAsync(Lc->rp.p + nc, h->cnt.p + 1, sizeof(int), hipMemcpyDeviceToDevice, c->s));
  hipLaunchKernelGGL(k_gal_rows, dim3(grid), dim3(64), 0, c->s, nc, Lf->mp.p, h->pos.p, h->gcol.p,
                     h->gval.p, 1, (int *)nullptr, (const int *)Lc->rp.p, Lc->col.p, Lc->val.p,
                     h->diag.p);
  hipLaunchKernelGGL(k_amg_dinv, dim3(nblocks(nc, 256)), dim3(256), 0, c->s, nc, h->diag.p,
                     Lc->dinv.p);
  APC_HIP(c, hipGetLastError());
  return 0;
}

// Galerkin product of the ELL level into Lc (CSR) by coarse rows, the
// entries formed in place (k_gal_rows_ell)
static int galerkin_rows_ell(apc_ctx *c, AmgHier *h, const Op<double> &A, AmgLevel *Lf, int n,
                             int nc, AmgLevel *Lc) {
  const size_t W = (size_t)A.lm + 1;
  if (W * (size_t)n >= (size_t)INT32_MAX)  // coarse-row offsets are int
    return fail(c, APC_ERR_UNSUPPORTED, "amg: %zu entries on the ELL level", W * (size_t)n);
  APC_ALLOC(c, h->gcol, W * n + 1);
  APC_ALLOC(c, h->gval, W * n + 1);
  APC_ALLOC(c, h->flag, (size_t)nc + 1);
  APC_ALLOC(c, h->cnt, 4);
  APC_ALLOC(c, h->diag, nc + 1);
  APC_HIP(c, hipMemsetAsync(h->diag.p, 0, sizeof(double) * nc, c->s));  // both passes write it
  const unsigned grid = (unsigned)std::min(nc, 1 << 16);
  hipLaunchKernelGGL(k_gal_rows_ell, dim3(grid), dim3(64), 0, c->s, nc, A, Lf->agg.p, Lf->mp.p,
                     Lf->ml.p, 0, h->flag.p, (const int *)nullptr, (int *)nullptr,
                     (double *)nullptr, h->gcol.p, h->gval.p, h->diag.p);
  APC_ALLOC(c, h->sb, (size_t)nc + 1);
  int r = scan_exclusive(c, h->flag.p, h->sb.p, nc, h->cnt.p + 1);
  if (r) return r;
  int nnz = 0;
  APC_HIP(c, hipMemcpyAsync(&nnz, h->cnt.p + 1, sizeof(int), hipMemcpyDeviceToHost, c->s));
  APC_HIP(c, hipStreamSynchronize(c->s));
  Lc->n = nc;
  Lc->nnz = nnz;
  APC_ALLOC(c, Lc->rp, nc + 1);
  APC_ALLOC(c, Lc->col, nnz + 1);
  APC_ALLOC(c, Lc->val, nnz + 1);
  APC_ALLOC(c, Lc->dinv, nc + 1);
  APC_HIP(c, hipMemcpyAsync(Lc->rp.p, h->sb.p, sizeof(int) * nc, hipMemcpyDeviceToDevice, c->s));
  APC_HIP(c, hipMemcpyAsync(Lc->rp.p + nc, h->cnt.p + 1, sizeof(int), hipMemcpyDeviceToDevice, c->s));
  hipLaunchKernelGGL(k_gal_rows_ell, dim3(grid), dim3(64), 0, c->s, nc, A, Lf->agg.p, Lf->mp.p,
                     Lf->ml.p, 1, (int *)nullptr, (const int *)Lc->rp.p, Lc->col.p, Lc->val.p,
                     h->gcol.p, h->gval.p, h->diag.p);
  hipLaunchKernelGGL(k_amg_dinv, dim3(nblocks(nc, 256)), dim3(256), 0, c->s, nc, h->diag.p,
                     Lc->dinv.p);
  APC_HIP(c, hipGetLastError());
  return 0;
}

// P^T (A P) of level 0 into Lc (CSR) from A P's compacted slices (k_gal_rows_ap)
static int galerkin_rows_ap(apc_ctx *c, AmgHier *h, AmgLevel *Lf, int n, int nc, AmgLevel *Lc) {
  const size_t W = (size_t)kApMax;
  if (W * (size_t)n >= (size_t)INT32_MAX)  // coarse-row offsets are int
    return fail(c, APC_ERR_UNSUPPORTED, "amg: %zu entries of A P", W * (size_t)n);
  APC_ALLOC(c, h->gcol, W * n + 1);
  APC_ALLOC(c, h->gval, W * n + 1);
  APC_ALLOC(c, h->flag, (size_t)nc + 1);
  APC_ALLOC(c, h->cnt, 4);
  APC_ALLOC(c, h->diag, nc + 1);
  APC_HIP(c, hipMemsetAsync(h->diag.p, 0, sizeof(double) * nc, c->s));  // both passes write it
  const unsigned grid = (unsigned)std::min(nc, 1 << 16);
  hipLaunchKernelGGL(k_gal_rows_ap, dim3(grid), dim3(64), 0, c->s, nc, (const int *)Lf->apoff.p,
                     (const int *)Lf->apcol.p, (const double *)Lf->apval.p,
                     (const unsigned char *)Lf->apn.p, (const int *)Lf->apss.p,
                     (const unsigned long long *)Lf->apmask.p, (const int *)Lf->mp.p,
                     (const int *)Lf->ml.p, 0,
                     h->flag.p, (const int *)nullptr, (int *)nullptr, (double *)nullptr,
                     h->gcol.p, h->gval.p, h->diag.p);
  APC_ALLOC(c, h->sb, (size_t)nc + 1);
  int r = scan_exclusive(c, h->flag.p, h->sb.p, nc, h->cnt.p + 1);
  if (r) return r;
  int nnz = 0;
  APC_HIP(c, hipMemcpyAsync(&nnz, h->cnt.p + 1, sizeof(int), hipMemcpyDeviceToHost, c->s));
  APC_HIP(c, hipStreamSynchronize(c->s));
  Lc->n = nc;
  Lc->nnz = nnz;
  APC_ALLOC(c, Lc->rp, nc + 1);
  APC_ALLOC(c, Lc->col, nnz + 1);
  APC_ALLOC(c, Lc->val, nnz + 1);
  APC_ALLOC(c, Lc->dinv, nc + 1);
  APC_HIP(c, hipMemcpyAsync(Lc->rp.p, h->sb.p, sizeof(int) * nc, hipMemcpyDeviceToDevice, c->s));
  APC_HIP(c, hipMemcpyAsync(Lc->rp.p + nc, h->cnt.p + 1, sizeof(int), hipMemcpyDeviceToDevice, c->s));
  hipLaunchKernelGGL(k_gal_rows_ap, dim3(grid), dim3(64), 0, c->s, nc, (const int *)Lf->apoff.p,
                     (const int *)Lf->apcol.p, (const double *)Lf->apval.p,
                     (const unsigned char *)Lf->apn.p, (const int *)Lf->apss.p,
                     (const unsigned long long *)Lf->apmask.p, (const int *)Lf->mp.p,
                     (const int *)Lf->ml.p, 1,
                     (int *)nullptr, (const int *)Lc->rp.p, Lc->col.p, Lc->val.p, h->gcol.p,
                     h->gval.p, h->diag.p);
  hipLaunchKernelGGL(k_amg_dinv, dim3(nblocks(nc, 256)), dim3(256), 0, c->s, nc, h->diag.p,
                     Lc->dinv.p);
  APC_HIP(c, hipGetLastError());
  return 0;
}

// Galerkin product into level Lc (CSR) from `ne` generated entries
static int galerkin_reduce(apc_ctx *c, AmgHier *h, size_t ne, int nc, AmgLevel *Lc) {
  if (ne >= (size_t)INT32_MAX) return fail(c, APC_ERR_UNSUPPORTED, "amg: too many entries");
  const int n = (int)ne;
  const uint64_t lim = (uint64_t)nc * (uint64_t)nc;
  int r = radix_sort_pairs(c, h->key.p, h->idx.p, h->key2.p, h->idx2.p, n, ilog2_ceil(lim + 1));
  if (r) return r;
  APC_ALLOC(c, h->flag, ne + 1);
  APC_ALLOC(c, h->pos, ne + 1);
  APC_ALLOC(c, h->cnt, 4);
  const unsigned g = nblocks(n, 256);
  hipLaunchKernelGGL(k_amg_heads, dim3(g), dim3(256), 0, c->s, n, lim, h->key.p, h->flag.p);
  r = scan_exclusive(c, h->flag.p, h->pos.p, n, h->cnt.p);
  if (r) return r;
  int nnz = 0;
  APC_HIP(c, hipMemcpyAsync(&nnz, h->cnt.p, sizeof(int), hipMemcpyDeviceToHost, c->s));
  APC_HIP(c, hipStreamSynchronize(c->s));
  Lc->n = nc;
  Lc->nnz = nnz;
  APC_ALLOC(c, Lc->rp, nc + 1);
  APC_ALLOC(c, Lc->col, nnz + 1);
  APC_ALLOC(c, Lc->val, nnz + 1);
  APC_ALLOC(c, Lc->dinv, nc + 1);
  APC_ALLOC(c, h->diag, nc + 1);
  APC_HIP(c, hipMemsetAsync(h->diag.p, 0, sizeof(double) * nc, c->s));
  hipLaunchKernelGGL(k_amg_reduce, dim3(g), dim3(256), 0, c->s, n, nc, h->key.p, h->idx.p, h->ev.p,
                     h->flag.p, h->pos.p, Lc->rp.p, Lc->col.p, Lc->val.p, h->diag.p);
  APC_HIP(c, hipMemcpyAsync(Lc->rp.p + nc, &Lc->nnz, sizeof(int), hipMemcpyHostToDevice, c->s));
  hipLaunchKernelGGL(k_amg_dinv, dim3(nblocks(nc, 256)), dim3(256), 0, c->s, nc, h->diag.p,
                     Lc->dinv.p);
  APC_HIP(c, hipStreamSynchronize(c->s));  // Lc->nnz copied from the host stack above
  return 0;
}

// Galerkin product by a global sort of every fine entry's (I, J) key (the
// first formulation; kept behind AmgOptions::galerkin_sorted as a check).
static int galerkin_sorted(apc_ctx *c, AmgHier *h, const Op<double> &A, AmgLevel *Lf, int l, int n,
                           int nc, AmgLevel *Lc) {
  const size_t ne = (l == 0) ? (size_t)n * (A.lm + 1) : (size_t)Lf->nnz;
  APC_ALLOC(c, h->key, ne + 1);
  APC_ALLOC(c, h->key2, ne + 1);
  APC_ALLOC(c, h->idx, ne + 1);
  APC_ALLOC(c, h->idx2, ne + 1);
  APC_ALLOC(c, h->ev, ne + 1);
  if (l == 0)
    hipLaunchKernelGGL(k_amg_entries_ell, dim3(nblocks(n, 256)), dim3(256), 0, c->s, A, Lf->agg.p,
                       nc, h->key.p, h->ev.p, h->idx.p);
  else
    hipLaunchKernelGGL(k_amg_entries_csr, dim3(nblocks(n, 256)), dim3(256), 0, c->s,
                       Csr<double>{Lf->rp.p, Lf->col.p, Lf->val.p, n}, Lf->agg.p, nc, h->key.p,
                       h->ev.p, h->idx.p);
  return galerkin_reduce(c, h, ne, nc, Lc);
}

static int ensure_work(apc_ctx *c, AmgLevel *L, int n, bool fp32) {
  if (fp32) {
    APC_ALLOC(c, L->x32, n + 1);
    APC_ALLOC(c, L->b32, n + 1);
    APC_ALLOC(c, L->r32, n + 1);
    APC_ALLOC(c, L->t32, n + 1);
  }
  APC_ALLOC(c, L->x, n + 1);
  APC_ALLOC(c, L->b, n + 1);
  APC_ALLOC(c, L->r, n + 1);
  APC_ALLOC(c, L->t, n + 1);
  return 0;
}

// Level 0's A P (k_ap_build) once its aggregates are known; fp32 adds the
// FP32 copy.  One read of the totals (the timers' byte counts).
static int build_ap(apc_ctx *c, AmgHier *h, const Op<double> &A, AmgLevel *L, bool fp32) {
  const int n = A.M, ns = (n + 63) / 64;
  const long long cap = (long long)ns * kApMax * 64;
  APC_ALLOC(c, L->apoff, ns + 1);
  APC_ALLOC(c, L->apwid, ns + 1);
  APC_ALLOC(c, h->apsn, ns + 1);
  APC_ALLOC(c, h->aptot, 2);
  APC_ALLOC(c, L->apn, (size_t)n + 1);
  APC_ALLOC(c, L->apss, (size_t)ns * kApMax + 1);
  APC_ALLOC(c, L->apmask, (size_t)ns * kApMax + 1);
  APC_ALLOC(c, L->apcol, cap + 1);
  APC_ALLOC(c, L->apval, cap + 1);
  if (fp32) APC_ALLOC(c, L->apval32, cap + 1);
  hipLaunchKernelGGL((k_ap_build<kApMax>), dim3(nblocks(n, 256)), dim3(256), 0, c->s, A, L->agg.p,
                     L->apoff.p, L->apwid.p, h->apsn.p, L->apcol.p, L->apval.p,
                     fp32 ? L->apval32.p : nullptr, L->apn.p, L->apss.p, L->apmask.p);
  hipLaunchKernelGGL(k_ap_total, dim3(1), dim3(1024), 0, c->s, ns, (const int *)L->apwid.p,
                     (const int *)h->apsn.p, h->aptot.p);
  long long tot[2] = {0, 0};
  APC_HIP(c, hipMemcpyAsync(tot, h->aptot.p, sizeof(tot), hipMemcpyDeviceToHost, c->s));
  APC_HIP(c, hipStreamSynchronize(c->s));
  L->ap_len = tot[0];
  L->ap_nnz = tot[1];
  L->ap = true;
  return 0;
}

// Sliced ELL (64-row slices, slot-major, each padded to its longest row)
// from a CSR matrix of n rows; padding takes column padc[r] (nullptr: r).
static int sell_from_csr(apc_ctx *c, AmgHier *h, int n, const int *rp, const int *col,
                         const double *val, DBuf<int> &soff, DBuf<int> &swid, DBuf<int> &scol,
                         DBuf<double> &sval, long long *len, const int *padc) {
  const int ns = (n + 63) / 64;
  APC_ALLOC(c, swid, ns + 1);
  APC_ALLOC(c, soff, ns + 1);
  hipLaunchKernelGGL(k_sell_width, dim3(nblocks(ns, 4)), dim3(256), 0, c->s, n, rp, swid.p);
  APC_ALLOC(c, h->cnt, 4);
  int r = scan_exclusive(c, swid.p, soff.p, ns, h->cnt.p);
  if (r) return r;
  int tot = 0;
  APC_HIP(c, hipMemcpyAsync(&tot, h->cnt.p, sizeof(int), hipMemcpyDeviceToHost, c->s));
  APC_HIP(c, hipStreamSynchronize(c->s));
  *len = (long long)tot * 64;
  APC_ALLOC(c, scol, *len + 1);
  APC_ALLOC(c, sval, *len + 1);
  hipLaunchKernelGGL(k_sell_fill, dim3(nblocks(n, 256)), dim3(256), 0, c->s, n, rp, col, val,
                     soff.p, swid.p, scol.p, sval.p, padc);
  return 0;
}

static int build_sell(apc_ctx *c, AmgHier *h, AmgLevel *L) {
  return sell_from_csr(c, h, L->n, L->rp.p, L->col.p, L->val.p, L->soff, L->swid, L->scol,
                       L->sval, &L->sell_len, nullptr);
}

__global__ void k_ap_iota(int n, int *__restrict__ a) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i <= n) a[i] = i;
}

// apass[agg[i]] = pass[i] (every row of an aggregate carries the same flag)
__global__ void k_ap_pass(int n, const int *__restrict__ agg, const unsigned char *__restrict__ pass,
                          unsigned char *__restrict__ apass) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) apass[agg[i]] = pass[i];
}

// A coarse level's A P (level L, CSR rows rp/col/val, aggregates agg into
// nc coarse rows): each row is its own group in the Galerkin row reduction
// (k_gal_rows), so its entries are summed per aggregate, distinct aggregates
// ascending; then sliced ELL like the level's own matrix.
static int build_ap_coarse(apc_ctx *c, AmgHier *h, AmgLevel *L, int nc, bool fp32) {
  const int n = L->n;
  const size_t nnz = (size_t)L->nnz;
  APC_ALLOC(c, h->gcol, nnz + 1);
  APC_ALLOC(c, h->gval, nnz + 1);
  APC_ALLOC(c, h->apiota, (size_t)n + 1);
  APC_ALLOC(c, h->apcnt, (size_t)n + 1);
  APC_ALLOC(c, h->aprp, (size_t)n + 1);
  APC_ALLOC(c, h->apdiag, (size_t)n + 1);
  APC_ALLOC(c, h->cnt, 4);
  const unsigned g = nblocks(n + 1, 256);
  hipLaunchKernelGGL(k_ap_iota, dim3(g), dim3(256), 0, c->s, n, h->apiota.p);
  const CsrView v{Csr<double>{L->rp.p, L->col.p, L->val.p, n}};
  hipLaunchKernelGGL((k_gal_emit<CsrView>), dim3(nblocks(n, 256)), dim3(256), 0, c->s, v, n,
                     L->agg.p, (const int *)L->rp.p, h->gcol.p, h->gval.p);
  const unsigned grid = (unsigned)std::min(n, 1 << 16);
  hipLaunchKernelGGL(k_gal_rows, dim3(grid), dim3(64), 0, c->s, n, (const int *)h->apiota.p,
                     (const int *)L->rp.p, h->gcol.p, h->gval.p, 0, h->apcnt.p,
                     (const int *)nullptr, (int *)nullptr, (double *)nullptr, h->apdiag.p);
  int r = scan_exclusive(c, h->apcnt.p, h->aprp.p, n, h->cnt.p);
  if (r) return r;
  int tot = 0;
  APC_HIP(c, hipMemcpyAsync(&tot, h->cnt.p, sizeof(int), hipMemcpyDeviceToHost, c->s));
  APC_HIP(c, hipStreamSynchronize(c->s));
  APC_HIP(c, hipMemcpyAsync(h->aprp.p + n, h->cnt.p, sizeof(int), hipMemcpyDeviceToDevice, c->s));
  APC_ALLOC(c, h->apc, (size_t)tot + 1);
  APC_ALLOC(c, h->apv, (size_t)tot + 1);
  hipLaunchKernelGGL(k_gal_rows, dim3(grid), dim3(64), 0, c->s, n, (const int *)h->apiota.p,
                     (const int *)L->rp.p, h->gcol.p, h->gval.p, 1, (int *)nullptr,
                     (const int *)h->aprp.p, h->apc.p, h->apv.p, h->apdiag.p);
  // padding: the row's own aggregate (a column of the coarse vector)
  r = sell_from_csr(c, h, n, h->aprp.p, h->apc.p, h->apv.p, L->apoff, L->apwid, L->apcol,
                    L->apval, &L->ap_len, L->agg.p);
  if (r) return r;
  if (fp32) {
    APC_ALLOC(c, L->apval32, L->ap_len + 1);
    hipLaunchKernelGGL((k_amg_convert<double, float>), dim3(1024), dim3(256), 0, c->s,
                       (size_t)L->ap_len, L->apval.p, L->apval32.p);
  }
  if (L->pass.p) {
    APC_ALLOC(c, L->apass, (size_t)nc + 1);
    hipLaunchKernelGGL(k_ap_pass, dim3(nblocks(n, 256)), dim3(256), 0, c->s, n, L->agg.p,
                       (const unsigned char *)L->pass.p, L->apass.p);
  }
  L->ap_nnz = tot;
  L->ap = true;
  return 0;
}

// Box of every coarse row: the box of any of its rows (aggregates never
// join rows of two boxes: the boxes' blocks share no entry).
__global__ void k_box_down(int n, const int *__restrict__ agg, const int *__restrict__ bf,
                           int *__restrict__ bc) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) bc[agg[i]] = bf[i];
}

// ---- packed boxes: per-box coarsening stops and the coarsest box blocks ----
// first row of every box on a box-major level; start[nbox] = n
__global__ void k_box_starts(int n, int nbox, const int *__restrict__ box, int *__restrict__ start) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  if (i == 0 || box[i - 1] != box[i]) start[box[i]] = i;
  if (i == n - 1) start[nbox] = n;
}

// rows of stopped boxes: flagged, and no smoothing (dinv 0)
__global__ void k_pass_rows(int n, const int *__restrict__ box, const int *__restrict__ frozen,
                            unsigned char *__restrict__ pass, double *__restrict__ dinv) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const int f = frozen[box[i]] ? 1 : 0;
  pass[i] = (unsigned char)f;
  if (f) dinv[i] = 0.0;
}

// shift of box b's block: the mean |a_ii| over its rows (periodic), else 0;
// one block per box
__global__ void __launch_bounds__(256) k_bgj_shift(Csr<double> A, const int *__restrict__ bstart,
                                                   int periodic, double *__restrict__ shift) {
  __shared__ double ws[4];
  const int b = blockIdx.x, r0 = bstart[b], r1 = bstart[b + 1];
  double a = 0.0;
  for (int i = r0 + threadIdx.x; i < r1; i += 256)
    for (int e = A.rp[i]; e < A.rp[i + 1]; ++e)
      if (A.col[e] == i) a += fabs(A.val[e]);
  a = block_sum256(a, ws);
  if (threadIdx.x == 0) shift[b] = periodic ? a / (r1 - r0) : 0.0;
}

// [A_b + shift_b / n_b 11^T | I] of row i's box b at goff[b] (n_b rows of
// 2 n_b); one block per row
__global__ void k_bgj_fill(Csr<double> A, const int *__restrict__ box,
                           const int *__restrict__ bstart, const long long *__restrict__ goff,
                           const double *__restrict__ shift, double *__restrict__ G) {
  const int i = blockIdx.x;
  const int b = box[i], r0 = bstart[b], n = bstart[b + 1] - r0, li = i - r0;
  double *g = G + goff[b] + (size_t)li * 2 * n;
  const double sh = shift[b] / n;
  for (int j = threadIdx.x; j < 2 * n; j += blockDim.x)
    g[j] = (j < n) ? sh : (j - n == li ? 1.0 : 0.0);
  __syncthreads();
  if (threadIdx.x == 0)
    for (int e = A.rp[i]; e < A.rp[i + 1]; ++e) g[A.col[e] - r0] += A.val[e];
}

// Gauss-Jordan with partial pivoting of every box block, one block per box
// (the steps of k_gj_pivot / k_gj_elim / k_gj_extract within one launch);
// the inverse to inv at ioff[b].  A zero pivot counts in bad.
__global__ void __launch_bounds__(256) k_bgj(const int *__restrict__ bstart,
                                             const long long *__restrict__ goff,
                                             const long long *__restrict__ ioff,
                                             double *__restrict__ G, double *__restrict__ inv,
                                             int *__restrict__ bad) {
  __shared__ double colk[1024];
  __shared__ double bv[256];
  __shared__ int bi[256];
  const int b = blockIdx.x, t = threadIdx.x;
  const int n = bstart[b + 1] - bstart[b], w = 2 * n;
  double *g = G + goff[b];
  for (int k = 0; k < n; ++k) {
    double v = -1.0;
    int vi = k;
    for (int i = k + t; i < n; i += 256) {
      const double a = fabs(g[(size_t)i * w + k]);
      if (a > v) { v = a; vi = i; }
    }
    bv[t] = v;
    bi[t] = vi;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
      if (t < s && (bv[t + s] > bv[t] || (bv[t + s] == bv[t] && bi[t + s] < bi[t]))) {
        bv[t] = bv[t + s];
        bi[t] = bi[t + s];
      }
      __syncthreads();
    }
    const int p = bi[0];
    if (!(bv[0] > 0.0)) {  // uniform: every thread read the same bv[0]
      if (t == 0) atomicAdd(bad, 1);
      return;
    }
    if (p != k)
      for (int j = t; j < w; j += 256) {
        const double a = g[(size_t)k * w + j];
        g[(size_t)k * w + j] = g[(size_t)p * w + j];
        g[(size_t)p * w + j] = a;
      }
    __syncthreads();
    const double ip = 1.0 / g[(size_t)k * w + k];
    __syncthreads();
    for (int j = t; j < w; j += 256) g[(size_t)k * w + j] *= ip;
    __syncthreads();
    for (int i = t; i < n; i += 256) colk[i] = g[(size_t)i * w + k];
    __syncthreads();
    for (int e = t; e < n * w; e += 256) {
      const int i = e / w, j = e - i * w;
      if (i != k) g[e] -= colk[i] * g[(size_t)k * w + j];
    }
    __syncthreads();
  }
  double *o = inv + ioff[b];
  for (int e = t; e < n * n; e += 256) {
    const int i = e / n, j = e - i * n;
    o[e] = g[(size_t)i * w + n + j];
  }
}

static int block_inverse(apc_ctx *c, AmgHier *h, AmgLevel *L, bool periodic,
                         const std::vector<int> &start) {
  const int n = L->n, nbox = h->nbox;
  std::vector<long long> go(nbox), io(nbox);
  long long gt = 0, it = 0;
  for (int b = 0; b < nbox; ++b) {
    const long long nb = start[b + 1] - start[b];
    if (nb < 1 || nb > 1024)
      return fail(c, APC_ERR_UNSUPPORTED, "amg: box %d keeps %lld rows on the coarsest level "
                  "(its coarsening stalled above the dense size)", b, nb);
    go[b] = gt;
    io[b] = it;
    gt += 2 * nb * nb;
    it += nb * nb;
  }
  Csr<double> A{L->rp.p, L->col.p, L->val.p, n};
  APC_ALLOC(c, h->bstart, nbox + 1);
  APC_ALLOC(c, h->goff, nbox);
  APC_ALLOC(c, h->ioff, nbox);
  APC_ALLOC(c, h->bshift, nbox);
  APC_ALLOC(c, h->gj, gt + 1);
  APC_ALLOC(c, h->cinv, it + 1);
  APC_ALLOC(c, h->cnt, 4);
  APC_HIP(c, hipMemcpyAsync(h->bstart.p, start.data(), sizeof(int) * (nbox + 1),
                            hipMemcpyHostToDevice, c->s));
  APC_HIP(c, hipMemcpyAsync(h->goff.p, go.data(), sizeof(long long) * nbox, hipMemcpyHostToDevice,
                            c->s));
  APC_HIP(c, hipMemcpyAsync(h->ioff.p, io.data(), sizeof(long long) * nbox, hipMemcpyHostToDevice,
                            c->s));
  APC_HIP(c, hipMemsetAsync(h->cnt.p, 0, sizeof(int), c->s));
  hipLaunchKernelGGL(k_bgj_shift, dim3(nbox), dim3(256), 0, c->s, A, h->bstart.p,
                     (periodic && h->o.shift) ? 1 : 0, h->bshift.p);
  hipLaunchKernelGGL(k_bgj_fill, dim3(n), dim3(256), 0, c->s, A, L->box.p, h->bstart.p, h->goff.p,
                     h->bshift.p, h->gj.p);
  hipLaunchKernelGGL(k_bgj, dim3(nbox), dim3(256), 0, c->s, h->bstart.p, h->goff.p, h->ioff.p,
                     h->gj.p, h->cinv.p, h->cnt.p);
  int bad = 0;
  APC_HIP(c, hipMemcpyAsync(&bad, h->cnt.p, sizeof(int), hipMemcpyDeviceToHost, c->s));
  APC_HIP(c, hipStreamSynchronize(c->s));  // the offsets' host vectors go out of scope
  if (bad) return fail(c, APC_ERR_UNSUPPORTED, "amg: %d singular coarsest box blocks", bad);
  if (h->fp32) {
    APC_ALLOC(c, h->cinv32, it + 1);
    hipLaunchKernelGGL((k_amg_convert<double, float>), dim3(256), dim3(256), 0, c->s, (size_t)it,
                       h->cinv.p, h->cinv32.p);
  }
  return 0;
}

static int dense_inverse(apc_ctx *c, AmgHier *h, AmgLevel *L, bool periodic) {
  const int n = L->n;
  periodic = periodic && h->o.shift;
  Csr<double> A{L->rp.p, L->col.p, L->val.p, n};
  double shift = 0.0;
  if (periodic) {  // mean |diagonal|, from the host copy of the level's dinv
    std::vector<double> dv(n);
    APC_HIP(c, hipMemcpyAsync(dv.data(), L->dinv.p, sizeof(double) * n, hipMemcpyDeviceToHost, c->s));
    APC_HIP(c, hipStreamSynchronize(c->s));
    double s = 0.0;
    int k = 0;
    for (double d : dv)
      if (d != 0.0) { s += 1.0 / fabs(d); ++k; }
    shift = k ? s / k : 1.0;
  }
  APC_ALLOC(c, h->gj, (size_t)2 * n * n + 1);
  APC_ALLOC(c, h->colk, n + 1);
  APC_ALLOC(c, h->cinv, (size_t)n * n + 1);
  APC_ALLOC(c, h->cnt, 4);
  APC_HIP(c, hipMemsetAsync(h->cnt.p, 0, sizeof(int), c->s));
  hipLaunchKernelGGL(k_gj_fill, dim3(n), dim3(256), 0, c->s, A, shift, h->gj.p);
  const unsigned ge = std::max(1u, std::min(1024u, nblocks((long)2 * n * n, 256)));
  for (int k = 0; k < n; ++k) {
    hipLaunchKernelGGL(k_gj_pivot, dim3(1), dim3(256), 0, c->s, n, k, h->gj.p, h->colk.p, h->cnt.p);
    hipLaunchKernelGGL(k_gj_elim, dim3(ge), dim3(256), 0, c->s, n, k, h->gj.p, h->colk.p);
  }
  hipLaunchKernelGGL(k_gj_extract, dim3(n), dim3(256), 0, c->s, n, h->gj.p, h->cinv.p);
  int bad = 0;
  APC_HIP(c, hipMemcpyAsync(&bad, h->cnt.p, sizeof(int), hipMemcpyDeviceToHost, c->s));
  APC_HIP(c, hipStreamSynchronize(c->s));
  if (bad) return fail(c, APC_ERR_UNSUPPORTED, "amg: singular coarsest matrix (%d rows)", n);
  if (h->fp32) {
    APC_ALLOC(c, h->cinv32, (size_t)n * n + 1);
    hipLaunchKernelGGL((k_amg_convert<double, float>), dim3(256), dim3(256), 0, c->s, (size_t)n * n,
                       h->cinv.p, h->cinv32.p);
  }
  return 0;
}

int amg_setup(apc_ctx *c, AmgHier **hp, const Op<double> &A, bool periodic, bool fp32,
              const AmgOptions &o) {
  if (!*hp) *hp = new AmgHier();
  AmgHier *h = *hp;
  h->s = c->s;
  for (auto &gr : h->graphs) gr.release(c->s);
  h->graphs.clear();
  ++h->gen;
  h->o = o;
  h->fp32 = fp32;
  h->nlev = 0;
  h->dense = false;
  auto level = [&](int l) -> AmgLevel * {
    while ((int)h->L.size() <= l) h->L.push_back(new AmgLevel());
    return h->L[l];
  };
  KScope ks(c, "amg_setup");
  // Packed boxes: every box stops coarsening where a context of that box
  // alone would (at most coarse_max rows, but level 0 coarsened once; or when
  // its coarsening stalls); its rows then pass through the deeper levels as
  // singleton aggregates, and the coarsest level solves every box's block
  // with its own dense inverse.  Levels stay box-major: coarse rows are
  // numbered by their roots' rows.
  const bool packed = c->nbox > 1;
  const int nbox = c->nbox;
  h->nbox = nbox;
  h->blocks = false;
  std::vector<int> frozen(nbox, 0), cnt(nbox, 0), start;
  if (packed) {
    if ((int)c->rstart_h.size() != nbox + 1 || c->rstart_h[nbox] != A.M)
      return fail(c, APC_ERR_ARG, "amg: the boxes' rows do not cover the system");
    for (int b = 0; b < nbox; ++b) cnt[b] = c->rstart_h[b + 1] - c->rstart_h[b];
    start = c->rstart_h;
    APC_ALLOC(c, h->frozen, nbox);
    APC_HIP(c, hipMemsetAsync(h->frozen.p, 0, sizeof(int) * nbox, c->s));
  }
  auto all_frozen = [&]() {
    for (int f : frozen)
      if (!f) return false;
    return true;
  };
  // the rows are the computational nodes, whose cells the geometric
  // aggregation reads
  const bool geo_rows = c->have_comp && c->nc + c->nv == A.M && c->cx.p && c->ch.p &&
                        c->prm.depth - o.geo0 <= kGeoRes;
  h->geo_prev = false;
  {  // correction scales (AmgOptions::alpha, alpha0)
    const bool g0 = o.geo0 > 0 && geo_rows;
    if (!(h->o.alpha > 0.0)) h->o.alpha = g0 ? 1.5 : 1.6;
    if (!(h->o.alpha0 > 0.0)) h->o.alpha0 = h->o.alpha;
  }
  AmgLevel *L0 = level(0);
  L0->n = A.M;
  L0->nnz = A.M * (A.lm + 1);
  L0->ap = false;
  int r;
  {
    KScope ks1(c, "amg_setup_work0");
    r = ensure_work(c, L0, A.M, fp32);
  }
  if (r) return r;
  int n = A.M;
  int k = 0;  // deepest level built by this setup
  for (int l = 0;; ++l) {
    AmgLevel *Lf = level(l);
    Lf->ap = false;
    // stop at the dense size (but coarsen level 0 at least once) or the cap
    if (packed ? (l > 0 && all_frozen()) : (l > 0 && n <= o.coarse_max)) break;
    if (l + 1 >= o.max_levels) break;
    const int *boxf = packed ? (l == 0 ? (const int *)c->rowbox.p : (const int *)Lf->box.p)
                             : nullptr;
    int nc = 0;
    // geometric aggregation where asked for and the rows are the context's
    // computational nodes (single box); a coarse level only below a
    // geometric one (its rows are then ancestor cells)
    bool geo = l == 0 ? (o.geo0 > 0 && geo_rows) : (!packed && o.geoc > 0 && h->geo_prev);
    {
      KScope ks1(c, "amg_setup_aggregate");
      if (geo) {
        r = aggregate_geo(c, h, l, n, l == 0 ? o.geo0 : o.geoc, Lf, &nc, boxf);
        if (!r && nc < 0) geo = false;  // too deep: independent sets instead
      }
      if (!r && !geo)
        r = (l == 0) ? aggregate_level(c, h, EllView{A}, n, Lf, &nc, boxf)
                     : aggregate_level(c, h, CsrView{Csr<double>{Lf->rp.p, Lf->col.p, Lf->val.p, n}},
                                       n, Lf, &nc, boxf);
    }
    if (r) return r;
    h->geo_prev = geo;
    if (nc < 1 || nc >= n) break;  // no coarsening: this level is the coarsest
    if (packed) {
      AmgLevel *Ln = level(l + 1);
      APC_ALLOC(c, Ln->box, nc + 1);
      APC_ALLOC(c, h->bstart, nbox + 1);
      hipLaunchKernelGGL(k_box_down, dim3(nblocks(n, 256)), dim3(256), 0, c->s, n, Lf->agg.p, boxf,
                         Ln->box.p);
      hipLaunchKernelGGL(k_box_starts, dim3(nblocks(nc, 256)), dim3(256), 0, c->s, nc, nbox,
                         (const int *)Ln->box.p, h->bstart.p);
      start.assign(nbox + 1, 0);
      APC_HIP(c, hipMemcpyAsync(start.data(), h->bstart.p, sizeof(int) * (nbox + 1),
                                hipMemcpyDeviceToHost, c->s));
      APC_HIP(c, hipStreamSynchronize(c->s));
      for (int b = 0; b < nbox; ++b) {
        const int m = start[b + 1] - start[b];
        if (m < 1) return fail(c, APC_ERR_ARG, "amg: box %d has no rows on level %d", b, l + 1);
        if (!frozen[b] && (m <= o.coarse_max || m >= cnt[b])) frozen[b] = 1;
        cnt[b] = m;
      }
      APC_HIP(c, hipMemcpyAsync(h->frozen.p, frozen.data(), sizeof(int) * nbox,
                                hipMemcpyHostToDevice, c->s));
      APC_HIP(c, hipStreamSynchronize(c->s));  // the flags' host vector changes on the next level
    }
    if (!geo) {  // (the geometric aggregation lists the members itself)
      KScope ks1(c, "amg_setup_group");
      r = group_members(c, h, Lf, n, nc);
    }
    if (r) return r;
    AmgLevel *Lc = level(l + 1);
    {
      if (o.ap && (l > 0 || A.lm + 1 <= kApMax)) {
        KScope ks1(c, "amg_setup_ap");
        r = (l == 0) ? build_ap(c, h, A, Lf, fp32) : build_ap_coarse(c, h, Lf, nc, fp32);
        if (r) return r;
      }
      // with A P built, the coarse operator is P^T (A P): its rows sum the
      // members' A P rows, a third of the entries of A's
      KScope ks1(c, "amg_setup_galerkin");
      if (o.galerkin_sorted) r = galerkin_sorted(c, h, A, Lf, l, n, nc, Lc);
      else if (Lf->ap && l == 0)
        r = galerkin_rows_ap(c, h, Lf, n, nc, Lc);
      else if (Lf->ap)
        r = galerkin_rows(c, h, CsrView{Csr<double>{h->aprp.p, h->apc.p, h->apv.p, n}}, Lf, n, nc,
                          Lc, true);
      else if (l == 0) r = galerkin_rows_ell(c, h, A, Lf, n, nc, Lc);
      else r = galerkin_rows(c, h, CsrView{Csr<double>{Lf->rp.p, Lf->col.p, Lf->val.p, n}}, Lf, n, nc, Lc);
    }
    if (r) return r;
    {
      KScope ks1(c, "amg_setup_work");
      r = ensure_work(c, Lc, nc, fp32);
    }
    if (r) return r;
    {
      KScope ks1(c, "amg_setup_sell");
      r = build_sell(c, h, Lc);
    }
    if (r) return r;
    if (packed) {
      APC_ALLOC(c, Lc->pass, nc + 1);
      hipLaunchKernelGGL(k_pass_rows, dim3(nblocks(nc, 256)), dim3(256), 0, c->s, nc,
                         (const int *)Lc->box.p, (const int *)h->frozen.p, Lc->pass.p, Lc->dinv.p);
    }
    if (fp32) {
      KScope ks1(c, "amg_setup_convert");
      APC_ALLOC(c, Lc->sval32, Lc->sell_len + 1);
      APC_ALLOC(c, Lc->dinv32, nc + 1);
      hipLaunchKernelGGL((k_amg_convert<double, float>), dim3(1024), dim3(256), 0, c->s,
                         (size_t)Lc->sell_len, Lc->sval.p, Lc->sval32.p);
      hipLaunchKernelGGL((k_amg_convert<double, float>), dim3(256), dim3(256), 0, c->s, (size_t)nc,
                         Lc->dinv.p, Lc->dinv32.p);
    }
    n = nc;
    k = l + 1;
  }
  h->nlev = k + 1;
  AmgLevel *Lc = h->L[k];
  h->nc = Lc->n;
  if (packed) {
    if (k == 0) return fail(c, APC_ERR_UNSUPPORTED, "amg: the packed boxes did not coarsen");
    KScope ks1(c, "amg_setup_dense");
    r = block_inverse(c, h, Lc, periodic, start);
    if (r) return r;
    h->blocks = true;
    h->dense = true;
  } else if (k > 0 && Lc->n <= 1024) {  // else (coarsening stalled) smoothing only
    KScope ks1(c, "amg_setup_dense");
    r = dense_inverse(c, h, Lc, periodic);
    if (r) return r;
    h->dense = true;
  }
  APC_HIP(c, hipGetLastError());
  return 0;
}

template <class T>
struct LevelPtrs {
  T *x, *b, *r, *t;
  const T *dinv, *val;
};

template <class T>
static LevelPtrs<T> ptrs(AmgLevel *L);
template <>
LevelPtrs<double> ptrs<double>(AmgLevel *L) {
  return {L->x.p, L->b.p, L->r.p, L->t.p, L->dinv.p, L->sval.p};
}
template <>
LevelPtrs<float> ptrs<float>(AmgLevel *L) {
  return {L->x32.p, L->b32.p, L->r32.p, L->t32.p, L->dinv32.p, L->sval32.p};
}

template <class T>
static const T *ap_val(AmgLevel *L);
template <>
const double *ap_val<double>(AmgLevel *L) { return L->apval.p; }
template <>
const float *ap_val<float>(AmgLevel *L) { return L->apval32.p; }

template <class T>
static const T *cinv_of(AmgHier *h);
template <>
const double *cinv_of<double>(AmgHier *h) { return h->cinv.p; }
template <>
const float *cinv_of<float>(AmgHier *h) { return h->cinv32.p; }

// x = cycle(b) on level l; x and b are device vectors of the level's size,
// tmp a scratch vector of that size (distinct from x and b).
// timer names of the coarse levels, one per level so profiles split by level
static const char *const kSweepName[] = {"amg_sweep_l0", "amg_sweep_l1", "amg_sweep_l2",
                                         "amg_sweep_l3", "amg_sweep_l4", "amg_sweep_l5",
                                         "amg_sweep_l6", "amg_sweep_l7", "amg_sweep_l8",
                                         "amg_sweep_l9", "amg_sweep_l10", "amg_sweep_l11"};
static const char *const kResidName[] = {"amg_resid_l0", "amg_resid_l1", "amg_resid_l2",
                                         "amg_resid_l3", "amg_resid_l4", "amg_resid_l5",
                                         "amg_resid_l6", "amg_resid_l7", "amg_resid_l8",
                                         "amg_resid_l9", "amg_resid_l10", "amg_resid_l11"};
static const char *const kRestrictName[] = {
    "amg_restrict_l0", "amg_restrict_l1", "amg_restrict_l2", "amg_restrict_l3",
    "amg_restrict_l4", "amg_restrict_l5", "amg_restrict_l6", "amg_restrict_l7",
    "amg_restrict_l8", "amg_restrict_l9", "amg_restrict_l10", "amg_restrict_l11"};
static const char *const kProlongName[] = {
    "amg_prolong_l0", "amg_prolong_l1", "amg_prolong_l2", "amg_prolong_l3",
    "amg_prolong_l4", "amg_prolong_l5", "amg_prolong_l6", "amg_prolong_l7",
    "amg_prolong_l8", "amg_prolong_l9", "amg_prolong_l10", "amg_prolong_l11"};
static const char *const kApResidName[] = {
    "amg_ap_resid_l0", "amg_ap_resid_l1", "amg_ap_resid_l2", "amg_ap_resid_l3",
    "amg_ap_resid_l4", "amg_ap_resid_l5", "amg_ap_resid_l6", "amg_ap_resid_l7",
    "amg_ap_resid_l8", "amg_ap_resid_l9", "amg_ap_resid_l10", "amg_ap_resid_l11"};
static const char *const kApSweepName[] = {
    "amg_ap_sweep_l0", "amg_ap_sweep_l1", "amg_ap_sweep_l2", "amg_ap_sweep_l3",
    "amg_ap_sweep_l4", "amg_ap_sweep_l5", "amg_ap_sweep_l6", "amg_ap_sweep_l7",
    "amg_ap_sweep_l8", "amg_ap_sweep_l9", "amg_ap_sweep_l10", "amg_ap_sweep_l11"};
static_assert(sizeof(kSweepName) / sizeof(kSweepName[0]) >= 12, "one name per level");
static_assert(sizeof(kApSweepName) / sizeof(kApSweepName[0]) >= 12, "one name per level");

template <class T>
static int cycle(apc_ctx *c, AmgHier *h, const Op<T> &A0, int l, const T *b, T *x, T *tmp,
                 const GState *st) {
  AmgLevel *L = h->L[l];
  const AmgOptions &o = h->o;
  const int n = (l == 0) ? A0.M : L->n;
  const unsigned g = nblocks(n, 256);
  LevelPtrs<T> P = ptrs<T>(L);
  const T *dinv = (l == 0) ? A0.dinv : P.dinv;
  Op<T> E = A0;
  Sell<T> C{L->soff.p, L->swid.p, L->scol.p, P.val, n};
  const bool ell = (l == 0);
  // level-0 algorithmic bytes per row, the SpMV convention (s + 4) nnz + M
  // (vectors): each of the lm slots reads its index and weight; the vector x
  // counts once per row like every other vector (d, f, b, x, r for the
  // residual; d, f, dinv, b, x in, x out for a sweep)
  const double sT = sizeof(T);
  const double lm0 = A0.lm;
  const bool f32 = sizeof(T) == 4;
  // lanes per row on this level's coarse kernels (and per coarse row on the
  // restriction into the next): more on the small, latency-bound levels
  const int gl = ell ? 1 : (n <= o.g8_rows ? 8 : (n <= o.g4_rows ? 4 : 1));
  auto sweep = [&](const T *xin, T *xout) {
    if (!ell && gl > 1) {
      KScope ks(c, kSweepName[l]);
      if (gl == 8)
        hipLaunchKernelGGL((k_cy_sweep_g<T, 8>), dim3(nblocks((long)n * 8, 256)), dim3(256), 0,
                           c->s, C, dinv, b, xin, xout, o.omega, st);
      else
        hipLaunchKernelGGL((k_cy_sweep_g<T, 4>), dim3(nblocks((long)n * 4, 256)), dim3(256), 0,
                           c->s, C, dinv, b, xin, xout, o.omega, st);
      return;
    }
    if (ell) {
      KScope ks(c, f32 ? "amg_sweep0_32" : "amg_sweep0", (double)n * (lm0 * (4 + sT) + 6 * sT));
      hipLaunchKernelGGL((k_cy_sweep<T, true>), dim3(g), dim3(256), 0, c->s, E, C, dinv, b, xin,
                         xout, o.omega, st);
    } else {
      KScope ks(c, kSweepName[l]);
      hipLaunchKernelGGL((k_cy_sweep<T, false>), dim3(g), dim3(256), 0, c->s, E, C, dinv, b, xin,
                         xout, o.omega, st);
    }
  };
  if (l == h->nlev - 1) {
    if (h->blocks) {
      KScope ks(c, "amg_dense");
      hipLaunchKernelGGL((k_cy_dense_blocks<T>), dim3(n), dim3(64), 0, c->s, L->box.p, h->bstart.p,
                         h->ioff.p, cinv_of<T>(h), b, x, st);
    } else if (h->dense) {
      KScope ks(c, "amg_dense");
      hipLaunchKernelGGL((k_cy_dense<T>), dim3(n), dim3(64), 0, c->s, n, cinv_of<T>(h), b, x, st);
    } else {  // no dense solve: smoothing only (odd sweep count ends in x)
      hipLaunchKernelGGL((k_cy_first<T>), dim3(g), dim3(256), 0, c->s, n, dinv, b, x, o.omega, st);
      for (int s = 0; s < 4 * o.nu; ++s) {
        sweep(x, tmp);
        sweep(tmp, x);
      }
    }
    return 0;
  }
  // smoothing writes alternate between tmp and x; with 2*nu - 1 sweeps after
  // the first (odd), starting in tmp ends in x
  T *cur = tmp, *alt = x;
  // coarse levels with one pre-smoothing step: the first step is formed
  // inside the first residual (k_cy_first_resid_g), one launch fewer
  const bool fuse_first = !ell && gl > 1 && o.nu == 1;
  // level 0 without pre-smoothing (A P form, W-cycle): x starts at 0, the
  // first restriction takes the input b, the first correction sets x
  const bool nopre = L->ap && l < o.nopre && (!ell || (l < o.wlevels && o.gamma == 2));
  if (nopre) {
  } else if (!fuse_first) {
    KScope ks(c, ell ? "amg_first0" : "amg_first_coarse");
    hipLaunchKernelGGL((k_cy_first<T>), dim3(g), dim3(256), 0, c->s, n, dinv, b, cur, o.omega, st);
  }
  for (int s = 1; s < o.nu && !nopre; ++s) {
    sweep(cur, alt);
    std::swap(cur, alt);
  }
  AmgLevel *Lc = h->L[l + 1];
  LevelPtrs<T> Pc = ptrs<T>(Lc);
  T *r = P.r;
  const int gam = (l < o.wlevels) ? o.gamma : 1;
  // with A P: after the first residual, every later one (and the first
  // post-smoothing step's) comes from r -= alpha (A P) e
  const bool apf = L->ap;
  const Sell<T> AP{L->apoff.p, L->apwid.p, L->apcol.p, apf ? ap_val<T>(L) : nullptr, n};
  const CSell<T> AP0{L->apoff.p, L->apwid.p, L->apn.p, L->apcol.p,
                     apf ? ap_val<T>(L) : nullptr, n};
  const double ap_bytes = apf ? (double)n * (4 + 4 * sT + (ell ? 1 : 0)) +
                                    (double)L->ap_nnz * (4 + sT) +
                                    (double)Lc->n * sT
                              : 0.0;
  // the correction scale of this level (level 0 may differ)
  const double alpha = (l == 0 && o.alpha0 > 0.0) ? o.alpha0 : o.alpha;
  int post_done = 0;
  for (int q = 0; q < gam; ++q) {
    if ((apf && q > 0) || (nopre && q == 0)) {
      // r = b - A cur already (k_cy_ap_resid after the previous correction),
      // or x = 0 and the residual is b (restricted directly)
    } else if (ell) {
      KScope ks(c, f32 ? "amg_resid0_32" : "amg_resid0", (double)n * (lm0 * (4 + sT) + 5 * sT));
      hipLaunchKernelGGL((k_cy_resid<T, true>), dim3(g), dim3(256), 0, c->s, E, C, b, cur, r, st);
    } else if (fuse_first && q == 0) {
      KScope ks(c, kResidName[l]);
      if (gl == 8)
        hipLaunchKernelGGL((k_cy_first_resid_g<T, 8>), dim3(nblocks((long)n * 8, 256)), dim3(256),
                           0, c->s, C, dinv, b, cur, r, o.omega, st);
      else
        hipLaunchKernelGGL((k_cy_first_resid_g<T, 4>), dim3(nblocks((long)n * 4, 256)), dim3(256),
                           0, c->s, C, dinv, b, cur, r, o.omega, st);
    } else if (gl == 8) {
      KScope ks(c, kResidName[l]);
      hipLaunchKernelGGL((k_cy_resid_g<T, 8>), dim3(nblocks((long)n * 8, 256)), dim3(256), 0, c->s,
                         C, b, cur, r, st);
    } else if (gl == 4) {
      KScope ks(c, kResidName[l]);
      hipLaunchKernelGGL((k_cy_resid_g<T, 4>), dim3(nblocks((long)n * 4, 256)), dim3(256), 0, c->s,
                         C, b, cur, r, st);
    } else {
      KScope ks(c, kResidName[l]);
      hipLaunchKernelGGL((k_cy_resid<T, false>), dim3(g), dim3(256), 0, c->s, E, C, b, cur, r, st);
    }
    {
      KScope ks(c, ell ? (f32 ? "amg_restrict0_32" : "amg_restrict0") : kRestrictName[l],
                ell ? (double)n * (sT + 4) + (double)Lc->n * (4 + sT) : 0.0);
      if (Lc->n <= o.rg8_rows)
        hipLaunchKernelGGL((k_cy_restrict_g<T, 8>), dim3(nblocks((long)Lc->n * 8, 256)), dim3(256),
                           0, c->s, Lc->n, L->mp.p, L->ml.p, (nopre && q == 0) ? b : r, Pc.b, st);
      else
        hipLaunchKernelGGL((k_cy_restrict<T>), dim3(nblocks(Lc->n, 256)), dim3(256), 0, c->s,
                           Lc->n, L->mp.p, L->ml.p, (nopre && q == 0) ? b : r, Pc.b, st);
    }
    int rr = cycle<T>(c, h, A0, l + 1, Pc.b, Pc.x, Pc.t, st);
    if (rr) return rr;
    // packed boxes: a stopped box's rows take the first coarse correction
    // unscaled and none after it (its coarsest solve, passed through once)
    const unsigned char *pass = (h->nbox > 1 && l > 0) ? L->pass.p : nullptr;
    const unsigned char *apass = pass ? L->apass.p : nullptr;
    const double apass_alpha = q == 0 ? 1.0 : 0.0;
    const unsigned gg = nblocks((long)n * gl, 256);
    const bool first = nopre && q == 0;  // x was 0, the residual is b
    if (ell && first) {
      KScope ks(c, f32 ? "amg_ap_resid0_32" : "amg_ap_resid0", ap_bytes);
      hipLaunchKernelGGL((k_cy_ap_resid_first<T>), dim3(g), dim3(256), 0, c->s, n,
                         (const int *)L->agg.p, AP0, (const T *)Pc.x, b, cur, r, alpha, st);
    } else if (apf && q + 1 < gam) {
      if (ell) {
        KScope ks(c, f32 ? "amg_ap_resid0_32" : "amg_ap_resid0", ap_bytes);
        hipLaunchKernelGGL((k_cy_ap_resid<T>), dim3(g), dim3(256), 0, c->s, n,
                           (const int *)L->agg.p, AP0, (const T *)Pc.x, cur, r, alpha, st);
      } else {
        KScope ks(c, kApResidName[l]);
        auto *kern = first ? (gl == 8 ? k_cy_ap_resid_g<T, 8, true>
                                      : gl == 4 ? k_cy_ap_resid_g<T, 4, true>
                                                : k_cy_ap_resid_g<T, 1, true>)
                           : (gl == 8 ? k_cy_ap_resid_g<T, 8, false>
                                      : gl == 4 ? k_cy_ap_resid_g<T, 4, false>
                                                : k_cy_ap_resid_g<T, 1, false>);
        hipLaunchKernelGGL(kern, dim3(gg), dim3(256), 0, c->s, AP, (const int *)L->agg.p,
                           (const T *)Pc.x, b, cur, r, alpha, pass, apass, apass_alpha, st);
      }
    } else if (apf && o.nu >= 1) {
      if (ell) {
        KScope ks(c, f32 ? "amg_ap_sweep0_32" : "amg_ap_sweep0", ap_bytes);
        hipLaunchKernelGGL((k_cy_ap_sweep<T>), dim3(g), dim3(256), 0, c->s, n,
                           (const int *)L->agg.p, AP0, (const T *)Pc.x, dinv, (const T *)cur,
                           (const T *)r, alt, alpha, o.omega, st);
      } else {
        KScope ks(c, kApSweepName[l]);
        auto *kern = first ? (gl == 8 ? k_cy_ap_sweep_g<T, 8, true>
                                      : gl == 4 ? k_cy_ap_sweep_g<T, 4, true>
                                                : k_cy_ap_sweep_g<T, 1, true>)
                           : (gl == 8 ? k_cy_ap_sweep_g<T, 8, false>
                                      : gl == 4 ? k_cy_ap_sweep_g<T, 4, false>
                                                : k_cy_ap_sweep_g<T, 1, false>);
        // first: no pre-smoothing, so the residual entering is b itself
        hipLaunchKernelGGL(kern, dim3(gg), dim3(256), 0, c->s, AP, (const int *)L->agg.p,
                           (const T *)Pc.x, dinv, (const T *)cur, first ? b : (const T *)r, alt,
                           alpha, o.omega, pass, apass, apass_alpha, st);
      }
      std::swap(cur, alt);
      post_done = 1;
    } else {
      KScope ks(c, ell ? (f32 ? "amg_prolong0_32" : "amg_prolong0") : kProlongName[l],
                ell ? (double)n * (4 + 3 * sT) : 0.0);
      hipLaunchKernelGGL((k_cy_prolong<T>), dim3(g), dim3(256), 0, c->s, n, L->agg.p, Pc.x, cur,
                         alpha, pass, q == 0 ? 1.0 : 0.0, st);
    }
  }
  for (int s = post_done; s < o.nu; ++s) {
    sweep(cur, alt);
    std::swap(cur, alt);
  }
  // a kernel, not a device-to-device memcpy: the captured cycle then holds
  // only this library's kernel nodes
  if (cur != x)
    hipLaunchKernelGGL((k_cy_copy<T>), dim3(g), dim3(256), 0, c->s, n, (const T *)cur, x, st);
  return 0;
}

// The cycle is a fixed sequence of a few hundred small launches (a W-cycle
// visits the deep levels 2^l times), so it is captured into a HIP graph on
// first use for a given (input, output, status word) and replayed; every
// kernel in it still checks the status word.  With kernel timing on, the
// launches are issued directly so that each one is timed.
template <class T>
int amg_apply(apc_ctx *c, AmgHier *h, const Op<T> &A, const T *in, T *out, const GState *st) {
  if (!h || h->nlev < 1) return fail(c, APC_ERR_ARG, "amg: no hierarchy");
  if (sizeof(T) == 4 && !h->fp32) return fail(c, APC_ERR_ARG, "amg: hierarchy has no FP32 copy");
  KScope ks(c, sizeof(T) == 4 ? "amg_cycle32" : "amg_cycle");
  T *tmp = ptrs<T>(h->L[0]).t;
  static const bool nograph = getenv("APC_AMG_NOGRAPH") != nullptr;
  if (c->timing || nograph) {
    int r = cycle<T>(c, h, A, 0, in, out, tmp, st);
    if (r) return r;
    APC_HIP(c, hipGetLastError());
    return 0;
  }
  const int f32 = sizeof(T) == 4;
  for (auto &g : h->graphs)
    if (g.exec && g.gen == h->gen && g.in == in && g.out == out && g.st == st && g.f32 == f32 &&
        g.a0 == (const void *)A.s) {
      APC_HIP(c, hipGraphLaunch(g.exec, c->s));
      return 0;
    }
  CycleGraph g;
  g.in = in;
  g.out = out;
  g.st = st;
  g.f32 = f32;
  g.gen = h->gen;
  g.a0 = A.s;
  APC_HIP(c, hipStreamBeginCapture(c->s, hipStreamCaptureModeThreadLocal));
  int r = cycle<T>(c, h, A, 0, in, out, tmp, st);
  hipGraph_t graph = nullptr;
  hipError_t e = hipStreamEndCapture(c->s, &graph);
  if (r) {
    if (graph) hipGraphDestroy(graph);
    return r;
  }
  APC_HIP(c, e);
  g.graph = graph;
  e = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    g.release(c->s);
    return hip_fail(c, e, "hipGraphInstantiate (multigrid cycle)");
  }
  if (h->graphs.size() >= 4) {
    h->graphs.front().release(c->s);
    h->graphs.erase(h->graphs.begin());
  }
  h->graphs.push_back(g);
  APC_HIP(c, hipGraphLaunch(g.exec, c->s));
  return 0;
}

template int amg_apply<double>(apc_ctx *, AmgHier *, const Op<double> &, const double *, double *,
                               const GState *);
template int amg_apply<float>(apc_ctx *, AmgHier *, const Op<float> &, const float *, float *,
                              const GState *);

void amg_free(AmgHier *h) { delete h; }

int amg_level_copy(apc_ctx *c, const AmgHier *h, int l, int *rp, int *col, double *val, int *agg) {
  if (!h || l < 0 || l >= h->nlev) return fail(c, APC_ERR_ARG, "amg level %d: no such level", l);
  const AmgLevel *L = h->L[l];
  if (l > 0 && (rp || col || val)) {
    if (rp) APC_HIP(c, hipMemcpyAsync(rp, L->rp.p, sizeof(int) * (L->n + 1), hipMemcpyDeviceToHost, c->s));
    if (col) APC_HIP(c, hipMemcpyAsync(col, L->col.p, sizeof(int) * L->nnz, hipMemcpyDeviceToHost, c->s));
    if (val) APC_HIP(c, hipMemcpyAsync(val, L->val.p, sizeof(double) * L->nnz, hipMemcpyDeviceToHost, c->s));
  }
  if (agg && l + 1 < h->nlev)
    APC_HIP(c, hipMemcpyAsync(agg, L->agg.p, sizeof(int) * L->n, hipMemcpyDeviceToHost, c->s));
  APC_HIP(c, hipStreamSynchronize(c->s));
  return 0;
}

int amg_levels(const AmgHier *h, int *rows, long long *nnz, int cap) {
  if (!h) return 0;
  for (int l = 0; l < h->nlev && l < cap; ++l) {
    if (rows) rows[l] = h->L[l]->n;
    if (nnz) nnz[l] = h->L[l]->nnz;
  }
  return h->nlev;
}

}  // namespace apc
