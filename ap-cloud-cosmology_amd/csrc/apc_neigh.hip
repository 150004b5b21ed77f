// GFD neighbour selection, solver.cpp:456-570 (searchneighboroctant) and its
// quicksort, solver.cpp:600-632.  One thread per computational/vacuum node;
// the candidate lists live in a chunked device scratch laid out element-major
// (element e of the chunk's thread t at e*rows + t) so a wave's accesses to
// the same element coalesce.  Built with -ffp-contract=off: the octant test
// and the squared distances must match the reference bit for bit.
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
  int rows, t;
  __device__ uint32_t &MO(int e) const { return mo[(size_t)e * rows + t]; }
  __device__ double &D2(int e) const { return d2[(size_t)e * rows + t]; }
  __device__ double &WD(int e) const { return wd[(size_t)e * rows + t]; }
  __device__ int &WM(int e) const { return wm[(size_t)e * rows + t]; }
};

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

__global__ void __launch_bounds__(256) k_search(int M, int S, int row0, int rows, int lm, int cap,
                                                Frame F, const double *__restrict__ cx,
                                                const double *__restrict__ cy,
                                                const double *__restrict__ cz,
                                                const int *__restrict__ n24,
                                                const int *__restrict__ n24cnt, uint32_t *mo,
                                                double *d2, double *wd, int *wm,
                                                int *__restrict__ nbr, double *__restrict__ dist,
                                                int *__restrict__ status) {
  int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= rows) return;
  const int i0 = row0 + t;
  Cand c{mo, d2, wd, wm, rows, t};
  const int nq = 10 * lm;
  const double x0 = cx[i0], y0 = cy[i0], z0 = cz[i0];
  auto octant = [&](int m) {
    double ex = pwrap(cx[m] - x0, F.len[0], F.bc);
    double ey = pwrap(cy[m] - y0, F.len[1], F.bc);
    double ez = pwrap(cz[m] - z0, F.len[2], F.bc);
    double xt = F.A[0] * ex + F.A[1] * ey + F.A[2] * ez;
    double yt = F.A[3] * ex + F.A[4] * ey + F.A[5] * ez;
    double zt = F.A[6] * ex + F.A[7] * ey + F.A[8] * ez;
    return (xt > 0) + 2 * (yt > 0) + 4 * (zt > 0);
  };
  auto sqd = [&](int m) {  // solver.cpp:595-598
    double dx = pwrap(cx[m] - x0, F.len[0], F.bc);
    double dy = pwrap(cy[m] - y0, F.len[1], F.bc);
    double dz = pwrap(cz[m] - z0, F.len[2], F.bc);
    return dx * dx + dy * dy + dz * dz;
  };
  int nn[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  int l = 0;
  // bucket the node's own Neigh24 (no de-duplication, solver.cpp:481-492)
  {
    const int cnt = n24cnt[i0];
    for (int q = 0; q < cnt; ++q) {
      int m = n24[(size_t)i0 * 24 + q];
      int o = octant(m);
      c.MO(l) = (uint32_t)m | ((uint32_t)o << 29);
      c.D2(l) = sqd(m);
      nn[o]++;
      l++;
    }
  }
  // expand through the neighbours' Neigh24 until >= lm (solver.cpp:495-551)
  int ite = 0;
  while (l < lm) {
    // snapshot of the current list in octant order
    int k = 0;
    for (int o = 0; o < 8; ++o)
      for (int e = 0; e < l; ++e)
        if ((int)(c.MO(e) >> 29) == o) c.WM(k++) = (int)(c.MO(e) & 0x1fffffff);
    const int lsnap = l;
    for (int a = 0; a < lsnap; ++a) {
      int src = c.WM(a);
      int cnt = (src < M) ? n24cnt[src] : 0;
      for (int q = 0; q < cnt; ++q) {
        int m = n24[(size_t)src * 24 + q];
        if (m >= S) { status[0] = APC_ERR_NEIGH_INVALID; return; }
        int o = octant(m);
        bool seen = false;
        for (int e = 0; e < l; ++e)
          if ((int)(c.MO(e) & 0x1fffffff) == m) { seen = true; break; }
        if (!seen && m != i0) {
          if (nn[o] >= nq) { status[0] = APC_ERR_NEIGH_MANY; return; }
          if (l >= cap) { status[0] = APC_ERR_UNSUPPORTED; status[1] = i0; return; }
          c.MO(l) = (uint32_t)m | ((uint32_t)o << 29);
          c.D2(l) = sqd(m);
          nn[o]++;
          l++;
        }
      }
    }
    ite++;
    if (ite > 5) { status[0] = APC_ERR_NEIGH_FEW; return; }
  }
  // The reference sorts td[k..k+nn[o]-1] but reads octant o's own slots
  // (solver.cpp:557-561), so only octant 0 is sorted; for o >= 1 the range
  // touches only octant 0's zero-initialised tail unless it reaches slot nq.
  {
    int ko = nn[0];
    for (int o = 1; o < 8; ++o) {
      if (nn[o] >= 2 && ko + nn[o] - 1 >= nq) { status[0] = APC_ERR_UNSUPPORTED; status[1] = i0; return; }
      ko += nn[o];
    }
  }
  // octant 0 in insertion order, quicksorted on the squared distance
  int k = 0;
  for (int e = 0; e < l; ++e)
    if ((c.MO(e) >> 29) == 0) { c.WD(k) = c.D2(e); c.WM(k) = (int)(c.MO(e) & 0x1fffffff); ++k; }
  if (nn[0] > 1) qsort_ref(c, 0, nn[0] - 1);
  {
    double pen = 1.0;
    for (int q = 0; q < nn[0]; ++q) { c.WD(q) = pen * c.WD(q); pen = pen * 10000.0; }
  }
  for (int o = 1; o < 8; ++o) {
    double pen = 1.0;
    for (int e = 0; e < l; ++e)
      if ((int)(c.MO(e) >> 29) == o) {
        c.WD(k) = pen * c.D2(e);
        c.WM(k) = (int)(c.MO(e) & 0x1fffffff);
        pen = pen * 10000.0;
        ++k;
      }
  }
  qsort_ref(c, 0, l - 1);
  for (int q = 0; q < lm; ++q) {
    int m = c.WM(q);
    nbr[(size_t)q * M + i0] = m;
    dist[(size_t)q * M + i0] = sqrt(sqd(m));
  }
}

int search_neighbors(apc_ctx *c) {
  if (!c->have_comp) return fail(c, APC_ERR_ARG, "neighbours: no computational particles");
  const int M = c->nc + c->nv, S = M + c->nb, lm = c->prm.lm;
  if (lm < 9) return fail(c, APC_ERR_ARG, "NumberOfNeighbors must be >= 9");
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
  // candidate capacity: an expansion round starts with fewer than lm
  // candidates, each adding at most 24 (the 2:1 bound), so 25*lm suffices
  const int cap = 25 * lm + 32;
  const int chunk = std::max(1, std::min(M, 131072));
  APC_ALLOC(c, c->nscratch, (size_t)2 * cap * chunk);
  APC_ALLOC(c, c->nscratch_i, (size_t)2 * cap * chunk);
  uint32_t *mo = (uint32_t *)c->nscratch_i.p;
  int *wm = c->nscratch_i.p + (size_t)cap * chunk;
  double *d2 = c->nscratch.p;
  double *wd = c->nscratch.p + (size_t)cap * chunk;
  for (int row0 = 0; row0 < M; row0 += chunk) {
    int rows = std::min(chunk, M - row0);
    KScope ks(c, "neighbor_search");
    hipLaunchKernelGGL(k_search, dim3(nblocks(rows, 256)), dim3(256), 0, c->s, M, S, row0, rows,
                       lm, cap, F, c->cx.p, c->cy.p, c->cz.p, c->n24.p, c->n24cnt.p, mo, d2, wd,
                       wm, c->nbr.p, c->dist.p, c->status_dev.p);
  }
  int st[2] = {0, 0};
  APC_HIP(c, hipMemcpyAsync(st, c->status_dev.p, sizeof(st), hipMemcpyDeviceToHost, c->s));
  APC_HIP(c, hipStreamSynchronize(c->s));
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
  APC_HIP(c, hipMemcpy(tn.data(), c->nbr.p, 4ull * M * lm, hipMemcpyDeviceToHost));
  APC_HIP(c, hipMemcpy(td.data(), c->dist.p, 8ull * M * lm, hipMemcpyDeviceToHost));
  for (int i = 0; i < M; ++i)
    for (int q = 0; q < lm; ++q) {
      if (nbr) nbr[(size_t)i * lm + q] = tn[(size_t)q * M + i];
      if (dist) dist[(size_t)i * lm + q] = td[(size_t)q * M + i];
    }
  return 0;
}
