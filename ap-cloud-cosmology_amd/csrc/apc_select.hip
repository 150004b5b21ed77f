// Computational-particle selection on the GPU (octree.cpp:1050-3001).
//
// The reference builds a temporary tree breadth first, then runs a serial
// bottom-up ripple that opens coarse leaves until every leaf's face
// neighbours are at most one level coarser, and emits leaves + Neigh24.
// Here the same tree is built level-synchronously:
//   * top-down: one launch per level decides which nodes open
//     ((n_p h) h > c, octree.cpp:1164-1183) and creates their eight children
//     with particle ranges found by binary search in the sorted keys;
//   * balance: for ln = deepest .. 2, every leaf at level ln looks up the
//     level ln-1 cell across each face; a coarser leaf found there is claimed
//     (atomicCAS) and opened, repeated until no claim is made — the same
//     minimal 2:1 face-balanced refinement the reference's ripple produces;
//   * output: computational leaves in creation (= reference) order, vacuum
//     leaves by (level, Morton), boundary nodes in leaf order; Neigh24 per
//     leaf in the reference's direction order x+ x- y+ y- z+ z- with the
//     four facing children of a finer neighbour in the reference's order.
// Built with -ffp-contract=off (node centres must match the reference).
#include "apc_internal.h"
#include "apc_prims.h"

namespace apc {

struct SelGeom {
  int L, bc, order, mindepth, maxdepth, leaf_cap, faces;
  double c;
  double hl[24], hx[24], hy[24], hz[24];
  double bb[6];
};

struct Tree {
  uint64_t *prefix;
  int *level, *first, *np, *status, *child, *parent;
  double *cx, *cy, *cz;
};

__device__ inline bool opens(const SelGeom &g, int np, int lv) {
  bool crit;
  if (g.leaf_cap > 0) {
    crit = np > g.leaf_cap;
  } else {
    double temp = 1.0 * np;
    double h = g.hl[lv];
    for (int q = 0; q < 2 * g.order - 2; ++q) temp = temp * h;
    crit = temp > g.c;
  }
  return (crit || lv < g.mindepth) && lv < g.maxdepth;
}

// First position in [lo, hi) whose level-`lv` digit is >= d (digits are
// non-decreasing there because the keys are sorted and share the prefix).
__device__ inline int lower_digit(const uint64_t *__restrict__ key, int lo, int hi, int sh, int d) {
  while (lo < hi) {
    int mid = lo + ((hi - lo) >> 1);
    int dm = (int)((key[mid] >> sh) & 7);
    if (dm < d) lo = mid + 1; else hi = mid;
  }
  return lo;
}

// Create the eight children of node p at base..base+7.
__device__ void make_children(const SelGeom &g, Tree t, const uint64_t *__restrict__ key, int p,
                              int base, int st) {
  const int lv = t.level[p];
  const int f = t.first[p], n = t.np[p];
  const int sh = 3 * (g.L - lv - 1);
  int b[9];
  b[0] = f;
  b[8] = f + n;
  if (n > 0) {
    for (int d = 1; d < 8; ++d) b[d] = lower_digit(key, b[d - 1], f + n, sh, d);
  } else {
    for (int d = 1; d < 8; ++d) b[d] = f;
  }
  const uint64_t P = t.prefix[p];
  const double px = t.cx[p], py = t.cy[p], pz = t.cz[p];
  const double qx = 0.25 * g.hx[lv], qy = 0.25 * g.hy[lv], qz = 0.25 * g.hz[lv];
  for (int d = 0; d < 8; ++d) {
    int k = base + d;
    t.prefix[k] = (P << 3) | (uint64_t)d;
    t.level[k] = lv + 1;
    t.first[k] = b[d];
    t.np[k] = b[d + 1] - b[d];
    t.status[k] = st;
    t.child[k] = -1;
    t.parent[k] = p;
    t.cx[k] = (d & 4) ? px + qx : px - qx;
    t.cy[k] = (d & 2) ? py + qy : py - qy;
    t.cz[k] = (d & 1) ? pz + qz : pz - qz;
  }
}

// Deepest existing node containing cell (tlv, tprefix), walking up from
// `start` and then down through opened nodes.
__device__ inline int find_cell(Tree t, int start, int tlv, uint64_t tprefix) {
  int a = start;
  for (;;) {
    int la = t.level[a];
    if (la <= tlv && (tprefix >> (3 * (tlv - la))) == t.prefix[a]) break;
    a = t.parent[a];
  }
  while (t.level[a] < tlv && t.status[a] == ST_OPENED) {
    int la = t.level[a];
    int d = (int)((tprefix >> (3 * (tlv - la - 1))) & 7);
    a = t.child[a] + d;
  }
  return a;
}

// Same-level neighbour of cell `prefix` at level lv across face dir
// (0 x+, 1 x-, 2 y-, 3 y+, 4 z+, 5 z-).  Returns false on a non-periodic
// domain face.
__device__ inline bool face_neighbour(uint64_t prefix, int lv, int dir, int bc, uint64_t *out) {
  uint32_t ix, iy, iz;
  morton_decode(prefix, ix, iy, iz);
  const int64_t m = (int64_t)1 << lv;
  int64_t c[3] = {ix, iy, iz};
  int ax = dir < 2 ? 0 : (dir < 4 ? 1 : 2);
  int up = (dir == 0 || dir == 3 || dir == 4) ? 1 : -1;
  int64_t v = c[ax] + up;
  if (v < 0 || v >= m) {
    if (bc != 1) return false;
    v = (v + m) % m;
  }
  c[ax] = v;
  *out = morton_encode((uint32_t)c[0], (uint32_t)c[1], (uint32_t)c[2]);
  return true;
}

// ------------------------------------------------------------- top-down --
__global__ void k_td_flags(SelGeom g, Tree t, int a, int b, int *__restrict__ flag) {
  int i = a + blockIdx.x * 256 + threadIdx.x;
  if (i >= b) return;
  flag[i - a] = opens(g, t.np[i], t.level[i]) ? 1 : 0;
}

__global__ void k_td_open(SelGeom g, Tree t, const uint64_t *__restrict__ key, int a, int b,
                          const int *__restrict__ flag, const int *__restrict__ rank) {
  int i = a + blockIdx.x * 256 + threadIdx.x;
  if (i >= b) return;
  if (!flag[i - a]) return;
  int base = b + 8 * rank[i - a];
  make_children(g, t, key, i, base, ST_LEAF_TD);
  t.status[i] = ST_OPENED;
  t.child[i] = base;
}

// The roots: one per box (nodes 0..nbox-1), each over its box's particles
// (pstart; NULL for a plain context: one root over all n).
__global__ void k_roots(SelGeom g, Tree t, int nbox, const int *__restrict__ pstart, int n) {
  const int b = threadIdx.x;
  if (b >= nbox) return;
  t.prefix[b] = 0;
  t.level[b] = 0;
  t.first[b] = pstart ? pstart[b] : 0;
  t.np[b] = pstart ? pstart[b + 1] - pstart[b] : n;
  t.status[b] = ST_LEAF_TD;
  t.child[b] = -1;
  t.parent[b] = -1;
  t.cx[b] = (g.bb[1] + g.bb[0]) / 2;
  t.cy[b] = (g.bb[3] + g.bb[2]) / 2;
  t.cz[b] = (g.bb[5] + g.bb[4]) / 2;
}

// Box of a node: the root above it (roots are nodes 0..nbox-1).
__device__ inline int node_box(const Tree &t, int i) {
  while (t.parent[i] >= 0) i = t.parent[i];
  return i;
}

// Packed boxes: sort keys (box) and values (current output position) of the
// rows [comp (creation order) | vacuum (level, Morton)]; a stable sort by
// box then gives every box its computational rows, then its vacuum rows, in
// the order a context of that box alone would give them.  Also flags a leaf
// the lowered depth keeps from opening (cut != 0).
__global__ void k_pack_keys(SelGeom g, Tree t, int M, const int *__restrict__ leaf_node,
                            int maxdepth_req, uint64_t *__restrict__ key, int *__restrict__ val,
                            int *__restrict__ cut) {
  const int r = blockIdx.x * 256 + threadIdx.x;
  if (r >= M) return;
  const int i = leaf_node[r];
  key[r] = (uint64_t)node_box(t, i);
  val[r] = r;
  const int lv = t.level[i];
  if (lv == g.maxdepth && g.maxdepth < maxdepth_req) {
    bool crit;
    if (g.leaf_cap > 0) {
      crit = t.np[i] > g.leaf_cap;
    } else {
      double temp = 1.0 * t.np[i];
      for (int q = 0; q < 2 * g.order - 2; ++q) temp = temp * g.hl[lv];
      crit = temp > g.c;
    }
    if (crit) atomicOr(cut, 1);
  }
}

// Rows in box-major order: leaf, output map, type, box of each row; the
// first row of every box and of its vacuum rows (vstart, -1 if none).
__global__ void k_pack_emit(int M, int nc, const uint64_t *__restrict__ key,
                            const int *__restrict__ val, const int *__restrict__ leaf_in,
                            int *__restrict__ leaf_out, int *__restrict__ out,
                            int *__restrict__ rtype, int *__restrict__ rowbox,
                            int *__restrict__ rstart, int *__restrict__ vstart) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p >= M) return;
  const int b = (int)key[p], v = val[p];
  const int i = leaf_in[v];
  leaf_out[p] = i;
  out[i] = p;
  rtype[p] = (v < nc) ? 0 : -1;
  rowbox[p] = b;
  const bool first = (p == 0) || (int)key[p - 1] != b;
  if (first) rstart[b] = p;
  if (v >= nc && (first || val[p - 1] < nc)) vstart[b] = p;
}

__global__ void k_fill_int(int n, int v, int *__restrict__ a) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) a[i] = v;
}

// -------------------------------------------------------------- balance --
// Leaves at level ln: the top-down range [a,b) plus all balance nodes
// [Ttd, T) (filtered by level).
__global__ void k_bal_check(SelGeom g, Tree t, int ln, int a, int b, int Ttd, int T,
                            int *__restrict__ claim_list, int *__restrict__ cnt /* [0]=claims [1]=pending */) {
  long q = (long)blockIdx.x * 256 + threadIdx.x;
  long span = b - a;
  int i;
  if (q < span) i = a + (int)q;
  else if (q < span + (T - Ttd)) i = Ttd + (int)(q - span);
  else return;
  if (t.level[i] != ln) return;
  int st = t.status[i];
  if (st != ST_LEAF_TD && st != ST_LEAF_VAC) return;
  const uint64_t P = t.prefix[i];
  for (int dir = 0; dir < 6; ++dir) {
    uint64_t nbp;
    if (!face_neighbour(P, ln, dir, g.bc, &nbp)) continue;
    uint64_t R = nbp >> 3;  // its parent cell at level ln-1
    int D = find_cell(t, i, ln - 1, R);
    if (t.level[D] >= ln - 1) continue;
    int sd = t.status[D];
    if (sd == ST_LEAF_TD || sd == ST_LEAF_VAC) {
      if (atomicCAS(&t.status[D], sd, ST_CLAIMED) == sd) {
        int k = atomicAdd(&cnt[0], 1);
        claim_list[k] = D;
        continue;
      }
    }
    atomicOr(&cnt[1], 1);
  }
}

__global__ void k_bal_open(SelGeom g, Tree t, const uint64_t *__restrict__ key,
                           const int *__restrict__ claim_list, int nclaim, int T) {
  int q = blockIdx.x * 256 + threadIdx.x;
  if (q >= nclaim) return;
  int D = claim_list[q];
  int base = T + 8 * q;
  make_children(g, t, key, D, base, ST_LEAF_VAC);
  t.child[D] = base;
  t.status[D] = ST_OPENED;
}

// --------------------------------------------------------------- output --
__global__ void k_leaf_flags(Tree t, int a, int b, int want, int *__restrict__ flag) {
  int i = a + blockIdx.x * 256 + threadIdx.x;
  if (i >= b) return;
  flag[i - a] = (t.status[i] == want) ? 1 : 0;
}

__global__ void k_comp_emit(Tree t, int T_td, const int *__restrict__ flag,
                            const int *__restrict__ rank, int *__restrict__ leaf_node,
                            int *__restrict__ out) {
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= T_td || !flag[i]) return;
  int r = rank[i];
  leaf_node[r] = i;
  out[i] = r;
}

__global__ void k_vac_codes(Tree t, int a, int b, const int *__restrict__ flag,
                            const int *__restrict__ rank, uint64_t *__restrict__ code,
                            int *__restrict__ node) {
  int i = a + blockIdx.x * 256 + threadIdx.x;
  if (i >= b || !flag[i - a]) return;
  int r = rank[i - a];
  int lv = t.level[i];
  code[r] = (((uint64_t)1) << (3 * lv)) | t.prefix[i];
  node[r] = i;
}

__global__ void k_vac_emit(int nv, int nc, const int *__restrict__ node, int *__restrict__ leaf_node,
                           int *__restrict__ out) {
  int r = blockIdx.x * 256 + threadIdx.x;
  if (r >= nv) return;
  int i = node[r];
  leaf_node[nc + r] = i;
  out[i] = nc + r;
}

__global__ void k_leaf_data(SelGeom g, Tree t, int M, const int *__restrict__ leaf_node, int nc,
                            const int *__restrict__ rtype,
                            double *__restrict__ x, double *__restrict__ y, double *__restrict__ z,
                            double *__restrict__ h, int *__restrict__ np, int *__restrict__ first,
                            int *__restrict__ type, int *__restrict__ bcount) {
  int r = blockIdx.x * 256 + threadIdx.x;
  if (r >= M) return;
  int i = leaf_node[r];
  int lv = t.level[i];
  x[r] = t.cx[i];
  y[r] = t.cy[i];
  z[r] = t.cz[i];
  h[r] = g.hl[lv];
  np[r] = t.np[i];
  first[r] = t.first[i];
  type[r] = rtype ? rtype[r] : ((r < nc) ? 0 : -1);
  int nbf = 0;
  if (g.bc != 1) {
    uint64_t dummy;
    for (int dir = 0; dir < 6; ++dir)
      if (!face_neighbour(t.prefix[i], lv, dir, g.bc, &dummy)) ++nbf;
  }
  bcount[r] = nbf;
}

// Neigh24 in the order x+ x- y+ y- z+ z-; a finer neighbour contributes its
// four facing children (octree.cpp:2770-2970).
__constant__ int c_dorder[6] = {0, 1, 3, 2, 4, 5};
__constant__ int c_facing[6][4] = {{0, 1, 2, 3}, {4, 5, 6, 7}, {2, 3, 6, 7},
                                   {0, 1, 4, 5}, {0, 2, 4, 6}, {1, 3, 5, 7}};

__global__ void k_neigh24(SelGeom g, Tree t, int M, const int *__restrict__ leaf_node,
                          const int *__restrict__ out, const int *__restrict__ boff,
                          int *__restrict__ n24, int *__restrict__ n24cnt,
                          double *__restrict__ x, double *__restrict__ y, double *__restrict__ z,
                          double *__restrict__ h, int *__restrict__ np, int *__restrict__ first,
                          int *__restrict__ type, int *__restrict__ err) {
  int r = blockIdx.x * 256 + threadIdx.x;
  if (r >= M) return;
  int i = leaf_node[r];
  int lv = t.level[i];
  uint64_t P = t.prefix[i];
  int cntn = 0;
  int bnext = M + boff[r];
  int *row = n24 + (size_t)r * 24;
  for (int e = 0; e < 6; ++e) {
    int dir = c_dorder[e];
    uint64_t nbp;
    if (!face_neighbour(P, lv, dir, g.bc, &nbp)) {
      int ax = dir < 2 ? 0 : (dir < 4 ? 1 : 2);
      bool up = (dir == 0 || dir == 3 || dir == 4);
      double f = g.faces ? (up ? g.bb[2 * ax + 1] : g.bb[2 * ax]) : (up ? 1.0 : -1.0);
      int b = bnext++;
      x[b] = (ax == 0) ? f : t.cx[i];
      y[b] = (ax == 1) ? f : t.cy[i];
      z[b] = (ax == 2) ? f : t.cz[i];
      h[b] = g.hl[lv];
      np[b] = 0;
      first[b] = -1;
      type[b] = 1;
      row[cntn++] = b;
      continue;
    }
    int D = find_cell(t, i, lv, nbp);
    if (t.level[D] == lv && t.status[D] == ST_OPENED) {
      int cb = t.child[D];
      for (int f = 0; f < 4; ++f) {
        int k = cb + c_facing[dir][f];
        int st = t.status[k];
        if (st != ST_LEAF_TD && st != ST_LEAF_VAC) atomicOr(err, 1);
        row[cntn++] = out[k];
      }
    } else {
      int st = t.status[D];
      if (st != ST_LEAF_TD && st != ST_LEAF_VAC) atomicOr(err, 2);
      row[cntn++] = out[D];
    }
  }
  n24cnt[r] = cntn;
}

// sorted position -> output leaf
__global__ void k_owner(int M, const int *__restrict__ first, const int *__restrict__ np,
                        int *__restrict__ owner) {
  int r = blockIdx.x * 256 + threadIdx.x;
  if (r >= M) return;
  int f = first[r], n = np[r];
  for (int s = f; s < f + n; ++s) owner[s] = r;
}

static SelGeom make_geom(apc_ctx *c, double cc, int order, int mindepth, int maxdepth) {
  SelGeom g;
  const apc_params &p = c->prm;
  g.L = p.depth;
  g.bc = p.bc;
  g.order = order;
  g.mindepth = mindepth;
  g.maxdepth = maxdepth;
  g.leaf_cap = p.leaf_capacity;
  g.faces = p.boundary_at_faces;
  g.c = cc;
  for (int q = 0; q < 6; ++q) g.bb[q] = p.bbox[q];
  // octree.cpp:1127-1133: h0 = cbrt of the root volume (host pow, like the
  // reference), widths halved per level
  double hx0 = p.bbox[1] - p.bbox[0], hy0 = p.bbox[3] - p.bbox[2], hz0 = p.bbox[5] - p.bbox[4];
  double h0 = std::pow(hx0 * hy0 * hz0, 1.0 / 3);
  c->h0 = h0;
  double h = h0, hx = hx0, hy = hy0, hz = hz0;
  c->hlev.assign(24, 0);
  for (int l = 0; l < 24; ++l) {
    g.hl[l] = h; g.hx[l] = hx; g.hy[l] = hy; g.hz[l] = hz;
    c->hlev[l] = h;
    h = 0.5 * h; hx = 0.5 * hx; hy = 0.5 * hy; hz = 0.5 * hz;
  }
  return g;
}

static int ensure_nodes(apc_ctx *c, size_t n) {
  APC_ALLOC(c, c->nprefix, n);
  APC_ALLOC(c, c->nlevel, n);
  APC_ALLOC(c, c->nfirst, n);
  APC_ALLOC(c, c->nnp, n);
  APC_ALLOC(c, c->nstatus, n);
  APC_ALLOC(c, c->nchild, n);
  APC_ALLOC(c, c->nparent, n);
  APC_ALLOC(c, c->ncx, n);
  APC_ALLOC(c, c->ncy, n);
  APC_ALLOC(c, c->ncz, n);
  APC_ALLOC(c, c->nout, n);
  return 0;
}

static Tree tree_of(apc_ctx *c) {
  Tree t;
  t.prefix = c->nprefix.p;
  t.level = c->nlevel.p;
  t.first = c->nfirst.p;
  t.np = c->nnp.p;
  t.status = c->nstatus.p;
  t.child = c->nchild.p;
  t.parent = c->nparent.p;
  t.cx = c->ncx.p;
  t.cy = c->ncy.p;
  t.cz = c->ncz.p;
  return t;
}

static int read_ints(apc_ctx *c, const int *dev, int *host, int n) {
  APC_HIP(c, hipMemcpyAsync(host, dev, sizeof(int) * n, hipMemcpyDeviceToHost, c->s));
  APC_HIP(c, hipStreamSynchronize(c->s));
  return 0;
}

int select_tree(apc_ctx *c, double cc, int order, int mindepth, int maxdepth) {
  if (!c->sorted) return fail(c, APC_ERR_ARG, "select: octree not built");
  const apc_params &p = c->prm;
  // a packed context's key budget may have lowered the depth below the
  // caller's: select to the lowered depth and refuse (below) a tree that the
  // caller's depth would have refined further
  int maxdepth_req = maxdepth;
  if (c->nbox > 1 && maxdepth > p.depth && maxdepth <= c->depth_req && mindepth <= p.depth)
    maxdepth = p.depth;
  if (maxdepth > p.depth || maxdepth < 0 || mindepth < 0 || order < 1)
    return fail(c, APC_ERR_UNSUPPORTED, "select: need 0 <= mindepth, maxdepth <= depth (%d)", p.depth);
  SelGeom g = make_geom(c, cc, order, mindepth, maxdepth);
  const int n = c->N;
  APC_ALLOC(c, c->dcount, 16);
  size_t cap = std::max<size_t>(1024, (size_t)c->T + (size_t)c->T / 4);
  if (c->nprefix.cap > cap) cap = c->nprefix.cap;
  int r = ensure_nodes(c, cap);
  if (r) return r;
  Tree t = tree_of(c);
  const int nbox = c->nbox;
  hipLaunchKernelGGL(k_roots, dim3(1), dim3(64), 0, c->s, g, t, nbox,
                     nbox > 1 ? (const int *)c->pstart.p : nullptr, n);
  c->lvl_start.assign(1, 0);
  c->lvl_start.push_back(nbox);
  int T = nbox;
  int hc[4];
  // ---- top-down, one level per pass
  for (int lv = 0;; ++lv) {
    int a = c->lvl_start[lv], b = c->lvl_start[lv + 1];
    int m = b - a;
    APC_ALLOC(c, c->nflag, m);
    APC_ALLOC(c, c->nrank, m);
    {
      KScope ks(c, "select_topdown");
      hipLaunchKernelGGL(k_td_flags, dim3(nblocks(m, 256)), dim3(256), 0, c->s, g, t, a, b,
                         c->nflag.p);
    }
    r = scan_exclusive(c, c->nflag.p, c->nrank.p, m, c->dcount.p);
    if (r) return r;
    r = read_ints(c, c->dcount.p, hc, 1);
    if (r) return r;
    int K = hc[0];
    if (K == 0) break;
    size_t need = (size_t)b + 8 * (size_t)K;
    if (need > c->nprefix.cap) {
      // grow, preserving the tree built so far
      Tree old = t;
      size_t ncap = need + need / 2;
      DBuf<uint64_t> pr; DBuf<int> i6[7]; DBuf<double> d3[3];
      if (pr.ensure(ncap, c->s) != hipSuccess) return fail(c, APC_ERR_NOMEM, "select: grow");
      for (auto &q : i6) if (q.ensure(ncap, c->s) != hipSuccess) return fail(c, APC_ERR_NOMEM, "select: grow");
      for (auto &q : d3) if (q.ensure(ncap, c->s) != hipSuccess) return fail(c, APC_ERR_NOMEM, "select: grow");
      APC_HIP(c, hipMemcpyAsync(pr.p, old.prefix, 8ull * b, hipMemcpyDeviceToDevice, c->s));
      const int *srcs[7] = {old.level, old.first, old.np, old.status, old.child, old.parent, c->nout.p};
      for (int q = 0; q < 7; ++q)
        APC_HIP(c, hipMemcpyAsync(i6[q].p, srcs[q], 4ull * b, hipMemcpyDeviceToDevice, c->s));
      const double *sd[3] = {old.cx, old.cy, old.cz};
      for (int q = 0; q < 3; ++q)
        APC_HIP(c, hipMemcpyAsync(d3[q].p, sd[q], 8ull * b, hipMemcpyDeviceToDevice, c->s));
      APC_HIP(c, hipStreamSynchronize(c->s));
      std::swap(c->nprefix, pr);
      std::swap(c->nlevel, i6[0]); std::swap(c->nfirst, i6[1]); std::swap(c->nnp, i6[2]);
      std::swap(c->nstatus, i6[3]); std::swap(c->nchild, i6[4]); std::swap(c->nparent, i6[5]);
      std::swap(c->nout, i6[6]);
      std::swap(c->ncx, d3[0]); std::swap(c->ncy, d3[1]); std::swap(c->ncz, d3[2]);
      pr.release(c->s);
      for (auto &q : i6) q.release(c->s);
      for (auto &q : d3) q.release(c->s);
      t = tree_of(c);
    }
    {
      KScope ks(c, "select_topdown");
      hipLaunchKernelGGL(k_td_open, dim3(nblocks(m, 256)), dim3(256), 0, c->s, g, t, c->key.p, a,
                         b, c->nflag.p, c->nrank.p);
    }
    T = b + 8 * K;
    c->lvl_start.push_back(T);
  }
  const int nlev = (int)c->lvl_start.size() - 1;  // levels 0..nlev-1 populated
  c->ml = nlev - 1;
  c->T_td = T;

  // ---- 2:1 face balance, deepest level first
  APC_ALLOC(c, c->nscratch_i, 1024);
  for (int ln = c->ml; ln > 1; --ln) {
    for (int iter = 0;; ++iter) {
      // each pass opens one more level of a refinement chain; chains are
      // shorter than the tree depth
      if (iter > p.depth + 2) return fail(c, APC_ERR_BALANCE, "select: balance did not converge");
      int a = c->lvl_start[ln], b = c->lvl_start[ln + 1];
      long span = (long)(b - a) + (T - c->T_td);
      APC_HIP(c, hipMemsetAsync(c->dcount.p, 0, 2 * sizeof(int), c->s));
      // worst case every candidate claims six nodes
      APC_ALLOC(c, c->nscratch_i, (size_t)6 * span + 16);
      {
        KScope ks(c, "select_balance");
        hipLaunchKernelGGL(k_bal_check, dim3(nblocks(span, 256)), dim3(256), 0, c->s, g, t, ln, a,
                           b, c->T_td, T, c->nscratch_i.p, c->dcount.p);
      }
      r = read_ints(c, c->dcount.p, hc, 2);
      if (r) return r;
      int nclaim = hc[0];
      if (nclaim == 0) {
        if (hc[1]) return fail(c, APC_ERR_BALANCE, "select: balance stalled at level %d", ln);
        break;
      }
      size_t need = (size_t)T + 8 * (size_t)nclaim;
      if (need > c->nprefix.cap) {
        // grow preserving content (rare: capacity carries over between steps)
        DBuf<uint64_t> pr; DBuf<int> i6[7]; DBuf<double> d3[3];
        size_t ncap = need + need / 2;
        if (pr.ensure(ncap, c->s) != hipSuccess) return fail(c, APC_ERR_NOMEM, "select: grow");
        for (auto &q : i6) if (q.ensure(ncap, c->s) != hipSuccess) return fail(c, APC_ERR_NOMEM, "select: grow");
        for (auto &q : d3) if (q.ensure(ncap, c->s) != hipSuccess) return fail(c, APC_ERR_NOMEM, "select: grow");
        APC_HIP(c, hipMemcpyAsync(pr.p, t.prefix, 8ull * T, hipMemcpyDeviceToDevice, c->s));
        const int *srcs[7] = {t.level, t.first, t.np, t.status, t.child, t.parent, c->nout.p};
        for (int q = 0; q < 7; ++q)
          APC_HIP(c, hipMemcpyAsync(i6[q].p, srcs[q], 4ull * T, hipMemcpyDeviceToDevice, c->s));
        const double *sd[3] = {t.cx, t.cy, t.cz};
        for (int q = 0; q < 3; ++q)
          APC_HIP(c, hipMemcpyAsync(d3[q].p, sd[q], 8ull * T, hipMemcpyDeviceToDevice, c->s));
        APC_HIP(c, hipStreamSynchronize(c->s));
        std::swap(c->nprefix, pr);
        std::swap(c->nlevel, i6[0]); std::swap(c->nfirst, i6[1]); std::swap(c->nnp, i6[2]);
        std::swap(c->nstatus, i6[3]); std::swap(c->nchild, i6[4]); std::swap(c->nparent, i6[5]);
        std::swap(c->nout, i6[6]);
        std::swap(c->ncx, d3[0]); std::swap(c->ncy, d3[1]); std::swap(c->ncz, d3[2]);
        pr.release(c->s);
        for (auto &q : i6) q.release(c->s);
        for (auto &q : d3) q.release(c->s);
        t = tree_of(c);
      }
      {
        KScope ks(c, "select_balance");
        hipLaunchKernelGGL(k_bal_open, dim3(nblocks(nclaim, 256)), dim3(256), 0, c->s, g, t,
                           c->key.p, c->nscratch_i.p, nclaim, T);
      }
      T += 8 * nclaim;
    }
  }
  c->T = T;

  // ---- output: computational leaves (creation order), vacuum (level, Morton)
  int Ttd = c->T_td;
  APC_ALLOC(c, c->nflag, std::max(Ttd, T - Ttd) + 1);
  APC_ALLOC(c, c->nrank, std::max(Ttd, T - Ttd) + 1);
  hipLaunchKernelGGL(k_leaf_flags, dim3(nblocks(Ttd, 256)), dim3(256), 0, c->s, t, 0, Ttd,
                     (int)ST_LEAF_TD, c->nflag.p);
  r = scan_exclusive(c, c->nflag.p, c->nrank.p, Ttd, c->dcount.p);
  if (r) return r;
  r = read_ints(c, c->dcount.p, hc, 1);
  if (r) return r;
  const int nc = hc[0];
  APC_ALLOC(c, c->leaf_node, (size_t)nc + (T - Ttd) + 1);
  hipLaunchKernelGGL(k_comp_emit, dim3(nblocks(Ttd, 256)), dim3(256), 0, c->s, t, Ttd, c->nflag.p,
                     c->nrank.p, c->leaf_node.p, c->nout.p);
  int nv = 0;
  if (T > Ttd) {
    int m = T - Ttd;
    hipLaunchKernelGGL(k_leaf_flags, dim3(nblocks(m, 256)), dim3(256), 0, c->s, t, Ttd, T,
                       (int)ST_LEAF_VAC, c->nflag.p);
    r = scan_exclusive(c, c->nflag.p, c->nrank.p, m, c->dcount.p);
    if (r) return r;
    r = read_ints(c, c->dcount.p, hc, 1);
    if (r) return r;
    nv = hc[0];
    if (nv > 0) {  // the sort buffers stay with the context (no allocation per step)
      APC_ALLOC(c, c->vcode, nv);
      APC_ALLOC(c, c->vcode2, nv);
      APC_ALLOC(c, c->vnode, nv);
      APC_ALLOC(c, c->vnode2, nv);
      hipLaunchKernelGGL(k_vac_codes, dim3(nblocks(m, 256)), dim3(256), 0, c->s, t, Ttd, T,
                         c->nflag.p, c->nrank.p, c->vcode.p, c->vnode.p);
      r = radix_sort_pairs(c, c->vcode.p, c->vnode.p, c->vcode2.p, c->vnode2.p, nv, 3 * p.depth + 1);
      if (r) return r;
      hipLaunchKernelGGL(k_vac_emit, dim3(nblocks(nv, 256)), dim3(256), 0, c->s, nv, nc,
                         c->vnode.p, c->leaf_node.p, c->nout.p);
    }
  }
  const int M = nc + nv;
  const int *rtype = nullptr;
  if (nbox > 1) {
    // box-major rows: one stable sort of [comp | vacuum] by box
    APC_ALLOC(c, c->pk_key, M + 1); APC_ALLOC(c, c->pk_key2, M + 1);
    APC_ALLOC(c, c->pk_val, M + 1); APC_ALLOC(c, c->pk_val2, M + 1);
    APC_ALLOC(c, c->pk_leaf, M + 1); APC_ALLOC(c, c->pk_type, M + 1);
    APC_ALLOC(c, c->rowbox, M + 1); APC_ALLOC(c, c->rstart, 2 * nbox + 2);
    APC_HIP(c, hipMemsetAsync(c->dcount.p + 8, 0, sizeof(int), c->s));
    hipLaunchKernelGGL(k_pack_keys, dim3(nblocks(M, 256)), dim3(256), 0, c->s, g, t, M,
                       c->leaf_node.p, maxdepth_req, c->pk_key.p, c->pk_val.p, c->dcount.p + 8);
    r = radix_sort_pairs(c, c->pk_key.p, c->pk_val.p, c->pk_key2.p, c->pk_val2.p, M, c->boxbits);
    if (r) return r;
    hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(256), 0, c->s, 2 * nbox + 2, -1, c->rstart.p);
    hipLaunchKernelGGL(k_pack_emit, dim3(nblocks(M, 256)), dim3(256), 0, c->s, M, nc,
                       c->pk_key.p, c->pk_val.p, c->leaf_node.p, c->pk_leaf.p, c->nout.p,
                       c->pk_type.p, c->rowbox.p, c->rstart.p, c->rstart.p + nbox + 1);
    std::swap(c->leaf_node, c->pk_leaf);
    rtype = c->pk_type.p;
    std::vector<int> st(2 * nbox + 2);
    int cut = 0;
    APC_HIP(c, hipMemcpyAsync(st.data(), c->rstart.p, sizeof(int) * (2 * nbox + 2),
                              hipMemcpyDeviceToHost, c->s));
    APC_HIP(c, hipMemcpyAsync(&cut, c->dcount.p + 8, sizeof(int), hipMemcpyDeviceToHost, c->s));
    APC_HIP(c, hipStreamSynchronize(c->s));
    if (cut)
      return fail(c, APC_ERR_UNSUPPORTED, "select: packed boxes cut the tree at depth %d (the "
                  "64-bit sort key leaves 3 bits per level below the box id); use fewer boxes",
                  p.depth);
    c->rstart_h.assign(nbox + 1, M);
    c->rnc_h.assign(nbox, 0);
    for (int b = 0; b < nbox; ++b) {
      if (st[b] < 0) return fail(c, APC_ERR_ARG, "select: box %d has no rows", b);
      c->rstart_h[b] = st[b];
    }
    for (int b = 0; b < nbox; ++b) {
      const int vs = st[nbox + 1 + b] >= 0 ? st[nbox + 1 + b] : c->rstart_h[b + 1];
      c->rnc_h[b] = vs - c->rstart_h[b];
    }
    st.assign(c->rstart_h.begin(), c->rstart_h.end());
    APC_HIP(c, hipMemcpyAsync(c->rstart.p, st.data(), sizeof(int) * (nbox + 1),
                              hipMemcpyHostToDevice, c->s));
  } else {
    c->rstart_h.assign({0, M});
    c->rnc_h.assign(1, nc);
  }
  // boundary faces per leaf -> boundary node ids
  APC_ALLOC(c, c->n24first, (size_t)M + 1);
  APC_ALLOC(c, c->n24cnt, (size_t)M + 1);
  // worst case six boundary nodes per leaf
  size_t Smax = (size_t)M + (p.bc == 1 ? 0 : 6 * (size_t)M);
  APC_ALLOC(c, c->cx, Smax + 1); APC_ALLOC(c, c->cy, Smax + 1); APC_ALLOC(c, c->cz, Smax + 1);
  APC_ALLOC(c, c->ch, Smax + 1);
  APC_ALLOC(c, c->cnp, Smax + 1); APC_ALLOC(c, c->cfirst, Smax + 1); APC_ALLOC(c, c->ctype, Smax + 1);
  {
    KScope ks(c, "select_output");
    hipLaunchKernelGGL(k_leaf_data, dim3(nblocks(M, 256)), dim3(256), 0, c->s, g, t, M,
                       c->leaf_node.p, nc, rtype, c->cx.p, c->cy.p, c->cz.p, c->ch.p, c->cnp.p,
                       c->cfirst.p, c->ctype.p, c->n24cnt.p);
  }
  r = scan_exclusive(c, c->n24cnt.p, c->n24first.p, M, c->dcount.p);
  if (r) return r;
  r = read_ints(c, c->dcount.p, hc, 1);
  if (r) return r;
  const int nb = hc[0];
  APC_ALLOC(c, c->n24, (size_t)M * 24 + 1);
  APC_HIP(c, hipMemsetAsync(c->dcount.p + 4, 0, sizeof(int), c->s));
  {
    KScope ks(c, "select_neigh24");
    hipLaunchKernelGGL(k_neigh24, dim3(nblocks(M, 256)), dim3(256), 0, c->s, g, t, M,
                       c->leaf_node.p, c->nout.p, c->n24first.p, c->n24.p, c->n24cnt.p, c->cx.p,
                       c->cy.p, c->cz.p, c->ch.p, c->cnp.p, c->cfirst.p, c->ctype.p,
                       c->dcount.p + 4);
  }
  APC_ALLOC(c, c->owner, n);
  {
    KScope ks(c, "select_owner");
    hipLaunchKernelGGL(k_owner, dim3(nblocks(M, 256)), dim3(256), 0, c->s, M, c->cfirst.p,
                       c->cnp.p, c->owner.p);
  }
  r = read_ints(c, c->dcount.p + 4, hc, 1);
  if (r) return r;
  if (hc[0]) return fail(c, APC_ERR_BALANCE, "select: Neigh24 met a non-leaf (code %d)", hc[0]);
  APC_HIP(c, hipGetLastError());
  c->nc = nc;
  c->nv = nv;
  c->nb = nb;
  c->have_comp = true;
  c->comp_external = 0;
  c->have_tr = 0;
  c->have_nbr = c->have_sys = c->have_fields = 0;
  return 0;
}

}  // namespace apc

using namespace apc;

extern "C" int apc_select(apc_ctx *c, double cc, int order, int mindepth, int maxdepth, int *S_out) {
  if (!c) return APC_ERR_ARG;
  if (c->frozen) return fail(c, APC_ERR_ARG, "select: the topology is frozen");
  int r = select_tree(c, cc, order, mindepth, maxdepth);
  if (r) return r;
  if (S_out) *S_out = c->nc + c->nv + c->nb;
  return 0;
}

extern "C" int apc_get_comp_sizes(apc_ctx *c, int *nc, int *nv, int *nb, int *nneigh) {
  if (!c) return APC_ERR_ARG;
  if (!c->have_comp) return fail(c, APC_ERR_ARG, "no computational particles selected");
  if (nc) *nc = c->nc;
  if (nv) *nv = c->nv;
  if (nb) *nb = c->nb;
  if (nneigh) {
    const int M = c->nc + c->nv;
    int last = 0, first = 0;
    if (M > 0) {
      APC_HIP(c, hipMemcpyAsync(&last, c->n24cnt.p + M - 1, 4, hipMemcpyDeviceToHost, c->s));
      APC_HIP(c, hipMemcpyAsync(&first, c->n24first.p + M - 1, 4, hipMemcpyDeviceToHost, c->s));
      APC_HIP(c, hipStreamSynchronize(c->s));
    }
    // n24first holds the boundary-count scan; the Neigh24 total is summed on the host
    std::vector<int> cnt(M);
    if (M > 0) {
      APC_HIP(c, hipMemcpyAsync(cnt.data(), c->n24cnt.p, 4ull * M, hipMemcpyDeviceToHost, c->s));
      APC_HIP(c, hipStreamSynchronize(c->s));
    }
    long tot = 0;
    for (int q = 0; q < M; ++q) tot += cnt[q];
    *nneigh = (int)tot;
    (void)last; (void)first;
  }
  return 0;
}

extern "C" int apc_get_comp(apc_ctx *c, const apc_comp_view *v) {
  if (!c || !v) return APC_ERR_ARG;
  if (!c->have_comp) return fail(c, APC_ERR_ARG, "no computational particles selected");
  const int M = c->nc + c->nv, S = M + c->nb;
  auto cpd = [&](double *dst, const double *src, int n) -> hipError_t {
    if (!dst || n <= 0) return hipSuccess;
    return hipMemcpyAsync(dst, src, 8ull * n, hipMemcpyDeviceToHost, c->s);
  };
  auto cpi = [&](int *dst, const int *src, int n) -> hipError_t {
    if (!dst || n <= 0) return hipSuccess;
    return hipMemcpyAsync(dst, src, 4ull * n, hipMemcpyDeviceToHost, c->s);
  };
  APC_HIP(c, cpd(v->x, c->cx.p, S));
  APC_HIP(c, cpd(v->y, c->cy.p, S));
  APC_HIP(c, cpd(v->z, c->cz.p, S));
  APC_HIP(c, cpd(v->cell_length, c->ch.p, S));
  APC_HIP(c, cpi(v->np, c->cnp.p, S));
  APC_HIP(c, cpi(v->first, c->cfirst.p, S));
  APC_HIP(c, cpi(v->type, c->ctype.p, S));
  if (v->first_neigh || v->number_neigh || v->neigh24) {
    std::vector<int> cnt(M), pad((size_t)M * 24);
    if (M > 0) {
      APC_HIP(c, hipMemcpyAsync(cnt.data(), c->n24cnt.p, 4ull * M, hipMemcpyDeviceToHost, c->s));
      APC_HIP(c, hipMemcpyAsync(pad.data(), c->n24.p, 4ull * M * 24, hipMemcpyDeviceToHost, c->s));
    }
    APC_HIP(c, hipStreamSynchronize(c->s));
    long tot = 0;
    for (int q = 0; q < M; ++q) {
      if (v->first_neigh) v->first_neigh[q] = (int)tot;
      if (v->number_neigh) v->number_neigh[q] = cnt[q];
      if (v->neigh24)
        for (int k = 0; k < cnt[q]; ++k) v->neigh24[tot + k] = pad[(size_t)q * 24 + k];
      tot += cnt[q];
    }
    for (int q = M; q < S; ++q) {
      if (v->first_neigh) v->first_neigh[q] = (int)tot;
      if (v->number_neigh) v->number_neigh[q] = 0;
    }
  }
  if (c->have_fields) {
    APC_HIP(c, cpd(v->density, c->rho.p, M));
    APC_HIP(c, cpd(v->potential, c->phi.p, M));
    double *dst[9] = {v->p_x, v->p_y, v->p_z, v->p_xx, v->p_yy, v->p_zz, v->p_xy, v->p_xz, v->p_yz};
    for (int q = 0; q < 9; ++q) APC_HIP(c, cpd(dst[q], c->grad.p + (size_t)q * M, M));
  }
  APC_HIP(c, hipStreamSynchronize(c->s));
  return 0;
}

// Solver::Solver takes a Computational_particle that its caller filled
// (solver.h:53, particle.h:23-53); the reference fills it by the octree
// selection, but the Solver does not depend on that.  This loads such a
// caller-filled set: rows nc computational (Type 0) then nv vacuum (Type -1),
// then nb boundary nodes (Type 1); Neigh24 of row i is
// neigh24[first_neigh[i] .. first_neigh[i] + number_neigh[i]).  The entries
// themselves are checked where the neighbour search reads them, in the
// reference's order (solver.cpp:511-518).  The set has no physical
// particles: the mover entry points refuse it until the next selection.
extern "C" int apc_set_comp(apc_ctx *c, int nc, int nv, int nb, const double *x, const double *y,
                            const double *z, const double *cell_length, const int *np,
                            const int *type, const int *first_neigh, const int *number_neigh,
                            const int *neigh24, int nneigh) {
  if (!c) return APC_ERR_ARG;
  if (c->frozen) return fail(c, APC_ERR_ARG, "set_comp: the topology is frozen");
  if (c->nbox > 1) return fail(c, APC_ERR_UNSUPPORTED, "set_comp: not on a packed multi-box context");
  if (nc < 0 || nv < 0 || nb < 0 || nc + nv < 1 || nneigh < 0)
    return fail(c, APC_ERR_ARG, "set_comp: bad sizes nc=%d nv=%d nb=%d", nc, nv, nb);
  if (!x || !y || !z || !cell_length || !np || !type || !first_neigh || !number_neigh ||
      (nneigh > 0 && !neigh24))
    return fail(c, APC_ERR_ARG, "set_comp: null array");
  const int M = nc + nv, S = M + nb;
  for (int i = 0; i < S; ++i) {
    const int want = i < nc ? 0 : (i < M ? -1 : 1);
    if (type[i] != want)
      return fail(c, APC_ERR_ARG, "set_comp: node %d has Type %d, expected %d", i, type[i], want);
    if (!(cell_length[i] > 0.0))
      return fail(c, APC_ERR_ARG, "set_comp: node %d has CellLength %g", i, cell_length[i]);
  }
  std::vector<int> cnt(M), pad((size_t)M * 24, 0);
  for (int i = 0; i < M; ++i) {
    const int f = first_neigh[i], k = number_neigh[i];
    if (k < 0 || f < 0 || (long)f + k > nneigh)
      return fail(c, APC_ERR_ARG, "set_comp: node %d Neigh24 range [%d, %d+%d) outside 0..%d", i,
                  f, f, k, nneigh);
    if (k > 24)
      return fail(c, APC_ERR_UNSUPPORTED, "set_comp: node %d has %d Neigh24 entries (at most 24)",
                  i, k);
    cnt[i] = k;
    for (int q = 0; q < k; ++q) pad[(size_t)i * 24 + q] = neigh24[f + q];
  }
  std::vector<int> zero(S, 0);
  APC_ALLOC(c, c->cx, S); APC_ALLOC(c, c->cy, S); APC_ALLOC(c, c->cz, S); APC_ALLOC(c, c->ch, S);
  APC_ALLOC(c, c->cnp, S); APC_ALLOC(c, c->cfirst, S); APC_ALLOC(c, c->ctype, S);
  APC_ALLOC(c, c->n24, (size_t)M * 24);
  APC_ALLOC(c, c->n24cnt, M);
  APC_ALLOC(c, c->n24first, (size_t)M + 1);
  APC_HIP(c, hipMemcpyAsync(c->cx.p, x, 8ull * S, hipMemcpyHostToDevice, c->s));
  APC_HIP(c, hipMemcpyAsync(c->cy.p, y, 8ull * S, hipMemcpyHostToDevice, c->s));
  APC_HIP(c, hipMemcpyAsync(c->cz.p, z, 8ull * S, hipMemcpyHostToDevice, c->s));
  APC_HIP(c, hipMemcpyAsync(c->ch.p, cell_length, 8ull * S, hipMemcpyHostToDevice, c->s));
  APC_HIP(c, hipMemcpyAsync(c->cnp.p, np, 4ull * S, hipMemcpyHostToDevice, c->s));
  APC_HIP(c, hipMemcpyAsync(c->cfirst.p, zero.data(), 4ull * S, hipMemcpyHostToDevice, c->s));
  APC_HIP(c, hipMemcpyAsync(c->ctype.p, type, 4ull * S, hipMemcpyHostToDevice, c->s));
  APC_HIP(c, hipMemcpyAsync(c->n24.p, pad.data(), 4ull * M * 24, hipMemcpyHostToDevice, c->s));
  APC_HIP(c, hipMemcpyAsync(c->n24cnt.p, cnt.data(), 4ull * M, hipMemcpyHostToDevice, c->s));
  APC_HIP(c, hipStreamSynchronize(c->s));  // the host vectors go out of scope
  c->nc = nc;
  c->nv = nv;
  c->nb = nb;
  c->have_comp = true;
  c->comp_external = 1;
  c->have_tr = 0;
  c->have_nbr = c->have_sys = c->have_fields = 0;
  return 0;
}
