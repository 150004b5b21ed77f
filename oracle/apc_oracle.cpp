/*
 * TEST INFRASTRUCTURE ONLY — the oracle (see apc_oracle.h).
 *
 * CPU restatement of the reference's per-timestep solve-and-move path.  Every
 * function cites the reference lines it follows.  Built with g++ -O2
 * -ffp-contract=off (no FMA), matching an x86-64 build of the reference.
 * The serial / OpenMP split follows the reference, so this file is also the
 * CPU baseline that bench.py times.
 *
 * Deliberate differences from the reference:
 *  - sort by (key, index): one legal outcome of the reference's unstable
 *    std::sort on key alone (octree.cpp:148);
 *  - selection arrays are heap vectors that grow, so the nmax/-1 retry of
 *    mover.cpp:84-88 never triggers (the result equals the retried one);
 *  - the UB read level[-1] (octree.cpp:1926...) is treated as "no back link";
 *  - LAPACKE_dgels is replaced by a Householder QR least-squares solve and
 *    PETSc KSPGMRES+PCJACOBI by a restatement of PETSc's GMRES (classical
 *    Gram-Schmidt, no refinement: PETSc's default orthogonalisation);
 *  - fatal exit() paths return negative status codes.
 */
#include "apc_oracle.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

const double kPiRef = 3.1415926;       /* apcloud_macro.h:1 */
const double kMpcCm = 3.085677581e24;  /* apcloud_macro.h:2 */

/* ---------------------------------------------------------------- keys -- */
/* octree.cpp:676-852 — FP64 bisection of the bounding box, 3 bits per level,
 * x=bit2, y=bit1, z=bit0, ">= midpoint" selects the upper half.  A NaN
 * coordinate matches none of the eight cases and leaves the key unshifted. */
uint64_t key64(double x, double y, double z, const double *bb, int depth) {
  double lx = bb[0], rx = bb[1], ly = bb[2], ry = bb[3], lz = bb[4], rz = bb[5];
  uint64_t key = 0;
  for (int i = 1; i <= depth; ++i) {
    double mx = (lx + rx) / 2.0, my = (ly + ry) / 2.0, mz = (lz + rz) / 2.0;
    bool xl = x < mx, xh = x >= mx, yl = y < my, yh = y >= my, zl = z < mz, zh = z >= mz;
    if (!((xl || xh) && (yl || yh) && (zl || zh))) continue;
    key = (key << 3) + (uint64_t)((xh ? 4 : 0) + (yh ? 2 : 0) + (zh ? 1 : 0));
    if (xh) lx = mx; else rx = mx;
    if (yh) ly = my; else ry = my;
    if (zh) lz = mz; else rz = mz;
  }
  return key;
}

/* --------------------------------------------------------------- octree -- */
struct Tree {
  int N = 0, L = 0;
  double bb[6];
  std::vector<uint64_t> skey;
  std::vector<int> sidx;
  std::vector<int> numnodes, base;
  int len = 0;
  std::vector<uint64_t> nkey;
  std::vector<int> depth, first, count, fchild, nchild;
  std::vector<double> lox, loy, loz, hix, hiy, hiz;
};

/* octree.cpp:111-494 */
void build_tree(Tree &T, int N, const double *x, const double *y, const double *z,
                const double *bb, int L) {
  T.N = N;
  T.L = L;
  std::memcpy(T.bb, bb, sizeof(T.bb));
  std::vector<std::pair<uint64_t, int>> ki(N);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < N; ++i) ki[i] = {key64(x[i], y[i], z[i], bb, L), i}; /* :133-139 */
  std::sort(ki.begin(), ki.end()); /* :148, (key,index) order */
  T.skey.resize(N);
  T.sidx.resize(N);
  for (int i = 0; i < N; ++i) { T.skey[i] = ki[i].first; T.sidx[i] = ki[i].second; }

  std::vector<uint64_t> mask(L + 1);
  mask[L] = (((uint64_t)1) << (3 * L)) - 1; /* :158-171 */
  mask[0] = 0;
  for (int k = L - 1; k >= 1; --k) {
    int s = 3 * (L - k);
    mask[k] = (mask[L] >> s) << s;
  }
  T.numnodes.assign(L + 1, 0);
  T.numnodes[0] = 1; /* :177-191 */
  for (int k = 1; k <= L; ++k) {
    int cnt = 1;
#pragma omp parallel for schedule(static) reduction(+ : cnt)
    for (int i = 1; i < N; ++i)
      if ((T.skey[i] & mask[k]) != (T.skey[i - 1] & mask[k])) ++cnt;
    T.numnodes[k] = cnt;
  }
  T.base.assign(L + 1, 0);
  T.base[0] = 0; /* :199-204 */
  if (L >= 1) T.base[1] = 1;
  for (int k = 2; k <= L; ++k) T.base[k] = T.base[k - 1] + T.numnodes[k - 1];
  T.len = 0;
  for (int k = 0; k <= L; ++k) T.len += T.numnodes[k];
  int n = T.len;
  T.nkey.assign(n, 0);
  T.depth.assign(n, 0);
  T.first.assign(n, 0);
  T.count.assign(n, 0);
  T.fchild.assign(n, -1);
  T.nchild.assign(n, 0);
  T.lox.assign(n, 0); T.loy.assign(n, 0); T.loz.assign(n, 0);
  T.hix.assign(n, 0); T.hiy.assign(n, 0); T.hiz.assign(n, 0);

  /* deepest level, :232-267 */
  int b = T.base[L];
  T.nkey[b] = T.skey[0];
  T.depth[b] = L;
  T.first[b] = 0;
  T.nchild[b] = 0;
  T.fchild[b] = -1;
  int j = 1, fix = 0;
  for (int i = 1; i < N; ++i) {
    if (T.skey[i] != T.skey[i - 1]) {
      T.nkey[b + j] = T.skey[i];
      T.depth[b + j] = L;
      T.first[b + j] = i;
      T.nchild[b + j] = 0;
      T.fchild[b + j] = -1;
      T.count[b + j - 1] = i - fix;
      ++j;
      fix = i;
    }
  }
  T.count[b + j - 1] = N - fix;

  /* upper levels, :276-313 */
  for (int level = L - 1; level >= 0; --level) {
    int bl = T.base[level], bc = T.base[level + 1];
    j = 1;
    fix = 0;
    T.nkey[bl] = T.nkey[bc] & mask[level];
    T.depth[bl] = level;
    T.first[bl] = 0;
    T.fchild[bl] = bc;
    int sum = T.count[bc];
    T.count[bl] = sum;
    for (int i = 1; i < T.numnodes[level + 1]; ++i) {
      if ((T.nkey[bc + i] & mask[level]) != (T.nkey[bc + i - 1] & mask[level])) {
        T.nkey[bl + j] = T.nkey[bc + i] & mask[level];
        T.depth[bl + j] = level;
        T.first[bl + j] = T.first[bc + i];
        T.count[bl + j - 1] = sum;
        T.fchild[bl + j] = bc + i;
        T.nchild[bl + j - 1] = i - fix;
        ++j;
        fix = i;
        sum = T.count[bc + i];
      } else {
        sum += T.count[bc + i];
      }
    }
    T.nchild[bl + j - 1] = T.numnodes[level + 1] - fix;
    T.count[bl + j - 1] = sum;
  }

  /* per-node bounding boxes by halving, :331-437 */
  double xw = bb[1] - bb[0], yw = bb[3] - bb[2], zw = bb[5] - bb[4];
  T.lox[0] = bb[0]; T.loy[0] = bb[2]; T.loz[0] = bb[4];
  T.hix[0] = bb[1]; T.hiy[0] = bb[3]; T.hiz[0] = bb[5];
  for (int level = 0; level < L; ++level) {
    xw = 0.5 * xw; yw = 0.5 * yw; zw = 0.5 * zw;
    int bl = T.base[level];
#pragma omp parallel for schedule(static)
    for (int i = 0; i < T.numnodes[level]; ++i) {
      int p = bl + i;
      for (int c = T.fchild[p]; c < T.fchild[p] + T.nchild[p]; ++c) {
        unsigned t = (unsigned)((T.nkey[c] >> (3 * (L - level - 1))) & 0x7);
        if (t & 4) { T.lox[c] = T.lox[p] + xw; T.hix[c] = T.hix[p]; }
        else       { T.lox[c] = T.lox[p];      T.hix[c] = T.hix[p] - xw; }
        if (t & 2) { T.loy[c] = T.loy[p] + yw; T.hiy[c] = T.hiy[p]; }
        else       { T.loy[c] = T.loy[p];      T.hiy[c] = T.hiy[p] - yw; }
        if (t & 1) { T.loz[c] = T.loz[p] + zw; T.hiz[c] = T.hiz[p]; }
        else       { T.loz[c] = T.loz[p];      T.hiz[c] = T.hiz[p] - zw; }
      }
    }
  }
}

/* ----------------------------------------------------------- selection -- */
struct Comp {
  int nc = 0, nv = 0, nb = 0;
  std::vector<double> x, y, z, h;
  std::vector<int> np, first, type, fneigh, nneigh, neigh24;
  void clear() {
    nc = nv = nb = 0;
    x.clear(); y.clear(); z.clear(); h.clear();
    np.clear(); first.clear(); type.clear(); fneigh.clear(); nneigh.clear(); neigh24.clear();
  }
  int size() const { return nc + nv + nb; }
};

/* direction indices of the reference's neigh[6*i+d] (octree.cpp:1135-1141):
 * 0 x+, 1 x-, 2 y-, 3 y+, 4 z+, 5 z- */
inline int dir_axis(int d) { return d < 2 ? 0 : (d < 4 ? 1 : 2); }
inline bool dir_up(int d) { return d == 0 || d == 3 || d == 4; }
inline int dir_opp(int d) { static const int o[6] = {1, 0, 3, 2, 5, 4}; return o[d]; }
inline int dir_of(int axis, bool up) {
  static const int t[3][2] = {{1, 0}, {2, 3}, {5, 4}};
  return t[axis][up ? 1 : 0];
}
inline int axis_bit(int axis) { return 4 >> axis; } /* x=4, y=2, z=1 */

double pwrap(double x, double len, int bc) { /* octree.cpp:3034-3053 */
  if (bc == 1) {
    while (x >= (0.5 * len)) x -= len;
    while (x < -(0.5 * len)) x += len;
  }
  return x;
}

struct Selector {
  const Tree &T;
  const double *px, *py, *pz;
  int bc;
  double lenx, leny, lenz;
  std::vector<double> tx, ty, tz;
  std::vector<int> leaf, n_p, old, pidx, parent, level, child, neigh;
  double h0, hx0, hy0, hz0;
  int j = 0;

  Selector(const Tree &t, const double *x, const double *y, const double *z, int bc_)
      : T(t), px(x), py(y), pz(z), bc(bc_) {
    lenx = T.bb[1] - T.bb[0];
    leny = T.bb[3] - T.bb[2];
    lenz = T.bb[5] - T.bb[4];
  }
  void ensure(int n) {
    if ((int)tx.size() >= n) return;
    int m = std::max(n, (int)(tx.size() * 2) + 64);
    tx.resize(m); ty.resize(m); tz.resize(m);
    leaf.resize(m); n_p.resize(m); old.resize(m); pidx.resize(m); parent.resize(m);
    level.resize(m);
    child.resize(8 * (size_t)m, 0);
    neigh.resize(6 * (size_t)m, -1);
  }
  double dist2(int a, int b) const { /* octree.cpp:1807-1821 */
    double dx = pwrap(tx[a] - tx[b], lenx, bc);
    double dy = pwrap(ty[a] - ty[b], leny, bc);
    double dz = pwrap(tz[a] - tz[b], lenz, bc);
    return dx * dx + dy * dy + dz * dz;
  }
  /* descend from i through opened nodes coarser than lvl, always to the
   * child closest to node `from` (first minimum wins), octree.cpp:1804-1825 */
  int descend(int i, int from, int lvl) const {
    while (((i > 0) && (leaf[i] == 0)) && (level[i] < lvl)) {
      double mindis = dist2(from, child[8 * i]);
      int i3 = 0;
      for (int i4 = 0; i4 < 8; ++i4) {
        double dis = dist2(from, child[8 * i + i4]);
        if (dis < mindis) { mindis = dis; i3 = i4; }
      }
      i = child[8 * i + i3];
    }
    return i;
  }
  /* open node i: create its eight children at j..j+7.  Top-down mode follows
   * octree.cpp:1185-1779, ripple mode octree.cpp:1826-2666. */
  void open(int i, bool ripple) {
    int inode = old[i];
    leaf[i] = 0;
    int l = 0;
    if (inode >= 0) l = T.nchild[inode];
    double hx = hx0, hy = hy0, hz = hz0;
    for (int i1 = 0; i1 < level[i]; ++i1) { hx = 0.5 * hx; hy = 0.5 * hy; hz = 0.5 * hz; }
    int np8[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (!l) { /* :1199-1220, dead for nonempty cells with maxdepth <= depth */
      for (int i1 = pidx[i]; i1 < pidx[i] + n_p[i]; ++i1) {
        int q = T.sidx[i1];
        int c = (px[q] >= tx[i] ? 4 : 0) + (py[q] >= ty[i] ? 2 : 0) + (pz[q] >= tz[i] ? 1 : 0);
        np8[c]++;
      }
    }
    ensure(j + 8);
    int fpi = (i < T.len) ? T.first[i] : -1; /* reference indexes octree arrays with a temp index */
    int acc = 0;
    for (int c = 0; c < 8; ++c) {
      int k = j + c;
      int bx = (c >> 2) & 1, by = (c >> 1) & 1, bz = c & 1;
      int i2 = -1;
      if (inode >= 0) {
        for (int i1 = T.fchild[inode]; i1 < T.fchild[inode] + l; ++i1) {
          double tempx = (T.hix[i1] + T.lox[i1]) / 2;
          double tempy = (T.hiy[i1] + T.loy[i1]) / 2;
          double tempz = (T.hiz[i1] + T.loz[i1]) / 2;
          bool okx = bx ? (tempx > tx[i]) : (tempx < tx[i]);
          bool oky = by ? (tempy > ty[i]) : (tempy < ty[i]);
          bool okz = bz ? (tempz > tz[i]) : (tempz < tz[i]);
          if (okx && oky && okz) { i2 = i1; break; }
        }
      }
      if (i2 > 0) {
        old[k] = i2;
        pidx[k] = T.first[i2];
        n_p[k] = T.count[i2];
      } else {
        old[k] = -1;
        pidx[k] = (fpi >= 0) ? fpi + acc : -1;
        n_p[k] = np8[c];
      }
      acc += np8[c];
      parent[k] = i;
      level[k] = level[i] + 1;
      leaf[k] = -1;
      child[8 * i + c] = k;
      tx[k] = bx ? tx[i] + 0.25 * hx : tx[i] - 0.25 * hx;
      ty[k] = by ? ty[i] + 0.25 * hy : ty[i] - 0.25 * hy;
      tz[k] = bz ? tz[i] + 0.25 * hz : tz[i] - 0.25 * hz;
      /* siblings */
      for (int ax = 0; ax < 3; ++ax) {
        int bit = axis_bit(ax);
        bool up = (c & bit) != 0;
        if (!up) neigh[6 * k + dir_of(ax, true)] = j + (c | bit);
        else neigh[6 * k + dir_of(ax, false)] = j + (c & ~bit);
      }
      /* outward faces, axis order x, y, z */
      for (int ax = 0; ax < 3; ++ax) {
        int bit = axis_bit(ax);
        bool up = (c & bit) != 0;
        int d = dir_of(ax, up);
        int cc = c ^ bit; /* the neighbour's child facing us */
        int i1 = neigh[6 * i + d];
        if (!ripple) {
          if ((i1 >= 0) && (leaf[i1] == 0)) {
            i1 = child[8 * i1 + cc];
            neigh[6 * k + d] = i1;
            neigh[6 * i1 + dir_opp(d)] = k;
          } else {
            neigh[6 * k + d] = i1;
          }
        } else {
          i1 = descend(i1, k, level[k]);
          neigh[6 * k + d] = i1;
          if (i1 >= 0 && level[i1] == level[k]) neigh[6 * i1 + dir_opp(d)] = k;
        }
      }
    }
    j += 8;
  }

  int leaf_capacity = 0;

  int run(Comp &out, double c, int order, int mindepth, int maxdepth, bool faces) {
    ensure(1024);
    int i = 0;
    j = 1;
    int inode = 0;
    tx[0] = (T.hix[inode] + T.lox[inode]) / 2; /* :1116-1158 */
    ty[0] = (T.hiy[inode] + T.loy[inode]) / 2;
    tz[0] = (T.hiz[inode] + T.loz[inode]) / 2;
    n_p[0] = T.count[inode];
    old[0] = 0;
    pidx[0] = T.first[inode];
    parent[0] = -1;
    level[0] = T.depth[inode];
    h0 = (T.hix[0] - T.lox[0]) * (T.hiy[0] - T.loy[0]) * (T.hiz[0] - T.loz[0]);
    h0 = std::pow(h0, 1.0 / 3);
    hx0 = T.hix[0] - T.lox[0];
    hy0 = T.hiy[0] - T.loy[0];
    hz0 = T.hiz[0] - T.loz[0];
    for (int d = 0; d < 6; ++d) neigh[d] = (bc == 1) ? 0 : -1;
    if (bc == 1)
      for (int q = 0; q < 8; ++q) child[q] = q + 1;

    /* top-down, :1161-1790 */
    while ((j - i) > 0) {
      double temp = 1.0 * n_p[i];
      double h = h0;
      for (int i1 = 0; i1 < level[i]; ++i1) h = 0.5 * h;
      for (int i1 = 0; i1 < (2 * order - 2); ++i1) temp = temp * h;
      bool crit = (leaf_capacity > 0) ? (n_p[i] > leaf_capacity) : (temp > c);
      if ((crit || (level[i] < mindepth)) && (level[i] < maxdepth)) {
        open(i, false);
      } else {
        leaf[i] = 1;
      }
      ++i;
    }

    /* bottom-up 2:1 ripple, :1793-2718 */
    int ml = level[i - 1];
    for (int ln = ml; ln > 1; --ln) {
      for (int i0 = 0; j > i0; ++i0) {
        if (level[i0] != ln) continue;
        for (int di = 0; di < 6; ++di) {
          int q = descend(neigh[6 * i0 + di], i0, ln);
          neigh[6 * i0 + di] = q;
          while ((q > 0) && (level[q] < (ln - 1))) {
            open(q, true);
            q = descend(neigh[6 * i0 + di], i0, ln);
            neigh[6 * i0 + di] = q;
          }
        }
      }
      for (int i0 = 0; j > i0; ++i0) {
        if (level[i0] != ln) continue;
        for (int di = 0; di < 6; ++di) neigh[6 * i0 + di] = descend(neigh[6 * i0 + di], i0, ln);
      }
    }

    /* output, :2720-3000 */
    out.clear();
    auto cell = [&](int q) {
      double t = h0;
      for (int i1 = 0; i1 < level[q]; ++i1) t = 0.5 * t;
      return t;
    };
    for (int q = 0; q < j; ++q)
      if (leaf[q] == 1) {
        out.x.push_back(tx[q]); out.y.push_back(ty[q]); out.z.push_back(tz[q]);
        out.h.push_back(cell(q));
        out.np.push_back(n_p[q]); out.first.push_back(pidx[q]); out.type.push_back(0);
        pidx[q] = out.nc;
        out.nc++;
      }
    for (int q = 0; q < j; ++q)
      if (leaf[q] == -1) {
        out.x.push_back(tx[q]); out.y.push_back(ty[q]); out.z.push_back(tz[q]);
        out.h.push_back(cell(q));
        out.np.push_back(n_p[q]); out.first.push_back(pidx[q]); out.type.push_back(-1);
        pidx[q] = out.nc + out.nv;
        out.nv++;
      }
    /* Neigh24 in the order x+, x-, y+, y-, z+, z-; the facing children of a
     * finer neighbour in fixed order (:2770-2970) */
    static const int dorder[6] = {0, 1, 3, 2, 4, 5};
    static const int facing[6][4] = {{0, 1, 2, 3}, {4, 5, 6, 7}, {2, 3, 6, 7},
                                     {0, 1, 4, 5}, {0, 2, 4, 6}, {1, 3, 5, 7}};
    int total = 0;
    for (int q = 0; q < j; ++q) {
      if (!leaf[q]) continue;
      out.fneigh.push_back(total);
      int i2 = 0;
      for (int e = 0; e < 6; ++e) {
        int d = dorder[e];
        int i3 = neigh[6 * q + d];
        if (i3 == -1) {
          double bxp, byp, bzp;
          int ax = dir_axis(d);
          bool up = dir_up(d);
          double fx = faces ? (up ? T.bb[1] : T.bb[0]) : (up ? 1.0 : -1.0);
          double fy = faces ? (up ? T.bb[3] : T.bb[2]) : (up ? 1.0 : -1.0);
          double fz = faces ? (up ? T.bb[5] : T.bb[4]) : (up ? 1.0 : -1.0);
          bxp = (ax == 0) ? fx : tx[q];
          byp = (ax == 1) ? fy : ty[q];
          bzp = (ax == 2) ? fz : tz[q];
          out.x.push_back(bxp); out.y.push_back(byp); out.z.push_back(bzp);
          out.h.push_back(cell(q));
          out.np.push_back(0);
          out.first.push_back(-1);
          out.neigh24.push_back(out.nc + out.nv + out.nb);
          i2 += 1;
          out.nb++;
        } else if (leaf[i3]) {
          out.neigh24.push_back(pidx[i3]);
          i2 += 1;
        } else {
          for (int f = 0; f < 4; ++f) out.neigh24.push_back(pidx[child[8 * i3 + facing[d][f]]]);
          i2 += 4;
        }
      }
      total += i2;
      out.nneigh.push_back(i2);
    }
    for (int q = out.nc + out.nv; q < out.nc + out.nv + out.nb; ++q) {
      out.nneigh.push_back(0);
      out.fneigh.push_back(total);
      out.type.push_back(1);
    }
    return out.size();
  }
};

/* --------------------------------------------------- neighbour search -- */
/* solver.cpp:600-632 — unstable quicksort with Hoare-style partition */
int apartition(double *a, int *b, int l, int r) {
  double pivot = a[l], t;
  int i = l, j = r + 1, ti;
  while (1) {
    do ++i; while (i <= r && a[i] <= pivot);
    do --j; while (a[j] > pivot);
    if (i >= j) break;
    t = a[i]; a[i] = a[j]; a[j] = t;
    ti = b[i]; b[i] = b[j]; b[j] = ti;
  }
  t = a[l]; a[l] = a[j]; a[j] = t;
  ti = b[l]; b[l] = b[j]; b[j] = ti;
  return j;
}
void aquicksort(double *a, int *b, int l, int r) {
  if (l < r) {
    int j = apartition(a, b, l, r);
    aquicksort(a, b, l, j - 1);
    aquicksort(a, b, j + 1, r);
  }
}

struct Frame { double A[9]; double len[3]; int bc; };

Frame make_frame(const double *bb, int bc) { /* solver.cpp:464-475 */
  Frame f;
  double alpha = 0.01, beta = 0.001, gamma = 0.0001;
  f.A[0] = cos(alpha) * cos(gamma) - sin(alpha) * sin(beta) * sin(gamma);
  f.A[1] = -sin(alpha) * cos(beta);
  f.A[2] = -cos(alpha) * sin(gamma) - sin(alpha) * sin(beta) * cos(gamma);
  f.A[3] = cos(alpha) * sin(beta) * sin(gamma) + sin(alpha) * cos(gamma);
  f.A[4] = cos(alpha) * cos(beta);
  f.A[5] = cos(alpha) * sin(beta) * cos(gamma) - sin(alpha) * sin(gamma);
  f.A[6] = cos(beta) * sin(gamma);
  f.A[7] = -sin(beta);
  f.A[8] = cos(beta) * cos(gamma);
  f.len[0] = bb[1] - bb[0];
  f.len[1] = bb[3] - bb[2];
  f.len[2] = bb[5] - bb[4];
  f.bc = bc;
  return f;
}

/* solver.cpp:456-570.  Returns 0, or -2 ("too many neighbours in an
 * octant", :535-539), -3 ("not enough neighbours", :544-550), -4 ("invalid
 * neighbour", :511-518).  Writes lm neighbours and true distances. */
int search_octant(const Comp &cp, const Frame &F, int lm, int i0, std::vector<int> &neigh,
                  std::vector<double> &distance) {
  const int nq = 10 * lm;
  std::vector<int> tn(8 * (size_t)nq, 0);
  std::vector<double> td(8 * (size_t)nq, 0.0);
  int nn[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  const double *A = F.A;
  auto w0 = [&](double v) { return pwrap(v, F.len[0], F.bc); };
  auto w1 = [&](double v) { return pwrap(v, F.len[1], F.bc); };
  auto w2 = [&](double v) { return pwrap(v, F.len[2], F.bc); };
  auto sqd = [&](double dx, double dy, double dz) { /* :595-598 */
    return w0(dx) * w0(dx) + w1(dy) * w1(dy) + w2(dz) * w2(dz);
  };
  auto octant = [&](int m) {
    double ex = w0(cp.x[m] - cp.x[i0]), ey = w1(cp.y[m] - cp.y[i0]), ez = w2(cp.z[m] - cp.z[i0]);
    double xt = A[0] * ex + A[1] * ey + A[2] * ez;
    double yt = A[3] * ex + A[4] * ey + A[5] * ez;
    double zt = A[6] * ex + A[7] * ey + A[8] * ez;
    return (xt > 0) + 2 * (yt > 0) + 4 * (zt > 0);
  };
  const int total = cp.size();
  for (int q = cp.fneigh[i0]; q < cp.fneigh[i0] + cp.nneigh[i0]; ++q) {
    int m = cp.neigh24[q];
    /* the reference does not check the node's own entries (:481-492) and
     * would read past its arrays; a caller-filled set can hold such entries */
    if (m < 0 || m >= total) return -4;
    int oct = octant(m);
    if (nn[oct] >= nq) return -2;
    tn[oct * nq + nn[oct]] = m;
    td[oct * nq + nn[oct]] = sqd(cp.x[m] - cp.x[i0], cp.y[m] - cp.y[i0], cp.z[m] - cp.z[i0]);
    nn[oct]++;
  }
  int ite = 0;
  int l;
  std::vector<int> snap;
  for (l = nn[0] + nn[1] + nn[2] + nn[3] + nn[4] + nn[5] + nn[6] + nn[7]; l < lm;
       l = nn[0] + nn[1] + nn[2] + nn[3] + nn[4] + nn[5] + nn[6] + nn[7]) {
    snap.clear();
    for (int o = 0; o < 8; ++o)
      for (int q = 0; q < nn[o]; ++q) snap.push_back(tn[o * nq + q]);
    for (int a = 0; a < l; ++a) {
      int src = snap[a];
      for (int q = cp.fneigh[src]; q < cp.fneigh[src] + cp.nneigh[src]; ++q) {
        int m = cp.neigh24[q];
        if (m > total || m < 0) return -4; /* :511-518 (m == total: unchecked there, UB) */
        if (m == total) return -4;
        int oct = octant(m);
        int flag = 0;
        for (int k = 0; k < nn[oct]; ++k)
          if (m == tn[oct * nq + k]) flag = 1;
        if ((flag == 0) && (m - i0)) {
          if (nn[oct] >= nq) return -2;
          tn[oct * nq + nn[oct]] = m;
          td[oct * nq + nn[oct]] = sqd(cp.x[m] - cp.x[i0], cp.y[m] - cp.y[i0], cp.z[m] - cp.z[i0]);
          nn[oct]++;
        }
      }
    }
    ite++;
    if (ite > 5) return -3;
  }
  if ((int)neigh.size() < l) { neigh.resize(l); distance.resize(l); }
  int k = 0;
  for (int o = 0; o < 8; ++o) {
    double penalty = 1.0;
    /* the reference sorts td[k..k+nn[o]-1] but reads octant o's own slots
     * (:557-561): only octant 0 is sorted while l < nq */
    aquicksort(td.data(), tn.data(), k, k + nn[o] - 1);
    for (int q = 0; q < nn[o]; ++q) {
      neigh[k] = tn[o * nq + q];
      distance[k] = penalty * td[o * nq + q];
      penalty = penalty * 10000.0;
      k = k + 1;
    }
  }
  aquicksort(distance.data(), neigh.data(), 0, l - 1);
  for (int q = 0; q < lm; ++q) {
    int m = neigh[q];
    distance[q] = sqrt(sqd(cp.x[m] - cp.x[i0], cp.y[m] - cp.y[i0], cp.z[m] - cp.z[i0]));
  }
  return 0;
}

/* -------------------------------------------------------- least squares -- */
/* Replacement for LAPACKE_dgels(ROW_MAJOR,'N',m,n,nrhs,B,n,b,nrhs) with m>=n
 * (solver.cpp:106,342): Householder QR (LAPACK dlarfg convention), Q^T b,
 * back substitution.  Returns 0 or the 1-based index of a zero diagonal of R
 * (dgels' info>0 rank-deficiency report). */
int wlsq(int m, int n, int nrhs, double *B, double *b) {
  std::vector<double> v(m);
  for (int k = 0; k < n; ++k) {
    double alpha = B[k * n + k];
    double xn = 0.0;
    for (int r = k + 1; r < m; ++r) xn += B[r * n + k] * B[r * n + k];
    xn = sqrt(xn);
    if (xn == 0.0) continue; /* tau = 0, H = I */
    double beta = -copysign(sqrt(alpha * alpha + xn * xn), alpha);
    double tau = (beta - alpha) / beta;
    double scal = 1.0 / (alpha - beta);
    v[k] = 1.0;
    for (int r = k + 1; r < m; ++r) v[r] = B[r * n + k] * scal;
    B[k * n + k] = beta;
    for (int r = k + 1; r < m; ++r) B[r * n + k] = 0.0;
    for (int c = k + 1; c < n; ++c) {
      double s = 0.0;
      for (int r = k; r < m; ++r) s += v[r] * B[r * n + c];
      s *= tau;
      for (int r = k; r < m; ++r) B[r * n + c] -= s * v[r];
    }
    for (int c = 0; c < nrhs; ++c) {
      double s = 0.0;
      for (int r = k; r < m; ++r) s += v[r] * b[r * nrhs + c];
      s *= tau;
      for (int r = k; r < m; ++r) b[r * nrhs + c] -= s * v[r];
    }
  }
  for (int k = 0; k < n; ++k)
    if (B[k * n + k] == 0.0) return k + 1;
  for (int c = 0; c < nrhs; ++c)
    for (int k = n - 1; k >= 0; --k) {
      double s = b[k * nrhs + c];
      for (int q = k + 1; q < n; ++q) s -= B[k * n + q] * b[q * nrhs + c];
      b[k * nrhs + c] = s / B[k * n + k];
    }
  return 0;
}

/* ---------------------------------------------------------------- GMRES -- */
/* Row-wise sparse matrix with INSERT_VALUES semantics (last write wins). */
struct SpMat {
  int n = 0;
  std::vector<std::vector<std::pair<int, double>>> rows;
  void init(int n_) { n = n_; rows.assign(n, {}); }
  void set(int i, int j, double v) {
    for (auto &e : rows[i])
      if (e.first == j) { e.second = v; return; }
    rows[i].push_back({j, v});
  }
};

struct Csr {
  int n = 0;
  std::vector<int> rp, col;
  std::vector<double> val, dinv;
};

Csr to_csr(const SpMat &A) {
  Csr C;
  C.n = A.n;
  C.rp.assign(A.n + 1, 0);
  for (int i = 0; i < A.n; ++i) C.rp[i + 1] = C.rp[i] + (int)A.rows[i].size();
  C.col.resize(C.rp[A.n]);
  C.val.resize(C.rp[A.n]);
  C.dinv.assign(A.n, 1.0);
  for (int i = 0; i < A.n; ++i) {
    int p = C.rp[i];
    double d = 0.0;
    for (auto &e : A.rows[i]) {
      C.col[p] = e.first;
      C.val[p] = e.second;
      if (e.first == i) d = e.second;
      ++p;
    }
    C.dinv[i] = (d == 0.0) ? 1.0 : 1.0 / d; /* PCJACOBI: zero diagonal -> 1 */
  }
  return C;
}

/* deterministic OpenMP helpers (static chunks, partials combined in order) */
double pdot(int n, const double *a, const double *b) {
#ifdef _OPENMP
  int nt = omp_get_max_threads();
#else
  int nt = 1;
#endif
  std::vector<double> part(nt, 0.0);
#pragma omp parallel num_threads(nt)
  {
#ifdef _OPENMP
    int t = omp_get_thread_num(), T = omp_get_num_threads();
#else
    int t = 0, T = 1;
#endif
    long lo = (long)n * t / T, hi = (long)n * (t + 1) / T;
    double s = 0.0;
    for (long i = lo; i < hi; ++i) s += a[i] * b[i];
    part[t] = s;
  }
  double s = 0.0;
  for (double v : part) s += v;
  return s;
}

void papply(const Csr &A, const double *x, double *y) { /* y = D^-1 A x */
#pragma omp parallel for schedule(static)
  for (int i = 0; i < A.n; ++i) {
    double s = 0.0;
    for (int p = A.rp[i]; p < A.rp[i + 1]; ++p) s += A.val[p] * x[A.col[p]];
    y[i] = A.dinv[i] * s;
  }
}

enum { KSP_CONVERGED_RTOL = 2, KSP_CONVERGED_ATOL = 3, KSP_DIVERGED_ITS = -3,
       KSP_DIVERGED_DTOL = -4, KSP_DIVERGED_BREAKDOWN = -5, KSP_DIVERGED_NANORINF = -9,
       KSP_DIVERGED_NULL = -2 };

/* Restatement of PETSc's KSPSolve_GMRES / KSPGMRESCycle /
 * KSPGMRESUpdateHessenberg / KSPGMRESBuildSoln with left PCJACOBI, zero
 * initial guess, KSPConvergedDefault and classical Gram-Schmidt (PETSc's
 * default orthogonalisation, no refinement). */
int gmres(const Csr &A, const double *b, double *x, double rtol, double atol, double dtol,
          int max_it, int restart, int *its_out, double *rnorm_out) {
  const int n = A.n;
  const int m = restart;
  std::vector<double> V((size_t)(m + 1) * n), w(n), HH((size_t)(m + 1) * m, 0.0);
  std::vector<double> cc(m), ss(m), grs(m + 1), nrs(m);
  for (int i = 0; i < n; ++i) x[i] = 0.0;
  int its = 0, reason = 0;
  double rnorm0 = 0.0, ttol = 0.0, res = 0.0;
  bool guess_zero = true;
  int itcount = 0;
  auto converged = [&](int nit, double rn) {
    if (nit == 0) {
      rnorm0 = rn;
      ttol = std::max(rtol * rnorm0, atol);
    }
    if (std::isnan(rn) || std::isinf(rn)) return (int)KSP_DIVERGED_NANORINF;
    if (rn <= ttol) return (rn < atol) ? (int)KSP_CONVERGED_ATOL : (int)KSP_CONVERGED_RTOL;
    if (rn >= dtol * rnorm0) return (int)KSP_DIVERGED_DTOL;
    return 0;
  };
  while (!reason) {
    double *v0 = V.data();
    if (guess_zero) { /* KSPInitialResidual: B b */
#pragma omp parallel for schedule(static)
      for (int i = 0; i < n; ++i) v0[i] = A.dinv[i] * b[i];
    } else {
#pragma omp parallel for schedule(static)
      for (int i = 0; i < n; ++i) {
        double s = 0.0;
        for (int p = A.rp[i]; p < A.rp[i + 1]; ++p) s += A.val[p] * x[A.col[p]];
        v0[i] = A.dinv[i] * (b[i] - s);
      }
    }
    double res_norm = sqrt(pdot(n, v0, v0));
    res = res_norm;
    grs[0] = res_norm;
    int it = 0;
    if (res == 0.0) { reason = KSP_CONVERGED_ATOL; break; }
    for (int i = 0; i < n; ++i) v0[i] *= 1.0 / res_norm;
    reason = converged(its, res);
    while (!reason && it < m && its < max_it) {
      double *vit = V.data() + (size_t)it * n;
      double *vn = V.data() + (size_t)(it + 1) * n;
      papply(A, vit, vn);
      /* classical Gram-Schmidt */
      for (int q = 0; q <= it; ++q) HH[(size_t)it * (m + 1) + q] = pdot(n, vn, V.data() + (size_t)q * n);
#pragma omp parallel for schedule(static)
      for (int i = 0; i < n; ++i) {
        double s = vn[i];
        for (int q = 0; q <= it; ++q) s -= HH[(size_t)it * (m + 1) + q] * V[(size_t)q * n + i];
        vn[i] = s;
      }
      double tt = sqrt(pdot(n, vn, vn));
      if (tt != 0.0) {
        double inv = 1.0 / tt;
        for (int i = 0; i < n; ++i) vn[i] *= inv;
      }
      double *hh = &HH[(size_t)it * (m + 1)];
      hh[it + 1] = tt;
      double hapbnd = fabs(tt / grs[it]);
      if (hapbnd > 1.0e-30) hapbnd = 1.0e-30;
      bool hapend = tt < hapbnd;
      /* KSPGMRESUpdateHessenberg */
      for (int q = 1; q <= it; ++q) {
        double t0 = hh[q - 1];
        hh[q - 1] = cc[q - 1] * t0 + ss[q - 1] * hh[q];
        hh[q] = cc[q - 1] * hh[q] - ss[q - 1] * t0;
      }
      if (!hapend) {
        double t1 = sqrt(hh[it] * hh[it] + hh[it + 1] * hh[it + 1]);
        if (t1 == 0.0) { reason = KSP_DIVERGED_NULL; break; }
        cc[it] = hh[it] / t1;
        ss[it] = hh[it + 1] / t1;
        grs[it + 1] = -(ss[it] * grs[it]);
        grs[it] = cc[it] * grs[it];
        hh[it] = cc[it] * hh[it] + ss[it] * hh[it + 1];
        res = fabs(grs[it + 1]);
      } else {
        res = 0.0;
      }
      it++;
      its++;
      reason = converged(its, res);
      if (hapend && !reason) { reason = KSP_DIVERGED_BREAKDOWN; }
    }
    /* KSPGMRESBuildSoln(it-1) */
    int top = it - 1;
    if (top >= 0) {
      auto H = [&](int r, int c) { return HH[(size_t)c * (m + 1) + r]; };
      bool ok = true;
      if (H(top, top) != 0.0) nrs[top] = grs[top] / H(top, top);
      else ok = false;
      for (int ii = 1; ok && ii <= top; ++ii) {
        int k = top - ii;
        double tt = grs[k];
        for (int q = k + 1; q <= top; ++q) tt = tt - H(k, q) * nrs[q];
        if (H(k, k) == 0.0) { ok = false; break; }
        nrs[k] = tt / H(k, k);
      }
      if (!ok) { reason = KSP_DIVERGED_BREAKDOWN; break; }
#pragma omp parallel for schedule(static)
      for (int i = 0; i < n; ++i) {
        double s = 0.0;
        for (int q = 0; q <= top; ++q) s += nrs[q] * V[(size_t)q * n + i];
        x[i] += s;
      }
    }
    itcount += it;
    if (itcount >= max_it) {
      if (!reason) reason = KSP_DIVERGED_ITS;
      break;
    }
    guess_zero = false;
  }
  if (its_out) *its_out = its;
  if (rnorm_out) *rnorm_out = res;
  return reason;
}

/* --------------------------------------------------------------- solver -- */
struct Solver {
  /* solver.cpp:8-22 defaults */
  int lm = 17, ln = 9, dirichlet_point = 1;
  double c_dens = 1.0, c_pot = 1.0, c_field = 1.0;
};

double bv_value(const OrcParams &p, double x, double y, double z) {
  if (p.bv_kind != 1) return 0.0;
  double dx = x - p.bv_center[0], dy = y - p.bv_center[1], dz = z - p.bv_center[2];
  double r = sqrt(dx * dx + dy * dy + dz * dz), R = p.bv_radius, GM = p.bv_G * p.bv_mass;
  if (r < R) return -GM * (3.0 * R * R - r * r) / (2.0 * R * R * R);
  return -GM / r;
}

}  // namespace

/* ------------------------------------------------------------ the sim -- */
struct OrcSim {
  OrcParams p;
  int N = 0;
  std::vector<double> x, y, z, vx, vy, vz, ax, ay, az;
  Tree T;
  bool have_tree = false;
  Comp cp;
  bool have_comp = false;
  Frame F;
  /* solver state */
  Solver sv;
  std::vector<int> nbr;       /* M*lm */
  std::vector<double> dist;   /* M*lm */
  SpMat A1, A2;
  std::vector<double> rhs1, pbnd, bv, rho, phi, g, H;
  OrcSolveInfo info{};
};

namespace {

int neighbors_all(OrcSim *s) {
  const int M = s->cp.nc + s->cp.nv, lm = s->p.lm;
  s->nbr.assign((size_t)M * lm, 0);
  s->dist.assign((size_t)M * lm, 0.0);
  int err = 0;
  std::vector<int> nb(16 * lm);
  std::vector<double> ds(16 * lm);
  for (int i = 0; i < M; ++i) { /* serial like solver.cpp:73 */
    int r = search_octant(s->cp, s->F, lm, i, nb, ds);
    if (r) { err = r; break; }
    for (int q = 0; q < lm; ++q) {
      s->nbr[(size_t)i * lm + q] = nb[q];
      s->dist[(size_t)i * lm + q] = ds[q];
    }
  }
  return err;
}

/* solver.cpp:24-180 (rank-0 part) */
int build_system(OrcSim *s) {
  Comp &cp = s->cp;
  const int M = cp.nc + cp.nv, lm = s->p.lm, ln = 9, S = cp.size();
  int r = neighbors_all(s);
  if (r) return r;
  s->A1.init(M);
  s->A2.init(M);
  s->rhs1.assign(M, 0.0);
  s->pbnd.assign(M, 0.0);
  s->bv.assign(cp.nb, 0.0);
  for (int q = 0; q < cp.nb; ++q) {
    int k = M + q;
    s->bv[q] = bv_value(s->p, cp.x[k], cp.y[k], cp.z[k]);
  }
  double gamma = 0.7;
  std::vector<double> B((size_t)lm * ln), bm((size_t)lm * lm);
  for (int i = 0; i < M; ++i) {
    const int *neigh = &s->nbr[(size_t)i * lm];
    const double *distance = &s->dist[(size_t)i * lm];
    double maxdis = 0;
    for (int j = 0; j < lm; ++j)
      if (distance[j] > maxdis) maxdis = distance[j];
    maxdis = maxdis * gamma;
    for (int j = 0; j < lm; ++j) {
      double weight = (exp(-distance[j] * distance[j] / maxdis / maxdis) - exp(-4)) / (1 - exp(-4));
      double p1 = pwrap(cp.x[neigh[j]] - cp.x[i], s->F.len[0], s->p.bc) / cp.h[i];
      double p2 = pwrap(cp.y[neigh[j]] - cp.y[i], s->F.len[1], s->p.bc) / cp.h[i];
      double p3 = pwrap(cp.z[neigh[j]] - cp.z[i], s->F.len[2], s->p.bc) / cp.h[i];
      double *Bj = &B[(size_t)j * ln];
      Bj[0] = p1; Bj[1] = p2; Bj[2] = p3;
      Bj[3] = p1 * p1 / 2.0; Bj[4] = p2 * p2 / 2.0; Bj[5] = p3 * p3 / 2.0;
      Bj[6] = p1 * p2; Bj[7] = p1 * p3; Bj[8] = p2 * p3;
      for (int k = 0; k < ln; ++k) Bj[k] = Bj[k] * weight;
      for (int k = 0; k < lm; ++k) bm[(size_t)j * lm + k] = 0.0;
      bm[(size_t)j * lm + j] = weight;
    }
    int info = wlsq(lm, ln, lm, B.data(), bm.data());
    if (info > 0) return -10; /* rank-deficient stencil, solver.cpp:107-115 */
    double p2 = 0.0;
    for (int j = 0; j < lm; ++j) {
      double p1 = (bm[3 * lm + j] + bm[4 * lm + j] + bm[5 * lm + j]) / 24.0;
      p2 = p2 - p1;
      if (neigh[j] < M) s->A1.set(i, neigh[j], p1);
    }
    p2 = p2 + 1;
    s->A1.set(i, i, p2);
    s->rhs1[i] = 1.0 * cp.np[i] / cp.h[i] / cp.h[i] / cp.h[i] * s->sv.c_dens;
    double p3 = 0.0;
    if ((i == 0) && (s->p.bc == 1) && s->sv.dirichlet_point) {
      s->A2.set(i, i, 1.0);
    } else {
      p2 = 0.0;
      for (int j = 0; j < lm; ++j) {
        double p1 = (bm[3 * lm + j] + bm[4 * lm + j] + bm[5 * lm + j]) / cp.h[i] / cp.h[i];
        p2 = p2 - p1;
        if (cp.type[neigh[j]] < 1) {
          s->A2.set(i, neigh[j], p1);
        } else {
          int k = neigh[j] - M;
          p3 = p3 - s->bv[k] * p1;
        }
      }
      s->A2.set(i, i, p2);
    }
    s->pbnd[i] = p3;
  }
  (void)S;
  return 0;
}

/* solver.cpp:295-362 */
int differentiate(OrcSim *s) {
  Comp &cp = s->cp;
  const int M = cp.nc + cp.nv, lm = s->p.lm, ln = 9;
  s->g.assign(3 * (size_t)M, 0.0);
  s->H.assign(6 * (size_t)M, 0.0);
  double gamma = 0.7;
  std::vector<double> B((size_t)lm * ln), bb(lm);
  for (int i = 0; i < M; ++i) {
    const int *neigh = &s->nbr[(size_t)i * lm];
    const double *distance = &s->dist[(size_t)i * lm];
    double maxdis = 0;
    for (int j = 0; j < lm; ++j)
      if (distance[j] > maxdis) maxdis = distance[j];
    maxdis = maxdis * gamma;
    double p4 = s->phi[i];
    for (int j = 0; j < lm; ++j) {
      double weight = (exp(-distance[j] * distance[j] / maxdis / maxdis) - exp(-4)) / (1 - exp(-4));
      double p1 = pwrap(cp.x[neigh[j]] - cp.x[i], s->F.len[0], s->p.bc) / cp.h[i];
      double p2 = pwrap(cp.y[neigh[j]] - cp.y[i], s->F.len[1], s->p.bc) / cp.h[i];
      double p3 = pwrap(cp.z[neigh[j]] - cp.z[i], s->F.len[2], s->p.bc) / cp.h[i];
      double *Bj = &B[(size_t)j * ln];
      Bj[0] = p1; Bj[1] = p2; Bj[2] = p3;
      Bj[3] = p1 * p1 / 2.0; Bj[4] = p2 * p2 / 2.0; Bj[5] = p3 * p3 / 2.0;
      Bj[6] = p1 * p2; Bj[7] = p1 * p3; Bj[8] = p2 * p3;
      for (int k = 0; k < ln; ++k) Bj[k] = Bj[k] * weight;
      int k = neigh[j];
      double v = (k < M) ? s->phi[k] : s->bv[k - M];
      bb[j] = weight * (v - p4) * s->sv.c_field;
    }
    int info = wlsq(lm, ln, 1, B.data(), bb.data());
    if (info > 0) return -10;
    double h = cp.h[i];
    s->g[3 * (size_t)i + 0] = bb[0] / h;
    s->g[3 * (size_t)i + 1] = bb[1] / h;
    s->g[3 * (size_t)i + 2] = bb[2] / h;
    for (int q = 0; q < 6; ++q) s->H[6 * (size_t)i + q] = bb[3 + q] / h / h;
  }
  return 0;
}

/* solver.cpp:382-454: build, rho, phi, gradient */
int solve_all(OrcSim *s) {
  Comp &cp = s->cp;
  const int M = cp.nc + cp.nv;
  int r = build_system(s);
  if (r) return r;
  Csr C1 = to_csr(s->A1), C2 = to_csr(s->A2);
  s->rho.assign(M, 0.0);
  s->phi.assign(M, 0.0);
  const OrcParams &p = s->p;
  s->info.reason_rho = gmres(C1, s->rhs1.data(), s->rho.data(), p.rtol, p.atol, p.dtol, p.max_it,
                             p.restart, &s->info.its_rho, &s->info.rnorm_rho);
  /* solver.cpp:233-261; with bv=0 (the reference) rho*C + p_x equals (rho+p_x)*C */
  std::vector<double> rhs2(M);
  if (p.bc == 0) {
    for (int i = 0; i < M; ++i) rhs2[i] = s->rho[i] * s->sv.c_pot + s->pbnd[i];
  } else {
    double p2 = 0, p3 = 0;
    for (int i = 0; i < M; ++i) {
      double h3 = cp.h[i] * cp.h[i] * cp.h[i];
      p2 = p2 + s->rho[i] * h3;
      p3 = p3 + h3;
    }
    p2 = p2 / p3;
    for (int i = 0; i < M; ++i) rhs2[i] = (s->rho[i] - p2) * s->sv.c_pot;
  }
  s->info.reason_phi = gmres(C2, rhs2.data(), s->phi.data(), p.rtol, p.atol, p.dtol, p.max_it,
                             p.restart, &s->info.its_phi, &s->info.rnorm_phi);
  return differentiate(s);
}

/* octree.cpp:3003-3031 */
void interpolate(OrcSim *s) {
  Comp &cp = s->cp;
  const int M = cp.nc + cp.nv;
  const Tree &T = s->T;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < M; ++i) {
    int j = cp.first[i];
    const double *g = &s->g[3 * (size_t)i];
    const double *H = &s->H[6 * (size_t)i]; /* xx yy zz xy xz yz */
    for (int k = j; k < j + cp.np[i]; ++k) {
      int l = T.sidx[k];
      double dx = s->x[l] - cp.x[i], dy = s->y[l] - cp.y[i], dz = s->z[l] - cp.z[i];
      s->ax[l] = g[0] + dx * H[0] + dy * H[3] + dz * H[4];
      s->ay[l] = g[1] + dx * H[3] + dy * H[1] + dz * H[5];
      s->az[l] = g[2] + dx * H[4] + dy * H[5] + dz * H[2];
    }
  }
}

double mwrap(double x, double lo, double hi, int bc) { /* mover.cpp:123-153 */
  if (bc == 1) {
    while (x >= hi) x -= hi - lo;
    while (x < lo) x += hi - lo;
  }
  return x;
}

}  // namespace

/* ----------------------------------------------------------- C API ------ */
extern "C" {

uint64_t orc_key64(double x, double y, double z, const double *bbox, int depth) {
  return key64(x, y, z, bbox, depth);
}

OrcSim *orc_new(const OrcParams *p) {
  OrcSim *s = new OrcSim();
  s->p = *p;
#ifdef _OPENMP
  if (p->nthreads > 0) omp_set_num_threads(p->nthreads);
#endif
  s->F = make_frame(p->bbox, p->bc);
  s->sv.lm = p->lm;
  s->sv.dirichlet_point = p->dirichlet_point;
  return s;
}

void orc_free(OrcSim *s) { delete s; }

int orc_set_particles(OrcSim *s, int n, const double *x, const double *y, const double *z,
                      const double *vx, const double *vy, const double *vz) {
  if (n < 1) return -1;
  s->N = n;
  s->x.assign(x, x + n); s->y.assign(y, y + n); s->z.assign(z, z + n);
  s->vx.assign(vx, vx + n); s->vy.assign(vy, vy + n); s->vz.assign(vz, vz + n);
  s->ax.assign(n, 0.0); s->ay.assign(n, 0.0); s->az.assign(n, 0.0);
  s->have_tree = s->have_comp = false;
  return 0;
}

int orc_get_particles(OrcSim *s, double *x, double *y, double *z, double *vx, double *vy,
                      double *vz, double *ax, double *ay, double *az) {
  size_t b = sizeof(double) * s->N;
  if (x) std::memcpy(x, s->x.data(), b);
  if (y) std::memcpy(y, s->y.data(), b);
  if (z) std::memcpy(z, s->z.data(), b);
  if (vx) std::memcpy(vx, s->vx.data(), b);
  if (vy) std::memcpy(vy, s->vy.data(), b);
  if (vz) std::memcpy(vz, s->vz.data(), b);
  if (ax) std::memcpy(ax, s->ax.data(), b);
  if (ay) std::memcpy(ay, s->ay.data(), b);
  if (az) std::memcpy(az, s->az.data(), b);
  return 0;
}

int orc_build_tree(OrcSim *s) {
  if (s->N < 1 || s->p.depth < 1 || s->p.depth > 21) return -1;
  build_tree(s->T, s->N, s->x.data(), s->y.data(), s->z.data(), s->p.bbox, s->p.depth);
  s->have_tree = true;
  return 0;
}

int orc_tree_info(OrcSim *s, int *len, int *numnodes) {
  if (!s->have_tree) return -1;
  *len = s->T.len;
  if (numnodes)
    for (int k = 0; k <= s->T.L; ++k) numnodes[k] = s->T.numnodes[k];
  return 0;
}

int orc_tree_sorted(OrcSim *s, uint64_t *key, int *idx) {
  if (!s->have_tree) return -1;
  std::memcpy(key, s->T.skey.data(), sizeof(uint64_t) * s->N);
  std::memcpy(idx, s->T.sidx.data(), sizeof(int) * s->N);
  return 0;
}

int orc_tree_nodes(OrcSim *s, uint64_t *key, int *depth, int *first, int *count, int *fchild,
                   int *nchild, double *lo, double *hi) {
  if (!s->have_tree) return -1;
  const Tree &T = s->T;
  for (int i = 0; i < T.len; ++i) {
    key[i] = T.nkey[i]; depth[i] = T.depth[i]; first[i] = T.first[i]; count[i] = T.count[i];
    fchild[i] = T.fchild[i]; nchild[i] = T.nchild[i];
    lo[3 * i] = T.lox[i]; lo[3 * i + 1] = T.loy[i]; lo[3 * i + 2] = T.loz[i];
    hi[3 * i] = T.hix[i]; hi[3 * i + 1] = T.hiy[i]; hi[3 * i + 2] = T.hiz[i];
  }
  return 0;
}

int orc_select(OrcSim *s, double c, int order, int mindepth, int maxdepth) {
  if (!s->have_tree) return -1;
  if (maxdepth > s->p.depth) return -1;
  Selector sel(s->T, s->x.data(), s->y.data(), s->z.data(), s->p.bc);
  sel.leaf_capacity = s->p.leaf_capacity;
  int S = sel.run(s->cp, c, order, mindepth, maxdepth, s->p.boundary_at_faces != 0);
  s->have_comp = true;
  return S;
}

int orc_set_comp(OrcSim *s, int nc, int nv, int nb, const double *x, const double *y,
                 const double *z, const double *h, const int *np, const int *fneigh,
                 const int *nneigh, const int *neigh24, int total_neigh) {
  Comp &c = s->cp;
  c.clear();
  c.nc = nc; c.nv = nv; c.nb = nb;
  const int S = nc + nv + nb;
  c.x.assign(x, x + S); c.y.assign(y, y + S); c.z.assign(z, z + S); c.h.assign(h, h + S);
  c.np.assign(np, np + S);
  c.first.assign(S, 0);
  c.type.resize(S);
  for (int i = 0; i < S; ++i) c.type[i] = i < nc ? 0 : (i < nc + nv ? -1 : 1);
  c.fneigh.assign(fneigh, fneigh + S);
  c.nneigh.assign(nneigh, nneigh + S);
  c.neigh24.assign(neigh24, neigh24 + total_neigh);
  s->F = make_frame(s->p.bbox, s->p.bc);
  s->have_comp = true;
  return 0;
}

int orc_comp_sizes(OrcSim *s, int *nc, int *nv, int *nb, int *nneigh) {
  if (!s->have_comp) return -1;
  *nc = s->cp.nc; *nv = s->cp.nv; *nb = s->cp.nb;
  *nneigh = (int)s->cp.neigh24.size();
  return 0;
}

int orc_comp_get(OrcSim *s, double *x, double *y, double *z, double *h, int *np, int *first,
                 int *type, int *fneigh, int *nneigh, int *neigh24) {
  if (!s->have_comp) return -1;
  const Comp &c = s->cp;
  int S = c.size();
  std::memcpy(x, c.x.data(), sizeof(double) * S);
  std::memcpy(y, c.y.data(), sizeof(double) * S);
  std::memcpy(z, c.z.data(), sizeof(double) * S);
  std::memcpy(h, c.h.data(), sizeof(double) * S);
  std::memcpy(np, c.np.data(), sizeof(int) * S);
  std::memcpy(first, c.first.data(), sizeof(int) * S);
  std::memcpy(type, c.type.data(), sizeof(int) * S);
  std::memcpy(fneigh, c.fneigh.data(), sizeof(int) * S);
  std::memcpy(nneigh, c.nneigh.data(), sizeof(int) * S);
  std::memcpy(neigh24, c.neigh24.data(), sizeof(int) * c.neigh24.size());
  return 0;
}

int orc_neighbors(OrcSim *s, int *nbr, double *dist) {
  if (!s->have_comp) return -1;
  int r = neighbors_all(s);
  if (r) return r;
  if (nbr) std::memcpy(nbr, s->nbr.data(), sizeof(int) * s->nbr.size());
  if (dist) std::memcpy(dist, s->dist.data(), sizeof(double) * s->dist.size());
  return 0;
}

int orc_set_constants(OrcSim *s, double c_dens, double c_pot, double c_field) {
  s->sv.c_dens = c_dens;
  s->sv.c_pot = c_pot;
  s->sv.c_field = c_field;
  return 0;
}

/* Exports the assembled rows densely: cols/a1/a2 hold lm entries per row in
 * neighbour order (a1 = 0 for boundary neighbours; a2 = the Laplacian
 * weight, boundary columns flagged by cols >= M). */
int orc_build_system(OrcSim *s, int *cols, double *a1, double *a2, double *d1, double *d2,
                     double *rhs1, double *pbnd) {
  if (!s->have_comp) return -1;
  int r = build_system(s);
  if (r) return r;
  const int M = s->cp.nc + s->cp.nv, lm = s->p.lm;
  for (int i = 0; i < M; ++i) {
    for (int q = 0; q < lm; ++q) {
      int j = s->nbr[(size_t)i * lm + q];
      cols[(size_t)i * lm + q] = j;
      double v1 = 0.0, v2 = 0.0;
      for (auto &e : s->A1.rows[i]) if (e.first == j && j != i) v1 = e.second;
      for (auto &e : s->A2.rows[i]) if (e.first == j && j != i) v2 = e.second;
      a1[(size_t)i * lm + q] = v1;
      a2[(size_t)i * lm + q] = v2;
    }
    double dd1 = 0, dd2 = 0;
    for (auto &e : s->A1.rows[i]) if (e.first == i) dd1 = e.second;
    for (auto &e : s->A2.rows[i]) if (e.first == i) dd2 = e.second;
    d1[i] = dd1;
    d2[i] = dd2;
    rhs1[i] = s->rhs1[i];
    pbnd[i] = s->pbnd[i];
  }
  return 0;
}

int orc_solve(OrcSim *s, OrcSolveInfo *info) {
  if (!s->have_comp) return -1;
  int r = solve_all(s);
  if (info) *info = s->info;
  return r;
}

int orc_comp_fields(OrcSim *s, double *rho, double *phi, double *g, double *H) {
  const int M = s->cp.nc + s->cp.nv;
  if ((int)s->g.size() < 3 * M) return -1;
  if (rho && (int)s->rho.size() >= M) std::memcpy(rho, s->rho.data(), sizeof(double) * M);
  if (phi && (int)s->phi.size() >= M) std::memcpy(phi, s->phi.data(), sizeof(double) * M);
  if (g) std::memcpy(g, s->g.data(), sizeof(double) * 3 * M);
  if (H) std::memcpy(H, s->H.data(), sizeof(double) * 6 * M);
  return 0;
}

int orc_set_gradients(OrcSim *s, const double *g, const double *H) {
  const int M = s->cp.nc + s->cp.nv;
  s->g.assign(g, g + 3 * (size_t)M);
  s->H.assign(H, H + 6 * (size_t)M);
  return 0;
}

int orc_set_potential(OrcSim *s, const double *phi) {
  const int M = s->cp.nc + s->cp.nv;
  s->phi.assign(phi, phi + M);
  return 0;
}

int orc_differentiate(OrcSim *s) {
  if (!s->have_comp) return -1;
  if (s->nbr.empty()) {
    int r = neighbors_all(s);
    if (r) return r;
  }
  if ((int)s->bv.size() != s->cp.nb) {
    s->bv.assign(s->cp.nb, 0.0);
    const int M = s->cp.nc + s->cp.nv;
    for (int q = 0; q < s->cp.nb; ++q)
      s->bv[q] = bv_value(s->p, s->cp.x[M + q], s->cp.y[M + q], s->cp.z[M + q]);
  }
  return differentiate(s);
}

/* mover.cpp:51-89 */
double orc_move_velocity(OrcSim *s, double Cv, double dt) {
  interpolate(s);
  double MAX_CHANGE = 0.0;
#pragma omp parallel
  {
    double max_change = 0.0, change;
#pragma omp for schedule(static)
    for (int i = 0; i < s->N; ++i) {
      s->vx[i] = s->vx[i] + s->ax[i] * Cv * dt;
      s->vy[i] = s->vy[i] + s->ay[i] * Cv * dt;
      s->vz[i] = s->vz[i] + s->az[i] * Cv * dt;
      change = (s->ax[i] * s->ax[i] + s->ay[i] * s->ay[i] + s->az[i] * s->az[i]) * Cv * Cv * dt * dt;
      if (change > max_change) max_change = change;
    }
#pragma omp critical
    if (max_change > MAX_CHANGE) MAX_CHANGE = max_change;
  }
  return sqrt(MAX_CHANGE);
}

/* mover.cpp:91-121 */
double orc_move_position(OrcSim *s, double Cx, double dt) {
  const double *b = s->p.bbox;
  int bc = s->p.bc;
  double MAX_CHANGE = 0.0;
#pragma omp parallel
  {
    double max_change = 0.0, change;
#pragma omp for schedule(static)
    for (int i = 0; i < s->N; ++i) {
      s->x[i] = mwrap(s->x[i] + s->vx[i] * Cx * dt, b[0], b[1], bc);
      s->y[i] = mwrap(s->y[i] + s->vy[i] * Cx * dt, b[2], b[3], bc);
      s->z[i] = mwrap(s->z[i] + s->vz[i] * Cx * dt, b[4], b[5], bc);
      change = (s->vx[i] * s->vx[i] + s->vy[i] * s->vy[i] + s->vz[i] * s->vz[i]) * Cx * Cx * dt * dt;
      if (change > max_change) max_change = change;
    }
#pragma omp critical
    if (max_change > MAX_CHANGE) MAX_CHANGE = max_change;
  }
  s->have_tree = s->have_comp = false;
  return sqrt(MAX_CHANGE);
}

/* controller.cpp:13-52 plus the build's normalised-comoving (2) and
 * Newtonian (3) unit modes */
void orc_coefficients(const OrcParams *p, double a, double *c_dens, double *c_pot,
                      double *c_field, double *cx, double *cv) {
  *c_field = -1.0;
  switch (p->units) {
    case 1: {
      *c_dens = 8.0 * a * a * a * kPiRef * p->G / 3.0 / 100000.0 / 100000.0 * kMpcCm * kMpcCm /
                p->omega * p->unit_mass / p->unit_length / p->unit_length / p->unit_length *
                p->mass_per_particle;
      *c_pot = 3.0 / 2.0 * p->omega / a;
      *cx = 1.0 / sqrt(p->omega / a + p->omega_lambda * a * a);
      *cv = 1.0 / sqrt(p->omega / a + p->omega_lambda * a * a) / a / a;
      break;
    }
    case 2: {
      double e = p->omega / a + p->omega_curv + p->omega_lambda * a * a;
      *c_dens = p->mass_per_particle;
      *c_pot = 3.0 / 2.0 * p->omega / a;
      *cx = 1.0 / sqrt(e);
      *cv = 1.0 / sqrt(e) / a / a;
      break;
    }
    case 3: {
      *c_dens = p->mass_per_particle;
      *c_pot = 4.0 * M_PI * p->G;
      *cx = 1.0;
      *cv = 1.0;
      break;
    }
    default: {
      *c_dens = 1.0;
      *c_pot = 1000.0;
      *cx = 1.0;
      *cv = 1.0;
    }
  }
}

/* controller.cpp:54-241 without Viewer output and MPI broadcasts */
int orc_run(OrcSim *s, double start_a, double end_a, double delta_a, int max_steps,
            double *timing, int *steps_done) {
  using clk = std::chrono::steady_clock;
  const OrcParams &p = s->p;
  double t = start_a;
  double cd, cpot, cf, cx, cv;
  double tsel = 0, tsol = 0, tmov = 0;
  orc_coefficients(&p, t, &cd, &cpot, &cf, &cx, &cv);
  double Cx = cx;
  orc_move_position(s, Cx, 0.5 * delta_a); /* :106-111 */
  t += 0.5 * delta_a;
  int ite = 0;
  double delta_t = delta_a;
  int err = 0;
  while (t < end_a && (max_steps <= 0 || ite < max_steps)) {
    auto t0 = clk::now();
    orc_build_tree(s);
    int S = orc_select(s, p.c, 2, 0, p.depth);
    auto t1 = clk::now();
    tsel += std::chrono::duration<double>(t1 - t0).count();
    if (S < 0) { err = -1; break; }
    if (p.adaptive_dt) { /* :146-162 */
      double max_density = 0.0;
      for (int i = 0; i < s->cp.size(); ++i) {
        double d = s->cp.np[i] / s->cp.h[i] / s->cp.h[i] / s->cp.h[i];
        if (max_density < d) max_density = d;
      }
      double mind = 1.0 / pow(max_density, 1.0 / 2);
      delta_t = 0.05 * delta_a * mind;
      if (delta_t > delta_a) delta_t = delta_a;
    } else {
      delta_t = delta_a;
    }
    if ((t + delta_t) > end_a) delta_t = end_a - t;
    orc_coefficients(&p, t, &cd, &cpot, &cf, &cx, &cv);
    double Cv = cv;
    double cx2, cv2, cd2, cp2, cf2;
    orc_coefficients(&p, t + delta_t / 2, &cd2, &cp2, &cf2, &cx2, &cv2);
    Cx = cx2;
    orc_set_constants(s, cd, cpot, cf);
    auto t2 = clk::now();
    err = solve_all(s);
    auto t3 = clk::now();
    tsol += std::chrono::duration<double>(t3 - t2).count();
    if (err) break;
    orc_move_velocity(s, Cv, delta_t);
    orc_move_position(s, Cx, delta_t);
    auto t4 = clk::now();
    tmov += std::chrono::duration<double>(t4 - t3).count();
    t += delta_t;
    ite++;
  }
  if (timing) { timing[0] = tsel; timing[1] = tsol; timing[2] = tmov; }
  if (steps_done) *steps_done = ite;
  return err;
}

int orc_wlsq(int m, int n, int nrhs, double *B, double *b) { return wlsq(m, n, nrhs, B, b); }

void orc_quicksort(int n, double *a, int *b) { aquicksort(a, b, 0, n - 1); }

int orc_gmres_csr(int n, const int *rowptr, const int *col, const double *val, const double *b,
                  double *x, double rtol, double atol, double dtol, int max_it, int restart,
                  int *its, int *reason, double *rnorm) {
  SpMat A;
  A.init(n);
  for (int i = 0; i < n; ++i)
    for (int q = rowptr[i]; q < rowptr[i + 1]; ++q) A.set(i, col[q], val[q]);
  Csr C = to_csr(A);
  *reason = gmres(C, b, x, rtol, atol, dtol, max_it, restart, its, rnorm);
  return 0;
}

}  // extern "C"
