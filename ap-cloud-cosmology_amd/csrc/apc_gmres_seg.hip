// Segmented GMRES for packed boxes (apc_create_packed).
//
// The boxes' systems are decoupled blocks of one matrix whose rows are
// box-major.  One GMRES over the whole matrix would have to resolve every
// box's near-null constant mode (the periodic Poisson operator's, pinned by
// one row per box) and the union of the boxes' spectra: measured on 48 boxes
// of 16^3..96^3 particles it did not reach rtol 1e-8 in 1000 steps, where no
// box alone needs 30.  Here every box keeps its own Krylov basis, Hessenberg
// matrix, Givens rotations, restart cycle and convergence test — the
// single-box algorithm of apc_gmres.hip (PETSc's KSPGMRES semantics) applied
// per box — while each Arnoldi step's SpMV, multi-dot, orthogonalisation and
// multigrid cycle run over every box in the same launches:
//   * rows are cut into tiles of TILE rows that never straddle two boxes;
//     the reductions write one partial per tile and a per-box kernel sums
//     the box's tiles in tile order (deterministic);
//   * a box whose solve has stopped (converged, or its restart cycle ended)
//     skips its tiles and rows; the multigrid cycle, which does not know the
//     boxes, runs until every box has stopped (it is linear, and entries of a
//     stopped box are never used);
//   * restart cycles run in lockstep: a box whose cycle ended waits for the
//     others, then every box builds its update from its own step count.
// A stopped box's state is that of the single-box solve at the same point,
// so each box reports its own iterations, residual and reason.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "apc_solve.h"

namespace apc {

constexpr int SG_TILE = 1024;  // rows per tile: 256 threads x 4 rows

struct SegDev {
  GState *st;   // [nseg]
  GState *all;  // reason: every box has stopped for good; cycle_done: every box has stopped
  double *hs;   // per box: HH (m+1)m, then cc, ss, grs, sig, h, coef (m+2 each)
  double *part;  // [(m+2) x ntile] partial sums, vector-major
  const int *tseg, *tlo, *thi, *tfirst, *rowseg;
  int nseg, ntile, m;
  size_t hstride;
  __device__ double *HH(int s) const { return hs + (size_t)s * hstride; }
  __device__ double *vec(int s, int k) const {
    return hs + (size_t)s * hstride + (size_t)(m + 1) * m + (size_t)k * (m + 2);
  }
  __device__ double *cc(int s) const { return vec(s, 0); }
  __device__ double *ss(int s) const { return vec(s, 1); }
  __device__ double *grs(int s) const { return vec(s, 2); }
  __device__ double *sig(int s) const { return vec(s, 3); }
  __device__ double *h(int s) const { return vec(s, 4); }
  __device__ double *coef(int s) const { return vec(s, 5); }
};

// sum of the partials of box s's tiles (vector q's row of part), in tile order
__device__ inline double seg_partials(const SegDev &g, int s, int q, double *ws) {
  const double *pq = g.part + (size_t)q * g.ntile;
  double a = 0.0;
  for (int t = g.tfirst[s] + threadIdx.x; t < g.tfirst[s + 1]; t += 256) a += pq[t];
  return block_sum256(a, ws);
}

// ---- cycle start: v0 = D^-1 (b - A x) (left Jacobi) or b - A x, per tile ||.||^2 ----
template <class T>
__global__ void __launch_bounds__(256) k_sg_resid(Op<double> A, const double *__restrict__ x,
                                                  const double *__restrict__ b, int zero_guess,
                                                  int left_pc, T *__restrict__ v0, SegDev g) {
  __shared__ double ws[4];
  if (g.all->reason) return;
  for (int t = blockIdx.x; t < g.ntile; t += gridDim.x) {
    if (g.st[g.tseg[t]].reason) continue;  // uniform over the block
    double acc = 0.0;
    for (int i = g.tlo[t] + threadIdx.x; i < g.thi[t]; i += 256) {
      double r = zero_guess ? b[i] : b[i] - row_apply(A, x, i);
      if (left_pc) r = A.dinv[i] * r;
      v0[i] = (T)r;
      acc += r * r;
    }
    acc = block_sum256(acc, ws);
    if (threadIdx.x == 0) g.part[t] = acc;
  }
}

// one block per box: ||v0||, the box's cycle state, PETSc's first check
__global__ void __launch_bounds__(256) k_sg_resid_fin(SegDev g) {
  __shared__ double ws[4];
  const int s = blockIdx.x;
  GState *st = g.st + s;
  if (st->reason) {  // stopped for good: no update from this cycle
    if (threadIdx.x == 0) st->it = 0;
    return;
  }
  const double nrm2 = seg_partials(g, s, 0, ws);
  if (threadIdx.x == 0) {
    const double res = sqrt(nrm2);
    st->res = res;
    st->it = 0;
    st->cycle_done = 0;
    st->cyc_res0 = res;
    g.grs(s)[0] = res;
    g.sig(s)[0] = (res != 0.0) ? 1.0 / res : 1.0;
    if (res == 0.0) st->reason = KSP_CONVERGED_ATOL;
    else st->reason = converged(st, st->its, res);
    if (!st->reason && st->its >= st->max_it) st->cycle_done = 1;
  }
}

// the summary state that gates the launches shared by every box
__global__ void __launch_bounds__(256) k_sg_status(SegDev g) {
  __shared__ int nr, nh;
  if (threadIdx.x == 0) { nr = 0; nh = 0; }
  __syncthreads();
  int r = 0, h = 0;
  for (int s = threadIdx.x; s < g.nseg; s += 256) {
    r += g.st[s].reason != 0;
    h += halted(g.st + s);
  }
  atomicAdd(&nr, r);
  atomicAdd(&nh, h);
  __syncthreads();
  if (threadIdx.x == 0) {
    g.all->reason = (nr == g.nseg) ? 1 : 0;
    g.all->cycle_done = (nh == g.nseg) ? 1 : 0;
  }
}

template <class T>
__global__ void k_sg_copy(int n, const T *__restrict__ a, T *__restrict__ b, const GState *all) {
  if (halted(all)) return;
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) b[i] = a[i];
}

__global__ void k_sg_to_float(int n, const double *__restrict__ a, float *__restrict__ b,
                              const GState *all) {
  if (halted(all)) return;
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) b[i] = (float)a[i];
}

// ---- w = sig_j D^-1 A w_j (left Jacobi) or w = sig_j A z_j, rows of running boxes ----
template <class T>
__global__ void __launch_bounds__(256) k_sg_spmv(Op<T> A, const T *__restrict__ vj,
                                                 T *__restrict__ w, int j, int left_pc, SegDev g) {
  if (halted(g.all)) return;
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= A.M) return;
  const int s = g.rowseg[i];
  if (halted(g.st + s)) return;
  const T sj = (T)g.sig(s)[j];
  const T av = (T)row_apply(A, vj, i);
  w[i] = left_pc ? (T)(sj * (T)(A.dinv[i] * av)) : (T)(sj * av);
}

// flexible form: w = sig_j A z_j with z_j = M^-1 w_j from the FP32 multigrid,
// kept as column j of Z
__global__ void __launch_bounds__(256) k_sg_spmv_z(Op<double> A, const float *__restrict__ z,
                                                   double *__restrict__ w, float *__restrict__ zj,
                                                   int j, SegDev g) {
  if (halted(g.all)) return;
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= A.M) return;
  const int s = g.rowseg[i];
  if (halted(g.st + s)) return;
  const double sj = g.sig(s)[j];
  const double acc = row_offdiag_sum<double, float>(A, z, i);
  const float zi = z[i];
  zj[i] = zi;
  w[i] = sj * (A.d[i] * (double)zi + A.f[i] * acc);
}

// ---- h_q = sig_q (w_q . w) per box: one partial per (vector, tile) ----
// A thread keeps its (up to) 4 entries of w in registers and walks the basis
// U vectors at a time; per-vector wave sums go to LDS per wave.
template <class T, int U>
__global__ void __launch_bounds__(256) k_sg_mdot(int nv, const T *__restrict__ V, size_t ld,
                                                 const T *__restrict__ w, SegDev g) {
  extern __shared__ double lacc[];  // [4][nv]
  if (halted(g.all)) return;
  const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
  double *la = lacc + wave * nv;
  for (int tile = blockIdx.x; tile < g.ntile; tile += gridDim.x) {
    if (halted(g.st + g.tseg[tile])) continue;  // uniform over the block
    const int lo = g.tlo[tile], hi = g.thi[tile];
    const int base = lo + t;
    double wr[4];
    if (hi - lo == SG_TILE) {
#pragma unroll
      for (int r = 0; r < 4; ++r) wr[r] = (double)w[base + r * 256];
      int q = 0;
      for (; q + U <= nv; q += U) {
        double a[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const T *vq = V + (size_t)(q + u) * ld + base;
          a[u] = 0.0;
#pragma unroll
          for (int r = 0; r < 4; ++r) a[u] += wr[r] * (double)vq[r * 256];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const double sum = wave_sum(a[u]);
          if (lane == 0) la[q + u] = sum;
        }
      }
      for (; q < nv; ++q) {
        const T *vq = V + (size_t)q * ld + base;
        double a = 0.0;
#pragma unroll
        for (int r = 0; r < 4; ++r) a += wr[r] * (double)vq[r * 256];
        a = wave_sum(a);
        if (lane == 0) la[q] = a;
      }
    } else {  // a box's last, short tile
#pragma unroll
      for (int r = 0; r < 4; ++r) wr[r] = (base + r * 256 < hi) ? (double)w[base + r * 256] : 0.0;
      for (int q = 0; q < nv; ++q) {
        const T *vq = V + (size_t)q * ld;
        double a = 0.0;
#pragma unroll
        for (int r = 0; r < 4; ++r)
          if (base + r * 256 < hi) a += wr[r] * (double)vq[base + r * 256];
        a = wave_sum(a);
        if (lane == 0) la[q] = a;
      }
    }
    __syncthreads();
    for (int q = t; q < nv; q += 256)
      g.part[(size_t)q * g.ntile + tile] =
          (lacc[q] + lacc[nv + q]) + (lacc[2 * nv + q] + lacc[3 * nv + q]);
    __syncthreads();  // the next tile overwrites lacc
  }
}

// h_q of box blockIdx.y = sig_q * its tiles' partials; one wave per vector
__global__ void __launch_bounds__(256) k_sg_mdot_fin(int nv, SegDev g) {
  const int s = blockIdx.y;
  if (halted(g.st + s)) return;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int q = blockIdx.x * 4 + wave;
  if (q >= nv) return;
  const double *pq = g.part + (size_t)q * g.ntile;
  double a = 0.0;
  for (int t = g.tfirst[s] + lane; t < g.tfirst[s + 1]; t += 64) a += pq[t];
  a = wave_sum(a);
  if (lane == 0) g.h(s)[q] = g.sig(s)[q] * a;
}

// ---- w -= sum_q h_q sig_q w_q over each running box's tiles; ||w||^2 per tile ----
template <class T, int U>
__global__ void __launch_bounds__(256) k_sg_maxpy(int nv, const T *__restrict__ V, size_t ld,
                                                  T *__restrict__ w, SegDev g) {
  extern __shared__ double cq[];  // nv coefficients of the tile's box
  __shared__ double ws[4];
  if (halted(g.all)) return;
  const int t = threadIdx.x;
  const int nvU = nv - nv % U;
  for (int tile = blockIdx.x; tile < g.ntile; tile += gridDim.x) {
    const int s = g.tseg[tile];
    if (halted(g.st + s)) continue;  // uniform over the block
    for (int q = t; q < nv; q += 256) cq[q] = g.h(s)[q] * g.sig(s)[q];
    __syncthreads();
    double acc = 0.0;
    for (int i = g.tlo[tile] + t; i < g.thi[tile]; i += 256) {
      double sacc = (double)w[i];
      const T *vi = V + i;
      int q = 0;
      for (; q < nvU; q += U) {
        double a[U];
#pragma unroll
        for (int u = 0; u < U; ++u) a[u] = (double)vi[(size_t)(q + u) * ld];
#pragma unroll
        for (int u = 0; u < U; ++u) sacc -= cq[q + u] * a[u];
      }
      for (; q < nv; ++q) sacc -= cq[q] * (double)vi[(size_t)q * ld];
      const T sv = (T)sacc;
      w[i] = sv;
      acc += (double)sv * (double)sv;
    }
    acc = block_sum256(acc, ws);  // its closing barrier also guards cq
    if (t == 0) g.part[tile] = acc;
  }
}

// ---- one block per running box: ||w||, Hessenberg column j, Givens, test ----
// The per-box form of the single-box Hessenberg kernel (KSPGMRESCycle).
__global__ void __launch_bounds__(256) k_sg_hess(int j, SegDev g) {
  extern __shared__ double sh[];  // column [j+2], cc [j+1], ss [j+1]
  __shared__ double ws[4];
  const int s = blockIdx.x;
  GState *st = g.st + s;
  if (halted(st)) return;
  const double nrm2 = seg_partials(g, s, 0, ws);
  if (st->it != j) {  // every running box is at step j (guards the LDS size)
    if (threadIdx.x == 0) st->reason = KSP_DIVERGED_BREAKDOWN;
    return;
  }
  const int m = st->m;
  double *hc = sh, *cs = sh + (j + 2), *sn = cs + (j + 1);
  double *gh = g.h(s), *gcc = g.cc(s), *gss = g.ss(s), *grs = g.grs(s), *sig = g.sig(s);
  for (int q = threadIdx.x; q <= j; q += 256) {
    hc[q] = gh[q];
    cs[q] = gcc[q];
    sn[q] = gss[q];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const double tt = sqrt(nrm2);
    hc[j + 1] = tt;
    sig[j + 1] = (tt != 0.0) ? 1.0 / tt : 1.0;
    double hapbnd = fabs(tt / grs[j]);
    if (hapbnd > 1.0e-30) hapbnd = 1.0e-30;
    const bool hapend = tt < hapbnd;
    double prev = hc[0];
    for (int q = 1; q <= j; ++q) {  // KSPGMRESUpdateHessenberg
      const double c = cs[q - 1], sq = sn[q - 1], cur = hc[q];
      hc[q - 1] = c * prev + sq * cur;
      prev = c * cur - sq * prev;
    }
    hc[j] = prev;
    double res = 0.0;
    int reason = 0;
    bool null_col = false;
    if (!hapend) {
      const double t1 = sqrt(hc[j] * hc[j] + hc[j + 1] * hc[j + 1]);
      if (t1 == 0.0) {
        null_col = true;
      } else {
        const double cj = hc[j] / t1, sj = hc[j + 1] / t1;
        gcc[j] = cj;
        gss[j] = sj;
        const double gj = grs[j];
        grs[j + 1] = -(sj * gj);
        grs[j] = cj * gj;
        hc[j] = cj * hc[j] + sj * hc[j + 1];
        res = fabs(grs[j + 1]);
      }
    }
    if (null_col) {
      st->reason = KSP_DIVERGED_NULL;
    } else {
      st->it = j + 1;
      st->its += 1;
      st->res = res;
      reason = converged(st, st->its, res);
      if (hapend && !reason) reason = KSP_DIVERGED_BREAKDOWN;
      if (st->fp32 && (reason == KSP_CONVERGED_RTOL || reason == KSP_CONVERGED_ATOL || hapend)) {
        reason = 0;  // an FP32 estimate only ends the cycle; the FP64 restart decides
        st->cycle_done = 1;
      }
      st->reason = reason;
      if (!reason) {
        if (st->fp32 && res <= 2e-6 * st->cyc_res0) st->cycle_done = 1;
        if (st->it >= m || st->its >= st->max_it) st->cycle_done = 1;
      }
    }
  }
  __syncthreads();
  double *hh = g.HH(s) + (size_t)j * (m + 1);
  for (int q = threadIdx.x; q <= j + 1; q += 256) hh[q] = hc[q];
}

// ---- KSPGMRESBuildSoln per box over its own it steps: coef_q = y_q sig_q ----
__global__ void k_sg_buildsoln(SegDev g) {
  if (threadIdx.x != 0) return;
  const int s = blockIdx.x;
  GState *st = g.st + s;
  const int m = st->m;
  const double *HHs = g.HH(s), *grs = g.grs(s), *sig = g.sig(s);
  auto H = [&](int r, int c) { return HHs[(size_t)c * (m + 1) + r]; };
  double *nrs = g.coef(s);
  int top = st->it - 1;
  if (top >= 0) {
    if (H(top, top) == 0.0) { st->reason = KSP_DIVERGED_BREAKDOWN; top = -1; }
    else nrs[top] = grs[top] / H(top, top);
  }
  for (int ii = 1; top >= 0 && ii <= top; ++ii) {
    const int k = top - ii;
    double tt = grs[k];
    for (int q = k + 1; q <= top; ++q) tt = tt - H(k, q) * nrs[q];
    if (H(k, k) == 0.0) { st->reason = KSP_DIVERGED_BREAKDOWN; top = -1; break; }
    nrs[k] = tt / H(k, k);
  }
  for (int q = 0; q <= top; ++q) nrs[q] *= sig[q];
  if (top < 0) st->it = 0;  // no update
}

// ---- x += sum_q coef_q V_q (add != 0), or u = sum_q coef_q V_q, per box ----
template <class T, class O>
__global__ void __launch_bounds__(256) k_sg_combine(const T *__restrict__ V, size_t ld, SegDev g,
                                                    O *__restrict__ out, int add) {
  extern __shared__ double cs[];
  for (int tile = blockIdx.x; tile < g.ntile; tile += gridDim.x) {
    const int s = g.tseg[tile];
    const int nv = g.st[s].it;
    if (nv <= 0 && add) continue;  // uniform over the block
    for (int q = threadIdx.x; q < nv; q += 256) cs[q] = g.coef(s)[q];
    __syncthreads();
    for (int i = g.tlo[tile] + threadIdx.x; i < g.thi[tile]; i += 256) {
      double a = 0.0;
      for (int q = 0; q < nv; ++q) a += cs[q] * (double)V[(size_t)q * ld + i];
      if (add) out[i] = (O)((double)out[i] + a);
      else out[i] = (O)a;
    }
    __syncthreads();  // the next tile overwrites cs
  }
}

template <class T>
__global__ void k_sg_add_to(int M, const T *__restrict__ z, double *__restrict__ x) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < M) x[i] += (double)z[i];
}

static int read_seg_state(apc_ctx *c, SolveState *S, int nseg) {
  APC_HIP(c, hipMemcpyAsync(S->hsst, S->sst.p, sizeof(GState) * (nseg + 1),
                            hipMemcpyDeviceToHost, c->s));
  APC_HIP(c, hipStreamSynchronize(c->s));
  return 0;
}

template <class T>
int run_cycles_seg(apc_ctx *c, const Op<double> &A64, const Op<T> &AT, T *V, const double *b,
                   double *x, const apc_ksp_opts &o, bool fp32, bool amg, T *zb, T *ub,
                   int *its_out, int *reason_out, double *rnorm_out, int *sweeps,
                   const Op<float> *A32, float *Z32, float *v32, float *z32) {
  const bool flex = Z32 != nullptr;
  const int M = A64.M, nseg = c->nbox;
  const int m = std::min(std::max(1, o.restart), 1000);
  SolveState *S = c->gm;
  if ((int)c->rstart_h.size() != nseg + 1 || c->rstart_h[nseg] != M || !c->rowbox.p)
    return fail(c, APC_ERR_ARG, "segmented gmres: the boxes' rows do not cover the system");
  // tiles of at most SG_TILE rows inside each box
  std::vector<int> tseg, tlo, thi, tfirst(nseg + 1);
  for (int s = 0; s < nseg; ++s) {
    tfirst[s] = (int)tseg.size();
    for (int a = c->rstart_h[s]; a < c->rstart_h[s + 1]; a += SG_TILE) {
      tseg.push_back(s);
      tlo.push_back(a);
      thi.push_back(std::min(a + SG_TILE, c->rstart_h[s + 1]));
    }
  }
  const int ntile = (int)tseg.size();
  tfirst[nseg] = ntile;
  std::vector<int> tab;
  tab.reserve(3 * (size_t)ntile + nseg + 1);
  tab.insert(tab.end(), tseg.begin(), tseg.end());
  tab.insert(tab.end(), tlo.begin(), tlo.end());
  tab.insert(tab.end(), thi.begin(), thi.end());
  tab.insert(tab.end(), tfirst.begin(), tfirst.end());
  APC_ALLOC(c, S->stile, tab.size() + 1);
  const size_t hstride = (size_t)(m + 1) * m + 6 * (size_t)(m + 2);
  APC_ALLOC(c, S->shess, hstride * nseg + 1);
  APC_ALLOC(c, S->spart, (size_t)(m + 2) * ntile + 1);
  APC_ALLOC(c, S->sst, nseg + 2);
  if (S->hsst_cap < nseg + 1) {
    if (S->hsst) APC_HIP(c, hipHostFree(S->hsst));
    S->hsst = nullptr;
    S->hsst_cap = 0;
    APC_HIP(c, hipHostMalloc((void **)&S->hsst, sizeof(GState) * (nseg + 1)));
    S->hsst_cap = nseg + 1;
  }
  GState h0;
  memset(&h0, 0, sizeof(h0));
  h0.m = m;
  h0.max_it = o.max_it;
  h0.fp32 = fp32 ? 1 : 0;
  h0.rtol = o.rtol;
  h0.atol = o.atol;
  h0.dtol = o.dtol;
  for (int s = 0; s < nseg; ++s) S->hsst[s] = h0;
  memset(&S->hsst[nseg], 0, sizeof(GState));
  APC_HIP(c, hipMemcpyAsync(S->stile.p, tab.data(), sizeof(int) * tab.size(),
                            hipMemcpyHostToDevice, c->s));
  APC_HIP(c, hipMemcpyAsync(S->sst.p, S->hsst, sizeof(GState) * (nseg + 1),
                            hipMemcpyHostToDevice, c->s));
  APC_HIP(c, hipStreamSynchronize(c->s));  // the table's host vector goes out of scope
  APC_HIP(c, hipMemsetAsync(x, 0, sizeof(double) * M, c->s));
  SegDev g;
  g.st = S->sst.p;
  g.all = S->sst.p + nseg;
  g.hs = S->shess.p;
  g.part = S->spart.p;
  g.tseg = S->stile.p;
  g.tlo = g.tseg + ntile;
  g.thi = g.tlo + ntile;
  g.tfirst = g.thi + ntile;
  g.rowseg = c->rowbox.p;
  g.nseg = nseg;
  g.ntile = ntile;
  g.m = m;
  g.hstride = hstride;
  const GState *all = g.all;
  constexpr int MU = (sizeof(T) == 4) ? 8 : 4;
  const size_t ld = (size_t)M;
  const unsigned gb = nblocks(M, 256);
  const unsigned gt = (unsigned)std::min(ntile, c->knob[APC_KNOB_SEG_BLOCKS]);
  const double nnz = (double)M * A64.lm;
  const double bspmv = fp32 ? 8.0 * nnz + 20.0 * M : 12.0 * nnz + 40.0 * M;
  int itcount = 0, ncyc = 0;
  bool zero_guess = true;
  auto next_read = [&]() {
    int k = 16;
    for (int s = 0; s < nseg; ++s)
      if (!S->hsst[s].reason && !S->hsst[s].cycle_done)
        k = std::min(k, steps_to_next_read(S->hsst[s], c->timing));
    return k;
  };
  for (;;) {
    {
      KScope ks(c, fp32 ? "sgmres_resid64" : "sgmres_resid", 12.0 * nnz + 48.0 * M + sizeof(T) * M);
      hipLaunchKernelGGL(k_sg_resid<T>, dim3(gt), dim3(256), 0, c->s, A64, x, b,
                         zero_guess ? 1 : 0, amg ? 0 : 1, V, g);
    }
    hipLaunchKernelGGL(k_sg_resid_fin, dim3(nseg), dim3(256), 0, c->s, g);
    hipLaunchKernelGGL(k_sg_status, dim3(1), dim3(256), 0, c->s, g);
    ++ncyc;
    const int budget = o.max_it - itcount;
    int r = read_seg_state(c, S, nseg);
    if (r) return r;
    if (S->hsst[nseg].reason) break;
    int to_read = next_read();
    for (int j = 0; j < m && j < budget; ++j) {
      const T *vj = V + (size_t)j * ld;
      T *w = V + (size_t)(j + 1) * ld;
      if (flex) {
        float *zj = Z32 + (size_t)j * ld;
        hipLaunchKernelGGL(k_sg_to_float, dim3(gb), dim3(256), 0, c->s, M, (const double *)vj, v32,
                           all);
        r = amg_apply<float>(c, S->amg, *A32, v32, z32, all);
        if (r) return r;
        KScope ks(c, "sgmres_spmv_z", 12.0 * nnz + 32.0 * M);
        hipLaunchKernelGGL(k_sg_spmv_z, dim3(gb), dim3(256), 0, c->s, A64, (const float *)z32,
                           (double *)w, zj, j, g);
      } else if (amg) {
        hipLaunchKernelGGL(k_sg_copy<T>, dim3(gb), dim3(256), 0, c->s, M, vj, ub, all);
        r = amg_apply<T>(c, S->amg, AT, ub, zb, all);
        if (r) return r;
        KScope ks(c, fp32 ? "sgmres_spmv32" : "sgmres_spmv", bspmv);
        hipLaunchKernelGGL(k_sg_spmv<T>, dim3(gb), dim3(256), 0, c->s, AT, (const T *)zb, w, j, 0,
                           g);
      } else {
        KScope ks(c, fp32 ? "sgmres_spmv32" : "sgmres_spmv", bspmv);
        hipLaunchKernelGGL(k_sg_spmv<T>, dim3(gb), dim3(256), 0, c->s, AT, vj, w, j, 1, g);
      }
      const int nv = j + 1;
      {
        KScope ks(c, fp32 ? "sgmres_mdot32" : "sgmres_mdot", (double)sizeof(T) * M * (nv + 1));
        hipLaunchKernelGGL((k_sg_mdot<T, 2>), dim3(gt), dim3(256), 4 * nv * sizeof(double), c->s,
                           nv, V, ld, w, g);
      }
      hipLaunchKernelGGL(k_sg_mdot_fin, dim3(nblocks(nv, 4), nseg), dim3(256), 0, c->s, nv, g);
      {
        KScope ks(c, fp32 ? "sgmres_maxpy32" : "sgmres_maxpy", (double)sizeof(T) * M * (nv + 2));
        hipLaunchKernelGGL((k_sg_maxpy<T, MU>), dim3(gt), dim3(256), nv * sizeof(double), c->s, nv,
                           V, ld, w, g);
      }
      hipLaunchKernelGGL(k_sg_hess, dim3(nseg), dim3(256), (3 * (size_t)j + 4) * sizeof(double),
                         c->s, j, g);
      hipLaunchKernelGGL(k_sg_status, dim3(1), dim3(256), 0, c->s, g);
      if (--to_read == 0) {
        r = read_seg_state(c, S, nseg);
        if (r) return r;
        if (S->hsst[nseg].reason || S->hsst[nseg].cycle_done) break;
        to_read = next_read();
      }
    }
    r = read_seg_state(c, S, nseg);
    if (r) return r;
    int it = 0;
    for (int s = 0; s < nseg; ++s) it = std::max(it, S->hsst[s].it);
    if (it > 0) {
      hipLaunchKernelGGL(k_sg_buildsoln, dim3(nseg), dim3(64), 0, c->s, g);
      if (flex) {  // x += sum_q y_q z_q
        KScope ks(c, "sgmres_update_z", 4.0 * M * it + 16.0 * M);
        hipLaunchKernelGGL((k_sg_combine<float, double>), dim3(gt), dim3(256), it * sizeof(double),
                           c->s, (const float *)Z32, ld, g, x, 1);
      } else if (amg) {  // x += M^-1 (sum_q y_q v_q), zero for a box without steps
        {
          KScope ks(c, fp32 ? "sgmres_update32" : "sgmres_update", (double)sizeof(T) * M * (it + 1));
          hipLaunchKernelGGL((k_sg_combine<T, T>), dim3(gt), dim3(256), it * sizeof(double), c->s,
                             (const T *)V, ld, g, ub, 0);
        }
        r = amg_apply<T>(c, S->amg, AT, ub, zb, nullptr);
        if (r) return r;
        hipLaunchKernelGGL(k_sg_add_to<T>, dim3(gb), dim3(256), 0, c->s, M, (const T *)zb, x);
      } else {
        KScope ks(c, fp32 ? "sgmres_update32" : "sgmres_update", (double)sizeof(T) * M * it + 16.0 * M);
        hipLaunchKernelGGL((k_sg_combine<T, double>), dim3(gt), dim3(256), it * sizeof(double),
                           c->s, (const T *)V, ld, g, x, 1);
      }
      hipLaunchKernelGGL(k_sg_status, dim3(1), dim3(256), 0, c->s, g);
      r = read_seg_state(c, S, nseg);
      if (r) return r;
    }
    itcount += it;
    if (S->hsst[nseg].reason) break;
    if (itcount >= o.max_it) break;
    zero_guess = false;
  }
  int its = 0, reason = 0, worst = 0;
  double rn = 0.0;
  c->seg_its.assign(nseg, 0);
  c->seg_reason.assign(nseg, 0);
  c->seg_res.assign(nseg, 0.0);
  for (int s = 0; s < nseg; ++s) {
    const GState &h = S->hsst[s];
    its = std::max(its, h.its);
    rn = std::max(rn, h.res);
    const int rs = h.reason ? h.reason : KSP_DIVERGED_ITS;
    c->seg_its[s] = h.its;
    c->seg_reason[s] = rs;
    c->seg_res[s] = h.res;
    if (rs < 0 && !worst) worst = rs;
    if (rs > 0 && (!reason || rs < reason)) reason = rs;
  }
  *its_out = its;
  *reason_out = worst ? worst : reason;
  *rnorm_out = rn;
  if (sweeps) *sweeps = ncyc;
  APC_HIP(c, hipGetLastError());
  return 0;
}

template int run_cycles_seg<double>(apc_ctx *, const Op<double> &, const Op<double> &, double *,
                                    const double *, double *, const apc_ksp_opts &, bool, bool,
                                    double *, double *, int *, int *, double *, int *,
                                    const Op<float> *, float *, float *, float *);
template int run_cycles_seg<float>(apc_ctx *, const Op<double> &, const Op<float> &, float *,
                                   const double *, double *, const apc_ksp_opts &, bool, bool,
                                   float *, float *, int *, int *, double *, int *,
                                   const Op<float> *, float *, float *, float *);

}  // namespace apc
