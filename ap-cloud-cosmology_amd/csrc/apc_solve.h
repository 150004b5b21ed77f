// Types shared by the Krylov solver (apc_gmres.hip) and the multigrid
// preconditioner (apc_amg.hip).
#pragma once
#include <algorithm>
#include <cmath>

#include "apc_internal.h"
#include "apc_prims.h"

namespace apc {

// One GFD system in slot-major ELL form: row i is
//   (A x)_i = d_i x_i + f_i * sum_j s_ij x_{nbr_ij}   over nbr_ij < M
// (neighbours at or beyond M are boundary nodes, already in the RHS).
// With trp set the operator is the transpose, applied by gathers through
// the column lists of A: column j's entries are e in [trp[j], trp[j+1]),
// each the ELL position tpos[e] = q*M + i of a_ij.
template <class T>
struct Op {
  const int *nbr;
  const T *s, *d, *f, *dinv;
  int M, lm;
  const int *trp = nullptr, *tpos = nullptr;
};

// Device-resident state of one GMRES solve; kernels launched ahead of the
// host's checks return at once when the solve has stopped or the cycle ended.
struct GState {
  int it, its, reason, cycle_done;
  int m, max_it, fp32, monitor;
  double res, rnorm0, ttol, cyc_res0;
  double rtol, atol, dtol;
};

__device__ inline bool halted(const GState *st) { return st->reason != 0 || st->cycle_done != 0; }

enum { KSP_CONVERGED_RTOL = 2, KSP_CONVERGED_ATOL = 3, KSP_DIVERGED_NULL = -2,
       KSP_DIVERGED_ITS = -3, KSP_DIVERGED_DTOL = -4, KSP_DIVERGED_BREAKDOWN = -5,
       KSP_DIVERGED_NANORINF = -9 };

// KSPConvergedDefault: the first residual sets the target rtol * r0 (atol
// below it); NaN/Inf and growth beyond dtol * r0 diverge.
__device__ inline int converged(GState *st, int nit, double rn) {
  if (nit == 0) {
    st->rnorm0 = rn;
    st->ttol = fmax(st->rtol * rn, st->atol);
  }
  if (isnan(rn) || isinf(rn)) return KSP_DIVERGED_NANORINF;
  if (rn <= st->ttol) return (rn < st->atol) ? KSP_CONVERGED_ATOL : KSP_CONVERGED_RTOL;
  if (rn >= st->dtol * st->rnorm0) return KSP_DIVERGED_DTOL;
  return 0;
}

struct GDev {
  GState *st;
  double *HH, *cc, *ss, *grs, *sig, *h, *coef, *part, *hist;
};

// Sum over the row's neighbour slots in slot order.  The slots are taken in
// groups of RG: the group's index and weight loads are issued together, then
// its gathers (an empty slot gathers the row's own entry and adds nothing), so
// a thread keeps a group's loads in flight instead of one slot's.  The
// additions are the slot-order sequence of the plain loop.  X: the vector's
// type (widened to T before the product).
template <class T, class X = T>
__device__ inline T row_offdiag_sum(const Op<T> &A, const X *__restrict__ x, int i) {
  constexpr int RG = 27;
  T acc = 0;
  int j = 0;
  for (; j + RG <= A.lm; j += RG) {
    int m[RG];
    T s[RG];
    X xv[RG];
#pragma unroll
    for (int u = 0; u < RG; ++u) {
      m[u] = A.nbr[(size_t)(j + u) * A.M + i];
      s[u] = A.s[(size_t)(j + u) * A.M + i];
    }
#pragma unroll
    for (int u = 0; u < RG; ++u) xv[u] = x[m[u] < A.M ? m[u] : i];
#pragma unroll
    for (int u = 0; u < RG; ++u)
      if (m[u] < A.M) acc += s[u] * (T)xv[u];
  }
  for (; j < A.lm; ++j) {
    int m = A.nbr[(size_t)j * A.M + i];
    if (m < A.M) acc += A.s[(size_t)j * A.M + i] * (T)x[m];
  }
  return acc;
}

template <class T>
__device__ inline double row_apply(const Op<T> &A, const T *__restrict__ x, int i) {
  if (A.trp) {  // (A^T x)_i = d_i x_i + sum over rows r with a_ri of f_r s_ri x_r
    T acc = 0;
    for (int e = A.trp[i]; e < A.trp[i + 1]; ++e) {
      const int p = A.tpos[e];
      const int r = p % A.M;
      acc += A.f[r] * A.s[p] * x[r];
    }
    return (double)(A.d[i] * x[i] + acc);
  }
  return (double)(A.d[i] * x[i] + A.f[i] * row_offdiag_sum(A, x, i));
}

struct AmgHier;  // apc_amg.hip

// Defaults from sweeps on the 256^3 Zel'dovich + Plummer system (7.5M rows):
// one Jacobi sweep each side (level 0: after the correction only, nopre),
// damping 0.85, a W-cycle on the three finest
// levels and V below, and a dense solve once a level has <= 256 rows (6
// levels: 7.5M, 611k, 58k, 6.7k, 838, ~100) took the step from 287 to 248 ms
// (40 -> 33 GMRES steps to rtol 1e-8) against a W-cycle everywhere with a
// dense level of <= 1024 rows.  The coarse correction is scaled by 1.6: per
// step 180 ms at 1.5 (33-34 GMRES steps), 174 at 1.6 (31-32), 173 at 1.65,
// 176 at 1.7 (30-35: a step needing more), and the cycle alone diverges at
// 1.9, so 1.6 keeps a margin.  Measured alternatives that were slower: the
// W-cycle from level 1 down (a V at level 0: 46 steps), a dense solve at the
// 851-row level, 8 lanes per row on level 1, damping 0.8 or 0.9 (no change).
struct AmgOptions {
  int nu = 1;             // Jacobi sweeps before and after the coarse correction
  double omega = 0.85;    // Jacobi damping
  // scaling of the coarse correction on levels >= 1 and on level 0; 0:
  // 1.5 with a geometric level 0 (at 256^3 1.6 on the coarse levels made
  // the schedule's cold first solve take 266 steps instead of 61; 1.6-1.8 on
  // level 0 alone are 2-4% faster, but the converged solves' true residual
  // then drifts to 2.2-2.7 rtol from the Krylov estimate), else 1.6;
  // alpha0 0: alpha
  double alpha = 0.0;
  double alpha0 = 0.0;
  int gamma = 2;          // 1: V-cycle, 2: W-cycle
  int wlevels = 3;        // gamma applies on the first wlevels levels, V below
  double theta = 0.0;     // strength threshold (fraction of the row's largest |a_ij|)
  int coarse_max = 256;   // dense solve at or below this many rows
  int max_levels = 12;
  int galerkin_sorted = 0;  // form every coarse operator by the global sort (tests)
  int shift = 1;            // shift the constant mode(s) of the periodic coarsest matrix
  // coarse levels with at most g8_rows / g4_rows rows run their sweeps and
  // residuals with 8 / 4 lanes per row; restrictions into at most rg8_rows
  // coarse rows use 8 lanes per coarse row.  The coarse levels are
  // latency-bound (a thread's chain of dependent gathers); at 256^3 (levels
  // of 611k, 58k, 6.7k rows) this took the Poisson solve from 348 to 304 ms.
  int g8_rows = 65536, g4_rows = 1 << 30, rg8_rows = 1 << 30;
  // level 0: after a coarse correction x += alpha P e, update the residual
  // by r -= alpha (A P) e instead of recomputing b - A x (and fold the
  // post-smoothing step's residual the same way); A P holds one entry per
  // aggregate a row's stencil meets, about a third of A's entries
  int ap = 1;
  // the first nopre levels without pre-smoothing (x = 0 before the first
  // coarse correction, so its residual is the level's input itself).  With
  // the A P updates level 0 then applies A only in the Krylov step's A z:
  // at 256^3 34-35 GMRES steps instead of 32, 138.2 -> 131.4 ms per FP32-form
  // step (FP64 154.3 -> 142.5); without pre-smoothing on levels 0-1, 0-2 or
  // all, 35-40 steps and no further gain
  int nopre = 1;
  // aggregates by octree cells instead of independent sets: level 0's rows
  // grouped by their ancestor geo0 levels up, each coarse level's by the
  // ancestor geoc levels up (0: independent-set aggregation; geoc only below
  // a geometric level)
  int geo0 = 1, geoc = 0;
};

// Build the hierarchy for A (FP64); `periodic` shifts the near-null constant
// mode of the pinned periodic operator in the coarsest solve.  FP32 copies
// of every level are made when `fp32` is set.
int amg_setup(apc_ctx *c, AmgHier **h, const Op<double> &A, bool periodic, bool fp32,
              const AmgOptions &o);
// out = M^-1 in (one cycle), level 0 given by A (same matrix as the setup, in
// the precision T).  Every kernel returns at once if st (device) says the
// solve has halted.
template <class T>
int amg_apply(apc_ctx *c, AmgHier *h, const Op<T> &A, const T *in, T *out, const GState *st);
void amg_free(AmgHier *h);

// Buffers of the Krylov solves, kept with the context between steps.
struct SolveState {
  DBuf<double> V64, dscratch, r64, part, z64, u64;
  DBuf<float> V32, s32, d32, f32, dinv32, z32, u32, Z32;
  DBuf<double> hess;  // HH, cc, ss, grs, sig, h, coef, hist
  DBuf<GState> st;
  GState *hst = nullptr;  // pinned mirror
  AmgHier *amg = nullptr;
  int mcap = 0, histcap = 0;
  // s32 holds the FP32 copy of these weights (the density and potential
  // systems share them): source, system generation, entries
  const double *s32_src = nullptr;
  long long s32_gen = -1;
  size_t s32_n = 0;
  // segmented solves (packed boxes): per-box Hessenberg data and states, the
  // tile table, and the pinned mirror of the states (nbox + 1: the summary)
  DBuf<double> shess, spart;
  DBuf<GState> sst;
  DBuf<int> stile;
  GState *hsst = nullptr;
  int hsst_cap = 0;
  ~SolveState() {  // the owner has drained its stream (apc_destroy)
    V64.release(); dscratch.release(); r64.release(); part.release(); z64.release();
    u64.release(); V32.release(); s32.release(); d32.release(); f32.release();
    dinv32.release(); z32.release(); u32.release(); Z32.release(); hess.release(); st.release();
    shess.release(); spart.release(); sst.release(); stile.release();
    if (hst) hipHostFree(hst);
    if (hsst) hipHostFree(hsst);
    amg_free(amg);
  }
};

// Arnoldi steps to launch before the host next reads the solve's state.
// Between reads the launches of a solve that has converged return at once (a
// few microseconds each, but a multigrid cycle is tens of launches), so the
// next read is placed where the residual, extrapolated at the rate observed
// in this restart cycle, reaches the target (at most 16 steps ahead).  With
// kernel timing on the state is read after every step, so that no timed
// launch is one that returned at once.
inline int steps_to_next_read(const GState &h, bool timing) {
  constexpr int kMax = 16;
  if (timing) return 1;
  double target = h.ttol;
  if (h.fp32) target = std::max(target, 2e-6 * h.cyc_res0);  // the FP32 cycle's floor
  if (h.it < 2 || !(h.res > 0.0) || !(h.cyc_res0 > h.res)) return kMax;
  if (h.res <= target) return 1;
  const double rate = std::pow(h.res / h.cyc_res0, 1.0 / h.it);
  const double need = std::log(target / h.res) / std::log(rate);
  if (!(need < kMax)) return kMax;
  return std::max(1, (int)std::ceil(need));
}

// GMRES cycles over packed boxes (apc_gmres_seg.hip): one Krylov space, one
// restart cycle and one convergence test per box (each box's solve follows
// the single-box GMRES of apc_gmres.hip), every box in the same launches.
// Arguments as the single-box cycles; per-box iterations and reasons go to
// its_box / reason_box (nbox entries, may be null); *its is the largest
// count, *reason the first failing box's reason or else the smallest
// positive one, *rnorm the largest final residual.
template <class T>
int run_cycles_seg(apc_ctx *c, const Op<double> &A64, const Op<T> &AT, T *V, const double *b,
                   double *x, const apc_ksp_opts &o, bool fp32, bool amg, T *zb, T *ub, int *its,
                   int *reason, double *rnorm, int *sweeps, const Op<float> *A32, float *Z32,
                   float *v32, float *z32);
int amg_levels(const AmgHier *h, int *rows, long long *nnz, int cap);
// Level l >= 1 as CSR (rp n+1, col/val nnz) and the aggregate map of level l
// (n entries, not for the coarsest); any pointer may be null.
int amg_level_copy(apc_ctx *c, const AmgHier *h, int l, int *rp, int *col, double *val, int *agg);

}  // namespace apc
