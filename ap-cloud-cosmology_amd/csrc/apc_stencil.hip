// GFD weighted-least-squares stencils (solver.cpp:24-180) and the gradient /
// Hessian fit (solver.cpp:295-362), one thread per computational/vacuum row.
//
// The reference solves each row's lm x 9 weighted least-squares problem with
// LAPACKE_dgels (QR).  Here every row keeps a 9x9 triangular factor R with
// B^T B = R^T R in registers:
//   Laplacian weights  s_j = w_j^2 * basis_j . y,  y = R^-1 R^-T (e3+e4+e5)
//   (the reference's X3+X4+X5 column sums of X = (WB)^+ W), and
//   gradient/Hessian   c = R^-1 R^-T B^T b (the semi-normal equations),
//   R from the Cholesky factorisation of the Gram matrix B^T B, formed from
//   31 weighted moments of the neighbours' offsets accumulated in one pass.
//   Both agree with the reference's QR to rounding of order cond(B)^2 u.
// Matrices are stored slot-major (slot j of all rows contiguous) so the
// SpMV and the stencil writes coalesce.
#include <algorithm>
#include <cstring>

#include "apc_gfd.h"

namespace apc {


// A row whose fit is ill conditioned goes to k_fit_qr: its condition estimate
// (cond_frob, within a factor of 9 above cond_2(B)) exceeds `thr`, or a
// Cholesky pivot is not positive.  It is appended to `qrows` (count in
// qcnt[0]) and its outputs are left to k_fit_qr.  The largest estimate of the
// launch is kept in qcnt's maximum word (IEEE bits of a non-negative double
// order like its value).
struct FitQueue {
  double thr;
  int *rows;
  unsigned int *cnt;
  unsigned long long *maxbits;
  double *cond;  // per row: the estimate (infinity where Cholesky failed)
};

__device__ inline bool fit_needs_qr(const FitQueue &Q, int i, const double *R, bool deficient,
                                    double *est) {
  const double c = deficient ? INFINITY : cond_frob(R);
  *est = deficient ? 0.0 : c;
  Q.cond[i] = c;
  if (deficient || !(c <= Q.thr)) {
    Q.rows[atomicAdd(Q.cnt, 1u)] = i;
    return true;
  }
  return false;
}

__device__ inline void fit_max(const FitQueue &Q, double est, double *ws) {
  const double m = block_max256(est, ws);
  if (threadIdx.x == 0 && m > 0.0) atomicMax(Q.maxbits, (unsigned long long)__double_as_longlong(m));
}

__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) k_stencil(
    int M, int lm, Geo G, int pin_row0, Boxes B, double c_dens, const double *__restrict__ cx,
    const double *__restrict__ cy, const double *__restrict__ cz, const double *__restrict__ ch,
    const int *__restrict__ cnp, const int *__restrict__ nbr, const double *__restrict__ dist,
    const double *__restrict__ bv, double *__restrict__ sco, double *__restrict__ d1,
    double *__restrict__ d2, double *__restrict__ f1, double *__restrict__ f2,
    double *__restrict__ dinv1, double *__restrict__ dinv2, double *__restrict__ rhs1,
    double *__restrict__ pbnd, FitQueue Q) {
  __shared__ double ws[4];
  const int i = blockIdx.x * 256 + threadIdx.x;
  double est = 0.0;
  if (i < M) {
    const FitRow f = fit_row(cx, cy, cz, ch, dist, M, i, lm);
    // The weights need only y = (B^T B)^-1 u (B: the weighted basis rows),
    // and for a well-conditioned B a solve with B^T B through its Cholesky
    // factor is as good as one through a QR of B (error cond(B)^2 u, small
    // below the threshold): nine square roots per row instead of a rotation
    // per basis entry.  Rows above the threshold go to k_fit_qr.
    double mo[kMoments];
#pragma unroll
    for (int q = 0; q < kMoments; ++q) mo[q] = 0.0;
    for_slots(M, lm, i, nbr, dist, cx, cy, cz, NoExtra{},
              [&](int, int, double d, double x, double y, double z, double) {
                double p[3];
                const double w = fit_offset(G, f, d, x, y, z, p);
                moments_add(mo, w * w, p[0], p[1], p[2]);
              });
    double R[45];  // the Gram matrix (upper, packed), then its factor in place
    gram_from_moments(mo, R);
    const bool deficient = chol_inplace(R);
    rhs1[i] = 1.0 * cnp[i] / ch[i] / ch[i] / ch[i] * (c_dens * box_mass(B, i));
    if (!fit_needs_qr(Q, i, R, deficient, &est)) {
      // y = R^-1 R^-T u, u = e3+e4+e5
      double z[9];
#pragma unroll
      for (int k = 0; k < 9; ++k) {
        double s = (k >= 3 && k <= 5) ? 1.0 : 0.0;
#pragma unroll
        for (int q = 0; q < k; ++q) s -= R[RI(q, k)] * z[q];
        z[k] = s / R[RI(k, k)];
      }
      double y[9];
#pragma unroll
      for (int k = 8; k >= 0; --k) {
        double s = z[k];
#pragma unroll
        for (int q = k + 1; q < 9; ++q) s -= R[RI(k, q)] * y[q];
        y[k] = s / R[RI(k, k)];
      }
      // coefficients, A1 / A2 rows (solver.cpp:116-172); s/24 and s/h^2 as
      // products with the reciprocals
      constexpr double I24 = 1.0 / 24.0;
      double p2a = 0.0, p2b = 0.0, p3 = 0.0;
      const double ih2 = f.ih * f.ih;
      const bool pinned = pin_row(B, pin_row0, i);
      for_slots(M, lm, i, nbr, dist, cx, cy, cz, NoExtra{},
                [&](int j, int m, double d, double x, double yy, double zz, double) {
                  double r[9];
                  const double w = fit_slot(G, f, d, x, yy, zz, r);
                  double s = 0.0;
#pragma unroll
                  for (int k = 0; k < 9; ++k) s += r[k] * y[k];
                  s *= w * w;
                  sco[(size_t)j * M + i] = s;
                  p2a = p2a - s * I24;
                  const double a2 = s * ih2;
                  p2b = p2b - a2;
                  if (m >= M) p3 = p3 - bv[m - M] * a2;
                });
      double dd1 = p2a + 1;
      double dd2 = pinned ? 1.0 : p2b;
      d1[i] = dd1;
      d2[i] = dd2;
      f1[i] = I24;
      f2[i] = pinned ? 0.0 : ih2;
      dinv1[i] = (dd1 == 0.0) ? 1.0 : 1.0 / dd1;
      dinv2[i] = (dd2 == 0.0) ? 1.0 : 1.0 / dd2;
      pbnd[i] = pinned ? 0.0 : p3;
    }
  }
  fit_max(Q, est, ws);
}

// Extra gather of the gradient fit: phi at interior nodes, the boundary value
// at boundary nodes.
struct PhiOrBv {
  const double *phi, *bv;
  int M;
  __device__ double operator()(int m) const { return (m < M) ? phi[m] : bv[m - M]; }
};

__global__ void __launch_bounds__(256) k_gradient(
    int M, int lm, Geo G, double c_field, const double *__restrict__ cx,
    const double *__restrict__ cy, const double *__restrict__ cz, const double *__restrict__ ch,
    const int *__restrict__ nbr, const double *__restrict__ dist, const double *__restrict__ phi,
    const double *__restrict__ bv, double *__restrict__ grad, FitQueue Q) {
  __shared__ double ws[4];
  const int i = blockIdx.x * 256 + threadIdx.x;
  double est = 0.0;
  if (i < M) {
    const FitRow f = fit_row(cx, cy, cz, ch, dist, M, i, lm);
    const double p4 = phi[i];
    // cf = (B^T B)^-1 B^T b (B: the weighted basis rows, b: the weighted
    // differences), by the semi-normal equations with R from the Cholesky
    // factorisation of the Gram matrix: one pass over the neighbours, nine
    // square roots per row.  The adjoint (k_grad_adjoint) transposes this
    // form.  Rows above the condition threshold go to k_fit_qr.
    double mo[kMoments], Btb[9];
#pragma unroll
    for (int q = 0; q < kMoments; ++q) mo[q] = 0.0;
#pragma unroll
    for (int q = 0; q < 9; ++q) Btb[q] = 0.0;
    for_slots(M, lm, i, nbr, dist, cx, cy, cz, PhiOrBv{phi, bv, M},
              [&](int, int, double d, double x, double y, double z, double v) {
                double p[3], r[9];
                const double w = fit_offset(G, f, d, x, y, z, p);
                const double W = w * w;
                moments_add(mo, W, p[0], p[1], p[2]);
                basis_row(p[0], p[1], p[2], r);
                const double t = W * (v - p4) * c_field;
#pragma unroll
                for (int k = 0; k < 9; ++k) Btb[k] += r[k] * t;
              });
    double R[45];
    gram_from_moments(mo, R);
    const bool deficient = chol_inplace(R);
    if (!fit_needs_qr(Q, i, R, deficient, &est)) {
      double cf[9];
      seminormal_solve(R, Btb, cf);
      const double ih2 = f.ih * f.ih;
      grad[i] = cf[0] * f.ih;
      grad[(size_t)M + i] = cf[1] * f.ih;
      grad[2 * (size_t)M + i] = cf[2] * f.ih;
#pragma unroll
      for (int q = 3; q < 9; ++q) grad[(size_t)q * M + i] = cf[q] * ih2;
    }
  }
  fit_max(Q, est, ws);
}

// The fits of the queued rows by Householder QR of the weighted basis B
// (lm x 9), as LAPACKE_dgels does it (solver.cpp:106,342): one wave per row,
// lane j holding neighbour slot j (lm <= 64).  Reflector k (the dlarfg
// convention: beta = -sign(alpha) ||(alpha, x)||, tau = (beta - alpha) /
// beta, v = (1, x / (alpha - beta))) is applied to the remaining columns and
// to the right-hand side with wave-wide sums.  The error of these fits grows
// with cond(B), not with its square.
//   stencil: s_j = sum over k = 3..5 of X_kj, X = (WB)^+ W, i.e.
//            s_j = w_j (Q R^-T (e3+e4+e5))_j;
//   gradient: c = R^-1 (Q^T b)_{0..8}.
// A zero diagonal of R (dgels' info > 0) is APC_ERR_RANK.
template <bool GRAD>
__global__ void __launch_bounds__(64) k_fit_qr(
    int nq, const int *__restrict__ qrows, int M, int lm, Geo G, int pin_row0, Boxes B,
    double c_field,
    const double *__restrict__ cx, const double *__restrict__ cy, const double *__restrict__ cz,
    const double *__restrict__ ch, const int *__restrict__ nbr, const double *__restrict__ dist,
    const double *__restrict__ phi, const double *__restrict__ bv, double *__restrict__ sco,
    double *__restrict__ d1, double *__restrict__ d2, double *__restrict__ f1,
    double *__restrict__ f2, double *__restrict__ dinv1, double *__restrict__ dinv2,
    double *__restrict__ pbnd, double *__restrict__ grad, int *__restrict__ status) {
  const int lane = threadIdx.x;
  for (int q = blockIdx.x; q < nq; q += gridDim.x) {
    const int i = qrows[q];
    const FitRow f = fit_row(cx, cy, cz, ch, dist, M, i, lm);
    const bool live = lane < lm;
    double a[9], b = 0.0, w = 0.0;
    int m = i;
#pragma unroll
    for (int k = 0; k < 9; ++k) a[k] = 0.0;
    if (live) {
      m = nbr[(size_t)lane * M + i];
      w = fit_slot(G, f, dist[(size_t)lane * M + i], cx[m], cy[m], cz[m], a);
#pragma unroll
      for (int k = 0; k < 9; ++k) a[k] *= w;
      if (GRAD) b = w * (((m < M) ? phi[m] : bv[m - M]) - phi[i]) * c_field;
    }
    double v[9], tau[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      const double alpha = __shfl(a[k], k, 64);
      const double xn = sqrt(wave_sum((lane > k && live) ? a[k] * a[k] : 0.0));
      v[k] = 0.0;
      tau[k] = 0.0;
      if (xn == 0.0) continue;  // H = I; R_kk = alpha
      const double beta = -copysign(sqrt(alpha * alpha + xn * xn), alpha);
      tau[k] = (beta - alpha) / beta;
      const double scal = 1.0 / (alpha - beta);
      v[k] = (lane == k) ? 1.0 : ((lane > k && live) ? a[k] * scal : 0.0);
      if (lane == k) a[k] = beta;
      else if (lane > k) a[k] = 0.0;
#pragma unroll
      for (int c = k + 1; c < 9; ++c) {
        const double sc = tau[k] * wave_sum(v[k] * a[c]);
        a[c] -= sc * v[k];
      }
      if (GRAD) b -= tau[k] * wave_sum(v[k] * b) * v[k];
    }
    // R_kc lives in lane k's a[c]
    bool zero = false;
#pragma unroll
    for (int k = 0; k < 9; ++k) zero |= (__shfl(a[k], k, 64) == 0.0);
    if (zero) {
      if (lane == 0) {
        status[0] = APC_ERR_RANK;
        status[1] = i;
      }
      continue;
    }
    if (!GRAD) {
      double y[9];  // R^T y = e3 + e4 + e5 (every lane, same values)
#pragma unroll
      for (int k = 0; k < 9; ++k) {
        double t = (k >= 3 && k <= 5) ? 1.0 : 0.0;
#pragma unroll
        for (int c = 0; c < k; ++c) t -= __shfl(a[k], c, 64) * y[c];
        y[k] = t / __shfl(a[k], k, 64);
      }
      double t = 0.0;  // Q (y, 0): H_0 H_1 ... H_8 applied from the right end
#pragma unroll
      for (int k = 0; k < 9; ++k) t = (lane == k) ? y[k] : t;
#pragma unroll
      for (int k = 8; k >= 0; --k) t -= tau[k] * wave_sum(v[k] * t) * v[k];
      const double sj = live ? w * t : 0.0;
      if (live) sco[(size_t)lane * M + i] = sj;
      constexpr double I24 = 1.0 / 24.0;
      const double ih2 = f.ih * f.ih;
      // the row sums in slot order, as the reference accumulates them
      double p2a = 0.0, p2b = 0.0, p3 = 0.0;
      for (int j = 0; j < lm; ++j) {
        const double s = __shfl(sj, j, 64);
        const int mj = __shfl(m, j, 64);
        p2a = p2a - s * I24;
        const double a2 = s * ih2;
        p2b = p2b - a2;
        if (mj >= M) p3 = p3 - bv[mj - M] * a2;
      }
      if (lane == 0) {
        const bool pinned = pin_row(B, pin_row0, i);
        const double dd1 = p2a + 1, dd2 = pinned ? 1.0 : p2b;
        d1[i] = dd1;
        d2[i] = dd2;
        f1[i] = I24;
        f2[i] = pinned ? 0.0 : ih2;
        dinv1[i] = (dd1 == 0.0) ? 1.0 : 1.0 / dd1;
        dinv2[i] = (dd2 == 0.0) ? 1.0 : 1.0 / dd2;
        pbnd[i] = pinned ? 0.0 : p3;
      }
    } else {
      double qb[9], cf[9];
#pragma unroll
      for (int k = 0; k < 9; ++k) qb[k] = __shfl(b, k, 64);
#pragma unroll
      for (int k = 8; k >= 0; --k) {
        double t = qb[k];
#pragma unroll
        for (int c = k + 1; c < 9; ++c) t -= __shfl(a[c], k, 64) * cf[c];
        cf[k] = t / __shfl(a[k], k, 64);
      }
      if (lane < 9) {
        double out = cf[0];
#pragma unroll
        for (int k = 1; k < 9; ++k) out = (lane == k) ? cf[k] : out;
        grad[(size_t)lane * M + i] = out * ((lane < 3) ? f.ih : f.ih * f.ih);
      }
    }
  }
}

// Boundary values at the boundary nodes (bv_kind 1: uniform sphere).
__global__ void k_bv(int M, int nb, int kind, double GM, double R0, double ccx, double ccy,
                     double ccz, const double *__restrict__ cx, const double *__restrict__ cy,
                     const double *__restrict__ cz, double *__restrict__ bv) {
  int q = blockIdx.x * 256 + threadIdx.x;
  if (q >= nb) return;
  double v = 0.0;
  if (kind == 1) {
    int k = M + q;
    double dx = cx[k] - ccx, dy = cy[k] - ccy, dz = cz[k] - ccz;
    double r = sqrt(dx * dx + dy * dy + dz * dz);
    v = (r < R0) ? -GM * (3.0 * R0 * R0 - r * r) / (2.0 * R0 * R0 * R0) : -GM / r;
  }
  bv[q] = v;
}


// The queue of k_stencil / k_gradient: rows, count and largest estimate.
static int fit_queue(apc_ctx *c, FitQueue *Q) {
  const int M = c->nc + c->nv;
  APC_ALLOC(c, c->fitq_rows, (size_t)M + 1);
  APC_ALLOC(c, c->fitq_stat, 2);
  APC_ALLOC(c, c->fit_cond, (size_t)M + 1);
  APC_ALLOC(c, c->status_dev, 4);
  APC_HIP(c, hipMemsetAsync(c->status_dev.p, 0, 4 * sizeof(int), c->s));
  APC_HIP(c, hipMemsetAsync(c->fitq_stat.p, 0, 2 * sizeof(unsigned long long), c->s));
  Q->thr = c->fit_cond_qr;
  Q->rows = c->fitq_rows.p;
  Q->cnt = (unsigned int *)c->fitq_stat.p;
  Q->maxbits = c->fitq_stat.p + 1;
  Q->cond = c->fit_cond.p;
  return 0;
}

// Reads the queue after the Cholesky pass, runs k_fit_qr over the queued
// rows and checks for rank deficiency.
static int fit_finish(apc_ctx *c, bool grad, const char *what) {
  const int M = c->nc + c->nv, lm = c->prm.lm;
  unsigned long long st2[2] = {0, 0};
  APC_HIP(c, hipMemcpyAsync(st2, c->fitq_stat.p, sizeof(st2), hipMemcpyDeviceToHost, c->s));
  APC_HIP(c, hipStreamSynchronize(c->s));
  const int nq = (int)(unsigned int)st2[0];
  double mx;
  memcpy(&mx, &st2[1], sizeof(mx));
  c->fit_qr_rows = nq;
  c->fit_max_cond = mx;
  if (nq > 0) {
    if (lm > 64)
      return fail(c, APC_ERR_UNSUPPORTED, "%s: %d ill-conditioned rows need the QR fit, which "
                  "handles at most 64 neighbours (lm = %d)", what, nq, lm);
    const unsigned g = (unsigned)std::min(nq, 8192);
    const int pin = (c->prm.bc == 1 && c->prm.dirichlet_point) ? 1 : 0;
    // per row: its lm slots' index, distance, centre (36 B) and, for the
    // gradient, the potential (8 B); the row's outputs
    KScope ks(c, grad ? "gradient_qr" : "stencil_qr",
              (double)nq * ((grad ? 44.0 : 44.0) * lm + 100.0));
    if (grad)
      hipLaunchKernelGGL((k_fit_qr<true>), dim3(g), dim3(64), 0, c->s, nq, c->fitq_rows.p, M, lm,
                         geo_of(c), pin, boxes_of(c), c->c_field, c->cx.p, c->cy.p, c->cz.p,
                         c->ch.p, c->nbr.p,
                         c->dist.p, c->phi.p, c->bv.p, nullptr, nullptr, nullptr, nullptr,
                         nullptr, nullptr, nullptr, nullptr, c->grad.p, c->status_dev.p);
    else
      hipLaunchKernelGGL((k_fit_qr<false>), dim3(g), dim3(64), 0, c->s, nq, c->fitq_rows.p, M,
                         lm, geo_of(c), pin, boxes_of(c), c->c_field, c->cx.p, c->cy.p, c->cz.p,
                         c->ch.p,
                         c->nbr.p, c->dist.p, (const double *)nullptr, c->bv.p, c->sco.p, c->d1.p,
                         c->d2.p, c->f1.p, c->f2.p, c->dinv1.p, c->dinv2.p, c->pbnd.p, nullptr,
                         c->status_dev.p);
    APC_HIP(c, hipGetLastError());
    int st[2] = {0, 0};
    APC_HIP(c, hipMemcpyAsync(st, c->status_dev.p, sizeof(st), hipMemcpyDeviceToHost, c->s));
    APC_HIP(c, hipStreamSynchronize(c->s));
    if (st[0]) return fail(c, st[0], "%s: rank-deficient least-squares fit at row %d", what, st[1]);
  }
  return 0;
}

int build_stencils(apc_ctx *c) {
  if (!c->have_nbr) {
    int r = search_neighbors(c);
    if (r) return r;
  }
  const int M = c->nc + c->nv, lm = c->prm.lm, nb = c->nb;
  APC_ALLOC(c, c->sco, (size_t)lm * M + 1);
  APC_ALLOC(c, c->d1, M + 1); APC_ALLOC(c, c->d2, M + 1);
  APC_ALLOC(c, c->f1, M + 1); APC_ALLOC(c, c->f2, M + 1);
  APC_ALLOC(c, c->dinv1, M + 1); APC_ALLOC(c, c->dinv2, M + 1);
  APC_ALLOC(c, c->rhs1, M + 1); APC_ALLOC(c, c->pbnd, M + 1);
  APC_ALLOC(c, c->bv, nb + 1);
  FitQueue Q;
  if (int r = fit_queue(c, &Q)) return r;
  const apc_params &p = c->prm;
  if (nb > 0)
    hipLaunchKernelGGL(k_bv, dim3(nblocks(nb, 256)), dim3(256), 0, c->s, M, nb, p.bv_kind,
                       p.bv_G * p.bv_mass, p.bv_radius, p.bv_center[0], p.bv_center[1],
                       p.bv_center[2], c->cx.p, c->cy.p, c->cz.p, c->bv.p);
  int pin = (p.bc == 1 && p.dirichlet_point) ? 1 : 0;
  {
    // per slot: index 4 + distance 8 + neighbour xyz 24 read, weight 8 written;
    // per row: own xyz, h, n_p 36 read, eight row arrays and the estimate 72 written
    KScope ks(c, "stencil", (double)M * (44.0 * lm + 108.0));
    hipLaunchKernelGGL(k_stencil, dim3(nblocks(M, 256)), dim3(256), 0, c->s, M, lm, geo_of(c), pin,
                       boxes_of(c), c->c_dens, c->cx.p, c->cy.p, c->cz.p, c->ch.p, c->cnp.p, c->nbr.p,
                       c->dist.p, c->bv.p, c->sco.p, c->d1.p, c->d2.p, c->f1.p, c->f2.p,
                       c->dinv1.p, c->dinv2.p, c->rhs1.p, c->pbnd.p, Q);
  }
  APC_HIP(c, hipGetLastError());
  int r = fit_finish(c, false, "stencil");
  if (r) return r;
  c->stencil_qr_rows = c->fit_qr_rows;
  c->stencil_max_cond = c->fit_max_cond;
  c->have_sys = 1;
  ++c->sys_gen;
  return 0;
}

int differentiate(apc_ctx *c) {
  const int M = c->nc + c->nv, lm = c->prm.lm;
  APC_ALLOC(c, c->grad, (size_t)9 * M + 1);
  FitQueue Q;
  if (int r = fit_queue(c, &Q)) return r;
  {
    // per slot: index 4 + distance 8 + xyz 24 + phi 8; per row: own 40 read, 80 written
    KScope ks(c, "gradient", (double)M * (44.0 * lm + 120.0));
    hipLaunchKernelGGL(k_gradient, dim3(nblocks(M, 256)), dim3(256), 0, c->s, M, lm, geo_of(c),
                       c->c_field, c->cx.p, c->cy.p, c->cz.p, c->ch.p, c->nbr.p, c->dist.p,
                       c->phi.p, c->bv.p, c->grad.p, Q);
  }
  APC_HIP(c, hipGetLastError());
  int r = fit_finish(c, true, "gradient");
  if (r) return r;
  c->have_fields = 1;
  return 0;
}

}  // namespace apc

using namespace apc;

extern "C" int apc_build_system(apc_ctx *c, int *cols, double *a1, double *a2, double *d1,
                                double *d2, double *rhs1, double *pbnd) {
  if (!c) return APC_ERR_ARG;
  if (!c->have_comp) return fail(c, APC_ERR_ARG, "build_system: no computational particles");
  int r = build_stencils(c);
  if (r) return r;
  const int M = c->nc + c->nv, lm = c->prm.lm;
  std::vector<int> tn((size_t)M * lm);
  std::vector<double> ts((size_t)M * lm), h(M), vd1(M), vd2(M), vr(M), vp(M), vf2(M);
  APC_HIP(c, hipMemcpyAsync(tn.data(), c->nbr.p, 4ull * M * lm, hipMemcpyDeviceToHost, c->s));
  APC_HIP(c, hipMemcpyAsync(ts.data(), c->sco.p, 8ull * M * lm, hipMemcpyDeviceToHost, c->s));
  APC_HIP(c, hipMemcpyAsync(h.data(), c->ch.p, 8ull * M, hipMemcpyDeviceToHost, c->s));
  APC_HIP(c, hipMemcpyAsync(vd1.data(), c->d1.p, 8ull * M, hipMemcpyDeviceToHost, c->s));
  APC_HIP(c, hipMemcpyAsync(vd2.data(), c->d2.p, 8ull * M, hipMemcpyDeviceToHost, c->s));
  APC_HIP(c, hipMemcpyAsync(vr.data(), c->rhs1.p, 8ull * M, hipMemcpyDeviceToHost, c->s));
  APC_HIP(c, hipMemcpyAsync(vp.data(), c->pbnd.p, 8ull * M, hipMemcpyDeviceToHost, c->s));
  APC_HIP(c, hipMemcpyAsync(vf2.data(), c->f2.p, 8ull * M, hipMemcpyDeviceToHost, c->s));
  APC_HIP(c, hipStreamSynchronize(c->s));
  for (int i = 0; i < M; ++i) {
    for (int q = 0; q < lm; ++q) {
      int j = tn[(size_t)q * M + i];
      double s = ts[(size_t)q * M + i];
      if (cols) cols[(size_t)i * lm + q] = j;
      if (a1) a1[(size_t)i * lm + q] = (j < M) ? s / 24.0 : 0.0;
      if (a2) a2[(size_t)i * lm + q] = (vf2[i] == 0.0 || j >= M) ? 0.0 : s / h[i] / h[i];
    }
    if (d1) d1[i] = vd1[i];
    if (d2) d2[i] = vd2[i];
    if (rhs1) rhs1[i] = vr[i];
    if (pbnd) pbnd[i] = vp[i];
  }
  return 0;
}

extern "C" int apc_solver_set_fit_threshold(apc_ctx *c, double cond_qr) {
  if (!c || !(cond_qr >= 0.0)) return APC_ERR_ARG;
  c->fit_cond_qr = cond_qr;
  c->have_sys = 0;
  return 0;
}

extern "C" int apc_fit_conditions(apc_ctx *c, double *cond, int *qr_rows, double *max_cond) {
  if (!c) return APC_ERR_ARG;
  if (!c->have_sys) return fail(c, APC_ERR_ARG, "fit_conditions: no system built");
  const int M = c->nc + c->nv;
  if (cond) {
    APC_HIP(c, hipMemcpyAsync(cond, c->fit_cond.p, 8ull * M, hipMemcpyDeviceToHost, c->s));
    APC_HIP(c, hipStreamSynchronize(c->s));
  }
  if (qr_rows) *qr_rows = c->stencil_qr_rows;
  if (max_cond) *max_cond = c->stencil_max_cond;
  return 0;
}
