// Differentiation form: adjoints of the solve-and-move step with the tree,
// the neighbour lists and the particle-to-leaf map held fixed.
//
// With the topology fixed after step 0 a step is, per particle s of leaf o,
//   v' = v + (g_o + H_o (x - c_o)) Cv dt          (mover.cpp:51-89,
//   x' = wrap(x + v' Cx dt)                          octree.cpp:3003-3031,
//                                                    mover.cpp:91-153)
// and the node fields come from the GFD systems of solver.cpp:24-362, whose
// right-hand sides depend on the fixed leaf counts only.  The adjoint of a
// step is therefore: drift adjoint (lambda_v += Cx dt lambda_x), kick adjoint
// (lambda_x += Cv dt H_o lambda_v), and, for the node fields, the transposes
// of the interpolation (dg, dH per leaf), of the gradient fit (Differentiate,
// solver.cpp:295-362) and of the two GMRES solves (A^T by gathers through an
// explicit transpose of the ELL matrix).  The loss is the sum of squared CIC
// density contrast on a periodic n^3 grid.
#include <algorithm>
#include <cmath>
#include <vector>

#include "apc_gfd.h"
#include "apc_solve.h"

namespace apc {

static __global__ void k_fill_zero(size_t n, double *__restrict__ p) {
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256)
    p[i] = 0.0;
}

// ---------------------------------------------------------------- CIC loss
__device__ inline void cic_axis(double x, int n, int &i0, int &i1, double &w0, double &w1) {
  double u = x * n - 0.5;
  double fl = floor(u);
  double f = u - fl;
  int i = (int)fl;
  i0 = ((i % n) + n) % n;
  i1 = (i0 + 1) % n;
  w0 = 1.0 - f;
  w1 = f;
}

// The deposit is a sorted, compensated gather rather than an atomic scatter:
// each particle emits its 8 (cell, corner) contributions, a stable radix sort
// groups them by cell in particle order, and every cell sums its group with
// error-free transformations.  The loss is then bitwise reproducible and its
// rounding error is about one unit in the last place per cell, which is what
// lets central differences of the loss resolve the gradient.
__device__ inline void two_sum(double a, double b, double &s, double &e) {
  s = a + b;
  const double bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}

__device__ inline double cic_weight(double x, double y, double z, int n, int q, int *cell) {
  int ix[2], iy[2], iz[2];
  double wx[2], wy[2], wz[2];
  cic_axis(x, n, ix[0], ix[1], wx[0], wx[1]);
  cic_axis(y, n, iy[0], iy[1], wy[0], wy[1]);
  cic_axis(z, n, iz[0], iz[1], wz[0], wz[1]);
  const int a = q >> 2, b = (q >> 1) & 1, c = q & 1;
  *cell = (ix[a] * n + iy[b]) * n + iz[c];
  return wx[a] * wy[b] * wz[c];
}

__global__ void __launch_bounds__(256) k_cic_keys(int N, int n, const double *__restrict__ x,
                                                  const double *__restrict__ y,
                                                  const double *__restrict__ z,
                                                  uint64_t *__restrict__ key, int *__restrict__ val) {
  const long e = (long)blockIdx.x * 256 + threadIdx.x;  // contribution 8 s + q
  if (e >= 8L * N) return;
  const int s = (int)(e >> 3), q = (int)(e & 7);
  int cell;
  cic_weight(x[s], y[s], z[s], n, q, &cell);
  key[e] = (uint64_t)cell;
  val[e] = (int)e;
}

// rng[2c], rng[2c+1] = first and one-past-last sorted contribution of cell c
// (cells without contributions keep 0, 0)
__global__ void k_cic_ranges(long ne, const uint64_t *__restrict__ key, int *__restrict__ rng) {
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i >= ne) return;
  const uint64_t k = key[i];
  if (i == 0 || key[i - 1] != k) rng[2 * k] = (int)i;
  if (i == ne - 1 || key[i + 1] != k) rng[2 * k + 1] = (int)(i + 1);
}

// delta_c = scale * sum of the cell's weights - 1, summed in sorted order with
// compensation; block partials of sum delta^2 as (value, error) pairs.
__global__ void __launch_bounds__(256) k_cic_cells(long ng, int n, const double *__restrict__ x,
                                                   const double *__restrict__ y,
                                                   const double *__restrict__ z, double scale,
                                                   const double *__restrict__ wts,
                                                   const int *__restrict__ val,
                                                   const int *__restrict__ rng,
                                                   double *__restrict__ delta,
                                                   double *__restrict__ part) {
  __shared__ double ws[4];
  double hi = 0.0, lo = 0.0;
  for (long c = (long)blockIdx.x * 256 + threadIdx.x; c < ng; c += (long)gridDim.x * 256) {
    double s = -1.0, err = 0.0;
    for (int i = rng[2 * c]; i < rng[2 * c + 1]; ++i) {
      const int e = val[i], p = e >> 3;
      int cell;
      const double w = (wts ? scale * wts[p] : scale) * cic_weight(x[p], y[p], z[p], n, e & 7, &cell);
      double t;
      two_sum(s, w, s, t);
      err += t;
    }
    const double d = s + err;
    delta[c] = d;
    const double sq = __dmul_rn(d, d), sqe = fma(d, d, -sq);  // no contraction into the sum
    double t;
    two_sum(hi, sq, hi, t);
    lo += t + sqe;
  }
  hi = block_sum256(hi, ws);
  lo = block_sum256(lo, ws);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = hi;
    part[2 * blockIdx.x + 1] = lo;
  }
}

// lambda_x = dL/dx = sum_c 2 delta_c scale w_s dW_c/dx, lambda_v = 0
__global__ void __launch_bounds__(256) k_cic_grad(int N, int n, const double *__restrict__ x,
                                                  const double *__restrict__ y,
                                                  const double *__restrict__ z, double scale,
                                                  const double *__restrict__ wts,
                                                  const double *__restrict__ delta,
                                                  double *__restrict__ lam) {
  int s = blockIdx.x * 256 + threadIdx.x;
  if (s >= N) return;
  if (wts) scale *= wts[s];
  int ix[2], iy[2], iz[2];
  double wx[2], wy[2], wz[2];
  cic_axis(x[s], n, ix[0], ix[1], wx[0], wx[1]);
  cic_axis(y[s], n, iy[0], iy[1], wy[0], wy[1]);
  cic_axis(z[s], n, iz[0], iz[1], wz[0], wz[1]);
  const double dw[2] = {-(double)n, (double)n};
  double gx = 0.0, gy = 0.0, gz = 0.0;
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        const double d = 2.0 * scale * delta[((size_t)ix[a] * n + iy[b]) * n + iz[c]];
        gx += d * dw[a] * wy[b] * wz[c];
        gy += d * wx[a] * dw[b] * wz[c];
        gz += d * wx[a] * wy[b] * dw[c];
      }
  lam[s] = gx;
  lam[(size_t)N + s] = gy;
  lam[2 * (size_t)N + s] = gz;
  lam[3 * (size_t)N + s] = 0.0;
  lam[4 * (size_t)N + s] = 0.0;
  lam[5 * (size_t)N + s] = 0.0;
}

// ---------------------------------------------------------------- leapfrog adjoints
__global__ void k_adj_drift(int N, double Cxdt, double *__restrict__ lam) {
  int s = blockIdx.x * 256 + threadIdx.x;
  if (s >= N) return;
  for (int d = 0; d < 3; ++d) lam[(size_t)(3 + d) * N + s] += Cxdt * lam[(size_t)d * N + s];
}

// lambda_x += Cv dt H_o lambda_v; with acc, dg_o += u and dH_o += du (u =
// Cv dt lambda_v, dx = x - c_o as in the interpolation)
__global__ void __launch_bounds__(256) k_adj_kick(int N, int M, const int *__restrict__ owner,
                                                  const double *__restrict__ grad,
                                                  const double *__restrict__ x,
                                                  const double *__restrict__ y,
                                                  const double *__restrict__ z,
                                                  const double *__restrict__ cx,
                                                  const double *__restrict__ cy,
                                                  const double *__restrict__ cz, double Cvdt,
                                                  double *__restrict__ lam, int acc,
                                                  double *__restrict__ dfield, int wrapdx,
                                                  double lx, double ly, double lz) {
  int s = blockIdx.x * 256 + threadIdx.x;
  if (s >= N) return;
  const int o = owner[s];
  const double hxx = grad[3 * (size_t)M + o], hyy = grad[4 * (size_t)M + o],
               hzz = grad[5 * (size_t)M + o];
  const double hxy = grad[6 * (size_t)M + o], hxz = grad[7 * (size_t)M + o],
               hyz = grad[8 * (size_t)M + o];
  const double u0 = Cvdt * lam[3 * (size_t)N + s], u1 = Cvdt * lam[4 * (size_t)N + s],
               u2 = Cvdt * lam[5 * (size_t)N + s];
  lam[s] += hxx * u0 + hxy * u1 + hxz * u2;
  lam[(size_t)N + s] += hxy * u0 + hyy * u1 + hyz * u2;
  lam[2 * (size_t)N + s] += hxz * u0 + hyz * u1 + hzz * u2;
  if (!acc) return;
  double dx = x[s] - cx[o], dy = y[s] - cy[o], dz = z[s] - cz[o];
  if (wrapdx) {  // as the forward interpolation with a fixed periodic topology
    dx -= lx * rint(dx / lx);
    dy -= ly * rint(dy / ly);
    dz -= lz * rint(dz / lz);
  }
  atomicAdd(&dfield[o], u0);
  atomicAdd(&dfield[(size_t)M + o], u1);
  atomicAdd(&dfield[2 * (size_t)M + o], u2);
  atomicAdd(&dfield[3 * (size_t)M + o], u0 * dx);
  atomicAdd(&dfield[4 * (size_t)M + o], u1 * dy);
  atomicAdd(&dfield[5 * (size_t)M + o], u2 * dz);
  atomicAdd(&dfield[6 * (size_t)M + o], u0 * dy + u1 * dx);
  atomicAdd(&dfield[7 * (size_t)M + o], u0 * dz + u2 * dx);
  atomicAdd(&dfield[8 * (size_t)M + o], u1 * dz + u2 * dy);
}

// ---------------------------------------------------------------- transposes
__global__ void k_tr_keys(long ne, int M, const int *__restrict__ nbr, uint64_t *__restrict__ key,
                          int *__restrict__ val) {
  long p = (long)blockIdx.x * 256 + threadIdx.x;
  if (p >= ne) return;
  int j = nbr[p];
  key[p] = (uint64_t)((j < M) ? j : M);
  val[p] = (int)p;
}

// trp[j] = first sorted position whose key is >= j, j = 0..M
__global__ void k_tr_rowptr(long ne, int M, const uint64_t *__restrict__ key, int *__restrict__ trp) {
  int j = blockIdx.x * 256 + threadIdx.x;
  if (j > M) return;
  long lo = 0, hi = ne;
  while (lo < hi) {
    long mid = (lo + hi) >> 1;
    if (key[mid] < (uint64_t)j) lo = mid + 1;
    else hi = mid;
  }
  trp[j] = (int)lo;
}

int build_transpose(apc_ctx *c) {
  if (c->have_tr) return 0;
  const int M = c->nc + c->nv, lm = c->prm.lm;
  const long ne = (long)M * lm;
  if (ne >= INT32_MAX) return fail(c, APC_ERR_UNSUPPORTED, "transpose: too many entries");
  APC_ALLOC(c, c->tr_key, ne + 1);
  APC_ALLOC(c, c->tr_key2, ne + 1);
  APC_ALLOC(c, c->tr_pos, ne + 1);
  APC_ALLOC(c, c->tr_pos2, ne + 1);
  APC_ALLOC(c, c->tr_rp, M + 2);
  hipLaunchKernelGGL(k_tr_keys, dim3(nblocks(ne, 256)), dim3(256), 0, c->s, ne, M, c->nbr.p,
                     c->tr_key.p, c->tr_pos.p);
  int bits = 1;
  while (bits < 63 && ((uint64_t)M >> bits) != 0) ++bits;
  int r = radix_sort_pairs(c, c->tr_key.p, c->tr_pos.p, c->tr_key2.p, c->tr_pos2.p, (int)ne, bits);
  if (r) return r;
  hipLaunchKernelGGL(k_tr_rowptr, dim3(nblocks(M + 1, 256)), dim3(256), 0, c->s, ne, M,
                     c->tr_key.p, c->tr_rp.p);
  APC_HIP(c, hipGetLastError());
  c->have_tr = 1;
  return 0;
}

// Transpose of Differentiate (solver.cpp:295-362): for row i the fit is
// cf = R^-1 R^-T B^T W^2 (v - phi_i) c_field with R^T R = (W B)^T (W B) (the
// Cholesky factor the forward kernel k_gradient uses), g =
// cf[0..2] (1/h), H = cf[3..8] (1/h)^2.  Given dG = (dg, dH) per row, z = R^-1 R^-T
// dcf and each neighbour j receives c_field w_j^2 (b_j . z); the row's own
// phi_i receives minus their sum.  Scattered with atomics.
__global__ void __launch_bounds__(256) k_grad_adjoint(
    int M, int lm, Geo G, double c_field, const double *__restrict__ cx,
    const double *__restrict__ cy, const double *__restrict__ cz, const double *__restrict__ ch,
    const int *__restrict__ nbr, const double *__restrict__ dist, const double *__restrict__ dG,
    double *__restrict__ dphi, int *__restrict__ status) {
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= M) return;
  const FitRow f = fit_row(cx, cy, cz, ch, dist, M, i, lm);
  double mo[kMoments];
#pragma unroll
  for (int q = 0; q < kMoments; ++q) mo[q] = 0.0;
  for_slots(M, lm, i, nbr, dist, cx, cy, cz, NoExtra{},
            [&](int, int, double d, double x, double y, double z, double) {
              double p[3];
              const double w = fit_offset(G, f, d, x, y, z, p);
              moments_add(mo, w * w, p[0], p[1], p[2]);
            });
  double R[45];
  gram_from_moments(mo, R);
  if (chol_inplace(R)) { status[0] = APC_ERR_RANK; status[1] = i; return; }
  const double ih2 = f.ih * f.ih;
  double dcf[9];
  dcf[0] = dG[i] * f.ih;
  dcf[1] = dG[(size_t)M + i] * f.ih;
  dcf[2] = dG[2 * (size_t)M + i] * f.ih;
#pragma unroll
  for (int q = 3; q < 9; ++q) dcf[q] = dG[(size_t)q * M + i] * ih2;
  // y = R^-T dcf (forward substitution), z = R^-1 y (back substitution)
  double yv[9], zv[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    double s = dcf[k];
#pragma unroll
    for (int q = 0; q < k; ++q) s -= R[RI(q, k)] * yv[q];
    yv[k] = s / R[RI(k, k)];
  }
#pragma unroll
  for (int k = 8; k >= 0; --k) {
    double s = yv[k];
#pragma unroll
    for (int q = k + 1; q < 9; ++q) s -= R[RI(k, q)] * zv[q];
    zv[k] = s / R[RI(k, k)];
  }
  double self = 0.0;
  for_slots(M, lm, i, nbr, dist, cx, cy, cz, NoExtra{},
            [&](int, int m, double d, double x, double y, double z, double) {
              double r[9];
              const double w = fit_slot(G, f, d, x, y, z, r);
              double bz = 0.0;
#pragma unroll
              for (int k = 0; k < 9; ++k) bz += r[k] * zv[k];
              const double coef = c_field * w * w * bz;
              if (m < M) atomicAdd(&dphi[m], coef);
              self -= coef;
            });
  atomicAdd(&dphi[i], self);
}

// gather caller-order arrays into device (sorted) order and back
__global__ void k_gather6(int N, const int *__restrict__ id, const double *__restrict__ src,
                          double *__restrict__ dst) {
  int s = blockIdx.x * 256 + threadIdx.x;
  if (s >= N) return;
  const int k = id[s];
  for (int d = 0; d < 6; ++d) dst[(size_t)d * N + s] = src[(size_t)d * N + k];
}
__global__ void k_scatter6(int N, const int *__restrict__ id, const double *__restrict__ src,
                           double *__restrict__ dst) {
  int s = blockIdx.x * 256 + threadIdx.x;
  if (s >= N) return;
  const int k = id[s];
  for (int d = 0; d < 6; ++d) dst[(size_t)d * N + k] = src[(size_t)d * N + s];
}

}  // namespace apc

using namespace apc;

static int need_frozen(apc_ctx *c, const char *what) {
  if (!c->frozen || !c->have_comp) return fail(c, APC_ERR_ARG, "%s: freeze the topology first", what);
  return 0;
}

extern "C" int apc_freeze_topology(apc_ctx *c, int on) {
  if (!c) return APC_ERR_ARG;
  if (on && (!c->have_comp || !c->sorted || c->comp_external))
    return fail(c, APC_ERR_ARG, "freeze_topology: build the octree and select first");
  if (on && c->nbox > 1)
    return fail(c, APC_ERR_UNSUPPORTED, "freeze_topology: not on a packed multi-box context");
  c->frozen = on ? 1 : 0;
  if (!on) c->have_tr = 0;
  return 0;
}

// caller-order (x, y, z, vx, vy, vz) into the frozen device order
extern "C" int apc_set_state(apc_ctx *c, const double *x, const double *y, const double *z,
                             const double *vx, const double *vy, const double *vz) {
  if (!c || !x || !y || !z || !vx || !vy || !vz) return APC_ERR_ARG;
  if (int r = need_frozen(c, "set_state")) return r;
  const int N = c->N;
  APC_ALLOC(c, c->adj_tmp, (size_t)6 * N + 6);
  APC_ALLOC(c, c->adj_tmp2, (size_t)6 * N + 6);
  const double *src[6] = {x, y, z, vx, vy, vz};
  for (int d = 0; d < 6; ++d)
    APC_HIP(c, hipMemcpyAsync(c->adj_tmp.p + (size_t)d * N, src[d], sizeof(double) * N,
                              hipMemcpyHostToDevice, c->s));
  hipLaunchKernelGGL(k_gather6, dim3(nblocks(N, 256)), dim3(256), 0, c->s, N, c->id.p,
                     c->adj_tmp.p, c->adj_tmp2.p);
  double *dst[6] = {c->x.p, c->y.p, c->z.p, c->vx.p, c->vy.p, c->vz.p};
  for (int d = 0; d < 6; ++d)
    APC_HIP(c, hipMemcpyAsync(dst[d], c->adj_tmp2.p + (size_t)d * N, sizeof(double) * N,
                              hipMemcpyDeviceToDevice, c->s));
  APC_HIP(c, hipStreamSynchronize(c->s));
  return 0;
}

// With the topology frozen the device order never changes, so (x, v) is the
// whole state.  Otherwise every step re-sorts the particles, and the slot
// also keeps their caller indices; a restore then drops the tree and
// everything built on it, as a drift does.
extern "C" int apc_checkpoint(apc_ctx *c, int slot, int restore) {
  if (!c || slot < 0 || slot >= 64) return APC_ERR_ARG;
  const int N = c->N;
  if (N < 1) return fail(c, APC_ERR_ARG, "checkpoint: no particles");
  if (c->frozen && !c->have_comp) return fail(c, APC_ERR_ARG, "checkpoint: frozen without a topology");
  double *st[6] = {c->x.p, c->y.p, c->z.p, c->vx.p, c->vy.p, c->vz.p};
  if (!restore) {
    if ((int)c->ckpt.size() <= slot) c->ckpt.resize(slot + 1);
    if ((int)c->ckpt_id.size() <= slot) c->ckpt_id.resize(slot + 1);
    APC_ALLOC(c, c->ckpt[slot], (size_t)6 * N + 6);
    APC_ALLOC(c, c->ckpt_id[slot], (size_t)N + 1);
    for (int d = 0; d < 6; ++d)
      APC_HIP(c, hipMemcpyAsync(c->ckpt[slot].p + (size_t)d * N, st[d], sizeof(double) * N,
                                hipMemcpyDeviceToDevice, c->s));
    APC_HIP(c, hipMemcpyAsync(c->ckpt_id[slot].p, c->id.p, sizeof(int) * N,
                              hipMemcpyDeviceToDevice, c->s));
    c->ckpt_n.resize(c->ckpt.size(), -1);
    c->ckpt_n[slot] = N;
  } else {
    if ((int)c->ckpt.size() <= slot || !c->ckpt[slot].p || c->ckpt_n[slot] != N)
      return fail(c, APC_ERR_ARG, "checkpoint %d was not saved", slot);
    for (int d = 0; d < 6; ++d)
      APC_HIP(c, hipMemcpyAsync(st[d], c->ckpt[slot].p + (size_t)d * N, sizeof(double) * N,
                                hipMemcpyDeviceToDevice, c->s));
    if (!c->frozen) {
      APC_HIP(c, hipMemcpyAsync(c->id.p, c->ckpt_id[slot].p, sizeof(int) * N,
                                hipMemcpyDeviceToDevice, c->s));
      c->sorted = false;
      c->have_comp = false;
      c->have_nbr = c->have_sys = 0;
    }
    c->have_fields = 0;
  }
  return 0;
}

extern "C" int apc_adjoint_seed_cic(apc_ctx *c, int n, double *loss) {
  if (!c || n < 2 || n > 1024 || !loss) return APC_ERR_ARG;
  if (int r = need_frozen(c, "adjoint_seed_cic")) return r;
  const int N = c->N;
  if (N > (1 << 28)) return fail(c, APC_ERR_ARG, "adjoint_seed_cic: more than 2^28 particles");
  const long ng = (long)n * n * n;
  const long ne = 8L * N;
  constexpr int NB = 1024;
  APC_ALLOC(c, c->cic, ng + 1);
  APC_ALLOC(c, c->lam, (size_t)6 * N + 6);
  APC_ALLOC(c, c->red_d, 2 * NB + 1);
  APC_ALLOC(c, c->cic_key, ne + 1);
  APC_ALLOC(c, c->cic_key2, ne + 1);
  APC_ALLOC(c, c->cic_v, ne + 1);
  APC_ALLOC(c, c->cic_v2, ne + 1);
  APC_ALLOC(c, c->cic_rng, 2 * ng + 1);
  const double *wts = c->have_cic_w ? c->cic_w.p : nullptr;
  const double scale = (double)ng / (wts ? c->cic_wsum : (double)N);  // unit mean density
  int bits = 1;
  while ((1L << bits) < ng) ++bits;
  {
    KScope ks(c, "cic");
    hipLaunchKernelGGL(k_cic_keys, dim3(nblocks(ne, 256)), dim3(256), 0, c->s, N, n, c->x.p,
                       c->y.p, c->z.p, c->cic_key.p, c->cic_v.p);
    int r = radix_sort_pairs(c, c->cic_key.p, c->cic_v.p, c->cic_key2.p, c->cic_v2.p, (int)ne, bits);
    if (r) return r;
    hipLaunchKernelGGL(k_fill_zero, dim3(1024), dim3(256), 0, c->s, (size_t)ng,
                       reinterpret_cast<double *>(c->cic_rng.p));  // 2 ints per cell = 1 double
    hipLaunchKernelGGL(k_cic_ranges, dim3(nblocks(ne, 256)), dim3(256), 0, c->s, ne,
                       c->cic_key.p, c->cic_rng.p);
    hipLaunchKernelGGL(k_cic_cells, dim3(NB), dim3(256), 0, c->s, ng, n, c->x.p, c->y.p, c->z.p,
                       scale, wts, c->cic_v.p, c->cic_rng.p, c->cic.p, c->red_d.p);
    hipLaunchKernelGGL(k_cic_grad, dim3(nblocks(N, 256)), dim3(256), 0, c->s, N, n, c->x.p, c->y.p,
                       c->z.p, scale, wts, c->cic.p, c->lam.p);
  }
  std::vector<double> part(2 * NB);
  APC_HIP(c, hipMemcpyAsync(part.data(), c->red_d.p, sizeof(double) * 2 * NB,
                            hipMemcpyDeviceToHost, c->s));
  APC_HIP(c, hipStreamSynchronize(c->s));
  double hi = 0.0, lo = 0.0;  // compensated sum of the block partials
  for (int b = 0; b < NB; ++b) {
    const double v = part[2 * b], s = hi + v, bb = s - hi;
    lo += (hi - (s - bb)) + (v - bb) + part[2 * b + 1];
    hi = s;
  }
  *loss = hi + lo;
  APC_HIP(c, hipGetLastError());
  return 0;
}

__global__ void k_gather1(int N, const int *__restrict__ id, const double *__restrict__ src,
                          double *__restrict__ dst) {
  int s = blockIdx.x * 256 + threadIdx.x;
  if (s < N) dst[s] = src[id[s]];
}

extern "C" int apc_adjoint_cic_weights(apc_ctx *c, const double *w) {
  if (!c) return APC_ERR_ARG;
  if (int r = need_frozen(c, "adjoint_cic_weights")) return r;
  if (!w) {
    c->have_cic_w = 0;
    return 0;
  }
  const int N = c->N;
  double sum = 0.0;
  for (int s = 0; s < N; ++s) {
    if (!(w[s] >= 0.0) || !std::isfinite(w[s]))
      return fail(c, APC_ERR_ARG, "adjoint_cic_weights: weight %d is negative or not finite", s);
    sum += w[s];
  }
  if (!(sum > 0.0)) return fail(c, APC_ERR_ARG, "adjoint_cic_weights: the weights sum to zero");
  APC_ALLOC(c, c->adj_tmp, (size_t)N + 1);
  APC_ALLOC(c, c->cic_w, (size_t)N + 1);
  APC_HIP(c, hipMemcpyAsync(c->adj_tmp.p, w, sizeof(double) * N, hipMemcpyHostToDevice, c->s));
  hipLaunchKernelGGL(k_gather1, dim3(nblocks(N, 256)), dim3(256), 0, c->s, N, c->id.p,
                     c->adj_tmp.p, c->cic_w.p);
  APC_HIP(c, hipStreamSynchronize(c->s));
  c->cic_wsum = sum;
  c->have_cic_w = 1;
  return 0;
}

extern "C" int apc_adjoint_get(apc_ctx *c, double *lx, double *ly, double *lz, double *lvx,
                               double *lvy, double *lvz) {
  if (!c || !lx || !ly || !lz || !lvx || !lvy || !lvz) return APC_ERR_ARG;
  if (int r = need_frozen(c, "adjoint_get")) return r;
  const int N = c->N;
  if (c->lam.cap < (size_t)6 * N) return fail(c, APC_ERR_ARG, "adjoint_get: no adjoint state");
  APC_ALLOC(c, c->adj_tmp, (size_t)6 * N + 6);
  hipLaunchKernelGGL(k_scatter6, dim3(nblocks(N, 256)), dim3(256), 0, c->s, N, c->id.p, c->lam.p,
                     c->adj_tmp.p);
  double *dst[6] = {lx, ly, lz, lvx, lvy, lvz};
  for (int d = 0; d < 6; ++d)
    APC_HIP(c, hipMemcpyAsync(dst[d], c->adj_tmp.p + (size_t)d * N, sizeof(double) * N,
                              hipMemcpyDeviceToHost, c->s));
  APC_HIP(c, hipStreamSynchronize(c->s));
  return 0;
}

extern "C" int apc_adjoint_set(apc_ctx *c, const double *lx, const double *ly, const double *lz,
                               const double *lvx, const double *lvy, const double *lvz) {
  if (!c || !lx || !ly || !lz || !lvx || !lvy || !lvz) return APC_ERR_ARG;
  if (int r = need_frozen(c, "adjoint_set")) return r;
  const int N = c->N;
  APC_ALLOC(c, c->lam, (size_t)6 * N + 6);
  APC_ALLOC(c, c->adj_tmp, (size_t)6 * N + 6);
  const double *src[6] = {lx, ly, lz, lvx, lvy, lvz};
  for (int d = 0; d < 6; ++d)
    APC_HIP(c, hipMemcpyAsync(c->adj_tmp.p + (size_t)d * N, src[d], sizeof(double) * N,
                              hipMemcpyHostToDevice, c->s));
  hipLaunchKernelGGL(k_gather6, dim3(nblocks(N, 256)), dim3(256), 0, c->s, N, c->id.p,
                     c->adj_tmp.p, c->lam.p);
  APC_HIP(c, hipStreamSynchronize(c->s));
  return 0;
}

extern "C" int apc_adjoint_drift(apc_ctx *c, double Cx, double dt) {
  if (!c) return APC_ERR_ARG;
  if (int r = need_frozen(c, "adjoint_drift")) return r;
  const int N = c->N;
  if (c->lam.cap < (size_t)6 * N) return fail(c, APC_ERR_ARG, "adjoint_drift: no adjoint state");
  KScope ks(c, "adjoint_drift", 72.0 * N);
  hipLaunchKernelGGL(k_adj_drift, dim3(nblocks(N, 256)), dim3(256), 0, c->s, N, Cx * dt, c->lam.p);
  APC_HIP(c, hipGetLastError());
  return 0;
}

// Kick adjoint with the node fields of the last solve; acc != 0 also adds the
// leaf-field adjoints (read with apc_adjoint_fields, cleared by clear != 0).
extern "C" int apc_adjoint_kick(apc_ctx *c, double Cv, double dt, int acc, int clear) {
  if (!c) return APC_ERR_ARG;
  if (int r = need_frozen(c, "adjoint_kick")) return r;
  if (!c->have_fields) return fail(c, APC_ERR_ARG, "adjoint_kick: no node fields (solve first)");
  const int N = c->N, M = c->nc + c->nv;
  if (c->lam.cap < (size_t)6 * N) return fail(c, APC_ERR_ARG, "adjoint_kick: no adjoint state");
  APC_ALLOC(c, c->dfield, (size_t)9 * M + 9);
  if (clear)
    hipLaunchKernelGGL(k_fill_zero, dim3(256), dim3(256), 0, c->s, (size_t)9 * M, c->dfield.p);
  KScope ks(c, "adjoint_kick", 100.0 * N + 96.0 * M);
  hipLaunchKernelGGL(k_adj_kick, dim3(nblocks(N, 256)), dim3(256), 0, c->s, N, M, c->owner.p,
                     c->grad.p, c->x.p, c->y.p, c->z.p, c->cx.p, c->cy.p, c->cz.p, Cv * dt,
                     c->lam.p, acc, c->dfield.p, c->prm.bc == 1 ? 1 : 0, c->prm.bbox[1] - c->prm.bbox[0],
                     c->prm.bbox[3] - c->prm.bbox[2], c->prm.bbox[5] - c->prm.bbox[4]);
  APC_HIP(c, hipGetLastError());
  return 0;
}

extern "C" int apc_adjoint_fields(apc_ctx *c, double *dg, double *dH) {
  if (!c || !dg || !dH) return APC_ERR_ARG;
  const int M = c->nc + c->nv;
  if (c->dfield.cap < (size_t)9 * M) return fail(c, APC_ERR_ARG, "adjoint_fields: none accumulated");
  std::vector<double> t((size_t)9 * M);
  APC_HIP(c, hipMemcpyAsync(t.data(), c->dfield.p, sizeof(double) * 9 * M, hipMemcpyDeviceToHost,
                            c->s));
  APC_HIP(c, hipStreamSynchronize(c->s));
  for (int i = 0; i < M; ++i) {
    for (int d = 0; d < 3; ++d) dg[(size_t)i * 3 + d] = t[(size_t)d * M + i];
    for (int d = 0; d < 6; ++d) dH[(size_t)i * 6 + d] = t[(size_t)(3 + d) * M + i];
  }
  return 0;
}

// dphi = (dG/dphi)^T (dg, dH): host layouts as apc_get_fields (M x 3, M x 6).
extern "C" int apc_gradient_adjoint(apc_ctx *c, const double *dg, const double *dH, double *dphi) {
  if (!c || !dg || !dH || !dphi) return APC_ERR_ARG;
  if (!c->have_sys) return fail(c, APC_ERR_ARG, "gradient_adjoint: no system (solve first)");
  const int M = c->nc + c->nv, lm = c->prm.lm;
  std::vector<double> t((size_t)9 * M);
  for (int i = 0; i < M; ++i) {
    for (int d = 0; d < 3; ++d) t[(size_t)d * M + i] = dg[(size_t)i * 3 + d];
    for (int d = 0; d < 6; ++d) t[(size_t)(3 + d) * M + i] = dH[(size_t)i * 6 + d];
  }
  APC_ALLOC(c, c->dfield, (size_t)9 * M + 9);
  APC_ALLOC(c, c->adj_vec, (size_t)M + 1);
  APC_ALLOC(c, c->status_dev, 4);
  APC_HIP(c, hipMemcpyAsync(c->dfield.p, t.data(), sizeof(double) * 9 * M, hipMemcpyHostToDevice,
                            c->s));
  APC_HIP(c, hipMemsetAsync(c->status_dev.p, 0, 4 * sizeof(int), c->s));
  hipLaunchKernelGGL(k_fill_zero, dim3(256), dim3(256), 0, c->s, (size_t)M, c->adj_vec.p);
  {
    KScope ks(c, "gradient_adjoint", (double)M * (44.0 * lm + 112.0));
    hipLaunchKernelGGL(k_grad_adjoint, dim3(nblocks(M, 256)), dim3(256), 0, c->s, M, lm, geo_of(c),
                       c->c_field, c->cx.p, c->cy.p, c->cz.p, c->ch.p, c->nbr.p, c->dist.p,
                       c->dfield.p, c->adj_vec.p, c->status_dev.p);
  }
  int st[2] = {0, 0};
  APC_HIP(c, hipMemcpyAsync(st, c->status_dev.p, sizeof(st), hipMemcpyDeviceToHost, c->s));
  APC_HIP(c, hipMemcpyAsync(dphi, c->adj_vec.p, sizeof(double) * M, hipMemcpyDeviceToHost, c->s));
  APC_HIP(c, hipStreamSynchronize(c->s));
  if (st[0]) return fail(c, st[0], "gradient_adjoint: rank-deficient fit at row %d", st[1]);
  return 0;
}

// out = A^-T rhs for the density (which = 1) or Poisson (which = 2) matrix of
// the last system build; Jacobi GMRES on the explicit transpose.
extern "C" int apc_solve_adjoint(apc_ctx *c, int which, const apc_ksp_opts *o, const double *rhs,
                                 double *out, int *its, int *reason) {
  if (!c || !o || !rhs || !out || (which != 1 && which != 2)) return APC_ERR_ARG;
  if (!c->have_sys) return fail(c, APC_ERR_ARG, "solve_adjoint: no system (solve first)");
  if (o->pc_type != 0 || o->precision != 0)
    return fail(c, APC_ERR_UNSUPPORTED, "solve_adjoint: FP64 Jacobi GMRES only");
  int r = build_transpose(c);
  if (r) return r;
  const int M = c->nc + c->nv;
  APC_ALLOC(c, c->adj_vec, (size_t)M + 1);
  APC_ALLOC(c, c->adj_vec2, (size_t)M + 1);
  APC_HIP(c, hipMemcpyAsync(c->adj_vec.p, rhs, sizeof(double) * M, hipMemcpyHostToDevice, c->s));
  r = solve_transposed(c, which, o, c->adj_vec.p, c->adj_vec2.p, its, reason);
  if (r) return r;
  APC_HIP(c, hipMemcpyAsync(out, c->adj_vec2.p, sizeof(double) * M, hipMemcpyDeviceToHost, c->s));
  APC_HIP(c, hipStreamSynchronize(c->s));
  return 0;
}
