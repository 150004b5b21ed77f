// Restarted GMRES replacing PETSc's KSPGMRES + PCJACOBI (solver.cpp:182-293).
//
// Semantics follow PETSc's KSPSolve_GMRES: left Jacobi preconditioning, zero
// initial guess, convergence on the preconditioned residual with
// KSPConvergedDefault (rtol relative to the first residual norm, atol, dtol),
// classical Gram-Schmidt (PETSc's default orthogonalisation), Givens update of
// the Hessenberg matrix, true residual at every restart, -ksp_max_it counted
// across restarts.
//
// Device-driven: an Arnoldi step is five launches — preconditioned SpMV,
// multi-dot partials (h = V^T w), their reduction, "subtract projections +
// norm partials", and a one-block Hessenberg/Givens/stopping-test kernel.
// Every reduction is two-stage with a fixed summation order (deterministic).
// The reductions end in kernels of their own rather than in the last block
// to arrive: on this part a device-scope release fence per block writes back
// that XCD's L2 and the arrival counter serialises every block on one
// address, which cost more than the extra launches.  The host only launches
// and checks a status word every few steps.  Basis vectors are kept
// unnormalised with scale factors sig[q] (v_q = sig[q] w_q), which removes
// the normalisation pass.
//
// Precision form: the same cycle with an FP32 basis and FP32 copies of the
// matrix, inside FP64 iterative refinement — every restart computes the true
// residual with the FP64 operator, the FP32 cycle solves for the correction
// and the correction is accumulated into the FP64 solution.
#include <algorithm>
#include <cmath>
#include <cstdlib>

#include "apc_solve.h"

namespace apc {


// Fixed-order sum of nb partials by the calling block; result to all threads.
__device__ inline double sum_partials(const double *part, int nb, double *ws) {
  double a = 0.0;
  for (int b = threadIdx.x; b < nb; b += 256) a += part[b];
  return block_sum256(a, ws);
}


// ---- cycle start: v0 = D^-1 (b - A x) in FP64 (left Jacobi) or b - A x
// (right preconditioning), ||v0||, PETSc's check ----
template <class T>
__global__ void __launch_bounds__(256) k_resid(Op<double> A, const double *__restrict__ x,
                                               const double *__restrict__ b, int zero_guess,
                                               int left_pc, T *__restrict__ v0, GDev g) {
  __shared__ double ws[4];
  if (g.st->reason) return;
  double acc = 0.0;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < A.M; i += gridDim.x * 256) {
    double r = zero_guess ? b[i] : b[i] - row_apply(A, x, i);
    if (left_pc) r = A.dinv[i] * r;
    v0[i] = (T)r;
    acc += r * r;
  }
  acc = block_sum256(acc, ws);
  if (threadIdx.x == 0) g.part[blockIdx.x] = acc;
}

// one block: ||v0|| from the partials, cycle state, PETSc's first check
__global__ void __launch_bounds__(256) k_resid_fin(int nb, GDev g) {
  __shared__ double ws[4];
  GState *st = g.st;
  if (st->reason) return;
  double nrm2 = sum_partials(g.part, nb, ws);
  if (threadIdx.x == 0) {
    double res = sqrt(nrm2);
    st->res = res;
    st->it = 0;
    st->cycle_done = 0;
    st->cyc_res0 = res;
    g.grs[0] = res;
    g.sig[0] = (res != 0.0) ? 1.0 / res : 1.0;
    if (st->monitor) g.hist[st->its] = res;
    if (res == 0.0) st->reason = KSP_CONVERGED_ATOL;
    else st->reason = converged(st, st->its, res);
    if (!st->reason && st->its >= st->max_it) st->cycle_done = 1;
  }
}

// ---- w = sig_j D^-1 A w_j (left Jacobi) or w = sig_j A z_j ----
template <class T>
__global__ void __launch_bounds__(256) k_spmv(Op<T> A, const T *__restrict__ vj,
                                              T *__restrict__ w, int j, int left_pc, GDev g) {
  if (halted(g.st)) return;
  const T s = (T)g.sig[j];
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= A.M) return;
  const T av = (T)row_apply(A, vj, i);
  w[i] = left_pc ? (T)(s * (T)(A.dinv[i] * av)) : (T)(s * av);
}

// ---- flexible form: w = sig_j A z_j, z_j = M^-1 w_j computed in FP32 ----
// (z is the cycle's fixed output buffer; it is also stored as z_j in zj)
__global__ void __launch_bounds__(256) k_spmv_z(Op<double> A, const float *__restrict__ z,
                                                double *__restrict__ w, float *__restrict__ zj,
                                                int j, GDev g) {
  if (halted(g.st)) return;
  const double s = g.sig[j];
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= A.M) return;
  const double acc = row_offdiag_sum<double, float>(A, z, i);
  const float zi = z[i];
  zj[i] = zi;
  w[i] = s * (A.d[i] * (double)zi + A.f[i] * acc);
}

template <class T>
__global__ void k_copy_gated(int n, const T *__restrict__ a, T *__restrict__ b, const GState *st) {
  if (halted(st)) return;
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) b[i] = a[i];
}

__global__ void k_to_float_gated(int n, const double *__restrict__ a, float *__restrict__ b,
                                 const GState *st) {
  if (halted(st)) return;
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) b[i] = (float)a[i];
}

// ---- h_q = sig_q (w_q . w), q = 0..nv-1 ----
// Grid-stride over row tiles of 256*R rows.  A thread keeps its R entries of
// w in registers and walks every basis vector over the tile, U vectors at a
// time (R*U independent loads in flight); per-vector wave sums accumulate in
// LDS per wave, in tile order.  Partials go to part[q*nb + block].
template <class T, int R, int U>
__global__ void __launch_bounds__(256) k_mdot(int M, int nv, const T *__restrict__ V, size_t ld,
                                              const T *__restrict__ w, GDev g) {
  extern __shared__ double lacc[];  // [4][nv]
  if (halted(g.st)) return;
  const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
  const int nb = gridDim.x;
  for (int q = t; q < 4 * nv; q += 256) lacc[q] = 0.0;
  __syncthreads();
  double *la = lacc + wave * nv;
  const long ntile = ((long)M + 256 * R - 1) / (256 * R);
  for (long tile = blockIdx.x; tile < ntile; tile += nb) {
    const long base = tile * 256 * R + t;
    double wr[R];
    if ((tile + 1) * 256 * R <= M) {
#pragma unroll
      for (int r = 0; r < R; ++r) wr[r] = (double)w[base + r * 256];
      int q = 0;
      for (; q + U <= nv; q += U) {
        double a[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const T *vq = V + (size_t)(q + u) * ld + base;
          a[u] = 0.0;
#pragma unroll
          for (int r = 0; r < R; ++r) a[u] += wr[r] * (double)vq[r * 256];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const double sum = wave_sum(a[u]);
          if (lane == 0) la[q + u] += sum;
        }
      }
      for (; q < nv; ++q) {
        const T *vq = V + (size_t)q * ld + base;
        double a = 0.0;
#pragma unroll
        for (int r = 0; r < R; ++r) a += wr[r] * (double)vq[r * 256];
        a = wave_sum(a);
        if (lane == 0) la[q] += a;
      }
    } else {  // ragged last tile
#pragma unroll
      for (int r = 0; r < R; ++r) wr[r] = (base + r * 256 < M) ? (double)w[base + r * 256] : 0.0;
      for (int q = 0; q < nv; ++q) {
        const T *vq = V + (size_t)q * ld;
        double a = 0.0;
#pragma unroll
        for (int r = 0; r < R; ++r)
          if (base + r * 256 < M) a += wr[r] * (double)vq[base + r * 256];
        a = wave_sum(a);
        if (lane == 0) la[q] += a;
      }
    }
  }
  __syncthreads();
  for (int q = t; q < nv; q += 256)
    g.part[(size_t)q * nb + blockIdx.x] =
        (lacc[q] + lacc[nv + q]) + (lacc[2 * nv + q] + lacc[3 * nv + q]);
}

// h_q = sig_q * sum of the nb partials of vector q; one wave per vector
__global__ void __launch_bounds__(256) k_mdot_fin(int nv, int nb, GDev g) {
  if (halted(g.st)) return;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int q = blockIdx.x * 4 + wave;
  if (q >= nv) return;
  const double *pq = g.part + (size_t)q * nb;
  double a = 0.0;
  for (int b = lane; b < nb; b += 64) a += pq[b];
  a = wave_sum(a);
  if (lane == 0) g.h[q] = g.sig[q] * a;
}

// ---- w -= sum_q h_q sig_q w_q ; ||w||^2 partials ----
// Grid-stride over rows; the update is accumulated in FP64 whatever the
// storage type, U basis vectors per step (U independent loads in flight per
// row).  w32 (flexible form, may be null): the FP32 copy of the new w, the
// next step's multigrid input.
template <class T, int U>
__global__ void __launch_bounds__(256) k_maxpy(int M, int nv, const T *__restrict__ V, size_t ld,
                                               T *__restrict__ w, float *__restrict__ w32,
                                               GDev g) {
  extern __shared__ double cq[];  // nv coefficients
  __shared__ double ws[4];
  if (halted(g.st)) return;
  const int t = threadIdx.x;
  for (int q = t; q < nv; q += 256) cq[q] = g.h[q] * g.sig[q];
  __syncthreads();
  const long stride = (long)gridDim.x * 256;
  const int nvU = nv - nv % U;
  double acc = 0.0;
  for (long i = (long)blockIdx.x * 256 + t; i < M; i += stride) {
    double s = (double)w[i];
    const T *vi = V + i;
    int q = 0;
    for (; q < nvU; q += U) {
      double a[U];
#pragma unroll
      for (int u = 0; u < U; ++u) a[u] = (double)vi[(size_t)(q + u) * ld];
#pragma unroll
      for (int u = 0; u < U; ++u) s -= cq[q + u] * a[u];
    }
    for (; q < nv; ++q) s -= cq[q] * (double)vi[(size_t)q * ld];
    const T sv = (T)s;
    w[i] = sv;
    if (w32) w32[i] = (float)sv;
    acc += (double)sv * (double)sv;
  }
  acc = block_sum256(acc, ws);
  if (t == 0) g.part[blockIdx.x] = acc;
}

// ---- one block: ||w||, Hessenberg column j, Givens, stopping test ----
// The column and the previous rotations are staged in LDS by the whole
// block; the rotation chain is sequential and runs on one thread out of LDS.
__global__ void __launch_bounds__(256) k_hess(int nb, GDev g) {
  extern __shared__ double sh[];  // column [j+2], cc [j+1], ss [j+1]
  __shared__ double ws[4];
  GState *st = g.st;
  if (halted(st)) return;
  const double nrm2 = sum_partials(g.part, nb, ws);
  const int m = st->m;
  const int j = st->it;
  double *hc = sh, *cs = sh + (j + 2), *sn = cs + (j + 1);
  for (int q = threadIdx.x; q <= j; q += 256) {
    hc[q] = g.h[q];
    cs[q] = g.cc[q];
    sn[q] = g.ss[q];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    // KSPGMRESCycle after the orthogonalisation (CGS coefficients in h)
    const double tt = sqrt(nrm2);
    hc[j + 1] = tt;
    g.sig[j + 1] = (tt != 0.0) ? 1.0 / tt : 1.0;
    double hapbnd = fabs(tt / g.grs[j]);
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
        g.cc[j] = cj;
        g.ss[j] = sj;
        const double gj = g.grs[j];
        g.grs[j + 1] = -(sj * gj);
        g.grs[j] = cj * gj;
        hc[j] = cj * hc[j] + sj * hc[j + 1];
        res = fabs(g.grs[j + 1]);
      }
    }
    if (null_col) {
      st->reason = KSP_DIVERGED_NULL;
    } else {
      st->it = j + 1;
      st->its += 1;
      st->res = res;
      if (st->monitor) g.hist[st->its] = res;
      reason = converged(st, st->its, res);
      if (hapend && !reason) reason = KSP_DIVERGED_BREAKDOWN;
      if (st->fp32 && (reason == KSP_CONVERGED_RTOL || reason == KSP_CONVERGED_ATOL || hapend)) {
        // an FP32 estimate only ends the cycle; the FP64 residual at the next
        // restart decides convergence
        reason = 0;
        st->cycle_done = 1;
      }
      st->reason = reason;
      if (!reason) {
        // FP32 cycles stop at their precision floor and hand back to the FP64
        // refinement; FP64 cycles run to the restart length like PETSc
        if (st->fp32 && res <= 2e-6 * st->cyc_res0) st->cycle_done = 1;
        if (st->it >= m || st->its >= st->max_it) st->cycle_done = 1;
      }
    }
  }
  __syncthreads();
  double *hh = g.HH + (size_t)j * (m + 1);
  for (int q = threadIdx.x; q <= j + 1; q += 256) hh[q] = hc[q];
}

// ---- KSPGMRESBuildSoln: coef_q = y_q sig_q ----
__global__ void k_buildsoln(GDev g, int top) {
  if (threadIdx.x != 0) return;
  GState *st = g.st;
  const int m = st->m;
  auto H = [&](int r, int c) { return g.HH[(size_t)c * (m + 1) + r]; };
  double *nrs = g.coef;
  if (H(top, top) == 0.0) { st->reason = KSP_DIVERGED_BREAKDOWN; top = -1; }
  else nrs[top] = g.grs[top] / H(top, top);
  for (int ii = 1; top >= 0 && ii <= top; ++ii) {
    int k = top - ii;
    double tt = g.grs[k];
    for (int q = k + 1; q <= top; ++q) tt = tt - H(k, q) * nrs[q];
    if (H(k, k) == 0.0) { st->reason = KSP_DIVERGED_BREAKDOWN; top = -1; break; }
    nrs[k] = tt / H(k, k);
  }
  for (int q = 0; q <= top; ++q) nrs[q] *= g.sig[q];
  for (int q = (top < 0 ? 0 : top + 1); q < m; ++q) nrs[q] = 0.0;
}

// ---- x += sum_q coef_q w_q ----
template <class T>
__global__ void __launch_bounds__(256) k_axpy_sol(int M, int nv, const T *__restrict__ V, size_t ld,
                                                  const double *__restrict__ coef,
                                                  double *__restrict__ x) {
  extern __shared__ double cs[];
  for (int q = threadIdx.x; q < nv; q += 256) cs[q] = coef[q];
  __syncthreads();
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < M; i += (long)gridDim.x * 256) {
    double s = 0.0;
    for (int q = 0; q < nv; ++q) s += cs[q] * (double)V[(size_t)q * ld + i];
    x[i] += s;
  }
}

// ---- u = sum_q coef_q w_q (right preconditioning: x += M^-1 u afterwards) ----
template <class T>
__global__ void __launch_bounds__(256) k_comb(int M, int nv, const T *__restrict__ V, size_t ld,
                                              const double *__restrict__ coef, T *__restrict__ u) {
  extern __shared__ double cs[];
  for (int q = threadIdx.x; q < nv; q += 256) cs[q] = coef[q];
  __syncthreads();
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < M; i += (long)gridDim.x * 256) {
    double s = 0.0;
    for (int q = 0; q < nv; ++q) s += cs[q] * (double)V[(size_t)q * ld + i];
    u[i] = (T)s;
  }
}

template <class T>
__global__ void k_add_to(int M, const T *__restrict__ z, double *__restrict__ x) {
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i < M) x[i] += (double)z[i];
}

__global__ void k_to_float(size_t n, const double *__restrict__ a, float *__restrict__ b) {
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256)
    b[i] = (float)a[i];
}

// sums of rho*h^3 and h^3 (solver.cpp:244-252), block partials
__global__ void __launch_bounds__(256) k_mass_sums(int M, const double *__restrict__ rho,
                                                   const double *__restrict__ h,
                                                   double *__restrict__ part) {
  __shared__ double ws[4];
  double a = 0.0, b = 0.0;
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < M; i += (long)gridDim.x * 256) {
    double h3 = h[i] * h[i] * h[i];
    a += rho[i] * h3;
    b += h3;
  }
  a = block_sum256(a, ws);
  b = block_sum256(b, ws);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = a;
    part[gridDim.x + blockIdx.x] = b;
  }
}

__global__ void __launch_bounds__(256) k_rhs2_periodic(int M, const double *__restrict__ rho,
                                                       const double *__restrict__ part, int nb,
                                                       double c_pot, double *__restrict__ rhs2) {
  __shared__ double ws[4];
  double a = 0.0, b = 0.0;
  for (int q = threadIdx.x; q < nb; q += 256) { a += part[q]; b += part[nb + q]; }
  a = block_sum256(a, ws);
  b = block_sum256(b, ws);
  double avg = a / b;
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < M; i += (long)gridDim.x * 256)
    rhs2[i] = (rho[i] - avg) * c_pot;
}

// Packed boxes (rows box-major, rstart[b] the first row of box b): the
// periodic right-hand side of every box with its own mean, one block per box.
__global__ void __launch_bounds__(256) k_rhs2_boxes(const int *__restrict__ rstart,
                                                    const double *__restrict__ rho,
                                                    const double *__restrict__ h, double c_pot,
                                                    double *__restrict__ rhs2) {
  __shared__ double ws[4];
  const int lo = rstart[blockIdx.x], hi = rstart[blockIdx.x + 1];
  double a = 0.0, b = 0.0;
  for (int i = lo + threadIdx.x; i < hi; i += 256) {
    const double h3 = h[i] * h[i] * h[i];
    a += rho[i] * h3;
    b += h3;
  }
  a = block_sum256(a, ws);
  b = block_sum256(b, ws);
  const double avg = a / b;
  for (int i = lo + threadIdx.x; i < hi; i += 256) rhs2[i] = (rho[i] - avg) * c_pot;
}

__global__ void k_rhs2_bounded(int M, const double *__restrict__ rho, const double *__restrict__ pbnd,
                               double c_pot, double *__restrict__ rhs2) {
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= M) return;
  rhs2[i] = rho[i] * c_pot + pbnd[i];
}

constexpr int RED_BLOCKS = 1024;
constexpr int MAX_RESTART = 1000;

// Launch shapes of the orthogonalisation kernels, from a sweep over grid
// sizes and unroll depths at 0.9M and 7.2M rows with 16-200 basis vectors
// (both ran at 5.2-5.8 TB/s on 7.2M rows).
template <class T>
struct Cfg;
template <>
struct Cfg<double> { static constexpr int MAXPY_U = 4; };
template <>
struct Cfg<float> { static constexpr int MAXPY_U = 8; };
constexpr int MDOT_R = 4, MDOT_U = 2;

static int env_int(const char *name, int dflt) {
  const char *v = getenv(name);
  return (v && *v) ? std::max(1, atoi(v)) : dflt;
}

static int ensure_state(apc_ctx *c, int m, int max_it, int M, bool fp32) {
  if (!c->gm) c->gm = new SolveState();
  SolveState *S = c->gm;
  if (fp32) APC_ALLOC(c, S->V32, (size_t)(m + 1) * M + 1);
  else APC_ALLOC(c, S->V64, (size_t)(m + 1) * M + 1);
  const int nbm = (int)nblocks(M, 256) + 8;
  APC_ALLOC(c, S->part, (size_t)(m + 2) * nbm + RED_BLOCKS * 2 + 64);
  size_t hsz = (size_t)(m + 1) * m + 6 * (size_t)(m + 2) + (size_t)max_it + 8;
  APC_ALLOC(c, S->hess, hsz);
  APC_ALLOC(c, S->st, 1);
  if (!S->hst) APC_HIP(c, hipHostMalloc((void **)&S->hst, sizeof(GState)));
  S->mcap = m;
  return 0;
}

static GDev gdev_of(SolveState *S, int m) {
  GDev g;
  g.st = S->st.p;
  double *p = S->hess.p;
  g.HH = p; p += (size_t)(m + 1) * m;
  g.cc = p; p += m + 2;
  g.ss = p; p += m + 2;
  g.grs = p; p += m + 2;
  g.sig = p; p += m + 2;
  g.h = p; p += m + 2;
  g.coef = p; p += m + 2;
  g.hist = p;
  g.part = S->part.p;
  return g;
}

static int read_state(apc_ctx *c, SolveState *S) {
  APC_HIP(c, hipMemcpyAsync(S->hst, S->st.p, sizeof(GState), hipMemcpyDeviceToHost, c->s));
  APC_HIP(c, hipStreamSynchronize(c->s));
  return 0;
}

// One GMRES solve.  amg: right preconditioning with the multigrid in
// S->amg; zb, ub are M-vectors of T for the preconditioned direction and the
// cycle's combination.  flex (T = double): the multigrid runs in FP32 on the
// FP32 operator A32 and its outputs are kept, one per Arnoldi step, in Z32
// ((m+1) x M floats) — flexible GMRES, so the inexact FP32 preconditioner
// costs no accuracy and the cycle's update is x += Z32 y; v32 and z32 are
// M-vectors (the cycle's fixed input and output, so it replays as a graph).
template <class T>
static int run_cycles(apc_ctx *c, const Op<double> &A64, const Op<T> &AT, T *V, const double *b,
                      double *x, const apc_ksp_opts &o, bool fp32, bool amg, T *zb, T *ub,
                      int *its_out, int *reason_out, double *rnorm_out, int *sweeps,
                      const Op<float> *A32 = nullptr, float *Z32 = nullptr, float *v32 = nullptr,
                      float *z32 = nullptr) {
  const bool flex = Z32 != nullptr;
  const int M = A64.M;
  const int m = std::min(std::max(1, o.restart), MAX_RESTART);
  SolveState *S = c->gm;
  GDev g = gdev_of(S, m);
  GState h0;
  memset(&h0, 0, sizeof(h0));
  h0.m = m;
  h0.max_it = o.max_it;
  h0.fp32 = fp32 ? 1 : 0;
  h0.monitor = o.monitor;
  h0.rtol = o.rtol;
  h0.atol = o.atol;
  h0.dtol = o.dtol;
  *S->hst = h0;
  APC_HIP(c, hipMemcpyAsync(S->st.p, S->hst, sizeof(GState), hipMemcpyHostToDevice, c->s));
  APC_HIP(c, hipMemsetAsync(x, 0, sizeof(double) * M, c->s));
  constexpr int MU = Cfg<T>::MAXPY_U;
  const size_t ld = (size_t)M;
  const unsigned gb = nblocks(M, 256);
  const long ntile = ((long)M + 256 * MDOT_R - 1) / (256 * MDOT_R);
  // grid sizes: the context's knobs (APC_KNOB_MDOT_BLOCKS / _MAXPY_BLOCKS)
  const int mdot_blocks = c->knob[APC_KNOB_MDOT_BLOCKS];
  const int maxpy_blocks = c->knob[APC_KNOB_MAXPY_BLOCKS];
  const unsigned gdot = (unsigned)std::min<long>(mdot_blocks, ntile);
  const unsigned gaxpy = (unsigned)std::min<long>(maxpy_blocks, gb);
  const unsigned gres = (unsigned)std::min<long>(RED_BLOCKS, gb);
  const double nnz = (double)M * A64.lm;
  const double bspmv = fp32 ? 8.0 * nnz + 20.0 * M : 12.0 * nnz + 40.0 * M;
  int itcount = 0, reason = 0, ncyc = 0, printed = 0;
  bool zero_guess = true;
  auto monitor_flush = [&](int its) {
    if (!o.monitor) return 0;
    std::vector<double> hist(its + 1);
    APC_HIP(c, hipMemcpyAsync(hist.data(), g.hist, sizeof(double) * (its + 1),
                              hipMemcpyDeviceToHost, c->s));
    APC_HIP(c, hipStreamSynchronize(c->s));
    for (int k = printed; k <= its; ++k) printf("%3d KSP Residual norm %14.12e \n", k, hist[k]);
    printed = its + 1;
    return 0;
  };
  for (;;) {
    {
      KScope ks(c, fp32 ? "gmres_resid64" : "gmres_resid", 12.0 * nnz + 48.0 * M + sizeof(T) * M);
      hipLaunchKernelGGL(k_resid<T>, dim3(gres), dim3(256), 0, c->s, A64, x, b, zero_guess ? 1 : 0,
                         amg ? 0 : 1, V, g);
    }
    hipLaunchKernelGGL(k_resid_fin, dim3(1), dim3(256), 0, c->s, (int)gres, g);
    ++ncyc;
    int it_launched = 0;
    const int budget = o.max_it - itcount;
    int r = read_state(c, S);
    if (r) return r;
    if (S->hst->reason) { reason = S->hst->reason; monitor_flush(S->hst->its); break; }
    int to_read = steps_to_next_read(*S->hst, c->timing);
    for (int j = 0; j < m && j < budget; ++j) {
      const T *vj = V + (size_t)j * ld;
      T *w = V + (size_t)(j + 1) * ld;
      if (flex) {
        float *zj = Z32 + (size_t)j * ld;
        if (j == 0)  // later steps: the previous step's k_maxpy wrote v32
          hipLaunchKernelGGL(k_to_float_gated, dim3(gb), dim3(256), 0, c->s, M,
                             (const double *)vj, v32, (const GState *)S->st.p);
        r = amg_apply<float>(c, S->amg, *A32, v32, z32, S->st.p);
        if (r) return r;
        KScope ks(c, "gmres_spmv_z", 12.0 * nnz + 32.0 * M);
        hipLaunchKernelGGL(k_spmv_z, dim3(gb), dim3(256), 0, c->s, A64, (const float *)z32,
                           (double *)w, zj, j, g);
      } else if (amg) {
        hipLaunchKernelGGL(k_copy_gated<T>, dim3(gb), dim3(256), 0, c->s, M, vj, ub,
                           (const GState *)S->st.p);
        r = amg_apply<T>(c, S->amg, AT, ub, zb, S->st.p);
        if (r) return r;
        KScope ks(c, fp32 ? "gmres_spmv32" : "gmres_spmv", bspmv);
        hipLaunchKernelGGL(k_spmv<T>, dim3(gb), dim3(256), 0, c->s, AT, (const T *)zb, w, j, 0, g);
      } else {
        KScope ks(c, fp32 ? "gmres_spmv32" : "gmres_spmv", bspmv);
        hipLaunchKernelGGL(k_spmv<T>, dim3(gb), dim3(256), 0, c->s, AT, vj, w, j, 1, g);
      }
      const int nv = j + 1;
      {
        KScope ks(c, fp32 ? "gmres_mdot32" : "gmres_mdot", (double)sizeof(T) * M * (nv + 1));
        hipLaunchKernelGGL((k_mdot<T, MDOT_R, MDOT_U>), dim3(gdot), dim3(256),
                           4 * nv * sizeof(double), c->s, M, nv, V, ld, w, g);
      }
      {
        KScope ks(c, "gmres_mdot_fin", 8.0 * nv * gdot + 16.0 * nv);
        hipLaunchKernelGGL(k_mdot_fin, dim3(nblocks(nv, 4)), dim3(256), 0, c->s, nv, (int)gdot, g);
      }
      {
        KScope ks(c, fp32 ? "gmres_maxpy32" : "gmres_maxpy", (double)sizeof(T) * M * (nv + 2) +
                                                                (flex ? 4.0 * M : 0.0));
        hipLaunchKernelGGL((k_maxpy<T, MU>), dim3(gaxpy), dim3(256), nv * sizeof(double), c->s, M,
                           nv, V, ld, w, flex ? v32 : nullptr, g);
      }
      {
        KScope ks(c, "gmres_hess");
        hipLaunchKernelGGL(k_hess, dim3(1), dim3(256), (3 * (size_t)j + 4) * sizeof(double), c->s,
                           (int)gaxpy, g);
      }
      ++it_launched;
      if (--to_read == 0) {
        r = read_state(c, S);
        if (r) return r;
        if (S->hst->reason || S->hst->cycle_done) break;
        to_read = steps_to_next_read(*S->hst, c->timing);
      }
    }
    r = read_state(c, S);
    if (r) return r;
    const int it = S->hst->it;
    reason = S->hst->reason;
    monitor_flush(S->hst->its);
    if (it > 0) {
      hipLaunchKernelGGL(k_buildsoln, dim3(1), dim3(64), 0, c->s, g, it - 1);
      if (flex) {  // x += sum_q y_q z_q
        KScope ks(c, "gmres_update_z", 4.0 * M * it + 16.0 * M);
        hipLaunchKernelGGL(k_axpy_sol<float>, dim3(gres), dim3(256), it * sizeof(double), c->s, M,
                           it, (const float *)Z32, ld, g.coef, x);
      } else if (amg) {  // x += M^-1 (sum_q y_q v_q); runs whatever the solve's state
        {
          KScope ks(c, fp32 ? "gmres_update32" : "gmres_update", (double)sizeof(T) * M * (it + 1));
          hipLaunchKernelGGL(k_comb<T>, dim3(gres), dim3(256), it * sizeof(double), c->s, M, it, V,
                             ld, g.coef, ub);
        }
        r = amg_apply<T>(c, S->amg, AT, ub, zb, nullptr);
        if (r) return r;
        hipLaunchKernelGGL(k_add_to<T>, dim3(gb), dim3(256), 0, c->s, M, (const T *)zb, x);
      } else {
        KScope ks(c, fp32 ? "gmres_update32" : "gmres_update", (double)sizeof(T) * M * it + 16.0 * M);
        hipLaunchKernelGGL(k_axpy_sol<T>, dim3(gres), dim3(256), it * sizeof(double), c->s, M, it,
                           V, ld, g.coef, x);
      }
      r = read_state(c, S);
      if (r) return r;
      if (!reason) reason = S->hst->reason;
    }
    itcount += it;
    (void)it_launched;
    if (reason) break;
    if (itcount >= o.max_it) { reason = KSP_DIVERGED_ITS; break; }
    zero_guess = false;
  }
  *its_out = S->hst->its;
  *reason_out = reason;
  *rnorm_out = S->hst->res;
  if (sweeps) *sweeps = ncyc;
  APC_HIP(c, hipGetLastError());
  return 0;
}

static AmgOptions amg_options(apc_ctx *c) {
  // APC_AMG_* override the multigrid settings (tuning runs); the lanes per
  // row of the coarse levels are the context's knobs
  AmgOptions a;
  a.nu = env_int("APC_AMG_NU", a.nu);
  a.gamma = env_int("APC_AMG_GAMMA", a.gamma);
  a.coarse_max = env_int("APC_AMG_COARSE", a.coarse_max);
  a.galerkin_sorted = env_int("APC_AMG_GALERKIN_SORTED", 0);
  if (const char *v = getenv("APC_AMG_SHIFT")) a.shift = atoi(v);
  a.g8_rows = c->knob[APC_KNOB_AMG_G8_ROWS];
  a.g4_rows = c->knob[APC_KNOB_AMG_G4_ROWS];
  a.rg8_rows = c->knob[APC_KNOB_AMG_RG8_ROWS];
  if (const char *v = getenv("APC_AMG_OMEGA")) a.omega = atof(v);
  if (const char *v = getenv("APC_AMG_ALPHA")) a.alpha = atof(v);
  if (const char *v = getenv("APC_AMG_ALPHA0")) a.alpha0 = atof(v);
  if (const char *v = getenv("APC_AMG_AP")) a.ap = atoi(v);
  if (const char *v = getenv("APC_AMG_GEO0")) a.geo0 = atoi(v);
  if (const char *v = getenv("APC_AMG_GEOC")) a.geoc = atoi(v);
  a.nopre = c->knob[APC_KNOB_AMG_NOPRE];
  a.wlevels = c->knob[APC_KNOB_AMG_WLEVELS];
  if (const char *v = getenv("APC_AMG_THETA")) a.theta = atof(v);
  return a;
}

// Solve A x = b (x overwritten) with the matrix (nbr, s, d, f, dinv).
// pinned: A is the periodic Poisson operator with its identity row.
static int gmres(apc_ctx *c, const Op<double> &A, const double *b, double *x,
                 const apc_ksp_opts &o, bool pinned, int *its, int *reason, double *rnorm,
                 int *sweeps, bool seg = false) {
  const int M = A.M;
  const int m = std::min(std::max(1, o.restart), MAX_RESTART);
  const bool fp32 = o.precision == 1;
  const bool amg = o.pc_type == 1;
  int r = ensure_state(c, m, std::max(1, o.max_it), M, fp32);
  if (r) return r;
  SolveState *S = c->gm;
  if (amg) {
    APC_ALLOC(c, S->z64, M + 1);
    APC_ALLOC(c, S->u64, M + 1);
    if (fp32) {
      APC_ALLOC(c, S->z32, M + 1);
      APC_ALLOC(c, S->u32, M + 1);
    }
  }
  if (!fp32) {
    if (amg) {
      r = amg_setup(c, &S->amg, A, pinned, false, amg_options(c));
      if (r) return r;
    }
    if (seg)
      return run_cycles_seg<double>(c, A, A, S->V64.p, b, x, o, false, amg, S->z64.p, S->u64.p,
                                    its, reason, rnorm, sweeps, nullptr, nullptr, nullptr, nullptr);
    return run_cycles<double>(c, A, A, S->V64.p, b, x, o, false, amg, S->z64.p, S->u64.p, its,
                              reason, rnorm, sweeps);
  }
  // FP32 copies of the operator for the inner cycles
  const size_t nnz = (size_t)M * A.lm;
  const float *s32_old = S->s32.p;
  APC_ALLOC(c, S->s32, nnz + 1);
  APC_ALLOC(c, S->d32, M + 1);
  APC_ALLOC(c, S->f32, M + 1);
  APC_ALLOC(c, S->dinv32, M + 1);
  // the density and potential systems share the weights: one copy per build
  if (!(S->s32.p == s32_old && S->s32_src == A.s && S->s32_gen == c->sys_gen &&
        S->s32_n == nnz && A.s == c->sco.p)) {
    hipLaunchKernelGGL(k_to_float, dim3(1024), dim3(256), 0, c->s, nnz, A.s, S->s32.p);
    S->s32_src = A.s;
    S->s32_gen = c->sys_gen;
    S->s32_n = nnz;
  }
  hipLaunchKernelGGL(k_to_float, dim3(256), dim3(256), 0, c->s, (size_t)M, A.d, S->d32.p);
  hipLaunchKernelGGL(k_to_float, dim3(256), dim3(256), 0, c->s, (size_t)M, A.f, S->f32.p);
  hipLaunchKernelGGL(k_to_float, dim3(256), dim3(256), 0, c->s, (size_t)M, A.dinv, S->dinv32.p);
  Op<float> A32{A.nbr, S->s32.p, S->d32.p, S->f32.p, S->dinv32.p, M, A.lm};
  if (amg) {
    // FP32 multigrid inside FP64 flexible GMRES (the FP64 Krylov cycle is the
    // refinement loop)
    r = amg_setup(c, &S->amg, A, pinned, true, amg_options(c));
    if (r) return r;
    APC_ALLOC(c, S->V64, (size_t)(m + 1) * M + 1);
    APC_ALLOC(c, S->Z32, (size_t)(m + 1) * M + 1);
    if (seg)
      return run_cycles_seg<double>(c, A, A, S->V64.p, b, x, o, false, true, S->z64.p, S->u64.p,
                                    its, reason, rnorm, sweeps, &A32, S->Z32.p, S->z32.p,
                                    S->u32.p);
    return run_cycles<double>(c, A, A, S->V64.p, b, x, o, false, true, S->z64.p, S->u64.p, its,
                              reason, rnorm, sweeps, &A32, S->Z32.p, S->z32.p, S->u32.p);
  }
  if (seg)
    return run_cycles_seg<float>(c, A, A32, S->V32.p, b, x, o, true, false, S->z32.p, S->u32.p,
                                 its, reason, rnorm, sweeps, nullptr, nullptr, nullptr, nullptr);
  return run_cycles<float>(c, A, A32, S->V32.p, b, x, o, true, false, S->z32.p, S->u32.p, its,
                           reason, rnorm, sweeps);
}

// gmres over the rows of packed boxes: one Krylov space and one convergence
// test per box (run_cycles_seg), so that every box's solve stops where a
// context of that box alone stops.
static int gmres_boxes(apc_ctx *c, int which, const Op<double> &A, const double *b, double *x,
                       const apc_ksp_opts &o, bool pinned, int *its, int *reason, double *rnorm,
                       int *sweeps) {
  if (c->nbox < 2) return gmres(c, A, b, x, o, pinned, its, reason, rnorm, sweeps);
  int r = gmres(c, A, b, x, o, pinned, its, reason, rnorm, sweeps, true);
  if (r) return r;
  c->box_its[which] = c->seg_its;
  c->box_reason[which] = c->seg_reason;
  c->box_res[which] = c->seg_res;
  return 0;
}

int solve_density_potential(apc_ctx *c, const apc_ksp_opts *o1, const apc_ksp_opts *o2,
                            apc_solve_info *info) {
  const int M = c->nc + c->nv;
  if (M < 1) return fail(c, APC_ERR_ARG, "solve: no computational particles");
  if (o1->precision < 0 || o1->precision > 1 || o2->precision < 0 || o2->precision > 1)
    return fail(c, APC_ERR_ARG, "GMRES precision must be 0 (FP64) or 1 (FP32 + FP64 refinement)");
  if (o1->pc_type < 0 || o1->pc_type > 1 || o2->pc_type < 0 || o2->pc_type > 1)
    return fail(c, APC_ERR_ARG, "pc_type must be 0 (Jacobi) or 1 (multigrid)");
  APC_ALLOC(c, c->rho, M + 1);
  APC_ALLOC(c, c->phi, M + 1);
  APC_ALLOC(c, c->rhs2, M + 1);
  Op<double> A1{c->nbr.p, c->sco.p, c->d1.p, c->f1.p, c->dinv1.p, M, c->prm.lm};
  Op<double> A2{c->nbr.p, c->sco.p, c->d2.p, c->f2.p, c->dinv2.p, M, c->prm.lm};
  apc_solve_info inf;
  memset(&inf, 0, sizeof(inf));
  int r = gmres_boxes(c, 0, A1, c->rhs1.p, c->rho.p, *o1, false, &inf.its_rho, &inf.reason_rho,
                      &inf.rnorm_rho, &inf.refine_rho);
  if (r) return r;
  // right-hand side of the Poisson solve (solver.cpp:233-261)
  if (c->prm.bc == 1 && c->nbox > 1) {
    hipLaunchKernelGGL(k_rhs2_boxes, dim3(c->nbox), dim3(256), 0, c->s, c->rstart.p, c->rho.p,
                       c->ch.p, c->c_pot, c->rhs2.p);
  } else if (c->prm.bc == 1) {
    const int nb = (int)std::min<long>(RED_BLOCKS, std::max<long>(1, nblocks(M, 256)));
    SolveState *S = c->gm;
    double *part = S->part.p;
    hipLaunchKernelGGL(k_mass_sums, dim3(nb), dim3(256), 0, c->s, M, c->rho.p, c->ch.p, part);
    hipLaunchKernelGGL(k_rhs2_periodic, dim3(nb), dim3(256), 0, c->s, M, c->rho.p, part, nb,
                       c->c_pot, c->rhs2.p);
  } else {
    hipLaunchKernelGGL(k_rhs2_bounded, dim3(nblocks(M, 256)), dim3(256), 0, c->s, M, c->rho.p,
                       c->pbnd.p, c->c_pot, c->rhs2.p);
  }
  r = gmres_boxes(c, 1, A2, c->rhs2.p, c->phi.p, *o2, c->prm.bc == 1, &inf.its_phi, &inf.reason_phi,
                  &inf.rnorm_phi, &inf.refine_phi);
  if (r) return r;
  inf.nc = c->nc;
  inf.nv = c->nv;
  inf.nb = c->nb;
  if (info) *info = inf;
  return 0;
}

int solve_transposed(apc_ctx *c, int which, const apc_ksp_opts *o, const double *b, double *out,
                     int *its, int *reason) {
  const int M = c->nc + c->nv;
  Op<double> A = (which == 1)
                     ? Op<double>{c->nbr.p, c->sco.p, c->d1.p, c->f1.p, c->dinv1.p, M, c->prm.lm}
                     : Op<double>{c->nbr.p, c->sco.p, c->d2.p, c->f2.p, c->dinv2.p, M, c->prm.lm};
  A.trp = c->tr_rp.p;
  A.tpos = c->tr_pos.p;
  double rn = 0.0;
  int sw = 0;
  return gmres(c, A, b, out, *o, false, its, reason, &rn, &sw);
}

int amg_level_info(apc_ctx *c, int l, int *rp, int *col, double *val, int *agg) {
  if (!c->gm || !c->gm->amg) return fail(c, APC_ERR_ARG, "amg level: no multigrid built");
  return amg_level_copy(c, c->gm->amg, l, rp, col, val, agg);
}

int amg_info(apc_ctx *c, int *rows, long long *nnz, int cap) {
  if (!c->gm || !c->gm->amg) return 0;
  return amg_levels(c->gm->amg, rows, nnz, cap);
}

void free_solve_state(apc_ctx *c) {
  delete c->gm;
  c->gm = nullptr;
}

}  // namespace apc
