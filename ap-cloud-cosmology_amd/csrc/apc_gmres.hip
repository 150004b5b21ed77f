// Restarted GMRES replacing PETSc's KSPGMRES + PCJACOBI (solver.cpp:182-293).
//
// Semantics follow PETSc's KSPSolve_GMRES: left Jacobi preconditioning, zero
// initial guess, convergence on the preconditioned residual with
// KSPConvergedDefault (rtol relative to the first residual norm, atol,
// dtol), classical Gram-Schmidt (PETSc's default orthogonalisation),
// Hessenberg updated with Givens rotations, true residual at each restart.
//
// Device side: one SpMV kernel over the slot-major GFD matrix
// (A x)_i = d_i x_i + f_i sum_j s_ij x_nbr(i,j), a multi-vector dot kernel
// and a fused "subtract projections + norm" kernel per Arnoldi step, all
// with fixed-order block reductions (deterministic).  The (j+2)-value
// Hessenberg column comes back to the host, which runs the Givens update.
#include "apc_internal.h"
#include "apc_prims.h"

namespace apc {

struct PMat {
  const int *nbr;
  const double *s, *d, *f, *dinv;
  int M, lm;
};

struct SolveState {
  DBuf<double> V, part, hd, work;
  double *hpin = nullptr;
  int hcap = 0;
  ~SolveState() {
    V.release(); part.release(); hd.release(); work.release();
    if (hpin) hipHostFree(hpin);
  }
};

constexpr int RED_BLOCKS = 1024;

__device__ inline double row_apply(const PMat &A, const double *__restrict__ x, int i) {
  double acc = 0.0;
  for (int j = 0; j < A.lm; ++j) {
    int m = A.nbr[(size_t)j * A.M + i];
    if (m < A.M) acc += A.s[(size_t)j * A.M + i] * x[m];
  }
  return A.d[i] * x[i] + A.f[i] * acc;
}

__global__ void __launch_bounds__(256) k_spmv_pc(PMat A, const double *__restrict__ x,
                                                 double *__restrict__ y) {
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= A.M) return;
  y[i] = A.dinv[i] * row_apply(A, x, i);
}

__global__ void __launch_bounds__(256) k_pc_resid(PMat A, const double *__restrict__ x,
                                                  const double *__restrict__ b,
                                                  double *__restrict__ r, int zero_guess) {
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= A.M) return;
  r[i] = zero_guess ? A.dinv[i] * b[i] : A.dinv[i] * (b[i] - row_apply(A, x, i));
}

// part[q*nb + block] = sum over the block's rows of w_i * V_q,i, q < nvec
__global__ void __launch_bounds__(256) k_mdot(int M, int nvec, const double *__restrict__ V,
                                              size_t ld, const double *__restrict__ w,
                                              double *__restrict__ part) {
  __shared__ double ws[4];
  const int nb = gridDim.x;
  for (int q0 = 0; q0 < nvec; q0 += 4) {
    double a0 = 0, a1 = 0, a2 = 0, a3 = 0;
    const int nq = min(4, nvec - q0);
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < M; i += (long)nb * 256) {
      double wi = w[i];
      a0 += wi * V[(size_t)q0 * ld + i];
      if (nq > 1) a1 += wi * V[(size_t)(q0 + 1) * ld + i];
      if (nq > 2) a2 += wi * V[(size_t)(q0 + 2) * ld + i];
      if (nq > 3) a3 += wi * V[(size_t)(q0 + 3) * ld + i];
    }
    double r0 = block_sum256(a0, ws);
    double r1 = (nq > 1) ? block_sum256(a1, ws) : 0.0;
    double r2 = (nq > 2) ? block_sum256(a2, ws) : 0.0;
    double r3 = (nq > 3) ? block_sum256(a3, ws) : 0.0;
    if (threadIdx.x == 0) {
      part[(size_t)q0 * nb + blockIdx.x] = r0;
      if (nq > 1) part[(size_t)(q0 + 1) * nb + blockIdx.x] = r1;
      if (nq > 2) part[(size_t)(q0 + 2) * nb + blockIdx.x] = r2;
      if (nq > 3) part[(size_t)(q0 + 3) * nb + blockIdx.x] = r3;
    }
  }
}

// out[q] = sum_b part[q*nb + b]; one block per q
__global__ void __launch_bounds__(256) k_reduce_cols(const double *__restrict__ part, int nb,
                                                     double *__restrict__ out) {
  __shared__ double ws[4];
  const int q = blockIdx.x;
  double a = 0.0;
  for (int b = threadIdx.x; b < nb; b += 256) a += part[(size_t)q * nb + b];
  a = block_sum256(a, ws);
  if (threadIdx.x == 0) out[q] = a;
}

// w -= sum_q h_q V_q ; part[block] = sum w^2
__global__ void __launch_bounds__(256) k_maxpy_norm(int M, int nvec, const double *__restrict__ V,
                                                    size_t ld, const double *__restrict__ h,
                                                    double *__restrict__ w,
                                                    double *__restrict__ part) {
  __shared__ double hs[256];
  __shared__ double ws[4];
  double acc = 0.0;
  for (long i0 = (long)blockIdx.x * 256; i0 < M; i0 += (long)gridDim.x * 256) {
    long i = i0 + threadIdx.x;
    double wi = (i < M) ? w[i] : 0.0;
    for (int q0 = 0; q0 < nvec; q0 += 256) {
      int nq = min(256, nvec - q0);
      __syncthreads();
      if (threadIdx.x < nq) hs[threadIdx.x] = h[q0 + threadIdx.x];
      __syncthreads();
      if (i < M)
        for (int q = 0; q < nq; ++q) wi -= hs[q] * V[(size_t)(q0 + q) * ld + i];
    }
    if (i < M) {
      w[i] = wi;
      acc += wi * wi;
    }
  }
  acc = block_sum256(acc, ws);
  if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

__global__ void k_scale(int M, double *__restrict__ w, const double *__restrict__ sc) {
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i < M) w[i] *= sc[0];
}

__global__ void k_scale_host(int M, double *__restrict__ w, double s) {
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i < M) w[i] *= s;
}

// x += sum_q c_q V_q
__global__ void __launch_bounds__(256) k_axpy_sol(int M, int nvec, const double *__restrict__ V,
                                                  size_t ld, const double *__restrict__ cf,
                                                  double *__restrict__ x) {
  __shared__ double hs[256];
  for (long i0 = (long)blockIdx.x * 256; i0 < M; i0 += (long)gridDim.x * 256) {
    long i = i0 + threadIdx.x;
    double s = 0.0;
    for (int q0 = 0; q0 < nvec; q0 += 256) {
      int nq = min(256, nvec - q0);
      __syncthreads();
      if (threadIdx.x < nq) hs[threadIdx.x] = cf[q0 + threadIdx.x];
      __syncthreads();
      if (i < M)
        for (int q = 0; q < nq; ++q) s += hs[q] * V[(size_t)(q0 + q) * ld + i];
    }
    if (i < M) x[i] += s;
  }
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

__global__ void k_rhs2_periodic(int M, const double *__restrict__ rho, const double *__restrict__ sums,
                                double c_pot, double *__restrict__ rhs2) {
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= M) return;
  double avg = sums[0] / sums[1];
  rhs2[i] = (rho[i] - avg) * c_pot;
}

__global__ void k_rhs2_bounded(int M, const double *__restrict__ rho, const double *__restrict__ pbnd,
                               double c_pot, double *__restrict__ rhs2) {
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= M) return;
  rhs2[i] = rho[i] * c_pot + pbnd[i];
}

enum { KSP_CONVERGED_RTOL = 2, KSP_CONVERGED_ATOL = 3, KSP_DIVERGED_NULL = -2,
       KSP_DIVERGED_ITS = -3, KSP_DIVERGED_DTOL = -4, KSP_DIVERGED_BREAKDOWN = -5,
       KSP_DIVERGED_NANORINF = -9 };

static int reduce_norm2(apc_ctx *c, SolveState *S, const double *v, int M, double *out) {
  const int nb = (int)std::min<long>(RED_BLOCKS, nblocks(M, 256));
  hipLaunchKernelGGL(k_mdot, dim3(nb), dim3(256), 0, c->s, M, 1, v, (size_t)M, v, S->part.p);
  hipLaunchKernelGGL(k_reduce_cols, dim3(1), dim3(256), 0, c->s, S->part.p, nb, S->hd.p);
  APC_HIP(c, hipMemcpyAsync(S->hpin, S->hd.p, sizeof(double), hipMemcpyDeviceToHost, c->s));
  APC_HIP(c, hipStreamSynchronize(c->s));
  *out = S->hpin[0];
  return 0;
}

// Solve A x = b (x overwritten).  Returns 0 and fills its/reason/rnorm.
static int gmres(apc_ctx *c, const PMat &A, const double *b, double *x, const apc_ksp_opts &o,
                 const char *tag, int *its_out, int *reason_out, double *rnorm_out) {
  const int M = A.M;
  const int m = std::max(1, o.restart);
  if (!c->gm) c->gm = new SolveState();
  SolveState *S = c->gm;
  const int nb = (int)std::min<long>(RED_BLOCKS, std::max<long>(1, nblocks(M, 256)));
  APC_ALLOC(c, S->V, (size_t)(m + 1) * M + 1);
  APC_ALLOC(c, S->part, (size_t)(m + 2) * nb + 16);
  APC_ALLOC(c, S->hd, (size_t)m + 8);
  if (S->hcap < m + 8) {
    if (S->hpin) hipHostFree(S->hpin);
    APC_HIP(c, hipHostMalloc((void **)&S->hpin, sizeof(double) * (m + 8)));
    S->hcap = m + 8;
  }
  std::vector<double> HH((size_t)(m + 1) * m, 0.0), cc(m), ss(m), grs(m + 1), nrs(m);
  const unsigned gb = nblocks(M, 256);
  APC_HIP(c, hipMemsetAsync(x, 0, sizeof(double) * M, c->s));
  int its = 0, reason = 0, itcount = 0;
  double rnorm0 = 0, ttol = 0, res = 0;
  bool guess_zero = true;
  auto converged = [&](int nit, double rn) {
    if (nit == 0) {
      rnorm0 = rn;
      ttol = std::max(o.rtol * rnorm0, o.atol);
    }
    if (std::isnan(rn) || std::isinf(rn)) return (int)KSP_DIVERGED_NANORINF;
    if (rn <= ttol) return (rn < o.atol) ? (int)KSP_CONVERGED_ATOL : (int)KSP_CONVERGED_RTOL;
    if (rn >= o.dtol * rnorm0) return (int)KSP_DIVERGED_DTOL;
    return 0;
  };
  double *V = S->V.p;
  const size_t ld = (size_t)M;
  while (!reason) {
    {
      KScope ks(c, "gmres_spmv");
      hipLaunchKernelGGL(k_pc_resid, dim3(gb), dim3(256), 0, c->s, A, x, b, V, guess_zero ? 1 : 0);
    }
    double nn2;
    int r = reduce_norm2(c, S, V, M, &nn2);
    if (r) return r;
    double res_norm = sqrt(nn2);
    res = res_norm;
    grs[0] = res_norm;
    int it = 0;
    if (o.monitor) printf("%3d KSP Residual norm %14.12e \n", its, res);
    if (res == 0.0) { reason = KSP_CONVERGED_ATOL; break; }
    hipLaunchKernelGGL(k_scale_host, dim3(gb), dim3(256), 0, c->s, M, V, 1.0 / res_norm);
    reason = converged(its, res);
    while (!reason && it < m && its < o.max_it) {
      double *vit = V + (size_t)it * ld, *vn = V + (size_t)(it + 1) * ld;
      {
        KScope ks(c, "gmres_spmv");
        hipLaunchKernelGGL(k_spmv_pc, dim3(gb), dim3(256), 0, c->s, A, vit, vn);
      }
      {
        KScope ks(c, "gmres_orth");
        hipLaunchKernelGGL(k_mdot, dim3(nb), dim3(256), 0, c->s, M, it + 1, V, ld, vn, S->part.p);
        hipLaunchKernelGGL(k_reduce_cols, dim3(it + 1), dim3(256), 0, c->s, S->part.p, nb, S->hd.p);
        hipLaunchKernelGGL(k_maxpy_norm, dim3(nb), dim3(256), 0, c->s, M, it + 1, V, ld, S->hd.p, vn,
                           S->part.p + (size_t)(m + 1) * nb);
        hipLaunchKernelGGL(k_reduce_cols, dim3(1), dim3(256), 0, c->s,
                           S->part.p + (size_t)(m + 1) * nb, nb, S->hd.p + it + 1);
      }
      APC_HIP(c, hipMemcpyAsync(S->hpin, S->hd.p, sizeof(double) * (it + 2), hipMemcpyDeviceToHost, c->s));
      APC_HIP(c, hipStreamSynchronize(c->s));
      double *hh = &HH[(size_t)it * (m + 1)];
      for (int q = 0; q <= it; ++q) hh[q] = S->hpin[q];
      double tt = sqrt(S->hpin[it + 1]);
      if (tt != 0.0) hipLaunchKernelGGL(k_scale_host, dim3(gb), dim3(256), 0, c->s, M, vn, 1.0 / tt);
      hh[it + 1] = tt;
      double hapbnd = fabs(tt / grs[it]);
      if (hapbnd > 1.0e-30) hapbnd = 1.0e-30;
      bool hapend = tt < hapbnd;
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
      if (o.monitor) printf("%3d KSP Residual norm %14.12e \n", its, res);
      reason = converged(its, res);
      if (hapend && !reason) reason = KSP_DIVERGED_BREAKDOWN;
    }
    int top = it - 1;
    if (top >= 0) {
      auto H = [&](int rr, int cl) { return HH[(size_t)cl * (m + 1) + rr]; };
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
      for (int q = 0; q <= top; ++q) S->hpin[q] = nrs[q];
      APC_HIP(c, hipMemcpyAsync(S->hd.p, S->hpin, sizeof(double) * (top + 1), hipMemcpyHostToDevice, c->s));
      {
        KScope ks(c, "gmres_update");
        hipLaunchKernelGGL(k_axpy_sol, dim3(nb), dim3(256), 0, c->s, M, top + 1, V, ld, S->hd.p, x);
      }
      APC_HIP(c, hipStreamSynchronize(c->s));  // hpin is reused next cycle
    }
    itcount += it;
    if (itcount >= o.max_it) {
      if (!reason) reason = KSP_DIVERGED_ITS;
      break;
    }
    guess_zero = false;
  }
  APC_HIP(c, hipGetLastError());
  (void)tag;
  *its_out = its;
  *reason_out = reason;
  *rnorm_out = res;
  return 0;
}

int solve_density_potential(apc_ctx *c, const apc_ksp_opts *o1, const apc_ksp_opts *o2,
                            apc_solve_info *info) {
  const int M = c->nc + c->nv;
  if (o1->precision != 0 || o2->precision != 0)
    return fail(c, APC_ERR_UNSUPPORTED, "mixed-precision GMRES is not built in this library");
  APC_ALLOC(c, c->rho, M + 1);
  APC_ALLOC(c, c->phi, M + 1);
  APC_ALLOC(c, c->rhs2, M + 1);
  PMat A1{c->nbr.p, c->sco.p, c->d1.p, c->f1.p, c->dinv1.p, M, c->prm.lm};
  PMat A2{c->nbr.p, c->sco.p, c->d2.p, c->f2.p, c->dinv2.p, M, c->prm.lm};
  apc_solve_info inf{};
  int r = gmres(c, A1, c->rhs1.p, c->rho.p, *o1, "rho", &inf.its_rho, &inf.reason_rho, &inf.rnorm_rho);
  if (r) return r;
  // right-hand side of the Poisson solve (solver.cpp:233-261)
  if (c->prm.bc == 1) {
    const int nb = (int)std::min<long>(RED_BLOCKS, std::max<long>(1, nblocks(M, 256)));
    SolveState *S = c->gm;
    hipLaunchKernelGGL(k_mass_sums, dim3(nb), dim3(256), 0, c->s, M, c->rho.p, c->ch.p, S->part.p);
    hipLaunchKernelGGL(k_reduce_cols, dim3(2), dim3(256), 0, c->s, S->part.p, nb, S->hd.p);
    hipLaunchKernelGGL(k_rhs2_periodic, dim3(nblocks(M, 256)), dim3(256), 0, c->s, M, c->rho.p,
                       S->hd.p, c->c_pot, c->rhs2.p);
  } else {
    hipLaunchKernelGGL(k_rhs2_bounded, dim3(nblocks(M, 256)), dim3(256), 0, c->s, M, c->rho.p,
                       c->pbnd.p, c->c_pot, c->rhs2.p);
  }
  r = gmres(c, A2, c->rhs2.p, c->phi.p, *o2, "phi", &inf.its_phi, &inf.reason_phi, &inf.rnorm_phi);
  if (r) return r;
  if (info) *info = inf;
  return 0;
}

void free_solve_state(apc_ctx *c) {
  delete c->gm;
  c->gm = nullptr;
}

}  // namespace apc
