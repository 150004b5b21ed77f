// Node-to-particle interpolation + kick (octree.cpp:3003-3031,
// mover.cpp:51-89) and drift (mover.cpp:91-153) over the key-sorted particle
// arrays: one thread per particle, the owning leaf from the owner map, so
// every particle array is read and written coalesced.  Built with
// -ffp-contract=off: with the same node fields these reproduce the
// reference's arithmetic bit for bit.
#include "apc_internal.h"
#include "apc_prims.h"

namespace apc {

__global__ void __launch_bounds__(256) k_interp_kick(
    int N, int M, const int *__restrict__ owner, const double *__restrict__ cx,
    const double *__restrict__ cy, const double *__restrict__ cz, const double *__restrict__ grad,
    const double *__restrict__ x, const double *__restrict__ y, const double *__restrict__ z,
    double *__restrict__ vx, double *__restrict__ vy, double *__restrict__ vz,
    double *__restrict__ ax, double *__restrict__ ay, double *__restrict__ az, double Cv,
    double dt, unsigned long long *__restrict__ maxbits, int wrapdx, double lx, double ly,
    double lz) {
  __shared__ double ws[4];
  double change = 0.0;
  for (int s = blockIdx.x * 256 + threadIdx.x; s < N; s += gridDim.x * 256) {
    int i = owner[s];
    double dx = x[s] - cx[i], dy = y[s] - cy[i], dz = z[s] - cz[i];
    if (wrapdx) {  // fixed topology on a periodic box: minimum image to the leaf
      dx -= lx * rint(dx / lx);
      dy -= ly * rint(dy / ly);
      dz -= lz * rint(dz / lz);
    }
    double gx = grad[i], gy = grad[(size_t)M + i], gz = grad[2 * (size_t)M + i];
    double hxx = grad[3 * (size_t)M + i], hyy = grad[4 * (size_t)M + i],
           hzz = grad[5 * (size_t)M + i];
    double hxy = grad[6 * (size_t)M + i], hxz = grad[7 * (size_t)M + i],
           hyz = grad[8 * (size_t)M + i];
    double a0 = gx + dx * hxx + dy * hxy + dz * hxz;
    double a1 = gy + dx * hxy + dy * hyy + dz * hyz;
    double a2 = gz + dx * hxz + dy * hyz + dz * hzz;
    ax[s] = a0;
    ay[s] = a1;
    az[s] = a2;
    vx[s] = vx[s] + a0 * Cv * dt;
    vy[s] = vy[s] + a1 * Cv * dt;
    vz[s] = vz[s] + a2 * Cv * dt;
    change = fmax(change, (a0 * a0 + a1 * a1 + a2 * a2) * Cv * Cv * dt * dt);
  }
  change = block_max256(change, ws);
  if (threadIdx.x == 0) atomicMax(maxbits, (unsigned long long)__double_as_longlong(change));
}

__device__ inline double mwrap(double x, double lo, double hi, int bc) {
  if (bc == 1 && isfinite(x)) {  // a non-finite value would never leave the loop
    while (x >= hi) x -= hi - lo;
    while (x < lo) x += hi - lo;
  }
  return x;
}

__global__ void __launch_bounds__(256) k_drift(int N, double *__restrict__ x, double *__restrict__ y,
                                               double *__restrict__ z, const double *__restrict__ vx,
                                               const double *__restrict__ vy,
                                               const double *__restrict__ vz, double Cx, double dt,
                                               double x0, double x1, double y0, double y1,
                                               double z0, double z1, int bc,
                                               unsigned long long *__restrict__ maxbits) {
  __shared__ double ws[4];
  double change = 0.0;
  for (int s = blockIdx.x * 256 + threadIdx.x; s < N; s += gridDim.x * 256) {
    double ux = vx[s], uy = vy[s], uz = vz[s];
    x[s] = mwrap(x[s] + ux * Cx * dt, x0, x1, bc);
    y[s] = mwrap(y[s] + uy * Cx * dt, y0, y1, bc);
    z[s] = mwrap(z[s] + uz * Cx * dt, z0, z1, bc);
    change = fmax(change, (ux * ux + uy * uy + uz * uz) * Cx * Cx * dt * dt);
  }
  change = block_max256(change, ws);
  if (threadIdx.x == 0) atomicMax(maxbits, (unsigned long long)__double_as_longlong(change));
}

// n_p / h^3 over all computational particles (controller.cpp:146-154)
__global__ void k_density(int S, const int *__restrict__ np, const double *__restrict__ h,
                          double *__restrict__ out) {
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i < S) out[i] = np[i] / h[i] / h[i] / h[i];
}

// The kick and the drift end in one atomic maximum per block on a single
// word, and those atomics serialise: one block per particle tile (65,536 at
// 16.8M particles) cost more than the particle traffic itself.  A grid of
// at most 2048 blocks (eight per CU) strides over the particles instead.
static unsigned mover_blocks(int N) { return std::max(1u, std::min(2048u, nblocks(N, 256))); }

static int read_max(apc_ctx *c, double *out) {
  unsigned long long bits = 0;
  APC_HIP(c, hipMemcpyAsync(&bits, c->red_u64.p, sizeof(bits), hipMemcpyDeviceToHost, c->s));
  APC_HIP(c, hipStreamSynchronize(c->s));
  double v;
  memcpy(&v, &bits, sizeof(v));
  *out = sqrt(v);
  return 0;
}

int interp_kick(apc_ctx *c, double Cv, double dt, double *maxdv) {
  if (!c->have_fields) return fail(c, APC_ERR_ARG, "move_velocity: no node fields (solve first)");
  if (c->comp_external)
    return fail(c, APC_ERR_ARG, "move_velocity: the nodes were set by apc_set_comp, not selected");
  const int N = c->N, M = c->nc + c->nv;
  APC_ALLOC(c, c->red_u64, 4);
  APC_HIP(c, hipMemsetAsync(c->red_u64.p, 0, sizeof(unsigned long long), c->s));
  {
    // per particle: owner 4, x 24, v 24 read + 24 written, a 24 written;
    // per leaf: centre 24 + g,H 72 read.  The reference's offset is not
    // wrapped (octree.cpp:3003-3031): its particles always sit in their
    // leaf.  With the topology held fixed a particle can cross the periodic
    // boundary away from its leaf, so there the minimum image is used.
    const double *b = c->prm.bbox;
    const int wrapdx = (c->frozen && c->prm.bc == 1) ? 1 : 0;
    KScope ks(c, "interp_kick", 100.0 * N + 96.0 * M);
    hipLaunchKernelGGL(k_interp_kick, dim3(mover_blocks(N)), dim3(256), 0, c->s, N, M, c->owner.p,
                       c->cx.p, c->cy.p, c->cz.p, c->grad.p, c->x.p, c->y.p, c->z.p, c->vx.p,
                       c->vy.p, c->vz.p, c->ax.p, c->ay.p, c->az.p, Cv, dt, c->red_u64.p, wrapdx,
                       b[1] - b[0], b[3] - b[2], b[5] - b[4]);
  }
  APC_HIP(c, hipGetLastError());
  if (maxdv) return read_max(c, maxdv);
  return 0;
}

int drift(apc_ctx *c, double Cx, double dt, double *maxdx) {
  const int N = c->N;
  const double *b = c->prm.bbox;
  APC_ALLOC(c, c->red_u64, 4);
  APC_HIP(c, hipMemsetAsync(c->red_u64.p, 0, sizeof(unsigned long long), c->s));
  {
    KScope ks(c, "drift", 72.0 * N);
    hipLaunchKernelGGL(k_drift, dim3(mover_blocks(N)), dim3(256), 0, c->s, N, c->x.p, c->y.p,
                       c->z.p, c->vx.p, c->vy.p, c->vz.p, Cx, dt, b[0], b[1], b[2], b[3], b[4],
                       b[5], c->prm.bc, c->red_u64.p);
  }
  APC_HIP(c, hipGetLastError());
  // positions moved: the tree and everything built on it are stale, unless
  // the topology is held fixed (then only the node fields are recomputed)
  c->have_fields = 0;
  if (!c->frozen) {
    c->sorted = false;
    c->have_comp = false;
    c->have_nbr = c->have_sys = 0;
  }
  if (maxdx) return read_max(c, maxdx);
  return 0;
}

}  // namespace apc

using namespace apc;

extern "C" int apc_interp_kick(apc_ctx *c, double Cv, double dt, double *max_dv) {
  if (!c) return APC_ERR_ARG;
  return interp_kick(c, Cv, dt, max_dv);
}

extern "C" int apc_drift(apc_ctx *c, double Cx, double dt, double *max_dx) {
  if (!c) return APC_ERR_ARG;
  return drift(c, Cx, dt, max_dx);
}

extern "C" int apc_max_density(apc_ctx *c, double *out) {
  if (!c || !out) return APC_ERR_ARG;
  if (!c->have_comp) return fail(c, APC_ERR_ARG, "max_density: no computational particles");
  const int S = c->nc + c->nv + c->nb;
  APC_ALLOC(c, c->red_d, S + 1);
  hipLaunchKernelGGL(k_density, dim3(nblocks(S, 256)), dim3(256), 0, c->s, S, c->cnp.p, c->ch.p,
                     c->red_d.p);
  return reduce_max_nonneg(c, c->red_d.p, S, out);
}
