// Context lifetime, particle transfer, solver entry points, per-kernel
// event timers and the one-call step.
#include <algorithm>
#include <chrono>
#include <climits>
#include <cstdarg>

#include "apc_internal.h"
#include "apc_prims.h"

namespace apc {

int fail(apc_ctx *c, int code, const char *fmt, ...) {
  if (c) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    c->err = buf;
  }
  return code;
}

int hip_fail(apc_ctx *c, hipError_t e, const char *what) {
  return fail(c, APC_ERR_HIP, "HIP error %d (%s) in %s", (int)e, hipGetErrorString(e), what);
}

int ktimer_index(apc_ctx *c, const char *name) {
  for (size_t q = 0; q < c->timers.size(); ++q)
    if (c->timers[q].name == name) return (int)q;
  c->timers.push_back(KTimer());
  c->timers.back().name = name;
  return (int)c->timers.size() - 1;
}

static hipEvent_t pooled_event(apc_ctx *c) {
  if (!c->evpool.empty()) {
    hipEvent_t e = c->evpool.back();
    c->evpool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

void ktime_begin(apc_ctx *c, int idx) {
  hipEvent_t e = pooled_event(c);
  if (!e) return;
  hipEventRecord(e, c->s);
  c->timers[idx].ev.push_back(e);
}

void ktime_end(apc_ctx *c, int idx) {
  hipEvent_t e = pooled_event(c);
  if (!e) return;
  hipEventRecord(e, c->s);
  c->timers[idx].ev.push_back(e);
  c->timers[idx].launches++;
}

void ktime_bytes(apc_ctx *c, int idx, double bytes) { c->timers[idx].bytes += bytes; }

static void fold_timers(apc_ctx *c) {
  hipStreamSynchronize(c->s);
  for (auto &t : c->timers) {
    for (size_t q = 0; q + 1 < t.ev.size(); q += 2) {
      float ms = 0;
      if (hipEventElapsedTime(&ms, t.ev[q], t.ev[q + 1]) == hipSuccess) t.total_ms += ms;
    }
    for (auto e : t.ev) c->evpool.push_back(e);
    t.ev.clear();
  }
}

__global__ void k_scatter_back(int n, const int *__restrict__ id, const double *__restrict__ a,
                               double *__restrict__ out) {
  int s = blockIdx.x * 256 + threadIdx.x;
  if (s < n) out[id[s]] = a[s];
}

// plain streaming copy (HBM bandwidth reference)
__global__ void __launch_bounds__(256) k_copy2(size_t n, const double2 *__restrict__ a,
                                               double2 *__restrict__ b) {
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256)
    b[i] = a[i];
}

__global__ void k_iota(int n, int *__restrict__ a) {
  int s = blockIdx.x * 256 + threadIdx.x;
  if (s < n) a[s] = s;
}

// (M x 3 / M x 6 row-major) <-> slot-major grad
__global__ void k_set_grad(int M, const double *__restrict__ g, const double *__restrict__ H,
                           double *__restrict__ grad) {
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= M) return;
  for (int q = 0; q < 3; ++q) grad[(size_t)q * M + i] = g[3 * (size_t)i + q];
  for (int q = 0; q < 6; ++q) grad[(size_t)(3 + q) * M + i] = H[6 * (size_t)i + q];
}

static thread_local std::string tl_err;
int fail_noctx(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  tl_err = buf;
  return code;
}

}  // namespace apc

using namespace apc;

extern "C" {

// NULL: the last error of a context-free call on this thread
const char *apc_last_error(const apc_ctx *c) {
  if (c) return c->err.c_str();
  return apc::tl_err.empty() ? "null context" : apc::tl_err.c_str();
}

const char *apc_version(void) { return "apcloud-mi355x 0.1 (gfx950)"; }

int apc_device_count(int *n) {
  int k = 0;
  hipError_t e = hipGetDeviceCount(&k);
  if (e != hipSuccess) k = 0;
  if (n) *n = k;
  return (k > 0) ? APC_OK : APC_ERR_NO_DEVICE;
}

// smallest valid value of a knob
static int knob_min(int k) { return k == APC_KNOB_AMG_NOPRE ? 0 : 1; }

int apc_create(const apc_params *p, apc_ctx **out) {
  if (!p || !out) return APC_ERR_ARG;
  *out = nullptr;
  if (p->depth < 1 || p->depth > 21) return APC_ERR_ARG;
  if (p->lm < 9) return APC_ERR_ARG;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return APC_ERR_NO_DEVICE;
  apc_ctx *c = new apc_ctx();
  c->prm = *p;
  // APC_* environment overrides of the launch-shape knobs (hand sweeps)
  static const char *const knob_env[APC_KNOB_COUNT] = {
      "APC_NEIGH_XCHUNK", "APC_MDOT_BLOCKS", "APC_MAXPY_BLOCKS", "APC_SEG_BLOCKS",
      "APC_AMG_G8",       "APC_AMG_G4",      "APC_AMG_RG8",      "APC_AMG_NOPRE",
      "APC_AMG_WLEVELS"};
  for (int k = 0; k < APC_KNOB_COUNT; ++k)
    if (const char *v = getenv(knob_env[k]))
      if (atoi(v) >= knob_min(k)) c->knob[k] = atoi(v);
  c->depth_req = p->depth;
  c->maxdepth_req = p->maxdepth;
  hipGetDevice(&c->dev);
  if (hipStreamCreateWithFlags(&c->s, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return APC_ERR_HIP;
  }
  *out = c;
  return APC_OK;
}

// nbox boxes of the same parameters in one context: the box id sits above
// the 3*depth Morton bits of the sort key, so the key budget of 64 bits may
// lower the depth (48 boxes: 6 bits, depth at most 19); the selection
// refuses a tree that the lowered depth would cut.
int apc_create_packed(const apc_params *p, int nbox, const int *counts, apc_ctx **out) {
  if (!p || !out || !counts || nbox < 1 || nbox > 64) return APC_ERR_ARG;
  *out = nullptr;
  if (p->bc != 1) return APC_ERR_UNSUPPORTED;  // periodic boxes only
  long total = 0;
  for (int b = 0; b < nbox; ++b) {
    if (counts[b] < 1) return APC_ERR_ARG;
    total += counts[b];
  }
  if (total > INT_MAX) return APC_ERR_ARG;
  int bits = 0;
  while ((1 << bits) < nbox) ++bits;
  apc_params q = *p;
  const int dmax = (64 - bits) / 3;
  q.depth = std::min(p->depth, dmax);
  q.maxdepth = std::min(p->maxdepth, q.depth);
  int r = apc_create(&q, out);
  if (r) return r;
  apc_ctx *c = *out;
  c->nbox = nbox;
  c->boxbits = bits;
  c->depth_req = p->depth;
  c->maxdepth_req = p->maxdepth;
  c->pstart_h.assign(nbox + 1, 0);
  for (int b = 0; b < nbox; ++b) c->pstart_h[b + 1] = c->pstart_h[b] + counts[b];
  std::vector<double> ones(nbox, 1.0);
  if (c->pstart.ensure(nbox + 1, c->s) != hipSuccess || c->boxmass.ensure(nbox, c->s) != hipSuccess ||
      hipMemcpy(c->pstart.p, c->pstart_h.data(), sizeof(int) * (nbox + 1), hipMemcpyHostToDevice) !=
          hipSuccess ||
      hipMemcpy(c->boxmass.p, ones.data(), sizeof(double) * nbox, hipMemcpyHostToDevice) != hipSuccess) {
    apc_destroy(c);
    *out = nullptr;
    return APC_ERR_NOMEM;
  }
  return APC_OK;
}

int apc_pack_set_masses(apc_ctx *c, const double *m) {
  if (!c || !m) return APC_ERR_ARG;
  if (c->nbox < 2) return fail(c, APC_ERR_ARG, "pack_set_masses: not a packed context");
  for (int b = 0; b < c->nbox; ++b)
    if (!(m[b] > 0.0)) return fail(c, APC_ERR_ARG, "pack_set_masses: box %d mass %g", b, m[b]);
  APC_HIP(c, hipMemcpyAsync(c->boxmass.p, m, sizeof(double) * c->nbox, hipMemcpyHostToDevice, c->s));
  APC_HIP(c, hipStreamSynchronize(c->s));
  c->have_sys = 0;  // rhs1 carries the masses
  return APC_OK;
}

int apc_pack_rows(apc_ctx *c, int *row_start, int *comp_rows) {
  if (!c) return APC_ERR_ARG;
  if ((int)c->rstart_h.size() != c->nbox + 1 || (int)c->rnc_h.size() != c->nbox)
    return fail(c, APC_ERR_ARG, "pack_rows: no selection yet");
  for (int b = 0; b <= c->nbox; ++b) {
    if (row_start) row_start[b] = c->rstart_h[b];
    if (comp_rows && b < c->nbox) comp_rows[b] = c->rnc_h[b];
  }
  return APC_OK;
}

int apc_set_knob(apc_ctx *c, int knob, int value) {
  if (!c) return APC_ERR_ARG;
  if (knob < 0 || knob >= APC_KNOB_COUNT || value < knob_min(knob))
    return fail(c, APC_ERR_ARG, "set_knob: knob %d value %d", knob, value);
  c->knob[knob] = value;
  return APC_OK;
}

int apc_get_knob(apc_ctx *c, int knob, int *value) {
  if (!c || !value) return APC_ERR_ARG;
  if (knob < 0 || knob >= APC_KNOB_COUNT) return fail(c, APC_ERR_ARG, "get_knob: knob %d", knob);
  *value = c->knob[knob];
  return APC_OK;
}

int apc_pack_solve_info(apc_ctx *c, int *its_rho, int *its_phi, int *reason_rho,
                        int *reason_phi, double *res_rho, double *res_phi) {
  if (!c) return APC_ERR_ARG;
  if (c->nbox < 2) return fail(c, APC_ERR_ARG, "pack_solve_info: not a packed context");
  if ((int)c->box_its[0].size() != c->nbox || (int)c->box_its[1].size() != c->nbox)
    return fail(c, APC_ERR_ARG, "pack_solve_info: no solve yet");
  for (int b = 0; b < c->nbox; ++b) {
    if (its_rho) its_rho[b] = c->box_its[0][b];
    if (its_phi) its_phi[b] = c->box_its[1][b];
    if (reason_rho) reason_rho[b] = c->box_reason[0][b];
    if (reason_phi) reason_phi[b] = c->box_reason[1][b];
    if (res_rho) res_rho[b] = c->box_res[0][b];
    if (res_phi) res_phi[b] = c->box_res[1][b];
  }
  return APC_OK;
}

int apc_destroy(apc_ctx *c) {
  if (!c) return APC_OK;
  hipStreamSynchronize(c->s);
  for (auto &t : c->timers)
    for (auto e : t.ev) hipEventDestroy(e);
  for (auto e : c->evpool) hipEventDestroy(e);
  free_solve_state(c);
  apc::DBuf<double> *ds[] = {&c->x, &c->y, &c->z, &c->vx, &c->vy, &c->vz, &c->ax, &c->ay, &c->az,
                             &c->t0, &c->t1, &c->t2, &c->t3, &c->t4, &c->t5, &c->ncx, &c->ncy,
                             &c->ncz, &c->cx, &c->cy, &c->cz, &c->ch, &c->dist, &c->sco, &c->d1,
                             &c->d2, &c->f1, &c->f2, &c->dinv1, &c->dinv2, &c->rhs1, &c->pbnd,
                             &c->bv, &c->rho, &c->phi, &c->rhs2, &c->grad, &c->nscratch, &c->red_d,
                             &c->nxyz, &c->lam, &c->dfield, &c->cic, &c->adj_tmp, &c->adj_tmp2,
                             &c->adj_vec, &c->adj_vec2, &c->cic_w};
  for (auto b : ds) b->release();
  apc::DBuf<int> *is[] = {&c->id, &c->tid, &c->perm, &c->perm2, &c->hist, &c->histx,
                          &c->scratch_i, &c->nlevel, &c->nfirst, &c->nnp, &c->nstatus, &c->nchild,
                          &c->nparent, &c->nflag, &c->nrank, &c->nout, &c->dcount, &c->leaf_node,
                          &c->cnp, &c->cfirst, &c->ctype, &c->n24, &c->n24cnt, &c->n24first,
                          &c->owner, &c->nbr, &c->nscratch_i, &c->nscratch_f, &c->novf,
                          &c->status_dev, &c->tr_pos, &c->tr_pos2, &c->tr_rp, &c->cic_v,
                          &c->cic_v2, &c->cic_rng, &c->vnode, &c->vnode2};
  for (auto b : is) b->release();
  for (auto &b : c->ckpt) b.release(c->s);
  for (auto &b : c->ckpt_id) b.release(c->s);
  c->key.release(c->s); c->key2.release(c->s); c->nprefix.release(c->s); c->red_u64.release(c->s);
  c->visits.release(c->s); c->pstart.release(c->s); c->rstart.release(c->s);
  c->rowbox.release(c->s); c->boxmass.release(c->s); c->pk_key.release(c->s);
  c->pk_key2.release(c->s); c->pk_val.release(c->s); c->pk_val2.release(c->s);
  c->pk_leaf.release(c->s); c->pk_type.release(c->s); c->fitq_rows.release(c->s); c->fitq_stat.release(c->s);
  c->fit_cond.release(c->s); c->scodes.release(c->s); c->slen.release(c->s); c->sidx.release(c->s);
  c->tr_key.release(c->s); c->tr_key2.release(c->s); c->cic_key.release(c->s); c->cic_key2.release(c->s);
  c->vcode.release(c->s); c->vcode2.release(c->s);
  hipStreamDestroy(c->s);
  delete c;
  return APC_OK;
}

int apc_set_particles(apc_ctx *c, int n, const double *x, const double *y, const double *z,
                      const double *vx, const double *vy, const double *vz) {
  if (!c || n < 1 || !x || !y || !z) return fail(c, APC_ERR_ARG, "set_particles: bad arguments");
  if (c->nbox > 1 && n != c->pstart_h[c->nbox])
    return fail(c, APC_ERR_ARG, "set_particles: %d particles, the boxes hold %d", n,
                c->pstart_h[c->nbox]);
  c->frozen = 0;  // new particles: a new topology
  c->have_tr = 0;
  c->have_cic_w = 0;
  c->N = n;
  APC_ALLOC(c, c->x, n); APC_ALLOC(c, c->y, n); APC_ALLOC(c, c->z, n);
  APC_ALLOC(c, c->vx, n); APC_ALLOC(c, c->vy, n); APC_ALLOC(c, c->vz, n);
  APC_ALLOC(c, c->ax, n); APC_ALLOC(c, c->ay, n); APC_ALLOC(c, c->az, n);
  APC_ALLOC(c, c->id, n);
  const size_t b = sizeof(double) * n;
  APC_HIP(c, hipMemcpyAsync(c->x.p, x, b, hipMemcpyHostToDevice, c->s));
  APC_HIP(c, hipMemcpyAsync(c->y.p, y, b, hipMemcpyHostToDevice, c->s));
  APC_HIP(c, hipMemcpyAsync(c->z.p, z, b, hipMemcpyHostToDevice, c->s));
  const double *vs[3] = {vx, vy, vz};
  double *vd[3] = {c->vx.p, c->vy.p, c->vz.p};
  for (int q = 0; q < 3; ++q) {
    if (vs[q]) APC_HIP(c, hipMemcpyAsync(vd[q], vs[q], b, hipMemcpyHostToDevice, c->s));
    else APC_HIP(c, hipMemsetAsync(vd[q], 0, b, c->s));
  }
  APC_HIP(c, hipMemsetAsync(c->ax.p, 0, b, c->s));
  APC_HIP(c, hipMemsetAsync(c->ay.p, 0, b, c->s));
  APC_HIP(c, hipMemsetAsync(c->az.p, 0, b, c->s));
  hipLaunchKernelGGL(k_iota, dim3(nblocks(n, 256)), dim3(256), 0, c->s, n, c->id.p);
  APC_HIP(c, hipStreamSynchronize(c->s));
  c->sorted = false;
  c->have_comp = false;
  c->have_nbr = c->have_sys = c->have_fields = 0;
  return APC_OK;
}

int apc_get_particles(apc_ctx *c, double *x, double *y, double *z, double *vx, double *vy,
                      double *vz, double *ax, double *ay, double *az) {
  if (!c || c->N < 1) return fail(c, APC_ERR_ARG, "get_particles: no particles");
  const int n = c->N;
  APC_ALLOC(c, c->t0, n);
  double *dst[9] = {x, y, z, vx, vy, vz, ax, ay, az};
  const double *src[9] = {c->x.p, c->y.p, c->z.p, c->vx.p, c->vy.p, c->vz.p, c->ax.p, c->ay.p, c->az.p};
  for (int q = 0; q < 9; ++q) {
    if (!dst[q]) continue;
    hipLaunchKernelGGL(k_scatter_back, dim3(nblocks(n, 256)), dim3(256), 0, c->s, n, c->id.p,
                       src[q], c->t0.p);
    APC_HIP(c, hipMemcpyAsync(dst[q], c->t0.p, sizeof(double) * n, hipMemcpyDeviceToHost, c->s));
    APC_HIP(c, hipStreamSynchronize(c->s));
  }
  return APC_OK;
}

int apc_solver_set_neighbors(apc_ctx *c, int lm) {
  if (!c) return APC_ERR_ARG;
  if (lm < 9) return fail(c, APC_ERR_ARG, "NumberOfNeighbors must be >= 9 (got %d)", lm);
  if (lm != c->prm.lm) {
    c->prm.lm = lm;
    c->have_nbr = c->have_sys = c->have_fields = 0;
  }
  return APC_OK;
}

int apc_solver_set_constants(apc_ctx *c, double c_dens, double c_pot, double c_field) {
  if (!c) return APC_ERR_ARG;
  c->c_dens = c_dens;
  c->c_pot = c_pot;
  c->c_field = c_field;
  c->have_sys = 0;  // rhs1 carries c_dens
  return APC_OK;
}

int apc_solver_solve(apc_ctx *c, const apc_ksp_opts *o1, const apc_ksp_opts *o2,
                     apc_solve_info *info) {
  if (!c || !o1 || !o2) return APC_ERR_ARG;
  if (!c->have_comp) return fail(c, APC_ERR_ARG, "solve: no computational particles");
  int r = build_stencils(c);
  if (r) return r;
  r = solve_density_potential(c, o1, o2, info);
  if (r) return r;
  r = differentiate(c);
  if (r) return r;
  if (info) {
    info->fit_qr_rows = c->stencil_qr_rows;
    info->fit_max_cond = c->stencil_max_cond;
  }
  return 0;
}

int apc_solver_amg_levels(apc_ctx *c, int *rows, long long *nnz, int cap, int *nlev) {
  if (!c || !nlev || cap < 0) return APC_ERR_ARG;
  *nlev = amg_info(c, rows, nnz, cap);
  return 0;
}

int apc_solver_amg_level(apc_ctx *c, int level, int *rp, int *col, double *val, int *agg) {
  if (!c) return APC_ERR_ARG;
  return amg_level_info(c, level, rp, col, val, agg);
}

int apc_differentiate(apc_ctx *c, const double *phi) {
  if (!c || !phi) return APC_ERR_ARG;
  if (!c->have_comp) return fail(c, APC_ERR_ARG, "differentiate: no computational particles");
  if (!c->have_sys) {
    int r = build_stencils(c);
    if (r) return r;
  }
  const int M = c->nc + c->nv;
  APC_ALLOC(c, c->phi, M + 1);
  APC_HIP(c, hipMemcpyAsync(c->phi.p, phi, sizeof(double) * M, hipMemcpyHostToDevice, c->s));
  return differentiate(c);
}

int apc_get_fields(apc_ctx *c, double *rho, double *phi, double *g, double *H) {
  if (!c) return APC_ERR_ARG;
  if (!c->have_fields) return fail(c, APC_ERR_ARG, "get_fields: no fields");
  const int M = c->nc + c->nv;
  std::vector<double> gr((size_t)9 * M);
  // the context's stream is non-blocking: copies go on it, then wait
  if (rho && c->rho.p)
    APC_HIP(c, hipMemcpyAsync(rho, c->rho.p, 8ull * M, hipMemcpyDeviceToHost, c->s));
  if (phi) APC_HIP(c, hipMemcpyAsync(phi, c->phi.p, 8ull * M, hipMemcpyDeviceToHost, c->s));
  APC_HIP(c, hipMemcpyAsync(gr.data(), c->grad.p, 8ull * 9 * M, hipMemcpyDeviceToHost, c->s));
  APC_HIP(c, hipStreamSynchronize(c->s));
  for (int i = 0; i < M; ++i) {
    if (g)
      for (int q = 0; q < 3; ++q) g[3 * (size_t)i + q] = gr[(size_t)q * M + i];
    if (H)
      for (int q = 0; q < 6; ++q) H[6 * (size_t)i + q] = gr[(size_t)(3 + q) * M + i];
  }
  return APC_OK;
}

int apc_set_gradients(apc_ctx *c, const double *g, const double *H) {
  if (!c || !g || !H) return APC_ERR_ARG;
  if (!c->have_comp) return fail(c, APC_ERR_ARG, "set_gradients: no computational particles");
  const int M = c->nc + c->nv;
  APC_ALLOC(c, c->grad, (size_t)9 * M + 1);
  apc::DBuf<double> tg, th;
  if (tg.ensure((size_t)3 * M + 1, c->s) || th.ensure((size_t)6 * M + 1, c->s)) {
    tg.release(c->s); th.release(c->s);
    return fail(c, APC_ERR_NOMEM, "set_gradients");
  }
  hipError_t e = hipMemcpyAsync(tg.p, g, 8ull * 3 * M, hipMemcpyHostToDevice, c->s);
  if (e == hipSuccess) e = hipMemcpyAsync(th.p, H, 8ull * 6 * M, hipMemcpyHostToDevice, c->s);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_set_grad, dim3(nblocks(M, 256)), dim3(256), 0, c->s, M, tg.p, th.p, c->grad.p);
    e = hipStreamSynchronize(c->s);
  }
  tg.release(c->s); th.release(c->s);
  if (e != hipSuccess) return hip_fail(c, e, "set_gradients");
  c->have_fields = 1;
  return APC_OK;
}

int apc_step(apc_ctx *c, double c_dens, double c_pot, double c_field, double Cv, double Cx,
             double dt, const apc_ksp_opts *o1, const apc_ksp_opts *o2, double *max_dv,
             double *max_dx, apc_solve_info *info, double *timing) {
  if (!c) return APC_ERR_ARG;
  using clk = std::chrono::steady_clock;
  auto t0 = clk::now();
  int r = 0;
  if (!c->frozen) {  // a fixed topology (differentiation form) keeps tree and leaves
    r = apc_octree_build(c);
    if (r) return r;
    int S = 0;
    r = apc_select(c, c->prm.c, c->prm.order, c->prm.mindepth, c->maxdepth_req, &S);
    if (r) return r;
  }
  APC_HIP(c, hipStreamSynchronize(c->s));
  auto t1 = clk::now();
  r = apc_solver_set_constants(c, c_dens, c_pot, c_field);
  if (r) return r;
  r = apc_solver_solve(c, o1, o2, info);
  if (r) return r;
  APC_HIP(c, hipStreamSynchronize(c->s));
  auto t2 = clk::now();
  r = interp_kick(c, Cv, dt, max_dv);
  if (r) return r;
  r = drift(c, Cx, dt, max_dx);
  if (r) return r;
  APC_HIP(c, hipStreamSynchronize(c->s));
  auto t3 = clk::now();
  if (timing) {
    timing[0] = std::chrono::duration<double>(t1 - t0).count();
    timing[1] = std::chrono::duration<double>(t2 - t1).count();
    timing[2] = std::chrono::duration<double>(t3 - t2).count();
  }
  return APC_OK;
}

int apc_set_timing(apc_ctx *c, int on) {
  if (!c) return APC_ERR_ARG;
  c->timing = on != 0;
  return APC_OK;
}

int apc_reset_kernel_times(apc_ctx *c) {
  if (!c) return APC_ERR_ARG;
  fold_timers(c);
  for (auto &t : c->timers) { t.total_ms = 0; t.launches = 0; t.bytes = 0; }
  return APC_OK;
}

int apc_kernel_times(apc_ctx *c, char *names, int names_len, double *avg_ms, int *launches,
                     double *avg_bytes, int cap) {
  if (!c) return APC_ERR_ARG;
  fold_timers(c);
  int n = 0, pos = 0;
  for (auto &t : c->timers) {
    if (n >= cap) break;
    if (names) {
      int L = (int)t.name.size() + 1;
      if (pos + L > names_len) break;
      memcpy(names + pos, t.name.c_str(), L);
      pos += L;
    }
    if (avg_ms) avg_ms[n] = t.launches ? t.total_ms / t.launches : 0.0;
    if (avg_bytes) avg_bytes[n] = t.launches ? t.bytes / t.launches : 0.0;
    if (launches) launches[n] = t.launches;
    ++n;
  }
  return n;
}

int apc_stream_copy(apc_ctx *c, double bytes, int reps, double *gbs) {
  if (!c || bytes < 1024 || reps < 1 || !gbs) return APC_ERR_ARG;
  size_t n = (size_t)(bytes / 16);  // double2 elements
  apc::DBuf<double> a, b;
  if (a.ensure(2 * n, c->s) != hipSuccess || b.ensure(2 * n, c->s) != hipSuccess) {
    a.release(c->s); b.release(c->s);
    return apc::fail(c, APC_ERR_NOMEM, "stream_copy");
  }
  hipEvent_t e0, e1;
  hipEventCreate(&e0);
  hipEventCreate(&e1);
  unsigned nb = (unsigned)std::min<size_t>((n + 255) / 256, 65535u * 16u);
  hipLaunchKernelGGL(k_copy2, dim3(nb), dim3(256), 0, c->s, n, (const double2 *)a.p, (double2 *)b.p);
  hipEventRecord(e0, c->s);
  for (int r = 0; r < reps; ++r) {
    KScope ks(c, "stream_copy", 2.0 * 16.0 * n);
    hipLaunchKernelGGL(k_copy2, dim3(nb), dim3(256), 0, c->s, n, (const double2 *)a.p, (double2 *)b.p);
  }
  hipEventRecord(e1, c->s);
  hipError_t e = hipEventSynchronize(e1);
  float ms = 0;
  hipEventElapsedTime(&ms, e0, e1);
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  a.release(c->s);
  b.release(c->s);
  if (e != hipSuccess) return apc::hip_fail(c, e, "stream_copy");
  *gbs = 2.0 * 16.0 * n * reps / (ms * 1e-3) / 1e9;
  return APC_OK;
}

int apc_sync(apc_ctx *c) {
  if (!c) return APC_ERR_ARG;
  APC_HIP(c, hipStreamSynchronize(c->s));
  return APC_OK;
}

}  // extern "C"
