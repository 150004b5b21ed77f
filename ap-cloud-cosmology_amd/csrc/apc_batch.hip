// Multi-box form: many independent periodic boxes stepped in one call.
//
// Each box is a full context (its own device buffers and its own non-blocking
// stream); a pool of host threads takes boxes largest first (LPT order) and
// runs the whole step schedule on each, with no barrier between boxes, so one
// box's host-side waits (selection rounds, solver restarts) overlap the other
// boxes' kernels on the GPU's hardware queues.  A box's trajectory is the one
// it has when stepped alone: the same kernels in the same order on its own
// stream.  The reference steps one box per process (controller.cpp:106-208).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdio>
#include <string>
#include <thread>
#include <vector>

#include "apcloud.h"

struct apc_batch {
  std::vector<apc_ctx *> box;
  std::vector<int> n;
  int nthreads = 1;
  std::string err;
};

extern "C" const char *apc_batch_last_error(const apc_batch *b) {
  return b ? b->err.c_str() : "null batch";
}

extern "C" int apc_batch_create(const apc_params *p, int nbox, int nthreads, apc_batch **out) {
  if (!p || !out || nbox < 1 || nthreads < 0) return APC_ERR_ARG;
  *out = nullptr;
  apc_batch *b = new apc_batch;
  b->nthreads = nthreads > 0 ? nthreads : std::min(nbox, 8);
  b->box.assign(nbox, nullptr);
  b->n.assign(nbox, 0);
  for (int i = 0; i < nbox; ++i) {
    int r = apc_create(p, &b->box[i]);
    if (r) {
      for (auto c : b->box)
        if (c) apc_destroy(c);
      delete b;
      return r;
    }
  }
  *out = b;
  return APC_OK;
}

extern "C" int apc_batch_destroy(apc_batch *b) {
  if (!b) return APC_OK;
  for (auto c : b->box)
    if (c) apc_destroy(c);
  delete b;
  return APC_OK;
}

extern "C" int apc_batch_size(const apc_batch *b) { return b ? (int)b->box.size() : 0; }

extern "C" apc_ctx *apc_batch_box(apc_batch *b, int i) {
  if (!b || i < 0 || i >= (int)b->box.size()) return nullptr;
  return b->box[i];
}

extern "C" int apc_batch_set_particles(apc_batch *b, int i, int n, const double *x,
                                       const double *y, const double *z, const double *vx,
                                       const double *vy, const double *vz) {
  apc_ctx *c = apc_batch_box(b, i);
  if (!c) return APC_ERR_ARG;
  int r = apc_set_particles(c, n, x, y, z, vx, vy, vz);
  if (r == APC_OK) b->n[i] = n;
  return r;
}

extern "C" int apc_batch_run(apc_batch *b, int nsteps, const double *coef, const double *c_dens,
                             const apc_ksp_opts *o1, const apc_ksp_opts *o2,
                             apc_solve_info *info, int *status) {
  if (!b || nsteps < 0 || (nsteps > 0 && !coef) || !c_dens || !o1 || !o2) return APC_ERR_ARG;
  const int nbox = (int)b->box.size();
  for (int i = 0; i < nbox; ++i)
    if (b->n[i] < 1) {
      b->err = "batch_run: box " + std::to_string(i) + " has no particles";
      return APC_ERR_ARG;
    }
  std::vector<int> order(nbox);
  for (int i = 0; i < nbox; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int a, int c) { return b->n[a] > b->n[c]; });
  std::vector<int> st(nbox, 0);
  std::vector<std::string> msg(nbox);
  std::atomic<int> next{0};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) dev = 0;
  auto worker = [&]() {
    hipSetDevice(dev);  // the caller's device, in every worker thread
    for (;;) {
      const int k = next.fetch_add(1);
      if (k >= nbox) return;
      const int i = order[k];
      apc_ctx *c = b->box[i];
      apc_solve_info last{};
      for (int s = 0; s < nsteps; ++s) {
        const double *q = coef + 5 * (size_t)s;  // c_pot, c_field, Cv, Cx, dt
        double mdv = 0.0, mdx = 0.0;
        int r = apc_step(c, c_dens[i], q[0], q[1], q[2], q[3], q[4], o1, o2, &mdv, &mdx, &last,
                         nullptr);
        if (r) {
          st[i] = r;
          msg[i] = apc_last_error(c);
          break;
        }
      }
      if (info) info[i] = last;
    }
  };
  const int nt = std::max(1, std::min(b->nthreads, nbox));
  std::vector<std::thread> pool;
  for (int t = 1; t < nt; ++t) pool.emplace_back(worker);
  worker();
  for (auto &t : pool) t.join();
  int first = 0;
  for (int i = 0; i < nbox; ++i) {
    if (status) status[i] = st[i];
    if (st[i] && !first) {
      first = st[i];
      b->err = "batch_run: box " + std::to_string(i) + ": " + msg[i];
    }
  }
  return first;
}
