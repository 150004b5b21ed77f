// Internal state of an apcloud context and helpers shared by the .hip files.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "apcloud.h"

namespace apc {

// Grow-only device buffer.  Growing frees the old allocation: `s`, the stream
// whose queued work may still use it, is drained first, so that memory is never
// handed to another context's thread while a kernel here still reads it.
template <class T>
struct DBuf {
  T *p = nullptr;
  size_t cap = 0;
  hipError_t ensure(size_t n, hipStream_t s) {
    if (n <= cap) return hipSuccess;
    size_t want = n + n / 4 + 256;
    if (p) {
      hipError_t e = hipStreamSynchronize(s);
      if (e != hipSuccess) return e;
      hipFree(p);
    }
    p = nullptr;
    cap = 0;
    hipError_t e = hipMalloc((void **)&p, want * sizeof(T));
    if (e != hipSuccess) { p = nullptr; return e; }
    cap = want;
    return hipSuccess;
  }
  // s: as for ensure; none where the owner has drained its stream already
  // (context destruction)
  void release(hipStream_t s = nullptr) {
    if (p) {
      if (s) (void)hipStreamSynchronize(s);
      hipFree(p);
    }
    p = nullptr;
    cap = 0;
  }
};

// Node status values of the selection tree.
enum : int { ST_LEAF_TD = 1, ST_LEAF_VAC = 2, ST_OPENED = 3, ST_CLAIMED = 4 };

struct KTimer {
  std::string name;
  std::vector<hipEvent_t> ev;  // start/stop pairs not yet folded in
  double total_ms = 0;
  int launches = 0;
  double bytes = 0;  // algorithmic bytes summed over the launches
};

struct SolveState;  // apc_gmres.hip

}  // namespace apc

struct apc_ctx {
  apc_params prm;
  // packed boxes (apc_create_packed): nbox independent boxes of the same
  // parameters in one context.  Box b owns the particles [pstart[b],
  // pstart[b+1]) in every device order (the box id is the top of the sort
  // key) and the rows [rstart[b], rstart[b+1]) of the selection, its
  // computational rows first.  A plain context is one box.
  int nbox = 1, boxbits = 0;
  // the caller's depth and maxdepth (a packed context's key budget may lower
  // prm.depth and prm.maxdepth)
  int depth_req = 0, maxdepth_req = 0;
  std::vector<int> pstart_h, rstart_h, rnc_h;
  apc::DBuf<int> pstart, rstart, rowbox;
  apc::DBuf<double> boxmass;            // per box: factor of the density constant
  apc::DBuf<uint64_t> pk_key, pk_key2;  // packed selection: rows sorted by box
  apc::DBuf<int> pk_val, pk_val2, pk_leaf, pk_type;
  // packed solves: per box, the last segmented solve's iterations, reasons
  // and final residuals (seg_*), and those of the density [0] and
  // potential [1] solves of the last apc_solver_solve
  std::vector<int> seg_its, seg_reason, box_its[2], box_reason[2];
  std::vector<double> seg_res, box_res[2];
  hipStream_t s = nullptr;
  int dev = 0;
  std::string err;
  int N = 0;
  // particles, current (key-sorted) order; id = caller index
  apc::DBuf<double> x, y, z, vx, vy, vz, ax, ay, az;
  apc::DBuf<int> id;
  apc::DBuf<double> t0, t1, t2, t3, t4, t5;  // permutation scratch
  apc::DBuf<int> tid;
  // sort
  apc::DBuf<uint64_t> key, key2;
  apc::DBuf<int> perm, perm2;
  apc::DBuf<int> hist, histx, scratch_i;
  bool sorted = false;
  // selection tree
  int T = 0, T_td = 0, ml = 0;
  std::vector<int> lvl_start;  // top-down level ranges
  apc::DBuf<uint64_t> nprefix;
  apc::DBuf<int> nlevel, nfirst, nnp, nstatus, nchild, nparent, nflag, nrank, nout;
  apc::DBuf<double> ncx, ncy, ncz;
  apc::DBuf<int> dcount;  // small device counters
  int *hcount = nullptr;  // pinned mirror
  // computational particles: S = nc+nv+nb, M = nc+nv
  int nc = 0, nv = 0, nb = 0, nneigh = 0;
  bool have_comp = false;
  int comp_external = 0;               // set by apc_set_comp: no particles belong to the nodes
  apc::DBuf<int> leaf_node;            // M: node id of each output leaf
  apc::DBuf<double> cx, cy, cz, ch;    // S
  apc::DBuf<int> cnp, cfirst, ctype;   // S
  apc::DBuf<int> n24, n24cnt, n24first;  // M*24 padded, M counts, S firsts
  apc::DBuf<int> owner;                // N: output leaf of each sorted particle
  double h0 = 0;
  std::vector<double> hlev, hxl, hyl, hzl;
  // solver
  double c_dens = 1, c_pot = 1, c_field = 1;
  int have_nbr = 0, have_sys = 0, have_fields = 0;
  // bumped by every build of the GFD system (the Laplacian weights sco)
  long long sys_gen = 0;
  apc::DBuf<int> nbr;      // lm x M (slot-major)
  apc::DBuf<double> dist;  // lm x M
  apc::DBuf<double> sco;   // lm x M Laplacian weights s_ij
  apc::DBuf<double> d1, d2, f1, f2, dinv1, dinv2, rhs1, pbnd, bv, rho, phi, rhs2;
  apc::DBuf<double> grad;  // 9 x M: gx gy gz hxx hyy hzz hxy hxz hyz
  // least-squares conditioning: rows whose estimate exceeds fit_cond_qr are
  // fitted by Householder QR
  double fit_cond_qr = 1e5;
  apc::DBuf<int> fitq_rows;
  apc::DBuf<unsigned long long> fitq_stat;  // count, largest estimate (bits)
  apc::DBuf<double> fit_cond;               // per row: the last fit's estimate
  int fit_qr_rows = 0, stencil_qr_rows = 0;
  double fit_max_cond = 0.0, stencil_max_cond = 0.0;
  apc::DBuf<double> nscratch;  // neighbour search scratch
  apc::DBuf<int> nscratch_i;
  apc::DBuf<int> nscratch_f;   // candidate fingerprints
  apc::DBuf<int> novf;         // nodes searched again with the full list bound
  apc::DBuf<uint64_t> vcode, vcode2;  // vacuum leaves sorted by (level, Morton)
  apc::DBuf<int> vnode, vnode2;
  apc::DBuf<double> nxyz;      // node centres packed 4 doubles per node
  // differentiation form (apc_adjoint.hip): topology held fixed
  int frozen = 0, have_tr = 0;
  apc::DBuf<uint64_t> tr_key, tr_key2;
  apc::DBuf<int> tr_pos, tr_pos2, tr_rp;   // column lists of the ELL matrix
  apc::DBuf<double> lam, dfield, cic, adj_tmp, adj_tmp2, adj_vec, adj_vec2;
  apc::DBuf<uint64_t> cic_key, cic_key2;   // CIC contributions sorted by cell
  apc::DBuf<int> cic_v, cic_v2, cic_rng;
  apc::DBuf<double> cic_w;                 // per-particle loss weights (device order)
  int have_cic_w = 0;
  double cic_wsum = 0.0;
  std::vector<apc::DBuf<double>> ckpt;
  std::vector<apc::DBuf<int>> ckpt_id;     // caller indices at the save (normal form)
  std::vector<int> ckpt_n;                 // particle count at the save, -1 if none
  int nsearch_long = 0;
  // launch-shape knobs (APC_KNOB_*): built-in defaults, the APC_* environment
  // overrides read at creation, apc_set_knob (the autotuner) after
  int knob[APC_KNOB_COUNT] = {256, 1024, 1024, 2048, 65536, 1 << 30, 1 << 30, 2, 3};
  int search_slots = 0;        // resident threads of the neighbour search
  int search_waves = 0;        // resident waves of the wave-per-node search
  int search_pen = 0;          // its penalty table is on the device
  apc::DBuf<int> status_dev;
  apc::DBuf<unsigned long long> visits;   // Neigh24 entries visited, list entries sorted
  apc::DBuf<uint8_t> scodes, slen;        // per node: sort codes of its list; list length,
                                          // then (at M) its octant-0 replay count
  apc::DBuf<int> sidx;                    // per node: its list (node indices)
  unsigned long long search_visits = 0;
  apc::SolveState *gm = nullptr;
  // reductions
  apc::DBuf<unsigned long long> red_u64;
  apc::DBuf<double> red_d;
  // timing
  bool timing = false;
  std::vector<apc::KTimer> timers;
  std::vector<hipEvent_t> evpool;
};

namespace apc {

int fail(apc_ctx *c, int code, const char *fmt, ...);
int hip_fail(apc_ctx *c, hipError_t e, const char *what);

#define APC_HIP(ctx, call)                                    \
  do {                                                        \
    hipError_t e__ = (call);                                  \
    if (e__ != hipSuccess) return apc::hip_fail(ctx, e__, #call); \
  } while (0)

#define APC_ALLOC(ctx, buf, n)                                          \
  do {                                                                  \
    if ((buf).ensure(n, (ctx)->s) != hipSuccess)                        \
      return apc::fail(ctx, APC_ERR_NOMEM, "hipMalloc failed for %s (%zu elements)", #buf, (size_t)(n)); \
  } while (0)

// Kernel timing: wrap a launch in ktime_begin/ktime_end.
int ktimer_index(apc_ctx *c, const char *name);
void ktime_begin(apc_ctx *c, int idx);
void ktime_end(apc_ctx *c, int idx);
void ktime_bytes(apc_ctx *c, int idx, double bytes);

struct KScope {
  apc_ctx *c;
  int idx;
  // `bytes`: algorithmic HBM bytes the scoped launch moves (each input read
  // once, each output written once, neighbour gathers counted per use)
  KScope(apc_ctx *c_, const char *name, double bytes = 0.0) : c(c_), idx(-1) {
    if (c->timing) {
      idx = ktimer_index(c, name);
      ktime_begin(c, idx);
      ktime_bytes(c, idx, bytes);
    }
  }
  ~KScope() {
    if (idx >= 0) ktime_end(c, idx);
  }
};

// Scan / reduction helpers (apc_scan.hip).  Exclusive scan of n ints; the
// total is written to *total_dev (device) if non-null.
int scan_exclusive(apc_ctx *c, const int *in, int *out, int n, int *total_dev);
// The same on stream s with a caller workspace `slab` of scan_slab_len(n)
// ints (no allocation, no total).
size_t scan_slab_len(long n);
hipError_t scan_exclusive_ws(hipStream_t s, const int *in, int *out, int n, int *slab);
// Max of n non-negative doubles (device pointer in), result to host.
int reduce_max_nonneg(apc_ctx *c, const double *in, int n, double *out_host);
// Stable LSD radix sort of (k, v) on the low `bits` bits of k; k2/v2 are
// same-sized scratch (apc_sort.hip).
int radix_sort_pairs(apc_ctx *c, uint64_t *k, int *v, uint64_t *k2, int *v2, int n, int bits);
// Morton keys and the stable sort of (key, index) on stream s with caller
// workspaces only (no allocation): keys/idx receive the sorted pairs; k2/v2
// hold n entries, hist/histx 256 * ceil(n / 4096) ints each, slab
// scan_slab_len(256 * ceil(n / 4096)) ints (apc_sort.hip).
hipError_t morton_sort_ws(hipStream_t s, int n, const double *x, const double *y, const double *z,
                          const double *bbox, int depth, uint64_t *keys, int *idx, uint64_t *k2,
                          int *v2, int *hist, int *histx, int *slab);
// Error text of the last failed context-free call on this thread.
int fail_noctx(int code, const char *fmt, ...);

// Stage entry points implemented across files.
int sort_particles(apc_ctx *c);                  // apc_sort.hip
int select_tree(apc_ctx *c, double cc, int order, int mindepth, int maxdepth);  // apc_select.hip
int search_neighbors(apc_ctx *c);                // apc_neigh.hip
int build_stencils(apc_ctx *c);                  // apc_stencil.hip
int differentiate(apc_ctx *c);                   // apc_stencil.hip
int solve_density_potential(apc_ctx *c, const apc_ksp_opts *o1, const apc_ksp_opts *o2,
                            apc_solve_info *info);  // apc_gmres.hip
void free_solve_state(apc_ctx *c);              // apc_gmres.hip
int amg_info(apc_ctx *c, int *rows, long long *nnz, int cap);  // apc_gmres.hip
int amg_level_info(apc_ctx *c, int l, int *rp, int *col, double *val, int *agg);
// A^T out = b (which 1: density, 2: Poisson), b and out device M-vectors
int solve_transposed(apc_ctx *c, int which, const apc_ksp_opts *o, const double *b, double *out,
                     int *its, int *reason);  // apc_gmres.hip
int build_transpose(apc_ctx *c);                // apc_adjoint.hip
int interp_kick(apc_ctx *c, double Cv, double dt, double *maxdv);  // apc_mover.hip
int drift(apc_ctx *c, double Cx, double dt, double *maxdx);        // apc_mover.hip

inline unsigned nblocks(long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

// Morton helpers on 3-bit-per-level prefixes.
__host__ __device__ inline uint64_t spread3(uint32_t v) {
  uint64_t x = v & 0x1fffff;
  x = (x | (x << 32)) & 0x1f00000000ffffULL;
  x = (x | (x << 16)) & 0x1f0000ff0000ffULL;
  x = (x | (x << 8)) & 0x100f00f00f00f00fULL;
  x = (x | (x << 4)) & 0x10c30c30c30c30c3ULL;
  x = (x | (x << 2)) & 0x1249249249249249ULL;
  return x;
}
__host__ __device__ inline uint32_t compact3(uint64_t x) {
  x &= 0x1249249249249249ULL;
  x = (x | (x >> 2)) & 0x10c30c30c30c30c3ULL;
  x = (x | (x >> 4)) & 0x100f00f00f00f00fULL;
  x = (x | (x >> 8)) & 0x1f0000ff0000ffULL;
  x = (x | (x >> 16)) & 0x1f00000000ffffULL;
  x = (x | (x >> 32)) & 0x1fffffULL;
  return (uint32_t)x;
}
// prefix digit order: x = bit 2, y = bit 1, z = bit 0 of every 3-bit group
__host__ __device__ inline uint64_t morton_encode(uint32_t ix, uint32_t iy, uint32_t iz) {
  return (spread3(ix) << 2) | (spread3(iy) << 1) | spread3(iz);
}
__host__ __device__ inline void morton_decode(uint64_t m, uint32_t &ix, uint32_t &iy, uint32_t &iz) {
  ix = compact3(m >> 2);
  iy = compact3(m >> 1);
  iz = compact3(m);
}

}  // namespace apc
