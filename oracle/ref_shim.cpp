/*
 * TEST INFRASTRUCTURE ONLY.  C entry points over the reference's own,
 * unmodified octree.cpp / mover.cpp / particle.cpp / parameter.cpp, compiled
 * from /root/reference by oracle/Makefile into oracle/_ref/libref.so.  Used
 * by tests/ to pin the oracle's restatement (keys, octree arrays, selection,
 * Neigh24, interpolation, kick, drift) on small inputs, and the build's
 * Viewer against the reference's viewer.cpp.
 *
 * The selection keeps its per-call arrays on the stack (octree.cpp:1112-1113,
 * ~104 B per temporary node), so every reference call runs on a worker thread
 * with a 1 GiB stack.
 */
#include <pthread.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <ctime>
#include <deque>
#include <fstream>
#include <functional>
#include <iomanip>
#include <iostream>
#include <numeric>
#include <sstream>
#include <vector>
#include <math.h>
#include <omp.h>
#include <stdio.h>
#include <stdlib.h>

/* the standard headers are already in; only the reference's own classes see
 * this, so the test can read the Octree's sorted keys and node arrays */
#define private public
#include "octree.h"
#include "mover.h"
#include "viewer.h"
#undef private

namespace {

void run_big_stack(std::function<void()> f) {
  pthread_attr_t a;
  pthread_attr_init(&a);
  pthread_attr_setstacksize(&a, (size_t)1 << 30);
  pthread_t t;
  auto tramp = [](void *p) -> void * {
    (*(std::function<void()> *)p)();
    return nullptr;
  };
  pthread_create(&t, &a, tramp, &f);
  pthread_join(t, nullptr);
  pthread_attr_destroy(&a);
}

struct RefRun {
  Physical_particle pp;
  Computational_particle cp;
  Mover *mover = nullptr;
  std::vector<double> x, y, z, vx, vy, vz, ax, ay, az;
  std::vector<int> id;
  double bb[6];
  int depth = 0, bc = 0;
  int S = -1;
};

}  // namespace

extern "C" {

/* Keys and sorted KeyIndex straight from Octree::buildOctree. */
int ref_tree(int n, const double *x, const double *y, const double *z, const double *bb,
             int depth, uint64_t *skey, int *sidx, int *len_out, int *numnodes_out,
             uint64_t *nkey, int *ndepth, int *nfirst, int *ncount, int *nfchild, int *nnchild,
             double *lo, double *hi, int cap) {
  std::vector<double> X(x, x + n), Y(y, y + n), Z(z, z + n);
  Physical_particle pp;
  memset(&pp, 0, sizeof(pp));
  pp.size = n;
  pp.x = X.data();
  pp.y = Y.data();
  pp.z = Z.data();
  int rc = 0;
  run_big_stack([&]() {
    Octree t(&pp, depth, bb[0], bb[1], bb[2], bb[3], bb[4], bb[5], 1);
    for (int i = 0; i < n; ++i) {
      skey[i] = t.m_vParticleKeyIndex[i].key;
      sidx[i] = t.m_vParticleKeyIndex[i].index;
    }
    *len_out = t.m_iTreeLength;
    /* numnodes by depth */
    for (int k = 0; k <= depth; ++k) numnodes_out[k] = 0;
    for (int i = 0; i < t.m_iTreeLength; ++i) numnodes_out[t.m_vDepth[i]]++;
    if (t.m_iTreeLength > cap) { rc = 1; return; }
    for (int i = 0; i < t.m_iTreeLength; ++i) {
      nkey[i] = t.m_vNodeKey[i];
      ndepth[i] = t.m_vDepth[i];
      nfirst[i] = t.m_vFirstParticleIndex[i];
      ncount[i] = t.m_vNumberOfContainedParticles[i];
      nfchild[i] = t.m_vFirstChildIndex[i];
      nnchild[i] = t.m_vNumberOfChildren[i];
      lo[3 * i] = t.m_vLowerLimitOfX[i];
      lo[3 * i + 1] = t.m_vLowerLimitOfY[i];
      lo[3 * i + 2] = t.m_vLowerLimitOfZ[i];
      hi[3 * i] = t.m_vUpperLimitOfX[i];
      hi[3 * i + 1] = t.m_vUpperLimitOfY[i];
      hi[3 * i + 2] = t.m_vUpperLimitOfZ[i];
    }
  });
  return rc;
}

/* A Mover over caller particles; select() runs Mover::select_computational_
 * particles (octree rebuild + nmax-doubling retry, mover.cpp:17-49). */
void *ref_new(int n, const double *x, const double *y, const double *z, const double *vx,
              const double *vy, const double *vz, const double *bb, int depth, int bc) {
  RefRun *r = new RefRun();
  r->x.assign(x, x + n); r->y.assign(y, y + n); r->z.assign(z, z + n);
  r->vx.assign(vx, vx + n); r->vy.assign(vy, vy + n); r->vz.assign(vz, vz + n);
  r->ax.assign(n, 0.0); r->ay.assign(n, 0.0); r->az.assign(n, 0.0);
  r->id.resize(n);
  for (int i = 0; i < n; ++i) r->id[i] = i;
  memset(&r->pp, 0, sizeof(r->pp));
  r->pp.size = n;
  r->pp.x = r->x.data(); r->pp.y = r->y.data(); r->pp.z = r->z.data();
  r->pp.v_x = r->vx.data(); r->pp.v_y = r->vy.data(); r->pp.v_z = r->vz.data();
  r->pp.a_x = r->ax.data(); r->pp.a_y = r->ay.data(); r->pp.a_z = r->az.data();
  r->pp.ID = r->id.data();
  memcpy(r->bb, bb, sizeof(r->bb));
  r->depth = depth;
  r->bc = bc;
  r->mover = new Mover(&r->pp, &r->cp);
  r->mover->Set_BoundingBox(bb[0], bb[1], bb[2], bb[3], bb[4], bb[5]);
  r->mover->Set_Octree_Depth(depth);
  r->mover->Set_BC(bc);
  return r;
}

int ref_select(void *h, double c, int maxnc) {
  RefRun *r = (RefRun *)h;
  r->mover->Set_Constant_For_GFD_Constant(c);
  r->mover->Set_maxnc(maxnc);
  run_big_stack([&]() { r->S = r->mover->select_computational_particles(); });
  return r->S;
}

int ref_comp_sizes(void *h, int *nc, int *nv, int *nb, int *nneigh) {
  RefRun *r = (RefRun *)h;
  *nc = r->cp.nc; *nv = r->cp.nv; *nb = r->cp.nb;
  *nneigh = (int)r->cp.Neigh24.size();
  return 0;
}

int ref_comp_get(void *h, double *x, double *y, double *z, double *hh, int *np, int *first,
                 int *type, int *fneigh, int *nneigh, int *neigh24) {
  RefRun *r = (RefRun *)h;
  Computational_particle &c = r->cp;
  int S = c.size;
  for (int i = 0; i < S; ++i) {
    x[i] = c.x[i]; y[i] = c.y[i]; z[i] = c.z[i]; hh[i] = c.CellLength[i];
    np[i] = c.NumberOfPhysicalParticles[i]; first[i] = c.IndexOfFirstPhysicalParticle[i];
    type[i] = c.Type[i]; fneigh[i] = c.FirstNeigh[i]; nneigh[i] = c.NumberNeigh[i];
  }
  for (size_t i = 0; i < c.Neigh24.size(); ++i) neigh24[i] = c.Neigh24[i];
  return 0;
}

/* Mover::move_velocity(dt) after the caller has supplied p_x..p_yz per
 * comp/vacuum node (order xx yy zz xy xz yz for H), then move_position(dt). */
int ref_move(void *h, const double *g, const double *H, double Cv, double Cx, double dt,
             double *maxdv, double *maxdx) {
  RefRun *r = (RefRun *)h;
  Computational_particle &c = r->cp;
  int M = c.nc + c.nv;
  for (int i = 0; i < M; ++i) {
    c.p_x[i] = g[3 * i]; c.p_y[i] = g[3 * i + 1]; c.p_z[i] = g[3 * i + 2];
    c.p_xx[i] = H[6 * i]; c.p_yy[i] = H[6 * i + 1]; c.p_zz[i] = H[6 * i + 2];
    c.p_xy[i] = H[6 * i + 3]; c.p_xz[i] = H[6 * i + 4]; c.p_yz[i] = H[6 * i + 5];
  }
  r->mover->Set_Constant_For_Mover_Velocity(Cv);
  r->mover->Set_Constant_For_Mover_Position(Cx);
  *maxdv = r->mover->move_velocity(dt);
  *maxdx = r->mover->move_position(dt);
  return 0;
}

int ref_particles(void *h, double *x, double *y, double *z, double *vx, double *vy, double *vz,
                  double *ax, double *ay, double *az) {
  RefRun *r = (RefRun *)h;
  size_t n = r->x.size(), b = n * sizeof(double);
  memcpy(x, r->x.data(), b); memcpy(y, r->y.data(), b); memcpy(z, r->z.data(), b);
  memcpy(vx, r->vx.data(), b); memcpy(vy, r->vy.data(), b); memcpy(vz, r->vz.data(), b);
  memcpy(ax, r->ax.data(), b); memcpy(ay, r->ay.data(), b); memcpy(az, r->az.data(), b);
  return 0;
}

/* The reference's Viewer (viewer.cpp:3-168) on the given particles: writes
 * <dir>/physical_<t0>.vtk by View_Physical_Particles_Limit with the box,
 * then <dir>/physical_<t1>.vtk and <dir>/computational_<t1>.vtk. */
int ref_view(const char *dir, double t0, double t1, const double *bb, int n, const double *x,
             const double *y, const double *z, const double *vx, const double *vy,
             const double *vz, const double *ax, const double *ay, const double *az, int S,
             const double *cx, const double *cy, const double *cz, const int *np,
             const int *type, const int *first, const double *dens, const double *pot,
             const double *px, const double *py, const double *pz) {
  std::vector<double> a[9];
  const double *src[9] = {x, y, z, vx, vy, vz, ax, ay, az};
  for (int k = 0; k < 9; ++k) a[k].assign(src[k], src[k] + n);
  Physical_particle pp;
  memset(&pp, 0, sizeof(pp));
  pp.size = n;
  pp.x = a[0].data(); pp.y = a[1].data(); pp.z = a[2].data();
  pp.v_x = a[3].data(); pp.v_y = a[4].data(); pp.v_z = a[5].data();
  pp.a_x = a[6].data(); pp.a_y = a[7].data(); pp.a_z = a[8].data();
  Computational_particle cp;
  cp.size = S;
  cp.nc = S; cp.nv = 0; cp.nb = 0;
  cp.x.assign(cx, cx + S); cp.y.assign(cy, cy + S); cp.z.assign(cz, cz + S);
  cp.NumberOfPhysicalParticles.assign(np, np + S);
  cp.Type.assign(type, type + S);
  cp.IndexOfFirstPhysicalParticle.assign(first, first + S);
  cp.Density.assign(dens, dens + S); cp.Potential.assign(pot, pot + S);
  cp.p_x.assign(px, px + S); cp.p_y.assign(py, py + S); cp.p_z.assign(pz, pz + S);
  snprintf(OutputFile, sizeof(OutputFile), "%s", dir);
  Viewer v(&pp, &cp);
  v.View_Physical_Particles_Limit(t0, bb[0], bb[2], bb[4], bb[1], bb[3], bb[5]);
  v.View_Physical_Particles(t1);
  v.View_Computational_Particles(t1);
  return 0;
}

void ref_free(void *h) {
  RefRun *r = (RefRun *)h;
  if (r->S >= 0 || r->mover->octree) delete r->mover;
  delete r;
}

}  // extern "C"
