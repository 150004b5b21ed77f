#include "octree.h"

#include <cstdio>
#include <vector>

#include "binding.h"

int Octree::leaf_capacity = 0;
int Octree::boundary_at_faces = 0;
int Octree::bv_kind = 0;
double Octree::bv_mass = 1.0;
double Octree::bv_radius = 0.5;
double Octree::bv_center[3] = {0.5, 0.5, 0.5};
double Octree::bv_G = 1.0;
int Octree::lm_hint = 17;

Octree::Octree(Physical_particle *physical_particle, int treedepth, double x_min, double x_max,
               double y_min, double y_max, double z_min, double z_max, int bc)
    : pp_(physical_particle), depth_(treedepth), bc_(bc), ctx_(nullptr) {
  bbox_[0] = x_min; bbox_[1] = x_max; bbox_[2] = y_min;
  bbox_[3] = y_max; bbox_[4] = z_min; bbox_[5] = z_max;
  buildOctree();
}

Octree::~Octree() {}

int Octree::buildOctree() {
  apc_params p;
  for (int q = 0; q < 6; ++q) p.bbox[q] = bbox_[q];
  p.depth = depth_;
  p.bc = bc_;
  p.c = 0.0;
  p.order = 2;
  p.mindepth = 0;
  p.maxdepth = depth_;
  p.leaf_capacity = leaf_capacity;
  p.lm = lm_hint;
  p.dirichlet_point = 1;
  p.boundary_at_faces = boundary_at_faces;
  p.bv_kind = bv_kind;
  p.bv_mass = bv_mass;
  p.bv_radius = bv_radius;
  p.bv_center[0] = bv_center[0];
  p.bv_center[1] = bv_center[1];
  p.bv_center[2] = bv_center[2];
  p.bv_G = bv_G;
  apc_host::Binding *b = apc_host::bind_particles(pp_, p);
  if (!b) return -1;
  ctx_ = b->ctx;
  int r = apc_octree_build(ctx_);
  if (r) apc_host::fatal(ctx_, r, "Octree::buildOctree");
  return r ? -1 : 0;
}

int Octree::select_computational_particles_2to1_3d(Computational_particle *cp, double c,
                                                   int order, int nmax, int mindepth,
                                                   int maxdepth, int bc) {
  (void)nmax;
  (void)bc;  // the periodicity is the tree's (the reference passes the same value)
  apc_host::Binding *b = apc_host::find_by_particles(pp_);
  if (!b) return -1;
  int S = 0;
  int r = apc_select(ctx_, c, order, mindepth, maxdepth, &S);
  if (r) {
    apc_host::fatal(ctx_, r, "select_computational_particles_2to1_3d");
    return -1;
  }
  apc_host::attach_comp(b, cp);
  r = apc_host::export_comp(b, cp);
  if (r) {
    apc_host::fatal(ctx_, r, "computational particle export");
    return -1;
  }
  return S;
}

int Octree::interpolation_from_node_to_particle_2to1_3d(Computational_particle *cp,
                                                        Physical_particle *pp) {
  (void)cp;
  // a = g + H dx on the device; dt = 0 leaves the velocities unchanged
  int r = apc_interp_kick(ctx_, 1.0, 0.0, nullptr);
  if (r) {
    apc_host::fatal(ctx_, r, "interpolation_from_node_to_particle_2to1_3d");
    return r;
  }
  return apc_get_particles(ctx_, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, pp->a_x,
                           pp->a_y, pp->a_z);
}

double Octree::periodic_wrap(double x, int dimension) {  // octree.cpp:3034-3053
  if (bc_ == 1) {
    double length = bbox_[2 * dimension + 1] - bbox_[2 * dimension];
    while (x >= (0.5 * length)) x -= length;
    while (x < -(0.5 * length)) x += length;
  }
  return x;
}
