// Octree with the reference's public interface (octree.h:224-252), backed by
// the device path: the keys, radix sort and selection run on the GPU.
#ifndef APC_HOST_OCTREE_H
#define APC_HOST_OCTREE_H
#include "apcloud.h"
#include "particle.h"

class Octree {
 public:
  // Builds the tree over the particles (octree.cpp:106-108).
  Octree(Physical_particle *physical_particle, int treedepth, double x_min, double x_max,
         double y_min, double y_max, double z_min, double z_max, int bc);
  ~Octree();

  int buildOctree();
  // octree.cpp:1050-3001.  Returns nc+nv+nb.  nmax is accepted for interface
  // compatibility; device storage grows as needed, so -1 is never returned
  // for lack of room.
  int select_computational_particles_2to1_3d(Computational_particle *computational_particle,
                                             double c, int order, int nmax, int mindepth,
                                             int maxdepth, int bc);
  // octree.cpp:3003-3031: a = g + H dx into pp->a_x..a_z.
  int interpolation_from_node_to_particle_2to1_3d(Computational_particle *cp,
                                                  Physical_particle *pp);
  double periodic_wrap(double x, int dimension);

  // build-only options (the reference has none): the added .param tags
  static int leaf_capacity, boundary_at_faces, bv_kind;
  static double bv_mass, bv_radius, bv_center[3], bv_G;
  static int lm_hint;

 private:
  Physical_particle *pp_;
  int depth_, bc_;
  double bbox_[6];
  apc_ctx *ctx_;
};

#endif
