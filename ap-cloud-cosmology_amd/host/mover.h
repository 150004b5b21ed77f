// Mover with the reference's public interface (mover.h:31-43).
#ifndef APC_HOST_MOVER_H
#define APC_HOST_MOVER_H
#include "octree.h"
#include "particle.h"

class Mover {
 public:
  Mover(Physical_particle *physical_particle1, Computational_particle *computational_particle1);
  ~Mover();
  int Set_Constant_For_Mover_Position(double c) { Constant_For_Mover_Position = c; return 0; }
  int Set_Constant_For_Mover_Velocity(double c) { Constant_For_Mover_Velocity = c; return 0; }
  int Set_Constant_For_GFD_Constant(double c) { GFD_Constant = c; return 0; }
  int Set_Octree_Depth(int c) { Octree_Depth = c; return 0; }
  int Set_BC(int c) { BC = c; return 0; }
  int Set_BoundingBox(double x1, double x2, double y1, double y2, double z1, double z2) {
    x_min = x1; x_max = x2; y_min = y1; y_max = y2; z_min = z1; z_max = z2;
    return 0;
  }
  int Set_maxnc(int c) { maxnc = c; return 0; }
  // mover.cpp:17-49: rebuilds the octree and selects; returns nc+nv+nb or -1
  int select_computational_particles();
  // mover.cpp:51-89 / 91-121: return the maximum change
  double move_position(double dt);
  double move_velocity(double dt);

 private:
  Physical_particle *physical_particle;
  Computational_particle *computational_particle;
  Octree *octree;
  int BC, Octree_Depth;
  double GFD_Constant, Constant_For_Mover_Position, Constant_For_Mover_Velocity;
  double x_min, x_max, y_min, y_max, z_min, z_max;
  int maxnc;
};

#endif
