#include "mover.h"

#include <cstdio>

#include "binding.h"

Mover::Mover(Physical_particle *physical_particle1, Computational_particle *computational_particle1)
    : physical_particle(physical_particle1), computational_particle(computational_particle1),
      octree(nullptr), BC(1), Octree_Depth(20), GFD_Constant(1.0),
      Constant_For_Mover_Position(1.0), Constant_For_Mover_Velocity(1.0), x_min(0), x_max(1),
      y_min(0), y_max(1), z_min(0), z_max(1), maxnc(10000) {}

Mover::~Mover() { delete octree; }

int Mover::select_computational_particles() {
  delete octree;  // the reference leaks the previous tree (mover.cpp:24-25)
  octree = new Octree(physical_particle, Octree_Depth, x_min, x_max, y_min, y_max, z_min, z_max, BC);
  int number = octree->select_computational_particles_2to1_3d(
      computational_particle, GFD_Constant, 2, maxnc, 0, Octree_Depth, BC);
  if (number < 0) printf("Not enough memory for computational particles.\n");
  return number;
}

double Mover::move_velocity(double dt) {
  apc_host::Binding *b = apc_host::find_by_particles(physical_particle);
  if (!b) return -1.0;
  double m = 0.0;
  int r = apc_interp_kick(b->ctx, Constant_For_Mover_Velocity, dt, &m);
  if (r) apc_host::fatal(b->ctx, r, "Mover::move_velocity");
  return m;
}

double Mover::move_position(double dt) {
  apc_host::Binding *b = apc_host::find_by_particles(physical_particle);
  if (!b) {
    // before the first tree the particles are still on the host (the
    // Controller's initial half drift): bind and drift on the device
    Octree t(physical_particle, Octree_Depth, x_min, x_max, y_min, y_max, z_min, z_max, BC);
    b = apc_host::find_by_particles(physical_particle);
    if (!b) return -1.0;
  }
  double m = 0.0;
  int r = apc_drift(b->ctx, Constant_For_Mover_Position, dt, &m);
  if (r) apc_host::fatal(b->ctx, r, "Mover::move_position");
  return m;
}
