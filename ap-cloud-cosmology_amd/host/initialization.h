// Initial conditions from the reference's ASCII VTK layout
// (initialization.cpp:4-101): 4 header lines, "POINTS N", N positions, then
// v_x, v_y, v_z and ID blocks each behind their header lines.
#ifndef APC_HOST_INITIALIZATION_H
#define APC_HOST_INITIALIZATION_H
#include "particle.h"

// Allocates and fills pp from ICFile (ICFileType 1).  Velocities are divided
// by UnitVelocity_in_cm_per_s, and for Units 1 by (1+z)^1.5 like the
// COSMOLOGY_UNIT build.  Returns 0 or 1.
int initialization(Physical_particle *physical_particle);
int initialization_vtk(Physical_particle *physical_particle, const char *path);
void free_particles(Physical_particle *physical_particle);

#endif
