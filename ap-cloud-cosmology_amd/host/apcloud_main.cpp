// apcloud: the drop-in driver with the reference's argv convention
// (test.cpp:16-80):  apcloud <ParameterFile> [-ksp_* options]
// Reads the .param file and the ASCII VTK initial conditions, builds Solver,
// Mover, Viewer and Controller and runs Controller::solve on GPU 0.
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "apcloud.h"
#include "binding.h"
#include "controller.h"
#include "initialization.h"
#include "octree.h"
#include "parameter.h"

int main(int argc, char *argv[]) {
  if (argc < 2) {
    fprintf(stdout, "\nParameters are missing.\n");
    fprintf(stdout, "Call with <ParameterFile>\n\n");
    return 1;
  }
  fprintf(stdout, "\nReading parameters...\n");
  if (read_parameterfile(argv[1])) return 1;
  fprintf(stdout, "\nDone with read parameters.\n");
  Physical_particle physical_particle;
  memset(&physical_particle, 0, sizeof(physical_particle));
  fprintf(stdout, "\nReading initial conditions...\n");
  if (initialization(&physical_particle)) return 1;
  fprintf(stdout, "\nDone with read initial conditions.\n");
  printf("Number of physical particles=%d\n", physical_particle.size);

  Octree::leaf_capacity = LeafCapacity;
  Octree::boundary_at_faces = BoundaryAtFaces;
  Octree::bv_kind = BoundaryValue;
  Octree::bv_mass = BoundaryMass;
  Octree::bv_radius = BoundaryRadius;
  Octree::bv_center[0] = BoundaryCenter_x;
  Octree::bv_center[1] = BoundaryCenter_y;
  Octree::bv_center[2] = BoundaryCenter_z;
  Octree::bv_G = GravitionalConstant;
  Octree::lm_hint = NumberOfNeighbors;

  Computational_particle computational_particle;
  computational_particle.size = computational_particle.nc = computational_particle.nv =
      computational_particle.nb = 0;
  Solver solver(&computational_particle, argc, argv);
  if (GmresPrecision) {
    // -ksp_gmres_precision on the command line wins over the tag
    apc_ksp_opts o = solver.ksp_options();
    if (!o.precision) solver.set_precision(GmresPrecision);
  }
  Mover mover(&physical_particle, &computational_particle);
  Viewer viewer(&physical_particle, &computational_particle);
  Controller controller(&physical_particle, &computational_particle, 1.0 / (1 + Redshift),
                        1.0 / (1 + EndRedShift), 1.0 * DeltaA, &solver, &viewer, &mover,
                        WriteInterval, Periodic, BoundingBox_min_x, BoundingBox_max_x,
                        BoundingBox_min_y, BoundingBox_max_y, BoundingBox_min_z,
                        BoundingBox_max_z, MaxOctreeLevel, CellSizeParameter, NumberOfNeighbors);
  const char *ms = getenv("APCLOUD_MAX_STEPS");
  if (ms) controller.max_steps = atoi(ms);
  int r = controller.solve();
  if (FinalStateFile[0]) {
    FILE *f = fopen(FinalStateFile, "wb");
    if (f) {
      const double *arrs[9] = {physical_particle.x, physical_particle.y, physical_particle.z,
                               physical_particle.v_x, physical_particle.v_y, physical_particle.v_z,
                               physical_particle.a_x, physical_particle.a_y, physical_particle.a_z};
      for (auto a : arrs) fwrite(a, sizeof(double), physical_particle.size, f);
      fclose(f);
    }
  }
  printf("\nExecution Finished!\n");
  apc_host::release(&physical_particle);
  free_particles(&physical_particle);
  return r ? 1 : 0;
}
