// ASCII VTK snapshots with the reference's file names and block layout
// (viewer.h:9-25, viewer.cpp:12-168): <OutputFile>/physical_<t>.vtk and
// <OutputFile>/computational_<t>.vtk.
#ifndef APC_HOST_VIEWER_H
#define APC_HOST_VIEWER_H
#include "particle.h"

class Viewer {
 public:
  Viewer(Physical_particle *physical_particle1, Computational_particle *computational_particle1);
  int Set_File_Name(const char *p_name, const char *c_name);
  void View_Physical_Particles(double t);
  void View_Physical_Particles_Limit(double t, double xmin, double ymin, double zmin, double xmax,
                                     double ymax, double zmax);
  void View_Computational_Particles(double t);

 private:
  char file_name_p[256];
  char file_name_c[256];
  Physical_particle *pp;
  Computational_particle *cp;
};

#endif
