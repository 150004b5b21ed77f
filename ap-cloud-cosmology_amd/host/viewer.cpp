#include "viewer.h"

#include <cstdio>
#include <cstring>

#include "binding.h"
#include "parameter.h"

namespace {

FILE *open_snapshot(const char *base, double t) {
  char name[320];
  snprintf(name, sizeof(name), "%s%f.vtk", base, t);
  FILE *f = fopen(name, "w");
  if (!f) perror("Error opening file");
  return f;
}

void header(FILE *f, const char *what, int n) {
  fprintf(f, "# vtk DataFile Version 3.0\n");
  fprintf(f, "Data on %s particles for cosmology problem\n", what);
  fprintf(f, "ASCII\n");
  fprintf(f, "DATASET POLYDATA\n");
  fprintf(f, "POINTS %d double\n\n", n);
}

template <class T>
void scalars(FILE *f, const char *name, const char *type, const char *fmt, const T *v, int n) {
  fprintf(f, "\nSCALARS %s %s\n", name, type);
  fprintf(f, "LOOKUP_TABLE default\n");
  for (int i = 0; i < n; ++i) fprintf(f, fmt, v[i]);
}

void vectors(FILE *f, const char *name, const double *a, const double *b, const double *c, int n) {
  fprintf(f, "\nVECTORS %s double\n", name);
  for (int i = 0; i < n; ++i) fprintf(f, "%e %e %e\n", a[i], b[i], c[i]);
}

}  // namespace

Viewer::Viewer(Physical_particle *p, Computational_particle *c) : pp(p), cp(c) {
  Set_File_Name(OutputFile, OutputFile);
  strncat(file_name_p, "/physical_", sizeof(file_name_p) - strlen(file_name_p) - 1);
  strncat(file_name_c, "/computational_", sizeof(file_name_c) - strlen(file_name_c) - 1);
}

int Viewer::Set_File_Name(const char *p_name, const char *c_name) {
  snprintf(file_name_p, sizeof(file_name_p), "%s", p_name);
  snprintf(file_name_c, sizeof(file_name_c), "%s", c_name);
  return 0;
}

void Viewer::View_Physical_Particles(double t) {
  FILE *f = open_snapshot(file_name_p, t);
  if (!f) return;
  const int n = pp->size;
  header(f, "physical", n);
  for (int i = 0; i < n; ++i) fprintf(f, "%f %f %f\n", pp->x[i], pp->y[i], pp->z[i]);
  fprintf(f, "\nPOINT_DATA %d", n);
  fprintf(f, "\nVECTORS velocity double\n");
  for (int i = 0; i < n; ++i) fprintf(f, "%e %e %e\n", pp->v_x[i], pp->v_y[i], pp->v_z[i]);
  scalars(f, "v_x", "double", "%e\n", pp->v_x, n);
  scalars(f, "v_y", "double", "%e\n", pp->v_y, n);
  scalars(f, "v_z", "double", "%e\n", pp->v_z, n);
  vectors(f, "acceleration", pp->a_x, pp->a_y, pp->a_z, n);
  scalars(f, "a_x", "double", "%e\n", pp->a_x, n);
  scalars(f, "a_y", "double", "%e\n", pp->a_y, n);
  scalars(f, "a_z", "double", "%e\n", pp->a_z, n);
  fclose(f);
}

void Viewer::View_Physical_Particles_Limit(double t, double xmin, double ymin, double zmin,
                                           double xmax, double ymax, double zmax) {
  FILE *f = open_snapshot(file_name_p, t);
  if (!f) return;
  header(f, "physical", 2);
  fprintf(f, "%f %f %f\n", xmin, ymin, zmin);
  fprintf(f, "%f %f %f\n", xmax, ymax, zmax);
  fprintf(f, "\nPOINT_DATA %d", 2);
  const double z2[2] = {0.0, 0.0};
  fprintf(f, "\nVECTORS velocity double\n");
  for (int i = 0; i < 2; ++i) fprintf(f, "%f %f %f\n", 0.0, 0.0, 0.0);
  scalars(f, "v_x", "double", "%f\n", z2, 2);
  scalars(f, "v_y", "double", "%f\n", z2, 2);
  scalars(f, "v_z", "double", "%f\n", z2, 2);
  fprintf(f, "\nVECTORS acceleration double\n");
  for (int i = 0; i < 2; ++i) fprintf(f, "%f %f %f\n", 0.0, 0.0, 0.0);
  scalars(f, "a_x", "double", "%f\n", z2, 2);
  scalars(f, "a_y", "double", "%f\n", z2, 2);
  scalars(f, "a_z", "double", "%f\n", z2, 2);
  fclose(f);
}

void Viewer::View_Computational_Particles(double t) {
  FILE *f = open_snapshot(file_name_c, t);
  if (!f) return;
  const int n = cp->size;
  header(f, "computational", n);
  for (int i = 0; i < n; ++i) fprintf(f, "%f %f %f\n", cp->x[i], cp->y[i], cp->z[i]);
  fprintf(f, "\nPOINT_DATA %d\n", n);
  fprintf(f, "SCALARS n_p int\nLOOKUP_TABLE default\n");
  for (int i = 0; i < n; ++i) fprintf(f, "%d\n", cp->NumberOfPhysicalParticles[i]);
  scalars(f, "type", "int", "%d\n", cp->Type.data(), n);
  scalars(f, "first_id", "int", "%d\n", cp->IndexOfFirstPhysicalParticle.data(), n);
  scalars(f, "density", "double", "%e\n", cp->Density.data(), n);
  scalars(f, "potential", "double", "%e\n", cp->Potential.data(), n);
  scalars(f, "p_x", "double", "%e\n", cp->p_x.data(), n);
  scalars(f, "p_y", "double", "%e\n", cp->p_y.data(), n);
  scalars(f, "p_z", "double", "%e\n", cp->p_z.data(), n);
  fclose(f);
}
