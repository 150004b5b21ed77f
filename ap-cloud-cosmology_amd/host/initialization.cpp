#include "initialization.h"

#include <cmath>
#include <cstdio>
#include <cstring>

#include "parameter.h"

namespace {
void skip_lines(FILE *f, int k) {
  char buf[512];
  for (int i = 0; i < k; ++i)
    if (!fgets(buf, sizeof(buf), f)) return;
}
int read_doubles(FILE *f, double *a, int n) {
  for (int i = 0; i < n; ++i)
    if (fscanf(f, "%lf", &a[i]) != 1) return 1;
  return 0;
}
}  // namespace

int initialization(Physical_particle *pp) {
  if (ICFileType != 1) {
    printf("ICFileType invalid!\n");
    return 1;
  }
  return initialization_vtk(pp, ICFile);
}

int initialization_vtk(Physical_particle *pp, const char *path) {
  FILE *f = fopen(path, "r");
  if (!f) {
    printf("Error in opening input file %s\n", path);
    return 1;
  }
  char tok[200];
  skip_lines(f, 4);
  int n = 0;
  if (fscanf(f, "%199s %d", tok, &n) != 2 || n < 1) {
    fclose(f);
    printf("Error in reading %s: no POINTS block\n", path);
    return 1;
  }
  pp->size = n;
  pp->x = new double[n]; pp->y = new double[n]; pp->z = new double[n];
  pp->v_x = new double[n]; pp->v_y = new double[n]; pp->v_z = new double[n];
  pp->a_x = new double[n](); pp->a_y = new double[n](); pp->a_z = new double[n]();
  pp->ID = new int[n];
  pp->mass = nullptr;
  pp->uniform_mass = MassPerParticle;
  skip_lines(f, 1);
  int err = 0;
  for (int i = 0; i < n && !err; ++i)
    err = fscanf(f, "%lf %lf %lf", &pp->x[i], &pp->y[i], &pp->z[i]) != 3;
  skip_lines(f, 5);
  err = err || read_doubles(f, pp->v_x, n);
  skip_lines(f, 4);
  err = err || read_doubles(f, pp->v_y, n);
  skip_lines(f, 4);
  err = err || read_doubles(f, pp->v_z, n);
  skip_lines(f, 4);
  for (int i = 0; i < n && !err; ++i) err = fscanf(f, "%d", &pp->ID[i]) != 1;
  fclose(f);
  if (err) {
    printf("Error in reading %s\n", path);
    return 1;
  }
  for (int i = 0; i < n; ++i) {
    pp->v_x[i] = pp->v_x[i] * 1.0 / UnitVelocity_in_cm_per_s;
    pp->v_y[i] = pp->v_y[i] * 1.0 / UnitVelocity_in_cm_per_s;
    pp->v_z[i] = pp->v_z[i] * 1.0 / UnitVelocity_in_cm_per_s;
  }
  if (Units == 1)
    for (int i = 0; i < n; ++i) {
      pp->v_x[i] = pp->v_x[i] * 1.0 / (1.0 + Redshift) / sqrt(1.0 + Redshift);
      pp->v_y[i] = pp->v_y[i] * 1.0 / (1.0 + Redshift) / sqrt(1.0 + Redshift);
      pp->v_z[i] = pp->v_z[i] * 1.0 / (1.0 + Redshift) / sqrt(1.0 + Redshift);
    }
  return 0;
}

void free_particles(Physical_particle *pp) {
  delete[] pp->x; delete[] pp->y; delete[] pp->z;
  delete[] pp->v_x; delete[] pp->v_y; delete[] pp->v_z;
  delete[] pp->a_x; delete[] pp->a_y; delete[] pp->a_z;
  delete[] pp->ID;
  pp->x = pp->y = pp->z = pp->v_x = pp->v_y = pp->v_z = pp->a_x = pp->a_y = pp->a_z = nullptr;
  pp->ID = nullptr;
}
