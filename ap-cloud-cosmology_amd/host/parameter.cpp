// .param reader (format and error behaviour of parameter.cpp:28-222).
#include "parameter.h"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

double Omega, OmegaLambda, Redshift, HubbleParam, GravitionalConstant;
double UnitLength_in_cm, UnitMass_in_g, UnitVelocity_in_cm_per_s;
char ICFile[200], FileBase[200], OutputFile[200];
int ICFileType;
double DeltaA, EndRedShift;
int MaxOctreeLevel;
double CellSizeParameter;
int NumberOfNeighbors, DensityEstimator;
double BoundingBox_min_x, BoundingBox_max_x, BoundingBox_min_y, BoundingBox_max_y;
double BoundingBox_min_z, BoundingBox_max_z;
int Periodic;
double MassPerParticle;

int Units = 0;
double OmegaCurv = 0.0;
int AdaptiveTimeStep = 1;
int LeafCapacity = 0;
int BoundaryAtFaces = 0;
int BoundaryValue = 0;
double BoundaryMass = 1.0, BoundaryRadius = 0.5;
double BoundaryCenter_x = 0.5, BoundaryCenter_y = 0.5, BoundaryCenter_z = 0.5;
int WriteInterval = 10;
int GmresPrecision = 0;
char FinalStateFile[200] = "";

namespace {
enum Kind { KFLOAT, KSTRING, KINT };
struct Tag {
  const char *name;
  void *addr;
  Kind kind;
  bool mandatory;
  bool seen;
};
}  // namespace

int read_parameterfile(const char *fname) {
  std::vector<Tag> tags = {
      {"Omega", &Omega, KFLOAT, true, false},
      {"OmegaLambda", &OmegaLambda, KFLOAT, true, false},
      {"HubbleParam", &HubbleParam, KFLOAT, true, false},
      {"Redshift", &Redshift, KFLOAT, true, false},
      {"GravitionalConstant", &GravitionalConstant, KFLOAT, true, false},
      {"UnitLength_in_cm", &UnitLength_in_cm, KFLOAT, true, false},
      {"UnitMass_in_g", &UnitMass_in_g, KFLOAT, true, false},
      {"UnitVelocity_in_cm_per_s", &UnitVelocity_in_cm_per_s, KFLOAT, true, false},
      {"ICFile", ICFile, KSTRING, true, false},
      {"ICFileType", &ICFileType, KINT, true, false},
      {"FileBase", FileBase, KSTRING, true, false},
      {"OutputFile", OutputFile, KSTRING, true, false},
      {"BoundingBox_min_x", &BoundingBox_min_x, KFLOAT, true, false},
      {"BoundingBox_max_x", &BoundingBox_max_x, KFLOAT, true, false},
      {"BoundingBox_min_y", &BoundingBox_min_y, KFLOAT, true, false},
      {"BoundingBox_max_y", &BoundingBox_max_y, KFLOAT, true, false},
      {"BoundingBox_min_z", &BoundingBox_min_z, KFLOAT, true, false},
      {"BoundingBox_max_z", &BoundingBox_max_z, KFLOAT, true, false},
      {"Periodic", &Periodic, KINT, true, false},
      {"DeltaA", &DeltaA, KFLOAT, true, false},
      {"EndRedShift", &EndRedShift, KFLOAT, true, false},
      {"MassPerParticle", &MassPerParticle, KFLOAT, true, false},
      {"MaxOctreeLevel", &MaxOctreeLevel, KINT, true, false},
      {"CellSizeParameter", &CellSizeParameter, KFLOAT, true, false},
      {"NumberOfNeighbors", &NumberOfNeighbors, KINT, true, false},
      {"DensityEstimator", &DensityEstimator, KINT, true, false},
      // optional tags of this build
      {"Units", &Units, KINT, false, false},
      {"OmegaCurv", &OmegaCurv, KFLOAT, false, false},
      {"AdaptiveTimeStep", &AdaptiveTimeStep, KINT, false, false},
      {"LeafCapacity", &LeafCapacity, KINT, false, false},
      {"BoundaryAtFaces", &BoundaryAtFaces, KINT, false, false},
      {"BoundaryValue", &BoundaryValue, KINT, false, false},
      {"BoundaryMass", &BoundaryMass, KFLOAT, false, false},
      {"BoundaryRadius", &BoundaryRadius, KFLOAT, false, false},
      {"BoundaryCenter_x", &BoundaryCenter_x, KFLOAT, false, false},
      {"BoundaryCenter_y", &BoundaryCenter_y, KFLOAT, false, false},
      {"BoundaryCenter_z", &BoundaryCenter_z, KFLOAT, false, false},
      {"WriteInterval", &WriteInterval, KINT, false, false},
      {"GmresPrecision", &GmresPrecision, KINT, false, false},
      {"FinalStateFile", FinalStateFile, KSTRING, false, false},
  };
  int errorFlag = 0;
  FILE *fd = fopen(fname, "r");
  if (!fd) {
    fprintf(stdout, "Parameter file %s not found.\n", fname);
    return 1;
  }
  char buf[400], b1[200], b2[200], b3[200];
  while (fgets(buf, sizeof(buf), fd)) {
    if (sscanf(buf, "%199s%199s%199s", b1, b2, b3) < 2) continue;
    if (b1[0] == '%') continue;
    Tag *hit = nullptr;
    for (auto &t : tags)
      if (!t.seen && strcmp(b1, t.name) == 0) { hit = &t; break; }
    if (!hit) {
      fprintf(stdout, "Error in file %s:   Tag '%s' not allowed or multiple defined.\n", fname, b1);
      errorFlag = 1;
      continue;
    }
    hit->seen = true;
    switch (hit->kind) {
      case KFLOAT: *(double *)hit->addr = atof(b2); break;
      case KSTRING: snprintf((char *)hit->addr, 200, "%s", b2); break;
      case KINT: *(int *)hit->addr = atoi(b2); break;
    }
  }
  fclose(fd);
  for (auto &t : tags)
    if (t.mandatory && !t.seen) {
      fprintf(stdout, "Error. I miss a value for tag '%s' in parameter file '%s'.\n", t.name, fname);
      errorFlag = 1;
    }
  return errorFlag;
}
