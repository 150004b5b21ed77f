// .param reader: the reference's Gadget-style "Tag value %comment" format
// (parameter.cpp:28-222) with its 21 mandatory tags and fatal errors, plus the
// optional tags this build adds (defaults reproduce the reference).
#ifndef APC_HOST_PARAMETER_H
#define APC_HOST_PARAMETER_H

// Reads the file into the globals below.  Returns 0, or 1 after printing the
// reference's messages for a missing file, an unknown or repeated tag, or a
// missing mandatory tag (the reference then calls exit(0)).
int read_parameterfile(const char *fname);

extern double Omega, OmegaLambda, Redshift, HubbleParam, GravitionalConstant;
extern double UnitLength_in_cm, UnitMass_in_g, UnitVelocity_in_cm_per_s;
extern char ICFile[200], FileBase[200], OutputFile[200];
extern int ICFileType;
extern double DeltaA, EndRedShift;
extern int MaxOctreeLevel;
extern double CellSizeParameter;
extern int NumberOfNeighbors, DensityEstimator;
extern double BoundingBox_min_x, BoundingBox_max_x, BoundingBox_min_y, BoundingBox_max_y;
extern double BoundingBox_min_z, BoundingBox_max_z;
extern int Periodic;
extern double MassPerParticle;

// optional tags (this build)
extern int Units;             // 0 reference default build, 1 COSMOLOGY_UNIT, 2 normalised comoving, 3 Newtonian
extern double OmegaCurv;
extern int AdaptiveTimeStep;  // reference: on (apcloud_macro.h:25-27)
extern int LeafCapacity;      // >0: open a cell iff it holds more particles
extern int BoundaryAtFaces;
extern int BoundaryValue;     // 0: bv = 0 (reference), 1: uniform-sphere potential
extern double BoundaryMass, BoundaryRadius, BoundaryCenter_x, BoundaryCenter_y, BoundaryCenter_z;
extern int WriteInterval;     // reference: 10 (test.cpp:66); 0 disables output
extern int GmresPrecision;    // 0 FP64, 1 FP32 inner / FP64 refinement
extern char FinalStateFile[200];  // raw FP64 dump of x,y,z,v,a after the run (empty: none)

#endif
