// Solver with the reference's public interface (solver.h:53-62).  solve()
// runs Build_Linear_System, Solve_Density, Solve_Potential and Differentiate
// on the device; the KSP options come from argv like KSPSetFromOptions.
#ifndef APC_HOST_SOLVER_H
#define APC_HOST_SOLVER_H
#include "apcloud.h"
#include "particle.h"

class Solver {
 public:
  Solver(Computational_particle *computational_particle1, int argc_o, char *args_o[]);
  ~Solver();
  int Set_Boundary_Condition(int bc) { BC = bc; return 0; }
  int Set_NumberOfGFDNeighbors(int n) { lm = n; return 0; }
  int Set_Constant_For_Density_Estimator(double c) { Constant_For_Density_Estimator = c; return 0; }
  int Set_Constant_For_Potential_Solver(double c) { Constant_For_Potential_Solver = c; return 0; }
  int Set_Constant_For_Field_Solver(double c) { Constant_For_Field_Solver = c; return 0; }
  int Set_BoundingBox(double x1, double x2, double y1, double y2, double z1, double z2) {
    x_min = x1; x_max = x2; y_min = y1; y_max = y2; z_min = z1; z_max = z2;
    return 0;
  }
  int solve();

  // -ksp_* options parsed from argv (both solves, like a global
  // KSPSetFromOptions); -ksp_gmres_precision selects the mixed-precision form
  apc_ksp_opts ksp_options() const { return opts; }
  void set_precision(int p) { opts.precision = p; }
  apc_solve_info last_info() const { return info; }
  int export_fields = 1;  // fill Density/Potential/p_* of the computational particles

 private:
  Computational_particle *cp;
  double Constant_For_Density_Estimator, Constant_For_Potential_Solver, Constant_For_Field_Solver;
  int BC, lm, ln, dirichlet_point;
  double x_min, x_max, y_min, y_max, z_min, z_max;
  int argc;
  char **args;
  apc_ksp_opts opts;
  int opts_status = 0;  // non-zero: argv named an option value this build does not support
  apc_solve_info info;
};

// Parses the PETSc options the path honours from argv.  Returns 0, or
// APC_ERR_UNSUPPORTED for a -pc_type other than jacobi / gamg / amg.
int apc_parse_ksp_options(int argc, char **argv, apc_ksp_opts *o);

#endif
