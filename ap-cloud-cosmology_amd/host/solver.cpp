#include "solver.h"

#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "binding.h"

int apc_parse_ksp_options(int argc, char **argv, apc_ksp_opts *o) {
  int status = 0;
  // PETSc defaults (KSPCreate / KSPGMRES)
  o->rtol = 1e-5;
  o->atol = 1e-50;
  o->dtol = 1e5;
  o->max_it = 10000;
  o->restart = 30;
  o->monitor = 0;
  o->precision = 0;
  o->pc_type = 0;  // PCJACOBI (solver.cpp:201,273)
  for (int i = 1; i < argc; ++i) {
    const char *a = argv[i];
    const char *v = (i + 1 < argc) ? argv[i + 1] : nullptr;
    if (!strcmp(a, "-ksp_rtol") && v) o->rtol = atof(v), ++i;
    else if (!strcmp(a, "-ksp_atol") && v) o->atol = atof(v), ++i;
    else if (!strcmp(a, "-ksp_divtol") && v) o->dtol = atof(v), ++i;
    else if (!strcmp(a, "-ksp_max_it") && v) o->max_it = atoi(v), ++i;
    else if (!strcmp(a, "-ksp_gmres_restart") && v) o->restart = atoi(v), ++i;
    else if (!strcmp(a, "-ksp_gmres_precision") && v) o->precision = atoi(v), ++i;
    else if (!strcmp(a, "-ksp_monitor")) o->monitor = 1;
    else if (!strcmp(a, "-pc_type") && v) {
      // jacobi: the reference's PCJACOBI; gamg/amg: the device multigrid
      // (right preconditioning, true-residual convergence test)
      if (!strcmp(v, "jacobi")) o->pc_type = 0;
      else if (!strcmp(v, "gamg") || !strcmp(v, "amg")) o->pc_type = 1;
      else {
        // PCSetType fails on a type it does not know, and KSPSetFromOptions
        // passes that error on (solver.cpp:202,274 CHKERRQ): an error, not a
        // silent Jacobi solve
        fprintf(stderr, "apcloud: -pc_type %s not supported (jacobi, gamg)\n", v);
        status = APC_ERR_UNSUPPORTED;
      }
      ++i;
    }
  }
  return status;
}

Solver::Solver(Computational_particle *computational_particle1, int argc_o, char *args_o[])
    : cp(computational_particle1), argc(argc_o), args(args_o) {
  // solver.cpp:8-18 defaults
  BC = 1;
  Constant_For_Density_Estimator = 1.0;
  Constant_For_Potential_Solver = 1.0;
  Constant_For_Field_Solver = 1.0;
  lm = 17;
  ln = 9;
  dirichlet_point = 1;
  x_min = y_min = z_min = 0.0;
  x_max = y_max = z_max = 1.0;
  opts_status = apc_parse_ksp_options(argc, args, &opts);
  memset(&info, 0, sizeof(info));
}

Solver::~Solver() {}

int Solver::solve() {
  if (opts_status) {
    fprintf(stdout, "Solver::solve: unsupported PETSc options (%d)\n", opts_status);
    return opts_status;
  }
  apc_host::Binding *b = apc_host::find_by_comp(cp);
  if (!b) {
    fprintf(stdout, "Solver::solve: no computational particles selected\n");
    return -1;
  }
  int r = apc_solver_set_neighbors(b->ctx, lm);
  if (!r)
    r = apc_solver_set_constants(b->ctx, Constant_For_Density_Estimator,
                                 Constant_For_Potential_Solver, Constant_For_Field_Solver);
  if (!r) r = apc_solver_solve(b->ctx, &opts, &opts, &info);
  if (r) {
    apc_host::fatal(b->ctx, r, "Solver::solve");
    return r;
  }
  if (export_fields) r = apc_host::export_fields(b, cp);
  return r;
}
