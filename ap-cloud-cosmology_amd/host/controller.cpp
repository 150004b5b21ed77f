#include "controller.h"

#include <chrono>
#include <cmath>
#include <cstdio>
#include <vector>

#include "apcloud.h"
#include "binding.h"
#include "parameter.h"

namespace {
const double kPiRef = 3.1415926;       // apcloud_macro.h:1
const double kMpcCm = 3.085677581e24;  // apcloud_macro.h:2
double now() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}
}  // namespace

Controller::Controller(Physical_particle *pp, Computational_particle *cp, double Start_A1,
                       double End_A1, double Delta_A1, Solver *solver1, Viewer *viewer1,
                       Mover *mover1, int Write_Interval1, int Boundary_Condition1, double x_min1,
                       double x_max1, double y_min1, double y_max1, double z_min1, double z_max1,
                       int Octree_Depth1, double GFD_Constant1, int lm1)
    : physical_particle(pp), computational_particle(cp), solver(solver1), viewer(viewer1),
      mover(mover1), x_min(x_min1), x_max(x_max1), y_min(y_min1), y_max(y_max1), z_min(z_min1),
      z_max(z_max1), BC(Boundary_Condition1), WriteInterval(Write_Interval1),
      Octree_Depth(Octree_Depth1), lm(lm1), GFD_Constant(GFD_Constant1), Start_A(Start_A1),
      End_A(End_A1), Delta_A(Delta_A1) {}

// controller.cpp:13-52.  Units 0 is the reference's default build, 1 its
// COSMOLOGY_UNIT build; 2 (normalised comoving: mean density 1 in the box,
// C_dens = MassPerParticle, curvature term OmegaCurv) and 3 (Newtonian, G) are
// this build's additions.
double Controller::Get_Constant_For_Density_Estimator(double a) {
  switch (Units) {
    case 1:
      return 8.0 * a * a * a * kPiRef * GravitionalConstant / 3.0 / 100000.0 / 100000.0 * kMpcCm *
             kMpcCm / Omega * UnitMass_in_g / UnitLength_in_cm / UnitLength_in_cm /
             UnitLength_in_cm * MassPerParticle;
    case 2:
    case 3:
      return MassPerParticle;
    default:
      return 1.0;
  }
}

double Controller::Get_Constant_For_Potential_Solver(double a) {
  switch (Units) {
    case 1:
    case 2:
      return 3.0 / 2.0 * Omega / a;
    case 3:
      return 4.0 * M_PI * GravitionalConstant;
    default:
      return 1000.0;
  }
}

double Controller::Get_Constant_For_Field_Solver(double) { return -1.0; }

double Controller::Get_Constant_For_Mover_Position(double a) {
  switch (Units) {
    case 1:
      return 1.0 / sqrt(Omega / a + OmegaLambda * a * a);
    case 2:
      return 1.0 / sqrt(Omega / a + OmegaCurv + OmegaLambda * a * a);
    default:
      return 1.0;
  }
}

double Controller::Get_Constant_For_Mover_Velocity(double a) {
  switch (Units) {
    case 1:
      return 1.0 / sqrt(Omega / a + OmegaLambda * a * a) / a / a;
    case 2:
      return 1.0 / sqrt(Omega / a + OmegaCurv + OmegaLambda * a * a) / a / a;
    default:
      return 1.0;
  }
}

// controller.cpp:54-241 on one process: the scalar broadcasts disappear, the
// per-phase timers stay; host particle arrays are refreshed before output.
int Controller::solve() {
  int errorflag = 0;
  const double dynamic_time_step_size = 0.05 * Delta_A;
  std::vector<double> dynamic_time;
  double t = Start_A;
  double max_pchange, max_vchange, delta_t;
  mover->Set_BoundingBox(x_min, x_max, y_min, y_max, z_min, z_max);
  mover->Set_maxnc(10000);
  mover->Set_Octree_Depth(Octree_Depth);
  mover->Set_Constant_For_GFD_Constant(GFD_Constant);
  mover->Set_BC(BC);
  mover->Set_Constant_For_Mover_Position(Get_Constant_For_Mover_Position(t));
  mover->Set_Constant_For_Mover_Velocity(Get_Constant_For_Mover_Velocity(t));
  if (viewer && WriteInterval > 0) viewer->View_Physical_Particles_Limit(t, x_min, y_min, z_min, x_max, y_max, z_max);
  solver->Set_BoundingBox(x_min, x_max, y_min, y_max, z_min, z_max);
  solver->Set_Boundary_Condition(BC);
  solver->Set_NumberOfGFDNeighbors(lm);
  max_pchange = mover->move_position(0.5 * Delta_A);  // :106-111
  if (verbose) printf("Maximum position change from %f to %f is: %f\n", t, t + 0.5 * Delta_A, max_pchange);
  t += 0.5 * Delta_A;
  int ite = 0;
  delta_t = Delta_A;
  t_select = t_solve = t_move = t_output = t_total = 0;
  while (t < End_A && (max_steps <= 0 || ite < max_steps)) {
    double t0 = now();
    if (viewer && WriteInterval > 0 && ite % WriteInterval == 0) {
      double to = now();
      apc_host::sync_host(physical_particle);
      viewer->View_Physical_Particles(t);
      viewer->View_Computational_Particles(t);
      t_output += now() - to;
    }
    mover->Set_Constant_For_Mover_Velocity(Get_Constant_For_Mover_Velocity(t));
    double ts = now();
    errorflag = mover->select_computational_particles();
    t_select += now() - ts;
    if (errorflag < 0) break;
    if (verbose) printf("Number of computational particles=%d\n", errorflag);
    if (AdaptiveTimeStep) {  // :146-162
      double max_density = 0.0;
      apc_host::Binding *b = apc_host::find_by_particles(physical_particle);
      if (!b) {
        fprintf(stdout, "Controller::solve: adaptive time step without selected particles\n");
        errorflag = APC_ERR_ARG;
        break;
      }
      int r = apc_max_density(b->ctx, &max_density);
      if (r) {
        apc_host::fatal(b->ctx, r, "Controller::solve (adaptive time step)");
        errorflag = r;
        break;
      }
      if (!(max_density > 0.0)) {  // every node empty: the reference would divide by zero
        fprintf(stdout, "Controller::solve: maximum node density is %g\n", max_density);
        errorflag = APC_ERR_ARG;
        break;
      }
      double min_interparticle_distance = 1.0 / pow(max_density, 1.0 / 2);
      delta_t = dynamic_time_step_size * min_interparticle_distance;
      if (delta_t > Delta_A) delta_t = Delta_A;
    } else {
      delta_t = Delta_A;
    }
    if ((t + delta_t) > End_A) delta_t = End_A - t;
    mover->Set_Constant_For_Mover_Velocity(Get_Constant_For_Mover_Velocity(t));
    mover->Set_Constant_For_Mover_Position(Get_Constant_For_Mover_Position(t + delta_t / 2));
    solver->Set_Constant_For_Density_Estimator(Get_Constant_For_Density_Estimator(t));
    solver->Set_Constant_For_Potential_Solver(Get_Constant_For_Potential_Solver(t));
    solver->Set_Constant_For_Field_Solver(Get_Constant_For_Field_Solver(t));
    double tv = now();
    errorflag = solver->solve();
    t_solve += now() - tv;
    if (errorflag) break;
    double tm = now();
    max_vchange = mover->move_velocity(delta_t);
    max_pchange = mover->move_position(delta_t);
    t_move += now() - tm;
    if (verbose) {
      printf("Maximum velocity change from %f to %f is: %e\n", t, t + delta_t, max_vchange);
      printf("Maximum position change from %f to %f is: %e\n", t, t + delta_t, max_pchange);
    }
    t_total += now() - t0;
    t += delta_t;
    dynamic_time.push_back(t);
    ite++;
  }
  steps_done = ite;
  apc_host::sync_host(physical_particle);
  if (viewer && WriteInterval > 0) {
    viewer->View_Physical_Particles(t);
    viewer->View_Computational_Particles(t);
  }
  if (verbose) {
    printf("Total running time: %f\n", t_total);
    printf("Running time for output: %f\n", t_output);
    printf("Running time for computational particles select: %f\n", t_select);
    printf("Running time for density estimation, potential solver and differentiation: %f\n", t_solve);
    printf("Running time for particle position and velocity update: %f\n", t_move);
    for (double v : dynamic_time) printf("%f\n", v);
  }
  return errorflag < 0 ? errorflag : 0;
}
