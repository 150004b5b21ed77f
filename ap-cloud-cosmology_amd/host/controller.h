// Controller with the reference's public interface (controller.h:12-56): the
// leapfrog in scale factor a driving Mover and Solver.
#ifndef APC_HOST_CONTROLLER_H
#define APC_HOST_CONTROLLER_H
#include "mover.h"
#include "particle.h"
#include "solver.h"
#include "viewer.h"

class Controller {
 public:
  Controller(Physical_particle *physical_particle1, Computational_particle *computational_particle1,
             double Start_A1, double End_A1, double Delta_A1, Solver *solver1, Viewer *viewer1,
             Mover *mover1, int Write_Interval1, int Boundary_Condition1, double x_min1,
             double x_max1, double y_min1, double y_max1, double z_min1, double z_max1,
             int Octree_Depth1, double GFD_Constant1, int lm1);
  virtual ~Controller() {}
  virtual int solve();

  // phase wall times of the last solve() (controller.cpp:229-233), seconds
  double t_select = 0, t_solve = 0, t_move = 0, t_output = 0, t_total = 0;
  int steps_done = 0;
  int max_steps = 0;  // 0: run to End_A
  int verbose = 1;

 private:
  double Get_Constant_For_Density_Estimator(double t);
  double Get_Constant_For_Potential_Solver(double t);
  double Get_Constant_For_Field_Solver(double t);
  double Get_Constant_For_Mover_Position(double t);
  double Get_Constant_For_Mover_Velocity(double t);

  Physical_particle *physical_particle;
  Computational_particle *computational_particle;
  Solver *solver;
  Viewer *viewer;
  Mover *mover;
  double x_min, x_max, y_min, y_max, z_min, z_max;
  int BC, WriteInterval, Octree_Depth, lm;
  double GFD_Constant, Start_A, End_A, Delta_A;
};

#endif
