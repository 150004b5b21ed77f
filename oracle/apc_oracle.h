/*
 * TEST INFRASTRUCTURE ONLY — the oracle.
 *
 * A CPU restatement (FP64, g++ -O2 -ffp-contract=off) of the AP-Cloud
 * per-timestep solve-and-move path of the reference (octree.cpp, solver.cpp,
 * mover.cpp, controller.cpp).  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it, and only as the checker or as the
 * timed CPU baseline — never as part of the product path.
 *
 * Parity pins (what each function is checked against) are listed in
 * oracle/README.md.
 */
#ifndef APC_ORACLE_H
#define APC_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  double bbox[6];          /* x_min,x_max,y_min,y_max,z_min,z_max */
  int depth;               /* MaxOctreeLevel */
  int bc;                  /* 1 periodic, 0 non-periodic */
  double c;                /* CellSizeParameter */
  int order;               /* criterion order (reference: 2) */
  int mindepth, maxdepth;  /* selection depth window (reference: 0, depth) */
  int lm;                  /* NumberOfNeighbors */
  int dirichlet_point;     /* periodic pin of row 0 (reference: 1) */
  int boundary_at_faces;   /* 0: boundary nodes at +-1 (reference), 1: at box faces */
  int bv_kind;             /* 0: bv=0 (reference); 1: uniform-sphere potential */
  double bv_mass, bv_radius, bv_center[3], bv_G;
  /* controller */
  int units;               /* 0 reference default, 1 COSMOLOGY_UNIT, 2 normalised comoving, 3 Newtonian */
  double omega, omega_lambda, omega_curv, G, unit_mass, unit_length, mass_per_particle;
  int adaptive_dt;         /* ADAPTIVE_TIME_STEP */
  /* GMRES (PETSc semantics) for rho and phi */
  double rtol, atol, dtol; int max_it, restart;
  int nthreads;            /* OpenMP threads for the parallel loops (0 = default) */
  int leaf_capacity;       /* >0: open a cell iff n_p > leaf_capacity (the build's added
                              criterion); 0: the reference's n*h^(2o-2) > c */
} OrcParams;

typedef struct {
  int its_rho, reason_rho; double rnorm_rho;
  int its_phi, reason_phi; double rnorm_phi;
} OrcSolveInfo;

typedef struct OrcSim OrcSim;

OrcSim *orc_new(const OrcParams *p);
void orc_free(OrcSim *s);
int orc_set_particles(OrcSim *s, int n, const double *x, const double *y, const double *z,
                      const double *vx, const double *vy, const double *vz);
int orc_get_particles(OrcSim *s, double *x, double *y, double *z, double *vx, double *vy,
                      double *vz, double *ax, double *ay, double *az);

uint64_t orc_key64(double x, double y, double z, const double *bbox, int depth);
int orc_build_tree(OrcSim *s);
int orc_tree_info(OrcSim *s, int *len, int *numnodes /* depth+1 */);
int orc_tree_sorted(OrcSim *s, uint64_t *key, int *idx);
int orc_tree_nodes(OrcSim *s, uint64_t *key, int *depth, int *first, int *count, int *fchild,
                   int *nchild, double *lo /* 3*len */, double *hi /* 3*len */);

int orc_select(OrcSim *s, double c, int order, int mindepth, int maxdepth);
int orc_comp_sizes(OrcSim *s, int *nc, int *nv, int *nb, int *nneigh);
int orc_comp_get(OrcSim *s, double *x, double *y, double *z, double *h, int *np, int *first,
                 int *type, int *fneigh, int *nneigh, int *neigh24);

/* A caller-filled computational-particle set (Types 0, -1, 1 by position). */
int orc_set_comp(OrcSim *s, int nc, int nv, int nb, const double *x, const double *y,
                 const double *z, const double *h, const int *np, const int *fneigh,
                 const int *nneigh, const int *neigh24, int total_neigh);
int orc_neighbors(OrcSim *s, int *nbr, double *dist);
int orc_build_system(OrcSim *s, int *cols, double *a1, double *a2, double *d1, double *d2,
                     double *rhs1, double *pbnd);
int orc_set_constants(OrcSim *s, double c_dens, double c_pot, double c_field);
int orc_solve(OrcSim *s, OrcSolveInfo *info);
int orc_comp_fields(OrcSim *s, double *rho, double *phi, double *g /* 3M */, double *H /* 6M */);
int orc_set_potential(OrcSim *s, const double *phi);
int orc_set_gradients(OrcSim *s, const double *g /* 3M */, const double *H /* 6M */);
int orc_differentiate(OrcSim *s);
double orc_move_velocity(OrcSim *s, double Cv, double dt);
double orc_move_position(OrcSim *s, double Cx, double dt);
int orc_run(OrcSim *s, double start_a, double end_a, double delta_a, int max_steps,
            double *timing /* [select, solve, move] seconds */, int *steps_done);

/* pieces pinned individually */
int orc_wlsq(int m, int n, int nrhs, double *B, double *b);
void orc_quicksort(int n, double *a, int *b);
int orc_gmres_csr(int n, const int *rowptr, const int *col, const double *val, const double *b,
                  double *x, double rtol, double atol, double dtol, int max_it, int restart,
                  int *its, int *reason, double *rnorm);
void orc_coefficients(const OrcParams *p, double a, double *c_dens, double *c_pot,
                      double *c_field, double *cx, double *cv);

#ifdef __cplusplus
}
#endif
#endif
