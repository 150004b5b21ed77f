// Particle containers with the reference's member names and meaning
// (particle.h:5-53): the drop-in C++ classes exchange data through these.
#ifndef APC_HOST_PARTICLE_H
#define APC_HOST_PARTICLE_H
#include <vector>

struct Physical_particle {
  int size;        // number of physical particles
  double *x, *y, *z;
  double *v_x, *v_y, *v_z;
  int *ID;
  double *a_x, *a_y, *a_z;
  double *mass;    // never allocated by the reference's IC reader
  double uniform_mass;
};

struct Computational_particle {
  int size;        // nc + nv + nb
  std::vector<double> x, y, z, CellLength;
  std::vector<int> IndexOfFirstPhysicalParticle;  // first particle in Morton order
  std::vector<int> NumberOfPhysicalParticles;
  std::vector<int> Type;                           // 0 comp, -1 vacuum, 1 boundary
  std::vector<int> FirstNeigh, NumberNeigh, Neigh24;
  std::vector<double> Density, Potential;
  std::vector<double> p_x, p_y, p_z, p_xx, p_yy, p_zz, p_xy, p_xz, p_yz;
  int nc, nv, nb;
};

int clear(Computational_particle *computational_particle);
int remove(Computational_particle *computational_particle);

#endif
