// Test harness: this build's Viewer (ap-cloud-cosmology_amd/host/viewer.cpp)
// on particles read from a binary file, writing the same three snapshots as
// the reference-side ref_view (oracle/ref_shim.cpp).  Input (native
// doubles/ints): n, S (int); t0, t1, bbox[6] (double); x y z vx vy vz ax ay
// az (n doubles each); cx cy cz (S doubles); np type first (S ints);
// density potential p_x p_y p_z (S doubles).
// Usage: viewer_harness <input> <output dir>
#include <cstdio>
#include <cstring>
#include <vector>

#include "parameter.h"
#include "particle.h"
#include "viewer.h"

template <class T>
static bool get(FILE *f, T *p, size_t n) {
  return fread(p, sizeof(T), n, f) == n;
}

int main(int argc, char **argv) {
  if (argc < 3) return 2;
  FILE *f = fopen(argv[1], "rb");
  if (!f) return 3;
  int n = 0, S = 0;
  double t0 = 0, t1 = 0, bb[6];
  if (!get(f, &n, 1) || !get(f, &S, 1) || !get(f, &t0, 1) || !get(f, &t1, 1) || !get(f, bb, 6))
    return 4;
  std::vector<double> a[9];
  for (auto &v : a) {
    v.resize(n);
    if (!get(f, v.data(), n)) return 5;
  }
  Computational_particle cp;
  cp.size = S;
  cp.nc = S;
  cp.nv = cp.nb = 0;
  std::vector<double> *dv[] = {&cp.x, &cp.y, &cp.z};
  for (auto v : dv) {
    v->resize(S);
    if (!get(f, v->data(), S)) return 6;
  }
  std::vector<int> *iv[] = {&cp.NumberOfPhysicalParticles, &cp.Type,
                            &cp.IndexOfFirstPhysicalParticle};
  for (auto v : iv) {
    v->resize(S);
    if (!get(f, v->data(), S)) return 7;
  }
  std::vector<double> *fv[] = {&cp.Density, &cp.Potential, &cp.p_x, &cp.p_y, &cp.p_z};
  for (auto v : fv) {
    v->resize(S);
    if (!get(f, v->data(), S)) return 8;
  }
  fclose(f);
  Physical_particle pp;
  memset(&pp, 0, sizeof(pp));
  pp.size = n;
  pp.x = a[0].data(); pp.y = a[1].data(); pp.z = a[2].data();
  pp.v_x = a[3].data(); pp.v_y = a[4].data(); pp.v_z = a[5].data();
  pp.a_x = a[6].data(); pp.a_y = a[7].data(); pp.a_z = a[8].data();
  snprintf(OutputFile, sizeof(OutputFile), "%s", argv[2]);
  Viewer v(&pp, &cp);
  v.View_Physical_Particles_Limit(t0, bb[0], bb[2], bb[4], bb[1], bb[3], bb[5]);
  v.View_Physical_Particles(t1);
  v.View_Computational_Particles(t1);
  return 0;
}
