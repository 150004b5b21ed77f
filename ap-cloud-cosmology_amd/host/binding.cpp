#include "binding.h"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <vector>

namespace apc_host {

namespace {
std::map<const Physical_particle *, Binding *> &by_pp() {
  static std::map<const Physical_particle *, Binding *> m;
  return m;
}
std::map<const Computational_particle *, Binding *> &by_cp() {
  static std::map<const Computational_particle *, Binding *> m;
  return m;
}
}  // namespace

void fatal(apc_ctx *ctx, int code, const char *what) {
  fprintf(stdout, "%s failed (%d): %s\n", what, code, ctx ? apc_last_error(ctx) : "");
  fflush(stdout);
}

Binding *bind_particles(Physical_particle *pp, const apc_params &prm) {
  auto it = by_pp().find(pp);
  if (it != by_pp().end()) {
    Binding *b = it->second;
    bool same = b->prm.depth == prm.depth && b->prm.bc == prm.bc &&
                memcmp(b->prm.bbox, prm.bbox, sizeof(prm.bbox)) == 0;
    if (same) return b;
    // the box changed: carry the device state back, then rebuild the context
    sync_host(pp);
    release(pp);
  }
  Binding *b = new Binding();
  b->pp = pp;
  b->prm = prm;
  int r = apc_create(&prm, &b->ctx);
  if (r) {
    fatal(nullptr, r, "apc_create");
    delete b;
    return nullptr;
  }
  r = apc_set_particles(b->ctx, pp->size, pp->x, pp->y, pp->z, pp->v_x, pp->v_y, pp->v_z);
  if (r) {
    fatal(b->ctx, r, "apc_set_particles");
    apc_destroy(b->ctx);
    delete b;
    return nullptr;
  }
  by_pp()[pp] = b;
  return b;
}

Binding *find_by_particles(const Physical_particle *pp) {
  auto it = by_pp().find(pp);
  return it == by_pp().end() ? nullptr : it->second;
}

Binding *find_by_comp(const Computational_particle *cp) {
  auto it = by_cp().find(cp);
  return it == by_cp().end() ? nullptr : it->second;
}

void attach_comp(Binding *b, Computational_particle *cp) {
  b->cp = cp;
  by_cp()[cp] = b;
}

int sync_host(Physical_particle *pp) {
  Binding *b = find_by_particles(pp);
  if (!b) return 0;
  return apc_get_particles(b->ctx, pp->x, pp->y, pp->z, pp->v_x, pp->v_y, pp->v_z, pp->a_x,
                           pp->a_y, pp->a_z);
}

void release(Physical_particle *pp) {
  Binding *b = find_by_particles(pp);
  if (!b) return;
  if (b->cp) by_cp().erase(b->cp);
  by_pp().erase(pp);
  apc_destroy(b->ctx);
  delete b;
}

int export_comp(Binding *b, Computational_particle *cp) {
  int nc, nv, nb, nn;
  int r = apc_get_comp_sizes(b->ctx, &nc, &nv, &nb, &nn);
  if (r) return r;
  const int S = nc + nv + nb;
  clear(cp);
  cp->x.resize(S); cp->y.resize(S); cp->z.resize(S); cp->CellLength.resize(S);
  cp->NumberOfPhysicalParticles.resize(S); cp->IndexOfFirstPhysicalParticle.resize(S);
  cp->Type.resize(S); cp->FirstNeigh.resize(S); cp->NumberNeigh.resize(S);
  cp->Neigh24.resize(nn);
  apc_comp_view v;
  memset(&v, 0, sizeof(v));
  v.x = cp->x.data(); v.y = cp->y.data(); v.z = cp->z.data(); v.cell_length = cp->CellLength.data();
  v.np = cp->NumberOfPhysicalParticles.data(); v.first = cp->IndexOfFirstPhysicalParticle.data();
  v.type = cp->Type.data(); v.first_neigh = cp->FirstNeigh.data();
  v.number_neigh = cp->NumberNeigh.data(); v.neigh24 = cp->Neigh24.data();
  r = apc_get_comp(b->ctx, &v);
  if (r) return r;
  cp->nc = nc; cp->nv = nv; cp->nb = nb; cp->size = S;
  // octree.cpp:2982-2996: field arrays sized for every node
  std::vector<double> *f[] = {&cp->Density, &cp->Potential, &cp->p_x, &cp->p_y, &cp->p_z,
                              &cp->p_xx, &cp->p_yy, &cp->p_zz, &cp->p_xy, &cp->p_xz, &cp->p_yz};
  for (auto q : f) q->assign(S, 0.0);
  return 0;
}

int export_fields(Binding *b, Computational_particle *cp) {
  apc_comp_view v;
  memset(&v, 0, sizeof(v));
  const int S = cp->nc + cp->nv + cp->nb;
  std::vector<double> *f[] = {&cp->Density, &cp->Potential, &cp->p_x, &cp->p_y, &cp->p_z,
                              &cp->p_xx, &cp->p_yy, &cp->p_zz, &cp->p_xy, &cp->p_xz, &cp->p_yz};
  for (auto q : f)
    if ((int)q->size() < S) q->assign(S, 0.0);
  v.density = cp->Density.data(); v.potential = cp->Potential.data();
  v.p_x = cp->p_x.data(); v.p_y = cp->p_y.data(); v.p_z = cp->p_z.data();
  v.p_xx = cp->p_xx.data(); v.p_yy = cp->p_yy.data(); v.p_zz = cp->p_zz.data();
  v.p_xy = cp->p_xy.data(); v.p_xz = cp->p_xz.data(); v.p_yz = cp->p_yz.data();
  return apc_get_comp(b->ctx, &v);
}

}  // namespace apc_host

// particle.cpp:3-51 equivalents
int clear(Computational_particle *c) {
  c->x.clear(); c->y.clear(); c->z.clear(); c->CellLength.clear();
  c->IndexOfFirstPhysicalParticle.clear(); c->NumberOfPhysicalParticles.clear(); c->Type.clear();
  c->FirstNeigh.clear(); c->NumberNeigh.clear(); c->Neigh24.clear();
  c->Density.clear(); c->Potential.clear();
  c->p_x.clear(); c->p_y.clear(); c->p_z.clear(); c->p_xx.clear(); c->p_yy.clear();
  c->p_zz.clear(); c->p_xy.clear(); c->p_xz.clear(); c->p_yz.clear();
  return 0;
}

int remove(Computational_particle *c) {
  clear(c);
  c->x.shrink_to_fit(); c->y.shrink_to_fit(); c->z.shrink_to_fit();
  c->Neigh24.shrink_to_fit();
  return 0;
}
