// Binds the reference-shaped host objects to one device context.
//
// The reference passes the same Physical_particle / Computational_particle
// objects to Octree, Mover and Solver; here those two objects name the device
// state: the first Octree built over a Physical_particle uploads it and owns
// an apc_ctx for the rest of the run.  From then on the device copy is
// authoritative — sync_host() copies positions, velocities and accelerations
// back into the caller's arrays (the Controller does so before output and at
// the end of a run).
#ifndef APC_HOST_BINDING_H
#define APC_HOST_BINDING_H
#include "apcloud.h"
#include "particle.h"

namespace apc_host {

struct Binding {
  apc_ctx *ctx = nullptr;
  Physical_particle *pp = nullptr;
  Computational_particle *cp = nullptr;
  apc_params prm;
};

// Returns the binding for pp, creating the context and uploading the
// particles when pp is new (or its size or box changed).
Binding *bind_particles(Physical_particle *pp, const apc_params &prm);
Binding *find_by_particles(const Physical_particle *pp);
Binding *find_by_comp(const Computational_particle *cp);
void attach_comp(Binding *b, Computational_particle *cp);
int sync_host(Physical_particle *pp);
void release(Physical_particle *pp);
// Fills cp from the device after a selection (sizes, geometry, Neigh24).
int export_comp(Binding *b, Computational_particle *cp);
// Copies Density, Potential and p_x..p_yz from the device into cp.
int export_fields(Binding *b, Computational_particle *cp);
void fatal(apc_ctx *ctx, int code, const char *what);

}  // namespace apc_host

#endif
