// Ensembles: many independent simulations advanced by one launch per phase of
// the fused skin-list step.  Each member keeps its own workspace; the batched
// kernels read a member's arguments from a device table and find their member
// and block in a block map (EnsBlock), then run the same device code as the
// single-simulation kernels, so a member's results are bitwise those of its
// own sph_step.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "sph_device.h"
#include "sph_internal.h"

namespace sph {

struct SweepPlan;  // sweep.h

// One member's fused-step arguments (those of sph_step on bank 2).
struct EnsMember {
  sph_ws* ws;
  sph_solver_params prm;
  float dt, ctvf;
  sph_bc_spec bc;
  sph_state_in in;
  float* r_out;
  sph_state_out out;
  const int64_t* prev_info;
  int64_t* info;
  int64_t max_occupancy;
};

// A block map on the device: workgroup b of a batched launch serves d[b].
struct EnsMap {
  EnsBlock* d = nullptr;
  int blocks = 0;
};

// nl.hip: device argument tables of the batched neighbour-search phases.
struct EnsNL;
int ens_nl_create(const std::vector<EnsMember>& m, EnsNL** out);
void ens_nl_destroy(EnsNL* t);
// kick/drift + rebuild test, reference-grid occupancy, and (where a member
// requested it) the skin-grid binning
int ens_skin_begin(EnsNL* t, int dim, const EnsMap& part, const EnsMap& cells, sph_ws* prof,
                   hipStream_t st);
// the skin pair lists of the members that requested a rebuild
int ens_skin_build(EnsNL* t, int dim, const EnsMap& lanes, sph_ws* prof, hipStream_t st);
// occupancy, largest list and overflow bit from the counting sweep's partials
int ens_skin_finish(EnsNL* t, sph_ws* prof, hipStream_t st);

// rhs.hip: device argument tables of the gather and the sweeps.
struct EnsRHS;
int ens_rhs_create(const std::vector<EnsMember>& m, EnsRHS** out);
void ens_rhs_destroy(EnsRHS* t);
int ens_gather(EnsRHS* t, int dim, const EnsMap& part, sph_ws* prof, hipStream_t st);
// the sweeps of `plan` over the members that `lanes` covers (all share the plan)
int ens_sweeps(EnsRHS* t, int dim, const SweepPlan& plan, const EnsMap& lanes, sph_ws* prof,
               hipStream_t st);

}  // namespace sph
