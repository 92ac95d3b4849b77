// Ensemble step: many independent simulations, each with its own workspace
// and Verlet-skin lists, advanced together by one launch per phase of the
// fused si_euler step (the batched kernels live beside their single-member
// forms in nl.hip and rhs.hip).
//
// The reference runs one simulation per process (simulate.py:19-136); an
// ensemble there is a host loop over independent runs.  Here every phase of
// the step covers all members at once, so small members no longer leave the
// chip idle between launches, and the whole step is a fixed launch sequence
// that a HIP graph can capture.
//
// Block maps: workgroup b of a batched launch serves block map[b].blk of
// member map[b].sim.  Members are laid out largest first (or in the given
// order, or interleaved), and with the XCD layout workgroups b, b+8, b+16, ...
// (which the hardware sends to one XCD) serve one contiguous run of the
// sequence, so neighbouring particles of one member share that XCD's L2.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "ens.h"
#include "sph_device.h"
#include "sph_internal.h"
#include "sweep.h"

using namespace sph;

struct sph_ensemble {
  std::vector<EnsMember> m;
  int dim = 2;
  EnsNL* nl = nullptr;
  EnsRHS* rhs = nullptr;
  EnsMap part, cells, lanes;  // particles, skin cells, particle lanes (all members)
  struct Group {
    SweepPlan plan;
    std::vector<int> members;
    EnsMap lanes;
  };
  std::vector<Group> groups;
  int order = 1;  // 0: given order, 1: largest member first, 2: blocks interleaved
  int xcd = 1;    // XCD-contiguous layout of the sequence
};

namespace {

int64_t cdiv64(int64_t a, int64_t b) { return (a + b - 1) / b; }

void free_map(EnsMap& m) {
  if (m.d) (void)hipFree(m.d);
  m.d = nullptr;
  m.blocks = 0;
}

// Block sequence of `who` (member indices) with blocks(k) blocks each, in the
// ensemble's order, then laid out for the XCDs; uploaded to the device.
template <typename F>
int make_map(const sph_ensemble* e, std::vector<int> who, F&& blocks, EnsMap* out) {
  if (e->order == 1)
    std::stable_sort(who.begin(), who.end(), [&](int a, int b) {
      return e->m[a].ws->cfg.n > e->m[b].ws->cfg.n;
    });
  std::vector<EnsBlock> seq;
  if (e->order == 2) {
    int64_t most = 0;
    for (int k : who) most = std::max<int64_t>(most, blocks(k));
    for (int64_t b = 0; b < most; ++b)
      for (int k : who)
        if (b < blocks(k)) seq.push_back(EnsBlock{k, (int32_t)b});
  } else {
    for (int k : who)
      for (int64_t b = 0; b < blocks(k); ++b) seq.push_back(EnsBlock{k, (int32_t)b});
  }
  if (seq.empty()) seq.push_back(EnsBlock{-1, 0});
  std::vector<EnsBlock> phys = seq;
  if (e->xcd) {
    while (seq.size() % kXcds) seq.push_back(EnsBlock{-1, 0});
    const size_t per = seq.size() / kXcds;
    phys.resize(seq.size());
    for (size_t b = 0; b < seq.size(); ++b) phys[b] = seq[(b % kXcds) * per + b / kXcds];
  }
  if (phys.size() > 0x7fffffff) {
    set_error("sph_ensemble: too many workgroups");
    return SPH_ERR_ARG;
  }
  free_map(*out);
  hipError_t err = hipMalloc((void**)&out->d, sizeof(EnsBlock) * phys.size());
  if (err == hipSuccess)
    err = hipMemcpy(out->d, phys.data(), sizeof(EnsBlock) * phys.size(), hipMemcpyHostToDevice);
  if (err != hipSuccess) return hip_fail(err, "sph_ensemble maps");
  out->blocks = (int)phys.size();
  return SPH_OK;
}

int build_maps(sph_ensemble* e) {
  std::vector<int> all(e->m.size());
  for (size_t k = 0; k < all.size(); ++k) all[k] = (int)k;
  const int tpp = tpp_of(e->dim);
  auto part = [&](int k) { return cdiv64(e->m[k].ws->cfg.n, kB); };
  auto lanes = [&](int k) { return cdiv64((int64_t)e->m[k].ws->cfg.n * tpp, kB); };
  auto cells = [&](int k) { return cdiv64(e->m[k].ws->sk.ncell * 64, 256); };
  int rc = make_map(e, all, part, &e->part);
  if (!rc) rc = make_map(e, all, cells, &e->cells);
  if (!rc) rc = make_map(e, all, lanes, &e->lanes);
  for (auto& g : e->groups)
    if (!rc) rc = make_map(e, g.members, lanes, &g.lanes);
  return rc;
}

void destroy(sph_ensemble* e) {
  if (!e) return;
  ens_nl_destroy(e->nl);
  ens_rhs_destroy(e->rhs);
  free_map(e->part);
  free_map(e->cells);
  free_map(e->lanes);
  for (auto& g : e->groups) free_map(g.lanes);
  delete e;
}

}  // namespace

extern "C" int sph_ensemble_create(const sph_ensemble_member* members, int32_t n,
                                   sph_ensemble** out) {
  if (!members || n < 1 || !out) {
    set_error("sph_ensemble_create: bad argument");
    return SPH_ERR_ARG;
  }
  *out = nullptr;
  sph_ensemble* e = new sph_ensemble;
  e->dim = members[0].prm ? members[0].prm->dim : 0;
  for (int k = 0; k < n; ++k) {
    const sph_ensemble_member& x = members[k];
    const std::string who = "sph_ensemble_create: member " + std::to_string(k);
    if (!x.ws || !x.prm || !x.in || !x.out || !x.r_out || !x.info) {
      set_error(who + ": null argument");
      destroy(e);
      return SPH_ERR_ARG;
    }
    int rc = check_state(x.prm, x.in, x.out);
    if (!rc) rc = check_bank(x.ws, kSkinBank, x.prm, "sph_step");
    if (rc) {
      destroy(e);
      return rc;
    }
    if (x.prm->dim != e->dim) {
      set_error(who + ": every member must have the same dimension");
      destroy(e);
      return SPH_ERR_ARG;
    }
    if (!x.in->dudt || !x.in->dvdt || x.in->g_ext) {
      set_error(who + ": needs dudt/dvdt and a constant body force");
      destroy(e);
      return SPH_ERR_ARG;
    }
    for (int j = 0; j < k; ++j)
      if (members[j].ws == x.ws) {
        set_error(who + ": shares its workspace with member " + std::to_string(j));
        destroy(e);
        return SPH_ERR_ARG;
      }
    EnsMember m;
    m.ws = x.ws;
    m.prm = *x.prm;
    m.dt = x.dt;
    m.ctvf = x.tvf_coef;
    std::memset(&m.bc, 0, sizeof(m.bc));
    if (x.bc) m.bc = *x.bc;
    m.in = *x.in;
    m.r_out = x.r_out;
    m.out = *x.out;
    m.prev_info = x.prev_info;
    m.info = x.info;
    m.max_occupancy = x.max_occupancy;
    e->m.push_back(m);
    const SweepPlan pl = plan_of(m.prm);
    bool found = false;
    for (auto& g : e->groups)
      if (g.plan == pl) {
        g.members.push_back(k);
        found = true;
      }
    if (!found) e->groups.push_back(sph_ensemble::Group{pl, {k}, EnsMap{}});
  }
  int rc = ens_nl_create(e->m, &e->nl);
  if (!rc) rc = ens_rhs_create(e->m, &e->rhs);
  if (!rc) rc = build_maps(e);
  if (rc) {
    destroy(e);
    return rc;
  }
  *out = e;
  return SPH_OK;
}

extern "C" int sph_ensemble_step(sph_ensemble* e, hipStream_t st) {
  if (!e) {
    set_error("sph_ensemble_step: null ensemble");
    return SPH_ERR_ARG;
  }
  for (auto& m : e->m) m.ws->sk.stepped = true;
  sph_ws* prof = e->m[0].ws;
  // 1. kick/drift + rebuild test + reference-grid occupancy; skin binning where needed
  int rc = ens_skin_begin(e->nl, e->dim, e->part, e->cells, prof, st);
  // 2. sorted records (and build positions), then the lists of members that rebuild
  if (!rc) rc = ens_gather(e->rhs, e->dim, e->part, prof, st);
  if (!rc) rc = ens_skin_build(e->nl, e->dim, e->lanes, prof, st);
  // 3. forward + case hook, one launch per sweep for each group of equal plans
  for (auto& g : e->groups)
    if (!rc) rc = ens_sweeps(e->rhs, e->dim, g.plan, g.lanes, prof, st);
  // 4. occupancy, largest list, overflow bit; rebuild requests consumed
  if (!rc) rc = ens_skin_finish(e->nl, prof, st);
  return rc;
}

extern "C" int sph_ensemble_tune(sph_ensemble* e, const char* key, int64_t value) {
  if (!e || !key) {
    set_error("sph_ensemble_tune: null argument");
    return SPH_ERR_ARG;
  }
  if (std::strcmp(key, "order") == 0) {
    if (value < 0 || value > 2) {
      set_error("sph_ensemble_tune: order is 0 (given), 1 (largest first) or 2 (interleaved)");
      return SPH_ERR_ARG;
    }
    e->order = (int)value;
  } else if (std::strcmp(key, "xcd") == 0) {
    e->xcd = value ? 1 : 0;
  } else {
    set_error(std::string("sph_ensemble_tune: unknown key ") + key);
    return SPH_ERR_ARG;
  }
  return build_maps(e);
}

extern "C" int sph_ensemble_info(const sph_ensemble* e, int64_t* members, int64_t* groups,
                                 int64_t* lane_blocks) {
  if (!e) {
    set_error("sph_ensemble_info: null ensemble");
    return SPH_ERR_ARG;
  }
  if (members) *members = (int64_t)e->m.size();
  if (groups) *groups = (int64_t)e->groups.size();
  if (lane_blocks) *lane_blocks = e->lanes.blocks;
  return SPH_OK;
}

extern "C" int sph_ensemble_destroy(sph_ensemble* e) {
  destroy(e);
  return SPH_OK;
}
