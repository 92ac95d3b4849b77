// Workspace lifetime, error reporting and the profiling hooks of libsph.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>

#include "skin_grid.h"
#include "sph_internal.h"

namespace sph {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

int hip_fail(hipError_t e, const char* where) {
  set_error(std::string(where) + ": " + hipGetErrorString(e));
  return SPH_ERR_HIP;
}

void prof_begin(sph_ws* ws, const char* name, hipStream_t s) {
  if (!ws || !ws->prof_on) return;
  ProfEntry e;
  e.name = name;
  for (int k = 0; k < 2; ++k) {
    hipEvent_t ev;
    if (!ws->prof_pool.empty()) {
      ev = ws->prof_pool.back();
      ws->prof_pool.pop_back();
    } else if (hipEventCreate(&ev) != hipSuccess) {
      return;
    }
    (k == 0 ? e.a : e.b) = ev;
  }
  hipEventRecord(e.a, s);
  ws->prof_pending.push_back(e);
}

void prof_end(sph_ws* ws, hipStream_t s) {
  if (!ws || !ws->prof_on || ws->prof_pending.empty()) return;
  hipEventRecord(ws->prof_pending.back().b, s);
}

static void prof_collect(sph_ws* ws) {
  for (auto& e : ws->prof_pending) {
    float ms = 0.0f;
    hipEventSynchronize(e.b);
    hipEventElapsedTime(&ms, e.a, e.b);
    ProfTotal* t = nullptr;
    for (auto& x : ws->prof_totals)
      if (x.name == e.name) t = &x;
    if (!t) {
      ws->prof_totals.push_back(ProfTotal{e.name, 0.0, 0});
      t = &ws->prof_totals.back();
    }
    t->ms += ms;
    t->count += 1;
    ws->prof_pool.push_back(e.a);
    ws->prof_pool.push_back(e.b);
  }
  ws->prof_pending.clear();
}

template <typename T>
static int alloc(sph_ws* ws, T** p, size_t count) {
  size_t bytes = sizeof(T) * (count ? count : 1);
  hipError_t e = hipMalloc((void**)p, bytes);
  if (e != hipSuccess) return hip_fail(e, "hipMalloc");
  ws->bytes += (int64_t)bytes;
  return SPH_OK;
}

}  // namespace sph

using namespace sph;

extern "C" const char* sph_last_error(void) { return g_last_error.c_str(); }

extern "C" int sph_version(void) { return 1; }

extern "C" int sph_ws_create(const sph_nl_config* cfg, sph_ws** out) {
  if (!cfg || !out) { set_error("sph_ws_create: null argument"); return SPH_ERR_ARG; }
  if (cfg->dim != 2 && cfg->dim != 3) { set_error("dim must be 2 or 3"); return SPH_ERR_ARG; }
  if (cfg->n <= 0) { set_error("n must be positive"); return SPH_ERR_ARG; }
  if (cfg->ell_width <= 0) { set_error("ell_width must be positive"); return SPH_ERR_ARG; }
  int64_t ncell = 1;
  if (cfg->use_cells) {
    for (int d = 0; d < cfg->dim; ++d) {
      if (cfg->cps[d] < 3) { set_error("cell list needs >= 3 cells per side"); return SPH_ERR_ARG; }
      ncell *= cfg->cps[d];
    }
  }
  const int64_t N = cfg->n;
  const int64_t npad = (N + 63) / 64 * 64;
  if ((double)cfg->ell_width * (double)npad > 4.0e10) {
    set_error("pair buffer would exceed 160 GB; reduce cell capacity");
    return SPH_ERR_SIZE;
  }
  sph_ws* ws = new sph_ws();
  ws->cfg = *cfg;
  ws->ncell = ncell;
  ws->npad = (int32_t)npad;
  // The pair build deals a particle's hits round-robin over its tpp
  // sub-lists, so ceil(width / tpp) entries per sub-list hold `width` pairs.
  ws->tpp = tpp_of(cfg->dim);
  {
    int64_t lw = (cfg->ell_width + ws->tpp - 1) / ws->tpp;
    const int64_t lw_n = (N + ws->tpp - 1) / ws->tpp + 1;
    ws->lane_width = (int32_t)(lw < lw_n ? lw : lw_n);
  }
  int rc = SPH_OK;
#define A(ptr, n) if (rc == SPH_OK) rc = alloc(ws, &(ptr), (size_t)(n))
  A(ws->cell_count, ncell);
  A(ws->pcell, N);
  A(ws->pslot, N);
  A(ws->unsorted, N);
  A(ws->off64, N + 1);
  A(ws->scan_tmp, 1024);
  A(ws->cnt64, N);
  // workgroups of the widest launch (npad * tpp lanes of 256, rounded to 8 XCDs)
  ws->part_blocks = (int32_t)(((npad * ws->tpp + 255) / 256 + 7) / 8 * 8);
  A(ws->part_build, 2 * (int64_t)ws->part_blocks);
  A(ws->part_sweep, 2 * (int64_t)ws->part_blocks);
  for (int b = 0; b < 2; ++b) {
    A(ws->bank[b].perm, N);
    A(ws->bank[b].iperm, N);
    A(ws->bank[b].cell_start, ncell + 1);
    A(ws->bank[b].pos, N);
    A(ws->bank[b].ell, (int64_t)ws->tpp * ws->lane_width * npad);
    A(ws->bank[b].cnt, (int64_t)ws->tpp * npad);
  }
  A(ws->P, N);
  A(ws->U0, N); A(ws->U1, N); A(ws->U2, N);
  A(ws->V0, N); A(ws->V2, N);
  A(ws->Q0, N); A(ws->Q1, N); A(ws->Q2, N);
  A(ws->NW, N); A(ws->UT, N);
  A(ws->drho, N);
  A(ws->LA, N); A(ws->AB0, N); A(ws->AB1, N);
  A(ws->LB, N); A(ws->AB2, N);
  A(ws->U3, N); A(ws->Q3, N);
  A(ws->W1, N); A(ws->W2, N); A(ws->W3, N);
  A(ws->eta_varies, 1);
  A(ws->DL0, N); A(ws->DL1, N); A(ws->DL2, N); A(ws->DG0, N); A(ws->DG1, N);
#undef A
  // the cell histogram is kept all-zero between launches (scans clear it)
  if (rc == SPH_OK) {
    hipError_t e = hipMemset(ws->cell_count, 0, sizeof(int32_t) * (size_t)(ncell ? ncell : 1));
    if (e == hipSuccess) e = hipMemset(ws->eta_varies, 0, sizeof(int32_t));
    if (e != hipSuccess) rc = hip_fail(e, "hipMemset");
  }
  if (rc != SPH_OK) { sph_ws_destroy(ws); return rc; }
  *out = ws;
  return SPH_OK;
}

extern "C" int sph_skin_enable(sph_ws* ws, float skin, int32_t width) {
  if (!ws) { set_error("sph_skin_enable: null workspace"); return SPH_ERR_ARG; }
  auto& S = ws->sk;
  if (S.on) { set_error("sph_skin_enable: skin lists are already enabled"); return SPH_ERR_ARG; }
  if (!(skin > 0.0f) || !std::isfinite(skin) || width <= 0) {
    set_error("sph_skin_enable: skin must be > 0 and width positive");
    return SPH_ERR_ARG;
  }
  const sph_nl_config& c = ws->cfg;
  if (!c.use_cells || c.ordered) return 0;
  double maxbox = 0.0;
  for (int d = 0; d < c.dim; ++d) maxbox = c.box[d] > maxbox ? c.box[d] : maxbox;
  // displacements are tracked to ~1e-7 box; the skin must dwarf that
  if ((double)skin < 1e-4 * maxbox) return 0;
  const double reach = (double)c.cutoff + (double)skin;
  int32_t cps[3] = {1, 1, 1};
  float cs[3] = {1.0f, 1.0f, 1.0f};
  int64_t ncell = 1;
  // The skin lists filter by distance alone, so they need every pair inside
  // the cutoff to lie in adjacent reference cells, as the reference's 3^d
  // candidate search requires; on a box of a whole number of cutoffs a pair
  // within rounding of the cutoff may not.
  for (int d = 0; d < c.dim; ++d) {
    if (c.cps[d] > 100000) return 0;
    if (!cutoff_pairs_in_adjacent_cells(c.box[d], c.cell_size[d], c.cps[d], c.cutoff_sq)) return 0;
  }
  for (int d = 0; d < c.dim; ++d) {
    // skin cells strictly wider than cutoff + skin, at least 3 per side
    const int k = (int)std::floor((double)c.box[d] / (reach * (1.0 + 1e-5)));
    if (k < 3) return 0;
    cps[d] = k;
    cs[d] = (float)((double)c.box[d] / k);
    ncell *= k;
  }
  const int64_t N = c.n, npad = ws->npad;
  int64_t lw = ((int64_t)width + ws->tpp - 1) / ws->tpp;  // hits dealt round-robin
  const int64_t lw_n = (N + ws->tpp - 1) / ws->tpp + 1;
  lw = lw < lw_n ? lw : lw_n;
  if ((double)lw * ws->tpp * (double)npad > 4.0e10) {
    set_error("sph_skin_enable: skin pair buffer would exceed 160 GB");
    return SPH_ERR_SIZE;
  }
  int rc = SPH_OK;
#define A(ptr, n) if (rc == SPH_OK) rc = alloc(ws, &(ptr), (size_t)(n))
  A(S.cell_count, ncell);
  A(S.flag, 1);
  A(S.posb, N);
  A(S.b.perm, N);
  A(S.b.iperm, N);
  A(S.b.cell_start, ncell + 1);
  A(S.b.ell, ws->tpp * lw * npad);
  A(S.b.cnt, ws->tpp * npad);
  A(S.b.pos, N + kSkinPosPad);  // last, so that its padding moves no other buffer
#undef A
  if (rc == SPH_OK) {
    const int32_t one = 1;  // the first step builds the lists
    hipError_t e = hipMemset(S.cell_count, 0, sizeof(int32_t) * (size_t)ncell);
    if (e == hipSuccess) e = hipMemcpy(S.flag, &one, sizeof(one), hipMemcpyHostToDevice);
    if (e != hipSuccess) rc = hip_fail(e, "sph_skin_enable");
  }
  if (rc != SPH_OK) {
    void* ps[] = {S.cell_count, S.flag, S.posb, S.b.perm, S.b.iperm, S.b.cell_start, S.b.pos,
                  S.b.ell, S.b.cnt};
    hipDeviceSynchronize();
    for (void* p : ps)
      if (p) hipFree(p);
    S = sph_ws::Skin();
    return rc;
  }
  for (int d = 0; d < 3; ++d) {
    S.cps[d] = cps[d];
    S.cell_size[d] = cs[d];
  }
  S.skin = skin;
  S.reach_sq = (float)(reach * reach * (1.0 + 1e-6));
  S.trigger_sq = (float)((0.49 * skin) * (0.49 * skin));
  S.ncell = ncell;
  S.lane_width = (int32_t)lw;
  S.on = true;
  return 1;
}

extern "C" int sph_ws_destroy(sph_ws* ws) {
  if (!ws) return SPH_OK;
  hipDeviceSynchronize();
  void* ptrs[] = {ws->cell_count, ws->pcell, ws->pslot, ws->unsorted, ws->off64, ws->scan_tmp,
                  ws->cnt64, ws->part_build, ws->part_sweep, ws->P, ws->U0, ws->U1, ws->U2, ws->V0, ws->V2, ws->Q0, ws->Q1,
                  ws->Q2, ws->NW, ws->UT, ws->drho, ws->LA, ws->AB0, ws->AB1, ws->LB, ws->AB2,
                  ws->U3, ws->Q3, ws->DL0, ws->DL1, ws->DL2, ws->DG0, ws->DG1,
                  ws->W1, ws->W2, ws->W3, ws->eta_varies};
  for (void* p : ptrs)
    if (p) hipFree(p);
  for (auto& b : ws->bank) {
    void* bp[] = {b.perm, b.iperm, b.cell_start, b.pos, b.ell, b.cnt};
    for (void* p : bp)
      if (p) hipFree(p);
  }
  {
    auto& S = ws->sk;
    void* ps[] = {S.cell_count, S.flag, S.posb, S.b.perm, S.b.iperm, S.b.cell_start, S.b.pos,
                  S.b.ell, S.b.cnt};
    for (void* p : ps)
      if (p) hipFree(p);
  }
  for (auto& e : ws->prof_pending) { ws->prof_pool.push_back(e.a); ws->prof_pool.push_back(e.b); }
  for (auto ev : ws->prof_pool) hipEventDestroy(ev);
  delete ws;
  return SPH_OK;
}

extern "C" int64_t sph_ws_bytes(const sph_ws* ws) { return ws ? ws->bytes : 0; }

extern "C" int sph_ws_tune(sph_ws* ws, const char* key, int64_t value) {
  if (!ws || !key) { set_error("sph_ws_tune: null argument"); return SPH_ERR_ARG; }
  if (std::strcmp(key, "xcd_remap") == 0) {
    ws->xcd_remap = value ? 1 : 0;
    return SPH_OK;
  }
  if (std::strcmp(key, "ell_overflow") == 0) {
    // exact pair buffers sized below 3^d x cell capacity per particle: a
    // sub-list that would be truncated raises SPH_NEIGHBOR_LIST_OVERFLOW
    ws->ell_overflow = value ? 1 : 0;
    return SPH_OK;
  }
  if (std::strcmp(key, "skin_rebuild") == 0) {
    // the next fused step rebuilds the skin lists, as the first step after
    // sph_skin_enable does (synchronous; not for use under graph capture)
    if (!ws->sk.on) { set_error("sph_ws_tune: skin lists are not enabled"); return SPH_ERR_ARG; }
    const int32_t one = value ? 1 : 0;
    hipError_t e = hipMemcpy(ws->sk.flag, &one, sizeof(one), hipMemcpyHostToDevice);
    return e == hipSuccess ? SPH_OK : hip_fail(e, "sph_ws_tune");
  }
  set_error(std::string("sph_ws_tune: unknown key ") + key);
  return SPH_ERR_ARG;
}

extern "C" int sph_prof_enable(sph_ws* ws, int on) {
  if (!ws) return SPH_ERR_ARG;
  ws->prof_on = on != 0;
  return SPH_OK;
}

extern "C" int sph_prof_reset(sph_ws* ws) {
  if (!ws) return SPH_ERR_ARG;
  prof_collect(ws);
  ws->prof_totals.clear();
  return SPH_OK;
}

extern "C" int sph_prof_read(sph_ws* ws, char* names, int names_cap, double* ms, int64_t* count,
                             int cap) {
  if (!ws) return SPH_ERR_ARG;
  prof_collect(ws);
  int n = (int)ws->prof_totals.size();
  int pos = 0;
  for (int i = 0; i < n && i < cap; ++i) {
    const auto& t = ws->prof_totals[i];
    if (ms) ms[i] = t.ms;
    if (count) count[i] = t.count;
    if (names) {
      int len = (int)t.name.size();
      if (pos + len + 1 <= names_cap) {
        std::memcpy(names + pos, t.name.c_str(), len + 1);
        pos += len + 1;
      }
    }
  }
  return n;
}
