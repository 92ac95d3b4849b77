// Host-side workspace layout shared by the neighbour-search and physics units.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/sph.h"

namespace sph {

// One bank = everything a NeighborList snapshot needs (two banks let the
// previous list stay readable while the next one is built).
struct Bank {
  int32_t* perm = nullptr;        // sorted slot -> original particle id
  int32_t* iperm = nullptr;       // original id -> sorted slot
  int32_t* cell_start = nullptr;  // ncell + 1
  float4* pos = nullptr;          // sorted positions the pairs were built from
  // Pair buffer: each sorted particle s owns kTpp sub-lists (one per lane that
  // serves it); entry k of sub-list h lives at ((h * lane_width + k) * npad + s).
  int32_t* ell = nullptr;
  int32_t* cnt = nullptr;         // [kTpp][npad] entries per sub-list
};

// Lanes that serve one particle in the pair build and the sweeps.  At the
// particle counts of interest one lane per particle leaves a few waves per
// SIMD; several lanes per particle (combined with cross-lane shuffles) give the
// scheduler enough independent gathers to hide memory latency.
__host__ __device__ constexpr int tpp_of(int dim) { return dim == 3 ? 4 : 2; }

// Spare entries after the skin lists' sorted build positions: the skin build
// reads each batch of candidates from one address with constant offsets, up
// to 31 TPP entries past a cell range, without clamping each one.
constexpr int kSkinPosPad = 128;

struct ProfEntry {
  const char* name;
  hipEvent_t a, b;
};

struct ProfTotal {
  std::string name;
  double ms = 0.0;
  int64_t count = 0;
};

}  // namespace sph

struct sph_ws {
  sph_nl_config cfg;
  int64_t ncell = 0;
  int32_t npad = 0;        // N rounded up to 64 (pair-buffer column stride)
  int32_t tpp = 1;         // lanes per particle (tpp_of(dim))
  int32_t lane_width = 0;  // capacity of one sub-list
  int32_t xcd_remap = 1;   // XCD-contiguous block order of the pair build and sweeps
  // exact lists narrower than 3^d x cell capacity (sph_ws_tune "ell_overflow"):
  // a truncated sub-list raises SPH_NEIGHBOR_LIST_OVERFLOW
  int32_t ell_overflow = 0;
  int64_t bytes = 0;
  // binning scratch (shared by both banks)
  int32_t* cell_count = nullptr;  // ncell
  int32_t* pcell = nullptr;       // N: cell of each original particle
  int32_t* pslot = nullptr;       // N: arrival slot inside the cell
  int32_t* unsorted = nullptr;    // N: ids scattered by arrival order
  int64_t* off64 = nullptr;       // N + 1: pair offsets of the idx writer
  int64_t* scan_tmp = nullptr;    // block carries of the scans
  int64_t* cnt64 = nullptr;       // N: pair count per original id
  // per-workgroup (pairs, largest list) partials of the pair build and of the
  // counting density sweep, reduced by one workgroup (no same-address atomics)
  int64_t* part_build = nullptr;  // 2 * part_blocks
  int64_t* part_sweep = nullptr;  // 2 * part_blocks
  int32_t part_blocks = 0;
  sph::Bank bank[2];
  // Verlet-skin lists of the fused step: pairs closer than cutoff + skin on a
  // grid whose cells are >= cutoff + skin, rebuilt only after some particle
  // has moved more than `trigger` since the last build.
  struct Skin {
    bool on = false;
    bool stepped = false;     // a step has built the lists (perm etc. are defined)
    float skin = 0.0f;
    float reach_sq = 0.0f;    // (cutoff + skin)^2 rounded up: build predicate
    float trigger_sq = 0.0f;  // squared displacement that forces a rebuild
    int32_t cps[3] = {1, 1, 1};
    float cell_size[3] = {1.0f, 1.0f, 1.0f};
    int64_t ncell = 0;
    int32_t lane_width = 0;
    int32_t* cell_count = nullptr;  // ncell, all zero between builds
    int32_t* flag = nullptr;        // device: 1 = rebuild the lists this step
    float4* posb = nullptr;         // N: position of each original id at the last build
    sph::Bank b;                    // perm, iperm, cell_start, pos, ell, cnt of the lists
  } sk;
  // sorted particle records used by the physics sweeps
  float4 *P = nullptr;   // x, y, z, mass
  float4 *U0 = nullptr, *U1 = nullptr, *U2 = nullptr;  // u, rho
  float4 *V0 = nullptr, *V2 = nullptr;                // v, kappa
  float4 *Q0 = nullptr, *Q1 = nullptr, *Q2 = nullptr;  // p, eta, tag, T
  float4* NW = nullptr;  // wall normal
  float4* UT = nullptr;  // Riemann wall velocity u_tilde
  float4 *U3 = nullptr, *Q3 = nullptr;  // after density renormalisation
  // (v, p) of each stage (density, walls, renormalisation): the force sweep's
  // transport velocity and pressure in one record
  float4 *W1 = nullptr, *W2 = nullptr, *W3 = nullptr;
  // device flag, sticky: 1 once a gathered state had particles of different
  // viscosity (until then the force sweep takes eta_j = eta_i)
  int32_t* eta_varies = nullptr;
  // delta-SPH: renormalisation matrix L (row-major, 4 or 9 floats) and the
  // two density-gradient terms
  float4 *DL0 = nullptr, *DL1 = nullptr, *DL2 = nullptr, *DG0 = nullptr, *DG1 = nullptr;
  float* drho = nullptr; // sorted drhodt
  // adjoint sweeps: sorted cotangent of dudt, then the per-particle partials
  float4 *LA = nullptr, *AB0 = nullptr, *AB1 = nullptr;
  float4 *LB = nullptr, *AB2 = nullptr;  // transport-velocity step: dvdt' and v' cotangents
  // profiling
  bool prof_on = false;
  std::vector<sph::ProfEntry> prof_pending;
  std::vector<hipEvent_t> prof_pool;
  std::vector<sph::ProfTotal> prof_totals;
};

namespace sph {
void set_error(const std::string& msg);
int hip_fail(hipError_t e, const char* where);
// profiling wrappers around a launch
void prof_begin(sph_ws* ws, const char* name, hipStream_t s);
void prof_end(sph_ws* ws, hipStream_t s);
}  // namespace sph

#define SPH_CHECK_LAUNCH(where)                                  \
  do {                                                           \
    hipError_t _e = hipGetLastError();                           \
    if (_e != hipSuccess) return sph::hip_fail(_e, where);       \
  } while (0)

#define SPH_LAUNCH(ws, name, stream, ...)                        \
  do {                                                           \
    sph::prof_begin(ws, name, stream);                           \
    __VA_ARGS__;                                                 \
    sph::prof_end(ws, stream);                                   \
    SPH_CHECK_LAUNCH(name);                                      \
  } while (0)
