// Shared pieces of the pair sweeps (forward RHS in rhs.hip, adjoint in adj.hip):
// kernel dispatch, small vector algebra, sweep arguments, the pair visitor and
// the host-side argument set-up.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <string>

#include "sph_device.h"
#include "sph_internal.h"

namespace sph {

void cut_band(const sph_nl_config& c, float* lo, float* hi);

// Bank 2 names the skin lists of the fused step.
constexpr int kSkinBank = 2;

constexpr int kB = 256;  // threads per workgroup of the sweeps
static inline int64_t cdivl(int64_t a, int64_t b) { return (a + b - 1) / b; }

// ---------------------------------------------------------- kernel dispatch
// KIND >= 0: compile-time kernel; KIND == -1: runtime switch over the rest.
template <int KIND, int DIM>
struct KernelEval {
  int kind;
  float s, ih;
  __device__ __forceinline__ float w(float q) const { return Kern<KIND, DIM>::w(q, s); }
  __device__ __forceinline__ float dw(float q) const { return Kern<KIND, DIM>::dw(q, s, ih); }
  __device__ __forceinline__ float d2w(float q) const { return Kern<KIND, DIM>::d2w(q, s, ih); }
};

template <int DIM>
struct KernelEval<-1, DIM> {
  int kind;
  float s, ih;
  __device__ __forceinline__ float w(float q) const {
    switch (kind) {
      case SPH_KERNEL_CSK: return Kern<SPH_KERNEL_CSK, DIM>::w(q, s);
      case SPH_KERNEL_QSK: return Kern<SPH_KERNEL_QSK, DIM>::w(q, s);
      case SPH_KERNEL_WC2K: return Kern<SPH_KERNEL_WC2K, DIM>::w(q, s);
      case SPH_KERNEL_WC4K: return Kern<SPH_KERNEL_WC4K, DIM>::w(q, s);
      case SPH_KERNEL_WC6K: return Kern<SPH_KERNEL_WC6K, DIM>::w(q, s);
      case SPH_KERNEL_GK: return Kern<SPH_KERNEL_GK, DIM>::w(q, s);
      default: return Kern<SPH_KERNEL_SGK, DIM>::w(q, s);
    }
  }
  __device__ __forceinline__ float dw(float q) const {
    switch (kind) {
      case SPH_KERNEL_CSK: return Kern<SPH_KERNEL_CSK, DIM>::dw(q, s, ih);
      case SPH_KERNEL_QSK: return Kern<SPH_KERNEL_QSK, DIM>::dw(q, s, ih);
      case SPH_KERNEL_WC2K: return Kern<SPH_KERNEL_WC2K, DIM>::dw(q, s, ih);
      case SPH_KERNEL_WC4K: return Kern<SPH_KERNEL_WC4K, DIM>::dw(q, s, ih);
      case SPH_KERNEL_WC6K: return Kern<SPH_KERNEL_WC6K, DIM>::dw(q, s, ih);
      case SPH_KERNEL_GK: return Kern<SPH_KERNEL_GK, DIM>::dw(q, s, ih);
      default: return Kern<SPH_KERNEL_SGK, DIM>::dw(q, s, ih);
    }
  }
};

// --------------------------------------------------------------- vectors
struct V3 { float x, y, z; };
__device__ __forceinline__ V3 v3(float4 a) { return V3{a.x, a.y, a.z}; }
__device__ __forceinline__ V3 operator-(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3 operator+(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3 operator*(float s, V3 a) { return V3{s * a.x, s * a.y, s * a.z}; }
template <int DIM>
__device__ __forceinline__ float dot(V3 a, V3 b) {
  float d = a.x * b.x + a.y * b.y;
  if (DIM == 3) d += a.z * b.z;
  return d;
}

struct Geo {
  float box[3], inv[3], half[3];
};

template <int DIM>
__device__ __forceinline__ V3 min_image(float4 a, float4 b, const Geo& g) {
  V3 d;
  d.x = disp_fast(a.x, b.x, g.box[0], g.inv[0]);
  d.y = disp_fast(a.y, b.y, g.box[1], g.inv[1]);
  d.z = DIM == 3 ? disp_fast(a.z, b.z, g.box[2], g.inv[2]) : 0.0f;
  return d;
}

template <int DIM>
__device__ __forceinline__ V3 load_vec(const float* __restrict__ a, int64_t i) {
  V3 v;
  v.x = a[i * DIM];
  v.y = a[i * DIM + 1];
  v.z = DIM == 3 ? a[i * DIM + 2] : 0.0f;
  return v;
}

template <int DIM>
__device__ __forceinline__ void store_vec(float* __restrict__ a, int64_t i, V3 v) {
  a[i * DIM] = v.x;
  a[i * DIM + 1] = v.y;
  if (DIM == 3) a[i * DIM + 2] = v.z;
}

__device__ __forceinline__ float4 f4(V3 v, float w) { return make_float4(v.x, v.y, v.z, w); }

// ------------------------------------------------------------- arguments
struct Recs {
  float4 *P, *U0, *V0, *Q0, *NW, *U1, *Q1, *U2, *V2, *Q2, *UT;
  float* drho;
  float4 *U3, *Q3;                    // renormalised density stage
  float4 *DL0, *DL1, *DL2, *DG0, *DG1;  // delta-SPH matrices and gradient terms
  float4 *W1, *W2, *W3;                // (v - u, p) after density, walls, renormalisation
};

struct SweepArgs {
  sph_solver_params prm;
  sph_state_in in;
  sph_state_out out;
  Recs R;
  const int32_t* perm;
  const int32_t* ell;
  const int32_t* cnt;
  Geo geo;
  int n, npad, lane_width;
  // which record set the force sweep reads
  const float4 *Uf, *Vf, *Qf;
  // which records hold the density stage's (u, rho) and (p, eta, tag, T):
  // U1/Q1, or U3/Q3 after renormalisation
  const float4 *Ud, *Qd;
  const float4* Wf;     // (v - u, p) records the force sweep reads
  int32_t* eta_varies;  // sticky device flag: the gathered viscosities differ
  int heat_in_density;  // apply T += dt*dTdt in the density epilogue
  sph_bc_spec bc;
  // Pair filter of skin lists: squared distances below cut_lo are inside the
  // cutoff, at or above cut_hi outside, in between the exact reference metric
  // decides.  Exact lists pass every pair (cut_lo = cut_hi = +inf).
  float cut_lo, cut_hi, cut_sq;
  int64_t* info;  // non-null: the density sweep counts pairs (per-workgroup partials in part)
  int64_t* part;  // [2 * gridDim.x] pairs and largest list of each workgroup
  int remap;      // XCD-contiguous block order (grid a multiple of 8)
};

// Pair (sweep centre s, listed j) of the neighbour set: the reference predicate
// metric_sq(r_s, r_j) < cutoff^2 with s in the sender role, as in the build.
template <int DIM>
__device__ __forceinline__ bool in_cutoff(const SweepArgs& a, float d2, float4 Pi, float4 Pj) {
  if (d2 < a.cut_lo) return true;
  if (d2 >= a.cut_hi) return false;
  return metric_sq_exact<DIM>(Pi, Pj, a.geo.box, a.geo.half) < a.cut_sq;
}

// Visits the pairs of sorted particle s.  The next pair index (a
// lane-coalesced load) is requested before the current pair's body runs, so
// its latency hides behind the body; indices live in scalars, not an array a
// rolled loop would have to index dynamically.
template <typename F>
__device__ __forceinline__ void for_each_pair(const SweepArgs& a, int s, int h, F&& f) {
  const int c = a.cnt[h * a.npad + s];
  const int32_t* lst = a.ell + ((int64_t)h * a.lane_width) * a.npad + s;
  int jn = c > 0 ? lst[0] : 0;
  for (int k = 0; k < c; ++k) {
    const int j = jn;
    __builtin_assume(j >= 0);  // sorted indices: record addresses need no sign extension
    if (k + 1 < c) jn = lst[(int64_t)(k + 1) * a.npad];
    f(j);
  }
}

// for_each_pair with the records of each pair requested one pair ahead:
// load(j) gathers the records of pair k + 1 (its index was requested a pair
// earlier still) before the body of pair k runs, so the gathers have a whole
// body to arrive.  The records are gathered for every listed pair, also those
// the body then skips; at the end the last pair's records are gathered again
// instead of reading past the list.
template <typename L, typename F>
__device__ __forceinline__ void for_each_pair_ahead(const SweepArgs& a, int s, int h, L&& load,
                                                    F&& f) {
  const int c = a.cnt[h * a.npad + s];
  if (c <= 0) return;
  const int32_t* lst = a.ell + ((int64_t)h * a.lane_width) * a.npad + s;
  int j = lst[0];
  int jn = c > 1 ? lst[a.npad] : j;
  __builtin_assume(j >= 0);
  auto rec = load(j);
  for (int k = 0; k < c; ++k) {
    const int jc = j;
    const auto rc = rec;
    j = jn;
    __builtin_assume(j >= 0);
    if (k + 2 < c) jn = lst[(int64_t)(k + 2) * a.npad];
    rec = load(j);
    f(jc, rc);
  }
}

// Sum of x over the TPP consecutive lanes serving one particle.
template <int TPP>
__device__ __forceinline__ float lane_sum(float x) {
#pragma unroll
  for (int o = 1; o < TPP; o <<= 1) x += __shfl_xor(x, o, 64);
  return x;
}

template <int TPP>
__device__ __forceinline__ int lane_sum_i(int x) {
#pragma unroll
  for (int o = 1; o < TPP; o <<= 1) x += __shfl_xor(x, o, 64);
  return x;
}

template <int TPP>
__device__ __forceinline__ V3 lane_sum(V3 v) {
  return V3{lane_sum<TPP>(v.x), lane_sum<TPP>(v.y), lane_sum<TPP>(v.z)};
}

// ------------------------------------------------------------- gather
// Original-order state -> cell-sorted records (P, U0, V0, Q0, NW).  nl_pos
// (nullable) receives the sorted positions of the pair build; with a gate it
// is written only when *gate != 0, together with posb (original order).
template <int DIM>
__device__ __forceinline__ void gather_state_run(const SweepArgs& a, const float* __restrict__ r,
                                                 const float* __restrict__ u,
                                                 const float* __restrict__ v,
                                                 float4* __restrict__ nl_pos,
                                                 const int32_t* __restrict__ gate,
                                                 float4* __restrict__ posb, const int blk) {
  int s = blk * blockDim.x + threadIdx.x;
  if (s >= a.n) return;
  const int i = a.perm[s];
  const sph_state_in& in = a.in;
  V3 rr = load_vec<DIM>(r, i);
  a.R.P[s] = f4(rr, in.mass[i]);
  if (nl_pos && (!gate || *gate)) {
    nl_pos[s] = f4(rr, 0.0f);
    if (posb) posb[i] = f4(rr, 0.0f);
  }
  a.R.U0[s] = f4(load_vec<DIM>(u, i), in.rho[i]);
  a.R.V0[s] = f4(load_vec<DIM>(v, i), in.kappa ? in.kappa[i] : 0.0f);
  const float eta = in.eta[i];
  a.R.Q0[s] = make_float4(in.p[i], eta, (float)in.tag[i], in.T ? in.T[i] : 0.0f);
  if (eta != in.eta[0]) *a.eta_varies = 1;  // same value from every writer
  if (in.nw || in.Cp) {
    V3 n = in.nw ? load_vec<DIM>(in.nw, i) : V3{0.f, 0.f, 0.f};
    a.R.NW[s] = f4(n, in.Cp ? in.Cp[i] : 0.0f);
  }
}

template <int DIM>
__global__ void k_gather_state(SweepArgs a, const float* __restrict__ r, const float* __restrict__ u,
                               const float* __restrict__ v, float4* __restrict__ nl_pos,
                               const int32_t* __restrict__ gate, float4* __restrict__ posb) {
  gather_state_run<DIM>(a, r, u, v, nl_pos, gate, posb, blockIdx.x);
}

// The single-list kernel NAME and its ensemble form NAME_ens around a sweep
// body RUN<TARGS>(args, blk).  TPARAMS and TARGS are parenthesised template
// parameter and argument lists.
#define SPH_UNPAREN(...) __VA_ARGS__
// MINW: waves per SIMD the register allocation must leave room for.  The
// sweeps wait on gathers, and 8 waves hide more of that than the registers a
// 7-wave allocation frees save; variants that would spill at 64 registers
// pass 1 (no bound).
#define SPH_SWEEP_KERNELS(NAME, RUN, TPARAMS, TARGS, MINW)                                \
  template <SPH_UNPAREN TPARAMS>                                                          \
  __global__ __launch_bounds__(kB, MINW) void NAME(SweepArgs a) {                         \
    RUN<SPH_UNPAREN TARGS>(a, logical_block(a.remap));                                    \
  }                                                                                       \
  template <SPH_UNPAREN TPARAMS>                                                          \
  __global__ __launch_bounds__(kB, MINW) void NAME##_ens(const SweepArgs* __restrict__ t, \
                                                         const EnsBlock* __restrict__ map) { \
    const EnsBlock m = map[blockIdx.x];                                                   \
    if (m.sim < 0) return;                                                                \
    RUN<SPH_UNPAREN TARGS>(t[m.sim], m.blk);                                              \
  }

// Arguments of one ensemble member's gather: the sweep arguments plus the
// step's new state and, for a list rebuild, the build positions.
struct GatherArgs {
  SweepArgs a;
  const float *r, *u, *v;
  float4* nl_pos;
  const int32_t* gate;
  float4* posb;
};

template <int DIM>
__global__ void k_gather_state_ens(const GatherArgs* __restrict__ t,
                                   const EnsBlock* __restrict__ map) {
  const EnsBlock m = map[blockIdx.x];
  if (m.sim < 0) return;
  const GatherArgs& g = t[m.sim];
  gather_state_run<DIM>(g.a, g.r, g.u, g.v, g.nl_pos, g.gate, g.posb, m.blk);
}

template <int DIM>
__device__ __forceinline__ V3 g_ext_of(const SweepArgs& a, int i) {
  if (a.in.g_ext) return load_vec<DIM>(a.in.g_ext, i);
  return V3{a.prm.g_const[0], a.prm.g_const[1], a.prm.g_const[2]};
}

// ------------------------------------------------------------- density
// MODE 0: summation (solver.py:176-178, 794-796)
// MODE 1: SPH continuity (solver.py:24-30, 750-756)
// MODE 2: Riemann continuity (solver.py:108-164, 772-793)
// MODE 3: delta-SPH continuity (solver.py:33-105, 757-769) from the gradient
//         terms of k_delta_grad
// plus the Riemann wall velocity u_tilde (solver.py:547-552) for wall targets,
// then the EOS (solver.py:801).
// Returns the pairs of s inside the cutoff (summed over its TPP lanes).
template <int DIM, int KIND, int MODE>
__device__ __forceinline__ int density_of(const SweepArgs& a, const int s, const int h) {
  constexpr int TPP = tpp_of(DIM);
  const sph_solver_params& p = a.prm;
  const KernelEval<KIND, DIM> K{p.kernel, p.sigma, p.one_over_h};
  const Eos eos = make_eos(p);
  const float4 Pi = a.R.P[s], Ui = a.R.U0[s], Qi = a.R.Q0[s];
  const int tag_i = (int)Qi.z;
  const bool fluid_i = tag_i == kFluid;
  const bool want_ut = p.solver == 1 && p.bc_trick && !p.free_slip && is_wall(tag_i);
  const V3 ui = v3(Ui);
  float wsum = 0.0f, acc = 0.0f, utw = 0.0f, diff = 0.0f;
  V3 ut = {0.f, 0.f, 0.f};
  V3 gi = {0.f, 0.f, 0.f};
  if (MODE == 2) gi = g_ext_of<DIM>(a, a.perm[s]);
  const V3 rgi = MODE == 3 ? v3(a.R.DG0[s]) : V3{0.f, 0.f, 0.f};
  int npair = 0;
  for_each_pair(a, s, h, [&](const int j) {
    const float4 Pj = a.R.P[j];
    const V3 dr = min_image<DIM>(Pi, Pj, a.geo);
    const float d2 = dist2<DIM>(dr.x, dr.y, dr.z);
    if (!in_cutoff<DIM>(a, d2, Pi, Pj)) return;
    ++npair;
    const float d = sqrt_fast(d2);
    const float q = d * p.one_over_h;
    if (MODE == 0) {
      wsum += K.w(q);
    } else {
      const float4 Uj = a.R.U0[j];
      const float gw = K.dw(q);
      const V3 e = rcp_fast(d + kEps) * dr;
      const V3 kg = gw * e;
      if (MODE == 1) {
        const float vj = Pj.w * rcp_fast(Uj.w);
        acc += vj * dot<DIM>(ui - v3(Uj), kg);
      } else if (MODE == 3) {
        const float vj = Pj.w * rcp_fast(Uj.w);
        acc += vj * dot<DIM>(ui - v3(Uj), kg);
        if ((int)a.R.Q0[j].z == kFluid) {
          // psi_ij = 2 (rho_j - rho_i) (-r_ij) / (d + eps)^2 - grad_i - grad_j
          const float de = d + kEps;
          const V3 psi = ((2.0f * (Uj.w - Ui.w)) * rcp_fast(de * de)) * ((-1.0f) * dr) - rgi -
                         v3(a.R.DG1[j]);
          diff += dot<DIM>(psi, kg) * vj;
        }
      } else {
        const float4 Qj = a.R.Q0[j];
        const bool wall_j = is_wall((int)Qj.z);
        const V3 uj = v3(Uj);
        float uL, uR, pR, rhoR;
        const float pL = Qi.x, rhoL = Ui.w;
        const V3 me = (-1.0f) * e;
        if (wall_j) {
          const float4 Nj = a.R.NW[j];
          const V3 nj = v3(Nj);
          uL = -dot<DIM>(ui, nj);
          uR = -uL + 2.0f * dot<DIM>(uj, nj);
          pR = pL + rhoL * dot<DIM>(gi, (-1.0f) * dr);
          rhoR = eos.rho_fn(pR);
        } else {
          uL = dot<DIM>(ui, me);
          uR = dot<DIM>(uj, me);
          pR = Qj.x;
          rhoR = Uj.w;
        }
        const float Uavg = (uL + uR) * 0.5f;
        const V3 vavg = 0.5f * (ui + uj);
        const float rhoavg = (rhoL + rhoR) * 0.5f;
        const float Ustar = Uavg + 0.5f * (pL - pR) * rcp_fast(rhoavg * p.c_ref);
        const V3 vstar = Ustar * me + (vavg - Uavg * me);
        acc += 2.0f * Ui.w * Pj.w * rcp_fast(Uj.w) * dot<DIM>(ui - vstar, kg);
      }
    }
    if (want_ut) {
      const float4 Qj = a.R.Q0[j];
      if ((int)Qj.z == kFluid) {
        const float w = K.w(q);
        utw += w;
        ut = ut + w * v3(a.R.U0[j]);
      }
    }
  });
  if (MODE == 0) {
    wsum = lane_sum<TPP>(wsum);
  } else {
    acc = lane_sum<TPP>(acc);
  }
  if (MODE == 3) diff = lane_sum<TPP>(diff);
  if (p.solver == 1 && p.bc_trick) {
    utw = lane_sum<TPP>(utw);
    ut = lane_sum<TPP>(ut);
  }
  if (a.info) npair = lane_sum_i<TPP>(npair);
  if (h != 0) return npair;
  float rho = Ui.w, drhodt;
  const int i = a.perm[s];
  if (MODE == 0) {
    if (fluid_i) rho = Pi.w * wsum;
    drhodt = a.in.drhodt ? a.in.drhodt[i] : 0.0f;
  } else if (MODE == 1) {
    drhodt = Ui.w * acc;
    rho = Ui.w + p.dt * drhodt;
  } else if (MODE == 3) {
    drhodt = Ui.w * acc + p.c_ref * p.diff_delta * p.support * diff;
    rho = Ui.w + p.dt * drhodt;
  } else {
    drhodt = acc * (fluid_i ? 1.0f : 0.0f);
    rho = Ui.w + p.dt * drhodt;
  }
  const float pr = eos.p_fn(rho);
  float T = Qi.w;
  if (a.heat_in_density) T = T + p.dt * a.in.dTdt[i];
  a.R.U1[s] = make_float4(Ui.x, Ui.y, Ui.z, rho);
  a.R.Q1[s] = make_float4(pr, Qi.y, Qi.z, T);
  const float4 Vi = a.R.V0[s];
  a.R.W1[s] = make_float4(Vi.x - Ui.x, Vi.y - Ui.y, Vi.z - Ui.z, pr);
  a.R.drho[s] = drhodt;
  if (p.solver == 1 && p.bc_trick) {
    // u_tilde of every particle: Shepard average for no-slip walls, u otherwise
    // (only wall neighbours read it, solver.py:378)
    V3 o = want_ut ? (1.0f / (utw + kEps)) * ut : ui;
    a.R.UT[s] = f4(o, 0.0f);
  }
  return npair;
}

// blk: the workgroup's block within this list's grid; it also indexes the
// workgroup's partials.
template <int DIM, int KIND, int MODE>
__device__ __forceinline__ void density_run(const SweepArgs& a, const int blk) {
  constexpr int TPP = tpp_of(DIM);
  const int tid = blk * blockDim.x + threadIdx.x;
  const int s = tid / TPP, h = tid % TPP;
  int npair = 0;
  if (s < a.n) npair = density_of<DIM, KIND, MODE>(a, s, h);  // lane groups stay together
  if (!a.info) return;
  // occupancy (pairs inside the cutoff) and the largest count of one particle;
  // every thread of the block reaches this point
  const int own = h == 0 ? npair : 0;
  unsigned long long tot = (unsigned long long)own;
  int mx = own;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    tot += __shfl_down(tot, o, 64);
    const int y = __shfl_down(mx, o, 64);
    mx = y > mx ? y : mx;
  }
  __shared__ unsigned long long btot[kB / 64];
  __shared__ int bmax[kB / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) { btot[wid] = tot; bmax[wid] = mx; }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    int m = 0;
    for (int w = 0; w < kB / 64; ++w) { t += btot[w]; m = bmax[w] > m ? bmax[w] : m; }
    a.part[2 * blk] = (int64_t)t;  // reduced by one workgroup afterwards
    a.part[2 * blk + 1] = m;
  }
}

SPH_SWEEP_KERNELS(k_density, density_run, (int DIM, int KIND, int MODE), (DIM, KIND, MODE), 8)

// ------------------------------------------------------------- walls
// MODE 0: generalized wall BC of the SPH solver (solver.py:431-530)
// MODE 1: heat BC of the Riemann solver (solver.py:554-565)
// followed by T += dt * dTdt when heat conduction is on (solver.py:834).
template <int DIM, int KIND, int MODE>
__device__ __forceinline__ void walls_run(const SweepArgs& a, const int blk) {
  constexpr int TPP = tpp_of(DIM);
  const int tid = blk * blockDim.x + threadIdx.x;
  const int s = tid / TPP, h = tid % TPP;
  if (s >= a.n) return;  // whole lane groups leave together
  const sph_solver_params& p = a.prm;
  const KernelEval<KIND, DIM> K{p.kernel, p.sigma, p.one_over_h};
  const Eos eos = make_eos(p);
  const float4 Pi = a.R.P[s], Ui = a.Ud[s], Vi = a.R.V0[s], Qi = a.Qd[s];
  const int tag_i = (int)Qi.z;
  const int i = a.perm[s];
  V3 u = v3(Ui), v = v3(Vi);
  float pr = Qi.x, T = Qi.w;
  if (is_wall(tag_i)) {
    float ws = 0.f, sp = 0.f, sT = 0.f;
    V3 su = {0.f, 0.f, 0.f}, sv = {0.f, 0.f, 0.f}, srr = {0.f, 0.f, 0.f};
    for_each_pair(a, s, h, [&](const int j) {
      const float4 Qj = a.Qd[j];
      if ((int)Qj.z != kFluid) return;
      const float4 Pj = a.R.P[j];
      const V3 dr = min_image<DIM>(Pi, Pj, a.geo);
      const float d2 = dist2<DIM>(dr.x, dr.y, dr.z);
      if (!in_cutoff<DIM>(a, d2, Pi, Pj)) return;
      const float d = sqrt_fast(d2);
      const float w = K.w(d * p.one_over_h);
      ws += w;
      if (MODE == 0) {
        const float4 Uj = a.Ud[j], Vj = a.R.V0[j];
        su = su + w * v3(Uj);
        sv = sv + w * v3(Vj);
        sp += w * Qj.x;
        srr = srr + (Uj.w * w) * dr;
      }
      if (p.heat_conduction) sT += w * Qj.w;
    });
    ws = lane_sum<TPP>(ws);
    if (MODE == 0) {
      su = lane_sum<TPP>(su);
      sv = lane_sum<TPP>(sv);
      sp = lane_sum<TPP>(sp);
      srr = lane_sum<TPP>(srr);
    }
    if (p.heat_conduction) sT = lane_sum<TPP>(sT);
    const float inv = 1.0f / (ws + kEps);
    if (MODE == 0) {
      V3 uw = inv * su, vw = inv * sv;
      if (p.free_slip) {
        const V3 n = (-1.0f) * v3(a.R.NW[s]);
        uw = dot<DIM>(uw, n) * n;
        vw = dot<DIM>(vw, n) * n;
      }
      u = 2.0f * u - uw;
      v = 2.0f * v - vw;
      const V3 g = g_ext_of<DIM>(a, i);
      pr = (sp + dot<DIM>(g, srr)) * inv;
    }
    if (p.heat_conduction && (tag_i == kSolidWall || tag_i == kMovingWall)) T = sT * inv;
  }
  if (h != 0) return;
  const float rho = MODE == 0 ? eos.rho_fn(pr) : Ui.w;
  if (p.heat_conduction) T = T + p.dt * a.in.dTdt[i];
  a.R.U2[s] = f4(u, rho);
  a.R.V2[s] = f4(v, Vi.w);
  a.R.Q2[s] = make_float4(pr, Qi.y, Qi.z, T);
  a.R.W2[s] = f4(v - u, pr);
}
SPH_SWEEP_KERNELS(k_walls, walls_run, (int DIM, int KIND, int MODE), (DIM, KIND, MODE), 8)


// ------------------------------------------------------------- sweep plans
// The sweeps of one forward evaluation as fixed by the solver parameters
// (solver.py:750-928): density variant, renormalisation, wall treatment,
// force variant, and the kernel template the sweeps are compiled for.
struct SweepPlan {
  int density;  // k_density MODE; MODE 3 runs k_delta_L and k_delta_grad first
  bool renorm;  // k_renorm after a continuity update
  int walls;    // -1: none, else k_walls MODE
  int force;    // k_force SOLVER
  bool tvf;     // k_force TVF
  bool extras;  // k_force EXTRAS: artificial viscosity or heat conduction on
  int kind;     // SPH_KERNEL_QSK, SPH_KERNEL_WC2K, or -1 (runtime switch)
  bool operator==(const SweepPlan& o) const {
    return density == o.density && renorm == o.renorm && walls == o.walls &&
           force == o.force && tvf == o.tvf && extras == o.extras && kind == o.kind;
  }
};

inline SweepPlan plan_of(const sph_solver_params& p) {
  SweepPlan pl;
  pl.density = !p.density_evolution ? 0 : p.solver == 0 ? 1 : p.solver == 2 ? 3 : 2;
  pl.renorm = p.density_evolution && p.rho_renorm;  // solver.py:755, 770, 792
  const bool gwbc = p.bc_trick && (p.solver == 0 || p.solver == 2);        // solver.py:806
  const bool heat_bc = p.bc_trick && p.solver == 1 && p.heat_conduction;  // solver.py:822
  pl.walls = gwbc ? 0 : heat_bc ? 1 : -1;
  pl.force = p.solver;
  pl.tvf = p.solver != 1 && p.tvf_stress;
  pl.extras = p.artificial_alpha != 0.0f || p.heat_conduction != 0;
  pl.kind = (p.kernel == SPH_KERNEL_QSK || p.kernel == SPH_KERNEL_WC2K) ? p.kernel : -1;
  return pl;
}

// The records the walls sweep (Ud, Qd) and the force sweep (Uf, Vf, Qf) read;
// the other sweeps read fixed records.
inline void select_records(SweepArgs& a, const SweepPlan& pl) {
  a.Ud = pl.renorm ? a.R.U3 : a.R.U1;
  a.Qd = pl.renorm ? a.R.Q3 : a.R.Q1;
  if (pl.walls >= 0) {
    a.Uf = a.R.U2; a.Vf = a.R.V2; a.Qf = a.R.Q2; a.Wf = a.R.W2;
  } else {
    a.Uf = a.Ud; a.Vf = a.R.V0; a.Qf = a.Qd; a.Wf = pl.renorm ? a.R.W3 : a.R.W1;
  }
}

using SweepKernel = void (*)(SweepArgs);
using EnsSweepKernel = void (*)(const SweepArgs*, const EnsBlock*);

// ------------------------------------------------------------- host set-up
inline Geo make_geo(const float* box) {
  Geo g;
  for (int d = 0; d < 3; ++d) {
    g.box[d] = box[d];
    g.inv[d] = box[d] != 0.0f ? 1.0f / box[d] : 0.0f;
    g.half[d] = box[d] * 0.5f;
  }
  return g;
}

inline int check_state(const sph_solver_params* prm, const sph_state_in* in, const sph_state_out* out) {
  if (!prm || !in || !out) { set_error("null argument"); return SPH_ERR_ARG; }
  if (!in->r || !in->u || !in->v || !in->rho || !in->p || !in->mass || !in->eta || !in->tag) {
    set_error("state is missing r/u/v/rho/p/mass/eta/tag");
    return SPH_ERR_ARG;
  }
  if (!out->u || !out->v || !out->dudt || !out->dvdt || !out->rho || !out->p || !out->drhodt) {
    set_error("output is missing u/v/dudt/dvdt/rho/p/drhodt");
    return SPH_ERR_ARG;
  }
  if (prm->heat_conduction && (!in->T || !in->dTdt || !in->kappa || !in->Cp || !out->T ||
                               !out->dTdt)) {
    set_error("heat conduction needs T, dTdt, kappa, Cp in and T, dTdt out");
    return SPH_ERR_ARG;
  }
  if ((prm->solver == 1 || (prm->free_slip && prm->bc_trick)) && !in->nw) {
    set_error("wall normals nw are required for this boundary treatment");
    return SPH_ERR_ARG;
  }
  if (prm->kernel < 0 || prm->kernel > 6) { set_error("unknown kernel"); return SPH_ERR_ARG; }
  if (prm->solver < 0 || prm->solver > 2) { set_error("unknown solver"); return SPH_ERR_ARG; }
  return SPH_OK;
}

inline SweepArgs make_args(sph_ws* ws, int bank, const sph_solver_params* prm, const sph_state_in* in,
                    sph_state_out* out) {
  SweepArgs a;
  a.prm = *prm;
  a.in = *in;
  a.out = *out;
  a.R = Recs{ws->P,  ws->U0, ws->V0, ws->Q0,  ws->NW,  ws->U1,  ws->Q1,  ws->U2,  ws->V2,
             ws->Q2, ws->UT, ws->drho, ws->U3, ws->Q3, ws->DL0, ws->DL1, ws->DL2, ws->DG0,
             ws->DG1, ws->W1, ws->W2, ws->W3};
  a.eta_varies = ws->eta_varies;
  const Bank& B = bank == kSkinBank ? ws->sk.b : ws->bank[bank];
  a.perm = B.perm;
  a.ell = B.ell;
  a.cnt = B.cnt;
  a.geo = make_geo(prm->box);
  a.n = ws->cfg.n;
  a.npad = ws->npad;
  a.lane_width = bank == kSkinBank ? ws->sk.lane_width : ws->lane_width;
  a.cut_sq = ws->cfg.cutoff_sq;
  if (bank == kSkinBank) {
    cut_band(ws->cfg, &a.cut_lo, &a.cut_hi);
  } else {
    a.cut_lo = a.cut_hi = INFINITY;
  }
  a.info = nullptr;
  a.part = ws->part_sweep;
  a.remap = ws->xcd_remap;
  a.Uf = a.R.U1; a.Vf = a.R.V0; a.Qf = a.R.Q1;
  a.Ud = a.R.U1; a.Qd = a.R.Q1;
  a.heat_in_density = prm->heat_conduction && !prm->bc_trick;
  a.bc.kind = 0;
  a.bc.u_lid[0] = a.bc.u_lid[1] = a.bc.u_lid[2] = 0.0f;
  return a;
}

// Bank 0/1: exact lists of sph_nl_build; bank 2: skin lists (needs them
// enabled and the solver box equal to the neighbour-search box).
inline int check_bank(sph_ws* ws, int bank, const sph_solver_params* prm, const char* who) {
  if (!ws || bank < 0 || bank > kSkinBank) {
    set_error(std::string(who) + ": bad workspace/bank");
    return SPH_ERR_ARG;
  }
  if (prm->dim != ws->cfg.dim) { set_error(std::string(who) + ": dim mismatch"); return SPH_ERR_ARG; }
  if (bank == kSkinBank) {
    if (!ws->sk.on) {
      set_error(std::string(who) + ": skin lists are not enabled on this workspace");
      return SPH_ERR_ARG;
    }
    if (!ws->sk.stepped && std::string(who) != "sph_step") {
      set_error(std::string(who) + ": the skin lists have not been built by sph_step yet");
      return SPH_ERR_ARG;
    }
    for (int d = 0; d < prm->dim; ++d)
      if (prm->box[d] != ws->cfg.box[d]) {
        set_error(std::string(who) + ": solver box differs from the neighbour-search box");
        return SPH_ERR_ARG;
      }
  }
  return SPH_OK;
}

}  // namespace sph
