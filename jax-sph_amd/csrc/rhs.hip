// WCSPH right-hand side and the fused si_euler step on gfx950.
//
// Particle data are gathered once per step into cell-sorted float4 records;
// every sweep is a gather over the particle's own pair list (no atomics, a
// fixed neighbour order, so results are bitwise reproducible run to run).
//
// Replaces solver.py:705-949 (WCSPH.forward), integrator.py:22-56 (si_euler)
// and the per-step case hooks tgv.py:56-57, ldc.py:130-144, db.py:299-307.
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>

#include "ens.h"
#include "sph_device.h"
#include "sph_internal.h"
#include "sweep.h"

namespace sph {

// Bank helpers from nl.hip
int bin_init(sph_ws* ws, const int64_t* prev_info, int64_t* info, hipStream_t st);
int integrate_and_hash(sph_ws* ws, const float* r, const float* u, const float* dudt,
                       const float* dvdt, float* r_out, float* u_out, float* v_out, float dt,
                       float ctvf, const float* box, hipStream_t st);
int bin_finish(sph_ws* ws, int bank, const float* r, int64_t* info, hipStream_t st);
int build_pairs(sph_ws* ws, int bank, int64_t max_occupancy, int64_t* info, hipStream_t st);
int skin_begin(sph_ws* ws, const float* r, const float* u, const float* dudt, const float* dvdt,
               float* r_out, float* u_out, float* v_out, float dt, float ctvf, const float* box,
               const int64_t* prev_info, int64_t* info, hipStream_t st);
int skin_build(sph_ws* ws, int64_t* info, hipStream_t st);
int skin_finish(sph_ws* ws, int64_t max_occupancy, int64_t* info, hipStream_t st);

// ------------------------------------------------------------- delta-SPH
// Renormalisation matrix of solver.py:42-58: M = sum_j (-r_ij) (x) grad W_ij V_j
// over the pairs of i, L = inv(M).  The reference's second matrix (grouped by
// the other index, solver.py:57) is the termwise negation of M, so its
// inverse is -L: k_delta_grad uses -L of the neighbour.
template <int DIM>
__device__ __forceinline__ void store_mat(const SweepArgs& a, int s, const float* m) {
  if (DIM == 2) {
    a.R.DL0[s] = make_float4(m[0], m[1], m[2], m[3]);
  } else {
    a.R.DL0[s] = make_float4(m[0], m[1], m[2], m[3]);
    a.R.DL1[s] = make_float4(m[4], m[5], m[6], m[7]);
    a.R.DL2[s] = make_float4(m[8], 0.f, 0.f, 0.f);
  }
}

template <int DIM>
__device__ __forceinline__ V3 mat_vec(const SweepArgs& a, int s, V3 x) {
  const float4 m0 = a.R.DL0[s];
  if (DIM == 2) return V3{m0.x * x.x + m0.y * x.y, m0.z * x.x + m0.w * x.y, 0.f};
  const float4 m1 = a.R.DL1[s], m2 = a.R.DL2[s];
  return V3{m0.x * x.x + m0.y * x.y + m0.z * x.z,  // row 0: m0.x m0.y m0.z
            m0.w * x.x + m1.x * x.y + m1.y * x.z,  // row 1: m0.w m1.x m1.y
            m1.z * x.x + m1.w * x.y + m2.x * x.z};  // row 2: m1.z m1.w m2.x
}

template <int DIM, int KIND>
__device__ __forceinline__ void delta_L_run(const SweepArgs& a, const int blk) {
  constexpr int TPP = tpp_of(DIM);
  const int tid = blk * blockDim.x + threadIdx.x;
  const int s = tid / TPP, h = tid % TPP;
  if (s >= a.n) return;
  const sph_solver_params& p = a.prm;
  const KernelEval<KIND, DIM> K{p.kernel, p.sigma, p.one_over_h};
  const float4 Pi = a.R.P[s];
  float M[9] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  for_each_pair(a, s, h, [&](const int j) {
    const float4 Pj = a.R.P[j];
    const V3 dr = min_image<DIM>(Pi, Pj, a.geo);
    const float d2 = dist2<DIM>(dr.x, dr.y, dr.z);
    if (!in_cutoff<DIM>(a, d2, Pi, Pj)) return;
    const float d = sqrt_fast(d2);
    const float vj = Pj.w * rcp_fast(a.R.U0[j].w);
    const V3 kg = (K.dw(d * p.one_over_h) * rcp_fast(d + kEps) * vj) * dr;  // grad W_ij V_j
    const float r[3] = {-dr.x, -dr.y, -dr.z}, g[3] = {kg.x, kg.y, kg.z};
#pragma unroll
    for (int x = 0; x < DIM; ++x)
#pragma unroll
      for (int y = 0; y < DIM; ++y) M[x * DIM + y] += r[x] * g[y];
  });
#pragma unroll
  for (int k = 0; k < DIM * DIM; ++k) M[k] = lane_sum<TPP>(M[k]);
  if (h != 0) return;
  float L[9];
  if (DIM == 2) {
    const float det = M[0] * M[3] - M[1] * M[2];
    const float id = 1.0f / det;
    L[0] = M[3] * id; L[1] = -M[1] * id; L[2] = -M[2] * id; L[3] = M[0] * id;
  } else {
    const float c00 = M[4] * M[8] - M[5] * M[7], c01 = M[5] * M[6] - M[3] * M[8],
                c02 = M[3] * M[7] - M[4] * M[6];
    const float det = M[0] * c00 + M[1] * c01 + M[2] * c02;
    const float id = 1.0f / det;
    L[0] = c00 * id;
    L[1] = (M[2] * M[7] - M[1] * M[8]) * id;
    L[2] = (M[1] * M[5] - M[2] * M[4]) * id;
    L[3] = c01 * id;
    L[4] = (M[0] * M[8] - M[2] * M[6]) * id;
    L[5] = (M[2] * M[3] - M[0] * M[5]) * id;
    L[6] = c02 * id;
    L[7] = (M[1] * M[6] - M[0] * M[7]) * id;
    L[8] = (M[0] * M[4] - M[1] * M[3]) * id;
  }
  store_mat<DIM>(a, s, L);
}
SPH_SWEEP_KERNELS(k_delta_L, delta_L_run, (int DIM, int KIND), (DIM, KIND), 8)

// Density-gradient terms of solver.py:60-66:
//   grad_i = L_i sum_j (rho_j - rho_i) grad W_ij V_j
//   grad_j = sum_j (rho_i - rho_j) (-L_j) (-grad W_ij V_i)
template <int DIM, int KIND>
__device__ __forceinline__ void delta_grad_run(const SweepArgs& a, const int blk) {
  constexpr int TPP = tpp_of(DIM);
  const int tid = blk * blockDim.x + threadIdx.x;
  const int s = tid / TPP, h = tid % TPP;
  if (s >= a.n) return;
  const sph_solver_params& p = a.prm;
  const KernelEval<KIND, DIM> K{p.kernel, p.sigma, p.one_over_h};
  const float4 Pi = a.R.P[s];
  const float rhoi = a.R.U0[s].w;
  const float vi = Pi.w / rhoi;
  V3 si = {0.f, 0.f, 0.f}, gj = {0.f, 0.f, 0.f};
  for_each_pair(a, s, h, [&](const int j) {
    const float4 Pj = a.R.P[j];
    const V3 dr = min_image<DIM>(Pi, Pj, a.geo);
    const float d2 = dist2<DIM>(dr.x, dr.y, dr.z);
    if (!in_cutoff<DIM>(a, d2, Pi, Pj)) return;
    const float d = sqrt_fast(d2);
    const float rhoj = a.R.U0[j].w;
    const V3 kg = (K.dw(d * p.one_over_h) * rcp_fast(d + kEps)) * dr;
    si = si + ((rhoj - rhoi) * (Pj.w * rcp_fast(rhoj))) * kg;
    gj = gj + ((rhoi - rhoj) * vi) * mat_vec<DIM>(a, j, kg);  // (-L_j)(-grad W V_i)
  });
  si = lane_sum<TPP>(si);
  gj = lane_sum<TPP>(gj);
  if (h != 0) return;
  a.R.DG0[s] = f4(mat_vec<DIM>(a, s, si), 0.0f);
  a.R.DG1[s] = f4(gj, 0.0f);
}
SPH_SWEEP_KERNELS(k_delta_grad, delta_grad_run, (int DIM, int KIND), (DIM, KIND), 8)

// ------------------------------------------------------------- renormalisation
// solver.py:167-173 after a continuity update: rho = sum m_j W / min(sum (m/rho)_j W, 1)
// over all particles, then the EOS; writes the U3/Q3 records.
template <int DIM, int KIND>
__device__ __forceinline__ void renorm_run(const SweepArgs& a, const int blk) {
  constexpr int TPP = tpp_of(DIM);
  const int tid = blk * blockDim.x + threadIdx.x;
  const int s = tid / TPP, h = tid % TPP;
  if (s >= a.n) return;
  const sph_solver_params& p = a.prm;
  const KernelEval<KIND, DIM> K{p.kernel, p.sigma, p.one_over_h};
  const Eos eos = make_eos(p);
  const float4 Pi = a.R.P[s];
  float nom = 0.f, den = 0.f;
  for_each_pair(a, s, h, [&](const int j) {
    const float4 Pj = a.R.P[j];
    const V3 dr = min_image<DIM>(Pi, Pj, a.geo);
    const float d2 = dist2<DIM>(dr.x, dr.y, dr.z);
    if (!in_cutoff<DIM>(a, d2, Pi, Pj)) return;
    const float w = K.w(sqrt_fast(d2) * p.one_over_h);
    nom += Pj.w * w;
    den += (Pj.w * rcp_fast(a.R.U1[j].w)) * w;
  });
  nom = lane_sum<TPP>(nom);
  den = lane_sum<TPP>(den);
  if (h != 0) return;
  const float rho = nom / (den > 1.0f ? 1.0f : den);
  const float4 Ui = a.R.U1[s], Qi = a.R.Q1[s];
  const float pr = eos.p_fn(rho);
  const float4 Vi = a.R.V0[s];
  a.R.U3[s] = make_float4(Ui.x, Ui.y, Ui.z, rho);
  a.R.Q3[s] = make_float4(pr, Qi.y, Qi.z, Qi.w);
  a.R.W3[s] = make_float4(Vi.x - Ui.x, Vi.y - Ui.y, Vi.z - Ui.z, pr);
}
SPH_SWEEP_KERNELS(k_renorm, renorm_run, (int DIM, int KIND), (DIM, KIND), 8)


// ------------------------------------------------------------- forces
// SOLVER 0: acceleration_standard_fn (solver.py:221-256) + transport velocity
//           (solver.py:199-213); TVF: include the A = rho u (v - u) stress.
// SOLVER 1: acceleration_fn_riemann (solver.py:316-401; its A term is zero).
// Plus artificial viscosity (solver.py:404-428), heat (solver.py:591-610),
// dudt + g_ext (solver.py:936) and the case boundary hook (integrator.py:54).
// EXTRAS: artificial viscosity and heat conduction are compiled in (otherwise
// the sweep cannot issue the extra per-pair loads they need).
template <int DIM, int KIND, int SOLVER, bool TVF, bool EXTRAS>
__device__ __forceinline__ void force_run(const SweepArgs& a, const int blk) {
  constexpr int TPP = tpp_of(DIM);
  const int tid = blk * blockDim.x + threadIdx.x;
  const int s = tid / TPP, h = tid % TPP;
  if (s >= a.n) return;  // whole lane groups leave together
  const sph_solver_params& p = a.prm;
  const KernelEval<KIND, DIM> K{p.kernel, p.sigma, p.one_over_h};
  const Eos eos = make_eos(p);
  const float4 Pi = a.R.P[s], Ui = a.Uf[s], Vi = a.Vf[s], Qi = a.Qf[s];
  const int i = a.perm[s];
  const int tag_i = (int)Qi.z;
  const float mi = Pi.w, rhoi = Ui.w, pi_ = Qi.x, etai = Qi.y, Ti = Qi.w;
  const float inv_mi = 1.0f / mi;
  const float inv_rhoi = 1.0f / rhoi;
  const float Vi2 = (mi / rhoi) * (mi / rhoi);
  const V3 ui = v3(Ui), vi = v3(Vi);
  const V3 hAi = (0.5f * rhoi) * ui;  // rho_i u_i / 2
  const V3 wi = vi - ui;            // v_i - u_i
  const float pbg = eos.p_fn(0.0f);  // background pressure (solver.py:802)
  const bool av = EXTRAS && p.artificial_alpha != 0.0f;
  const bool fluid_i = tag_i == kFluid;
  const bool heat = EXTRAS && p.heat_conduction != 0;
  const float kappa_i = Vi.w;
  const float Cp_i = heat ? a.R.NW[s].w : 1.0f;
  const float av_pref = p.artificial_alpha * p.av_h * 10.0f;  // alpha * h_ab * c_ab
  const float delta_pref = p.diff_alpha * p.support * p.c_ref * p.rho_ref;  // alpha h c rho_0
  const float av_h2 = 0.01f * p.av_h * p.av_h;
  V3 gi = g_ext_of<DIM>(a, i);
  V3 acc = {0.f, 0.f, 0.f}, accv = {0.f, 0.f, 0.f};
  float dT = 0.0f;
  // The standard solver without extras reads (p, eta) of j only: p (and v for
  // the transport-velocity stress) from the (v, p) record, eta from Q unless
  // every particle has the same viscosity (then eta_j = eta_i, the same value).
  constexpr bool kLean = SOLVER == 0 && !EXTRAS;
  const bool need_q = !kLean || *a.eta_varies != 0;
  const float eta_ii = 2.0f * etai * etai * rcp_fast(etai + etai + kEps);
  struct Rec { float4 P; };
  const auto load = [&](const int j) { Rec r; r.P = a.R.P[j]; return r; };
  for_each_pair_ahead(a, s, h, load, [&](const int j, const Rec& rr) {
    const float4 Pj = rr.P;
    const V3 dr = min_image<DIM>(Pi, Pj, a.geo);
    const float d2 = dist2<DIM>(dr.x, dr.y, dr.z);
    if (d2 == 0.0f) return;  // self pair / coincident: every term is exactly 0
    if (!in_cutoff<DIM>(a, d2, Pi, Pj)) return;
    // 1/(d + EPS) is 1/d in float32 for every d above ~4e-9 (EPS = 2.2e-16)
    const float inv_d = rsq_fast(d2);
    const float d = d2 * inv_d;
    const float q = d * p.one_over_h;
    const float gw = K.dw(q);
    const float4 Uj = a.Uf[j];
    float4 Qj = make_float4(0.0f, etai, 0.0f, 0.0f);
    if (need_q) Qj = a.Qf[j];
    float4 Wj = make_float4(0.0f, 0.0f, 0.0f, Qj.x);  // (v - u, p) of j
    if (SOLVER != 1 && (TVF || !need_q)) Wj = a.Wf[j];
    const float mj = Pj.w, rhoj = Uj.w, pj = Wj.w, etaj = Qj.y;
    const V3 uj = v3(Uj);
    const float Vj = mj * rcp_fast(rhoj);
    const float cpre = (Vi2 + Vj * Vj) * inv_mi * gw * inv_d;
    // with one viscosity for all particles eta_ij is the particle's own value
    float eta_ij = eta_ii;
    if (need_q) eta_ij = 2.0f * etai * etaj * rcp_fast(etai + etaj + kEps);
    if (SOLVER == 0 || SOLVER == 2) {
      const float pij = (rhoj * pi_ + rhoi * pj) * rcp_fast(rhoi + rhoj);
      V3 t = (-pij) * dr + eta_ij * (ui - uj);
      if (TVF) {
        // 0.5 (A_i + A_j) . r with A = rho u (v - u)^T
        const V3 Ar = dot<DIM>(wi, dr) * hAi + (0.5f * dot<DIM>(v3(Wj), dr) * rhoj) * uj;
        t = t + Ar;
      }
      acc = acc + cpre * t;
      if (SOLVER == 2 && (int)Qj.z == kFluid) {
        // delta-SPH velocity diffusion (solver.py:300-311)
        const float pij_d = dot<DIM>(uj - ui, (-1.0f) * dr) * inv_d * inv_d;
        acc = acc + (Vj * pij_d * gw * inv_d * delta_pref * inv_rhoi) * dr;
      }
    } else {
      const int tag_j = (int)Qj.z;
      const bool wall_j = is_wall(tag_j);
      const V3 e = inv_d * dr;
      const V3 kg = gw * e;
      const V3 me = (-1.0f) * e;
      float uL, uR, pR, rhoR;
      V3 vij;
      if (wall_j) {
        const V3 nj = v3(a.R.NW[j]);
        uL = -dot<DIM>(ui, nj);
        uR = -uL + 2.0f * dot<DIM>(uj, nj);
        pR = pi_ + rhoi * dot<DIM>(gi, (-1.0f) * dr);
        rhoR = eos.rho_fn(pR);
        // u_tilde_j: Shepard wall velocity when the Riemann wall BC is on, else u_j
        const V3 ud = p.bc_trick ? 2.0f * uj - v3(a.R.UT[j]) : uj;
        vij = ui - ud;
      } else {
        uL = dot<DIM>(ui, me);
        uR = dot<DIM>(uj, me);
        pR = pj;
        rhoR = rhoj;
        vij = ui - uj;
      }
      const float Pavg = (pi_ + pR) * 0.5f;
      const float rhoavg = (rhoi + rhoR) * 0.5f;
      float beta;
      if (p.eta_limiter == -1.0f) {
        beta = p.c_ref;
      } else {
        beta = fminf(p.eta_limiter * fmaxf(uL - uR, 0.0f), p.c_ref);
      }
      const float Pstar = Pavg + 0.5f * rhoavg * (uL - uR) * beta;
      const float irr = rcp_fast(rhoi * rhoj);
      const V3 eq9 = (-2.0f * mj * (Pstar * irr)) * kg;
      float mask = 1.0f;
      if (p.bc_trick && p.free_slip) mask = fluid_i ? 1.0f : 0.0f;
      const V3 eq6 = (2.0f * mj * eta_ij * irr * inv_d * gw * mask) * vij;
      acc = acc + eq9 + eq6;
    }
    accv = accv + cpre * dr;  // times p_bg after the sum
    if (av) {
      const bool fluid_j = (int)Qj.z == kFluid;
      if (fluid_i && fluid_j) {
        const float rho_ab = (rhoi + rhoj) * 0.5f;
        const float num = mj * av_pref * dot<DIM>(ui - uj, dr);
        const float den = rho_ab * (d2 + av_h2);
        acc = acc + (num * gw * inv_d * rcp_fast(den)) * dr;
      }
    }
    if (heat) {
      const float kj = a.Vf[j].w;
      const float keff = (kappa_i * kj) * rcp_fast(kappa_i + kj);
      const float F = dot<DIM>(dr, (gw * inv_d) * dr) * rcp_fast(d2 + kEps);
      dT += 4.0f * mj * keff * (Ti - Qj.w) * F * rcp_fast(Cp_i * rhoi * rhoj);
    }
  });
  acc = lane_sum<TPP>(acc);
  accv = lane_sum<TPP>(accv);
  if (heat) dT = lane_sum<TPP>(dT);
  if (h != 0) return;
  V3 dudt = acc + gi;
  V3 dvdt = pbg * accv;
  V3 u_out = ui, v_out = vi;
  float p_out = pi_;
  // case boundary hook (applied to the output state, integrator.py:54)
  if (a.bc.kind == 1) {
    if (tag_i == kSolidWall) {
      u_out = v_out = V3{0.f, 0.f, 0.f};
      dudt = dvdt = V3{0.f, 0.f, 0.f};
    } else if (tag_i == kMovingWall) {
      u_out = v_out = V3{a.bc.u_lid[0], a.bc.u_lid[1], a.bc.u_lid[2]};
      dudt = dvdt = V3{0.f, 0.f, 0.f};
    }
  } else if (a.bc.kind == 2) {
    if (tag_i == kSolidWall) {
      u_out = v_out = V3{0.f, 0.f, 0.f};
      dudt = dvdt = V3{0.f, 0.f, 0.f};
      p_out = 0.0f;
    }
  } else if (a.bc.kind == 3) {
    if (is_wall(tag_i)) {
      u_out = v_out = V3{0.f, 0.f, 0.f};
      dudt = dvdt = V3{0.f, 0.f, 0.f};
    }
  }
  const sph_state_out& o = a.out;
  store_vec<DIM>(o.dudt, i, dudt);
  store_vec<DIM>(o.dvdt, i, dvdt);
  store_vec<DIM>(o.u, i, u_out);
  store_vec<DIM>(o.v, i, v_out);
  o.rho[i] = rhoi;
  o.p[i] = p_out;
  o.drhodt[i] = a.R.drho[s];
  if (o.T) o.T[i] = Ti;
  if (o.dTdt) o.dTdt[i] = heat ? dT : (a.in.dTdt ? a.in.dTdt[i] : 0.0f);
}
SPH_SWEEP_KERNELS(k_force, force_run, (int DIM, int KIND, int SOLVER, bool TVF, bool EXTRAS),
                  (DIM, KIND, SOLVER, TVF, EXTRAS),
                  (SOLVER == 0 && !EXTRAS && KIND >= 0 ? 8 : 1))

// ------------------------------------------------------------- integrate
// si_euler kick + transport velocity + drift + periodic wrap (integrator.py:26-30)
template <int DIM>
__global__ void k_integrate(const float* __restrict__ r, const float* __restrict__ u,
                            const float* __restrict__ dudt, const float* __restrict__ dvdt,
                            float* __restrict__ r_out, float* __restrict__ u_out,
                            float* __restrict__ v_out, int n, float dt, float ctvf, float bx,
                            float by, float bz) {
#pragma clang fp contract(off)
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (int64_t)n * DIM) return;
  const int d = (int)(k % DIM);
  const float side = d == 0 ? bx : (d == 1 ? by : bz);
  float uu = u[k] + dt * dudt[k];
  float vv = uu + ctvf * dvdt[k];
  float x = r[k] + dt * vv;
  if (x >= side || x < 0.0f) x = jmod(x, side);
  u_out[k] = uu;
  v_out[k] = vv;
  r_out[k] = x;
}

// kernel.py: W(r), dW/dr elementwise
template <int DIM>
__global__ void k_kernel_eval(int kind, float ih, float sg, const float* __restrict__ r, int64_t n,
                              float* __restrict__ w, float* __restrict__ dw) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  KernelEval<-1, DIM> K{kind, sg, ih};
  float q = r[k] * ih;
  if (w) w[k] = K.w(q);
  if (dw) dw[k] = K.dw(q);
}

}  // namespace sph

using namespace sph;

namespace sph {

template <int DIM, int KIND, bool X, typename F>
int for_force(const SweepPlan& pl, F&& f) {
  if (pl.force == 0)
    return pl.tvf ? f("force", k_force<DIM, KIND, 0, true, X>, k_force_ens<DIM, KIND, 0, true, X>)
                  : f("force", k_force<DIM, KIND, 0, false, X>,
                      k_force_ens<DIM, KIND, 0, false, X>);
  if (pl.force == 2)
    return pl.tvf ? f("force", k_force<DIM, KIND, 2, true, X>, k_force_ens<DIM, KIND, 2, true, X>)
                  : f("force", k_force<DIM, KIND, 2, false, X>,
                      k_force_ens<DIM, KIND, 2, false, X>);
  return f("force", k_force<DIM, KIND, 1, false, X>, k_force_ens<DIM, KIND, 1, false, X>);
}

// Calls f(name, kernel, ensemble kernel) for each sweep of the plan, in order;
// stops at the first non-zero return.
template <int DIM, int KIND, typename F>
int for_plan(const SweepPlan& pl, F&& f) {
  int rc = 0;
  switch (pl.density) {
    case 0: rc = f("density", k_density<DIM, KIND, 0>, k_density_ens<DIM, KIND, 0>); break;
    case 1: rc = f("density", k_density<DIM, KIND, 1>, k_density_ens<DIM, KIND, 1>); break;
    case 3:
      if ((rc = f("delta_L", k_delta_L<DIM, KIND>, k_delta_L_ens<DIM, KIND>))) return rc;
      if ((rc = f("delta_grad", k_delta_grad<DIM, KIND>, k_delta_grad_ens<DIM, KIND>))) return rc;
      rc = f("density", k_density<DIM, KIND, 3>, k_density_ens<DIM, KIND, 3>);
      break;
    default: rc = f("density", k_density<DIM, KIND, 2>, k_density_ens<DIM, KIND, 2>); break;
  }
  if (rc) return rc;
  if (pl.renorm && (rc = f("renorm", k_renorm<DIM, KIND>, k_renorm_ens<DIM, KIND>))) return rc;
  if (pl.walls == 0) rc = f("walls", k_walls<DIM, KIND, 0>, k_walls_ens<DIM, KIND, 0>);
  if (pl.walls == 1) rc = f("walls", k_walls<DIM, KIND, 1>, k_walls_ens<DIM, KIND, 1>);
  if (rc) return rc;
  return pl.extras ? for_force<DIM, KIND, true>(pl, f) : for_force<DIM, KIND, false>(pl, f);
}

template <int DIM, typename F>
int for_plan_kind(const SweepPlan& pl, F&& f) {
  switch (pl.kind) {
    case SPH_KERNEL_QSK: return for_plan<DIM, SPH_KERNEL_QSK>(pl, f);
    case SPH_KERNEL_WC2K: return for_plan<DIM, SPH_KERNEL_WC2K>(pl, f);
    default: return for_plan<DIM, -1>(pl, f);
  }
}

// ------------------------------------------------------------ ensembles
// Per-member gather and sweep arguments, exactly those sph_step passes on
// bank 2 (the gather reads the step's new state; the sweeps count pairs).
struct EnsRHS {
  int n = 0;
  GatherArgs* gather = nullptr;
  SweepArgs* sweep = nullptr;
};

int ens_rhs_create(const std::vector<EnsMember>& members, EnsRHS** out) {
  const int n = (int)members.size();
  std::vector<GatherArgs> g(n);
  std::vector<SweepArgs> s(n);
  for (int k = 0; k < n; ++k) {
    const EnsMember& m = members[k];
    sph_ws* ws = m.ws;
    SweepArgs a = make_args(ws, kSkinBank, &m.prm, &m.in, const_cast<sph_state_out*>(&m.out));
    a.bc = m.bc;
    g[k] = GatherArgs{a, m.r_out, m.out.u, m.out.v, ws->sk.b.pos, ws->sk.flag, ws->sk.posb};
    a.in.r = m.r_out;
    a.in.u = m.out.u;
    a.in.v = m.out.v;
    a.info = m.info;
    select_records(a, plan_of(m.prm));
    s[k] = a;
  }
  EnsRHS* t = new EnsRHS;
  t->n = n;
  hipError_t e = hipMalloc((void**)&t->gather, sizeof(GatherArgs) * n);
  if (e == hipSuccess) e = hipMalloc((void**)&t->sweep, sizeof(SweepArgs) * n);
  if (e == hipSuccess)
    e = hipMemcpy(t->gather, g.data(), sizeof(GatherArgs) * n, hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(t->sweep, s.data(), sizeof(SweepArgs) * n, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    ens_rhs_destroy(t);
    return hip_fail(e, "ensemble tables");
  }
  *out = t;
  return SPH_OK;
}

void ens_rhs_destroy(EnsRHS* t) {
  if (!t) return;
  (void)hipFree(t->gather);
  (void)hipFree(t->sweep);
  delete t;
}

int ens_gather(EnsRHS* t, int dim, const EnsMap& part, sph_ws* prof, hipStream_t st) {
  if (dim == 2)
    SPH_LAUNCH(prof, "ens_gather_state", st,
               k_gather_state_ens<2><<<part.blocks, kB, 0, st>>>(t->gather, part.d));
  else
    SPH_LAUNCH(prof, "ens_gather_state", st,
               k_gather_state_ens<3><<<part.blocks, kB, 0, st>>>(t->gather, part.d));
  return SPH_OK;
}

int ens_sweeps(EnsRHS* t, int dim, const SweepPlan& plan, const EnsMap& lanes, sph_ws* prof,
               hipStream_t st) {
  auto f = [&](const char* name, SweepKernel, EnsSweepKernel k) -> int {
    SPH_LAUNCH(prof, name, st, (k<<<lanes.blocks, kB, 0, st>>>(t->sweep, lanes.d)));
    return 0;
  };
  return dim == 2 ? for_plan_kind<2>(plan, f) : for_plan_kind<3>(plan, f);
}

}  // namespace sph

namespace {

// The sweeps of one list (density ... force), one launch each.
template <int DIM>
int launch_kind(sph_ws* ws, SweepArgs& a, hipStream_t st) {
  const SweepPlan pl = plan_of(a.prm);
  select_records(a, pl);
  const dim3 grid(xcd_grid(cdivl((int64_t)a.n * tpp_of(DIM), kB), a.remap));
  return for_plan_kind<DIM>(pl, [&](const char* name, SweepKernel k, EnsSweepKernel) -> int {
    SPH_LAUNCH(ws, name, st, (k<<<grid, kB, 0, st>>>(a)));
    return 0;
  });
}

}  // namespace

extern "C" int sph_forward(sph_ws* ws, int bank, const sph_solver_params* prm,
                           const sph_state_in* in, sph_state_out* out, hipStream_t st) {
  int rc = check_state(prm, in, out);
  if (rc) return rc;
  rc = check_bank(ws, bank, prm, "sph_forward");
  if (rc) return rc;
  SweepArgs a = make_args(ws, bank, prm, in, out);
  dim3 grid((unsigned)cdivl(a.n, kB));
  if (prm->dim == 2) {
    SPH_LAUNCH(ws, "gather_state", st,
               k_gather_state<2><<<grid, kB, 0, st>>>(a, in->r, in->u, in->v, nullptr, nullptr,
                                                      nullptr));
    return launch_kind<2>(ws, a, st);
  }
  SPH_LAUNCH(ws, "gather_state", st,
             k_gather_state<3><<<grid, kB, 0, st>>>(a, in->r, in->u, in->v, nullptr, nullptr,
                                                    nullptr));
  return launch_kind<3>(ws, a, st);
}

extern "C" int sph_step(sph_ws* ws, int bank, const sph_solver_params* prm, float dt, float tvf_coef,
                        const sph_bc_spec* bc, const sph_state_in* in, float* r_out,
                        sph_state_out* out, const int64_t* prev_info, int64_t max_occupancy,
                        int64_t* info, hipStream_t st) {
  if (!r_out || !info) {
    set_error("sph_step: bad argument");
    return SPH_ERR_ARG;
  }
  int rc = check_state(prm, in, out);
  if (rc) return rc;
  rc = check_bank(ws, bank, prm, "sph_step");
  if (rc) return rc;
  if (!in->dudt || !in->dvdt) { set_error("sph_step: dudt/dvdt required"); return SPH_ERR_ARG; }
  if (in->g_ext) { set_error("sph_step: position-dependent g_ext is not fused"); return SPH_ERR_ARG; }
  const int n = ws->cfg.n, dim = prm->dim;
  if (bank == kSkinBank) {
    // 1. kick/drift + rebuild test + reference-grid occupancy; skin binning if needed
    rc = skin_begin(ws, in->r, in->u, in->dudt, in->dvdt, r_out, out->u, out->v, dt, tvf_coef,
                    prm->box, prev_info, info, st);
    if (rc) return rc;
    ws->sk.stepped = true;
    // 2. sorted records (and, on a rebuild, the build positions), then the lists
    SweepArgs a = make_args(ws, bank, prm, in, out);
    if (bc) a.bc = *bc;
    dim3 grid((unsigned)cdivl(n, kB));
    if (dim == 2)
      SPH_LAUNCH(ws, "gather_state", st,
                 k_gather_state<2><<<grid, kB, 0, st>>>(a, r_out, out->u, out->v, ws->sk.b.pos,
                                                        ws->sk.flag, ws->sk.posb));
    else
      SPH_LAUNCH(ws, "gather_state", st,
                 k_gather_state<3><<<grid, kB, 0, st>>>(a, r_out, out->u, out->v, ws->sk.b.pos,
                                                        ws->sk.flag, ws->sk.posb));
    rc = skin_build(ws, info, st);
    if (rc) return rc;
    // 3. forward + case hook over the pairs inside the cutoff, counted on the way
    a.in.r = r_out;
    a.in.u = out->u;
    a.in.v = out->v;
    a.info = info;
    rc = dim == 2 ? launch_kind<2>(ws, a, st) : launch_kind<3>(ws, a, st);
    if (rc) return rc;
    return skin_finish(ws, max_occupancy, info, st);
  }
  // 1. kick/drift into the output buffers, binned at the new positions
  rc = bin_init(ws, prev_info, info, st);
  if (rc) return rc;
  rc = integrate_and_hash(ws, in->r, in->u, in->dudt, in->dvdt, r_out, out->u, out->v, dt,
                          tvf_coef, prm->box, st);
  if (rc) return rc;
  rc = bin_finish(ws, bank, nullptr, info, st);
  if (rc) return rc;
  // 2. sorted records (also the positions of the pair build), then the pairs
  SweepArgs a = make_args(ws, bank, prm, in, out);
  if (bc) a.bc = *bc;
  dim3 grid((unsigned)cdivl(n, kB));
  float4* nl_pos = ws->bank[bank].pos;
  if (dim == 2)
    SPH_LAUNCH(ws, "gather_state", st,
               k_gather_state<2><<<grid, kB, 0, st>>>(a, r_out, out->u, out->v, nl_pos, nullptr,
                                                      nullptr));
  else
    SPH_LAUNCH(ws, "gather_state", st,
               k_gather_state<3><<<grid, kB, 0, st>>>(a, r_out, out->u, out->v, nl_pos, nullptr,
                                                      nullptr));
  rc = build_pairs(ws, bank, max_occupancy, info, st);
  if (rc) return rc;
  // 3. forward + case hook
  a.in.r = r_out;
  a.in.u = out->u;
  a.in.v = out->v;
  return dim == 2 ? launch_kind<2>(ws, a, st) : launch_kind<3>(ws, a, st);
}

extern "C" int sph_kernel_eval(int32_t kind, int32_t dim, float one_over_h, float sigma,
                               const float* r, int64_t n, float* w, float* dw, hipStream_t st) {
  if (!r || n < 0 || kind < 0 || kind > 6 || dim < 1 || dim > 3) {
    set_error("sph_kernel_eval: bad argument");
    return SPH_ERR_ARG;
  }
  if (n == 0) return SPH_OK;
  unsigned g = (unsigned)cdivl(n, kB);
  if (dim == 1) k_kernel_eval<1><<<g, kB, 0, st>>>(kind, one_over_h, sigma, r, n, w, dw);
  else if (dim == 2) k_kernel_eval<2><<<g, kB, 0, st>>>(kind, one_over_h, sigma, r, n, w, dw);
  else k_kernel_eval<3><<<g, kB, 0, st>>>(kind, one_over_h, sigma, r, n, w, dw);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(e, "kernel_eval");
  return SPH_OK;
}

extern "C" int sph_integrate(int32_t dim, int64_t n, const float* r, const float* u,
                             const float* dudt, const float* dvdt, float* r_out, float* u_out,
                             float* v_out, float dt, float tvf_coef, const float* box,
                             hipStream_t st) {
  if (!r || !u || !dudt || !dvdt || !r_out || !u_out || !v_out || !box || n < 0 ||
      (dim != 2 && dim != 3)) {
    set_error("sph_integrate: bad argument");
    return SPH_ERR_ARG;
  }
  if (n == 0) return SPH_OK;
  const int64_t nk = n * dim;
  unsigned g = (unsigned)cdivl(nk, kB);
  if (dim == 2)
    k_integrate<2><<<g, kB, 0, st>>>(r, u, dudt, dvdt, r_out, u_out, v_out, (int)n, dt, tvf_coef,
                                     box[0], box[1], 0.0f);
  else
    k_integrate<3><<<g, kB, 0, st>>>(r, u, dudt, dvdt, r_out, u_out, v_out, (int)n, dt, tvf_coef,
                                     box[0], box[1], box[2]);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(e, "integrate");
  return SPH_OK;
}
