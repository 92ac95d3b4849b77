// Hand-written adjoint (vector-Jacobian product) of one si_euler step on gfx950.
//
// The reference differentiates its step with jax.grad (reverse mode through
// XLA).  Here the transpose of the step is written out.  Scope: the standard
// SPH solver with density summation, no walls, no heat, no artificial
// viscosity and tvf = 0 (v = u), which is the 2D Taylor-Green inverse problem's
// configuration.  State n -> n+1 (integrator.py:26-54, solver.py:176-178,
// 221-256, 801):
//
//   u' = u + dt a,   r' = wrap(r + dt u'),   a' = F(r', u')
//   rho_i = m_i sum_j W(d_ij),  p_i = P(rho_i)
//   F_i   = sum_j c_ij [ -p_ij r_ij + eta_ij (u_i - u_j) ] + g
//   c_ij  = (V_i^2 + V_j^2) / m_i * W'(d) / (d + eps),  V = m / rho
//   p_ij  = (rho_j p_i + rho_i p_j) / (rho_i + rho_j)
//
// Given the cotangents (rb', ub', ab') of state n+1 it returns those of state
// n:  rb' += J_r^T ab',  ub' += J_u^T ab',  ub' += dt rb',  then
// rb = rb', ub = ub', ab = dt ub'.  Both Jacobian products are gathers over the
// pair list of r' (two sweeps: the second needs the density cotangent of the
// neighbours), with no atomics, like the forward sweeps.
#include <hip/hip_runtime.h>

#include "sph_device.h"
#include "sph_internal.h"
#include "sweep.h"

namespace sph {

constexpr int kAB = 256;

struct AdjArgs {
  SweepArgs s;               // records of state n+1 (P, U1 = (u, rho), Q1 = (p, eta, tag, .))
  const float* rb_in;        // (N, d) cotangent of r'  (orig order)
  const float* ub_in;        // (N, d) cotangent of u'
  const float* ab_in;        // (N, d) cotangent of a' = dudt'
  float* rb_out;             // (N, d) cotangents of state n
  float* ub_out;
  float* ab_out;
  float4* la;                // sorted cotangent of a'
  float dt;
};

template <int DIM>
__global__ void k_adj_gather(AdjArgs a) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= a.s.n) return;
  a.la[s] = f4(load_vec<DIM>(a.ab_in, a.s.perm[s]), 0.0f);
}

// Sweep A: density cotangent (times m, fluid only), the u' cotangent of F and
// the direct r' cotangent of F, per sorted particle k.
template <int DIM, int KIND>
__global__ __launch_bounds__(kAB) void k_adj_forces(AdjArgs aa, float4* __restrict__ ab0,
                                                    float4* __restrict__ ab1) {
  constexpr int TPP = tpp_of(DIM);
  const SweepArgs& a = aa.s;
  const int tid = logical_block(a.remap) * blockDim.x + threadIdx.x;
  const int k = tid / TPP, h = tid % TPP;
  if (k >= a.n) return;  // whole lane groups leave together
  const sph_solver_params& p = a.prm;
  const KernelEval<KIND, DIM> K{p.kernel, p.sigma, p.one_over_h};
  const Eos eos = make_eos(p);
  const float4 Pk = a.R.P[k], Uk = a.R.U1[k], Qk = a.R.Q1[k], Lk4 = aa.la[k];
  const float mk = Pk.w, rhok = Uk.w, pk = Qk.x, etak = Qk.y;
  const float Vk2 = (mk / rhok) * (mk / rhok);
  const float dPk = eos.dp_drho(rhok);
  const float inv_mk = 1.0f / mk;
  const V3 uk = v3(Uk), lk = v3(Lk4);
  float rhob = 0.0f;
  V3 ub = {0.f, 0.f, 0.f}, rb = {0.f, 0.f, 0.f};
  for_each_pair(a, k, h, [&](const int j) {
    const float4 Pj = a.R.P[j];
    const V3 dr = min_image<DIM>(Pk, Pj, a.geo);
    const float d2 = dist2<DIM>(dr.x, dr.y, dr.z);
    if (d2 == 0.0f) return;  // self pair: every term and derivative vanishes
    if (!in_cutoff<DIM>(a, d2, Pk, Pj)) return;
    const float4 Uj = a.R.U1[j], Qj = a.R.Q1[j], Lj4 = aa.la[j];
    const float mj = Pj.w, rhoj = Uj.w, pj = Qj.x, etaj = Qj.y;
    const float d = sqrtf(d2);
    const float q = d * p.one_over_h;
    const float w1 = K.dw(q);
    const float w2 = K.d2w(q);
    const float inv = 1.0f / (d + kEps);
    const float A0 = w1 * inv;                      // W'(d) / (d + eps)
    const float Ad0 = w2 * inv - w1 * inv * inv;    // d/dd of the above
    const float inv_mj = 1.0f / mj;
    const float Akj = A0 * inv_mk, Ajk = A0 * inv_mj;
    const float Adkj = Ad0 * inv_mk, Adjk = Ad0 * inv_mj;
    const float Vj2 = (mj / rhoj) * (mj / rhoj);
    const float Vs = Vk2 + Vj2;
    const float ckj = Vs * Akj, cjk = Vs * Ajk;
    const float rs = rhok + rhoj;
    const float pkj = (rhoj * pk + rhok * pj) / rs;
    const float etakj = 2.0f * etak * etaj / (etak + etaj + kEps);
    const V3 T = (-pkj) * dr + etakj * (uk - v3(Uj));
    const float Dp = rhoj * (dPk * rs + (pj - pk)) / (rs * rs);  // d p_kj / d rho_k (via p_k)
    const V3 lj = v3(Lj4);
    const V3 lc = ckj * lk - cjk * lj;
    const V3 lA = Akj * lk - Ajk * lj;
    const V3 lAd = Adkj * lk - Adjk * lj;
    rhob += (-2.0f * Vk2 / rhok) * dot<DIM>(lA, T) - Dp * dot<DIM>(lc, dr);
    ub = ub + etakj * lc;
    rb = rb + Vs * ((dot<DIM>(lAd, T) / d) * dr - pkj * lA);
  });
  rhob = lane_sum<TPP>(rhob);
  ub = lane_sum<TPP>(ub);
  rb = lane_sum<TPP>(rb);
  if (h != 0) return;
  const bool fluid = (int)Qk.z == kFluid;
  ab0[k] = f4(ub, fluid ? mk * rhob : 0.0f);  // rho = m sum W only for fluid (solver.py:796)
  ab1[k] = f4(rb, 0.0f);
}

// Sweep B: the r' cotangent through the density sums, then the kinematic
// transpose of the kick/drift, written in original order.
template <int DIM, int KIND>
__global__ __launch_bounds__(kAB) void k_adj_density(AdjArgs aa, const float4* __restrict__ ab0,
                                                     const float4* __restrict__ ab1) {
  constexpr int TPP = tpp_of(DIM);
  const SweepArgs& a = aa.s;
  const int tid = logical_block(a.remap) * blockDim.x + threadIdx.x;
  const int k = tid / TPP, h = tid % TPP;
  if (k >= a.n) return;
  const sph_solver_params& p = a.prm;
  const KernelEval<KIND, DIM> K{p.kernel, p.sigma, p.one_over_h};
  const float4 Pk = a.R.P[k];
  const float Bk = ab0[k].w;
  V3 rb = {0.f, 0.f, 0.f};
  for_each_pair(a, k, h, [&](const int j) {
    const float4 Pj = a.R.P[j];
    const V3 dr = min_image<DIM>(Pk, Pj, a.geo);
    const float d2 = dist2<DIM>(dr.x, dr.y, dr.z);
    if (d2 == 0.0f) return;
    if (!in_cutoff<DIM>(a, d2, Pk, Pj)) return;
    const float d = sqrtf(d2);
    const float w1 = K.dw(d * p.one_over_h);
    rb = rb + ((Bk + ab0[j].w) * w1 / d) * dr;
  });
  rb = lane_sum<TPP>(rb);
  if (h != 0) return;
  const int i = a.perm[k];
  const float4 A0 = ab0[k], A1 = ab1[k];
  const V3 rbt = load_vec<DIM>(aa.rb_in, i) + v3(A1) + rb;
  V3 ubt = load_vec<DIM>(aa.ub_in, i) + v3(A0);
  ubt = ubt + aa.dt * rbt;  // r' = r + dt u'
  store_vec<DIM>(aa.rb_out, i, rbt);
  store_vec<DIM>(aa.ub_out, i, ubt);
  store_vec<DIM>(aa.ab_out, i, aa.dt * ubt);  // u' = u + dt a
}

template <int DIM, int KIND>
int launch_adjoint(sph_ws* ws, AdjArgs& a, bool records_ready, hipStream_t st) {
  const int n = a.s.n;
  dim3 grid(xcd_grid(cdivl((int64_t)n * tpp_of(DIM), kAB), a.s.remap));
  if (!records_ready)
    SPH_LAUNCH(ws, "density", st, (k_density<DIM, KIND, 0><<<grid, kAB, 0, st>>>(a.s)));
  SPH_LAUNCH(ws, "adj_gather", st,
             k_adj_gather<DIM><<<(unsigned)cdivl(n, kAB), kAB, 0, st>>>(a));
  SPH_LAUNCH(ws, "adj_forces", st,
             (k_adj_forces<DIM, KIND><<<grid, kAB, 0, st>>>(a, ws->AB0, ws->AB1)));
  SPH_LAUNCH(ws, "adj_density", st,
             (k_adj_density<DIM, KIND><<<grid, kAB, 0, st>>>(a, ws->AB0, ws->AB1)));
  return SPH_OK;
}

}  // namespace sph

using namespace sph;

extern "C" int sph_step_vjp(sph_ws* ws, int bank, const sph_solver_params* prm, float dt,
                            const sph_state_in* next, const sph_cotangent_in* bar_next,
                            sph_cotangent_out* bar, int32_t records_ready, hipStream_t st) {
  if (!prm || !next || !bar_next || !bar) { set_error("sph_step_vjp: null argument"); return SPH_ERR_ARG; }
  int rc = check_bank(ws, bank, prm, "sph_step_vjp");
  if (rc) return rc;
  if (!next->r || !next->u || !next->mass || !next->eta || !next->tag || !next->rho || !next->p) {
    set_error("sph_step_vjp: state n+1 needs r, u, rho, p, mass, eta, tag");
    return SPH_ERR_ARG;
  }
  if (!bar_next->r || !bar_next->u || !bar_next->a || !bar->r || !bar->u || !bar->a) {
    set_error("sph_step_vjp: every cotangent array is required");
    return SPH_ERR_ARG;
  }
  if (prm->solver != 0 || prm->density_evolution || prm->bc_trick || prm->heat_conduction ||
      prm->artificial_alpha != 0.0f || prm->tvf_stress) {
    set_error("sph_step_vjp: only the standard SPH solver with density summation, no walls, "
              "no heat, no artificial viscosity and tvf = 0 is differentiated");
    return SPH_ERR_ARG;
  }
  if (prm->kernel != SPH_KERNEL_QSK && prm->kernel != SPH_KERNEL_WC2K) {
    set_error("sph_step_vjp: kernel must be QSK or WC2K");
    return SPH_ERR_ARG;
  }
  sph_state_in in = *next;
  in.v = next->u;   // tvf = 0: v == u
  in.g_ext = nullptr;
  sph_state_out none{};
  AdjArgs a;
  a.s = make_args(ws, bank, prm, &in, &none);
  a.rb_in = bar_next->r; a.ub_in = bar_next->u; a.ab_in = bar_next->a;
  a.rb_out = bar->r; a.ub_out = bar->u; a.ab_out = bar->a;
  a.la = ws->LA;
  a.dt = dt;
  const int n = a.s.n;
  const bool ready = records_ready != 0;
  dim3 g1((unsigned)cdivl(n, kAB));
  if (!ready) {
    if (prm->dim == 2)
      SPH_LAUNCH(ws, "gather_state", st,
                 k_gather_state<2><<<g1, kAB, 0, st>>>(a.s, in.r, in.u, in.v, nullptr, nullptr,
                                                       nullptr));
    else
      SPH_LAUNCH(ws, "gather_state", st,
                 k_gather_state<3><<<g1, kAB, 0, st>>>(a.s, in.r, in.u, in.v, nullptr, nullptr,
                                                       nullptr));
  }
  if (prm->dim == 2)
    return prm->kernel == SPH_KERNEL_QSK ? launch_adjoint<2, SPH_KERNEL_QSK>(ws, a, ready, st)
                                         : launch_adjoint<2, SPH_KERNEL_WC2K>(ws, a, ready, st);
  return prm->kernel == SPH_KERNEL_QSK ? launch_adjoint<3, SPH_KERNEL_QSK>(ws, a, ready, st)
                                       : launch_adjoint<3, SPH_KERNEL_WC2K>(ws, a, ready, st);
}
