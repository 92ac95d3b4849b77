// Hand-written adjoint (vector-Jacobian product) of one si_euler step on gfx950.
//
// The reference differentiates its step with jax.grad (reverse mode through
// XLA).  Here the transpose of the step is written out.  Scope: the Riemann
// solver with density summation or continuity and tvf = 0, walls held by the
// case hook (k_adj_forces_rie, k_adj_rie_walls, k_adj_rie_continuity), and
// the standard SPH solver, density summation or continuity (sweep C), with or
// without the transport-velocity formulation (tvf), artificial viscosity
// (k_adj_forces) and the generalized wall boundary condition for no-slip
// walls (below); delta-SPH with density summation (the standard solver plus
// its velocity diffusion, in k_adj_forces).  Heat conduction (solver.py:832-850) may be on: the
// temperature enters neither the EOS nor the forces, so u' and r' do not
// depend on it and the transpose below is the whole one for them.
// State n -> n+1 (integrator.py:26-54, solver.py:176-178, 199-256, 801-802):
//
//   u' = u + dt a,   v' = u' + c b  (c = tvf dt / 2),   r' = wrap(r + dt v')
//   a' = F(r', u', v'),   b' = G(r')
//   rho_i = m_i sum_j W(d_ij),  p_i = P(rho_i)
//   F_i   = sum_j c_ij T_ij + g,   G_i = p_bg sum_j c_ij r_ij
//   T_ij  = -p_ij r_ij + eta_ij (u_i - u_j)
//           + [rho_i u_i (w_i . r_ij) + rho_j u_j (w_j . r_ij)] / 2,  w = v - u
//   c_ij  = (V_i^2 + V_j^2) / m_i * W'(d) / (d + eps),  V = m / rho
//   p_ij  = (rho_j p_i + rho_i p_j) / (rho_i + rho_j)
//
// (without tvf, v' = u', w = 0 and G is not evaluated).  Since m_i c_ij is
// symmetric and T_ij, r_ij antisymmetric, the cotangent-weighted sums read
//   <abar', F> + <bbar', G> = 1/2 sum_i sum_j S_ij,
//   S_ij = (V_i^2 + V_j^2) W'/(d + eps) [l_ij . T_ij + p_bg n_ij . r_ij],
//   l_ij = abar'_i / m_i - abar'_j / m_j,  n_ij = bbar'_i / m_i - bbar'_j / m_j,
// with S_ij = S_ji, so the derivative with respect to particle k's own
// quantities is sum_j dS_kj / dx_k: one gather sweep over k's pairs.
//
// Given the cotangents (rbar', ubar', vbar', abar', bbar') of state n+1 it
// returns those of state n: rbar' += J_r^T (abar', bbar'), ubar' += J_u^T abar',
// vbar' += J_v^T abar', vbar' += dt rbar', ubar' += vbar', then rbar = rbar',
// ubar = ubar', abar = dt ubar', bbar = c vbar'.  Both Jacobian products are
// gathers over the pair list of r' (two sweeps: the second needs the density
// cotangent of the neighbours), with no atomics, like the forward sweeps.
//
// Generalized wall boundary condition (solver.py:431-530; for free-slip walls
// alpha and beta below take only their normal parts):
// before the force, every wall particle i takes Shepard averages over its
// fluid neighbours j (w_ij = W(d_ij), S_i = sum_j w_ij + eps):
//   u^_i = 2 u_i - sum_j w_ij u_j / S_i   (v^ likewise)
//   p^_i = (sum_j w_ij p_j + g . sum_j rho_j w_ij r_ij) / S_i,  rho^_i = R(p^_i)
// (R the inverse EOS; fluid particles keep rho, as R(P(rho)) = rho).  The force
// transpose gives the cotangents of u^, v^ and rho^ (through p^ = P(rho^)); with
// alpha_i = ubar^_i / S_i, beta_i = vbar^_i / S_i, gamma_i = pbar^_i / S_i,
// pbar^_i = rhobar^_i / P'(rho^_i), the walls pass to their fluid neighbours
//   ubar_j -= w_ij alpha_i,  vbar_j -= w_ij beta_i,
//   pbar_j += w_ij gamma_i,  rhobar_j += w_ij (g . r_ij) gamma_i,
// keep ubar_i = 2 ubar^_i, and both sides take the position derivative of
// L_ij = -alpha_i . (u_j - uw_i) - beta_i . (v_j - vw_i)
//        + gamma_i (p_j + rho_j g . r_ij - p^_i)        (times dw_ij)
// plus the explicit r_ij of the pressure sum, gamma_i rho_j w_ij g.
#include <hip/hip_runtime.h>

#include "sph_device.h"
#include "sph_internal.h"
#include "sweep.h"

namespace sph {

constexpr int kAB = 256;

struct AdjArgs {
  SweepArgs s;               // records of state n+1 (P, U1 = (u, rho), Q1 = (p, eta, tag, .),
                             // V0 = (v, .))
  const float* rb_in;        // (N, d) cotangent of r'  (orig order)
  const float* ub_in;        // (N, d) cotangent of u'
  const float* ab_in;        // (N, d) cotangent of a' = dudt'
  const float* vb_in;        // (N, d) cotangent of v', or null (zero)
  const float* bb_in;        // (N, d) cotangent of b' = dvdt', or null (zero)
  const float* rhob_in;      // (N) continuity: cotangent of rho', or null (zero)
  float* rb_out;             // (N, d) cotangents of state n
  float* ub_out;
  float* ab_out;
  float* bb_out;             // with tvf: cotangent of b = dvdt (null: not written)
  float* rhob_out;           // (N) continuity: cotangent of rho (null: not written)
  float4* la;                // sorted cotangent of a'
  float4* lb;                // sorted cotangent of b' (tvf)
  float dt;
  float c_tvf;               // v' = u' + c_tvf b
  int bc_kind;               // case boundary hook of the step (sph_bc_spec::kind)
  const float4* rho_c;       // with renormalisation: records whose .w is the continuity
                             // density rho' (the sweeps then read the renormalised ones)
};

// Whether the case boundary hook (integrator.py:54; ldc.py:130-144, db.py:299-307,
// case_setup.py:393-402) sets particle k's u, v, dudt and dvdt to constants:
// their cotangents then flow nowhere.
__device__ __forceinline__ bool bc_holds(int kind, int tag) {
  if (kind == 1) return tag == kSolidWall || tag == kMovingWall;
  if (kind == 2) return tag == kSolidWall;
  if (kind == 3) return is_wall(tag);
  return false;
}

template <int DIM, bool TVF>
__global__ void k_adj_gather(AdjArgs a) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= a.s.n) return;
  const int i = a.s.perm[s];
  // the hook replaces the accelerations of the particles it holds
  const bool held = bc_holds(a.bc_kind, a.s.in.tag[i]);
  const V3 zero = {0.f, 0.f, 0.f};
  a.la[s] = f4(held ? zero : load_vec<DIM>(a.ab_in, i), 0.0f);
  if (TVF) a.lb[s] = f4(held || !a.bb_in ? zero : load_vec<DIM>(a.bb_in, i), 0.0f);
}

// Sweep A: density cotangent (times m, fluid only), the u' and v' cotangents
// of F and the direct r' cotangent of F and G, per sorted particle k.
// WALLS: the records after the generalized wall BC (U2, V2, Q2), and for wall
// particles the raw cotangent of rho^ in ab0.w (the walls sweeps consume it).
// CONT (density continuity): ab0.w = B, the whole cotangent of rho', the next
// step's included (for a wall with WALLS, that of rho^).
// AV: the artificial viscosity of fluid-fluid pairs (solver.py:404-428),
//   a_k += K m_j phi_kj (du . r_kj) r_kj,  phi = W'(d) / (d rho_ab (d^2 + 0.01 h^2)),
// K = alpha h c_ab, rho_ab = (rho_k + rho_j) / 2, du = u_k - u_j; its
// cotangent-weighted sum is 1/2 sum_kj K phi (du . r)(r . l^AV_kj) with
// l^AV_kj = m_j abar'_k - m_k abar'_j, symmetric in (k, j) like S_kj.
// The delta-SPH velocity diffusion (solver.py:259-311, solver == DELTA):
//   a_k += C m_j psi_kj (du . r_kj) r_kj for fluid j,  psi = W'(d) / (rho_k rho_j d^3),
// C = alpha h c_ref rho_ref; the same transpose with l = m_j [j fluid] abar'_k -
// m_k [k fluid] abar'_j, and d psi / d rho_k = -psi / rho_k.
template <int DIM, int KIND, bool TVF, bool WALLS, bool CONT, bool AV>
__global__ __launch_bounds__(kAB) void k_adj_forces(AdjArgs aa, float4* __restrict__ ab0,
                                                    float4* __restrict__ ab1,
                                                    float4* __restrict__ ab2) {
  constexpr int TPP = tpp_of(DIM);
  const SweepArgs& a = aa.s;
  const int tid = logical_block(a.remap) * blockDim.x + threadIdx.x;
  const int k = tid / TPP, h = tid % TPP;
  if (k >= a.n) return;  // whole lane groups leave together
  const sph_solver_params& p = a.prm;
  const KernelEval<KIND, DIM> K{p.kernel, p.sigma, p.one_over_h};
  const Eos eos = make_eos(p);
  const float4* __restrict__ RU = WALLS ? a.R.U2 : a.R.U1;
  const float4* __restrict__ RQ = WALLS ? a.R.Q2 : a.R.Q1;
  const float4* __restrict__ RV = WALLS ? a.R.V2 : a.R.V0;
  const float4 Pk = a.R.P[k], Uk = RU[k], Qk = RQ[k], Lk4 = aa.la[k];
  const float mk = Pk.w, rhok = Uk.w, pk = Qk.x, etak = Qk.y;
  const float Vk2 = (mk / rhok) * (mk / rhok);
  const float dPk = eos.dp_drho(rhok);
  const float inv_mk = 1.0f / mk;
  const V3 uk = v3(Uk), lk = v3(Lk4);
  // transport-velocity terms: w = v - u, the background pressure of G and the
  // sorted cotangent of b'
  const V3 wk = TVF ? v3(RV[k]) - uk : V3{0.f, 0.f, 0.f};
  const V3 nk = TVF ? v3(aa.lb[k]) : V3{0.f, 0.f, 0.f};
  const float pbg = TVF ? eos.p_fn(0.0f) : 0.0f;
  const bool fluid = (int)Qk.z == kFluid;
  const float avK = AV ? p.artificial_alpha * p.av_h * 10.0f : 0.0f;  // alpha h c_ab
  const float av_h2 = AV ? 0.01f * p.av_h * p.av_h : 0.0f;
  const float dK = p.solver == 2 ? p.diff_alpha * p.support * p.c_ref * p.rho_ref : 0.0f;
  float rhob = 0.0f;
  V3 ub = {0.f, 0.f, 0.f}, rb = {0.f, 0.f, 0.f}, vb = {0.f, 0.f, 0.f};
  for_each_pair(a, k, h, [&](const int j) {
    const float4 Pj = a.R.P[j];
    const V3 dr = min_image<DIM>(Pk, Pj, a.geo);
    const float d2 = dist2<DIM>(dr.x, dr.y, dr.z);
    if (d2 == 0.0f) return;  // self pair: every term and derivative vanishes
    if (!in_cutoff<DIM>(a, d2, Pk, Pj)) return;
    const float4 Uj = RU[j], Qj = RQ[j], Lj4 = aa.la[j];
    const float mj = Pj.w, rhoj = Uj.w, pj = Qj.x, etaj = Qj.y;
    // hardware rsqrt/rcp (about 1 ulp), as in the forward sweeps; 1/(d + eps)
    // is 1/d in float32 for every d above ~4e-9 (eps = 2.2e-16)
    const float inv = rsq_fast(d2);
    const float d = d2 * inv;
    const float q = d * p.one_over_h;
    const float w1 = K.dw(q);
    const float w2 = K.d2w(q);
    const float A0 = w1 * inv;                      // W'(d) / (d + eps)
    const float Ad0 = (w2 - w1 * inv) * inv;        // d/dd of the above
    const float inv_mj = rcp_fast(mj);
    const float Akj = A0 * inv_mk, Ajk = A0 * inv_mj;
    const float Adkj = Ad0 * inv_mk, Adjk = Ad0 * inv_mj;
    const float Vj = mj * rcp_fast(rhoj);
    const float Vj2 = Vj * Vj;
    const float Vs = Vk2 + Vj2;
    const float rs = rhok + rhoj;
    const float irs = rcp_fast(rs);
    const float pkj = (rhoj * pk + rhok * pj) * irs;
    const float etakj = 2.0f * etak * etaj * rcp_fast(etak + etaj + kEps);
    const V3 uj = v3(Uj);
    V3 T = (-pkj) * dr + etakj * (uk - uj);
    V3 wj = {0.f, 0.f, 0.f};
    if (TVF) {
      wj = v3(RV[j]) - uj;
      T = T + (0.5f * rhok * dot<DIM>(wk, dr)) * uk + (0.5f * rhoj * dot<DIM>(wj, dr)) * uj;
    }
    const float Dp = rhoj * (dPk * rs + (pj - pk)) * (irs * irs);  // d p_kj / d rho_k (via p_k)
    const V3 lj = v3(Lj4);
    const V3 lA = Akj * lk - Ajk * lj;   // W'/(d + eps) l_kj
    const V3 lc = Vs * lA;               // (V_k^2 + V_j^2) W'/(d + eps) l_kj
    const V3 lAd = Adkj * lk - Adjk * lj;
    rhob += (-2.0f * Vk2 / rhok) * dot<DIM>(lA, T) - Dp * dot<DIM>(lc, dr);
    ub = ub + etakj * lc;
    rb = rb + Vs * ((dot<DIM>(lAd, T) * inv) * dr - pkj * lA);
    if (TVF) {
      const V3 nj = v3(aa.lb[j]);
      const V3 nA = Akj * nk - Ajk * nj;     // W'/(d + eps) n_kj
      const V3 nAd = Adkj * nk - Adjk * nj;
      const float uk_lc = dot<DIM>(uk, lc), wk_r = dot<DIM>(wk, dr);
      // G: the volume factor and the pair geometry
      rhob += (-2.0f * Vk2 / rhok) * pbg * dot<DIM>(nA, dr);
      rb = rb + (Vs * pbg) * ((dot<DIM>(nAd, dr) * inv) * dr + nA);
      // the stress A_k = rho_k u_k w_k^T of T: through rho_k, u_k, v_k and r
      rhob += 0.5f * wk_r * uk_lc;
      ub = ub + (0.5f * rhok) * (wk_r * lc - uk_lc * dr);
      vb = vb + (0.5f * rhok * uk_lc) * dr;
      rb = rb + (0.5f * rhok * uk_lc) * wk + (0.5f * rhoj * dot<DIM>(uj, lc)) * wj;
    }
    if (AV && fluid && (int)Qj.z == kFluid) {
      const float D = d2 + av_h2;
      const float iD = rcp_fast(D);
      const float irab = 2.0f * irs;              // 1 / rho_ab
      const float Kphi = avK * A0 * irab * iD;    // K phi
      const V3 du = uk - uj;
      const float e = dot<DIM>(du, dr);
      const V3 l = mj * lk - mk * lj;
      const float sl = dot<DIM>(dr, l);
      // d phi / dd, over d
      const float dphi = (Ad0 * D - 2.0f * A0 * d) * irab * iD * iD * inv;
      ub = ub + (Kphi * sl) * dr;
      rhob -= Kphi * e * sl * irs;                // d phi / d rho_k = -phi / (rho_k + rho_j)
      rb = rb + (avK * dphi * e * sl) * dr + Kphi * (sl * du + e * l);
    }
    const bool fluid_j = (int)Qj.z == kFluid;
    if (dK != 0.0f && (fluid || fluid_j)) {
      const float ipp = rcp_fast(rhok * rhoj);
      const float psi = A0 * inv * inv * ipp;                      // W' / (rho_k rho_j d^3)
      const float dpsi = (Ad0 * d - 2.0f * A0) * (inv * inv) * (inv * inv) * ipp;  // psi' / d
      const V3 du = uk - uj;
      const float e = dot<DIM>(du, dr);
      const V3 l = (fluid_j ? mj : 0.0f) * lk - (fluid ? mk : 0.0f) * lj;
      const float sl = dot<DIM>(dr, l);
      const float Kpsi = dK * psi;
      ub = ub + (Kpsi * sl) * dr;
      rhob -= Kpsi * e * sl / rhok;
      rb = rb + (dK * dpsi * e * sl) * dr + Kpsi * (sl * du + e * l);
    }
  });
  rhob = lane_sum<TPP>(rhob);
  ub = lane_sum<TPP>(ub);
  rb = lane_sum<TPP>(rb);
  if (TVF) vb = lane_sum<TPP>(vb);
  if (h != 0) return;
  if (CONT) {
    ab0[k] = f4(ub, rhob + (aa.rhob_in ? aa.rhob_in[a.perm[k]] : 0.0f));
  } else {
    // rho = m sum W only for fluid (solver.py:796); a wall's rho^ comes from its p^
    ab0[k] = f4(ub, fluid ? mk * rhob : (WALLS ? rhob : 0.0f));
  }
  ab1[k] = f4(rb, 0.0f);
  if (TVF) ab2[k] = f4(vb, 0.0f);
}

// The Riemann limiter (solver.py:574-588) and its derivative in D = u_L - u_R.
__device__ __forceinline__ void rie_beta(const sph_solver_params& p, float D, float* beta,
                                         float* dbeta) {
  *beta = p.c_ref;
  *dbeta = 0.0f;
  if (p.eta_limiter != -1.0f) {
    const float t = p.eta_limiter * fmaxf(D, 0.0f);
    *beta = fminf(t, p.c_ref);
    *dbeta = (D > 0.0f && t < p.c_ref) ? p.eta_limiter : 0.0f;
  }
}

// Row term of fluid particle k and wall neighbour j in the Riemann
// acceleration (solver.py:345-388): the wall enters through mirrored states,
// u_L = -u_k . n_j, u_R = -u_L + 2 u_j . n_j, p_R = p_k - rho_k g_k . r_kj,
// rho_R = R(p_R), and the viscous velocity u_d = 2 u_j - ut_j (ut_j the wall's
// Shepard velocity with the wall condition, else u_j).  T = abar_k . f_kj with
//   f_kj = -c9 P* e + c6 (u_k - u_d),  c9 = 2 m_j W'(d) / (rho_k rho_j),
//   c6 = c9 eta_kj / d.
// Returns dT/du_k, dT/drho_k (p_k = P(rho_k) included), dT/dr_kj and c6 (the
// coefficient of ut_j).  The wall's own u, n, rho and m are constants here.
template <int DIM>
__device__ __forceinline__ void rie_wall_row(const sph_solver_params& p, const Eos& eos, V3 ak,
                                             V3 uk, float rhok, float pk, float etak, V3 gk,
                                             float mj, float rhoj, float etaj, V3 uj, V3 nj,
                                             V3 utj, V3 dr, float d, float inv, float w1,
                                             float w2, V3* dU, float* dRho, V3* dR,
                                             float* c6o, float* To) {
  const V3 e = inv * dr;
  const float ae = dot<DIM>(ak, e);
  const float gr = dot<DIM>(gk, dr);
  const float pR = pk - rhok * gr;
  const float rhoR = eos.rho_fn(pR);
  const float dPR = eos.dp_drho(rhoR);  // P'(rho_R); R'(p_R) = 1 / P'(rho_R)
  const float D = -2.0f * dot<DIM>(uk + uj, nj);
  float beta, dbeta;
  rie_beta(p, D, &beta, &dbeta);
  const float rho_avg = 0.5f * (rhok + rhoR);
  const float hDb = 0.5f * D * beta;
  const float Ps = 0.5f * (pk + pR) + rho_avg * hDb;
  const float PsD = 0.5f * rho_avg * (beta + D * dbeta);
  const float etakj = 2.0f * etak * etaj * rcp_fast(etak + etaj + kEps);
  const float irr = rcp_fast(rhok * rhoj);
  const float c9 = 2.0f * mj * w1 * irr;
  const float c6 = c9 * etakj * inv;
  const V3 vk = uk - (2.0f * uj - utj);
  const float av = dot<DIM>(ak, vk);
  const float T = -c9 * Ps * ae + c6 * av;
  *dU = (2.0f * c9 * ae * PsD) * nj + c6 * ak;
  const float dPs_dp = 1.0f + 0.5f * hDb / dPR;
  const float dPs_drho = -0.5f * gr + 0.5f * hDb * (1.0f - gr / dPR) + dPs_dp * eos.dp_drho(rhok);
  *dRho = -T / rhok - c9 * ae * dPs_drho;
  const float c9d = 2.0f * mj * w2 * irr;
  const float c6d = 2.0f * mj * etakj * irr * (w2 * inv - w1 * inv * inv);
  const V3 dPs_dr = (-0.5f * rhok * (1.0f + hDb / dPR)) * gk;
  *dR = (-c9d * Ps * ae + c6d * av) * e - (c9 * ae) * dPs_dr - (c9 * Ps * inv) * (ak - ae * e);
  *c6o = c6;
  *To = T;  // dT/drho_j = -T / rho_j (the wall's density, when it changes)
}

// Sweep A of the Riemann solver (solver.py:316-401, 574-588) with density
// summation and tvf = 0.  Fluid pairs: with e = r_kj / d, D = u_L - u_R =
// -(u_k - u_j) . e, the limiter beta = min(eta_lim max(D, 0), c_ref) (c_ref
// when eta_lim = -1), P* = (p_k + p_j)/2 + (rho_k + rho_j)/4 D beta and
// K = 2 m_k m_j W'(d) / (rho_k rho_j), the pair force times m_k is
// F_kj = K (-P* e + eta_kj du / d) (the stress term A vanishes: v = u there),
// antisymmetric, so as for the standard solver <abar', a'> = 1/2 sum_kj S_kj
// with S_kj = F_kj . l_kj symmetric, l_kj = abar'_k / m_k - abar'_j / m_j, and
// particle k's derivatives are sum_j dS_kj / dx_k.  Fluid-wall pairs: the
// fluid's row term (rie_wall_row); the walls, held by the case hook, carry no
// row of their own and gather the position derivative of their fluid
// neighbours' rows, and (with the wall condition) the cotangent of their
// Shepard velocity, passed on as wf = (ubar_t / S, S).  AV: as in
// k_adj_forces.
template <int DIM, int KIND, bool AV, bool CONT, bool TVF>
__global__ __launch_bounds__(kAB) void k_adj_forces_rie(AdjArgs aa, float4* __restrict__ ab0,
                                                        float4* __restrict__ ab1,
                                                        float4* __restrict__ wf) {
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
  const float dPk = eos.dp_drho(rhok);
  const float inv_mk = 1.0f / mk;
  const V3 uk = v3(Uk), lk = v3(Lk4);
  const bool fluid = (int)Qk.z == kFluid;
  const bool wall_k = is_wall((int)Qk.z);
  const V3 gk = g_ext_of<DIM>(a, a.perm[k]);
  const V3 nk = wall_k ? v3(a.R.NW[k]) : V3{0.f, 0.f, 0.f};
  // the wall's velocity in u_d: its Shepard velocity with the no-slip wall
  // condition, else its own (solver.py:378, 547-552); the records hold that
  const bool shepard = p.bc_trick && !p.free_slip;
  const V3 utk = wall_k ? (p.bc_trick ? v3(a.R.UT[k]) : uk) : V3{0.f, 0.f, 0.f};
  // transport velocity: the background-pressure acceleration G of
  // solver.py:199-213, the same pair sum as the standard solver's (the
  // Riemann stress term vanishes)
  const V3 bk = TVF ? v3(aa.lb[k]) : V3{0.f, 0.f, 0.f};
  const float pbg = TVF ? eos.p_fn(0.0f) : 0.0f;
  const float avK = AV ? p.artificial_alpha * p.av_h * 10.0f : 0.0f;
  const float av_h2 = AV ? 0.01f * p.av_h * p.av_h : 0.0f;
  float rhob = 0.0f, Sw = 0.0f;
  V3 ub = {0.f, 0.f, 0.f}, rb = {0.f, 0.f, 0.f}, utb = {0.f, 0.f, 0.f};
  for_each_pair(a, k, h, [&](const int j) {
    const float4 Pj = a.R.P[j];
    const V3 dr = min_image<DIM>(Pk, Pj, a.geo);
    const float d2 = dist2<DIM>(dr.x, dr.y, dr.z);
    if (d2 == 0.0f) return;
    if (!in_cutoff<DIM>(a, d2, Pk, Pj)) return;
    const float4 Uj = a.R.U1[j], Qj = a.R.Q1[j], Lj4 = aa.la[j];
    const int tag_j = (int)Qj.z;
    const float mj = Pj.w, rhoj = Uj.w, pj = Qj.x, etaj = Qj.y;
    const float inv = rsq_fast(d2);
    const float d = d2 * inv;
    const float q = d * p.one_over_h;
    const float w1 = K.dw(q), w2 = K.d2w(q);
    if (TVF) {
      const float Vk = mk / rhok, Vj = mj * rcp_fast(rhoj);
      const float Vk2 = Vk * Vk, Vs = Vk2 + Vj * Vj;
      const float A0 = w1 * inv, Ad0 = (w2 - w1 * inv) * inv;
      const float imj = rcp_fast(mj);
      const V3 bj = v3(aa.lb[j]);
      const V3 nA = (A0 * inv_mk) * bk - (A0 * imj) * bj;
      const V3 nAd = (Ad0 * inv_mk) * bk - (Ad0 * imj) * bj;
      rhob += (-2.0f * Vk2 / rhok) * pbg * dot<DIM>(nA, dr);
      rb = rb + (Vs * pbg) * ((dot<DIM>(nAd, dr) * inv) * dr + nA);
    }
    if (!fluid) {
      // a wall (held by the hook): the rows of its fluid neighbours
      if (!wall_k || tag_j != kFluid) return;
      V3 dU, dR;
      float dRho, c6, T;
      rie_wall_row<DIM>(p, eos, v3(Lj4), v3(Uj), rhoj, pj, etaj, g_ext_of<DIM>(a, a.perm[j]),
                        mk, rhok, etak, uk, nk, utk, (-1.0f) * dr, d, inv, w1, w2, &dU, &dRho,
                        &dR, &c6, &T);
      rb = rb - dR;  // the row depends on r_jk = r_j - r_k
      rhob -= T / rhok;  // and on the wall's density through 1 / rho_k
      if (shepard) {
        utb = utb + c6 * v3(Lj4);
        Sw += K.w(q);
      }
      return;
    }
    if (is_wall(tag_j)) {
      V3 dU, dR;
      float dRho, c6, T;
      const V3 utj = p.bc_trick ? v3(a.R.UT[j]) : v3(Uj);
      rie_wall_row<DIM>(p, eos, lk, uk, rhok, pk, etak, gk, mj, rhoj, etaj, v3(Uj),
                        v3(a.R.NW[j]), utj, dr, d, inv, w1, w2, &dU, &dRho, &dR, &c6, &T);
      ub = ub + dU;
      rhob += dRho;
      rb = rb + dR;
      return;
    }
    const V3 e = inv * dr;
    const V3 du = uk - v3(Uj);
    const float s_ = dot<DIM>(du, e);
    const float D = -s_;
    float beta, dbeta;
    rie_beta(p, D, &beta, &dbeta);
    const float rho_avg = 0.5f * (rhok + rhoj);
    const float Ps = 0.5f * (pk + pj) + 0.5f * rho_avg * D * beta;
    const float PsD = 0.5f * rho_avg * (beta + D * dbeta);   // d P* / d D
    const float etakj = 2.0f * etak * etaj * rcp_fast(etak + etaj + kEps);
    const V3 l = inv_mk * lk - rcp_fast(mj) * v3(Lj4);
    const float el = dot<DIM>(e, l), ul = dot<DIM>(du, l);
    const float k0 = 2.0f * mk * mj * rcp_fast(rhok * rhoj);
    const float Kw = k0 * w1;
    const float Q = -Ps * el + etakj * ul * inv;
    const float S = Kw * Q;
    ub = ub + Kw * ((el * PsD) * e + (etakj * inv) * l);
    rhob += -S / rhok - Kw * el * (0.5f * dPk + 0.25f * D * beta);
    const V3 dQ = ((el * PsD) * inv) * (du - s_ * e) - (Ps * inv) * (l - el * e) -
                  (etakj * ul * inv * inv) * e;
    rb = rb + (k0 * w2 * Q) * e + Kw * dQ;
    if (AV && tag_j == kFluid) {
      const V3 lkj = mj * lk - mk * v3(Lj4);
      const float A0 = w1 * inv, Ad0 = (w2 - w1 * inv) * inv;
      const float irs = rcp_fast(rhok + rhoj);
      const float Dd = d2 + av_h2;
      const float iD = rcp_fast(Dd);
      const float irab = 2.0f * irs;
      const float Kphi = avK * A0 * irab * iD;
      const float eu = dot<DIM>(du, dr);
      const float sl = dot<DIM>(dr, lkj);
      const float dphi = (Ad0 * Dd - 2.0f * A0 * d) * irab * iD * iD * inv;
      ub = ub + (Kphi * sl) * dr;
      rhob -= Kphi * eu * sl * irs;
      rb = rb + (avK * dphi * eu * sl) * dr + Kphi * (sl * du + eu * lkj);
    }
  });
  rhob = lane_sum<TPP>(rhob);
  ub = lane_sum<TPP>(ub);
  rb = lane_sum<TPP>(rb);
  utb = lane_sum<TPP>(utb);
  Sw = lane_sum<TPP>(Sw);
  if (h != 0) return;
  if (CONT)   // the whole cotangent of rho' (the next step's included)
    ab0[k] = f4(ub, rhob + (aa.rhob_in ? aa.rhob_in[a.perm[k]] : 0.0f));
  else
    ab0[k] = f4(ub, fluid ? mk * rhob : 0.0f);
  ab1[k] = f4(rb, 0.0f);
  if (wall_k) wf[k] = f4((1.0f / (Sw + kEps)) * utb, Sw);
}

// Row of fluid particle k and wall neighbour j in the Riemann continuity
// (solver.py:108-164): with mirrored states the bracket of
// R_kj = (rho_k m_j / rho_j) W'(d) [s + rho_k G / (rho_avg c)], s = (u_k - u_j) . e,
// G = g_k . r_kj, rho_avg = (rho_k + R(p_k - rho_k G)) / 2.  Returns dR/du_k
// (a multiple of e: *cu), dR/drho_k, dR/dp_k and dR/dr_kj.
template <int DIM>
__device__ __forceinline__ void rie_cont_wall_row(const sph_solver_params& p, const Eos& eos,
                                                  V3 uk, float rhok, float pk, V3 gk, float mj,
                                                  float rhoj, V3 uj, V3 dr, float inv, float w1,
                                                  float w2, float* cu, float* dRho, float* dP,
                                                  V3* dR, float* Ro) {
  const V3 e = inv * dr;
  const V3 du = uk - uj;
  const float s_ = dot<DIM>(du, e);
  const float G = dot<DIM>(gk, dr);
  const float rhoR = eos.rho_fn(pk - rhok * G);
  const float dPR = eos.dp_drho(rhoR);
  const float ra = 0.5f * (rhok + rhoR);
  const float ic = 1.0f / (ra * p.c_ref);
  const float Y = s_ + rhok * G * ic;
  const float al = rhok * mj * rcp_fast(rhoj);
  const float R = al * w1 * Y;
  const float t2 = rhok * G * ic / ra;  // rho_k G / (rho_avg^2 c)
  *cu = al * w1;
  *dRho = R / rhok + al * w1 * (G * ic - t2 * 0.5f * (1.0f - G / dPR));
  *dP = -al * w1 * t2 * 0.5f / dPR;
  *dR = (al * w2 * Y) * e +
        (al * w1) * (inv * (du - s_ * e) + (rhok * ic + t2 * 0.5f * rhok / dPR) * gk);
  *Ro = R;  // dR/drho_j = -R / rho_j (the wall's density, when it changes)
}

// Sweep C of the Riemann continuity (solver.py:108-164, 772-793): the
// transpose of rho'_k = rho_k + dt drhodt_k (fluid k), with
//   drhodt_k = sum_j (rho_k m_j / rho_j) W'(d) [s + (p_k - p_j) / (rho_avg c)]
// for fluid j (s = (u'_k - u'_j) . e, rho_avg = (rho_k + rho_j) / 2, u' the
// kicked velocity, rho and p those of the state the step starts from) and the
// mirrored row of rie_cont_wall_row for wall j.  B = ab0.w (the cotangent of
// rho').  Particle k gathers its own rows (times B_k) and its fluid
// neighbours' rows (times B_j); p = P(rho) of the state turns the pressure
// cotangent into one of rho.  Walls (rho' = rho) pass B on and gather the
// position derivative of their fluid neighbours' rows.  Then the kinematic
// transpose of sweep B.
template <int DIM, int KIND>
__global__ __launch_bounds__(kAB) void k_adj_rie_continuity(AdjArgs aa,
                                                            const float4* __restrict__ ab0,
                                                            const float4* __restrict__ ab1) {
  constexpr int TPP = tpp_of(DIM);
  const SweepArgs& a = aa.s;
  const int tid = logical_block(a.remap) * blockDim.x + threadIdx.x;
  const int k = tid / TPP, h = tid % TPP;
  if (k >= a.n) return;
  const sph_solver_params& p = a.prm;
  const KernelEval<KIND, DIM> K{p.kernel, p.sigma, p.one_over_h};
  const Eos eos = make_eos(p);
  const float4 Pk = a.R.P[k], Uk = a.R.U0[k], Qk = a.R.Q0[k];
  const float mk = Pk.w, rhok = Uk.w, pk = Qk.x;
  const V3 uk = v3(Uk);
  const bool fluid = (int)Qk.z == kFluid, wall_k = is_wall((int)Qk.z);
  const V3 gk = g_ext_of<DIM>(a, a.perm[k]);
  const float Bk = ab0[k].w;
  float rho_c = 0.f, p_c = 0.f;
  V3 ub = {0.f, 0.f, 0.f}, rb = {0.f, 0.f, 0.f};
  for_each_pair(a, k, h, [&](const int j) {
    const float4 Pj = a.R.P[j];
    const V3 dr = min_image<DIM>(Pk, Pj, a.geo);
    const float d2 = dist2<DIM>(dr.x, dr.y, dr.z);
    if (d2 == 0.0f) return;
    if (!in_cutoff<DIM>(a, d2, Pk, Pj)) return;
    const float4 Uj = a.R.U0[j], Qj = a.R.Q0[j];
    const int tag_j = (int)Qj.z;
    const float inv = rsq_fast(d2);
    const float q = d2 * inv * p.one_over_h;
    const float w1 = K.dw(q), w2 = K.d2w(q);
    const float Bj = ab0[j].w;
    if (!fluid) {
      if (!wall_k || tag_j != kFluid) return;
      float cu, dRho, dP, R;
      V3 dR;
      rie_cont_wall_row<DIM>(p, eos, v3(Uj), Uj.w, Qj.x, g_ext_of<DIM>(a, a.perm[j]), mk,
                             rhok, uk, (-1.0f) * dr, inv, w1, w2, &cu, &dRho, &dP, &dR, &R);
      rb = rb - Bj * dR;
      rho_c -= Bj * R / rhok;
      return;
    }
    if (is_wall(tag_j)) {
      float cu, dRho, dP, R;
      V3 dR;
      rie_cont_wall_row<DIM>(p, eos, uk, rhok, pk, gk, Pj.w, Uj.w, v3(Uj), dr, inv, w1, w2,
                             &cu, &dRho, &dP, &dR, &R);
      ub = ub + (Bk * cu * inv) * dr;
      rho_c += Bk * dRho;
      p_c += Bk * dP;
      rb = rb + Bk * dR;
      return;
    }
    if (tag_j != kFluid) return;
    const V3 e = inv * dr;
    const V3 du = uk - v3(Uj);
    const float s_ = dot<DIM>(du, e);
    const float dp_ = pk - Qj.x;
    const float ra = 0.5f * (rhok + Uj.w);
    const float ic = 1.0f / (ra * p.c_ref);
    const float X = s_ + dp_ * ic, Xn = s_ - dp_ * ic;
    const float al = rhok * Pj.w * rcp_fast(Uj.w);     // own row rho_k m_j / rho_j
    const float be = Uj.w * mk * rcp_fast(rhok);       // neighbour's row rho_j m_k / rho_k
    const float ca = Bk * al, cb = Bj * be;
    ub = ub + ((ca + cb) * w1) * e;
    rho_c += ca * w1 * X / rhok - cb * w1 * Xn / rhok +
             (cb - ca) * w1 * dp_ * ic / (2.0f * ra);
    p_c += (ca - cb) * w1 * ic;
    rb = rb + (w2 * (ca * X + cb * Xn)) * e + ((ca + cb) * w1 * inv) * (du - s_ * e);
  });
  rho_c = lane_sum<TPP>(rho_c);
  p_c = lane_sum<TPP>(p_c);
  ub = lane_sum<TPP>(ub);
  rb = lane_sum<TPP>(rb);
  if (h != 0) return;
  const float dt = aa.dt;
  const int i = a.perm[k];
  const float4 A0r = ab0[k], A1 = ab1[k];
  const V3 rbt = load_vec<DIM>(aa.rb_in, i) + v3(A1) + dt * rb;
  V3 vbt = dt * rbt;  // r' = r + dt v'
  if (aa.vb_in) vbt = vbt + load_vec<DIM>(aa.vb_in, i);
  V3 ubt = load_vec<DIM>(aa.ub_in, i) + v3(A0r) + dt * ub + vbt;
  if (bc_holds(aa.bc_kind, a.in.tag[i])) {
    ubt = V3{0.f, 0.f, 0.f};
    vbt = ubt;
  }
  store_vec<DIM>(aa.rb_out, i, rbt);
  store_vec<DIM>(aa.ub_out, i, ubt);
  store_vec<DIM>(aa.ab_out, i, dt * ubt);  // u' = u + dt a
  if (aa.bb_out) store_vec<DIM>(aa.bb_out, i, aa.c_tvf * vbt);
  // (a wall's density is constant unless renormalised; its cotangent is kept either way)
  if (aa.rhob_out)
    aa.rhob_out[i] = Bk + dt * (fluid ? rho_c + eos.dp_drho(rhok) * p_c : rho_c);
}

// Sweep W of the Riemann solver's wall condition (solver.py:547-552): the
// walls' Shepard velocity ut_j = sum_k w_jk u_k / (S_j + eps) over fluid k;
// with alpha_j = ubar_t_j / (S_j + eps) (wf), fluid k gathers
// ubar_k += w_jk alpha_j, and both sides take the position derivative
// alpha_j . (u_k - ut_j) dw_jk.
template <int DIM, int KIND>
__global__ __launch_bounds__(kAB) void k_adj_rie_walls(AdjArgs aa, float4* __restrict__ ab0,
                                                       float4* __restrict__ ab1,
                                                       const float4* __restrict__ wf) {
  constexpr int TPP = tpp_of(DIM);
  const SweepArgs& a = aa.s;
  const int tid = logical_block(a.remap) * blockDim.x + threadIdx.x;
  const int k = tid / TPP, h = tid % TPP;
  if (k >= a.n) return;
  const sph_solver_params& p = a.prm;
  const KernelEval<KIND, DIM> K{p.kernel, p.sigma, p.one_over_h};
  const int tag_k = (int)a.R.Q1[k].z;
  const bool fluid_k = tag_k == kFluid;
  if (!fluid_k && !is_wall(tag_k)) return;
  const float4 Pk = a.R.P[k];
  const V3 uk = v3(a.R.U1[k]);
  const V3 alk = fluid_k ? V3{0.f, 0.f, 0.f} : v3(wf[k]);
  const V3 utk = fluid_k ? V3{0.f, 0.f, 0.f} : v3(a.R.UT[k]);
  V3 ub = {0.f, 0.f, 0.f}, rb = {0.f, 0.f, 0.f};
  for_each_pair(a, k, h, [&](const int j) {
    const int tag_j = (int)a.R.Q1[j].z;
    if (fluid_k ? !is_wall(tag_j) : tag_j != kFluid) return;
    const float4 Pj = a.R.P[j];
    const V3 dr = min_image<DIM>(Pk, Pj, a.geo);
    const float d2 = dist2<DIM>(dr.x, dr.y, dr.z);
    if (d2 == 0.0f) return;
    if (!in_cutoff<DIM>(a, d2, Pk, Pj)) return;
    const float inv = rsq_fast(d2);
    const float q = d2 * inv * p.one_over_h;
    const float w = K.w(q), wd = K.dw(q);
    if (fluid_k) {
      const V3 al = v3(wf[j]);
      ub = ub + w * al;
      rb = rb + (dot<DIM>(al, uk - v3(a.R.UT[j])) * wd * inv) * dr;
    } else {
      rb = rb + (dot<DIM>(alk, v3(a.R.U1[j]) - utk) * wd * inv) * dr;
    }
  });
  rb = lane_sum<TPP>(rb);
  if (fluid_k) ub = lane_sum<TPP>(ub);
  if (h != 0) return;
  const float4 B1 = ab1[k];
  ab1[k] = f4(v3(B1) + rb, B1.w);
  if (fluid_k) {
    const float4 B0 = ab0[k];
    ab0[k] = f4(v3(B0) + ub, B0.w);
  }
}

// Sweep W1, wall particles: their Shepard sums recomputed as the forward
// walls sweep takes them (fluid neighbours in the cutoff), and their
// cotangents turned into the factors the fluid side gathers:
// wf0 = (alpha, gamma), wf1 = (beta, p^), wf2 = uw, wf3 = vw.  The wall's own
// u (and v) then carry 2 ubar^ (2 vbar^), and its density cotangent is spent.
template <int DIM, int KIND, bool TVF>
__global__ __launch_bounds__(kAB) void k_adj_walls_prep(AdjArgs aa, float4* __restrict__ ab0,
                                                        float4* __restrict__ ab2,
                                                        float4* __restrict__ wf0,
                                                        float4* __restrict__ wf1,
                                                        float4* __restrict__ wf2,
                                                        float4* __restrict__ wf3) {
  constexpr int TPP = tpp_of(DIM);
  const SweepArgs& a = aa.s;
  const int tid = logical_block(a.remap) * blockDim.x + threadIdx.x;
  const int k = tid / TPP, h = tid % TPP;
  if (k >= a.n) return;
  if (!is_wall((int)a.R.Q1[k].z)) return;  // a particle's lanes leave together
  const sph_solver_params& p = a.prm;
  const KernelEval<KIND, DIM> K{p.kernel, p.sigma, p.one_over_h};
  const Eos eos = make_eos(p);
  const float4 Pk = a.R.P[k];
  float S = 0.f, Cp = 0.f;
  V3 A = {0.f, 0.f, 0.f}, B = {0.f, 0.f, 0.f}, Cr = {0.f, 0.f, 0.f};
  for_each_pair(a, k, h, [&](const int j) {
    const float4 Qj = a.R.Q1[j];
    if ((int)Qj.z != kFluid) return;
    const float4 Pj = a.R.P[j];
    const V3 dr = min_image<DIM>(Pk, Pj, a.geo);
    const float d2 = dist2<DIM>(dr.x, dr.y, dr.z);
    if (!in_cutoff<DIM>(a, d2, Pk, Pj)) return;
    const float w = K.w(sqrt_fast(d2) * p.one_over_h);
    const float4 Uj = a.R.U1[j];
    S += w;
    A = A + w * v3(Uj);
    if (TVF) B = B + w * v3(a.R.V0[j]);
    Cp += w * Qj.x;
    Cr = Cr + (Uj.w * w) * dr;
  });
  S = lane_sum<TPP>(S);
  A = lane_sum<TPP>(A);
  if (TVF) B = lane_sum<TPP>(B);
  Cp = lane_sum<TPP>(Cp);
  Cr = lane_sum<TPP>(Cr);
  if (h != 0) return;
  const float inv = 1.0f / (S + kEps);
  const V3 g = g_ext_of<DIM>(a, a.perm[k]);
  const float phat = (Cp + dot<DIM>(g, Cr)) * inv;
  const float4 A0 = ab0[k];
  const float pbar = A0.w / eos.dp_drho(a.R.U2[k].w);  // rho^ = R(p^): R' = 1 / P'(rho^)
  const V3 vbh = TVF ? v3(ab2[k]) : V3{0.f, 0.f, 0.f};
  // free-slip walls take only the normal part of the average (u^ = 2 u - n n^T
  // ubar_w with n the wall normal, solver.py:505-510): the fluid side sees the
  // projected cotangents
  V3 ubp = v3(A0), vbp = vbh;
  if (p.free_slip) {
    const V3 nn = v3(a.R.NW[k]);
    ubp = dot<DIM>(nn, ubp) * nn;
    vbp = dot<DIM>(nn, vbp) * nn;
  }
  wf0[k] = f4(inv * ubp, inv * pbar);
  wf1[k] = f4(inv * vbp, phat);
  wf2[k] = f4(inv * A, 0.0f);
  wf3[k] = f4(inv * B, 0.0f);
  ab0[k] = f4(2.0f * v3(A0), 0.0f);
  if (TVF) ab2[k] = f4(2.0f * vbh, 0.0f);
}

// Sweep W2: fluid particles gather the cotangents their wall neighbours pass
// them (u, v, p -> rho, rho, and r); wall particles gather the derivative of
// their averages with respect to their own position.  Each particle updates
// only its own ab0/ab1/ab2 entries.
// CONT: the fluid's density cotangent is that of rho' itself (no m_k of the
// summation's chain).
template <int DIM, int KIND, bool TVF, bool CONT>
__global__ __launch_bounds__(kAB) void k_adj_walls(AdjArgs aa, float4* __restrict__ ab0,
                                                   float4* __restrict__ ab1,
                                                   float4* __restrict__ ab2,
                                                   const float4* __restrict__ wf0,
                                                   const float4* __restrict__ wf1,
                                                   const float4* __restrict__ wf2,
                                                   const float4* __restrict__ wf3) {
  constexpr int TPP = tpp_of(DIM);
  const SweepArgs& a = aa.s;
  const int tid = logical_block(a.remap) * blockDim.x + threadIdx.x;
  const int k = tid / TPP, h = tid % TPP;
  if (k >= a.n) return;
  const sph_solver_params& p = a.prm;
  const KernelEval<KIND, DIM> K{p.kernel, p.sigma, p.one_over_h};
  const Eos eos = make_eos(p);
  const int tag_k = (int)a.R.Q1[k].z;
  const bool fluid_k = tag_k == kFluid;
  if (!fluid_k && !is_wall(tag_k)) return;
  const float4 Pk = a.R.P[k], Uk = a.R.U1[k];
  const V3 g = g_ext_of<DIM>(a, a.perm[k]);
  const V3 uk = v3(Uk), vk = TVF ? v3(a.R.V0[k]) : V3{0.f, 0.f, 0.f};
  const float pk = a.R.Q1[k].x, rhok = Uk.w;
  // a wall's own factors (fluid particles use their neighbours')
  const float4 F0 = fluid_k ? make_float4(0.f, 0.f, 0.f, 0.f) : wf0[k];
  const float4 F1 = fluid_k ? make_float4(0.f, 0.f, 0.f, 0.f) : wf1[k];
  const V3 uwk = fluid_k ? V3{0.f, 0.f, 0.f} : v3(wf2[k]);
  const V3 vwk = fluid_k || !TVF ? V3{0.f, 0.f, 0.f} : v3(wf3[k]);
  V3 ub = {0.f, 0.f, 0.f}, vb = {0.f, 0.f, 0.f}, rb = {0.f, 0.f, 0.f};
  float pacc = 0.f, racc = 0.f;
  for_each_pair(a, k, h, [&](const int j) {
    const float4 Qj = a.R.Q1[j];
    const int tag_j = (int)Qj.z;
    // fluid-wall pairs only, seen from either side
    if (fluid_k ? !is_wall(tag_j) : tag_j != kFluid) return;
    const float4 Pj = a.R.P[j];
    const V3 dr = min_image<DIM>(Pk, Pj, a.geo);  // r_k - r_j
    const float d2 = dist2<DIM>(dr.x, dr.y, dr.z);
    if (d2 == 0.0f) return;
    if (!in_cutoff<DIM>(a, d2, Pk, Pj)) return;
    const float inv_d = rsq_fast(d2);
    const float q = d2 * inv_d * p.one_over_h;
    const float w = K.w(q), wd = K.dw(q);
    if (fluid_k) {
      // wall neighbour j: r_jk = -dr; its pressure extrapolation uses its own g
      const float4 G0 = wf0[j], G1 = wf1[j];
      const V3 al = v3(G0), be = v3(G1);
      const V3 gj = g_ext_of<DIM>(a, a.perm[j]);
      const float ga = G0.w, gr = -dot<DIM>(gj, dr);
      ub = ub - w * al;
      if (TVF) vb = vb - w * be;
      pacc += w * ga;
      racc += w * gr * ga;
      float L = -dot<DIM>(al, uk - v3(wf2[j])) + ga * (pk + rhok * gr - G1.w);
      if (TVF) L -= dot<DIM>(be, vk - v3(wf3[j]));
      rb = rb + (L * wd * inv_d) * dr - (ga * rhok * w) * gj;
    } else {
      // fluid neighbour j: r_kj = dr
      const float4 Uj = a.R.U1[j];
      float L = -dot<DIM>(v3(F0), v3(Uj) - uwk) + F0.w * (Qj.x + Uj.w * dot<DIM>(g, dr) - F1.w);
      if (TVF) L -= dot<DIM>(v3(F1), v3(a.R.V0[j]) - vwk);
      rb = rb + (L * wd * inv_d) * dr + (F0.w * Uj.w * w) * g;
    }
  });
  rb = lane_sum<TPP>(rb);
  if (fluid_k) {
    ub = lane_sum<TPP>(ub);
    if (TVF) vb = lane_sum<TPP>(vb);
    pacc = lane_sum<TPP>(pacc);
    racc = lane_sum<TPP>(racc);
  }
  if (h != 0) return;
  float4 B1 = ab1[k];
  ab1[k] = f4(v3(B1) + rb, B1.w);
  if (fluid_k) {
    const float4 B0 = ab0[k];
    const float mk = CONT ? 1.0f : a.R.P[k].w;
    ab0[k] = f4(v3(B0) + ub, B0.w + mk * (eos.dp_drho(rhok) * pacc + racc));
    if (TVF) {
      const float4 B2 = ab2[k];
      ab2[k] = f4(v3(B2) + vb, B2.w);
    }
  }
}

// Sweep C (density continuity, solver.py:24-30, 750-756): the transpose of
//   rho'_k = rho_k + dt rho_k sum_j V_j (u'_k - u'_j) . g_kj,
// V_j = m_j / rho_j, g_kj = W'(d) r_kj / (d + eps), with rho the step's input
// density and u' the kicked velocity (the U0 records).  With B = ab0.w the
// cotangent of rho', c_kj = B_k rho_k V_j + B_j rho_j V_k is symmetric, so
//   ubar'_k += dt sum_j c_kj g_kj,
//   rbar'_k += dt sum_j c_kj H_kj (u'_k - u'_j),  H = d g / d r_k,
//   rhobar_k = B_k (1 + dt sum_j V_j (u'_k - u'_j) . g_kj)
//              - dt m_k / rho_k^2 sum_j B_j rho_j (u'_k - u'_j) . g_kj,
// then the kinematic transpose as in sweep B.
template <int DIM, int KIND, bool TVF>
__global__ __launch_bounds__(kAB) void k_adj_continuity(AdjArgs aa,
                                                        const float4* __restrict__ ab0,
                                                        const float4* __restrict__ ab1,
                                                        const float4* __restrict__ ab2) {
  constexpr int TPP = tpp_of(DIM);
  const SweepArgs& a = aa.s;
  const int tid = logical_block(a.remap) * blockDim.x + threadIdx.x;
  const int k = tid / TPP, h = tid % TPP;
  if (k >= a.n) return;
  const sph_solver_params& p = a.prm;
  const KernelEval<KIND, DIM> K{p.kernel, p.sigma, p.one_over_h};
  const float4 Pk = a.R.P[k], Uk = a.R.U0[k];
  const float mk = Pk.w, rhok = Uk.w, Vk = mk / rhok;
  const V3 uk = v3(Uk);
  const float Bk = ab0[k].w;
  float accD = 0.f, accR = 0.f;
  V3 ub = {0.f, 0.f, 0.f}, rb = {0.f, 0.f, 0.f};
  for_each_pair(a, k, h, [&](const int j) {
    const float4 Pj = a.R.P[j];
    const V3 dr = min_image<DIM>(Pk, Pj, a.geo);
    const float d2 = dist2<DIM>(dr.x, dr.y, dr.z);
    if (d2 == 0.0f) return;
    if (!in_cutoff<DIM>(a, d2, Pk, Pj)) return;
    const float4 Uj = a.R.U0[j];
    const float inv = rsq_fast(d2);
    const float d = d2 * inv;
    const float q = d * p.one_over_h;
    const float w1 = K.dw(q), w2 = K.d2w(q);
    const float A0 = w1 * inv;                // W'(d) / (d + eps)
    const float Ad0 = (w2 - w1 * inv) * inv;  // d/dd of the above
    const V3 du = uk - v3(Uj);
    const float Vj = Pj.w * rcp_fast(Uj.w);
    const float Bj = ab0[j].w;
    const float rdu = dot<DIM>(dr, du);
    const float dg = A0 * rdu;                // (u'_k - u'_j) . g_kj
    const float c = Bk * rhok * Vj + Bj * Uj.w * Vk;
    accD += Vj * dg;
    accR += Bj * Uj.w * dg;
    ub = ub + (c * A0) * dr;
    rb = rb + c * (A0 * du + (Ad0 * rdu * inv) * dr);
  });
  accD = lane_sum<TPP>(accD);
  accR = lane_sum<TPP>(accR);
  ub = lane_sum<TPP>(ub);
  rb = lane_sum<TPP>(rb);
  if (h != 0) return;
  const float dt = aa.dt;
  const int i = a.perm[k];
  const float4 A0r = ab0[k], A1 = ab1[k];
  const V3 rbt = load_vec<DIM>(aa.rb_in, i) + v3(A1) + dt * rb;
  V3 vbt = dt * rbt;  // r' = r + dt v'
  if (aa.vb_in) vbt = vbt + load_vec<DIM>(aa.vb_in, i);
  if (TVF) vbt = vbt + v3(ab2[k]);
  V3 ubt = load_vec<DIM>(aa.ub_in, i) + v3(A0r) + dt * ub + vbt;  // v' = u' + c b
  if (bc_holds(aa.bc_kind, a.in.tag[i])) {
    ubt = V3{0.f, 0.f, 0.f};
    vbt = ubt;
  }
  store_vec<DIM>(aa.rb_out, i, rbt);
  store_vec<DIM>(aa.ub_out, i, ubt);
  store_vec<DIM>(aa.ab_out, i, aa.dt * ubt);  // u' = u + dt a
  if (aa.bb_out) store_vec<DIM>(aa.bb_out, i, aa.c_tvf * vbt);
  if (aa.rhob_out) aa.rhob_out[i] = Bk * (1.0f + dt * accD) - dt * (mk / (rhok * rhok)) * accR;
}

// Sweeps N1/N2: the transpose of the density renormalisation after continuity
// (solver.py:167-173), rho~_k = N_k / min(D_k, 1) with N_k = sum_j m_j W_kj and
// D_k = sum_j (m_j / rho'_j) W_kj over all pairs (the self pair included).
// Given B~ = ab0.w, the cotangent of rho~, N1 stores a_k = B~_k / min(D_k, 1)
// and b_k = -B~_k N_k / D_k^2 (0 where D_k > 1 is clamped); N2 gathers
//   rhobar'_k = -(m_k / rho'_k^2) sum_j b_j W_jk  (into ab0.w: continuity's B),
//   rbar_k   += sum_j (a_k m_j + a_j m_k + b_k V_j + b_j V_k) W'(d) r_kj / d,
// V = m / rho'.
template <int DIM, int KIND>
__global__ __launch_bounds__(kAB) void k_adj_renorm_prep(AdjArgs aa,
                                                         const float4* __restrict__ ab0,
                                                         float4* __restrict__ nb) {
  constexpr int TPP = tpp_of(DIM);
  const SweepArgs& a = aa.s;
  const int tid = logical_block(a.remap) * blockDim.x + threadIdx.x;
  const int k = tid / TPP, h = tid % TPP;
  if (k >= a.n) return;
  const sph_solver_params& p = a.prm;
  const KernelEval<KIND, DIM> K{p.kernel, p.sigma, p.one_over_h};
  const float4 Pk = a.R.P[k];
  float nom = 0.f, den = 0.f;
  for_each_pair(a, k, h, [&](const int j) {
    const float4 Pj = a.R.P[j];
    const V3 dr = min_image<DIM>(Pk, Pj, a.geo);
    const float d2 = dist2<DIM>(dr.x, dr.y, dr.z);
    if (!in_cutoff<DIM>(a, d2, Pk, Pj)) return;
    const float w = K.w(sqrt_fast(d2) * p.one_over_h);
    nom += Pj.w * w;
    den += (Pj.w * rcp_fast(aa.rho_c[j].w)) * w;
  });
  nom = lane_sum<TPP>(nom);
  den = lane_sum<TPP>(den);
  if (h != 0) return;
  const float Bt = ab0[k].w;
  const float Dc = den > 1.0f ? 1.0f : den;
  nb[k] = make_float4(Bt / Dc, den > 1.0f ? 0.0f : -Bt * nom / (den * den), 0.0f, 0.0f);
}

template <int DIM, int KIND>
__global__ __launch_bounds__(kAB) void k_adj_renorm(AdjArgs aa, float4* __restrict__ ab0,
                                                    float4* __restrict__ ab1,
                                                    const float4* __restrict__ nb) {
  constexpr int TPP = tpp_of(DIM);
  const SweepArgs& a = aa.s;
  const int tid = logical_block(a.remap) * blockDim.x + threadIdx.x;
  const int k = tid / TPP, h = tid % TPP;
  if (k >= a.n) return;
  const sph_solver_params& p = a.prm;
  const KernelEval<KIND, DIM> K{p.kernel, p.sigma, p.one_over_h};
  const float4 Pk = a.R.P[k];
  const float mk = Pk.w, rk = aa.rho_c[k].w, Vk = mk / rk;
  const float4 Nk = nb[k];
  float bw = 0.f;
  V3 rb = {0.f, 0.f, 0.f};
  for_each_pair(a, k, h, [&](const int j) {
    const float4 Pj = a.R.P[j];
    const V3 dr = min_image<DIM>(Pk, Pj, a.geo);
    const float d2 = dist2<DIM>(dr.x, dr.y, dr.z);
    if (!in_cutoff<DIM>(a, d2, Pk, Pj)) return;
    const float4 Nj = nb[j];
    if (d2 == 0.0f) {  // the self pair: W(0), no position derivative
      bw += Nj.y * K.w(0.0f);
      return;
    }
    const float inv = rsq_fast(d2);
    const float q = d2 * inv * p.one_over_h;
    bw += Nj.y * K.w(q);
    const float mj = Pj.w, Vj = mj * rcp_fast(aa.rho_c[j].w);
    const float c = Nk.x * mj + Nj.x * mk + Nk.y * Vj + Nj.y * Vk;
    rb = rb + (c * K.dw(q) * inv) * dr;
  });
  bw = lane_sum<TPP>(bw);
  rb = lane_sum<TPP>(rb);
  if (h != 0) return;
  const float4 A0 = ab0[k], A1 = ab1[k];
  ab0[k] = f4(v3(A0), -(mk / (rk * rk)) * bw);
  ab1[k] = f4(v3(A1) + rb, A1.w);
}

// Sweep B: the r' cotangent through the density sums, then the kinematic
// transpose of the kick, transport velocity and drift, written in original
// order.
template <int DIM, int KIND, bool TVF>
__global__ __launch_bounds__(kAB) void k_adj_density(AdjArgs aa, const float4* __restrict__ ab0,
                                                     const float4* __restrict__ ab1,
                                                     const float4* __restrict__ ab2) {
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
    const float inv_d = rsq_fast(d2);
    const float w1 = K.dw(d2 * inv_d * p.one_over_h);
    rb = rb + ((Bk + ab0[j].w) * w1 * inv_d) * dr;
  });
  rb = lane_sum<TPP>(rb);
  if (h != 0) return;
  const int i = a.perm[k];
  const float4 A0 = ab0[k], A1 = ab1[k];
  const V3 rbt = load_vec<DIM>(aa.rb_in, i) + v3(A1) + rb;
  V3 vbt = aa.dt * rbt;  // r' = r + dt v'
  if (aa.vb_in) vbt = vbt + load_vec<DIM>(aa.vb_in, i);
  if (TVF) vbt = vbt + v3(ab2[k]);
  V3 ubt = load_vec<DIM>(aa.ub_in, i) + v3(A0) + vbt;  // v' = u' + c b
  if (bc_holds(aa.bc_kind, a.in.tag[i])) {
    // held by the case hook: u, v, dudt and dvdt of this particle are constants
    // in every step, so only its position carries a cotangent
    ubt = V3{0.f, 0.f, 0.f};
    vbt = ubt;
  }
  store_vec<DIM>(aa.rb_out, i, rbt);
  store_vec<DIM>(aa.ub_out, i, ubt);
  store_vec<DIM>(aa.ab_out, i, aa.dt * ubt);  // u' = u + dt a
  if (aa.bb_out) store_vec<DIM>(aa.bb_out, i, aa.c_tvf * vbt);
}

// The renormalisation's transpose (sweeps N1, N2), between the force and
// wall transposes and the continuity transpose.
template <int DIM, int KIND>
int launch_adj_renorm(sph_ws* ws, AdjArgs& a, dim3 grid, hipStream_t st) {
  SPH_LAUNCH(ws, "adj_renorm_prep", st,
             (k_adj_renorm_prep<DIM, KIND><<<grid, kAB, 0, st>>>(a, ws->AB0, ws->DG1)));
  SPH_LAUNCH(ws, "adj_renorm", st,
             (k_adj_renorm<DIM, KIND><<<grid, kAB, 0, st>>>(a, ws->AB0, ws->AB1, ws->DG1)));
  return SPH_OK;
}

// The Riemann solver with density summation and tvf = 0: the density sweep
// (its records, and the walls' Shepard velocity with the wall condition),
// the force transpose, the Shepard velocity's transpose, then sweep B.
template <int DIM, int KIND, bool AV, bool CONT, bool TVF>
int launch_adjoint_rie(sph_ws* ws, AdjArgs& a, bool records_ready, hipStream_t st) {
  const int n = a.s.n;
  dim3 grid(xcd_grid(cdivl((int64_t)n * tpp_of(DIM), kAB), a.s.remap));
  if (!records_ready)
    SPH_LAUNCH(ws, "density", st,
               (k_density<DIM, KIND, CONT ? 2 : 0><<<grid, kAB, 0, st>>>(a.s)));
  SPH_LAUNCH(ws, "adj_gather", st,
             (k_adj_gather<DIM, TVF><<<(unsigned)cdivl(n, kAB), kAB, 0, st>>>(a)));
  SPH_LAUNCH(ws, "adj_forces", st,
             (k_adj_forces_rie<DIM, KIND, AV, CONT, TVF><<<grid, kAB, 0, st>>>(
                 a, ws->AB0, ws->AB1, ws->DL0)));
  if (a.s.prm.bc_trick && !a.s.prm.free_slip)
    SPH_LAUNCH(ws, "adj_rie_walls", st,
               (k_adj_rie_walls<DIM, KIND><<<grid, kAB, 0, st>>>(a, ws->AB0, ws->AB1, ws->DL0)));
  if (CONT) {
    if (a.s.prm.rho_renorm) {
      const int rc = launch_adj_renorm<DIM, KIND>(ws, a, grid, st);
      if (rc) return rc;
    }
    SPH_LAUNCH(ws, "adj_continuity", st,
               (k_adj_rie_continuity<DIM, KIND><<<grid, kAB, 0, st>>>(a, ws->AB0, ws->AB1)));
    return SPH_OK;
  }
  SPH_LAUNCH(ws, "adj_density", st,
             (k_adj_density<DIM, KIND, false><<<grid, kAB, 0, st>>>(a, ws->AB0, ws->AB1,
                                                                     ws->AB2)));
  return SPH_OK;
}

template <int DIM, int KIND, bool TVF, bool WALLS, bool CONT, bool AV>
int launch_adjoint(sph_ws* ws, AdjArgs& a, bool records_ready, hipStream_t st) {
  if (a.s.prm.solver == 1)
    return launch_adjoint_rie<DIM, KIND, AV, CONT, TVF>(ws, a, records_ready, st);
  const int n = a.s.n;
  dim3 grid(xcd_grid(cdivl((int64_t)n * tpp_of(DIM), kAB), a.s.remap));
  if (!records_ready) {
    SPH_LAUNCH(ws, "density", st,
               (k_density<DIM, KIND, CONT ? 1 : 0><<<grid, kAB, 0, st>>>(a.s)));
    if (WALLS) SPH_LAUNCH(ws, "walls", st, (k_walls<DIM, KIND, 0><<<grid, kAB, 0, st>>>(a.s)));
  }
  SPH_LAUNCH(ws, "adj_gather", st,
             (k_adj_gather<DIM, TVF><<<(unsigned)cdivl(n, kAB), kAB, 0, st>>>(a)));
  SPH_LAUNCH(ws, "adj_forces", st,
             (k_adj_forces<DIM, KIND, TVF, WALLS, CONT, AV><<<grid, kAB, 0, st>>>(
                 a, ws->AB0, ws->AB1, ws->AB2)));
  if (WALLS) {
    // the per-wall factors go to the delta-SPH continuity's buffers (that
    // continuity is not differentiated, so they are free)
    SPH_LAUNCH(ws, "adj_walls_prep", st,
               (k_adj_walls_prep<DIM, KIND, TVF><<<grid, kAB, 0, st>>>(
                   a, ws->AB0, ws->AB2, ws->DL0, ws->DL1, ws->DL2, ws->DG0)));
    SPH_LAUNCH(ws, "adj_walls", st,
               (k_adj_walls<DIM, KIND, TVF, CONT><<<grid, kAB, 0, st>>>(
                   a, ws->AB0, ws->AB1, ws->AB2, ws->DL0, ws->DL1, ws->DL2, ws->DG0)));
  }
  if (CONT) {
    // the walls' rho^ cotangent is spent above (their rho' is overwritten by rho^)
    if (a.s.prm.rho_renorm) {
      const int rc = launch_adj_renorm<DIM, KIND>(ws, a, grid, st);
      if (rc) return rc;
    }
    SPH_LAUNCH(ws, "adj_continuity", st,
               (k_adj_continuity<DIM, KIND, TVF><<<grid, kAB, 0, st>>>(a, ws->AB0, ws->AB1,
                                                                        ws->AB2)));
    return SPH_OK;
  }
  SPH_LAUNCH(ws, "adj_density", st,
             (k_adj_density<DIM, KIND, TVF><<<grid, kAB, 0, st>>>(a, ws->AB0, ws->AB1, ws->AB2)));
  return SPH_OK;
}

template <int DIM, int KIND, bool TVF, bool AV>
int launch_adjoint_kind(sph_ws* ws, AdjArgs& a, bool walls, bool cont, bool records_ready,
                        hipStream_t st) {
  if (cont)
    return walls ? launch_adjoint<DIM, KIND, TVF, true, true, AV>(ws, a, records_ready, st)
                 : launch_adjoint<DIM, KIND, TVF, false, true, AV>(ws, a, records_ready, st);
  return walls ? launch_adjoint<DIM, KIND, TVF, true, false, AV>(ws, a, records_ready, st)
               : launch_adjoint<DIM, KIND, TVF, false, false, AV>(ws, a, records_ready, st);
}

template <int DIM, int KIND>
int launch_adjoint_tvf(sph_ws* ws, AdjArgs& a, bool tvf, bool walls, bool cont, bool av,
                       bool records_ready, hipStream_t st) {
  if (av)
    return tvf ? launch_adjoint_kind<DIM, KIND, true, true>(ws, a, walls, cont, records_ready, st)
               : launch_adjoint_kind<DIM, KIND, false, true>(ws, a, walls, cont, records_ready,
                                                             st);
  return tvf ? launch_adjoint_kind<DIM, KIND, true, false>(ws, a, walls, cont, records_ready, st)
             : launch_adjoint_kind<DIM, KIND, false, false>(ws, a, walls, cont, records_ready,
                                                            st);
}

}  // namespace sph

using namespace sph;

extern "C" int sph_step_vjp(sph_ws* ws, int bank, const sph_solver_params* prm, float dt,
                            float tvf_coef, const sph_bc_spec* bc, const sph_state_in* next,
                            const sph_cotangent_in* bar_next, sph_cotangent_out* bar,
                            int32_t records_ready, hipStream_t st) {
  if (!prm || !next || !bar_next || !bar) { set_error("sph_step_vjp: null argument"); return SPH_ERR_ARG; }
  int rc = check_bank(ws, bank, prm, "sph_step_vjp");
  if (rc) return rc;
  if (!next->r || !next->u || !next->mass || !next->eta || !next->tag || !next->rho || !next->p) {
    set_error("sph_step_vjp: state n+1 needs r, u, rho, p, mass, eta, tag");
    return SPH_ERR_ARG;
  }
  if (!bar_next->r || !bar_next->u || !bar_next->a || !bar->r || !bar->u || !bar->a) {
    set_error("sph_step_vjp: the r, u and a cotangent arrays are required");
    return SPH_ERR_ARG;
  }
  if (prm->solver != 0 && prm->solver != 1 && prm->solver != 2) {
    set_error("sph_step_vjp: unknown solver");
    return SPH_ERR_ARG;
  }
  if (prm->solver == 2 && prm->density_evolution) {
    set_error("sph_step_vjp: delta-SPH is differentiated with density summation only "
              "(not its density diffusion)");
    return SPH_ERR_ARG;
  }
  if (prm->bc_trick && prm->free_slip && !next->nw) {
    set_error("sph_step_vjp: free-slip walls need the wall normals nw");
    return SPH_ERR_ARG;
  }
  if (prm->kernel != SPH_KERNEL_QSK && prm->kernel != SPH_KERNEL_WC2K) {
    set_error("sph_step_vjp: kernel must be QSK or WC2K");
    return SPH_ERR_ARG;
  }
  const bool tvf = prm->tvf_stress != 0;
  if (tvf && !next->v) {
    set_error("sph_step_vjp: the transport-velocity step needs v of state n+1");
    return SPH_ERR_ARG;
  }
  if (!tvf && tvf_coef != 0.0f) {
    set_error("sph_step_vjp: tvf_coef != 0 needs tvf_stress (the transport-velocity step)");
    return SPH_ERR_ARG;
  }
  sph_state_in in = *next;
  if (!tvf) in.v = next->u;   // tvf = 0: v == u
  // a position-dependent body force (next->g_ext, at next->r) enters only the
  // walls' pressure extrapolation; its own derivative is the caller's
  sph_state_out none{};
  AdjArgs a;
  a.s = make_args(ws, bank, prm, &in, &none);
  a.rb_in = bar_next->r; a.ub_in = bar_next->u; a.ab_in = bar_next->a;
  a.vb_in = bar_next->v; a.bb_in = bar_next->b; a.rhob_in = bar_next->rho;
  a.rb_out = bar->r; a.ub_out = bar->u; a.ab_out = bar->a; a.bb_out = bar->b;
  a.rhob_out = bar->rho;
  a.la = ws->LA;
  a.lb = ws->LB;
  a.rho_c = ws->U1;
  if (prm->density_evolution && prm->rho_renorm) {
    // the sweeps read the renormalised density and pressure (as the forward
    // force and wall sweeps do); rho_c keeps the continuity density
    a.s.R.U1 = ws->U3;
    a.s.R.Q1 = ws->Q3;
  }
  a.dt = dt;
  a.c_tvf = tvf_coef;
  a.bc_kind = bc ? bc->kind : 0;
  if (a.bc_kind < 0 || a.bc_kind > 3) {
    set_error("sph_step_vjp: unknown case boundary hook");
    return SPH_ERR_ARG;
  }
  const int n = a.s.n;
  const bool ready = records_ready != 0;
  if (prm->density_evolution && !ready) {
    // continuity advances the density of state n, which `next` does not hold
    set_error("sph_step_vjp: density continuity needs the step's own records "
              "(records_ready: call right after the sph_step that produced `next`)");
    return SPH_ERR_ARG;
  }
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
  const bool walls = prm->bc_trick != 0, cont = prm->density_evolution != 0;
  const bool av = prm->artificial_alpha != 0.0f;
  if (prm->dim == 2)
    return prm->kernel == SPH_KERNEL_QSK
               ? launch_adjoint_tvf<2, SPH_KERNEL_QSK>(ws, a, tvf, walls, cont, av, ready, st)
               : launch_adjoint_tvf<2, SPH_KERNEL_WC2K>(ws, a, tvf, walls, cont, av, ready, st);
  return prm->kernel == SPH_KERNEL_QSK
             ? launch_adjoint_tvf<3, SPH_KERNEL_QSK>(ws, a, tvf, walls, cont, av, ready, st)
             : launch_adjoint_tvf<3, SPH_KERNEL_WC2K>(ws, a, tvf, walls, cont, av, ready, st);
}
