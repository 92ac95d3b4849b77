// Device-side building blocks of the SPH step: periodic geometry, the exact
// neighbour predicate, SPH kernels and equations of state.  gfx950 only.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/sph.h"

namespace sph {

// solver.py:21 uses jnp.finfo(float).eps (the float64 epsilon) as a weak
// scalar, i.e. rounded to float32 in float32 arithmetic.
constexpr float kEps = 2.220446049250313e-16f;

// -------------------------------------------------------------- geometry
// jnp.mod(x, side) for x in (-side/2, 3*side/2): fmod is exact there, so the
// two branches below are bit-identical to fmod + sign fix (jax_md/space.py:181).
__device__ __forceinline__ float mod_near(float t, float side) {
#pragma clang fp contract(off)
  float m = t;
  if (t >= side) m = t - side;        // exact (Sterbenz)
  if (t < 0.0f) m = t + side;          // rounded, as jnp.mod
  return m;
}

// General jnp.mod for arbitrary finite x (position shift with large steps).
__device__ __forceinline__ float jmod(float x, float side) {
  float r = fmodf(x, side);
  if (r != 0.0f && ((r < 0.0f) != (side < 0.0f))) r = r + side;
  return r;
}

// periodic_displacement (jax_md/space.py:170-181): mod(dR + side/2, side) - side/2
// with dR = fl(a - b), evaluated exactly like the reference in float32.
__device__ __forceinline__ float disp_exact(float a, float b, float side, float half) {
#pragma clang fp contract(off)
  float d = a - b;
  float t = d + half;
  return mod_near(t, side) - half;
}

// Exact reference predicate metric_sq(r_sender, r_receiver) < cutoff^2
// (jax_md/partition.py:895-897): squares summed x, then y, then z, no FMA.
template <int DIM>
__device__ __forceinline__ float metric_sq_exact(const float4 s, const float4 r, const float* box,
                                                 const float* half) {
#pragma clang fp contract(off)
  float dx = disp_exact(s.x, r.x, box[0], half[0]);
  float dy = disp_exact(s.y, r.y, box[1], half[1]);
  float acc = dx * dx;
  acc = acc + dy * dy;
  if (DIM == 3) {
    float dz = disp_exact(s.z, r.z, box[2], half[2]);
    acc = acc + dz * dz;
  }
  return acc;
}

// Minimum-image displacement for the physics sweeps: d - k*side (one
// rounding) with d = a - b and k = rint(d/side).  For every pair inside the
// cutoff this is the same image as the exact formula, within about 1 ulp(side)
// of it; the sweeps' cutoff test decides pairs that close to the cutoff with
// the exact formula, so pair sets do not depend on this rounding.
__device__ __forceinline__ float disp_fast(float a, float b, float side, float inv_side) {
  const float d = a - b;
  return __builtin_fmaf(-rintf(d * inv_side), side, d);
}

// Hardware reciprocal, square root and reciprocal square root (v_rcp_f32,
// v_sqrt_f32, v_rsq_f32, about 1 ulp) for
// the pair sums, whose results are held to the reference at float32 rounding
// noise; the compiler's IEEE forms cost 5-10 instructions each.  The
// neighbour predicate never divides or takes roots, so pair sets stay exact.
__device__ __forceinline__ float rcp_fast(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float sqrt_fast(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float rsq_fast(float x) { return __builtin_amdgcn_rsqf(x); }

// x^2 + y^2 (+ z^2) with explicit fused multiply-adds (same bits everywhere).
template <int DIM>
__device__ __forceinline__ float dist2(float dx, float dy, float dz) {
  float d2 = fmaf(dy, dy, dx * dx);
  if (DIM == 3) d2 = fmaf(dz, dz, d2);
  return d2;
}

// -------------------------------------------------------------- kernels
// W(r) and dW/dr of kernel.py:26-201; q = r * fl32(1/h).  dW/dr is the closed
// form of jax.grad(w) (identical away from the knots, where both are 0 or equal).
template <int KIND, int DIM>
struct Kern;

__device__ __forceinline__ float p4(float x) { float x2 = x * x; return x2 * x2; }
__device__ __forceinline__ float p5(float x) { return p4(x) * x; }

template <int DIM>
struct Kern<SPH_KERNEL_QSK, DIM> {
  __device__ __forceinline__ static float w(float q, float s) {
    float q1 = fmaxf(0.0f, 1.0f - q), q2 = fmaxf(0.0f, 2.0f - q), q3 = fmaxf(0.0f, 3.0f - q);
    return s * (p5(q3) - 6.0f * p5(q2) + 15.0f * p5(q1));
  }
  // returns dW/dr given c = 5*sigma/h (precomputed)
  __device__ __forceinline__ static float dw(float q, float s, float ih) {
    float q1 = fmaxf(0.0f, 1.0f - q), q2 = fmaxf(0.0f, 2.0f - q), q3 = fmaxf(0.0f, 3.0f - q);
    return -(5.0f * s * ih) * (p4(q3) - 6.0f * p4(q2) + 15.0f * p4(q1));
  }
  // d^2W/dr^2 = 20 sigma / h^2 [(3-q)^3 - 6 (2-q)^3 + 15 (1-q)^3]_+ (for the adjoint)
  __device__ __forceinline__ static float d2w(float q, float s, float ih) {
    float q1 = fmaxf(0.0f, 1.0f - q), q2 = fmaxf(0.0f, 2.0f - q), q3 = fmaxf(0.0f, 3.0f - q);
    return (20.0f * s * ih * ih) * (q3 * q3 * q3 - 6.0f * q2 * q2 * q2 + 15.0f * q1 * q1 * q1);
  }
  static constexpr float kSupport = 3.0f;
};

template <int DIM>
struct Kern<SPH_KERNEL_WC2K, DIM> {
  __device__ __forceinline__ static float w(float q, float s) {
    float q1 = fmaxf(0.0f, 1.0f - 0.5f * q);
    if (DIM == 1) return s * (q1 * q1 * q1 * (1.5f * q + 1.0f));
    return s * (p4(q1) * (2.0f * q + 1.0f));
  }
  __device__ __forceinline__ static float dw(float q, float s, float ih) {
    float q1 = fmaxf(0.0f, 1.0f - 0.5f * q);
    if (DIM == 1) return s * ih * (-3.0f * q * q1 * q1);
    return -(5.0f * s * ih) * q * (q1 * q1 * q1);
  }
  // d^2W/dr^2 (dim >= 2) = -5 sigma / h^2 (1 - q/2)_+^2 (1 - 2q) (for the adjoint)
  __device__ __forceinline__ static float d2w(float q, float s, float ih) {
    float q1 = fmaxf(0.0f, 1.0f - 0.5f * q);
    return -(5.0f * s * ih * ih) * (q1 * q1) * (1.0f - 2.0f * q);
  }
  static constexpr float kSupport = 2.0f;
};

template <int DIM>
struct Kern<SPH_KERNEL_CSK, DIM> {
  __device__ __forceinline__ static float w(float q, float s) {
    float c1 = (1.0f - q >= 0.0f) ? 1.0f : 0.0f;
    float t = 2.0f - q;
    float c2 = (t < 1.0f && t >= 0.0f) ? 1.0f : 0.0f;
    float a = 1.0f - 1.5f * q * q * (1.0f - q / 2.0f);
    float b = 0.25f * t * t * t;
    return s * (a * c1 + b * c2);
  }
  __device__ __forceinline__ static float dw(float q, float s, float ih) {
    float c1 = (1.0f - q >= 0.0f) ? 1.0f : 0.0f;
    float t = 2.0f - q;
    float c2 = (t < 1.0f && t >= 0.0f) ? 1.0f : 0.0f;
    float d1 = -3.0f * q + 2.25f * q * q;
    float d2 = -0.75f * t * t;
    return s * ih * (d1 * c1 + d2 * c2);
  }
  static constexpr float kSupport = 2.0f;
};

template <int DIM>
struct Kern<SPH_KERNEL_WC4K, DIM> {
  __device__ __forceinline__ static float w(float q, float s) {
    float q1 = fmaxf(0.0f, 1.0f - 0.5f * q);
    if (DIM == 1) return s * (p5(q1) * (2.0f * q * q + 2.5f * q + 1.0f));
    float q16 = p4(q1) * q1 * q1;
    return s * (q16 * ((35.0f / 12.0f) * q * q + 3.0f * q + 1.0f));
  }
  __device__ __forceinline__ static float dw(float q, float s, float ih) {
    float q1 = fmaxf(0.0f, 1.0f - 0.5f * q);
    if (DIM == 1) {
      float p = 2.0f * q * q + 2.5f * q + 1.0f, dp = 4.0f * q + 2.5f;
      return s * ih * (-2.5f * p4(q1) * p + p5(q1) * dp);
    }
    float p = (35.0f / 12.0f) * q * q + 3.0f * q + 1.0f, dp = (35.0f / 6.0f) * q + 3.0f;
    return s * ih * (-3.0f * p5(q1) * p + p5(q1) * q1 * dp);
  }
  static constexpr float kSupport = 2.0f;
};

template <int DIM>
struct Kern<SPH_KERNEL_WC6K, DIM> {
  __device__ __forceinline__ static float w(float q, float s) {
    float q1 = fmaxf(0.0f, 1.0f - 0.5f * q);
    if (DIM == 1) {
      float q17 = p5(q1) * q1 * q1;
      return s * (q17 * ((21.0f / 8.0f) * q * q * q + (19.0f / 4.0f) * q * q + 3.5f * q + 1.0f));
    }
    float q18 = p4(q1) * p4(q1);
    return s * (q18 * (4.0f * q * q * q + 6.25f * q * q + 4.0f * q + 1.0f));
  }
  __device__ __forceinline__ static float dw(float q, float s, float ih) {
    float q1 = fmaxf(0.0f, 1.0f - 0.5f * q);
    if (DIM == 1) {
      float p = (21.0f / 8.0f) * q * q * q + (19.0f / 4.0f) * q * q + 3.5f * q + 1.0f;
      float dp = (63.0f / 8.0f) * q * q + 9.5f * q + 3.5f;
      float q16 = p5(q1) * q1;
      return s * ih * (-3.5f * q16 * p + q16 * q1 * dp);
    }
    float p = 4.0f * q * q * q + 6.25f * q * q + 4.0f * q + 1.0f;
    float dp = 12.0f * q * q + 12.5f * q + 4.0f;
    float q17 = p4(q1) * q1 * q1 * q1;
    return s * ih * (-4.0f * q17 * p + q17 * q1 * dp);
  }
  static constexpr float kSupport = 2.0f;
};

template <int DIM>
struct Kern<SPH_KERNEL_GK, DIM> {
  __device__ __forceinline__ static float w(float q, float s) {
    float c = (3.0f - q >= 0.0f) ? 1.0f : 0.0f;
    return s * c * expf(-(q * q));
  }
  __device__ __forceinline__ static float dw(float q, float s, float ih) {
    float c = (3.0f - q >= 0.0f) ? 1.0f : 0.0f;
    return s * ih * c * (-2.0f * q) * expf(-(q * q));
  }
  static constexpr float kSupport = 3.0f;
};

template <int DIM>
struct Kern<SPH_KERNEL_SGK, DIM> {
  __device__ __forceinline__ static float w(float q, float s) {
    float c = (3.0f - q >= 0.0f) ? 1.0f : 0.0f;
    return s * c * expf(-(q * q)) * ((float)DIM / 2.0f + 1.0f - q * q);
  }
  __device__ __forceinline__ static float dw(float q, float s, float ih) {
    float c = (3.0f - q >= 0.0f) ? 1.0f : 0.0f;
    float e = expf(-(q * q));
    float a = (float)DIM / 2.0f + 1.0f;
    return s * ih * c * e * (-2.0f * q * (a - q * q) - 2.0f * q);
  }
  static constexpr float kSupport = 3.0f;
};

// ---------------------------------------------------------------- EOS
// eos.py:20-57.  Tait with gamma == 1 skips pow (pow(x, 1) == x exactly).
struct Eos {
  int kind;
  float p_ref, rho_ref, p_bg, gamma, inv_gamma, rk;
  __device__ __forceinline__ float p_fn(float rho) const {
    if (kind == 1) return rk * (rho - rho_ref) + p_bg;
    float x = rho / rho_ref;
    if (gamma != 1.0f) x = powf(x, gamma);
    return p_ref * (x - 1.0f) + p_bg;
  }
  // dp/drho (for the adjoint)
  __device__ __forceinline__ float dp_drho(float rho) const {
    if (kind == 1) return rk;
    if (gamma == 1.0f) return p_ref / rho_ref;
    return p_ref * gamma * powf(rho / rho_ref, gamma - 1.0f) / rho_ref;
  }
  __device__ __forceinline__ float rho_fn(float p) const {
    if (kind == 1) return (p - p_bg) / rk + rho_ref;
    float t = p + p_ref - p_bg;
    float x = t / p_ref;
    if (gamma != 1.0f) x = powf(x, inv_gamma);
    return rho_ref * x;
  }
};

__device__ __forceinline__ Eos make_eos(const sph_solver_params& p) {
  Eos e;
  e.kind = p.eos_kind;
  e.p_ref = p.p_ref; e.rho_ref = p.rho_ref; e.p_bg = p.p_bg;
  e.gamma = p.gamma; e.inv_gamma = p.inv_gamma; e.rk = p.riemann_k;
  return e;
}

// ------------------------------------------------------------ XCD mapping
// MI355X dispatches the workgroups of a launch to its 8 XCDs round-robin by
// workgroup id, and each XCD has its own L2.  With remap on (and a grid that
// is a multiple of 8), XCD x runs the contiguous logical blocks
// [x * per, (x + 1) * per): cell-sorted particles that share neighbours then
// share one L2 instead of being spread over all eight.
constexpr int kXcds = 8;
__device__ __forceinline__ int logical_block(int remap) {
  const int b = blockIdx.x;
  if (!remap) return b;
  const int per = gridDim.x / kXcds;
  return (b % kXcds) * per + b / kXcds;
}
__host__ __device__ inline unsigned xcd_grid(long long blocks, int remap) {
  return (unsigned)(remap ? (blocks + kXcds - 1) / kXcds * kXcds : blocks);
}

// Batched launches over the members of an ensemble: workgroup b of the launch
// serves block `blk` of member `sim` (sim < 0 pads the grid; it exits at once).
struct EnsBlock {
  int32_t sim, blk;
};

// tags (utils.py:22-32)
constexpr int kFluid = 0, kSolidWall = 1, kMovingWall = 2, kDirichletWall = 3;
__device__ __forceinline__ bool is_wall(int t) { return t >= 1 && t <= 3; }

}  // namespace sph
