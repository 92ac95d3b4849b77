/*
 * CPU restatement of one jax-sph si_euler step, for the CPU baseline of the
 * benchmark: TEST INFRASTRUCTURE ONLY (the benchmark's cpu_baseline leg and
 * tests/ load it; the product never does).
 *
 * Configuration: standard WCSPH with density summation, Tait equation of
 * state, Quintic or Wendland-C2 kernel, optional transport-velocity
 * formulation, constant body force, fully periodic box, no walls.  That is the
 * 3D Taylor-Green benchmark and the 2D Taylor-Green case.  Follows:
 *   integrator.py:26-30   kick u += dt dudt, v = u + tvf dt/2 dvdt, drift
 *                         r = mod(r + dt v, box)
 *   jax_md/partition.py:294-413, 832-909   cells of fl32(box / floor(box /
 *                         cutoff)), the 3^d neighbouring cells, predicate
 *                         mod(dR + side/2, side) - side/2, d^2 < cutoff^2
 *   solver.py:176-178     rho_i = m_i sum_j W(d_ij) (self pair included)
 *   eos.py:20-38          Tait p = p_ref ((rho / rho_ref)^gamma - 1) + p_bg
 *   solver.py:199-256     standard + transport-velocity acceleration and the
 *                         background-pressure dvdt (p_bg of the TVF is
 *                         eos.p_fn(0), solver.py:802)
 *   kernel.py:51-103      W and dW/dr of the Quintic and Wendland-C2 kernels
 * in float32, one OpenMP thread per group of particles: each particle gathers
 * its own sums (no atomics), so results do not depend on the thread count.
 * The summation order inside a particle differs from the numpy oracle's
 * segment sums, so the two agree to float32 rounding.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct {
  int32_t dim, n;
  float box[3];
  float cutoff;        /* kernel support: 3h (QSK) or 2h (WC2K) */
  int32_t kernel;      /* 1 = QSK, 2 = WC2K */
  float one_over_h;    /* fl32(1 / h) */
  float sigma;         /* fl32(kernel normalisation) */
  float p_ref, rho_ref, p_bg, gamma;
  float dt;
  float ctvf;          /* fl32(fl32(tvf / 2) dt): v = u + ctvf dvdt */
  float g[3];
  int32_t threads;     /* <= 0: all available */
} sph_cpu_params;

static const float kEpsF = 2.220446049250313e-16f; /* solver.py:21, as float32 */

static inline float kern_w(int kind, float q, float s) {
  if (kind == 1) {
    const float q1 = fmaxf(0.0f, 1.0f - q), q2 = fmaxf(0.0f, 2.0f - q), q3 = fmaxf(0.0f, 3.0f - q);
    const float a = q1 * q1 * q1 * q1 * q1, b = q2 * q2 * q2 * q2 * q2, c = q3 * q3 * q3 * q3 * q3;
    return s * (c - 6.0f * b + 15.0f * a);
  }
  const float q1 = fmaxf(0.0f, 1.0f - 0.5f * q);
  return s * (q1 * q1 * q1 * q1 * (2.0f * q + 1.0f));
}

static inline float kern_dw(int kind, float q, float s, float ih) {
  if (kind == 1) {
    const float q1 = fmaxf(0.0f, 1.0f - q), q2 = fmaxf(0.0f, 2.0f - q), q3 = fmaxf(0.0f, 3.0f - q);
    const float a = q1 * q1 * q1 * q1, b = q2 * q2 * q2 * q2, c = q3 * q3 * q3 * q3;
    return -(5.0f * s * ih) * (c - 6.0f * b + 15.0f * a);
  }
  const float q1 = fmaxf(0.0f, 1.0f - 0.5f * q);
  return -(5.0f * s * ih) * q * q1 * q1 * q1;
}

/* jnp.mod(x, side) */
static inline float jmodf(float x, float side) {
  float r = fmodf(x, side);
  if (r != 0.0f && ((r < 0.0f) != (side < 0.0f))) r += side;
  return r;
}

/* periodic_displacement: mod(a - b + side/2, side) - side/2.  Both positions
 * lie in [0, side], so the argument of mod lies in (-side/2, 3 side/2), where
 * one conditional subtraction or addition is jnp.mod exactly. */
static inline float pdisp(float a, float b, float side) {
  const float half = side * 0.5f;
  float t = (a - b) + half;
  if (t >= side) t = t - side;
  else if (t < 0.0f) t = t + side;
  return t - half;
}

/* Cell list of the step: particles sorted by cell (counting sort). */
typedef struct {
  int cps[3];
  float cs[3];
  int ncell;
  int* start;   /* ncell + 1 */
  int* order;   /* n: original index of each sorted slot */
  int* cell;    /* n: cell of each original particle */
} cells_t;

static void cells_build(const sph_cpu_params* p, const float* r, cells_t* c) {
  const int n = p->n, dim = p->dim;
  c->ncell = 1;
  for (int d = 0; d < 3; ++d) {
    if (d < dim) {
      c->cps[d] = (int)floorf(p->box[d] / p->cutoff);
      c->cs[d] = p->box[d] / (float)c->cps[d];
    } else {
      c->cps[d] = 1;
      c->cs[d] = 1.0f;
    }
    c->ncell *= c->cps[d];
  }
  int* count = (int*)calloc((size_t)c->ncell + 1, sizeof(int));
  for (int i = 0; i < n; ++i) {
    int h = 0, stride = 1;
    for (int d = 0; d < dim; ++d) {
      int q = (int)(r[i * dim + d] / c->cs[d]);
      q = q < 0 ? 0 : (q >= c->cps[d] ? c->cps[d] - 1 : q);
      h += q * stride;
      stride *= c->cps[d];
    }
    c->cell[i] = h;
    count[h + 1]++;
  }
  for (int k = 0; k < c->ncell; ++k) count[k + 1] += count[k];
  memcpy(c->start, count, sizeof(int) * ((size_t)c->ncell + 1));
  for (int i = 0; i < n; ++i) c->order[count[c->cell[i]]++] = i;
  free(count);
}

/* Runs the statements BODY (the trailing arguments) with j, dr[] and d2 set for
 * every j with d^2 < cutoff^2 in the 3^d cells around particle i (self
 * included); a macro so that the loop body inlines. */
#define FOR_NEIGHBOURS(P, C, R, I, ...)                                              \
  do {                                                                             \
    const int dim_ = (P)->dim;                                                     \
    const float c2_ = (P)->cutoff * (P)->cutoff;                                   \
    const int hi_ = (C)->cell[I];                                                  \
    int ci_[3] = {hi_ % (C)->cps[0], (hi_ / (C)->cps[0]) % (C)->cps[1],            \
                  hi_ / ((C)->cps[0] * (C)->cps[1])};                              \
    const int zr_ = dim_ == 3 ? 1 : 0;                                             \
    for (int oz_ = -zr_; oz_ <= zr_; ++oz_)                                        \
      for (int oy_ = -1; oy_ <= 1; ++oy_)                                          \
        for (int ox_ = -1; ox_ <= 1; ++ox_) {                                      \
          const int cx_ = (ci_[0] + ox_ + (C)->cps[0]) % (C)->cps[0];              \
          const int cy_ = (ci_[1] + oy_ + (C)->cps[1]) % (C)->cps[1];              \
          const int cz_ = dim_ == 3 ? (ci_[2] + oz_ + (C)->cps[2]) % (C)->cps[2] : 0; \
          const int hc_ = cx_ + (C)->cps[0] * (cy_ + (C)->cps[1] * cz_);          \
          for (int t_ = (C)->start[hc_]; t_ < (C)->start[hc_ + 1]; ++t_) {        \
            const int j = (C)->order[t_];                                          \
            float dr[3] = {0.0f, 0.0f, 0.0f};                                      \
            float d2 = 0.0f;                                                       \
            for (int d_ = 0; d_ < dim_; ++d_) {                                    \
              dr[d_] = pdisp((R)[(I) * dim_ + d_], (R)[j * dim_ + d_], (P)->box[d_]); \
              d2 += dr[d_] * dr[d_];                                               \
            }                                                                      \
            if (!(d2 < c2_)) continue;                                             \
            __VA_ARGS__                                                            \
          }                                                                        \
        }                                                                          \
  } while (0)

static float tait_p(const sph_cpu_params* p, float rho) {
  float x = rho / p->rho_ref;
  if (p->gamma != 1.0f) x = powf(x, p->gamma);
  return p->p_ref * (x - 1.0f) + p->p_bg;
}

/* One si_euler step in place.  Arrays are (n, dim) or (n,), float32. */
int sph_cpu_step(const sph_cpu_params* p, float* r, float* u, float* v, float* dudt,
                 float* dvdt, float* rho, float* pr, const float* mass, const float* eta) {
  const int n = p->n, dim = p->dim;
  if (n <= 0 || (dim != 2 && dim != 3) || (p->kernel != 1 && p->kernel != 2)) return -1;
  for (int d = 0; d < dim; ++d)
    if ((int)floorf(p->box[d] / p->cutoff) < 3) return -1;
#ifdef _OPENMP
  if (p->threads > 0) omp_set_num_threads(p->threads);
#endif
  const float ctvf = p->ctvf;
  /* 1. kick, transport velocity, drift (integrator.py:26-30) */
#pragma omp parallel for schedule(static)
  for (int k = 0; k < n * dim; ++k) {
    const int d = k % dim;
    u[k] = u[k] + p->dt * dudt[k];
    v[k] = u[k] + ctvf * dvdt[k];
    r[k] = jmodf(r[k] + p->dt * v[k], p->box[d]);
  }
  /* 2. cell list at the new positions */
  cells_t c;
  c.start = NULL;
  {
    int ncell = 1;
    for (int d = 0; d < dim; ++d) ncell *= (int)floorf(p->box[d] / p->cutoff);
    c.start = (int*)malloc(sizeof(int) * ((size_t)ncell + 1));
  }
  c.order = (int*)malloc(sizeof(int) * (size_t)n);
  c.cell = (int*)malloc(sizeof(int) * (size_t)n);
  cells_build(p, r, &c);
  const int kind = p->kernel;
  const float s = p->sigma, ih = p->one_over_h;
  /* 3. density summation and EOS (solver.py:176-178, 801) */
#pragma omp parallel for schedule(dynamic, 256)
  for (int t = 0; t < n; ++t) {
    const int i = c.order[t];
    float wsum = 0.0f;
    FOR_NEIGHBOURS(p, &c, r, i, { wsum += kern_w(kind, sqrtf(d2) * ih, s); });
    rho[i] = mass[i] * wsum;
  }
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) pr[i] = tait_p(p, rho[i]);
  const float pbg = tait_p(p, 0.0f); /* background pressure of the TVF, solver.py:802 */
  /* 4. accelerations (solver.py:199-256) */
#pragma omp parallel for schedule(dynamic, 256)
  for (int t = 0; t < n; ++t) {
    const int i = c.order[t];
    const float mi = mass[i], rhoi = rho[i], pi_ = pr[i], etai = eta[i];
    const float Vi = mi / rhoi;
    float a[3] = {0.0f, 0.0f, 0.0f}, b[3] = {0.0f, 0.0f, 0.0f};
    float wi[3] = {0.0f, 0.0f, 0.0f};
    for (int d = 0; d < dim; ++d) wi[d] = v[i * dim + d] - u[i * dim + d];
    FOR_NEIGHBOURS(p, &c, r, i, {
      if (j != i) {
        const float dist = sqrtf(d2);
        const float gw = kern_dw(kind, dist * ih, s, ih);
        const float Vj = mass[j] / rho[j];
        const float cpre = (Vi * Vi + Vj * Vj) / mi * gw / (dist + kEpsF);
        const float rhoj = rho[j];
        const float pij = (rhoj * pi_ + rhoi * pr[j]) / (rhoi + rhoj);
        const float eij = 2.0f * etai * eta[j] / (etai + eta[j] + kEpsF);
        float wr_i = 0.0f, wr_j = 0.0f;
        for (int d = 0; d < dim; ++d) {
          wr_i += wi[d] * dr[d];
          wr_j += (v[j * dim + d] - u[j * dim + d]) * dr[d];
        }
        for (int d = 0; d < dim; ++d) {
          const float A = 0.5f * (rhoi * u[i * dim + d] * wr_i + rhoj * u[j * dim + d] * wr_j);
          a[d] += cpre * (-pij * dr[d] + A + eij * (u[i * dim + d] - u[j * dim + d]));
          b[d] += cpre * pbg * dr[d];
        }
      }
    });
    for (int d = 0; d < dim; ++d) {
      dudt[i * dim + d] = a[d] + p->g[d];
      dvdt[i * dim + d] = b[d];
    }
  }
  free(c.start);
  free(c.order);
  free(c.cell);
  return 0;
}

int sph_cpu_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
