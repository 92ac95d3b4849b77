/*
 * C-ABI of the MI355X SPH step library (libsph.so, gfx950).
 *
 * The reference (tumaer/jax-sph) has no native layer: its hot path is a set of
 * Python closures compiled by XLA.  Each entry point below replaces one of
 * those closures and is what a Python/ctypes (or any FFI) binding calls with
 * raw device pointers from contiguous tensors.
 *
 * Conventions
 *   - every function returns SPH_OK (0) or a negative SPH_ERR_* code;
 *     sph_last_error() returns a message for the last failure on this thread;
 *   - arrays are device pointers, float32, particle-major (N, dim) row-major
 *     like the reference's state dict; tags are int32;
 *   - the caller owns every state/output buffer and the int64 `info` block;
 *     the workspace owns its scratch, sized once at sph_ws_create: update,
 *     forward and step never allocate;
 *   - all work is enqueued on `stream`; a workspace is not thread-safe
 *     (one workspace per stream).
 *
 * info block (int64[8], device):
 *   [0] PartitionErrorCode bits      [1] occupancy (real pairs)
 *   [2] max particles in one cell    [3] max pairs of one particle
 *   [4] rebuild flag (dr_threshold)  [5] internal: pair-buffer truncations
 *   [6] chunk overflow count (scan)  [7] reserved
 */
#ifndef SPH_H
#define SPH_H

#include <stddef.h>
#include <stdint.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SPH_OK 0
#define SPH_ERR_ARG (-1)
#define SPH_ERR_HIP (-2)
#define SPH_ERR_SIZE (-3)

/* PartitionErrorCode, jax_sph/jax_md/partition.py:434-457 */
#define SPH_NEIGHBOR_LIST_OVERFLOW 1
#define SPH_CELL_LIST_OVERFLOW 2
#define SPH_CELL_SIZE_TOO_SMALL 4
#define SPH_MALFORMED_BOX 8

/* kernel.py:26-201 */
enum sph_kernel_kind {
  SPH_KERNEL_CSK = 0, SPH_KERNEL_QSK = 1, SPH_KERNEL_WC2K = 2, SPH_KERNEL_WC4K = 3,
  SPH_KERNEL_WC6K = 4, SPH_KERNEL_GK = 5, SPH_KERNEL_SGK = 6
};

typedef struct sph_ws sph_ws;

/* Geometry of one neighbour search (jax_md/partition.py:114-156, 723-822).
 * The host derives cps/cell_size with the reference's float32 arithmetic. */
typedef struct {
  int32_t dim;            /* 2 or 3 */
  int32_t n;              /* particles */
  float box[3];           /* periodic side lengths, float32 */
  float cutoff;           /* fl32(r_cutoff) + fl32(dr_threshold) */
  float cutoff_sq;        /* fl32(cutoff * cutoff) */
  int32_t use_cells;      /* 0: every particle is a candidate (no cell list) */
  int32_t cps[3];         /* cells per side */
  float cell_size[3];     /* fl32(box / cps) */
  int32_t cell_capacity;  /* emulated fixed cell capacity: sets SPH_CELL_LIST_OVERFLOW */
  int32_t mask_self;      /* drop (i, i) pairs */
  int32_t ordered;        /* OrderedSparse: keep receiver < sender only */
  int32_t ell_width;      /* pairs stored per particle; >= 3^dim * cell_capacity */
} sph_nl_config;

/* WCSPH configuration (solver.py:616-700 plus case constants). */
typedef struct {
  int32_t dim;
  int32_t kernel;            /* enum sph_kernel_kind */
  int32_t solver;            /* 0 = SPH, 1 = RIE, 2 = DELTA (solver.py:684-694) */
  int32_t density_evolution; /* 0 = summation, 1 = continuity */
  int32_t bc_trick;
  int32_t free_slip;
  int32_t heat_conduction;
  int32_t eos_kind;          /* 0 = Tait, 1 = RIEMANN */
  float one_over_h, sigma;   /* fl32(1/h), fl32(kernel normalisation) */
  float dt, c_ref, eta_limiter, artificial_alpha, av_h;  /* av_h = dx (solver.py:406) */
  float p_ref, rho_ref, p_bg, gamma, inv_gamma, riemann_k; /* riemann_k = 100 u_ref^2 */
  float box[3];
  float g_const[3];          /* body force used when g_ext == NULL */
  int32_t tvf_stress;        /* 1: evaluate the A = rho u (v-u) term; 0: caller proves v == u */
  int32_t rho_renorm;        /* density renormalisation after continuity (solver.py:167-173) */
  float diff_delta;          /* DELTA: density diffusion coefficient (solver.py:693) */
  float diff_alpha;          /* DELTA: velocity diffusion coefficient (solver.py:687) */
  float support;             /* DELTA: h_fac * dx (solver.py:688, 693) */
  int32_t pad_;
} sph_solver_params;

/* Input state (orig particle order). nw/T/... may be NULL when unused. */
typedef struct {
  const float *r, *u, *v, *dudt, *dvdt, *rho, *p, *mass, *eta;
  const float *T, *dTdt, *kappa, *Cp, *nw, *drhodt;
  const float *g_ext;        /* (N, dim) body force, or NULL for g_const */
  const int32_t *tag;
} sph_state_in;

/* Outputs of WCSPH.forward (solver.py:930-947); all (orig order). */
typedef struct {
  float *u, *v, *dudt, *dvdt, *rho, *p, *T, *dTdt, *drhodt;
} sph_state_out;

/* Case boundary hook applied after the RHS (integrator.py:54):
 * 0 none (tgv.py:56-57), 1 lid-driven cavity (ldc.py:130-144),
 * 2 dam break (db.py:299-307), 3 relaxation walls (case_setup.py:393-402). */
typedef struct {
  int32_t kind;
  float u_lid[3];
} sph_bc_spec;

/* ---- library ---------------------------------------------------------- */
const char *sph_last_error(void);
int sph_version(void);

/* ---- workspace: replaces the buffers XLA materialises per call ---------- */
/* Allocates the cell arrays, two banks of sorted order + pair buffer, and the
 * sorted particle records.  Replaces the shape decisions of
 * NeighborListFns.allocate (jax_md/partition.py:956-971). */
int sph_ws_create(const sph_nl_config *cfg, sph_ws **out);
int sph_ws_destroy(sph_ws *ws);
/* bytes of device memory held by the workspace */
int64_t sph_ws_bytes(const sph_ws *ws);
/* Launch-shape knobs of a workspace (no counterpart in the reference, whose
 * launches XLA decides).  "xcd_remap" (default 1): the pair build and the
 * sweeps hand each of the 8 XCDs a contiguous range of cell-sorted particles.
 * "skin_rebuild" (1): the next fused step rebuilds the skin lists, as after
 * sph_skin_enable (synchronous; outside graph capture).
 * "ell_overflow" (1): for a workspace whose ell_width is below 3^dim x
 * cell_capacity, a particle with more pairs than fit raises
 * SPH_NEIGHBOR_LIST_OVERFLOW (default 0: such a width is not expected). */
int sph_ws_tune(sph_ws *ws, const char *key, int64_t value);

/* ---- neighbour search: jax_sph/partition.py:492-571 + jax_md/partition.py:723-1026 */
/* Max particles in one cell on a grid (count_cell_filling, partition.py:159-174);
 * writes info[2].  Used by allocate, before a workspace exists, to size the
 * emulated cell capacity; `counts` is caller scratch of prod(cps) int32. */
int sph_nl_fill_max(int32_t dim, int32_t n, const float *r, const float *cell_size,
                    const int32_t *cps, int32_t *counts, int64_t *info, hipStream_t stream);
/* sph_nl_fill_max without caller scratch and without any allocation.  No
 * counterpart in the reference: it is for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The cell counts live in a static device array of
 * SPH_FILL_MAX_STATIC_CELLS int32 (8 MiB) per GPU, held until the process
 * ends; a grid with more cells returns SPH_ERR_SIZE.  Every call on one GPU
 * shares that array (cleared by a kernel at the start of each call), so calls
 * that overlap on one GPU (different streams or threads) race. */
#define SPH_FILL_MAX_STATIC_CELLS (1 << 21)
int sph_nl_fill_max_static(int32_t dim, int32_t n, const float *r, const float *cell_size,
                           const int32_t *cps, int64_t *info, hipStream_t stream);
/* Dense cell list of CellListFns.allocate/update (jax_md/partition.py:243-428,
 * cell_list_fn at :294-413).  Buffers are cell-major over the flat cell hash
 * x + nx*(y + ny*z), `capacity` slots per cell: position_buffer
 * [cells*capacity*dim] f32, id_buffer [cells*capacity] i32, and n_named side
 * arrays (per particle named_width[b] 32-bit words, read from named[b]) into
 * named_buffer[b] [cells*capacity*named_width[b]].  Empty slots hold position
 * 0, id n and named_fill[b] in every word.  The particle at cell-sorted
 * position s (cells ascending, ids ascending inside a cell: the stable argsort
 * of :367-373) goes to slot cell*capacity + s mod capacity (:379-385).  In a
 * cell holding more than `capacity` particles the later one of two that share
 * a slot keeps it, and info[0] |= SPH_CELL_LIST_OVERFLOW (info[0] is not
 * cleared, so overflow is sticky as in update at :396).  info[2] = the
 * largest cell occupancy.  Cells are int32(r / cell_size) clamped into the
 * grid (the reference hashes r == side into a wrong cell).  `work` is caller
 * scratch of sph_cell_list_work_bytes(n, cells) bytes, 4-byte aligned. */
size_t sph_cell_list_work_bytes(int64_t n, int64_t cell_count);
int sph_cell_list(int32_t dim, int32_t n, const float *r, const float *cell_size,
                  const int32_t *cps, int64_t capacity, float *position_buffer,
                  int32_t *id_buffer, int32_t n_named, const uint32_t *const *named,
                  const int32_t *named_width, const uint32_t *named_fill,
                  uint32_t *const *named_buffer, void *work, size_t work_bytes, int64_t *info,
                  hipStream_t stream);
/* Rebuild test of update (partition.py:1010-1018): info[4] = any(|r - r_ref|^2 > thr_sq). */
int sph_nl_needs_rebuild(sph_ws *ws, const float *r, const float *r_ref, float threshold_sq,
                         int64_t *info, hipStream_t stream);
/* Cell list + candidate prune (partition.py:294-413, 832-909) into `bank`:
 * bins r, sorts by cell, stores every pair with fp32 min-image d^2 < cutoff^2,
 * writes info[0..3,5].  prev_info (nullable) supplies the sticky error bits.
 * max_occupancy < 0 leaves SPH_NEIGHBOR_LIST_OVERFLOW for sph_nl_finalize. */
int sph_nl_build(sph_ws *ws, int bank, const float *r, const int64_t *prev_info,
                 int64_t max_occupancy, int64_t *info, hipStream_t stream);
/* Sets SPH_NEIGHBOR_LIST_OVERFLOW iff info[1] > max_occupancy (partition.py:976). */
int sph_nl_finalize(sph_ws *ws, int64_t max_occupancy, int64_t *info, hipStream_t stream);
/* jaxmd_scan chunk budgets (jax_sph/partition.py:233-291): adds
 * SPH_CELL_LIST_OVERFLOW when a chunk of n_sub senders holds > budget pairs. */
int sph_nl_scan_budget(sph_ws *ws, int bank, int64_t n_sub, int64_t budget, int64_t *info,
                       hipStream_t stream);
/* Materialises NeighborList.idx: int32[2, max_occupancy], row 0 receivers,
 * row 1 senders (sender-sorted), padded with N (partition.py:885-909, 971). */
int sph_nl_write_idx(sph_ws *ws, int bank, int32_t *idx, int64_t max_occupancy,
                     const int64_t *info, hipStream_t stream);

/* ---- Verlet-skin lists of the fused step (bank 2) ------------------------
 * The reference rebuilds the neighbour list on every si_euler step
 * (integrator.py:46, dr_threshold = 0).  With skin lists, sph_step on bank 2
 * keeps pairs closer than cutoff + skin, rebuilds them on the device only
 * after some particle has moved more than 0.49 * skin since the last build,
 * and applies the reference predicate (partition.py:895-897) to every pair on
 * every step.  Pair sets, occupancy info[1], info[3] and the error bits are
 * those of a rebuild at every step; info[7] holds the largest skin list of one
 * particle at the last rebuild.  A skin list that outgrows `width` raises
 * SPH_NEIGHBOR_LIST_OVERFLOW, so the caller re-allocates as for any overflow.
 * Returns 1 when enabled, 0 when this geometry cannot use them (no cell list,
 * OrderedSparse, fewer than 3 skin cells per side, cells too close to the
 * cutoff), < 0 on error.  Allocates; call once, after sph_ws_create. */
int sph_skin_enable(sph_ws *ws, float skin, int32_t width);
/* NeighborList.idx of the skin lists at positions r (orig order, those of the
 * step that produced the list), same layout as sph_nl_write_idx. */
int sph_nl_write_idx_skin(sph_ws *ws, const float *r, int32_t *idx, int64_t max_occupancy,
                          hipStream_t stream);

/* ---- physics ------------------------------------------------------------ */
/* W(r) and dW/dr of one kernel, elementwise (kernel.py:9-201). */
int sph_kernel_eval(int32_t kind, int32_t dim, float one_over_h, float sigma, const float *r,
                    int64_t n, float *w, float *dw, hipStream_t stream);
/* WCSPH.forward (solver.py:705-949) over the pairs of `bank` (0, 1, or 2 for
 * the skin lists, filtered at in->r). */
int sph_forward(sph_ws *ws, int bank, const sph_solver_params *prm, const sph_state_in *in,
                sph_state_out *out, hipStream_t stream);
/* One si_euler step (integrator.py:22-56) fused with the neighbour update,
 * forward and the case boundary hook: reads `in`, writes r/u/v to r_out/out
 * and the rest of the state to `out`; the new pair set goes to `bank`
 * (0 or 1: exact lists rebuilt now; 2: the skin lists, see sph_skin_enable). */
int sph_step(sph_ws *ws, int bank, const sph_solver_params *prm, float dt, float tvf_coef,
             const sph_bc_spec *bc, const sph_state_in *in, float *r_out, sph_state_out *out,
             const int64_t *prev_info, int64_t max_occupancy, int64_t *info,
             hipStream_t stream);

/* ---- adjoint ------------------------------------------------------------- */
/* Cotangents of the step's state, (N, dim), orig order: positions r, velocity
 * u, acceleration a = dudt, and for the transport-velocity step the transport
 * velocity v and its acceleration b = dvdt; with density continuity also the
 * density rho, (N,).  v, b and rho of bar_next may be NULL (zero); b and rho
 * of bar may be NULL (not written). */
typedef struct {
  const float *r, *u, *a, *v, *b, *rho;
} sph_cotangent_in;
typedef struct {
  float *r, *u, *a, *b, *rho;
} sph_cotangent_out;
/* Vector-Jacobian product of one si_euler step (integrator.py:22-56), the
 * transpose that jax.grad builds for the reference's step.  The Riemann
 * solver (solver.py:108-164, 316-401) with density summation or its
 * continuity, walls if any held by `bc`; delta-SPH with density summation
 * (standard SPH plus its velocity diffusion, solver.py:259-311; its density
 * diffusion, with continuity, is refused); or standard SPH
 * (heat conduction may be on: u', r' do not depend on the temperature, whose
 * cotangent is not formed); density summation or continuity (solver.py:24-30,
 * with or without renormalisation); with or without artificial viscosity (solver.py:404-428)
 * and the generalized wall boundary condition (bc_trick, solver.py:431-530)
 * for no-slip or free-slip walls (free slip needs next->nw); QSK or WC2K;
 * with prm->tvf_stress the
 * transport-velocity step
 * (solver.py:199-256: the A = rho u (v - u) stress and the background-pressure
 * dvdt), whose v' = u' + tvf_coef * dvdt (tvf_coef = tvf dt / 2, 0 without
 * tvf).  `bc` (nullable) is the step's case boundary hook as in sph_step:
 * the velocities and accelerations it sets are constants of the step, so the
 * particles it holds carry only a position cotangent.
 * `next` is state n+1 (r, u, v with tvf, rho, p, mass, eta, tag; g_ext, the
 * body force at next->r, or NULL for prm->g_const) and `bank`
 * must hold the pairs of next->r; given the cotangents of state n+1 writes
 * those of state n (bar may alias nothing in bar_next).  With density
 * continuity records_ready must be set (the density the step advanced is not
 * in `next`).  records_ready != 0
 * promises that the last kernels on this workspace were the sph_step that
 * produced `next` into `bank` (its sorted records are reused instead of
 * gathered again).  A position-dependent body force's own derivative,
 * (dg/dr')^T applied to the cotangent of dudt', is not included: the caller
 * adds it to bar_next->r (it is zero for piecewise-constant forces). */
int sph_step_vjp(sph_ws *ws, int bank, const sph_solver_params *prm, float dt, float tvf_coef,
                 const sph_bc_spec *bc, const sph_state_in *next,
                 const sph_cotangent_in *bar_next, sph_cotangent_out *bar,
                 int32_t records_ready, hipStream_t stream);

/* si_euler kick/drift alone (integrator.py:26-30), for callers that compose
 * the step from separate update/forward calls: u' = u + dt dudt,
 * v' = u' + tvf_coef dvdt, r' = mod(r + dt v', box). */
int sph_integrate(int32_t dim, int64_t n, const float *r, const float *u, const float *dudt,
                  const float *dvdt, float *r_out, float *u_out, float *v_out, float dt,
                  float tvf_coef, const float *box, hipStream_t stream);

/* ---- ensembles ------------------------------------------------------------ */
/* Many independent simulations advanced by one launch per phase of the fused
 * step.  The reference has no counterpart: it runs one simulation per process
 * (simulate.py:19-136), so an ensemble there is a host loop over sph_step-like
 * steps, one member after the other.  Each member is exactly one sph_step call
 * on bank 2 (its own workspace with skin lists enabled, its own parameters,
 * state buffers and error block); a member's results are bitwise those of that
 * call.  All members have the same dimension and distinct workspaces. */
typedef struct sph_ensemble sph_ensemble;
typedef struct {
  sph_ws *ws;
  const sph_solver_params *prm;
  float dt;
  float tvf_coef;
  const sph_bc_spec *bc; /* nullable: no case hook */
  const sph_state_in *in;
  float *r_out;
  sph_state_out *out;
  const int64_t *prev_info; /* nullable */
  int64_t max_occupancy;
  int64_t *info;
} sph_ensemble_member;
/* Copies the member descriptions (pointers are kept, not the data) and builds
 * the device argument tables and block maps.  Allocates. */
int sph_ensemble_create(const sph_ensemble_member *members, int32_t n, sph_ensemble **out);
/* One fused step of every member on `stream`; allocates nothing and makes no
 * host decisions, so a HIP graph can capture it. */
int sph_ensemble_step(sph_ensemble *e, hipStream_t stream);
/* "order": 0 given, 1 largest member first (default), 2 blocks interleaved;
 * "xcd": 1 (default) XCD-contiguous workgroup layout.  Rebuilds the block
 * maps (allocates): tune before capturing steps in a graph. */
int sph_ensemble_tune(sph_ensemble *e, const char *key, int64_t value);
/* members, groups of equal sweep plans, workgroups of one lane-parallel launch */
int sph_ensemble_info(const sph_ensemble *e, int64_t *members, int64_t *groups,
                      int64_t *lane_blocks);
int sph_ensemble_destroy(sph_ensemble *e);

/* ---- profiling hooks ---------------------------------------------------- */
/* When on, the library records a HIP event pair around every kernel it
 * launches and accumulates per-kernel device time (ms) and launch counts. */
int sph_prof_enable(sph_ws *ws, int on);
/* Copies up to `cap` entries: names (NUL-separated into `names`, size names_cap),
 * total ms and counts.  Returns the number of kernels. */
int sph_prof_read(sph_ws *ws, char *names, int names_cap, double *ms, int64_t *count, int cap);
int sph_prof_reset(sph_ws *ws);

#ifdef __cplusplus
}
#endif
#endif /* SPH_H */
