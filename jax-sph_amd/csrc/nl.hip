// Neighbour search on gfx950: counting-sort cell binning, a wavefront-scan
// prefix sum, per-cell id sort, and the exact float32 pair predicate.
//
// Replaces (all XLA-generated in the reference):
//   jax_md/partition.py:294-413   cell_list_fn (hash, argsort, slot scatter, occupancy)
//   jax_md/partition.py:832-909   3^d candidate copies + prune + cumsum compaction
//   jax_md/partition.py:956-976   capacity and NEIGHBOR_LIST_OVERFLOW
//   jax_md/partition.py:1010-1018 rebuild test of update
//   jax_sph/partition.py:208-291  jaxmd_scan chunk budgets
#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>
#include <string>

#include "ens.h"
#include "sph_device.h"
#include "sph_internal.h"

namespace sph {

constexpr int kBlock = 256;

__host__ __device__ inline int64_t cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }

// ------------------------------------------------------------------ helpers
template <typename T>
__global__ void k_fill(T* __restrict__ a, int64_t n, T v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

struct Grid {
  float cs[3];   // cell size (fp32)
  int cps[3];    // cells per side
  int use_cells;
};

void cut_band(const sph_nl_config& c, float* lo, float* hi);

// int32(r / cell_size) per dimension (partition.py:367), clamped into the grid:
// the reference aliases r == side into a wrong cell; clamping keeps it in the last one.
template <int DIM>
__device__ __forceinline__ void cell_coords(const float4 p, const Grid& g, int* c) {
  float x[3] = {p.x, p.y, p.z};
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    // fl32(x / cs), correctly rounded: a float64 quotient rounded to float32 is
    // the correctly rounded float32 quotient (53 >= 2*24 + 2 bits)
    int q = (int)(float)((double)x[d] / (double)g.cs[d]);
    q = q < 0 ? 0 : q;
    q = q >= g.cps[d] ? g.cps[d] - 1 : q;
    c[d] = q;
  }
}

template <int DIM>
__device__ __forceinline__ int cell_hash(const int* c, const Grid& g) {
  int h = c[0] + c[1] * g.cps[0];
  if (DIM == 3) h += c[2] * g.cps[0] * g.cps[1];
  return h;
}

template <int DIM>
__device__ __forceinline__ float4 load_pos(const float* __restrict__ r, int64_t i) {
  float4 p;
  p.x = r[i * DIM + 0];
  p.y = r[i * DIM + 1];
  p.z = DIM == 3 ? r[i * DIM + 2] : 0.0f;
  p.w = 0.0f;
  return p;
}

// ----------------------------------------------------------- binning kernels
// atomicAdd(&counter[key], 1) for every active lane, with one atomic per
// distinct key in the wavefront (neighbouring particles share cells, so a
// wave holds only a few keys).  Returns the lane's slot.
__device__ __forceinline__ int wave_aggregated_inc(int32_t* __restrict__ counter, int key) {
  const int lane = threadIdx.x & 63;
  const unsigned long long lt = (1ull << lane) - 1ull;
  unsigned long long active = __ballot(1);
  int slot = 0;
  while (active) {
    const int leader = __ffsll((long long)active) - 1;
    const int lkey = __shfl(key, leader, 64);
    const unsigned long long same = __ballot(key == lkey) & active;
    int base = 0;
    if (lane == leader) base = atomicAdd(&counter[lkey], __popcll(same));
    base = __shfl(base, leader, 64);
    if (key == lkey) slot = base + __popcll(same & lt);
    active &= ~same;
  }
  return slot;
}

// atomicAdd(&counter[key], 1) for every active lane, with one atomic per run
// of equal keys in consecutive lanes (particles in id order come in runs of
// one cell), issued by all runs at once.  The active lanes must be a prefix
// of the wavefront.  Returns nothing: the counts only.
__device__ __forceinline__ void wave_run_inc(int32_t* __restrict__ counter, int key) {
  const int lane = threadIdx.x & 63;
  const int prev = __shfl_up(key, 1, 64);
  const bool start = lane == 0 || prev != key;
  const unsigned long long starts = __ballot(start);
  const unsigned long long active = __ballot(1);
  if (start) {
    const unsigned long long later = lane == 63 ? 0ull : (starts >> (lane + 1)) << (lane + 1);
    const int end = later ? __ffsll((long long)later) - 1 : 64 - __clzll((long long)active);
    atomicAdd(&counter[key], end - lane);
  }
}

template <int DIM>
__device__ __forceinline__ void hash_count_body(const float* __restrict__ r, int n, const Grid& g,
                                                int32_t* __restrict__ cell_count,
                                                int32_t* __restrict__ pcell,
                                                int32_t* __restrict__ pslot,
                                                const int32_t* __restrict__ gate, int blk) {
  if (gate && *gate == 0) return;
  int i = blk * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int h = 0;
  if (g.use_cells) {
    int c[3];
    cell_coords<DIM>(load_pos<DIM>(r, i), g, c);
    h = cell_hash<DIM>(c, g);
  }
  const int slot = wave_aggregated_inc(cell_count, h);
  if (pcell) {
    pcell[i] = h;
    pslot[i] = slot;
  }
}

template <int DIM>
__global__ void k_hash_count(const float* __restrict__ r, int n, Grid g,
                             int32_t* __restrict__ cell_count, int32_t* __restrict__ pcell,
                             int32_t* __restrict__ pslot, const int32_t* __restrict__ gate) {
  hash_count_body<DIM>(r, n, g, cell_count, pcell, pslot, gate, blockIdx.x);
}

// si_euler kick + transport velocity + drift + periodic wrap (integrator.py:26-30)
// fused with the cell hash and histogram of the new position.
template <int DIM>
__global__ void k_integrate_hash(const float* __restrict__ r, const float* __restrict__ u,
                                 const float* __restrict__ dudt, const float* __restrict__ dvdt,
                                 float* __restrict__ r_out, float* __restrict__ u_out,
                                 float* __restrict__ v_out, int n, float dt, float ctvf, Grid g,
                                 float bx, float by, float bz, int32_t* __restrict__ cell_count,
                                 int32_t* __restrict__ pcell, int32_t* __restrict__ pslot) {
#pragma clang fp contract(off)
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float side[3] = {bx, by, bz};
  float x[3] = {0.f, 0.f, 0.f};
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    const int64_t k = (int64_t)i * DIM + d;
    const float uu = u[k] + dt * dudt[k];
    const float vv = uu + ctvf * dvdt[k];
    float xx = r[k] + dt * vv;
    if (xx >= side[d] || xx < 0.0f) xx = jmod(xx, side[d]);
    u_out[k] = uu;
    v_out[k] = vv;
    r_out[k] = xx;
    x[d] = xx;
  }
  int h = 0;
  if (g.use_cells) {
    int c[3];
    cell_coords<DIM>(make_float4(x[0], x[1], x[2], 0.0f), g, c);
    h = cell_hash<DIM>(c, g);
  }
  const int slot = wave_aggregated_inc(cell_count, h);
  pcell[i] = h;
  pslot[i] = slot;
}

// Skin-list step: kick/drift/wrap as k_integrate_hash, the histogram of the
// new positions on the reference grid (for its cell-capacity bit), and the
// rebuild test: a particle that moved more than the trigger distance since
// the last build (minimum image) raises *flag.
template <int DIM>
__device__ __forceinline__ void integrate_skin_body(
    const float* __restrict__ r, const float* __restrict__ u, const float* __restrict__ dudt,
    const float* __restrict__ dvdt, float* __restrict__ r_out, float* __restrict__ u_out,
    float* __restrict__ v_out, int n, float dt, float ctvf, const Grid& g, float bx, float by,
    float bz, int32_t* __restrict__ cell_count, const float4* __restrict__ posb, float trigger_sq,
    int32_t* __restrict__ flag, int blk) {
#pragma clang fp contract(off)
  const float side[3] = {bx, by, bz};
  float x[3] = {0.f, 0.f, 0.f};
  int i;
  if constexpr (DIM == 3) {
    // The block's 3 x kBlock floats of each array pass through LDS, so every
    // wave-wide load and store covers 64 consecutive floats; per-particle
    // stride-3 accesses touch three times as many cache lines per instruction.
    // Every block of a launch starts below n (launches use cdiv(n, kBlock)
    // blocks), and every thread reaches the barriers.
    __shared__ float s[4][3 * kBlock];
    const int i0 = blk * kBlock;
    i = i0 + threadIdx.x;
    const int64_t base = (int64_t)i0 * 3;
    const int lim = 3 * min(kBlock, n - i0);
    for (int k = threadIdx.x; k < lim; k += kBlock) {
      s[0][k] = r[base + k];
      s[1][k] = u[base + k];
      s[2][k] = dudt[base + k];
      s[3][k] = dvdt[base + k];
    }
    __syncthreads();
    float uo[3], vo[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const int k = 3 * threadIdx.x + d;
      uo[d] = s[1][k] + dt * s[2][k];
      vo[d] = uo[d] + ctvf * s[3][k];
      float xx = s[0][k] + dt * vo[d];
      if (xx >= side[d] || xx < 0.0f) xx = jmod(xx, side[d]);
      x[d] = xx;
    }
    __syncthreads();
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const int k = 3 * threadIdx.x + d;
      s[0][k] = x[d];
      s[1][k] = uo[d];
      s[2][k] = vo[d];
    }
    __syncthreads();
    for (int k = threadIdx.x; k < lim; k += kBlock) {
      r_out[base + k] = s[0][k];
      u_out[base + k] = s[1][k];
      v_out[base + k] = s[2][k];
    }
    if (i >= n) return;
  } else {
    i = blk * blockDim.x + threadIdx.x;
    if (i >= n) return;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      const int64_t k = (int64_t)i * DIM + d;
      const float uu = u[k] + dt * dudt[k];
      const float vv = uu + ctvf * dvdt[k];
      float xx = r[k] + dt * vv;
      if (xx >= side[d] || xx < 0.0f) xx = jmod(xx, side[d]);
      u_out[k] = uu;
      v_out[k] = vv;
      r_out[k] = xx;
      x[d] = xx;
    }
  }
  int c[3];
  cell_coords<DIM>(make_float4(x[0], x[1], x[2], 0.0f), g, c);
  wave_run_inc(cell_count, cell_hash<DIM>(c, g));
  const float4 b = posb[i];
  const float pb[3] = {b.x, b.y, b.z};
  float m2 = 0.0f;
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    const float dd = disp_fast(x[d], pb[d], side[d], 1.0f / side[d]);
    m2 += dd * dd;
  }
  if (m2 > trigger_sq) *flag = 1;  // same value from every writer
}

template <int DIM>
__global__ void k_integrate_skin(const float* __restrict__ r, const float* __restrict__ u,
                                 const float* __restrict__ dudt, const float* __restrict__ dvdt,
                                 float* __restrict__ r_out, float* __restrict__ u_out,
                                 float* __restrict__ v_out, int n, float dt, float ctvf, Grid g,
                                 float bx, float by, float bz, int32_t* __restrict__ cell_count,
                                 const float4* __restrict__ posb, float trigger_sq,
                                 int32_t* __restrict__ flag) {
  integrate_skin_body<DIM>(r, u, dudt, dvdt, r_out, u_out, v_out, n, dt, ctvf, g, bx, by, bz,
                           cell_count, posb, trigger_sq, flag, blockIdx.x);
}

// Largest reference-grid cell occupancy of this step (one workgroup): sets
// info[2] and the cell-capacity bit, resets the other info entries (keeping
// the sticky bits of prev[0]) and clears the histogram for the next step.
constexpr int kCellsThreads = 1024;
__device__ __forceinline__ void cells_max_body(int32_t* __restrict__ count, int64_t ncell,
                                               int64_t* __restrict__ info,
                                               const int64_t* __restrict__ prev,
                                               int64_t capacity) {
  __shared__ int wm[kCellsThreads / 64];
  int m = 0;
  for (int64_t i = threadIdx.x; i < ncell; i += kCellsThreads) {
    const int c = count[i];
    m = c > m ? c : m;
    count[i] = 0;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const int y = __shfl_down(m, o, 64);
    m = y > m ? y : m;
  }
  if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    int mm = 0;
    for (int w = 0; w < kCellsThreads / 64; ++w) mm = wm[w] > mm ? wm[w] : mm;
    int64_t bits = prev ? prev[0] : 0;
    if (capacity >= 0 && mm > capacity) bits |= SPH_CELL_LIST_OVERFLOW;
    info[0] = bits;
    for (int k = 1; k < 8; ++k) info[k] = 0;
    info[2] = mm;
  }
}

__global__ __launch_bounds__(kCellsThreads) void k_cells_max(int32_t* __restrict__ count,
                                                             int64_t ncell,
                                                             int64_t* __restrict__ info,
                                                             const int64_t* __restrict__ prev,
                                                             int64_t capacity) {
  cells_max_body(count, ncell, info, prev, capacity);
}

__global__ void k_max_reduce(const int32_t* __restrict__ a, int64_t n, int64_t* __restrict__ out) {
  int m = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    m = a[i] > m ? a[i] : m;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    int y = __shfl_down(m, o, 64);
    m = y > m ? y : m;
  }
  if ((threadIdx.x & 63) == 0) atomicMax((unsigned long long*)out, (unsigned long long)m);
}

// Single-workgroup exclusive scan with wavefront (64-lane) inclusive scans.
// out[k] = sum(in[0:k]) for k in [0, n]; also the max of in (for the cell capacity).
constexpr int kScanThreads = 1024;
constexpr int kScanItems = 8;

template <typename TIn, typename TOut>
__device__ __forceinline__ void scan_single_body(const TIn* in, int64_t n, TOut* __restrict__ out,
                                                 int64_t* __restrict__ max_out,
                                                 int64_t* __restrict__ info, int64_t capacity,
                                                 TIn* clear, const int32_t* __restrict__ gate) {
  // `clear` (the input array itself, or null) is zeroed as it is read;
  // `gate` (or null) skips the whole scan when *gate == 0.
  __shared__ TOut wave_tot[kScanThreads / 64];
  __shared__ TOut carry_s;
  __shared__ int64_t wmax[kScanThreads / 64];
  if (gate && *gate == 0) return;
  const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
  TOut carry = 0;
  int64_t mx = 0;
  const int64_t chunk = (int64_t)kScanThreads * kScanItems;
  for (int64_t base = 0; base < n; base += chunk) {
    TOut v[kScanItems];
    TOut s = 0;
    int64_t i0 = base + (int64_t)t * kScanItems;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
      int64_t i = i0 + k;
      TOut x = i < n ? (TOut)in[i] : (TOut)0;
      if (clear && i < n) clear[i] = 0;
      v[k] = x;
      s += x;
      mx = (int64_t)x > mx ? (int64_t)x : mx;
    }
    // inclusive wave scan of the per-thread sums
    TOut inc = s;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      TOut y = __shfl_up(inc, o, 64);
      if (lane >= o) inc += y;
    }
    if (lane == 63) wave_tot[wid] = inc;
    __syncthreads();
    if (wid == 0) {
      TOut w = lane < kScanThreads / 64 ? wave_tot[lane] : (TOut)0;
      TOut winc = w;
#pragma unroll
      for (int o = 1; o < 16; o <<= 1) {
        TOut y = __shfl_up(winc, o, 64);
        if (lane >= o) winc += y;
      }
      if (lane < kScanThreads / 64) wave_tot[lane] = winc - w;  // exclusive
      if (lane == kScanThreads / 64 - 1) carry_s = winc;
    }
    __syncthreads();
    TOut run = carry + wave_tot[wid] + (inc - s);
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
      int64_t i = i0 + k;
      if (i < n) out[i] = run;
      run += v[k];
    }
    carry += carry_s;
    __syncthreads();
  }
  // max reduction
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    int64_t y = __shfl_down(mx, o, 64);
    mx = y > mx ? y : mx;
  }
  if (lane == 0) wmax[wid] = mx;
  __syncthreads();
  if (t == 0) {
    out[n] = carry;
    int64_t m = 0;
    for (int w = 0; w < kScanThreads / 64; ++w) m = wmax[w] > m ? wmax[w] : m;
    if (max_out) *max_out = m;
    if (info && capacity >= 0 && m > capacity) info[0] |= SPH_CELL_LIST_OVERFLOW;
  }
}

template <typename TIn, typename TOut>
__global__ __launch_bounds__(kScanThreads) void k_scan_single(const TIn* in, int64_t n,
                                                              TOut* __restrict__ out,
                                                              int64_t* __restrict__ max_out,
                                                              int64_t* __restrict__ info,
                                                              int64_t capacity,
                                                              TIn* clear = nullptr,
                                                              const int32_t* __restrict__ gate =
                                                                  nullptr) {
  scan_single_body<TIn, TOut>(in, n, out, max_out, info, capacity, clear, gate);
}

__device__ __forceinline__ void scatter_body(const int32_t* __restrict__ pcell,
                                             const int32_t* __restrict__ pslot,
                                             const int32_t* __restrict__ cell_start,
                                             int32_t* __restrict__ unsorted, int n,
                                             const int32_t* __restrict__ gate, int blk) {
  if (gate && *gate == 0) return;
  int i = blk * blockDim.x + threadIdx.x;
  if (i >= n) return;
  unsorted[cell_start[pcell[i]] + pslot[i]] = i;
}

__global__ void k_scatter(const int32_t* __restrict__ pcell, const int32_t* __restrict__ pslot,
                          const int32_t* __restrict__ cell_start, int32_t* __restrict__ unsorted,
                          int n, const int32_t* __restrict__ gate = nullptr) {
  scatter_body(pcell, pslot, cell_start, unsorted, n, gate, blockIdx.x);
}

// Morton code of the particle's 4^DIM sub-cell inside its cell (2 bits per
// axis, the octant's bits above the quarter's).
template <int DIM>
__device__ __forceinline__ unsigned sub_cell_code(const float4 p, const Grid& g) {
  const float x[3] = {p.x, p.y, p.z};
  unsigned code = 0u;
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    const float q = (float)((double)x[d] / (double)g.cs[d]);
    int c = (int)q;
    c = c < 0 ? 0 : (c >= g.cps[d] ? g.cps[d] - 1 : c);
    int f = (int)((q - (float)c) * 4.0f);
    f = f < 0 ? 0 : (f > 3 ? 3 : f);
    code |= ((unsigned)(f >> 1) << (DIM + d)) | ((unsigned)(f & 1) << d);
  }
  return code;
}

// Sort key of particle `id` within its cell: the id alone, or (with positions)
// the sub-cell code first and the id second.
template <int DIM>
__device__ __forceinline__ unsigned long long cell_sort_key(int id, const float* __restrict__ r,
                                                            const Grid& g) {
  unsigned long long k = (unsigned)id;
  if (r) k |= (unsigned long long)sub_cell_code<DIM>(load_pos<DIM>(r, id), g) << 32;
  return k;
}

// One wavefront per cell: rank-sort the ids that arrived in the cell.  Without
// positions (r == nullptr) a cell lists its particles in ascending id, like the
// reference's stable argsort; with positions (skin lists, whose order is free)
// the particles of a cell follow a Morton curve over its sub-cells, so that
// consecutive particles, which share wavefronts in the sweeps, are close.
template <int DIM>
__device__ __forceinline__ void cell_sort_body(const int32_t* __restrict__ cell_start,
                                               const int32_t* __restrict__ unsorted,
                                               int32_t* __restrict__ perm,
                                               int32_t* __restrict__ iperm, int64_t ncell,
                                               const int32_t* __restrict__ gate, int blk,
                                               const float* __restrict__ r, const Grid& g) {
  if (gate && *gate == 0) return;
  const int lane = threadIdx.x & 63;
  const int64_t cell = ((int64_t)blk * blockDim.x + threadIdx.x) >> 6;
  if (cell >= ncell) return;
  const int b = cell_start[cell], e = cell_start[cell + 1], n = e - b;
  if (n <= 64) {
    const unsigned long long key = lane < n ? cell_sort_key<DIM>(unsorted[b + lane], r, g)
                                            : ~0ull;
    int rank = 0;
    for (int k = 0; k < n; ++k) {
      const unsigned long long o = __shfl(key, k, 64);
      rank += (o < key) ? 1 : 0;
    }
    if (lane < n) {
      const int id = (int)(unsigned)(key & 0xffffffffull);
      perm[b + rank] = id;
      iperm[id] = b + rank;
    }
  } else {
    for (int k0 = lane; k0 < n; k0 += 64) {
      const int id = unsorted[b + k0];
      const unsigned long long key = cell_sort_key<DIM>(id, r, g);
      int rank = 0;
      for (int k = 0; k < n; ++k) rank += (cell_sort_key<DIM>(unsorted[b + k], r, g) < key) ? 1 : 0;
      perm[b + rank] = id;
      iperm[id] = b + rank;
    }
  }
}

template <int DIM>
__global__ __launch_bounds__(256) void k_cell_sort(const int32_t* __restrict__ cell_start,
                                                   const int32_t* __restrict__ unsorted,
                                                   int32_t* __restrict__ perm,
                                                   int32_t* __restrict__ iperm, int64_t ncell,
                                                   const int32_t* __restrict__ gate,
                                                   const float* __restrict__ r, Grid g) {
  cell_sort_body<DIM>(cell_start, unsorted, perm, iperm, ncell, gate, blockIdx.x, r, g);
}

template <int DIM>
__global__ void k_gather_pos(const float* __restrict__ r, const int32_t* __restrict__ perm,
                             float4* __restrict__ pos, int n) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  pos[s] = load_pos<DIM>(r, perm[s]);
}

// ----------------------------------------------------------- pair build
struct BuildArgs {
  const float4* pos;
  const int32_t* cell_start;
  const int32_t* perm;
  int32_t* ell;
  int32_t* cnt;
  int64_t* info;
  Grid g;
  float box[3], half[3];
  float cutoff_sq;
  // fast filter: squared distances of the cell-adjacent image below `lo` are
  // in, above `hi` out; the band [lo, hi) bounds the difference between that
  // and the exact reference metric, and is decided by the exact predicate.
  float lo, hi;
  int fast;
  int n, npad, lane_width, mask_self;
  int64_t* part;          // [2 * gridDim.x]: pairs found and largest list, per workgroup
  int64_t trunc_bits;     // error bits raised when a sub-list overflows
  const int32_t* gate;    // nullable: skip the build when *gate == 0
  int remap;              // XCD-contiguous block order (grid a multiple of 8)
};

// Lane h (of TPP) tests candidates base + h, base + h + TPP, ... of one cell
// range [lo, hi) whose image is shifted by `img`; the TPP lanes of a particle
// step through the range together, 32 candidates each per chunk (so each load
// of the particle's lanes reads TPP consecutive positions), and collect their
// hits in a bit mask.  The particle's hits are numbered in candidate order
// over all its lanes, and hit k goes to sub-list k mod TPP at position k / TPP:
// the sub-lists differ in length by at most one, so the lanes that later sweep
// them finish together, and at each step of a sweep the particles of a
// wavefront read neighbours at similar places of the sorted order.
// The first kStage entries of each sub-list are staged in LDS (in the column
// of the thread that owns the sub-list) and written out row by row at the end,
// so every pair-buffer row is written by whole wavefronts at once; later
// entries go straight to the buffer.  24 staged entries (24 KB per workgroup)
// leave LDS for 6 workgroups per CU, which matches the 6 waves per SIMD the
// build's registers allow.
constexpr int kStage = 24;

// BAND = false (skin lists): pairs are kept by the fast squared distance
// alone.  A skin list must hold every pair within cutoff + 0.98 skin (the
// most two particles can close before the next rebuild), so rounding at its
// outer edge, cutoff + skin, cannot drop a pair that matters.
template <int DIM, int TPP, bool ORDERED, bool BAND>
__device__ __forceinline__ void test_range(const BuildArgs& a, int s, int h, const float4 ps,
                                           int sid, int lo, int hi, const float* img,
                                           int32_t* __restrict__ out0, int32_t* stage0,
                                           int& count) {
  // out0: entry 0 of the particle's sub-list 0; stage0: the stage column of
  // the particle's lane 0 (sub-list q is column stage0 + q)
  const float sx = ps.x - img[0], sy = ps.y - img[1], sz = ps.z - img[2];
  constexpr int G = 8;  // candidates whose loads are issued together
  for (int base = lo; base < hi; base += 32 * TPP) {  // same trip count on all TPP lanes
    const int c0 = base + h;
    int nb = c0 < hi ? (hi - c0 + TPP - 1) / TPP : 0;
    nb = nb < 32 ? nb : 32;
    unsigned mask = 0u;
    for (int t0 = 0; t0 < nb; t0 += G) {
      float4 pg[G];
      if (BAND) {
#pragma unroll
        for (int u = 0; u < G; ++u) {
          const int b = c0 + (t0 + u) * TPP;
          const int bc = b < hi ? b : hi - 1;
          __builtin_assume(bc >= 0);  // sorted slots: no sign extension in the address
          pg[u] = a.pos[bc];
        }
      } else {
        // skin builds read past the range into the padded tail (kSkinPosPad),
        // from one address with constant offsets; those candidates are masked
        // off below
        __builtin_assume(c0 >= 0);
        const float4* __restrict__ pb = a.pos + (c0 + t0 * TPP);
#pragma unroll
        for (int u = 0; u < G; ++u) pg[u] = pb[u * TPP];
      }
      unsigned got = 0u;  // bit u: candidate t0 + u is a pair
      if (!BAND) {
#pragma unroll
        for (int u = 0; u < G; ++u) {
          const float4 pb = pg[u];
          got |= dist2<DIM>(sx - pb.x, sy - pb.y, sz - pb.z) < a.lo ? 1u << u : 0u;
        }
      } else if (a.fast) {
        bool near = false;  // some candidate lies in the band around the cutoff
#pragma unroll
        for (int u = 0; u < G; ++u) {
          const float4 pb = pg[u];
          const float d2 = dist2<DIM>(sx - pb.x, sy - pb.y, sz - pb.z);
          got |= d2 < a.lo ? 1u << u : 0u;
          near = near || (d2 >= a.lo && d2 < a.hi);
        }
        if (near) {  // rare: the exact reference metric decides the band
#pragma unroll
          for (int u = 0; u < G; ++u) {
            const float4 pb = pg[u];
            const float d2 = dist2<DIM>(sx - pb.x, sy - pb.y, sz - pb.z);
            if (d2 >= a.lo && d2 < a.hi &&
                metric_sq_exact<DIM>(ps, pb, a.box, a.half) < a.cutoff_sq)
              got |= 1u << u;
          }
        }
      } else {
#pragma unroll
        for (int u = 0; u < G; ++u)
          if (metric_sq_exact<DIM>(ps, pg[u], a.box, a.half) < a.cutoff_sq) got |= 1u << u;
      }
      if (ORDERED) {  // candidates past the range are dropped before perm is read
#pragma unroll
        for (int u = 0; u < G; ++u) {
          const int b = c0 + (t0 + u) * TPP;
          if ((got >> u) & 1u && (t0 + u >= nb || !(a.perm[b] < sid))) got &= ~(1u << u);
        }
      }
      mask |= got << t0;
    }
    // drop the candidates past the range, and the particle itself
    if (nb < 32) mask &= (1u << nb) - 1u;
    if (a.mask_self) {
      const int ds = s - c0;
      if (ds >= 0 && ds % TPP == 0 && ds / TPP < 32) mask &= ~(1u << (ds / TPP));
    }
    // candidate base + t TPP + h follows the particle's hits at smaller t on
    // every lane and at the same t on the lanes below h
    unsigned m[TPP];
    int total = 0;
#pragma unroll
    for (int q = 0; q < TPP; ++q) {
      m[q] = __shfl(mask, q, TPP);
      total += __popc(m[q]);
    }
    while (mask) {
      const int t = __ffs(mask) - 1;
      mask &= mask - 1u;
      const unsigned upto = (2u << t) - 1u;  // bits 0..t (all 32 for t = 31)
      const unsigned below = upto >> 1;      // bits 0..t-1
      unsigned k = (unsigned)count;
#pragma unroll
      for (int q = 0; q < TPP; ++q) k += __popc(m[q] & (q < h ? upto : below));
      const int32_t b = c0 + t * TPP;
      const unsigned q = k % TPP, pos = k / TPP;
      if (pos < (unsigned)kStage) stage0[pos * kBlock + q] = b;
      else if (pos < (unsigned)a.lane_width)
        out0[(int64_t)(q * (unsigned)a.lane_width + pos) * a.npad] = b;
    }
    count += total;
  }
}

// TPP lanes per sorted particle (the sender): each visits the 3^d cells around
// the particle's own cell as 3^(d-1) contiguous x-strips, keeps its share of
// the receivers with d^2 < cutoff^2 and stores them in its own sub-list.
// blk: the workgroup's block within this list's grid; it also indexes the
// workgroup's partials.
template <int DIM, int TPP, bool ORDERED, bool BAND>
__device__ __forceinline__ void nl_build_body(const BuildArgs& a, const int blk) {
  __shared__ int32_t stage_buf[kStage][kBlock];
  if (a.gate && *a.gate == 0) return;  // uniform over the list's grid
  const int tid = blk * blockDim.x + threadIdx.x;
  const int s = tid / TPP, h = tid % TPP;
  int count = 0;  // hits of the particle (the same in its TPP lanes)
  if (s < a.n) {
    const float4 ps = a.pos[s];
    const int sid = ORDERED ? a.perm[s] : 0;
    int32_t* out = a.ell + s;                           // sub-list 0 of s
    int32_t* stage = &stage_buf[0][threadIdx.x - h];    // stage column of lane 0 of s
    float img[3] = {0.0f, 0.0f, 0.0f};
    if (!a.g.use_cells) {
      test_range<DIM, TPP, ORDERED, BAND>(a, s, h, ps, sid, 0, a.n, img, out, stage, count);
    } else {
      int c[3];
      cell_coords<DIM>(ps, a.g, c);
      const int nx = a.g.cps[0], ny = a.g.cps[1];
      const int nz = DIM == 3 ? a.g.cps[2] : 1;
      const int z0 = DIM == 3 ? -1 : 0, z1 = DIM == 3 ? 1 : 0;
      for (int dz = z0; dz <= z1; ++dz) {
        int cz = DIM == 3 ? c[2] + dz : 0;
        img[2] = 0.0f;
        if (cz < 0) { cz += nz; img[2] = -a.box[2]; }
        else if (cz >= nz) { cz -= nz; img[2] = a.box[2]; }
        for (int dy = -1; dy <= 1; ++dy) {
          int cy = c[1] + dy;
          img[1] = 0.0f;
          if (cy < 0) { cy += ny; img[1] = -a.box[1]; }
          else if (cy >= ny) { cy -= ny; img[1] = a.box[1]; }
          const int row = (cz * ny + cy) * nx;
          if (c[0] >= 1 && c[0] + 1 < nx) {
            img[0] = 0.0f;
            test_range<DIM, TPP, ORDERED, BAND>(a, s, h, ps, sid, a.cell_start[row + c[0] - 1],
                                          a.cell_start[row + c[0] + 2], img, out, stage, count);
          } else {
            for (int dx = -1; dx <= 1; ++dx) {
              int cx = c[0] + dx;
              img[0] = 0.0f;
              if (cx < 0) { cx += nx; img[0] = -a.box[0]; }
              else if (cx >= nx) { cx -= nx; img[0] = a.box[0]; }
              test_range<DIM, TPP, ORDERED, BAND>(a, s, h, ps, sid, a.cell_start[row + cx],
                                            a.cell_start[row + cx + 1], img, out, stage, count);
            }
          }
        }
      }
    }
  }
  // the particle's other lanes wrote into this thread's stage column
  __syncthreads();
  // this lane's sub-list: hits h, h + TPP, ... of the particle
  const int share = s < a.n ? (count + TPP - 1 - h) / TPP : 0;
  if (s < a.n) {
    int32_t* out = a.ell + ((int64_t)h * a.lane_width) * a.npad + s;
    const int32_t* stage = &stage_buf[0][threadIdx.x];
    int staged = share < kStage ? share : kStage;
    staged = staged < a.lane_width ? staged : a.lane_width;  // never past the sub-list
    for (int k = 0; k < staged; ++k) out[(int64_t)k * a.npad] = stage[k * kBlock];
    a.cnt[h * a.npad + s] = share < a.lane_width ? share : a.lane_width;
    if (h == 0 && count > TPP * a.lane_width) {  // some sub-list of s was truncated
      atomicAdd((unsigned long long*)&a.info[5], 1ull);
      if (a.trunc_bits) atomicOr((unsigned long long*)&a.info[0], (unsigned long long)a.trunc_bits);
    }
  }
  // block reductions: total pairs (sum of the lanes' shares) and the largest
  // per-particle count
  __shared__ unsigned long long tot[kBlock / 64];
  __shared__ int mx[kBlock / 64];
  unsigned long long v = (unsigned long long)share;
  int m = count;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    v += __shfl_down(v, o, 64);
    int y = __shfl_down(m, o, 64);
    m = y > m ? y : m;
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) { tot[wid] = v; mx[wid] = m; }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    int mm = 0;
    for (int w = 0; w < kBlock / 64; ++w) { t += tot[w]; mm = mx[w] > mm ? mx[w] : mm; }
    a.part[2 * blk] = (int64_t)t;
    a.part[2 * blk + 1] = mm;
  }
}

template <int DIM, int TPP, bool ORDERED, bool BAND>
__global__ __launch_bounds__(kBlock, ORDERED ? 1 : 6) void k_nl_build(BuildArgs a) {
  nl_build_body<DIM, TPP, ORDERED, BAND>(a, logical_block(a.remap));
}

__global__ void k_finalize(int64_t* __restrict__ info, int64_t max_occ) {
  if (threadIdx.x == 0 && info[1] > max_occ) info[0] |= SPH_NEIGHBOR_LIST_OVERFLOW;
}

// Reduces the per-workgroup (pairs, largest list) partials of one launch:
// *tot_out = sum, *max_out = max (each nullable); with max_occ >= 0 also the
// NEIGHBOR_LIST_OVERFLOW bit (partition.py:976).  With a gate, runs only when
// *gate != 0; with clear_flag, zeroes *clear_flag at the end (the skin step's
// rebuild request is consumed).
constexpr int kReduceThreads = 1024;
__device__ __forceinline__ void reduce_parts_body(const int64_t* __restrict__ part, int nblocks,
                                                  int64_t* __restrict__ tot_out,
                                                  int64_t* __restrict__ max_out,
                                                  int64_t* __restrict__ info, int64_t max_occ,
                                                  const int32_t* __restrict__ gate,
                                                  int32_t* __restrict__ clear_flag) {
  __shared__ int64_t ws_t[kReduceThreads / 64], ws_m[kReduceThreads / 64];
  const bool run = !gate || *gate != 0;  // uniform
  if (run) {
    int64_t t = 0, m = 0;
    for (int b = threadIdx.x; b < nblocks; b += kReduceThreads) {
      t += part[2 * b];
      m = part[2 * b + 1] > m ? part[2 * b + 1] : m;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      t += __shfl_down(t, o, 64);
      const int64_t y = __shfl_down(m, o, 64);
      m = y > m ? y : m;
    }
    if ((threadIdx.x & 63) == 0) { ws_t[threadIdx.x >> 6] = t; ws_m[threadIdx.x >> 6] = m; }
    __syncthreads();
    if (threadIdx.x == 0) {
      int64_t tt = 0, mm = 0;
      for (int w = 0; w < kReduceThreads / 64; ++w) {
        tt += ws_t[w];
        mm = ws_m[w] > mm ? ws_m[w] : mm;
      }
      if (tot_out) *tot_out = tt;
      if (max_out) *max_out = mm;
      if (max_occ >= 0 && tt > max_occ) info[0] |= SPH_NEIGHBOR_LIST_OVERFLOW;
    }
  }
  if (clear_flag && threadIdx.x == 0) *clear_flag = 0;
}

__global__ __launch_bounds__(kReduceThreads) void k_reduce_parts(
    const int64_t* __restrict__ part, int nblocks, int64_t* __restrict__ tot_out,
    int64_t* __restrict__ max_out, int64_t* __restrict__ info, int64_t max_occ,
    const int32_t* __restrict__ gate, int32_t* __restrict__ clear_flag) {
  reduce_parts_body(part, nblocks, tot_out, max_out, info, max_occ, gate, clear_flag);
}

template <int DIM>
__global__ void k_needs_rebuild(const float* __restrict__ r, const float* __restrict__ ref, int n,
                                Grid g, float bx, float by, float bz, float thr_sq,
                                int64_t* __restrict__ info) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float box[3] = {bx, by, bz};
  float half[3] = {bx * 0.5f, by * 0.5f, bz * 0.5f};
  float d2 = metric_sq_exact<DIM>(load_pos<DIM>(r, i), load_pos<DIM>(ref, i), box, half);
  if (d2 > thr_sq) atomicOr((unsigned long long*)&info[4], 1ull);
}

// ----------------------------------------------------------- idx writer
__global__ void k_counts_orig(const int32_t* __restrict__ perm, const int32_t* __restrict__ cnt,
                              int64_t* __restrict__ cnt64, int n, int npad, int tpp) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  int64_t c = 0;
  for (int h = 0; h < tpp; ++h) c += cnt[h * npad + s];
  cnt64[perm[s]] = c;
}

// Receivers of sender a in sub-list order (lane 0's list, then lane 1's, ...).
__global__ void k_write_idx(const int32_t* __restrict__ perm, const int32_t* __restrict__ cnt,
                            const int32_t* __restrict__ ell, const int64_t* __restrict__ off,
                            int32_t* __restrict__ idx, int64_t max_occ, int n, int npad, int tpp,
                            int lane_width) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const int a = perm[s];
  int64_t p = off[a];
  for (int h = 0; h < tpp; ++h) {
    const int c = cnt[h * npad + s];
    const int32_t* lst = ell + ((int64_t)h * lane_width) * npad + s;
    for (int k = 0; k < c && p < max_occ; ++k, ++p) {
      idx[p] = perm[lst[(int64_t)k * npad]];
      idx[max_occ + p] = a;
    }
  }
}

__global__ void k_pad_idx(int32_t* __restrict__ idx, int64_t max_occ,
                          const int64_t* __restrict__ total, int n) {
  int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < max_occ && p >= *total) {
    idx[p] = n;
    idx[max_occ + p] = n;
  }
}

// idx of a skin list: the pairs of each sender (sub-list order) that pass the
// exact reference predicate at positions r.  WRITE = false counts them per
// original id; WRITE = true stores them from the scanned offsets.
struct SkinIdxArgs {
  const float* r;
  const int32_t *perm, *cnt, *ell;
  int64_t* cnt64;
  const int64_t* off;
  int32_t* idx;
  int64_t max_occ;
  float box[3], half[3];
  float cutoff_sq;
  int n, npad, tpp, lane_width;
  // reference cells no wider than the cutoff: the pair's reference cells must
  // be adjacent (the grid of the reference's candidate search)
};

template <int DIM, bool WRITE>
__global__ void k_skin_pairs(SkinIdxArgs a) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= a.n) return;
  const int i = a.perm[s];
  const float4 ri = load_pos<DIM>(a.r, i);
  int64_t p = WRITE ? a.off[i] : 0;
  for (int h = 0; h < a.tpp; ++h) {
    const int c = a.cnt[h * a.npad + s];
    const int32_t* lst = a.ell + ((int64_t)h * a.lane_width) * a.npad + s;
    for (int k = 0; k < c; ++k) {
      const int j = a.perm[lst[(int64_t)k * a.npad]];
      if (!(metric_sq_exact<DIM>(ri, load_pos<DIM>(a.r, j), a.box, a.half) < a.cutoff_sq))
        continue;
      if (WRITE && p < a.max_occ) {
        a.idx[p] = j;
        a.idx[a.max_occ + p] = i;
      }
      ++p;
    }
  }
  if (!WRITE) a.cnt64[i] = p;
}

__global__ void k_chunk_budget(const int64_t* __restrict__ off, int n, int64_t n_sub,
                               int64_t budget, int64_t* __restrict__ info) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t lo = k * n_sub;
  if (lo >= n) return;
  int64_t hi = lo + n_sub < n ? lo + n_sub : n;
  if (off[hi] - off[lo] > budget) {
    atomicOr((unsigned long long*)&info[0], (unsigned long long)SPH_CELL_LIST_OVERFLOW);
    atomicAdd((unsigned long long*)&info[6], 1ull);
  }
}

Grid make_grid(const sph_nl_config& c) {
  Grid g;
  for (int d = 0; d < 3; ++d) {
    g.cs[d] = c.cell_size[d];
    g.cps[d] = c.cps[d] > 0 ? c.cps[d] : 1;
  }
  g.use_cells = c.use_cells;
  return g;
}

// Bins r into `bank`: hash + histogram, scan, scatter, per-cell sort, positions.
__global__ void k_bin_init(int64_t* __restrict__ info, const int64_t* __restrict__ prev,
                           int32_t* __restrict__ count, int64_t ncell) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < ncell) count[i] = 0;
  if (i == 0) {
    info[0] = prev ? prev[0] : 0;
    for (int k = 1; k < 8; ++k) info[k] = 0;
  }
}

// Resets the error block and the cell histogram.
int bin_init(sph_ws* ws, const int64_t* prev_info, int64_t* info, hipStream_t st) {
  SPH_LAUNCH(ws, "bin_init", st,
             k_bin_init<<<cdiv(ws->ncell, kBlock), kBlock, 0, st>>>(info, prev_info,
                                                                   ws->cell_count, ws->ncell));
  return SPH_OK;
}

// Kick/drift/wrap of the state fused with the histogram of the new positions.
int integrate_and_hash(sph_ws* ws, const float* r, const float* u, const float* dudt,
                       const float* dvdt, float* r_out, float* u_out, float* v_out, float dt,
                       float ctvf, const float* box, hipStream_t st) {
  const auto& c = ws->cfg;
  const int n = c.n;
  Grid g = make_grid(c);
  if (c.dim == 2)
    SPH_LAUNCH(ws, "integrate_hash", st,
               k_integrate_hash<2><<<cdiv(n, kBlock), kBlock, 0, st>>>(
                   r, u, dudt, dvdt, r_out, u_out, v_out, n, dt, ctvf, g, box[0], box[1], 0.0f,
                   ws->cell_count, ws->pcell, ws->pslot));
  else
    SPH_LAUNCH(ws, "integrate_hash", st,
               k_integrate_hash<3><<<cdiv(n, kBlock), kBlock, 0, st>>>(
                   r, u, dudt, dvdt, r_out, u_out, v_out, n, dt, ctvf, g, box[0], box[1], box[2],
                   ws->cell_count, ws->pcell, ws->pslot));
  return SPH_OK;
}

// Scan of the histogram, scatter, per-cell id sort; then (if r is given) the
// sorted positions the pair build reads.
int bin_finish(sph_ws* ws, int bank, const float* r, int64_t* info, hipStream_t st) {
  const auto& c = ws->cfg;
  const int n = c.n;
  Bank& B = ws->bank[bank];
  SPH_LAUNCH(ws, "scan_cells", st,
             (k_scan_single<int32_t, int32_t><<<1, kScanThreads, 0, st>>>(
                 ws->cell_count, ws->ncell, B.cell_start, &info[2], info,
                 c.use_cells ? (int64_t)c.cell_capacity : (int64_t)-1, ws->cell_count)));
  SPH_LAUNCH(ws, "scatter", st,
             k_scatter<<<cdiv(n, kBlock), kBlock, 0, st>>>(ws->pcell, ws->pslot, B.cell_start,
                                                         ws->unsorted, n));
  SPH_LAUNCH(ws, "cell_sort", st,
             k_cell_sort<3><<<cdiv(ws->ncell * 64, 256), 256, 0, st>>>(
                 B.cell_start, ws->unsorted, B.perm, B.iperm, ws->ncell, nullptr, nullptr,
                 Grid{}));
  if (!r) return SPH_OK;
  if (c.dim == 2)
    SPH_LAUNCH(ws, "gather_pos", st,
               k_gather_pos<2><<<cdiv(n, kBlock), kBlock, 0, st>>>(r, B.perm, B.pos, n));
  else
    SPH_LAUNCH(ws, "gather_pos", st,
               k_gather_pos<3><<<cdiv(n, kBlock), kBlock, 0, st>>>(r, B.perm, B.pos, n));
  return SPH_OK;
}

int bin_particles(sph_ws* ws, int bank, const float* r, const int64_t* prev_info, int64_t* info,
                  hipStream_t st) {
  const auto& c = ws->cfg;
  const int n = c.n;
  Grid g = make_grid(c);
  int rc = bin_init(ws, prev_info, info, st);
  if (rc) return rc;
  if (c.dim == 2)
    SPH_LAUNCH(ws, "hash_count", st,
               k_hash_count<2><<<cdiv(n, kBlock), kBlock, 0, st>>>(r, n, g, ws->cell_count,
                                                                 ws->pcell, ws->pslot, nullptr));
  else
    SPH_LAUNCH(ws, "hash_count", st,
               k_hash_count<3><<<cdiv(n, kBlock), kBlock, 0, st>>>(r, n, g, ws->cell_count,
                                                                 ws->pcell, ws->pslot, nullptr));
  return bin_finish(ws, bank, r, info, st);
}

int build_pairs(sph_ws* ws, int bank, int64_t max_occupancy, int64_t* info, hipStream_t st) {
  const auto& c = ws->cfg;
  Bank& B = ws->bank[bank];
  BuildArgs a;
  a.pos = B.pos;
  a.cell_start = B.cell_start;
  a.perm = B.perm;
  a.ell = B.ell;
  a.cnt = B.cnt;
  a.info = info;
  a.g = make_grid(c);
  for (int d = 0; d < 3; ++d) {
    a.box[d] = c.box[d];
    a.half[d] = c.box[d] * 0.5f;
  }
  a.cutoff_sq = c.cutoff_sq;
  a.n = c.n;
  a.npad = ws->npad;
  a.lane_width = ws->lane_width;
  a.mask_self = c.mask_self;
  a.part = ws->part_build;
  // 3^d x cell capacity entries per particle never truncate; a narrower
  // buffer reports a truncated sub-list as an overflow
  a.trunc_bits = ws->ell_overflow ? SPH_NEIGHBOR_LIST_OVERFLOW : 0;
  a.gate = nullptr;
  a.remap = ws->xcd_remap;
  // The cell-adjacent image is the minimum image of every pair within the
  // cutoff when each side has >= 5 cells (2 cells < box/2).  Both the fast and
  // the exact squared distance are then within a few ulp(box) per component of
  // the true one (see cut_band).
  a.fast = c.use_cells ? 1 : 0;
  for (int d = 0; d < c.dim; ++d)
    if (c.cps[d] < 5) a.fast = 0;
  cut_band(c, &a.lo, &a.hi);
  constexpr int T2 = tpp_of(2), T3 = tpp_of(3);
  const dim3 grid(xcd_grid(cdiv((int64_t)c.n * (c.dim == 2 ? T2 : T3), kBlock), a.remap));
  if (c.dim == 2) {
    if (c.ordered)
      SPH_LAUNCH(ws, "nl_build", st, (k_nl_build<2, T2, true, true><<<grid, kBlock, 0, st>>>(a)));
    else
      SPH_LAUNCH(ws, "nl_build", st, (k_nl_build<2, T2, false, true><<<grid, kBlock, 0, st>>>(a)));
  } else {
    if (c.ordered)
      SPH_LAUNCH(ws, "nl_build", st, (k_nl_build<3, T3, true, true><<<grid, kBlock, 0, st>>>(a)));
    else
      SPH_LAUNCH(ws, "nl_build", st, (k_nl_build<3, T3, false, true><<<grid, kBlock, 0, st>>>(a)));
  }
  // occupancy info[1], largest list info[3], and the overflow bit when max_occupancy >= 0
  SPH_LAUNCH(ws, "reduce_pairs", st,
             k_reduce_parts<<<1, kReduceThreads, 0, st>>>(ws->part_build, (int)grid.x, &info[1],
                                                          &info[3], info, max_occupancy,
                                                          nullptr, nullptr));
  return SPH_OK;
}

// The exact-predicate band of the pair build: squared distances (fp32, minimum
// image, explicit FMA) below lo are inside the cutoff, at or above hi outside;
// in between the reference metric decides.  The two differ by a few ulp(box)
// per component, which `band` bounds with a wide margin.
void cut_band(const sph_nl_config& c, float* lo, float* hi) {
  float maxbox = 0.0f;
  for (int d = 0; d < c.dim; ++d) maxbox = c.box[d] > maxbox ? c.box[d] : maxbox;
  const float ulp_box = std::nextafter(maxbox, INFINITY) - maxbox;
  const float ulp_c2 = std::nextafter(c.cutoff_sq, INFINITY) - c.cutoff_sq;
  const float band = 64.0f * c.cutoff * ulp_box + 16.0f * ulp_c2;
  *lo = c.cutoff_sq - band;
  *hi = c.cutoff_sq + band;
}

static Grid skin_grid(const sph_ws* ws) {
  Grid g;
  for (int d = 0; d < 3; ++d) {
    g.cs[d] = ws->sk.cell_size[d];
    g.cps[d] = ws->sk.cps[d];
  }
  g.use_cells = 1;
  return g;
}

// First half of a skin-list step: kick/drift with the rebuild test and the
// reference-grid occupancy; then, only if a rebuild was requested, the
// skin-grid binning (hash, scan, scatter, per-cell sort).
int skin_begin(sph_ws* ws, const float* r, const float* u, const float* dudt, const float* dvdt,
               float* r_out, float* u_out, float* v_out, float dt, float ctvf, const float* box,
               const int64_t* prev_info, int64_t* info, hipStream_t st) {
  const auto& c = ws->cfg;
  auto& S = ws->sk;
  const int n = c.n;
  const Grid g = make_grid(c);
  const Grid gs = skin_grid(ws);
  const unsigned nb = (unsigned)cdiv(n, kBlock);
  if (c.dim == 2)
    SPH_LAUNCH(ws, "integrate_skin", st,
               k_integrate_skin<2><<<nb, kBlock, 0, st>>>(
                   r, u, dudt, dvdt, r_out, u_out, v_out, n, dt, ctvf, g, box[0], box[1], 0.0f,
                   ws->cell_count, S.posb, S.trigger_sq, S.flag));
  else
    SPH_LAUNCH(ws, "integrate_skin", st,
               k_integrate_skin<3><<<nb, kBlock, 0, st>>>(
                   r, u, dudt, dvdt, r_out, u_out, v_out, n, dt, ctvf, g, box[0], box[1], box[2],
                   ws->cell_count, S.posb, S.trigger_sq, S.flag));
  SPH_LAUNCH(ws, "cells_max", st,
             k_cells_max<<<1, kCellsThreads, 0, st>>>(ws->cell_count, ws->ncell, info, prev_info,
                                                      (int64_t)c.cell_capacity));
  if (c.dim == 2)
    SPH_LAUNCH(ws, "skin_hash", st,
               k_hash_count<2><<<nb, kBlock, 0, st>>>(r_out, n, gs, S.cell_count, ws->pcell,
                                                      ws->pslot, S.flag));
  else
    SPH_LAUNCH(ws, "skin_hash", st,
               k_hash_count<3><<<nb, kBlock, 0, st>>>(r_out, n, gs, S.cell_count, ws->pcell,
                                                      ws->pslot, S.flag));
  SPH_LAUNCH(ws, "skin_scan", st,
             (k_scan_single<int32_t, int32_t><<<1, kScanThreads, 0, st>>>(
                 S.cell_count, S.ncell, S.b.cell_start, nullptr, nullptr, (int64_t)-1,
                 S.cell_count, S.flag)));
  SPH_LAUNCH(ws, "skin_scatter", st,
             k_scatter<<<nb, kBlock, 0, st>>>(ws->pcell, ws->pslot, S.b.cell_start, ws->unsorted,
                                              n, S.flag));
  if (c.dim == 2)
    SPH_LAUNCH(ws, "skin_sort", st,
               k_cell_sort<2><<<cdiv(S.ncell * 64, 256), 256, 0, st>>>(
                   S.b.cell_start, ws->unsorted, S.b.perm, S.b.iperm, S.ncell, S.flag, r_out, gs));
  else
    SPH_LAUNCH(ws, "skin_sort", st,
               k_cell_sort<3><<<cdiv(S.ncell * 64, 256), 256, 0, st>>>(
                   S.b.cell_start, ws->unsorted, S.b.perm, S.b.iperm, S.ncell, S.flag, r_out, gs));
  return SPH_OK;
}

// Second half: the skin pair lists (cutoff + skin) from the sorted build
// positions, only if a rebuild was requested.  Pair totals are not counted
// here (the sweeps count the pairs inside the cutoff); info[7] receives the
// largest list of one particle, and an overflowing sub-list raises
// SPH_NEIGHBOR_LIST_OVERFLOW so the caller re-allocates.
static BuildArgs skin_build_args(sph_ws* ws, int64_t* info) {
  const auto& c = ws->cfg;
  auto& S = ws->sk;
  BuildArgs a;
  a.pos = S.b.pos;
  a.cell_start = S.b.cell_start;
  a.perm = S.b.perm;
  a.ell = S.b.ell;
  a.cnt = S.b.cnt;
  a.info = info;
  a.g = skin_grid(ws);
  for (int d = 0; d < 3; ++d) {
    a.box[d] = c.box[d];
    a.half[d] = c.box[d] * 0.5f;
  }
  a.cutoff_sq = S.reach_sq;
  a.fast = 1;
  a.lo = a.hi = S.reach_sq;
  a.n = c.n;
  a.npad = ws->npad;
  a.lane_width = S.lane_width;
  a.mask_self = c.mask_self;
  a.part = ws->part_build;
  a.trunc_bits = SPH_NEIGHBOR_LIST_OVERFLOW;
  a.gate = S.flag;
  a.remap = ws->xcd_remap;
  return a;
}

int skin_build(sph_ws* ws, int64_t* info, hipStream_t st) {
  const auto& c = ws->cfg;
  auto& S = ws->sk;
  const BuildArgs a = skin_build_args(ws, info);
  constexpr int T2 = tpp_of(2), T3 = tpp_of(3);
  const dim3 grid(xcd_grid(cdiv((int64_t)c.n * (c.dim == 2 ? T2 : T3), kBlock), a.remap));
  if (c.dim == 2)
    SPH_LAUNCH(ws, "skin_build", st, (k_nl_build<2, T2, false, false><<<grid, kBlock, 0, st>>>(a)));
  else
    SPH_LAUNCH(ws, "skin_build", st, (k_nl_build<3, T3, false, false><<<grid, kBlock, 0, st>>>(a)));
  // largest skin list of one particle -> info[7] (only on a rebuild)
  SPH_LAUNCH(ws, "reduce_skin", st,
             k_reduce_parts<<<1, kReduceThreads, 0, st>>>(ws->part_build, (int)grid.x, nullptr,
                                                          &info[7], info, -1, S.flag, nullptr));
  return SPH_OK;
}

// End of a skin-list step: occupancy and largest list from the counting
// density sweep's partials, the overflow bit, and the rebuild request is consumed.
int skin_finish(sph_ws* ws, int64_t max_occupancy, int64_t* info, hipStream_t st) {
  const int64_t blocks = xcd_grid(cdiv((int64_t)ws->cfg.n * ws->tpp, kBlock), ws->xcd_remap);
  SPH_LAUNCH(ws, "reduce_pairs", st,
             k_reduce_parts<<<1, kReduceThreads, 0, st>>>(ws->part_sweep, (int)blocks, &info[1],
                                                          &info[3], info, max_occupancy,
                                                          nullptr, ws->sk.flag));
  return SPH_OK;
}

// pair counts per original id and their exclusive scan (off64[0..N])
int pair_offsets(sph_ws* ws, int bank, hipStream_t st) {
  const int n = ws->cfg.n;
  Bank& B = ws->bank[bank];
  SPH_LAUNCH(ws, "counts_orig", st,
             k_counts_orig<<<cdiv(n, kBlock), kBlock, 0, st>>>(B.perm, B.cnt, ws->cnt64, n,
                                                             ws->npad, ws->tpp));
  SPH_LAUNCH(ws, "scan_pairs", st,
             (k_scan_single<int64_t, int64_t><<<1, kScanThreads, 0, st>>>(
                 ws->cnt64, n, ws->off64, nullptr, nullptr, -1)));
  return SPH_OK;
}

// ------------------------------------------------------------ ensembles
// Batched forms of the skin-step phases above (skin_begin, skin_build,
// skin_finish): one launch serves every member.  Per-member arguments are
// those the single-simulation launches pass; the bodies are shared.
struct EnsInteg {
  const float *r, *u, *dudt, *dvdt;
  float *r_out, *u_out, *v_out;
  int n;
  float dt, ctvf;
  Grid g;
  float box[3];
  int32_t* cell_count;
  const float4* posb;
  float trigger_sq;
  int32_t* flag;
};

struct EnsCells {
  int32_t* count;
  int64_t ncell;
  int64_t* info;
  const int64_t* prev;
  int64_t capacity;
};

struct EnsBin {  // skin-grid binning, all gated on the member's rebuild flag
  const float* r;
  int n;
  Grid g;
  int32_t *cell_count, *pcell, *pslot, *flag, *cell_start;
  int64_t ncell;
  int32_t *unsorted, *perm, *iperm;
};

struct EnsReduce {
  const int64_t* part;
  int nblocks;
  int64_t *tot_out, *max_out, *info;
  int64_t max_occ;
  const int32_t* gate;
  int32_t* clear_flag;
};

struct EnsNL {
  int n = 0;
  EnsInteg* integ = nullptr;
  EnsCells* cells = nullptr;
  EnsBin* bin = nullptr;
  BuildArgs* build = nullptr;
  EnsReduce* red_build = nullptr;
  EnsReduce* red_sweep = nullptr;
};

template <int DIM>
__global__ void k_integrate_skin_ens(const EnsInteg* __restrict__ t,
                                     const EnsBlock* __restrict__ map) {
  const EnsBlock m = map[blockIdx.x];
  if (m.sim < 0) return;
  const EnsInteg& a = t[m.sim];
  integrate_skin_body<DIM>(a.r, a.u, a.dudt, a.dvdt, a.r_out, a.u_out, a.v_out, a.n, a.dt,
                           a.ctvf, a.g, a.box[0], a.box[1], a.box[2], a.cell_count, a.posb,
                           a.trigger_sq, a.flag, m.blk);
}

// one workgroup per member
__global__ __launch_bounds__(kCellsThreads) void k_cells_max_ens(const EnsCells* __restrict__ t) {
  const EnsCells& a = t[blockIdx.x];
  cells_max_body(a.count, a.ncell, a.info, a.prev, a.capacity);
}

template <int DIM>
__global__ void k_hash_count_ens(const EnsBin* __restrict__ t, const EnsBlock* __restrict__ map) {
  const EnsBlock m = map[blockIdx.x];
  if (m.sim < 0) return;
  const EnsBin& a = t[m.sim];
  hash_count_body<DIM>(a.r, a.n, a.g, a.cell_count, a.pcell, a.pslot, a.flag, m.blk);
}

// one workgroup per member
__global__ __launch_bounds__(kScanThreads) void k_scan_ens(const EnsBin* __restrict__ t) {
  const EnsBin& a = t[blockIdx.x];
  scan_single_body<int32_t, int32_t>(a.cell_count, a.ncell, a.cell_start, nullptr, nullptr, -1,
                                     a.cell_count, a.flag);
}

__global__ void k_scatter_ens(const EnsBin* __restrict__ t, const EnsBlock* __restrict__ map) {
  const EnsBlock m = map[blockIdx.x];
  if (m.sim < 0) return;
  const EnsBin& a = t[m.sim];
  scatter_body(a.pcell, a.pslot, a.cell_start, a.unsorted, a.n, a.flag, m.blk);
}

template <int DIM>
__global__ __launch_bounds__(256) void k_cell_sort_ens(const EnsBin* __restrict__ t,
                                                       const EnsBlock* __restrict__ map) {
  const EnsBlock m = map[blockIdx.x];
  if (m.sim < 0) return;
  const EnsBin& a = t[m.sim];
  cell_sort_body<DIM>(a.cell_start, a.unsorted, a.perm, a.iperm, a.ncell, a.flag, m.blk, a.r,
                      a.g);
}

template <int DIM, int TPP>
__global__ __launch_bounds__(kBlock, 6) void k_nl_build_ens(const BuildArgs* __restrict__ t,
                                                         const EnsBlock* __restrict__ map) {
  const EnsBlock m = map[blockIdx.x];
  if (m.sim < 0) return;
  nl_build_body<DIM, TPP, false, false>(t[m.sim], m.blk);
}

// one workgroup per member
__global__ __launch_bounds__(kReduceThreads) void k_reduce_parts_ens(
    const EnsReduce* __restrict__ t) {
  const EnsReduce& a = t[blockIdx.x];
  reduce_parts_body(a.part, a.nblocks, a.tot_out, a.max_out, a.info, a.max_occ, a.gate,
                    a.clear_flag);
}

template <typename T>
static int upload(const std::vector<T>& h, T** d) {
  hipError_t e = hipMalloc((void**)d, sizeof(T) * h.size());
  if (e != hipSuccess) return hip_fail(e, "ensemble tables");
  e = hipMemcpy(*d, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice);
  if (e != hipSuccess) return hip_fail(e, "ensemble tables");
  return SPH_OK;
}

int ens_nl_create(const std::vector<EnsMember>& members, EnsNL** out) {
  const int n = (int)members.size();
  std::vector<EnsInteg> integ(n);
  std::vector<EnsCells> cells(n);
  std::vector<EnsBin> bin(n);
  std::vector<BuildArgs> build(n);
  std::vector<EnsReduce> rb(n), rs(n);
  for (int k = 0; k < n; ++k) {
    const EnsMember& m = members[k];
    sph_ws* ws = m.ws;
    const auto& c = ws->cfg;
    auto& S = ws->sk;
    const int64_t lane_blocks = cdiv((int64_t)c.n * ws->tpp, kBlock);
    integ[k] = EnsInteg{m.in.r, m.in.u, m.in.dudt, m.in.dvdt, m.r_out, m.out.u, m.out.v, c.n,
                        m.dt, m.ctvf, make_grid(c), {m.prm.box[0], m.prm.box[1], m.prm.box[2]},
                        ws->cell_count, S.posb, S.trigger_sq, S.flag};
    cells[k] = EnsCells{ws->cell_count, ws->ncell, m.info, m.prev_info,
                        (int64_t)c.cell_capacity};
    bin[k] = EnsBin{m.r_out, c.n, skin_grid(ws), S.cell_count, ws->pcell, ws->pslot, S.flag,
                    S.b.cell_start, S.ncell, ws->unsorted, S.b.perm, S.b.iperm};
    build[k] = skin_build_args(ws, m.info);
    rb[k] = EnsReduce{ws->part_build, (int)lane_blocks, nullptr, &m.info[7], m.info, -1, S.flag,
                      nullptr};
    rs[k] = EnsReduce{ws->part_sweep, (int)lane_blocks, &m.info[1], &m.info[3], m.info,
                      m.max_occupancy, nullptr, S.flag};
  }
  EnsNL* t = new EnsNL;
  t->n = n;
  int rc = upload(integ, &t->integ);
  if (!rc) rc = upload(cells, &t->cells);
  if (!rc) rc = upload(bin, &t->bin);
  if (!rc) rc = upload(build, &t->build);
  if (!rc) rc = upload(rb, &t->red_build);
  if (!rc) rc = upload(rs, &t->red_sweep);
  if (rc) {
    ens_nl_destroy(t);
    return rc;
  }
  *out = t;
  return SPH_OK;
}

void ens_nl_destroy(EnsNL* t) {
  if (!t) return;
  (void)hipFree(t->integ);
  (void)hipFree(t->cells);
  (void)hipFree(t->bin);
  (void)hipFree(t->build);
  (void)hipFree(t->red_build);
  (void)hipFree(t->red_sweep);
  delete t;
}

int ens_skin_begin(EnsNL* t, int dim, const EnsMap& part, const EnsMap& cells, sph_ws* prof,
                   hipStream_t st) {
  if (dim == 2)
    SPH_LAUNCH(prof, "ens_integrate_skin", st,
               k_integrate_skin_ens<2><<<part.blocks, kBlock, 0, st>>>(t->integ, part.d));
  else
    SPH_LAUNCH(prof, "ens_integrate_skin", st,
               k_integrate_skin_ens<3><<<part.blocks, kBlock, 0, st>>>(t->integ, part.d));
  SPH_LAUNCH(prof, "ens_cells_max", st,
             k_cells_max_ens<<<t->n, kCellsThreads, 0, st>>>(t->cells));
  if (dim == 2)
    SPH_LAUNCH(prof, "ens_skin_hash", st,
               k_hash_count_ens<2><<<part.blocks, kBlock, 0, st>>>(t->bin, part.d));
  else
    SPH_LAUNCH(prof, "ens_skin_hash", st,
               k_hash_count_ens<3><<<part.blocks, kBlock, 0, st>>>(t->bin, part.d));
  SPH_LAUNCH(prof, "ens_skin_scan", st, k_scan_ens<<<t->n, kScanThreads, 0, st>>>(t->bin));
  SPH_LAUNCH(prof, "ens_skin_scatter", st,
             k_scatter_ens<<<part.blocks, kBlock, 0, st>>>(t->bin, part.d));
  if (dim == 2)
    SPH_LAUNCH(prof, "ens_skin_sort", st,
               k_cell_sort_ens<2><<<cells.blocks, 256, 0, st>>>(t->bin, cells.d));
  else
    SPH_LAUNCH(prof, "ens_skin_sort", st,
               k_cell_sort_ens<3><<<cells.blocks, 256, 0, st>>>(t->bin, cells.d));
  return SPH_OK;
}

int ens_skin_build(EnsNL* t, int dim, const EnsMap& lanes, sph_ws* prof, hipStream_t st) {
  constexpr int T2 = tpp_of(2), T3 = tpp_of(3);
  if (dim == 2)
    SPH_LAUNCH(prof, "ens_skin_build", st,
               (k_nl_build_ens<2, T2><<<lanes.blocks, kBlock, 0, st>>>(t->build, lanes.d)));
  else
    SPH_LAUNCH(prof, "ens_skin_build", st,
               (k_nl_build_ens<3, T3><<<lanes.blocks, kBlock, 0, st>>>(t->build, lanes.d)));
  SPH_LAUNCH(prof, "ens_reduce_skin", st,
             k_reduce_parts_ens<<<t->n, kReduceThreads, 0, st>>>(t->red_build));
  return SPH_OK;
}

int ens_skin_finish(EnsNL* t, sph_ws* prof, hipStream_t st) {
  SPH_LAUNCH(prof, "ens_reduce_pairs", st,
             k_reduce_parts_ens<<<t->n, kReduceThreads, 0, st>>>(t->red_sweep));
  return SPH_OK;
}

}  // namespace sph

using namespace sph;

static int check_ws(sph_ws* ws, int bank) {
  if (!ws) { set_error("null workspace"); return SPH_ERR_ARG; }
  if (bank < 0 || bank > 1) { set_error("bank must be 0 or 1"); return SPH_ERR_ARG; }
  return SPH_OK;
}

namespace sph {
// Grid of a cell-occupancy count; false (with the error set) on a bad argument.
static bool fill_grid(const char* who, int32_t dim, int32_t n, const float* r,
                      const float* cell_size, const int32_t* cps, const int64_t* info, Grid* g,
                      int64_t* nc) {
  if (!r || !cell_size || !cps || !info || n < 0 || (dim != 2 && dim != 3)) {
    set_error(std::string(who) + ": bad argument");
    return false;
  }
  *nc = 1;
  for (int d = 0; d < 3; ++d) {
    g->cs[d] = d < dim ? cell_size[d] : 1.0f;
    g->cps[d] = d < dim ? cps[d] : 1;
    if (g->cps[d] < 1) { set_error("cps must be >= 1"); return false; }
    *nc *= g->cps[d];
  }
  g->use_cells = 1;
  return true;
}

// Clears `counts` (nc cells), histograms r into it and writes the largest
// count to info[2]; every buffer is cleared by a kernel, none by a memset.
static int fill_max_run(int32_t dim, int32_t n, const float* r, const Grid& g, int64_t nc,
                        int32_t* counts, int64_t* info, hipStream_t st) {
  sph_ws* ws = nullptr;  // no profiling without a workspace
  SPH_LAUNCH(ws, "fill", st, k_fill<int32_t><<<cdiv(nc, kBlock), kBlock, 0, st>>>(counts, nc, 0));
  SPH_LAUNCH(ws, "fill", st, k_fill<int64_t><<<1, 64, 0, st>>>(info + 2, 1, 0));
  if (n > 0) {
    if (dim == 2)
      SPH_LAUNCH(ws, "hash_count", st,
                 k_hash_count<2><<<cdiv(n, kBlock), kBlock, 0, st>>>(r, n, g, counts, nullptr,
                                                                   nullptr, nullptr));
    else
      SPH_LAUNCH(ws, "hash_count", st,
                 k_hash_count<3><<<cdiv(n, kBlock), kBlock, 0, st>>>(r, n, g, counts, nullptr,
                                                                   nullptr, nullptr));
  }
  int64_t blocks = cdiv(nc, kBlock);
  blocks = blocks > 1024 ? 1024 : blocks;
  SPH_LAUNCH(ws, "max_reduce", st, k_max_reduce<<<blocks, kBlock, 0, st>>>(counts, nc, info + 2));
  return SPH_OK;
}

// Cell counts of sph_nl_fill_max_static: one copy per GPU, held until the
// process ends.  Host code never passes this name to a kernel (on the host it
// is a stand-in, not a device address); it uses the per-device address that
// hipGetSymbolAddress returns.
static __device__ __align__(256) int32_t scratch[SPH_FILL_MAX_STATIC_CELLS];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static int32_t* g_scratch_ptr[kMaxDevices];  // nullptr until looked up on that device

// Device address of `scratch` on the current device, looked up once per device.
static int scratch_of_current_device(int32_t** out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
    set_error("sph_nl_fill_max_static: no current HIP device");
    return SPH_ERR_HIP;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
      set_error(std::string("sph_nl_fill_max_static: hipGetSymbolAddress: ") +
                hipGetErrorString(e));
      return SPH_ERR_HIP;
    }
    g_scratch_ptr[dev] = static_cast<int32_t*>(ptr);
  }
  *out = g_scratch_ptr[dev];
  return SPH_OK;
}

// ----------------------------------------------------------- dense cell list
// Slot of the particle at sorted position s (cell-major, ascending id inside a
// cell): cell * capacity + s mod capacity, the scatter index of
// partition.py:368-385.  A cell's particles take consecutive slots from its
// first sorted position mod capacity, wrapping inside the cell, so up to
// `capacity` of them never collide.  In an overflowing cell the particle
// capacity places later in the cell takes the slot (the later scatter update
// wins), so every slot has exactly one writer.
__device__ __forceinline__ int64_t cell_list_slot(const int32_t* __restrict__ perm,
                                                  const int32_t* __restrict__ pcell,
                                                  const int32_t* __restrict__ cell_start, int s,
                                                  int64_t cap, int* id) {
  *id = perm[s];
  const int h = pcell[*id];
  if ((int64_t)s + cap < (int64_t)cell_start[h + 1]) return -1;  // overwritten later
  return (int64_t)h * cap + s % cap;
}

template <int DIM>
__global__ void k_cell_list_fill(const float* __restrict__ r, const int32_t* __restrict__ perm,
                                 const int32_t* __restrict__ pcell,
                                 const int32_t* __restrict__ cell_start, int n, int64_t cap,
                                 float* __restrict__ pos_buf, int32_t* __restrict__ id_buf) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  int id;
  const int64_t slot = cell_list_slot(perm, pcell, cell_start, s, cap, &id);
  if (slot < 0) return;
  id_buf[slot] = id;
#pragma unroll
  for (int d = 0; d < DIM; ++d) pos_buf[slot * DIM + d] = r[(int64_t)id * DIM + d];
}

// One 32-bit word of a per-particle side array (`width` words per particle).
__global__ void k_cell_list_named(const uint32_t* __restrict__ src, int width,
                                  const int32_t* __restrict__ perm,
                                  const int32_t* __restrict__ pcell,
                                  const int32_t* __restrict__ cell_start, int n, int64_t cap,
                                  uint32_t* __restrict__ dst) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)n * width) return;
  const int s = (int)(t / width), w = (int)(t - (int64_t)s * width);
  int id;
  const int64_t slot = cell_list_slot(perm, pcell, cell_start, s, cap, &id);
  if (slot < 0) return;
  dst[slot * width + w] = src[(int64_t)id * width + w];
}
}  // namespace sph

extern "C" int sph_nl_fill_max(int32_t dim, int32_t n, const float* r, const float* cell_size,
                               const int32_t* cps, int32_t* counts, int64_t* info,
                               hipStream_t st) {
  Grid g;
  int64_t nc = 0;
  if (!counts) { set_error("sph_nl_fill_max: bad argument"); return SPH_ERR_ARG; }
  if (!fill_grid("sph_nl_fill_max", dim, n, r, cell_size, cps, info, &g, &nc)) return SPH_ERR_ARG;
  return fill_max_run(dim, n, r, g, nc, counts, info, st);
}

extern "C" int sph_nl_fill_max_static(int32_t dim, int32_t n, const float* r,
                                      const float* cell_size, const int32_t* cps, int64_t* info,
                                      hipStream_t st) {
  Grid g;
  int64_t nc = 0;
  if (!fill_grid("sph_nl_fill_max_static", dim, n, r, cell_size, cps, info, &g, &nc))
    return SPH_ERR_ARG;
  if (nc > SPH_FILL_MAX_STATIC_CELLS) {
    set_error("sph_nl_fill_max_static: the grid has " + std::to_string(nc) +
              " cells, more than the " + std::to_string(SPH_FILL_MAX_STATIC_CELLS) +
              " its static scratch holds (use sph_nl_fill_max with caller scratch)");
    return SPH_ERR_SIZE;
  }
  int32_t* counts = nullptr;
  const int rc = scratch_of_current_device(&counts);
  if (rc != SPH_OK) return rc;
  return fill_max_run(dim, n, r, g, nc, counts, info, st);
}

extern "C" size_t sph_cell_list_work_bytes(int64_t n, int64_t cell_count) {
  if (n < 0 || cell_count < 1) return 0;
  // counts, cell starts, cell and arrival slot per particle, arrivals, sorted ids, inverse
  return (size_t)(2 * cell_count + 1 + 5 * n) * sizeof(int32_t);
}

extern "C" int sph_cell_list(int32_t dim, int32_t n, const float* r, const float* cell_size,
                             const int32_t* cps, int64_t capacity, float* position_buffer,
                             int32_t* id_buffer, int32_t n_named, const uint32_t* const* named,
                             const int32_t* named_width, const uint32_t* named_fill,
                             uint32_t* const* named_buffer, void* work, size_t work_bytes,
                             int64_t* info, hipStream_t st) {
  Grid g;
  int64_t nc = 0;
  if (!fill_grid("sph_cell_list", dim, n, r, cell_size, cps, info, &g, &nc)) return SPH_ERR_ARG;
  if (!position_buffer || !id_buffer || !work || n_named < 0 ||
      (n_named > 0 && (!named || !named_width || !named_fill || !named_buffer))) {
    set_error("sph_cell_list: bad argument");
    return SPH_ERR_ARG;
  }
  if (nc > INT32_MAX) { set_error("sph_cell_list: more than 2^31 - 1 cells"); return SPH_ERR_SIZE; }
  if (capacity < 1 || nc * capacity > INT64_MAX / 4 / dim) {
    set_error("sph_cell_list: capacity must be >= 1 and the buffers addressable");
    return SPH_ERR_ARG;
  }
  if (work_bytes < sph_cell_list_work_bytes(n, nc)) {
    set_error("sph_cell_list: work buffer smaller than sph_cell_list_work_bytes");
    return SPH_ERR_SIZE;
  }
  for (int b = 0; b < n_named; ++b)
    if (!named[b] || !named_buffer[b] || named_width[b] < 1) {
      set_error("sph_cell_list: bad side array " + std::to_string(b));
      return SPH_ERR_ARG;
    }
  int32_t* counts = static_cast<int32_t*>(work);
  int32_t* start = counts + nc;
  int32_t* pcell = start + nc + 1;
  int32_t* pslot = pcell + n;
  int32_t* unsorted = pslot + n;
  int32_t* perm = unsorted + n;
  int32_t* iperm = perm + n;
  const int64_t slots = nc * capacity;
  sph_ws* ws = nullptr;  // no profiling without a workspace
  // empty slots: position 0, id N, side arrays their fill word (partition.py:334-358)
  SPH_LAUNCH(ws, "fill", st,
             k_fill<float><<<cdiv(slots * dim, kBlock), kBlock, 0, st>>>(position_buffer,
                                                                        slots * dim, 0.0f));
  SPH_LAUNCH(ws, "fill", st,
             k_fill<int32_t><<<cdiv(slots, kBlock), kBlock, 0, st>>>(id_buffer, slots, n));
  for (int b = 0; b < n_named; ++b) {
    const int64_t words = slots * named_width[b];
    SPH_LAUNCH(ws, "fill", st,
               k_fill<uint32_t><<<cdiv(words, kBlock), kBlock, 0, st>>>(named_buffer[b], words,
                                                                       named_fill[b]));
  }
  SPH_LAUNCH(ws, "fill", st, k_fill<int32_t><<<cdiv(nc, kBlock), kBlock, 0, st>>>(counts, nc, 0));
  if (n > 0) {
    if (dim == 2)
      SPH_LAUNCH(ws, "hash_count", st,
                 k_hash_count<2><<<cdiv(n, kBlock), kBlock, 0, st>>>(r, n, g, counts, pcell,
                                                                   pslot, nullptr));
    else
      SPH_LAUNCH(ws, "hash_count", st,
                 k_hash_count<3><<<cdiv(n, kBlock), kBlock, 0, st>>>(r, n, g, counts, pcell,
                                                                   pslot, nullptr));
  }
  // cell starts, the largest occupancy into info[2], CELL_LIST_OVERFLOW into info[0]
  SPH_LAUNCH(ws, "scan_cells", st,
             (k_scan_single<int32_t, int32_t><<<1, kScanThreads, 0, st>>>(
                 counts, nc, start, &info[2], info, capacity)));
  if (n == 0) return SPH_OK;
  SPH_LAUNCH(ws, "scatter", st,
             k_scatter<<<cdiv(n, kBlock), kBlock, 0, st>>>(pcell, pslot, start, unsorted, n));
  // ascending id inside each cell: the order of the reference's stable argsort
  SPH_LAUNCH(ws, "cell_sort", st,
             k_cell_sort<3><<<cdiv(nc * 64, 256), 256, 0, st>>>(start, unsorted, perm, iperm, nc,
                                                               nullptr, nullptr, Grid{}));
  if (dim == 2)
    SPH_LAUNCH(ws, "cell_list_fill", st,
               k_cell_list_fill<2><<<cdiv(n, kBlock), kBlock, 0, st>>>(
                   r, perm, pcell, start, n, capacity, position_buffer, id_buffer));
  else
    SPH_LAUNCH(ws, "cell_list_fill", st,
               k_cell_list_fill<3><<<cdiv(n, kBlock), kBlock, 0, st>>>(
                   r, perm, pcell, start, n, capacity, position_buffer, id_buffer));
  for (int b = 0; b < n_named; ++b)
    SPH_LAUNCH(ws, "cell_list_named", st,
               k_cell_list_named<<<cdiv((int64_t)n * named_width[b], kBlock), kBlock, 0, st>>>(
                   named[b], named_width[b], perm, pcell, start, n, capacity, named_buffer[b]));
  return SPH_OK;
}

extern "C" int sph_nl_needs_rebuild(sph_ws* ws, const float* r, const float* r_ref,
                                    float threshold_sq, int64_t* info, hipStream_t st) {
  if (!ws || !r || !r_ref || !info) { set_error("sph_nl_needs_rebuild: null"); return SPH_ERR_ARG; }
  const auto& c = ws->cfg;
  Grid g = make_grid(c);
  SPH_LAUNCH(ws, "fill", st, k_fill<int64_t><<<1, 64, 0, st>>>(info + 4, 1, 0));
  if (c.dim == 2)
    SPH_LAUNCH(ws, "needs_rebuild", st,
               k_needs_rebuild<2><<<cdiv(c.n, kBlock), kBlock, 0, st>>>(
                   r, r_ref, c.n, g, c.box[0], c.box[1], c.box[2], threshold_sq, info));
  else
    SPH_LAUNCH(ws, "needs_rebuild", st,
               k_needs_rebuild<3><<<cdiv(c.n, kBlock), kBlock, 0, st>>>(
                   r, r_ref, c.n, g, c.box[0], c.box[1], c.box[2], threshold_sq, info));
  return SPH_OK;
}

extern "C" int sph_nl_build(sph_ws* ws, int bank, const float* r, const int64_t* prev_info,
                            int64_t max_occupancy, int64_t* info, hipStream_t st) {
  int rc = check_ws(ws, bank);
  if (rc) return rc;
  if (!r || !info) { set_error("sph_nl_build: null"); return SPH_ERR_ARG; }
  rc = bin_particles(ws, bank, r, prev_info, info, st);
  if (rc) return rc;
  return build_pairs(ws, bank, max_occupancy, info, st);
}

extern "C" int sph_nl_finalize(sph_ws* ws, int64_t max_occupancy, int64_t* info, hipStream_t st) {
  if (!ws || !info) { set_error("sph_nl_finalize: null"); return SPH_ERR_ARG; }
  SPH_LAUNCH(ws, "finalize", st, k_finalize<<<1, 64, 0, st>>>(info, max_occupancy));
  return SPH_OK;
}

extern "C" int sph_nl_scan_budget(sph_ws* ws, int bank, int64_t n_sub, int64_t budget,
                                  int64_t* info, hipStream_t st) {
  int rc = check_ws(ws, bank);
  if (rc) return rc;
  if (n_sub <= 0 || !info) { set_error("sph_nl_scan_budget: bad argument"); return SPH_ERR_ARG; }
  rc = pair_offsets(ws, bank, st);
  if (rc) return rc;
  int64_t chunks = cdiv(ws->cfg.n, n_sub);
  SPH_LAUNCH(ws, "chunk_budget", st,
             k_chunk_budget<<<cdiv(chunks, kBlock), kBlock, 0, st>>>(ws->off64, ws->cfg.n, n_sub,
                                                                     budget, info));
  return SPH_OK;
}

extern "C" int sph_nl_write_idx(sph_ws* ws, int bank, int32_t* idx, int64_t max_occupancy,
                                const int64_t* info, hipStream_t st) {
  int rc = check_ws(ws, bank);
  if (rc) return rc;
  if (!idx || max_occupancy < 0) { set_error("sph_nl_write_idx: bad argument"); return SPH_ERR_ARG; }
  if (max_occupancy == 0) return SPH_OK;
  rc = pair_offsets(ws, bank, st);
  if (rc) return rc;
  const int n = ws->cfg.n;
  Bank& B = ws->bank[bank];
  SPH_LAUNCH(ws, "write_idx", st,
             k_write_idx<<<cdiv(n, kBlock), kBlock, 0, st>>>(B.perm, B.cnt, B.ell, ws->off64, idx,
                                                           max_occupancy, n, ws->npad, ws->tpp,
                                                           ws->lane_width));
  SPH_LAUNCH(ws, "pad_idx", st,
             k_pad_idx<<<cdiv(max_occupancy, kBlock), kBlock, 0, st>>>(idx, max_occupancy,
                                                                      ws->off64 + n, n));
  (void)info;
  return SPH_OK;
}

extern "C" int sph_nl_write_idx_skin(sph_ws* ws, const float* r, int32_t* idx,
                                     int64_t max_occupancy, hipStream_t st) {
  if (!ws || !r || !idx || max_occupancy < 0) {
    set_error("sph_nl_write_idx_skin: bad argument");
    return SPH_ERR_ARG;
  }
  if (!ws->sk.on || !ws->sk.stepped) {
    set_error("sph_nl_write_idx_skin: no skin lists built by sph_step on this workspace");
    return SPH_ERR_ARG;
  }
  if (max_occupancy == 0) return SPH_OK;
  const auto& c = ws->cfg;
  const int n = c.n;
  SkinIdxArgs a;
  a.r = r;
  a.perm = ws->sk.b.perm;
  a.cnt = ws->sk.b.cnt;
  a.ell = ws->sk.b.ell;
  a.cnt64 = ws->cnt64;
  a.off = ws->off64;
  a.idx = idx;
  a.max_occ = max_occupancy;
  for (int d = 0; d < 3; ++d) {
    a.box[d] = c.box[d];
    a.half[d] = c.box[d] * 0.5f;
  }
  a.cutoff_sq = c.cutoff_sq;
  a.n = n;
  a.npad = ws->npad;
  a.tpp = ws->tpp;
  a.lane_width = ws->sk.lane_width;
  const unsigned nb = (unsigned)cdiv(n, kBlock);
  if (c.dim == 2)
    SPH_LAUNCH(ws, "skin_idx_count", st, (k_skin_pairs<2, false><<<nb, kBlock, 0, st>>>(a)));
  else
    SPH_LAUNCH(ws, "skin_idx_count", st, (k_skin_pairs<3, false><<<nb, kBlock, 0, st>>>(a)));
  SPH_LAUNCH(ws, "scan_pairs", st,
             (k_scan_single<int64_t, int64_t><<<1, kScanThreads, 0, st>>>(
                 ws->cnt64, n, ws->off64, nullptr, nullptr, -1)));
  if (c.dim == 2)
    SPH_LAUNCH(ws, "skin_idx_write", st, (k_skin_pairs<2, true><<<nb, kBlock, 0, st>>>(a)));
  else
    SPH_LAUNCH(ws, "skin_idx_write", st, (k_skin_pairs<3, true><<<nb, kBlock, 0, st>>>(a)));
  SPH_LAUNCH(ws, "pad_idx", st,
             k_pad_idx<<<cdiv(max_occupancy, kBlock), kBlock, 0, st>>>(idx, max_occupancy,
                                                                      ws->off64 + n, n));
  return SPH_OK;
}
