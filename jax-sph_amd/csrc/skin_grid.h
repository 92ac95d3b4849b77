// Host-side check that a reference grid keeps every pair inside the cutoff in
// adjacent cells, the condition under which the Verlet-skin lists (filtered by
// distance alone) equal the reference's cell-list search.  Plain C++ (no HIP),
// so that the CPU tests compile it on its own.
#pragma once

#include <stdint.h>

#include <cmath>
#include <cstring>
#include <vector>

namespace sph {

// Reference cell of coordinate x >= 0 along one axis: int32(fl32(x / cell))
// (partition.py:367), not clamped (cps and above: hashed out of the grid).
inline int ref_cell_1d(float x, float cell) {
  const int q = (int)(float)((double)x / (double)cell);
  return q < 0 ? 0 : q;
}

// The reference's periodic displacement along one axis in float32
// (jax_md/space.py:170-181), as disp_exact evaluates it on the device.
inline float ref_disp_1d(float a, float b, float side) {
#pragma clang fp contract(off)
  const float half = side * 0.5f;
  const float t = (a - b) + half;
  float m = t;
  if (t >= side) m = t - side;
  if (t < 0.0f) m = t + side;
  return m - half;
}

inline uint32_t f2u(float x) { uint32_t u; std::memcpy(&u, &x, 4); return u; }
inline float u2f(uint32_t u) { float x; std::memcpy(&x, &u, 4); return x; }

// Whether, along one periodic axis, every pair of coordinates in [0, side)
// whose squared displacement is below cutoff_sq lies in the same or adjacent
// reference cells.  The squared 1D displacement bounds the summed metric from
// below (adding non-negative terms never rounds below an addend), so a pair
// in cells two apart along any axis can be inside the cutoff only if this
// test fails for that axis.  Coordinates two cells apart are at least one
// whole cell apart, so the candidates closest to each other are the floats at
// the facing edges of the two cells: a cell's lowest and highest floats (non-
// negative floats order as their bit patterns), probed 4 floats inwards, in
// both sender/receiver orders and both ways round the box.  Cells three or
// more apart lie two cells apart at least.  Coordinates that the reference
// hashes out of the grid (int32(fl32(x / cell)) == cps: r == side, or a float
// just below side whose quotient rounds up) are left out, as the reference
// has no cell for them.
inline bool cutoff_pairs_in_adjacent_cells(float side, float cell, int cps, float cutoff_sq) {
  if (cps < 3) return false;
  if (cps == 3) return true;  // every cell is adjacent to every other
  if (!((double)cell * 2.0 > std::sqrt((double)cutoff_sq) * 1.001)) return false;
  std::vector<uint32_t> lo(cps + 1);
  for (int k = 0; k <= cps; ++k) {  // lowest float of [0, side) in cell >= k, else side
    uint32_t a = 0, b = f2u(side);
    while (a < b) {
      const uint32_t m = a + (b - a) / 2;
      if (ref_cell_1d(u2f(m), cell) >= k) b = m; else a = m + 1;
    }
    lo[k] = a;
  }
  for (int k = 0; k < cps; ++k) {
    if (lo[k + 1] <= lo[k]) return false;  // an empty cell: no edges to probe
  }
  constexpr int kProbe = 4;
  for (int k = 0; k < cps; ++k) {
    const int m = (k + 2) % cps;
    float ca[2 * kProbe], cb[2 * kProbe];
    int na = 0, nb = 0;
    for (int i = 0; i < kProbe; ++i) {
      if (lo[k] + i < lo[k + 1]) ca[na++] = u2f(lo[k] + i);
      if (lo[k + 1] - 1 - i >= lo[k]) ca[na++] = u2f(lo[k + 1] - 1 - i);
      if (lo[m] + i < lo[m + 1]) cb[nb++] = u2f(lo[m] + i);
      if (lo[m + 1] - 1 - i >= lo[m]) cb[nb++] = u2f(lo[m + 1] - 1 - i);
    }
    for (int i = 0; i < na; ++i) {
      for (int j = 0; j < nb; ++j) {
        const float d1 = ref_disp_1d(ca[i], cb[j], side);
        const float d2 = ref_disp_1d(cb[j], ca[i], side);
        if (d1 * d1 < cutoff_sq || d2 * d2 < cutoff_sq) return false;
      }
    }
  }
  return true;
}

}  // namespace sph
