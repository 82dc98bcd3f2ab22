// The f64 neighbour grid: the device replacement for the CPU executor's KD-tree
// (PointTree, crates/sherd-core/src/spatial/kdtree.rs) on its bounded queries.
//
// Points are bucketed by cell (side >= the largest search reach), sorted by cell key, and the cells
// are found through an open-addressed hash table. A query scans the 27 cells around it and keeps
// the global nearest with kiddo's arithmetic ((dx^2 + dy^2) + dz^2 in f64) and tie rule (lowest
// index), so the answer does not depend on the cell layout: any point within the reach of a query
// lies in those 27 cells.
#pragma once

#include "common.hpp"

namespace sherd {

constexpr uint64_t GRID_EMPTY = ~0ull;
constexpr int GRID_BITS = 21;  // per axis in the packed key
constexpr int64_t GRID_MAXC = (1ll << 20);  // cell coordinates stay below this

// One slot of the open-addressed cell table: the packed cell key (GRID_EMPTY when free) and the
// cell's run of sorted points, read together in one 16-byte load.
struct alignas(16) GridSlot {
  uint64_t key;
  uint32_t start, count;
};

struct GridView {
  const double *pts;       // sorted by cell, 3 doubles each
  const uint32_t *idx;     // original index of each sorted point
  const GridSlot *table;   // mask + 1 slots
  uint32_t mask;           // table capacity - 1
  int n;
  double origin[3];  // the low corner of the points' box
  double hi[3];      // its high corner
  double cell;
};

// Whether every point of the grid is farther than `reach` from (x, y, z): the distance to the
// points' box exceeds the reach by more than the rounding of any distance computed near it.
SH_HD bool grid_far(const GridView &g, double x, double y, double z, double reach) {
  double q[3] = {x, y, z}, s = 0.0, m = 1.0;
  for (int k = 0; k < 3; ++k) {
    double gap = fmax(fmax(g.origin[k] - q[k], q[k] - g.hi[k]), 0.0);
    s += gap * gap;
    m = fmax(m, fmax(fabs(q[k]), fmax(fabs(g.origin[k]), fabs(g.hi[k]))));
  }
  double r = reach * (1.0 + 1e-9) + 1e-9 * m;
  return s > r * r;
}

SH_HD uint64_t grid_pack(int64_t x, int64_t y, int64_t z) {
  return (uint64_t)x | ((uint64_t)y << GRID_BITS) | ((uint64_t)z << (2 * GRID_BITS));
}

SH_HD uint32_t grid_hash(uint64_t k, uint32_t mask) {
  k ^= k >> 31;
  k *= 0x9E3779B97F4A7C15ull;
  k ^= k >> 29;
  return (uint32_t)k & mask;
}

// A table slot in one load: (key, start, count).
__device__ __forceinline__ GridSlot load_slot(const GridSlot *t, uint32_t i) {
  typedef uint32_t u32x4_t __attribute__((ext_vector_type(4)));
  u32x4_t v = *(__attribute__((address_space(1))) const u32x4_t *)(t + i);
  GridSlot s;
  s.key = (uint64_t)v.x | ((uint64_t)v.y << 32);
  s.start = v.z;
  s.count = v.w;
  return s;
}

// The points of the 27 cells around cell (cx, cy, cz), a slab of nine cells at a time: the nine
// table probes are issued together, each bringing its key and its run of points at once, so a
// slab costs one memory round trip rather than eighteen (a probe that lands on another key walks
// on alone, exactly as the table was filled). visit(j) is called for every sorted point j of every occupied cell; the
// callers keep an order-independent minimum, so the order of the calls does not matter.
template <class Visit>
__device__ __forceinline__ void grid_scan27(const GridView &g, int64_t cx, int64_t cy, int64_t cz, Visit &&visit) {
  for (int dx = -1; dx <= 1; ++dx) {
    const int64_t x = cx + dx;
    if (x < 0 || x >= GRID_MAXC) continue;
    uint64_t want[9];
    GridSlot got[9];
#pragma unroll
    for (int t = 0; t < 9; ++t) {
      const int64_t y = cy + (t / 3 - 1), z = cz + (t % 3 - 1);
      const bool ok = y >= 0 && y < GRID_MAXC && z >= 0 && z < GRID_MAXC;
      want[t] = ok ? grid_pack(x, y, z) : GRID_EMPTY;  // a packed cell is never GRID_EMPTY
      got[t] = GridSlot{GRID_EMPTY, 0u, 0u};
      if (ok) got[t] = load_slot(g.table, grid_hash(want[t], g.mask));
    }
#pragma unroll
    for (int t = 0; t < 9; ++t) {
      uint32_t slot = grid_hash(want[t], g.mask);
      for (uint32_t probe = 0;
           want[t] != GRID_EMPTY && got[t].key != want[t] && got[t].key != GRID_EMPTY && probe < g.mask; ++probe) {
        slot = (slot + 1) & g.mask;
        got[t] = load_slot(g.table, slot);
      }
      if (got[t].key != want[t]) got[t].count = 0;
    }
#pragma unroll
    for (int t = 0; t < 9; ++t)
      for (uint32_t k = got[t].start; k < got[t].start + got[t].count; ++k) visit(k);
  }
}

// The global nearest point to q among the 27 cells around it, ties to the lowest original index.
// Returns false when those cells hold no point.
__device__ __forceinline__ bool grid_nearest(const GridView &g, double qx, double qy, double qz, uint32_t *out_idx,
                                             double *out_d2) {
  if (g.n <= 0) return false;
  double fx = floor((qx - g.origin[0]) / g.cell);
  double fy = floor((qy - g.origin[1]) / g.cell);
  double fz = floor((qz - g.origin[2]) / g.cell);
  // Far outside the occupied cells (or NaN): nothing within one cell.
  if (!(fx >= -1.0 && fx <= (double)GRID_MAXC && fy >= -1.0 && fy <= (double)GRID_MAXC && fz >= -1.0 &&
        fz <= (double)GRID_MAXC))
    return false;
  double best = __builtin_inf();
  uint32_t bi = 0xFFFFFFFFu;
  grid_scan27(g, (int64_t)fx, (int64_t)fy, (int64_t)fz, [&](uint32_t j) {
    const double *p = g.pts + 3 * (size_t)j;
    double d2 = d2_f64(qx, qy, qz, gldd(p), gldd(p + 1), gldd(p + 2));
    uint32_t oi = gldu(g.idx + j);
    if (d2 < best || (d2 == best && oi < bi)) {
      best = d2;
      bi = oi;
    }
  });
  if (bi == 0xFFFFFFFFu) return false;
  *out_idx = bi;
  *out_d2 = best;
  return true;
}

// PointTree::nearest_below (kdtree.rs:182-192): the nearest point when sqrt(d^2) < bound. The grid's
// cell must be at least `bound`.
__device__ __forceinline__ bool grid_nearest_below(const GridView &g, double qx, double qy, double qz, double bound,
                                                   uint32_t *idx, double *dist) {
  if (!(bound > 0.0)) return false;
  double d2;
  if (!grid_nearest(g, qx, qy, qz, idx, &d2)) return false;
  double d = sqrt(d2);
  if (!(d < bound)) return false;
  *dist = d;
  return true;
}

// cpu.rs:136-169 `agreeing` for one pose (R row-major then tau, 12 doubles) by one wave, lanes
// striding the probe: move the point and its normal by the pose in f64, the nearest target strictly
// inside the radius, the normal test. Every lane of the wave calls it and gets the count.
__device__ __forceinline__ uint32_t coarse_agree_wave(const GridView &g, const double *tn, const double *pp,
                                                      const double *pn, int np, const double *r, double radius,
                                                      double agree) {
  const int lane = threadIdx.x & 63;
  const double *tau = r + 9;
  uint32_t cnt = 0;
  for (int base = 0; base < np; base += 64) {
    int i = base + lane;
    bool ok = false;
    if (i < np) {
      double px = pp[3 * i], py = pp[3 * i + 1], pz = pp[3 * i + 2];
      double mx = r[0] * px + r[1] * py + r[2] * pz + tau[0];
      double my = r[3] * px + r[4] * py + r[5] * pz + tau[1];
      double mz = r[6] * px + r[7] * py + r[8] * pz + tau[2];
      uint32_t near;
      double dist;
      if (grid_nearest_below(g, mx, my, mz, radius, &near, &dist)) {
        double nx = pn[3 * i], ny = pn[3 * i + 1], nz = pn[3 * i + 2];
        double ux = r[0] * nx + r[1] * ny + r[2] * nz;
        double uy = r[3] * nx + r[4] * ny + r[5] * nz;
        double uz = r[6] * nx + r[7] * ny + r[8] * nz;
        const double *th = tn + 3 * (size_t)near;
        double dd = th[0] * ux + th[1] * uy + th[2] * uz;
        ok = dd > agree;
      }
    }
    cnt += (uint32_t)__popcll(__ballot(ok));
  }
  return cnt;
}

}  // namespace sherd
