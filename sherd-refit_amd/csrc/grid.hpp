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

struct GridView {
  const double *pts;    // sorted by cell, 3 doubles each
  const uint32_t *idx;  // original index of each sorted point
  const uint64_t *keys; // hash table keys, GRID_EMPTY when free
  const uint2 *cells;   // (start, count) per table slot
  uint32_t mask;        // table capacity - 1
  int n;
  double origin[3];
  double cell;
};

SH_HD uint64_t grid_pack(int64_t x, int64_t y, int64_t z) {
  return (uint64_t)x | ((uint64_t)y << GRID_BITS) | ((uint64_t)z << (2 * GRID_BITS));
}

SH_HD uint32_t grid_hash(uint64_t k, uint32_t mask) {
  k ^= k >> 31;
  k *= 0x9E3779B97F4A7C15ull;
  k ^= k >> 29;
  return (uint32_t)k & mask;
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
  int64_t cx = (int64_t)fx, cy = (int64_t)fy, cz = (int64_t)fz;
  double best = __builtin_inf();
  uint32_t bi = 0xFFFFFFFFu;
  for (int dx = -1; dx <= 1; ++dx) {
    int64_t x = cx + dx;
    if (x < 0 || x >= GRID_MAXC) continue;
    for (int dy = -1; dy <= 1; ++dy) {
      int64_t y = cy + dy;
      if (y < 0 || y >= GRID_MAXC) continue;
      for (int dz = -1; dz <= 1; ++dz) {
        int64_t z = cz + dz;
        if (z < 0 || z >= GRID_MAXC) continue;
        uint64_t key = grid_pack(x, y, z);
        uint32_t slot = grid_hash(key, g.mask);
        for (uint32_t probe = 0; probe <= g.mask; ++probe) {
          uint64_t k = g.keys[slot];
          if (k == GRID_EMPTY) break;
          if (k == key) {
            uint2 c = g.cells[slot];
            for (uint32_t j = c.x; j < c.x + c.y; ++j) {
              double d2 = d2_f64(qx, qy, qz, g.pts[3 * j], g.pts[3 * j + 1], g.pts[3 * j + 2]);
              uint32_t oi = g.idx[j];
              if (d2 < best || (d2 == best && oi < bi)) {
                best = d2;
                bi = oi;
              }
            }
            break;
          }
          slot = (slot + 1) & g.mask;
        }
      }
    }
  }
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
