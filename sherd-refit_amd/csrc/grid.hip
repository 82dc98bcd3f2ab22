// Device build of the f64 neighbour grid (grid.hpp), the bounded nearest query, and the coarse
// breakline score (CpuExecutor::coarse_scores, crates/sherd-core/src/executor/cpu.rs:46-68, 136-169).
#include <hip/hip_runtime.h>

#include <cmath>

#include "grid.hpp"
#include "kernels.hpp"
#include "sort.hpp"

namespace sherd {
namespace {

constexpr int BT = 256;
inline int blocks_for(long n, int t) { return (int)((n + t - 1) / t); }

__global__ void k_bbox_partial(const double *p, int n, double *part) {
  __shared__ double sh[6][BT];
  int t = threadIdx.x;
  double v[6] = {__builtin_inf(), __builtin_inf(), __builtin_inf(), -__builtin_inf(), -__builtin_inf(),
                 -__builtin_inf()};
  for (int i = blockIdx.x * BT + t; i < n; i += gridDim.x * BT)
    for (int k = 0; k < 3; ++k) {
      double c = p[3 * i + k];
      v[k] = fmin(v[k], c);
      v[3 + k] = fmax(v[3 + k], c);
    }
  for (int k = 0; k < 6; ++k) sh[k][t] = v[k];
  __syncthreads();
  for (int s = BT / 2; s > 0; s >>= 1) {
    if (t < s)
      for (int k = 0; k < 6; ++k) sh[k][t] = k < 3 ? fmin(sh[k][t], sh[k][t + s]) : fmax(sh[k][t], sh[k][t + s]);
    __syncthreads();
  }
  if (t < 6) part[blockIdx.x * 6 + t] = sh[t][0];
}

__global__ void k_bbox_final(const double *part, int blocks, double *out) {
  int k = threadIdx.x;
  if (k >= 6) return;
  double v = k < 3 ? __builtin_inf() : -__builtin_inf();
  for (int b = 0; b < blocks; ++b) v = k < 3 ? fmin(v, part[b * 6 + k]) : fmax(v, part[b * 6 + k]);
  out[k] = v;
}

__global__ void k_keys(const double *p, int n, double ox, double oy, double oz, double cell, uint64_t *key,
                       uint32_t *val) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int64_t x = (int64_t)floor((p[3 * i] - ox) / cell);
  int64_t y = (int64_t)floor((p[3 * i + 1] - oy) / cell);
  int64_t z = (int64_t)floor((p[3 * i + 2] - oz) / cell);
  x = x < 0 ? 0 : (x >= GRID_MAXC ? GRID_MAXC - 1 : x);
  y = y < 0 ? 0 : (y >= GRID_MAXC ? GRID_MAXC - 1 : y);
  z = z < 0 ? 0 : (z >= GRID_MAXC ? GRID_MAXC - 1 : z);
  key[i] = grid_pack(x, y, z);
  val[i] = (uint32_t)i;
}

__global__ void k_gather(const double *p, const uint32_t *order, int n, double *sp, uint32_t *sidx) {
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  uint32_t i = order[j];
  sp[3 * j] = p[3 * i];
  sp[3 * j + 1] = p[3 * i + 1];
  sp[3 * j + 2] = p[3 * i + 2];
  sidx[j] = i;
}

__global__ void k_insert(const uint64_t *key, int n, GridSlot *table, uint32_t mask) {
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  uint64_t k = key[j];
  if (j > 0 && key[j - 1] == k) return;
  uint32_t slot = grid_hash(k, mask);
  for (uint32_t probe = 0; probe <= mask; ++probe) {
    unsigned long long prev = atomicCAS((unsigned long long *)&table[slot].key, (unsigned long long)GRID_EMPTY,
                                        (unsigned long long)k);
    if (prev == GRID_EMPTY) {
      table[slot].start = (uint32_t)j;
      return;
    }
    slot = (slot + 1) & mask;
  }
}

__global__ void k_counts(const uint64_t *key, int n, GridSlot *table, uint32_t mask) {
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  uint64_t k = key[j];
  if (j + 1 < n && key[j + 1] == k) return;
  uint32_t slot = grid_hash(k, mask);
  for (uint32_t probe = 0; probe <= mask; ++probe) {
    if (table[slot].key == k) {
      table[slot].count = (uint32_t)(j + 1) - table[slot].start;
      return;
    }
    slot = (slot + 1) & mask;
  }
}

__global__ void __launch_bounds__(256) k_nearest(GridView g, const double *q, int m, double bound, int64_t *idx,
                                                 double *dist) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  uint32_t j;
  double d;
  if (grid_nearest_below(g, q[3 * i], q[3 * i + 1], q[3 * i + 2], bound, &j, &d)) {
    idx[i] = (int64_t)j;
    dist[i] = d;
  } else {
    idx[i] = -1;
    dist[i] = __builtin_inf();
  }
}

// One wave per pose (coarse_agree_wave).
__global__ void __launch_bounds__(256) k_coarse(GridView g, const double *tn, const double *pp, const double *pn,
                                                int np, const double *poses, int k, double radius, double agree,
                                                uint32_t *counts, double *scores) {
  int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  if (wave >= k) return;
  uint32_t cnt = coarse_agree_wave(g, tn, pp, pn, np, poses + 12 * (size_t)wave, radius, agree);
  if ((threadIdx.x & 63) == 0) {
    counts[wave] = cnt;
    scores[wave] = (double)cnt / (double)np;
  }
}

}  // namespace

hipError_t grid_build(const double *dp, int n, double reach, DevGrid *g, hipStream_t s) {
  g->n = n;
  if (n <= 0) return hipSuccess;
  int pb = blocks_for(n, BT);
  if (pb > 512) pb = 512;
  double *part = nullptr, *bb = nullptr;
  uint64_t *key = nullptr, *tkey = nullptr;
  uint32_t *val = nullptr, *tval = nullptr, *hist = nullptr;
  hipError_t e = hipSuccess;
  double hb[6];
  do {
    if ((e = hipMalloc(&part, sizeof(double) * 6 * pb)) != hipSuccess) break;
    if ((e = hipMalloc(&bb, sizeof(double) * 6)) != hipSuccess) break;
    hipLaunchKernelGGL(k_bbox_partial, dim3(pb), dim3(BT), 0, s, dp, n, part);
    hipLaunchKernelGGL(k_bbox_final, dim3(1), dim3(64), 0, s, part, pb, bb);
    if ((e = hipMemcpyAsync(hb, bb, sizeof hb, hipMemcpyDeviceToHost, s)) != hipSuccess) break;
    if ((e = hipStreamSynchronize(s)) != hipSuccess) break;
    double ext = 0.0;
    for (int k = 0; k < 3; ++k) {
      g->origin[k] = hb[k];
      g->hi[k] = hb[3 + k];
      ext = std::fmax(ext, hb[3 + k] - hb[k]);
    }
    // At most 2^19 cells per axis, and a little slack so that a point and a query one reach apart
    // are never two cells apart after rounding.
    double cell = std::fmax(reach, ext / (double)(1 << 19)) * (1.0 + 1e-9);
    if (!(cell > 0.0) || !std::isfinite(cell)) cell = 1.0;
    g->cell = cell;
    uint32_t cap = 1;
    while (cap < 2u * (uint32_t)n) cap <<= 1;
    g->mask = cap - 1;
    if ((e = hipMalloc(&key, sizeof(uint64_t) * n)) != hipSuccess) break;
    if ((e = hipMalloc(&tkey, sizeof(uint64_t) * n)) != hipSuccess) break;
    if ((e = hipMalloc(&val, sizeof(uint32_t) * n)) != hipSuccess) break;
    if ((e = hipMalloc(&tval, sizeof(uint32_t) * n)) != hipSuccess) break;
    if ((e = hipMalloc(&hist, sizeof(uint32_t) * rsort::hist_words(n))) != hipSuccess) break;
    if ((e = hipMalloc(&g->pts, sizeof(double) * 3 * n)) != hipSuccess) break;
    if ((e = hipMalloc(&g->idx, sizeof(uint32_t) * n)) != hipSuccess) break;
    if ((e = hipMalloc(&g->table, sizeof(GridSlot) * cap)) != hipSuccess) break;
    hipLaunchKernelGGL(k_keys, dim3(blocks_for(n, BT)), dim3(BT), 0, s, dp, n, g->origin[0], g->origin[1],
                       g->origin[2], cell, key, val);
    if ((e = hipGetLastError()) != hipSuccess) break;
    if ((e = rsort::sort_pairs<uint64_t>(key, val, tkey, tval, hist, n, 0, 3 * GRID_BITS, s)) != hipSuccess) break;
    hipLaunchKernelGGL(k_gather, dim3(blocks_for(n, BT)), dim3(BT), 0, s, dp, val, n, g->pts, g->idx);
    // Every byte 0xFF: free keys (and runs that no probe reads).
    if ((e = hipMemsetAsync(g->table, 0xFF, sizeof(GridSlot) * cap, s)) != hipSuccess) break;
    hipLaunchKernelGGL(k_insert, dim3(blocks_for(n, BT)), dim3(BT), 0, s, key, n, g->table, g->mask);
    hipLaunchKernelGGL(k_counts, dim3(blocks_for(n, BT)), dim3(BT), 0, s, key, n, g->table, g->mask);
    e = hipGetLastError();
  } while (0);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  (void)hipFree(part);
  (void)hipFree(bb);
  (void)hipFree(key);
  (void)hipFree(tkey);
  (void)hipFree(val);
  (void)hipFree(tval);
  (void)hipFree(hist);
  if (e != hipSuccess) grid_free(g);
  return e;
}

void grid_free(DevGrid *g) {
  if (g->pts) (void)hipFree(g->pts);
  if (g->idx) (void)hipFree(g->idx);
  if (g->table) (void)hipFree(g->table);
  g->pts = nullptr;
  g->idx = nullptr;
  g->table = nullptr;
}

GridView view_of(const DevGrid &d) {
  GridView v;
  v.pts = d.pts;
  v.idx = d.idx;
  v.table = d.table;
  v.mask = d.mask;
  v.n = d.n;
  for (int k = 0; k < 3; ++k) {
    v.origin[k] = d.origin[k];
    v.hi[k] = d.hi[k];
  }
  v.cell = d.cell;
  return v;
}

hipError_t launch_nearest(const GridView &g, const double *q, int m, double bound, int64_t *idx, double *dist,
                          hipStream_t s) {
  if (m <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_nearest, dim3(blocks_for(m, 256)), dim3(256), 0, s, g, q, m, bound, idx, dist);
  return hipGetLastError();
}

hipError_t launch_coarse(const GridView &g, const double *tn, const double *pp, const double *pn, int np,
                         const double *poses, int k, double radius, double agree, uint32_t *counts, double *scores,
                         hipStream_t s) {
  if (k <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_coarse, dim3(blocks_for((long)k * 64, 256)), dim3(256), 0, s, g, tn, pp, pn, np, poses, k,
                     radius, agree, counts, scores);
  return hipGetLastError();
}

}  // namespace sherd
