// Stable LSD radix sort of (key, value) pairs on the device: 8-bit digits, 256-thread tiles of
// 4096 keys, ranks inside a tile from wave64 ballots. Used by the LBVH build (30-bit Morton codes)
// and the neighbour grid build (packed cell keys). Stability makes equal keys keep their input
// order, which is what gives both builds a deterministic layout.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sherd {
namespace rsort {

constexpr int THREADS = 256;
constexpr int ITEMS = 16;
constexpr int TILE = THREADS * ITEMS;
constexpr int BINS = 256;
constexpr int WAVES = THREADS / 64;

template <class K>
__global__ void __launch_bounds__(THREADS) k_hist(const K *keys, int n, int shift, uint32_t *hist, int tiles) {
  __shared__ uint32_t h[BINS];
  int t = threadIdx.x;
  h[t] = 0;
  __syncthreads();
  size_t base = (size_t)blockIdx.x * TILE;
  for (int r = 0; r < ITEMS; ++r) {
    size_t i = base + (size_t)r * THREADS + t;
    if (i < (size_t)n) atomicAdd(&h[(uint32_t)(keys[i] >> shift) & 255u], 1u);
  }
  __syncthreads();
  hist[(size_t)t * tiles + blockIdx.x] = h[t];  // digit-major
}

// Exclusive scan of `len` counters in one workgroup of 1024 threads, each owning a contiguous run.
static __global__ void __launch_bounds__(1024) k_scan(uint32_t *data, int len) {
  __shared__ uint32_t part[1024];
  int t = threadIdx.x;
  int per = (len + 1023) / 1024;
  int beg = t * per, end = beg + per < len ? beg + per : len;
  uint32_t s = 0;
  for (int i = beg; i < end; ++i) s += data[i];
  part[t] = s;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    uint32_t v = t >= off ? part[t - off] : 0u;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  uint32_t run = part[t] - s;
  for (int i = beg; i < end; ++i) {
    uint32_t c = data[i];
    data[i] = run;
    run += c;
  }
}

// ---- device-wide exclusive scan of u32 counters -------------------------------------------------
// Chunks of SCAN_CHUNK counters: each chunk's total, an exclusive scan of the totals (one
// workgroup), then each chunk scanned from its offset.
constexpr int SCAN_PER = 16;
constexpr int SCAN_CHUNK = THREADS * SCAN_PER;

static __global__ void __launch_bounds__(THREADS) k_scan_reduce(const uint32_t *data, size_t n, uint32_t *part) {
  __shared__ uint32_t w[WAVES];
  size_t base = (size_t)blockIdx.x * SCAN_CHUNK;
  uint32_t s = 0;
  for (int r = 0; r < SCAN_PER; ++r) {
    size_t i = base + (size_t)r * THREADS + threadIdx.x;
    if (i < n) s += data[i];
  }
  for (int o = 32; o > 0; o >>= 1) s += (uint32_t)__shfl_xor((int)s, o);
  if ((threadIdx.x & 63) == 0) w[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
    for (int k = 0; k < WAVES; ++k) t += w[k];
    part[blockIdx.x] = t;
  }
}

static __global__ void __launch_bounds__(THREADS) k_scan_apply(uint32_t *data, size_t n, const uint32_t *part) {
  __shared__ uint32_t sums[THREADS];
  size_t base = (size_t)blockIdx.x * SCAN_CHUNK + (size_t)threadIdx.x * SCAN_PER;
  uint32_t v[SCAN_PER];
  uint32_t s = 0;
#pragma unroll
  for (int r = 0; r < SCAN_PER; ++r) {
    size_t i = base + r;
    v[r] = i < n ? data[i] : 0u;
    s += v[r];
  }
  sums[threadIdx.x] = s;
  __syncthreads();
  for (int off = 1; off < THREADS; off <<= 1) {
    uint32_t x = threadIdx.x >= (unsigned)off ? sums[threadIdx.x - off] : 0u;
    __syncthreads();
    sums[threadIdx.x] += x;
    __syncthreads();
  }
  uint32_t run = part[blockIdx.x] + sums[threadIdx.x] - s;
#pragma unroll
  for (int r = 0; r < SCAN_PER; ++r) {
    size_t i = base + r;
    if (i < n) data[i] = run;
    run += v[r];
  }
}

inline size_t scan_parts(size_t n) { return (n + SCAN_CHUNK - 1) / SCAN_CHUNK; }

// Exclusive scan of n counters in place; `part` holds scan_parts(n) counters of scratch.
inline hipError_t scan_exclusive(uint32_t *data, size_t n, uint32_t *part, hipStream_t s) {
  if (n == 0) return hipSuccess;
  size_t nb = scan_parts(n);
  hipLaunchKernelGGL(k_scan_reduce, dim3((unsigned)nb), dim3(THREADS), 0, s, data, n, part);
  hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, s, part, (int)nb);
  hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)nb), dim3(THREADS), 0, s, data, n, part);
  return hipGetLastError();
}

template <class K>
__global__ void __launch_bounds__(THREADS) k_scatter(const K *keys, const uint32_t *vals, K *okeys, uint32_t *ovals,
                                                     int n, int shift, const uint32_t *offs, int tiles) {
  __shared__ uint32_t running[BINS];
  __shared__ uint32_t wcount[WAVES][BINS];
  __shared__ uint32_t goff[BINS];
  int t = threadIdx.x;
  int lane = t & 63, wave = t >> 6;
  running[t] = 0;
  goff[t] = offs[(size_t)t * tiles + blockIdx.x];
  size_t base = (size_t)blockIdx.x * TILE;
  uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  for (int r = 0; r < ITEMS; ++r) {
    for (int w = 0; w < WAVES; ++w) wcount[w][t] = 0;
    __syncthreads();
    size_t i = base + (size_t)r * THREADS + t;
    bool valid = i < (size_t)n;
    K key = valid ? keys[i] : (K)0;
    uint32_t d = (uint32_t)(key >> shift) & 255u;
    // Lanes of this wave holding the same digit.
    uint64_t peers = __ballot(valid);
    for (int b = 0; b < 8; ++b) {
      uint64_t bits = __ballot(valid && ((d >> b) & 1u));
      peers &= ((d >> b) & 1u) ? bits : ~bits;
    }
    uint32_t before = (uint32_t)__popcll(peers & lt);
    if (valid && before == 0) wcount[wave][d] = (uint32_t)__popcll(peers);
    __syncthreads();
    if (valid) {
      uint32_t rank = running[d] + before;
      for (int w = 0; w < wave; ++w) rank += wcount[w][d];
      okeys[goff[d] + rank] = key;
      ovals[goff[d] + rank] = vals[i];
    }
    __syncthreads();
    uint32_t add = 0;
    for (int w = 0; w < WAVES; ++w) add += wcount[w][t];
    running[t] += add;
  }
}

// Sorts n pairs by bits [begin_bit, end_bit) of the key. `tk`/`tv` are scratch of n entries,
// `hist` holds hist_words(n) counters. The sorted pairs end in `keys`/`vals`.
template <class K>
hipError_t sort_pairs(K *keys, uint32_t *vals, K *tk, uint32_t *tv, uint32_t *hist, int n, int begin_bit,
                      int end_bit, hipStream_t s) {
  if (n <= 1) return hipSuccess;
  int tiles = (n + TILE - 1) / TILE;
  K *ik = keys, *ok = tk;
  uint32_t *iv = vals, *ov = tv;
  int passes = 0;
  for (int shift = begin_bit; shift < end_bit; shift += 8, ++passes) {
    hipLaunchKernelGGL(k_hist<K>, dim3(tiles), dim3(THREADS), 0, s, ik, n, shift, hist, tiles);
    hipError_t e = scan_exclusive(hist, (size_t)tiles * BINS, hist + (size_t)tiles * BINS, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_scatter<K>, dim3(tiles), dim3(THREADS), 0, s, ik, iv, ok, ov, n, shift, hist, tiles);
    K *xk = ik; ik = ok; ok = xk;
    uint32_t *xv = iv; iv = ov; ov = xv;
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  if (passes & 1) {
    e = hipMemcpyAsync(keys, ik, sizeof(K) * n, hipMemcpyDeviceToDevice, s);
    if (e != hipSuccess) return e;
    e = hipMemcpyAsync(vals, iv, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, s);
  }
  return e;
}

// Counters for sort_pairs: the per-tile digit histograms and the scan's chunk totals.
inline size_t hist_words(int n) {
  size_t h = (size_t)BINS * (size_t)((n + TILE - 1) / TILE);
  return h + scan_parts(h);
}

}  // namespace rsort
}  // namespace sherd
