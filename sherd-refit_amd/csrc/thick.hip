// R §3.2's wall thickness on the device: thickness::estimate_thickness and thickness_from_hits
// (crates/sherd-core/src/fragment/thickness.rs:100-161, looking_back :184-203, hist_mode
// :207-246, bin_edges :249-262, percentile90 :267-290).
//
//   k_thick_normals  the f64 face normals (mesh/geometry.rs:70-96), one thread per face
//   k_thick_rays     one inward ray per face of R §3.2's stride: origin c + d 1e-3 and direction
//                    d = -n in f64, narrowed to f32 once; the first hit through the LBVH
//   k_thick_filter   the valid hits (finite t, a face of the mesh) and, of those, the ones whose
//                    face looks back along the ray (f64 normals, > 0.7), each gathered into its own
//                    list (their order does not matter to a percentile or a histogram)
//   k_sel_digits,    per list, numpy's percentile: the two order statistics it interpolates
//   k_sel_next       between, by a device-wide 8-bit MSD radix select over the values'
//                    order-preserving keys (four digit passes), then the smallest key above the
//                    lower one when the upper is not a copy of it
//   k_sel_hist       the 60 bin edges in f32 and the histogram with numpy's two edge corrections
//   k_sel_mode       the centre of the first fullest bin
//
// Every float step is the reference's f32 (or, for the normals and the look-back test, f64)
// operation in its order, so the two modes are its bits.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "bvh_query.hpp"
#include "kernels.hpp"

namespace sherd {
namespace {

constexpr int HIST_BINS = 60;
static_assert(HIST_BINS == THICK_BINS, "the state's bin count");

struct HistLists {
  const float *x[2];
};

// A face with a vertex index outside V raises *bad and gets a zero normal (its ray is a miss).
__device__ __forceinline__ bool face_ok(const uint32_t *F, uint32_t f, int nv) {
  return F[3 * (size_t)f] < (uint32_t)nv && F[3 * (size_t)f + 1] < (uint32_t)nv && F[3 * (size_t)f + 2] < (uint32_t)nv;
}

__global__ void k_thick_normals(const double *V, int nv, const uint32_t *F, int nf, double *N, uint32_t *bad) {
  int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  if (!face_ok(F, (uint32_t)f, nv)) {
    double *n = N + 3 * (size_t)f;
    n[0] = n[1] = n[2] = 0.0;
    atomicOr(bad, 1u);
    return;
  }
  face_normal(V, F, (uint32_t)f, N + 3 * (size_t)f);
}

__global__ void __launch_bounds__(256) k_thick_rays(SceneView sv, const double *V, int nv, const uint32_t *F,
                                                    const double *N, int stride, int m, float *t_hit, uint32_t *prim) {
  __shared__ uint2 s_stk[QSTACK][256];
  LdsStack<256, QSTACK> st{&s_stk[0][threadIdx.x]};
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= m) return;
  uint32_t i = (uint32_t)r * (uint32_t)stride;
  if (!face_ok(F, i, nv)) {
    t_hit[r] = __builtin_inff();
    prim[r] = 0xFFFFFFFFu;
    return;
  }
  double c[3];
  face_centroid(V, F, i, c);
  const double *n = N + 3 * (size_t)i;
  double d0 = -n[0], d1 = -n[1], d2 = -n[2];
  V3 o = V3{(float)(c[0] + d0 * 1e-3), (float)(c[1] + d1 * 1e-3), (float)(c[2] + d2 * 1e-3)};
  V3 d = V3{(float)d0, (float)d1, (float)d2};
  float t;
  uint32_t face;
  first_hit_with(st, sv, o, d, 0, &t, &face);
  t_hit[r] = t;
  prim[r] = face;
}

// idx: each ray's face, or null for the stride rule (ray r from face r * stride).
__global__ void __launch_bounds__(256) k_thick_filter(const double *N, int nf, const uint32_t *idx, int stride,
                                                      const float *t_hit, const uint32_t *prim, int m, float *ok_d,
                                                      float *far_d, uint32_t *counts) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  bool ok = false, far = false;
  float d = 0.0f;
  if (r < m) {
    d = t_hit[r];
    uint32_t p = prim[r];
    ok = (d - d == 0.0f) && p < (uint32_t)nf;
    if (ok) {
      uint32_t from = idx ? idx[r] : (uint32_t)r * (uint32_t)stride;
      const double *h = N + 3 * (size_t)p, *f = N + 3 * (size_t)from;
      double dot = -((h[0] * f[0] + h[1] * f[1]) + h[2] * f[2]);
      far = dot > 0.7;
    }
  }
  uint32_t so = wave_slot(&counts[0], ok);
  if (ok) ok_d[so] = d;
  uint32_t sf = wave_slot(&counts[1], far);
  if (far) far_d[sf] = d;
}

// An order-preserving key of an f32: unsigned comparison of keys is the float order (total order,
// -0 below +0).
__device__ __forceinline__ uint32_t fkey(float v) {
  uint32_t b = __float_as_uint(v);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float from_fkey(uint32_t k) {
  return __uint_as_float((k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k);
}

// Each list's selection and histogram are device-wide passes over SEL_BLOCKS workgroups per list
// (blockIdx.y is the list). Counts go to per-wave LDS copies, then one global atomic per non-empty
// bin per workgroup; a wave whose counted lanes share one bin (the common case: the values crowd
// round the mode) adds them with one atomic.
constexpr int SEL_THREADS = 256;
constexpr int SEL_WAVES = SEL_THREADS / 64;

__device__ __forceinline__ void wave_count(uint32_t *h, uint32_t bin, bool valid) {
  const uint64_t act = __ballot(valid);
  if (act == 0) return;
  const int lead = __ffsll((unsigned long long)act) - 1;
  const uint32_t b0 = (uint32_t)__shfl((int)bin, lead);
  const uint64_t same = __ballot(valid && bin == b0);
  if (same == act) {
    if ((int)(threadIdx.x & 63) == lead) atomicAdd(&h[b0], (uint32_t)__popcll(act));
  } else if (valid) {
    atomicAdd(&h[bin], 1u);
  }
}

// numpy's percentile(x, 90): the virtual index 0.9 (n - 1) in f64, split into the lower order
// statistic's rank and the interpolation weight.
struct P90Index {
  uint32_t i;
  double gamma;
};
__device__ __forceinline__ P90Index p90_index(uint32_t n) {
  const double vi = (double)(n - 1) * (90.0 / 100.0);
  const double prev = floor(vi);
  return P90Index{(uint32_t)prev, vi - prev};
}

// The selection after `passes` digit passes: the key prefix of the rank-k value, its rank among the
// keys with that prefix, and how many keys share the prefix. Every thread calls it; lds holds 256
// words. Every workgroup computes the same answer from the same counts.
struct SelPos {
  uint32_t prefix, rank, eq;
};
__device__ SelPos sel_walk(const ThickSel *s, int passes, uint32_t k, uint32_t *lds, SelPos *s_pos) {
  uint32_t prefix = 0, kk = k, eq = 0;
  for (int q = 0; q < passes; ++q) {
    __syncthreads();
    for (int b = threadIdx.x; b < 256; b += blockDim.x) lds[b] = s->hist[q][b];
    __syncthreads();
    if (threadIdx.x == 0) {
      int d = 0;
      while (d < 255 && kk >= lds[d]) kk -= lds[d++];
      prefix |= (uint32_t)d << (24 - 8 * q);
      eq = lds[d];
    }
  }
  if (threadIdx.x == 0) *s_pos = SelPos{prefix, kk, eq};
  __syncthreads();
  return *s_pos;
}

// One 8-bit digit pass of the MSD radix select of the rank-i key.
__global__ void __launch_bounds__(SEL_THREADS) k_sel_digits(HistLists L, ThickState *st, int pass) {
  __shared__ uint32_t h[SEL_WAVES][256];
  __shared__ SelPos s_pos;
  const int w = blockIdx.y;
  const uint32_t n = st->counts[w];
  if (n == 0) return;
  ThickSel *s = &st->sel[w];
  const SelPos pos = sel_walk(s, pass, p90_index(n).i, &h[0][0], &s_pos);
  const uint32_t pmask = pass ? ~0u << (32 - 8 * pass) : 0u;
  const int shift = 24 - 8 * pass;
  for (int b = threadIdx.x; b < SEL_WAVES * 256; b += blockDim.x) (&h[0][0])[b] = 0;
  __syncthreads();
  const float *x = L.x[w];
  uint32_t *hw = h[threadIdx.x >> 6];
  for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += gridDim.x * blockDim.x) {
    const uint32_t j = base + threadIdx.x;
    uint32_t key = 0;
    if (j < n) key = fkey(x[j]);
    wave_count(hw, (key >> shift) & 255u, j < n && (key & pmask) == pos.prefix);
  }
  __syncthreads();
  for (int b = threadIdx.x; b < 256; b += blockDim.x) {
    uint32_t c = 0;
    for (int q = 0; q < SEL_WAVES; ++q) c += h[q][b];
    if (c) atomicAdd(&s->hist[pass][b], c);
  }
}

// The rank-(i+1) value when it is not a copy of the rank-i one: the smallest key above it, kept as
// the largest complemented key.
__global__ void __launch_bounds__(SEL_THREADS) k_sel_next(HistLists L, ThickState *st) {
  __shared__ uint32_t lds[256];
  __shared__ SelPos s_pos;
  __shared__ uint32_t s_max;
  const int w = blockIdx.y;
  const uint32_t n = st->counts[w];
  if (n == 0) return;
  ThickSel *s = &st->sel[w];
  const uint32_t i = p90_index(n).i;
  const SelPos pos = sel_walk(s, 4, i, lds, &s_pos);
  if (i + 1 >= n || pos.rank + 1 < pos.eq) return;  // the next one is a copy (or there is none)
  if (threadIdx.x == 0) s_max = 0;
  __syncthreads();
  const float *x = L.x[w];
  uint32_t best = 0;
  for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
    const uint32_t key = fkey(x[j]);
    if (key > pos.prefix) best = max(best, ~key);
  }
  best = wave_max(best);
  if ((threadIdx.x & 63) == 0 && best) atomicMax(&s_max, best);
  __syncthreads();
  if (threadIdx.x == 0 && s_max) atomicMax(&s->gt_inv, s_max);
}

// The histogram's range from the two order statistics: numpy's percentile interpolation in f32,
// and _get_outer_edges' widening of an empty range.
__device__ void sel_limits(const ThickSel *s, uint32_t n, const SelPos &pos, float *first, float *last) {
  const P90Index pi = p90_index(n);
  const float a = from_fkey(pos.prefix);
  const float b = (pi.i + 1 >= n || pos.rank + 1 < pos.eq) ? a : from_fkey(~s->gt_inv);
  const float diff = b - a;
  const float p90 = n == 1 ? a : (pi.gamma >= 0.5 ? b - diff * (float)(1.0 - pi.gamma) : a + diff * (float)pi.gamma);
  *first = p90 == 0.0f ? -0.5f : 0.0f;
  *last = p90 == 0.0f ? 0.5f : p90;
}

// np.linspace(first, last, 61) in f32.
__device__ __forceinline__ float bin_edge(int e, float first, float last) {
  const float step = (last - first) / (float)HIST_BINS;
  return e == HIST_BINS ? last : (float)e * step + first;
}

// np.histogram(x, 60, (first, last)): the bin from the scaled offset, then numpy's two corrections
// for values within a ULP of a bin edge.
__global__ void __launch_bounds__(SEL_THREADS) k_sel_hist(HistLists L, ThickState *st) {
  __shared__ uint32_t lds[256];
  __shared__ SelPos s_pos;
  __shared__ float edges[HIST_BINS + 1];
  __shared__ uint32_t bins[SEL_WAVES][HIST_BINS];
  const int w = blockIdx.y;
  const uint32_t n = st->counts[w];
  if (n == 0) return;
  ThickSel *s = &st->sel[w];
  const SelPos pos = sel_walk(s, 4, p90_index(n).i, lds, &s_pos);
  float first, last;
  sel_limits(s, n, pos, &first, &last);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    s->first = first;
    s->last = last;
  }
  for (int e = threadIdx.x; e <= HIST_BINS; e += blockDim.x) edges[e] = bin_edge(e, first, last);
  for (int e = threadIdx.x; e < SEL_WAVES * HIST_BINS; e += blockDim.x) (&bins[0][0])[e] = 0;
  __syncthreads();
  const float denom = last - first;
  const float *x = L.x[w];
  uint32_t *bw = bins[threadIdx.x >> 6];
  for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += gridDim.x * blockDim.x) {
    const uint32_t j = base + threadIdx.x;
    bool in = false;
    int k = 0;
    if (j < n) {
      const float v = x[j];
      in = v >= first && v <= last;
      if (in) {
        k = (int)(((v - first) / denom) * (float)HIST_BINS);
        if (k == HIST_BINS) k -= 1;
        if (v < edges[k]) k -= 1;
        if (k != HIST_BINS - 1 && v >= edges[k + 1]) k += 1;
      }
    }
    wave_count(bw, (uint32_t)k, in);
  }
  __syncthreads();
  for (int e = threadIdx.x; e < HIST_BINS; e += blockDim.x) {
    uint32_t c = 0;
    for (int q = 0; q < SEL_WAVES; ++q) c += bins[q][e];
    if (c) atomicAdd(&s->bins[e], c);
  }
}

// The centre of the first fullest bin, per list (NaN for an empty list).
__global__ void k_sel_mode(ThickState *st) {
  const int w = threadIdx.x;
  if (w >= 2) return;
  if (st->counts[w] == 0) {
    st->modes[w] = __builtin_nanf("");
    return;
  }
  const ThickSel *s = &st->sel[w];
  int best = 0;
  for (int e = 1; e < HIST_BINS; ++e)
    if (s->bins[e] > s->bins[best]) best = e;
  st->modes[w] = 0.5f * (bin_edge(best, s->first, s->last) + bin_edge(best + 1, s->first, s->last));
}

}  // namespace

hipError_t launch_thick_normals(const double *V, int nv, const uint32_t *F, int nf, double *N, uint32_t *bad,
                                hipStream_t s) {
  if (nf <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_thick_normals, dim3((nf + 255) / 256), dim3(256), 0, s, V, nv, F, nf, N, bad);
  return hipGetLastError();
}

hipError_t launch_thick_rays(const SceneView &sv, const double *V, int nv, const uint32_t *F, const double *N,
                             int stride, int m, float *t_hit, uint32_t *prim, hipStream_t s) {
  if (m <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_thick_rays, dim3((m + 255) / 256), dim3(256), 0, s, sv, V, nv, F, N, stride, m, t_hit, prim);
  return hipGetLastError();
}

hipError_t launch_thick_modes(const double *N, int nf, const uint32_t *idx, int stride, const float *t_hit,
                              const uint32_t *prim, int m, float *ok_d, float *far_d, ThickState *st, hipStream_t s) {
  hipLaunchKernelGGL(k_thick_filter, dim3(m > 0 ? (m + 255) / 256 : 1), dim3(256), 0, s, N, nf, idx, stride, t_hit,
                     prim, m, ok_d, far_d, st->counts);
  HistLists L;
  L.x[0] = ok_d;
  L.x[1] = far_d;
  // About eight values per thread per pass; a list's count (at most m) is read on the device.
  const int per = SEL_THREADS * 8;
  const dim3 grid((unsigned)std::min(std::max((m + per - 1) / per, 1), 512), 2);
  for (int pass = 0; pass < 4; ++pass) hipLaunchKernelGGL(k_sel_digits, grid, dim3(SEL_THREADS), 0, s, L, st, pass);
  hipLaunchKernelGGL(k_sel_next, grid, dim3(SEL_THREADS), 0, s, L, st);
  hipLaunchKernelGGL(k_sel_hist, grid, dim3(SEL_THREADS), 0, s, L, st);
  hipLaunchKernelGGL(k_sel_mode, dim3(1), dim3(64), 0, s, st);
  return hipGetLastError();
}

}  // namespace sherd
