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
//   k_hist_mode      per list, one workgroup: the two order statistics numpy's percentile
//                    interpolates between, by an 8-bit MSD radix select over the values'
//                    order-preserving keys; then the 60 bin edges in f32, the histogram with
//                    numpy's two edge corrections, and the centre of the first fullest bin
//
// Every float step is the reference's f32 (or, for the normals and the look-back test, f64)
// operation in its order, so the two modes are its bits.
#include <hip/hip_runtime.h>

#include "bvh_query.hpp"
#include "kernels.hpp"

namespace sherd {
namespace {

constexpr int HIST_BINS = 60;
constexpr int HM_THREADS = 1024;

__global__ void k_thick_normals(const double *V, const uint32_t *F, int nf, double *N) {
  int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  face_normal(V, F, (uint32_t)f, N + 3 * (size_t)f);
}

__global__ void __launch_bounds__(256) k_thick_rays(SceneView sv, const double *V, const uint32_t *F, const double *N,
                                                    int stride, int m, float *t_hit, uint32_t *prim) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= m) return;
  uint32_t i = (uint32_t)r * (uint32_t)stride;
  double c[3];
  face_centroid(V, F, i, c);
  const double *n = N + 3 * (size_t)i;
  double d0 = -n[0], d1 = -n[1], d2 = -n[2];
  V3 o = V3{(float)(c[0] + d0 * 1e-3), (float)(c[1] + d1 * 1e-3), (float)(c[2] + d2 * 1e-3)};
  V3 d = V3{(float)d0, (float)d1, (float)d2};
  float t;
  uint32_t face;
  first_hit(sv, o, d, 0, &t, &face);
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

// The k-th smallest (0-based) of x[0..n) by four 8-bit passes over the keys, most significant
// first; every thread of the workgroup calls it.
__device__ float kth_smallest(const float *x, int n, int k, uint32_t *hist, uint32_t *s_prefix, int *s_k) {
  uint32_t prefix = 0, pmask = 0;
  if (threadIdx.x == 0) *s_k = k;
  for (int pass = 0; pass < 4; ++pass) {
    const int shift = 24 - 8 * pass;
    for (int b = threadIdx.x; b < 256; b += blockDim.x) hist[b] = 0;
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
      uint32_t key = fkey(x[i]);
      if ((key & pmask) == prefix) atomicAdd(&hist[(key >> shift) & 255u], 1u);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      int kk = *s_k, d = 0;
      while (d < 255 && kk >= (int)hist[d]) kk -= (int)hist[d++];
      *s_k = kk;
      *s_prefix = prefix | ((uint32_t)d << shift);
    }
    __syncthreads();
    prefix = *s_prefix;
    pmask |= 255u << shift;
    __syncthreads();
  }
  return from_fkey(prefix);
}

// hist_mode over each of the lists (workgroup w takes list w): x[w], counts[w] values, into out[w].
struct HistLists {
  const float *x[2];
};

__global__ void __launch_bounds__(HM_THREADS) k_hist_mode(HistLists L, const uint32_t *counts, float *out) {
  __shared__ uint32_t hist[256];
  __shared__ uint32_t s_prefix;
  __shared__ int s_k;
  __shared__ float edges[HIST_BINS + 1];
  __shared__ uint32_t bins[HIST_BINS];
  __shared__ float s_lim[2];
  const int w = blockIdx.x;
  const float *x = L.x[w];
  const int n = (int)counts[w];
  if (n == 0) {
    if (threadIdx.x == 0) out[w] = __builtin_nanf("");
    return;
  }
  // percentile90: the virtual index 0.9 (n - 1) in f64, interpolated in f32.
  const double vi = (double)(n - 1) * (90.0 / 100.0);
  const double prev = floor(vi);
  const int i = (int)prev;
  const double gamma = vi - prev;
  const float a = kth_smallest(x, n, i, hist, &s_prefix, &s_k);
  const float b = i + 1 < n ? kth_smallest(x, n, i + 1, hist, &s_prefix, &s_k) : a;
  if (threadIdx.x == 0) {
    const float diff = b - a;
    const float p90 = n == 1 ? a : (gamma >= 0.5 ? b - diff * (float)(1.0 - gamma) : a + diff * (float)gamma);
    // numpy's _get_outer_edges widens an empty range.
    const float first = p90 == 0.0f ? -0.5f : 0.0f, last = p90 == 0.0f ? 0.5f : p90;
    s_lim[0] = first;
    s_lim[1] = last;
    const float step = (last - first) / (float)HIST_BINS;
    for (int e = 0; e < HIST_BINS; ++e) edges[e] = (float)e * step + first;
    edges[HIST_BINS] = last;
  }
  for (int e = threadIdx.x; e < HIST_BINS; e += blockDim.x) bins[e] = 0;
  __syncthreads();
  const float first = s_lim[0], last = s_lim[1], denom = last - first;
  for (int j = threadIdx.x; j < n; j += blockDim.x) {
    const float v = x[j];
    if (!(v >= first && v <= last)) continue;
    int k = (int)(((v - first) / denom) * (float)HIST_BINS);
    if (k == HIST_BINS) k -= 1;
    // numpy's two corrections for values within a ULP of a bin edge.
    if (v < edges[k]) k -= 1;
    if (k != HIST_BINS - 1 && v >= edges[k + 1]) k += 1;
    atomicAdd(&bins[k], 1u);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int best = 0;
    for (int e = 1; e < HIST_BINS; ++e)
      if (bins[e] > bins[best]) best = e;  // the first of the fullest
    out[w] = 0.5f * (edges[best] + edges[best + 1]);
  }
}

}  // namespace

hipError_t launch_thick_normals(const double *V, const uint32_t *F, int nf, double *N, hipStream_t s) {
  if (nf <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_thick_normals, dim3((nf + 255) / 256), dim3(256), 0, s, V, F, nf, N);
  return hipGetLastError();
}

hipError_t launch_thick_rays(const SceneView &sv, const double *V, const uint32_t *F, const double *N, int stride, int m,
                             float *t_hit, uint32_t *prim, hipStream_t s) {
  if (m <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_thick_rays, dim3((m + 255) / 256), dim3(256), 0, s, sv, V, F, N, stride, m, t_hit, prim);
  return hipGetLastError();
}

hipError_t launch_thick_modes(const double *N, int nf, const uint32_t *idx, int stride, const float *t_hit,
                              const uint32_t *prim, int m, float *ok_d, float *far_d, uint32_t *counts,
                              float *modes, hipStream_t s) {
  hipLaunchKernelGGL(k_thick_filter, dim3(m > 0 ? (m + 255) / 256 : 1), dim3(256), 0, s, N, nf, idx, stride, t_hit,
                     prim, m, ok_d, far_d, counts);
  HistLists L;
  L.x[0] = ok_d;
  L.x[1] = far_d;
  hipLaunchKernelGGL(k_hist_mode, dim3(2), dim3(HM_THREADS), 0, s, L, counts, modes);
  return hipGetLastError();
}

}  // namespace sherd
