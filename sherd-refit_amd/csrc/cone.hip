// R §3.4.3's cone vote on the device: segment::classify_faces
// (crates/sherd-core/src/fragment/segment.rs:318-381, cone_basis :383-396, constants :53-62).
//
// One lane per ray, seven per face: the face normal and centroid in f64 (mesh/geometry.rs:70-96),
// the origin pushed 1e-3 along the raw normal and narrowed to f32, seven rays (straight down -NS,
// then six on a 15 degree cone), each a first hit through the LBVH; a hit votes when 0.1 t < dh,
// 0.5 < dh/t < 1.8 in f32, and the hit face looks back along the ray (h . d > 0.7 in f64). The
// cone's sines and cosines are computed once on the host and passed in, so every lane uses the
// same doubles.
//
// Beside it, the per-vertex form the benchmark's slab plate asks for: one ray per vertex along the
// inward vertex normal, with the cone's first ray's test, and a face label from its vertices.
#include <hip/hip_runtime.h>

#include "bvh_query.hpp"
#include "kernels.hpp"

namespace sherd {
namespace {

// Every face's f64 unit normal (face_normal, geometry.rs:70-96), once per call: the cone's own
// face normal and every hit's look-back test read 24 bytes instead of three vertex rows.
__global__ void __launch_bounds__(256) k_face_normals(const double *V, const uint32_t *F, int nf, double *fn) {
  int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  face_normal(V, F, (uint32_t)f, fn + 3 * (size_t)f);
}

constexpr int CONE_RAYS = 7;
constexpr int CONE_FACES_PER_WAVE = 64 / CONE_RAYS;  // 9 faces, 63 lanes; the last lane idles
constexpr int CONE_LSTACK = 11;                      // deferred nodes per ray, in LDS (22 KB: 7 workgroups per CU)

// One lane per ray: a wave takes nine faces, its lanes 7j..7j+6 cast face j's seven rays (straight
// down -NS, then the six of the 15 degree cone) from the same origin, so a face's traversals run
// side by side through the same nodes. Each lane builds the face's basis itself (the same doubles
// on every lane); a face's votes are a popcount of its seven bits of the wave's ballot.
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(7))) k_cone(SceneView sv, const double *V, const uint32_t *F, const double *fnv,
                                              int nf, const int64_t *faces, const double *ns, int n, ConeConst cc,
                                              int retest, uint8_t *good) {
  __shared__ uint2 s_stk[CONE_LSTACK][256];
  const int lane = threadIdx.x & 63;
  const int wave = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  const int j = lane / CONE_RAYS, k = lane - CONE_RAYS * j;
  const int q = wave * CONE_FACES_PER_WAVE + j;
  bool vote = false;
  if (j < CONE_FACES_PER_WAVE && q < n) {
    uint32_t i = faces ? (uint32_t)faces[q] : (uint32_t)q;
    const double *nv = ns + 3 * (size_t)q;
    double n0 = nv[0], n1 = nv[1], n2 = nv[2];
    if (n0 - n0 == 0.0 && n1 - n1 == 0.0 && n2 - n2 == 0.0) {
      double a0 = 0.0, a1 = 0.0, a2 = 0.0;
      if (fabs(n0) < 0.9) a0 = 1.0; else a1 = 1.0;
      double e0 = n1 * a2 - n2 * a1, e1 = n2 * a0 - n0 * a2, e2 = n0 * a1 - n1 * a0;
      double len = sqrt((e0 * e0 + e1 * e1) + e2 * e2);
      if (len > 0.0) {
        e0 /= len;
        e1 /= len;
        e2 /= len;
        double f0 = n1 * e2 - n2 * e1, f1 = n2 * e0 - n0 * e2, f2 = n0 * e1 - n1 * e0;
        const double *fn = fnv + 3 * (size_t)i;
        const double *pa = V + 3 * (size_t)F[3 * i], *pb = V + 3 * (size_t)F[3 * i + 1],
                     *pc = V + 3 * (size_t)F[3 * i + 2];
        double c0 = ((pa[0] + pb[0]) + pc[0]) / 3.0, c1 = ((pa[1] + pb[1]) + pc[1]) / 3.0,
               c2 = ((pa[2] + pb[2]) + pc[2]) / 3.0;
        V3 o = V3{(float)(c0 - fn[0] * 1e-3), (float)(c1 - fn[1] * 1e-3), (float)(c2 - fn[2] * 1e-3)};
        double d0, d1, d2;
        if (k == 0) {
          d0 = -n0;
          d1 = -n1;
          d2 = -n2;
        } else {
          double cp = cc.cp[k - 1], sp = cc.sp[k - 1];
          d0 = -(cc.cos_a * n0) + cc.sin_a * (cp * e0 + sp * f0);
          d1 = -(cc.cos_a * n1) + cc.sin_a * (cp * e1 + sp * f1);
          d2 = -(cc.cos_a * n2) + cc.sin_a * (cp * e2 + sp * f2);
        }
        V3 dir = V3{(float)d0, (float)d1, (float)d2};
        float t;
        uint32_t prim;
        LdsStack<256, CONE_LSTACK> st{&s_stk[0][threadIdx.x]};
        first_hit_with(st, sv, o, dir, retest, &t, &prim);
        if (prim != 0xFFFFFFFFu && t - t == 0.0f && prim < (uint32_t)nf && t > cc.min_hit) {
          float ratio = t / cc.thick;
          if (ratio > 0.5f && ratio < 1.8f) {
            const double *h = fnv + 3 * (size_t)prim;
            double al = (h[0] * d0 + h[1] * d1) + h[2] * d2;
            vote = al > 0.7;
          }
        }
      }
    }
  }
  const uint64_t bits = __ballot(vote);
  if (k == 0 && j < CONE_FACES_PER_WAVE && q < n)
    good[q] = (uint8_t)__popcll((bits >> (CONE_RAYS * j)) & ((1ull << CONE_RAYS) - 1));
}

// The per-vertex form of the same test (one ray per vertex along the inward normal): the vertex
// normal n (Open3D's area-weighted normal, refine.rs:118-150), the origin pushed 1e-3 along n and
// narrowed to f32, the ray straight down -n, and the cone's first ray's test on its first hit.
__global__ void __launch_bounds__(256) k_vertex_votes(SceneView sv, const double *V, int nv, const uint32_t *F,
                                                      int nf, const double *nrm, float min_hit, float thick,
                                                      int retest, uint8_t *good) {
  __shared__ uint2 s_stk[CONE_LSTACK][256];
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nv) return;
  const double *p = V + 3 * (size_t)q, *n = nrm + 3 * (size_t)q;
  double n0 = n[0], n1 = n[1], n2 = n[2];
  uint8_t g = 0;
  if (n0 - n0 == 0.0 && n1 - n1 == 0.0 && n2 - n2 == 0.0 && (n0 != 0.0 || n1 != 0.0 || n2 != 0.0)) {
    V3 o = V3{(float)(p[0] - n0 * 1e-3), (float)(p[1] - n1 * 1e-3), (float)(p[2] - n2 * 1e-3)};
    double d0 = -n0, d1 = -n1, d2 = -n2;
    V3 dir = V3{(float)d0, (float)d1, (float)d2};
    float t;
    uint32_t prim;
    LdsStack<256, CONE_LSTACK> st{&s_stk[0][threadIdx.x]};
    first_hit_with(st, sv, o, dir, retest, &t, &prim);
    if (prim != 0xFFFFFFFFu && t - t == 0.0f && prim < (uint32_t)nf && t > min_hit) {
      float ratio = t / thick;
      if (ratio > 0.5f && ratio < 1.8f) {
        double h[3];
        face_normal(V, F, prim, h);
        g = ((h[0] * d0 + h[1] * d1) + h[2] * d2) > 0.7 ? 1 : 0;
      }
    }
  }
  good[q] = g;
}

// A face is fracture when fewer than two of its vertices voted shell.
__global__ void k_vertex_labels(const uint32_t *F, int nf, const uint8_t *good, uint8_t *frac) {
  int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  int s = (int)good[F[3 * f]] + (int)good[F[3 * f + 1]] + (int)good[F[3 * f + 2]];
  frac[f] = s < 2 ? 1 : 0;
}

__global__ void k_iota(uint32_t *x, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = (uint32_t)i;
}

}  // namespace

hipError_t launch_vertex_votes(const SceneView &sv, const double *V, int nv, const uint32_t *F, int nf,
                               const double *nrm, double thick, int retest, uint8_t *good, uint8_t *frac,
                               hipStream_t s) {
  if (nv <= 0) return hipSuccess;
  // segment.rs:337-338: the hit floor and the wall in f32.
  float min_hit = (float)(0.1 * thick), th = (float)thick;
  hipLaunchKernelGGL(k_vertex_votes, dim3((nv + 255) / 256), dim3(256), 0, s, sv, V, nv, F, nf, nrm, min_hit, th,
                     retest, good);
  if (frac && nf > 0) hipLaunchKernelGGL(k_vertex_labels, dim3((nf + 255) / 256), dim3(256), 0, s, F, nf, good, frac);
  return hipGetLastError();
}

hipError_t launch_iota(uint32_t *x, int n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_iota, dim3((n + 255) / 256), dim3(256), 0, s, x, n);
  return hipGetLastError();
}

hipError_t launch_cone(const SceneView &sv, const double *V, const uint32_t *F, int nf, const int64_t *faces,
                       const double *ns, int n, const ConeConst &cc, int retest, uint8_t *good, double *fn,
                       hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_face_normals, dim3((nf + 255) / 256), dim3(256), 0, s, V, F, nf, fn);
  const long waves = ((long)n + CONE_FACES_PER_WAVE - 1) / CONE_FACES_PER_WAVE;
  hipLaunchKernelGGL(k_cone, dim3((unsigned)((waves + 3) / 4)), dim3(256), 0, s, sv, V, F, fn, nf, faces, ns, n, cc,
                     retest, good);
  return hipGetLastError();
}

}  // namespace sherd
