// R §9's full-resolution refinement on the device (crates/sherd-core/src/refine.rs): the fracture
// cloud of an original mesh (which vertices, where, and Open3D's vertex normals) and the placement
// of a cloud through a pose. The registrations themselves are the batched ICP ladder's (icp.hip).
//
//   k_select         refine.rs:163-182 select_candidates: vertex i is kept when its nearest
//                    working-mesh face centroid (f64, ties to the lowest index) is a fracture face
//                    strictly closer than max(0.15 t, 1.5 res); an exclusive scan then lists the
//                    kept vertices in ascending order (np.where(sel)[0]).
//   k_face_cross     refine.rs:118-150 vertex_normals, the per-face term (v1 - v0) x (v2 - v0);
//   k_corners        the (vertex, corner) pairs, stably sorted by vertex so that each vertex
//   k_corner_starts  meets its faces in face order, as Open3D's accumulation loop does;
//   k_gather         a cloud's points and normals for a list of vertices: the normal is the sum of
//                    the incident faces' terms in face order, divided once by its length
//                    ((x^2 + y^2) + z^2), or (0, 0, 1) when that length is not positive.
//   k_place          FractureCloud::placed (refine.rs:99-104): points through the 4x4 and normals
//                    through its rotation, each row a fused chain (types.rs:202-214).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "cloud.hpp"
#include "kernels.hpp"
#include "sort.hpp"

namespace sherd {
namespace {

constexpr int RT = 256;

__global__ void __launch_bounds__(RT) k_select(const double *V, int nv, SceneView tree, const double *C,
                                               const uint8_t *fracture, double radius, uint8_t *sel,
                                               uint32_t *pos) {
  __shared__ uint2 s_stk[QSTACK][RT];
  LdsStack<RT, QSTACK> st{&s_stk[0][threadIdx.x]};
  int i = blockIdx.x * RT + threadIdx.x;
  if (i >= nv) return;
  uint32_t j = NONE;
  double d2 = 0.0;
  bool keep = false;
  if (cloud_nearest_with(st, tree, C, V[3 * (size_t)i], V[3 * (size_t)i + 1], V[3 * (size_t)i + 2], __builtin_inf(),
                         &j, &d2))
    keep = fracture[j] != 0 && sqrt(d2) < radius;
  sel[i] = keep ? 1 : 0;
  pos[i] = keep ? 1u : 0u;
}

__global__ void __launch_bounds__(RT) k_compact(const uint8_t *sel, const uint32_t *pos, int nv, uint32_t *idx) {
  int i = blockIdx.x * RT + threadIdx.x;
  if (i < nv && sel[i]) idx[pos[i]] = (uint32_t)i;
}

__global__ void __launch_bounds__(RT) k_face_cross(const double *V, const uint32_t *F, int nf, double *fn) {
  int f = blockIdx.x * RT + threadIdx.x;
  if (f >= nf) return;
  const double *o = V + 3 * (size_t)F[3 * (size_t)f];
  const double *b = V + 3 * (size_t)F[3 * (size_t)f + 1];
  const double *c = V + 3 * (size_t)F[3 * (size_t)f + 2];
  double e1[3] = {b[0] - o[0], b[1] - o[1], b[2] - o[2]};
  double e2[3] = {c[0] - o[0], c[1] - o[1], c[2] - o[2]};
  fn[3 * (size_t)f] = e1[1] * e2[2] - e1[2] * e2[1];
  fn[3 * (size_t)f + 1] = e1[2] * e2[0] - e1[0] * e2[2];
  fn[3 * (size_t)f + 2] = e1[0] * e2[1] - e1[1] * e2[0];
}

__global__ void __launch_bounds__(RT) k_corners(const uint32_t *F, int nc, uint32_t *key, uint32_t *val) {
  int k = blockIdx.x * RT + threadIdx.x;
  if (k >= nc) return;
  key[k] = F[k];
  val[k] = (uint32_t)k;
}

__global__ void __launch_bounds__(RT) k_clear(uint32_t *a, int n, uint32_t v) {
  int k = blockIdx.x * RT + threadIdx.x;
  if (k < n) a[k] = v;
}

// start[v] = the first sorted corner of vertex v (NONE for a vertex no face references).
__global__ void __launch_bounds__(RT) k_corner_starts(const uint32_t *key, int nc, uint32_t *start) {
  int k = blockIdx.x * RT + threadIdx.x;
  if (k >= nc) return;
  if (k == 0 || key[k - 1] != key[k]) start[key[k]] = (uint32_t)k;
}

__global__ void __launch_bounds__(RT) k_gather(const double *V, const uint32_t *idx, int n, const uint32_t *key,
                                               const uint32_t *val, int nc, const uint32_t *start, const double *fn,
                                               double *pts, double *nrm) {
  int k = blockIdx.x * RT + threadIdx.x;
  if (k >= n) return;
  uint32_t v = idx[k];
  for (int a = 0; a < 3; ++a) pts[3 * (size_t)k + a] = V[3 * (size_t)v + a];
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
  for (uint32_t q = start[v]; q != NONE && q < (uint32_t)nc && key[q] == v; ++q) {
    const double *t = fn + 3 * (size_t)(val[q] / 3u);
    s0 += t[0];
    s1 += t[1];
    s2 += t[2];
  }
  double len = sqrt(s0 * s0 + s1 * s1 + s2 * s2);
  if (len > 0.0) {
    s0 /= len;
    s1 /= len;
    s2 /= len;
  } else {
    s0 = 0.0;
    s1 = 0.0;
    s2 = 1.0;
  }
  nrm[3 * (size_t)k] = s0;
  nrm[3 * (size_t)k + 1] = s1;
  nrm[3 * (size_t)k + 2] = s2;
}

__global__ void __launch_bounds__(RT) k_place(const double *p, const double *n, int cnt, Pose T, double *po,
                                              double *no) {
  int k = blockIdx.x * RT + threadIdx.x;
  if (k >= cnt) return;
  const double *m = T.m;
  double x = p[3 * (size_t)k], y = p[3 * (size_t)k + 1], z = p[3 * (size_t)k + 2];
  for (int r = 0; r < 3; ++r)
    po[3 * (size_t)k + r] = fma(m[4 * r + 2], z, fma(m[4 * r + 1], y, m[4 * r] * x)) + m[4 * r + 3];
  if (n && no) {
    double a = n[3 * (size_t)k], b = n[3 * (size_t)k + 1], c = n[3 * (size_t)k + 2];
    for (int r = 0; r < 3; ++r) no[3 * (size_t)k + r] = fma(m[4 * r + 2], c, fma(m[4 * r + 1], b, m[4 * r] * a));
  }
}

unsigned blocks(size_t n) { return (unsigned)((n + RT - 1) / RT); }

}  // namespace

hipError_t launch_select(const double *V, int nv, const SceneView &tree, const double *C, const uint8_t *fracture,
                         double radius, uint8_t *sel, uint32_t *pos, uint32_t *scan_part, hipStream_t s) {
  if (nv <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_select, dim3(blocks(nv)), dim3(RT), 0, s, V, nv, tree, C, fracture, radius, sel, pos);
  return rsort::scan_exclusive(pos, (size_t)nv, scan_part, s);
}

size_t select_scan_words(int nv) { return rsort::scan_parts((size_t)std::max(nv, 1)); }

hipError_t launch_compact(const uint8_t *sel, const uint32_t *pos, int nv, uint32_t *idx, hipStream_t s) {
  if (nv <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_compact, dim3(blocks(nv)), dim3(RT), 0, s, sel, pos, nv, idx);
  return hipGetLastError();
}

size_t normals_scratch_words(int nv, int nf) {
  size_t nc = 3 * (size_t)nf;
  // key, val, their sort buffers, the histogram, the starts
  return 4 * nc + rsort::hist_words((int)std::max<size_t>(nc, 1)) + (size_t)nv;
}

hipError_t launch_vertex_cloud(const double *V, int nv, const uint32_t *F, int nf, const uint32_t *idx, int n,
                               double *fn, uint32_t *scratch, double *pts, double *nrm, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  int nc = 3 * nf;
  uint32_t *key = scratch, *val = key + nc, *tk = val + nc, *tv = tk + nc, *hist = tv + nc;
  uint32_t *start = hist + rsort::hist_words(std::max(nc, 1));
  hipLaunchKernelGGL(k_clear, dim3(blocks(nv)), dim3(RT), 0, s, start, nv, NONE);
  if (nf > 0) {
    hipLaunchKernelGGL(k_face_cross, dim3(blocks(nf)), dim3(RT), 0, s, V, F, nf, fn);
    hipLaunchKernelGGL(k_corners, dim3(blocks(nc)), dim3(RT), 0, s, F, nc, key, val);
    int bits = 8;
    while (bits < 32 && ((uint64_t)1 << bits) < (uint64_t)nv) bits += 8;
    hipError_t e = rsort::sort_pairs<uint32_t>(key, val, tk, tv, hist, nc, 0, bits, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_corner_starts, dim3(blocks(nc)), dim3(RT), 0, s, key, nc, start);
  }
  hipLaunchKernelGGL(k_gather, dim3(blocks(n)), dim3(RT), 0, s, V, idx, n, key, val, nc, start, fn, pts, nrm);
  return hipGetLastError();
}

hipError_t launch_place(const double *p, const double *n, int cnt, const Pose &T, double *po, double *no,
                        hipStream_t s) {
  if (cnt <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_place, dim3(blocks(cnt)), dim3(RT), 0, s, p, n, cnt, T, po, no);
  return hipGetLastError();
}

}  // namespace sherd
