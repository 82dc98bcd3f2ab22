// Shared device arithmetic and layouts for the MI355X backend.
//
// Every parity-relevant expression here is written in one fixed order and the library is built
// with -ffp-contract=off, so no multiply-add is fused: the f32 ray and distance arithmetic, the f64
// pose application and the f64 neighbour distances are the same IEEE operations as the CPU
// executor's (crates/sherd-core/src/spatial/bvh.rs:199-221, executor/cpu.rs:192-200,
// spatial/kdtree.rs), one operation at a time.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define SH_HD __host__ __device__ __forceinline__

namespace sherd {

constexpr int WAVE = 64;

// ---- f32 triples with glam's association (parry3d's `Vector`) ---------------------------------
struct V3 {
  float x, y, z;
};
SH_HD V3 v3(float x, float y, float z) { return V3{x, y, z}; }
SH_HD V3 sub(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
SH_HD V3 scale(V3 a, float s) { return V3{a.x * s, a.y * s, a.z * s}; }
SH_HD float dot(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
SH_HD V3 cross(V3 a, V3 b) { return V3{a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y}; }
SH_HD float len2(V3 a) { return (a.x * a.x + a.y * a.y) + a.z * a.z; }
// std::min's definition: the first operand unless the second is strictly smaller.
SH_HD float mnf(float a, float b) { return (b < a) ? b : a; }
SH_HD float absf(float a) { return __builtin_fabsf(a); }

constexpr float MT_EPS = 1e-9f;       // bvh.rs:204
constexpr float BARY_TOL = 1e-6f;     // barycentric margin that triggers the FP64 retest
constexpr float DET_TOL = 1e-6f;      // relative determinant that triggers the FP64 retest

// Moller-Trumbore exactly as bvh.rs:199-221, with the optional FP64 retest of a near-degenerate
// case (|det| < 1e-6 |e1|_1 |e2|_1 |d|_1, or min(u, v, 1-u-v) within 1e-6 of zero).
// Returns 1 and writes *t for an accepted hit.
SH_HD int mt_hit(V3 o, V3 d, V3 v0, V3 v1, V3 v2, int retest, float *t) {
  V3 e1 = sub(v1, v0);
  V3 e2 = sub(v2, v0);
  V3 pvec = cross(d, e2);
  float det = dot(e1, pvec);
  bool near = false;
  if (retest) {
    float sc = (absf(e1.x) + absf(e1.y) + absf(e1.z)) * (absf(e2.x) + absf(e2.y) + absf(e2.z)) *
               (absf(d.x) + absf(d.y) + absf(d.z));
    near = absf(det) < DET_TOL * sc;
  }
  if (!near) {
    if (absf(det) < MT_EPS) return 0;
    float inv = 1.0f / det;
    V3 tvec = sub(o, v0);
    float u = inv * dot(tvec, pvec);
    V3 qvec = cross(tvec, e1);
    float v = inv * dot(d, qvec);
    if (retest) {
      float w = (1.0f - u) - v;
      float m = mnf(u, mnf(v, w));
      if (!(absf(m) < BARY_TOL)) {
        if (!(m >= BARY_TOL)) return 0;
        float tt = inv * dot(e2, qvec);
        if (!(tt > MT_EPS)) return 0;
        *t = tt;
        return 1;
      }
      near = true;
    } else {
      if (!(u >= 0.0f && u <= 1.0f)) return 0;
      if (v < 0.0f || u + v > 1.0f) return 0;
      float tt = inv * dot(e2, qvec);
      if (!(tt > MT_EPS)) return 0;
      *t = tt;
      return 1;
    }
  }
  // FP64 retest from the same f32 data.
  double ox = o.x, oy = o.y, oz = o.z, dx = d.x, dy = d.y, dz = d.z;
  double ax = v0.x, ay = v0.y, az = v0.z;
  double e1x = (double)v1.x - ax, e1y = (double)v1.y - ay, e1z = (double)v1.z - az;
  double e2x = (double)v2.x - ax, e2y = (double)v2.y - ay, e2z = (double)v2.z - az;
  double px = dy * e2z - e2y * dz, py = dz * e2x - e2z * dx, pz = dx * e2y - e2x * dy;
  double dt = e1x * px + e1y * py + e1z * pz;
  if (dt == 0.0 || !(dt - dt == 0.0)) return 0;
  double inv = 1.0 / dt;
  double tx = ox - ax, ty = oy - ay, tz = oz - az;
  double u = inv * (tx * px + ty * py + tz * pz);
  if (!(u >= 0.0 && u <= 1.0)) return 0;
  double qx = ty * e1z - e1y * tz, qy = tz * e1x - e1z * tx, qz = tx * e1y - e1x * ty;
  double v = inv * (dx * qx + dy * qy + dz * qz);
  if (v < 0.0 || u + v > 1.0) return 0;
  double tt = inv * (e2x * qx + e2y * qy + e2z * qz);
  if (!(tt > 1e-9)) return 0;
  *t = (float)tt;
  return 1;
}

// Squared distance from p to the segment a + s (b - a), s in [0, 1], given ap = p - a, ab = b - a.
SH_HD float seg_d2(V3 ap, V3 ab) {
  float t = dot(ab, ap);
  if (t <= 0.0f) return len2(ap);
  float l = dot(ab, ab);
  if (t >= l) return len2(sub(ap, ab));
  float s = t / l;
  return len2(sub(ap, scale(ab, s)));
}

// Squared point-triangle distance in f32: the Voronoi regions of the triangle (after Ericson,
// Real-Time Collision Detection 5.1.5), offsets from the nearest vertex, the interior through the
// plane equation s^2 / |n|^2.
SH_HD float tri_d2(V3 p, V3 a, V3 b, V3 c) {
  V3 ab = sub(b, a), ac = sub(c, a), ap = sub(p, a);
  float d1 = dot(ab, ap), d2 = dot(ac, ap);
  if (d1 <= 0.0f && d2 <= 0.0f) return len2(ap);
  V3 bp = sub(p, b);
  float d3 = dot(ab, bp), d4 = dot(ac, bp);
  if (d3 >= 0.0f && d4 <= d3) return len2(bp);
  float vc = d1 * d4 - d3 * d2;
  if (vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f) {
    float v = d1 / (d1 - d3);
    return len2(sub(ap, scale(ab, v)));
  }
  V3 cp = sub(p, c);
  float d5 = dot(ab, cp), d6 = dot(ac, cp);
  if (d6 >= 0.0f && d5 <= d6) return len2(cp);
  float vb = d5 * d2 - d1 * d6;
  if (vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f) {
    float w = d2 / (d2 - d6);
    return len2(sub(ap, scale(ac, w)));
  }
  float va = d3 * d6 - d5 * d4;
  if (va <= 0.0f && (d4 - d3) >= 0.0f && (d5 - d6) >= 0.0f) {
    float w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
    return len2(sub(bp, scale(sub(c, b), w)));
  }
  V3 n = cross(ab, ac);
  float nn = dot(n, n);
  if (!(nn > 0.0f)) {
    float e = mnf(seg_d2(ap, ab), seg_d2(ap, ac));
    return mnf(e, seg_d2(bp, sub(c, b)));
  }
  float s = dot(ap, n);
  return (s * s) / nn;
}

// mesh/geometry.rs:70-96: a face's unit normal in f64, the cross product of its two edges from
// the first vertex over max(|n|, 1e-12).
SH_HD void face_normal(const double *V, const uint32_t *F, uint32_t f, double *n) {
  const double *a = V + 3 * (size_t)F[3 * f], *b = V + 3 * (size_t)F[3 * f + 1], *c = V + 3 * (size_t)F[3 * f + 2];
  double e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
  double e2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
  double x = e1[1] * e2[2] - e1[2] * e2[1], y = e1[2] * e2[0] - e1[0] * e2[2], z = e1[0] * e2[1] - e1[1] * e2[0];
  double len = sqrt((x * x + y * y) + z * z);
  double den = len < 1e-12 ? 1e-12 : len;  // f64::max(len, 1e-12)
  n[0] = x / den;
  n[1] = y / den;
  n[2] = z / den;
}

// The same face's centroid, ((a + b) + c) / 3 per coordinate.
SH_HD void face_centroid(const double *V, const uint32_t *F, uint32_t f, double *c) {
  const double *a = V + 3 * (size_t)F[3 * f], *b = V + 3 * (size_t)F[3 * f + 1], *d = V + 3 * (size_t)F[3 * f + 2];
  for (int k = 0; k < 3; ++k) c[k] = ((a[k] + b[k]) + d[k]) / 3.0;
}

// executor/cpu.rs:192-200: T.p in f64, left to right, then narrowed once to f32.
SH_HD V3 moved_f32(const double *T, double px, double py, double pz) {
  float x = (float)(T[0] * px + T[1] * py + T[2] * pz + T[3]);
  float y = (float)(T[4] * px + T[5] * py + T[6] * pz + T[7]);
  float z = (float)(T[8] * px + T[9] * py + T[10] * pz + T[11]);
  return V3{x, y, z};
}

// ---- BVH layout --------------------------------------------------------------------------------
// 2n-1 nodes in pre-order, the root at 0. `a` >= 0 marks an internal node and is its right child
// (the left child is always the next node); a leaf holds a = ~(its triangle slot). `esc` is the
// next node in pre-order after this node's subtree, or -1. A descent thus knows both children's
// addresses from the parent and loads them together. 32 bytes, one dwordx4 pair per node.
struct Node {
  float lo[3];
  int a;
  float hi[3];
  int esc;
};

// A triangle in leaf order: the three f32 vertices and the scene-local face index.
struct Tri {
  float v[9];
  uint32_t face;
  uint32_t pad[2];
};

// The triangles an axis-parallel ray can cross, by the cell of the ray's two other coordinates:
// cell (cu, cv) of a uniform grid over the scene's box lists every triangle whose box, grown by
// the pruning slack, overlaps it. A +axis ray's crossings are the hits among its cell's triangles.
struct AxisGrid {
  uint32_t *start;  // nu * nv + 1 offsets into the item arrays
  // Each listed triangle's coordinates, copied into its cell's run so that a ray reads its cell
  // contiguously: v0.xyz v1.x | v1.yz v2.xy | v2.z (the Tri record's v[0..8]).
  float4 *ix, *iy;
  float *iz;
  float org_u, org_v, inv;
  int nu, nv;
};

// A lower bound of the distance to the scene from anywhere: a uniform grid over the scene's box
// (one cell of margin) holding, per cell, a bound no point of the cell can beat; outside the grid
// the distance to the grid's box bounds it.
struct DistBound {
  float *lb;  // n[0] * n[1] * n[2] bounds, x fastest
  float org[3], h;
  int n[3];
  float ub_add;  // what dist_upper adds back: two cell diagonals, the largest triangle box, the pads
  // A cut of the tree: up to CUT_BOXES nodes whose subtrees together hold every triangle once, as
  // 8 floats each (lo xyz, the node index's bits, hi xyz, 0). Far outside the grid, where its
  // bound degrades to the distance to the grid's box, the nearest of these boxes bounds the
  // distance far more tightly (a curved shell's boxes hug it; its box does not).
  float *cut;
  int n_cut;
};
constexpr int CUT_BOXES = 64;

struct SceneView {
  const Node *nodes;
  const Tri *tris;
  AxisGrid pg[3];  // parity grids, one per ray axis (null start until built)
  DistBound db;    // distance bounds (null lb until built)
  int n;         // triangles
  int root;      // root node index
  float lo[3];   // aabb of the scene's vertices (RayScene::aabb)
  float hi[3];
  float absmax;  // largest |coordinate| of the scene
};

SH_HD V3 tri_v(const Tri &t, int k) { return V3{t.v[3 * k], t.v[3 * k + 1], t.v[3 * k + 2]}; }

// Cell of coordinate x on a grid axis: floor((x - org) * inv) clamped to [0, n - 1], computed the
// same way when a grid is filled and when it is queried, and monotone in x.
SH_HD int grid_cell(float x, float org, float inv, int n) {
  float f = (x - org) * inv;
  if (!(f >= 0.0f)) return 0;
  if (f >= (float)n) return n - 1;
  int c = (int)f;
  return c < n ? c : n - 1;
}

// A lower bound of the distance from p to the scene (see DistBound); 0 where nothing is known.
// Outside the grid's box, p' (p clamped to the box) has the bound of its cell, and for any scene
// point q inside the box |p - q|^2 >= |p - p'|^2 + |p' - q|^2 (projection onto a convex set).
SH_HD float dist_bound(const DistBound &g, V3 p) {
  if (!g.lb) return 0.0f;
  float q[3] = {p.x, p.y, p.z};
  float out2 = 0.0f;
  int c[3];
  for (int k = 0; k < 3; ++k) {
    float lo = g.org[k], hi = g.org[k] + g.h * (float)g.n[k];
    float e = q[k] < lo ? lo - q[k] : (q[k] > hi ? q[k] - hi : 0.0f);
    out2 += e * e;
    c[k] = grid_cell(q[k], g.org[k], 1.0f / g.h, g.n[k]);
  }
  float in = g.lb[((size_t)c[2] * g.n[1] + c[1]) * g.n[0] + c[0]];
  if (out2 == 0.0f) return in;
  return sqrtf(out2 + in * in) * 0.999f;
}

// An upper bound of the distance from p to the mesh: the stored bound is (e - sqrt 3) h for the
// distance e h between its cell's centre and the nearest marked cell's, and a marked cell meets the
// (padded) box of a triangle, so a point of that triangle lies within the cell's half diagonal plus
// the triangle box's diagonal of its centre. Outside the grid, plus the distance to the grid. +inf
// without bounds.
SH_HD float dist_upper(const DistBound &g, V3 p) {
  if (!g.lb) return __builtin_inff();
  float q[3] = {p.x, p.y, p.z};
  float out2 = 0.0f;
  int c[3];
  for (int k = 0; k < 3; ++k) {
    float lo = g.org[k], hi = g.org[k] + g.h * (float)g.n[k];
    float e = q[k] < lo ? lo - q[k] : (q[k] > hi ? q[k] - hi : 0.0f);
    out2 += e * e;
    c[k] = grid_cell(q[k], g.org[k], 1.0f / g.h, g.n[k]);
  }
  float in = g.lb[((size_t)c[2] * g.n[1] + c[1]) * g.n[0] + c[0]];
  float u = in / 0.999f + g.ub_add;
  return (out2 == 0.0f ? u : sqrtf(out2) + u) * 1.0001f;
}

// The pruning slack: ~80 ulps of the largest coordinate in play, more than the f32 rounding of
// any distance or hit parameter computed above, so a box can never hide the triangle that attains
// the computed minimum (the oracle prunes with the same bound).
SH_HD float slack_of(float absmax, V3 p) {
  float m = absmax;
  float ax = absf(p.x), ay = absf(p.y), az = absf(p.z);
  m = m < ax ? ax : m;
  m = m < ay ? ay : m;
  m = m < az ? az : m;
  return 1e-5f * m + 1e-30f;
}

// Squared distance from p to the node's box grown by pad (conservative). An axis's gap is the
// larger of lo - x and x - hi, or 0 inside the slab: the values of the case split (lo - x when x is
// below, x - hi when above), taken without branches.
SH_HD float box_gap(float lo, float hi, float x) { return fmaxf(fmaxf(lo - x, x - hi), 0.0f); }
SH_HD float box_d2(const Node &nd, V3 p, float pad) {
  float g0 = box_gap(nd.lo[0] - pad, nd.hi[0] + pad, p.x);
  float g1 = box_gap(nd.lo[1] - pad, nd.hi[1] + pad, p.y);
  float g2 = box_gap(nd.lo[2] - pad, nd.hi[2] + pad, p.z);
  return (g0 * g0 + g1 * g1) + g2 * g2;
}

// ---- f64 neighbour distance (kiddo's SquaredEuclidean fold: (dx^2 + dy^2) + dz^2) --------------
SH_HD double d2_f64(double qx, double qy, double qz, double px, double py, double pz) {
  double ex = qx - px, ey = qy - py, ez = qz - pz;
  return (ex * ex + ey * ey) + ez * ez;
}

// f64::total_cmp order as an unsigned key, and back.
SH_HD uint64_t total_key(double v) {
  uint64_t u = __builtin_bit_cast(uint64_t, v);
  return (u & 0x8000000000000000ull) ? ~u : (u | 0x8000000000000000ull);
}
SH_HD double from_total_key(uint64_t k) {
  uint64_t u = (k & 0x8000000000000000ull) ? (k & 0x7fffffffffffffffull) : ~k;
  return __builtin_bit_cast(double, u);
}

// ---- loads from device memory ------------------------------------------------------------------
// Scene, tree and cloud pointers reach the kernels inside structures, so the compiler cannot tell
// they address global memory and emits flat loads, which also wait on the LDS counter. These read
// through the global address space instead (every such buffer is device or pinned memory).
typedef float f32x4_t __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(1))) const f32x4_t g_f32x4;
typedef __attribute__((address_space(1))) const float g_f32;
typedef __attribute__((address_space(1))) const uint32_t g_u32;
typedef __attribute__((address_space(1))) const double g_f64;

__device__ __forceinline__ float4 gld4(const void *p) {
  f32x4_t v = *(g_f32x4 *)p;
  return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ float gldf(const float *p) { return *(g_f32 *)p; }
__device__ __forceinline__ uint32_t gldu(const uint32_t *p) { return *(g_u32 *)p; }
__device__ __forceinline__ double gldd(const double *p) { return *(g_f64 *)p; }

// ---- wave64 helpers ----------------------------------------------------------------------------
// How many lanes below this one are set in m.
__device__ __forceinline__ uint32_t lanes_below(uint64_t m) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// Consecutive slots of *counter for the lanes that take one, with one atomic per wave. Every
// active lane of the wave calls it.
__device__ __forceinline__ uint32_t wave_slot(uint32_t *counter, bool take) {
  uint64_t m = __ballot(take);
  if (m == 0) return 0;
  uint32_t rank = lanes_below(m);
  uint32_t base = 0;
  if (take && rank == 0) base = atomicAdd(counter, (uint32_t)__popcll(m));
  base = __shfl(base, __ffsll((unsigned long long)m) - 1);
  return base + rank;
}

// Maximum and minimum over the 64 lanes; every lane of the wave must be active.
__device__ __forceinline__ uint32_t wave_max(uint32_t v) {
  for (int o = 32; o > 0; o >>= 1) {
    uint32_t w = (uint32_t)__shfl_xor((int)v, o);
    v = v > w ? v : w;
  }
  return v;
}
__device__ __forceinline__ uint32_t wave_min(uint32_t v) {
  for (int o = 32; o > 0; o >>= 1) {
    uint32_t w = (uint32_t)__shfl_xor((int)v, o);
    v = v < w ? v : w;
  }
  return v;
}

// MI355X dispatches consecutive workgroups of a launch to its 8 XCDs in turn, and each XCD has
// its own L2. This maps the physical workgroup id to a logical one so that each XCD runs one
// contiguous run of the logical order: work laid out by shared data then shares one L2.
constexpr uint32_t N_XCD = 8;
__device__ __forceinline__ uint32_t xcd_logical(uint32_t phys, uint32_t n) {
  uint32_t q = n / N_XCD, r = n % N_XCD, x = phys % N_XCD, k = phys / N_XCD;
  return x * q + (x < r ? x : r) + k;
}

}  // namespace sherd
