// Oracle, geometry half: the f32 triangle scene and the queries CpuExecutor makes of it.
//
// TEST INFRASTRUCTURE ONLY (see oracle.h). The acceleration structure here is a plain CPU BVH
// whose pruning is deliberately conservative, so every query answers exactly what a scan over all
// triangles answers; `brute` runs that scan and pins the BVH.
//
// Arithmetic contract shared with the device kernels (compiled with -ffp-contract=off on both
// sides, so no multiply-add is fused):
//   * Moller-Trumbore exactly as bvh.rs:199-221 writes it, with glam's cross/dot association;
//   * the closest point: own algorithm (parry's is not vendored), Voronoi regions of the triangle
//     measured relative to the nearest vertex, interior distance through the plane equation;
//   * the FP64 retest of a near-degenerate ray hit: |det| < 1e-6 * |e1|_1 |e2|_1 |d|_1, or a
//     barycentric margin min(u, v, 1-u-v) inside (-1e-6, 1e-6).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <vector>

#include "oracle.h"
#include "internal.hpp"

namespace {

using oracle::V3;

const float EPS = 1e-9f;        // bvh.rs:204, the E4 harness constant
const float BARY_TOL = 1e-6f;   // barycentric margin of the FP64 retest
const float DET_TOL = 1e-6f;    // relative determinant of the FP64 retest

struct Node {
  float lo[3], hi[3];
  int left, right;   // children, or -1
  int first, count;  // leaf range into `order`
};

}  // namespace

struct oracle_scene {
  std::vector<V3> v;                    // f32 vertices
  std::vector<uint32_t> f;              // 3 per scene face
  std::vector<Node> nodes;
  std::vector<int> order;               // leaf order of scene faces
  float lo[3], hi[3];                   // aabb of the scene's vertices
  float absmax;                         // largest |coordinate| in the scene
};

namespace {

inline V3 vert(const oracle_scene *s, size_t face, int k) { return s->v[s->f[3 * face + k]]; }

int build_node(oracle_scene *s, std::vector<V3> &cent, int first, int count) {
  Node n;
  for (int k = 0; k < 3; ++k) {
    n.lo[k] = std::numeric_limits<float>::infinity();
    n.hi[k] = -std::numeric_limits<float>::infinity();
  }
  for (int i = first; i < first + count; ++i) {
    for (int c = 0; c < 3; ++c) {
      V3 p = vert(s, s->order[i], c);
      for (int k = 0; k < 3; ++k) {
        n.lo[k] = std::min(n.lo[k], p[k]);
        n.hi[k] = std::max(n.hi[k], p[k]);
      }
    }
  }
  n.left = n.right = -1;
  n.first = first;
  n.count = count;
  int id = (int)s->nodes.size();
  s->nodes.push_back(n);
  if (count <= 4) return id;
  float clo[3], chi[3];
  for (int k = 0; k < 3; ++k) {
    clo[k] = std::numeric_limits<float>::infinity();
    chi[k] = -std::numeric_limits<float>::infinity();
  }
  for (int i = first; i < first + count; ++i)
    for (int k = 0; k < 3; ++k) {
      clo[k] = std::min(clo[k], cent[s->order[i]][k]);
      chi[k] = std::max(chi[k], cent[s->order[i]][k]);
    }
  int axis = 0;
  for (int k = 1; k < 3; ++k)
    if (chi[k] - clo[k] > chi[axis] - clo[axis]) axis = k;
  int mid = first + count / 2;
  std::nth_element(s->order.begin() + first, s->order.begin() + mid, s->order.begin() + first + count,
                   [&](int a, int b) {
                     if (cent[a][axis] != cent[b][axis]) return cent[a][axis] < cent[b][axis];
                     return a < b;
                   });
  int l = build_node(s, cent, first, mid - first);
  int r = build_node(s, cent, mid, first + count - mid);
  s->nodes[id].left = l;
  s->nodes[id].right = r;
  return id;
}

// Moller-Trumbore, bvh.rs:199-221, returning the parameter of an accepted hit.
// Returns 0 for no hit, 1 for a hit at *t. With retest, a near-degenerate case is recomputed in
// f64 from the same f32 inputs and that answer replaces the f32 one.
int mt_hit(V3 o, V3 d, V3 v0, V3 v1, V3 v2, int retest, float *t) {
  V3 e1 = oracle::sub(v1, v0);
  V3 e2 = oracle::sub(v2, v0);
  V3 pvec = oracle::cross(d, e2);
  float det = oracle::dot(e1, pvec);
  bool near = false;
  if (retest) {
    float scale = (std::fabs(e1.x) + std::fabs(e1.y) + std::fabs(e1.z)) *
                  (std::fabs(e2.x) + std::fabs(e2.y) + std::fabs(e2.z)) *
                  (std::fabs(d.x) + std::fabs(d.y) + std::fabs(d.z));
    near = std::fabs(det) < DET_TOL * scale;
  }
  if (!near) {
    if (std::fabs(det) < EPS) return 0;
    float inv = 1.0f / det;
    V3 tvec = oracle::sub(o, v0);
    float u = inv * oracle::dot(tvec, pvec);
    V3 qvec = oracle::cross(tvec, e1);
    float v = inv * oracle::dot(d, qvec);
    if (retest) {
      float w = (1.0f - u) - v;
      float m = std::min(u, std::min(v, w));
      if (!(std::fabs(m) < BARY_TOL)) {
        if (!(m >= BARY_TOL)) return 0;
        float tt = inv * oracle::dot(e2, qvec);
        if (!(tt > EPS)) return 0;
        *t = tt;
        return 1;
      }
      near = true;
    } else {
      if (!(u >= 0.0f && u <= 1.0f)) return 0;
      if (v < 0.0f || u + v > 1.0f) return 0;
      float tt = inv * oracle::dot(e2, qvec);
      if (!(tt > EPS)) return 0;
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
  if (dt == 0.0 || !std::isfinite(dt)) return 0;
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

inline float len2(V3 a) { return (a.x * a.x + a.y * a.y) + a.z * a.z; }

// Squared distance from p to the segment a + s*(b - a), s in [0, 1], relative to a.
inline float seg_d2(V3 ap, V3 ab) {
  float t = oracle::dot(ab, ap);
  if (t <= 0.0f) return len2(ap);
  float l = oracle::dot(ab, ab);
  if (t >= l) return len2(oracle::sub(ap, ab));
  float s = t / l;
  return len2(oracle::sub(ap, oracle::scale(ab, s)));
}

}  // namespace

namespace oracle {

// Squared point-triangle distance in f32: Voronoi regions (after Ericson, Real-Time Collision
// Detection, 5.1.5), every offset taken from the nearest vertex, the interior through the plane
// equation (s^2 / |n|^2). The device kernels execute these operations in this order.
float tri_d2(V3 p, V3 a, V3 b, V3 c) {
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
    float e = std::min(seg_d2(ap, ab), seg_d2(ap, ac));
    return std::min(e, seg_d2(bp, sub(c, b)));
  }
  float s = dot(ap, n);
  return (s * s) / nn;
}

}  // namespace oracle

namespace {

// Conservative slack for the pruning tests: the f32 arithmetic above is accurate to a few ulps of
// the largest coordinate, and the slack is ~80 of them, so a box can never hide the triangle that
// attains the computed minimum.
inline double slack_of(const oracle_scene *s, V3 p) {
  double m = std::max<double>(s->absmax, std::max(std::fabs(p.x), std::max(std::fabs(p.y), std::fabs(p.z))));
  return 1e-5 * m + 1e-30;
}

inline double box_d2(const Node &n, V3 p, double pad) {
  double d2 = 0.0;
  double q[3] = {p.x, p.y, p.z};
  for (int k = 0; k < 3; ++k) {
    double lo = (double)n.lo[k] - pad, hi = (double)n.hi[k] + pad, g = 0.0;
    if (q[k] < lo) g = lo - q[k];
    else if (q[k] > hi) g = q[k] - hi;
    d2 += g * g;
  }
  return d2;
}

// Slab test in f64 against the padded box: the entry parameter, or +inf when the ray misses.
inline double box_enter(const Node &n, V3 o, V3 d, double pad) {
  double oo[3] = {o.x, o.y, o.z}, dd[3] = {d.x, d.y, d.z};
  double t0 = 0.0, t1 = std::numeric_limits<double>::infinity();
  for (int k = 0; k < 3; ++k) {
    double lo = (double)n.lo[k] - pad, hi = (double)n.hi[k] + pad;
    if (dd[k] == 0.0) {
      if (oo[k] < lo || oo[k] > hi) return std::numeric_limits<double>::infinity();
      continue;
    }
    double a = (lo - oo[k]) / dd[k], b = (hi - oo[k]) / dd[k];
    if (a > b) std::swap(a, b);
    t0 = std::max(t0, a);
    t1 = std::min(t1, b);
    if (t0 > t1) return std::numeric_limits<double>::infinity();
  }
  return t0;
}

template <class Leaf, class Enter>
void traverse(const oracle_scene *s, Enter enter, Leaf leaf) {
  int stack[128];
  int sp = 0;
  stack[sp++] = 0;
  while (sp) {
    int id = stack[--sp];
    const Node &n = s->nodes[id];
    if (!enter(n)) continue;
    if (n.left < 0) {
      for (int i = n.first; i < n.first + n.count; ++i) leaf(s->order[i]);
      continue;
    }
    stack[sp++] = n.right;
    stack[sp++] = n.left;
  }
}

float min_d2(const oracle_scene *s, V3 p, float window_d, int brute) {
  float best = std::numeric_limits<float>::infinity();
  size_t nf = s->f.size() / 3;
  if (brute) {
    for (size_t i = 0; i < nf; ++i) best = std::min(best, oracle::tri_d2(p, vert(s, i, 0), vert(s, i, 1), vert(s, i, 2)));
    return best;
  }
  double slack = slack_of(s, p);
  traverse(
      s,
      [&](const Node &n) {
        double lim = std::min<double>(std::sqrt((double)best), (double)window_d) + slack;
        return box_d2(n, p, slack) <= lim * lim;
      },
      [&](int i) { best = std::min(best, oracle::tri_d2(p, vert(s, i, 0), vert(s, i, 1), vert(s, i, 2))); });
  return best;
}

}  // namespace

extern "C" {

oracle_scene *oracle_scene_new(const float *V, size_t nv, const uint32_t *F, size_t nf,
                               const uint8_t *mask) {
  if (nf == 0) return nullptr;
  oracle_scene *s = new oracle_scene();
  s->v.resize(nv);
  for (size_t i = 0; i < nv; ++i) s->v[i] = V3{V[3 * i], V[3 * i + 1], V[3 * i + 2]};
  for (size_t i = 0; i < nf; ++i)
    if (!mask || mask[i]) s->f.insert(s->f.end(), F + 3 * i, F + 3 * i + 3);
  if (s->f.empty()) s->f.insert(s->f.end(), F, F + 3);  // bvh.rs:71-76: first face
  size_t m = s->f.size() / 3;
  for (int k = 0; k < 3; ++k) {
    s->lo[k] = std::numeric_limits<float>::infinity();
    s->hi[k] = -std::numeric_limits<float>::infinity();
  }
  s->absmax = 0.0f;
  for (uint32_t idx : s->f) {
    V3 p = s->v[idx];
    for (int k = 0; k < 3; ++k) {
      s->lo[k] = std::min(s->lo[k], p[k]);
      s->hi[k] = std::max(s->hi[k], p[k]);
      s->absmax = std::max(s->absmax, std::fabs(p[k]));
    }
  }
  std::vector<V3> cent(m);
  for (size_t i = 0; i < m; ++i) {
    V3 a = vert(s, i, 0), b = vert(s, i, 1), c = vert(s, i, 2);
    cent[i] = V3{(a.x + b.x + c.x) / 3.0f, (a.y + b.y + c.y) / 3.0f, (a.z + b.z + c.z) / 3.0f};
  }
  s->order.resize(m);
  for (size_t i = 0; i < m; ++i) s->order[i] = (int)i;
  s->nodes.reserve(2 * m / 2 + 8);
  build_node(s, cent, 0, (int)m);
  return s;
}

void oracle_scene_free(oracle_scene *s) { delete s; }
size_t oracle_scene_faces(const oracle_scene *s) { return s->f.size() / 3; }
void oracle_scene_aabb(const oracle_scene *s, float lo[3], float hi[3]) {
  for (int k = 0; k < 3; ++k) {
    lo[k] = s->lo[k];
    hi[k] = s->hi[k];
  }
}

void oracle_first_hit(const oracle_scene *s, const float *orig, const float *dir, size_t n,
                      int retest, int brute, float *t_out, uint32_t *prim_out) {
#pragma omp parallel for schedule(dynamic, 256)
  for (long r = 0; r < (long)n; ++r) {
    V3 o{orig[3 * r], orig[3 * r + 1], orig[3 * r + 2]};
    V3 d{dir[3 * r], dir[3 * r + 1], dir[3 * r + 2]};
    float best = std::numeric_limits<float>::infinity();
    uint32_t prim = 0xFFFFFFFFu;
    auto leaf = [&](int i) {
      float t;
      if (mt_hit(o, d, vert(s, i, 0), vert(s, i, 1), vert(s, i, 2), retest, &t)) {
        if (t < best || (t == best && (uint32_t)i < prim)) {
          best = t;
          prim = (uint32_t)i;
        }
      }
    };
    if (brute) {
      for (size_t i = 0; i < s->f.size() / 3; ++i) leaf((int)i);
    } else {
      double dm = std::max(std::fabs(d.x), std::max(std::fabs(d.y), std::fabs(d.z)));
      double pad = slack_of(s, o);
      traverse(
          s,
          [&](const Node &nd) {
            double te = box_enter(nd, o, d, pad);
            return te < std::numeric_limits<double>::infinity() &&
                   te <= (double)best + pad / (dm > 0 ? dm : 1.0);
          },
          leaf);
    }
    t_out[r] = best;
    prim_out[r] = prim;
  }
}

uint32_t oracle_crossings(const oracle_scene *s, const float o_[3], const float d_[3], int retest,
                          int brute) {
  V3 o{o_[0], o_[1], o_[2]}, d{d_[0], d_[1], d_[2]};
  uint32_t hits = 0;
  auto leaf = [&](int i) {
    float t;
    hits += (uint32_t)mt_hit(o, d, vert(s, i, 0), vert(s, i, 1), vert(s, i, 2), retest, &t);
  };
  if (brute) {
    for (size_t i = 0; i < s->f.size() / 3; ++i) leaf((int)i);
  } else {
    double pad = slack_of(s, o);
    traverse(s, [&](const Node &nd) { return box_enter(nd, o, d, pad) < std::numeric_limits<double>::infinity(); }, leaf);
  }
  return hits;
}

float oracle_distance(const oracle_scene *s, const float p[3], int brute) {
  return std::sqrt(min_d2(s, V3{p[0], p[1], p[2]}, std::numeric_limits<float>::infinity(), brute));
}

}  // extern "C"

namespace oracle {

// cpu.rs:192-200: T.p in f64 (left to right), then narrowed to f32.
V3 moved_f32(const double T[16], const double *p) {
  float out[3];
  for (int i = 0; i < 3; ++i)
    out[i] = (float)(T[4 * i + 0] * p[0] + T[4 * i + 1] * p[1] + T[4 * i + 2] * p[2] + T[4 * i + 3]);
  return V3{out[0], out[1], out[2]};
}

// RayScene::bounded_distance (bvh.rs:123-135): the distance when it is at most max_dist.
bool bounded(const oracle_scene *s, V3 p, float max_dist, float *d) {
  if (!(max_dist > 0.0f)) return false;
  float d2 = min_d2(s, p, max_dist, 0);
  float dist = std::sqrt(d2);
  if (!(dist <= max_dist)) return false;
  *d = dist;
  return true;
}

// RayScene::inside (bvh.rs:146-158): three axis rays, majority of odd crossing counts.
bool inside(const oracle_scene *s, V3 p, int retest) {
  int votes = 0;
  for (int axis = 0; axis < 3; ++axis) {
    float d[3] = {0.0f, 0.0f, 0.0f};
    d[axis] = 1.0f;
    float o[3] = {p.x, p.y, p.z};
    if (oracle_crossings(s, o, d, retest, 0) % 2 == 1) votes += 1;
  }
  return votes >= 2;
}

}  // namespace oracle

extern "C" {

void oracle_exec_bounded_distance(const oracle_scene *s, const double *pts, size_t n,
                                  const double T[16], double max_dist, int reduce_min,
                                  double *out) {
  float window = (float)max_dist;
  if (!reduce_min) {
#pragma omp parallel for schedule(dynamic, 256)
    for (long i = 0; i < (long)n; ++i) {
      float d;
      out[i] = oracle::bounded(s, oracle::moved_f32(T, pts + 3 * i), window, &d)
                   ? (double)d
                   : std::numeric_limits<double>::infinity();
    }
    return;
  }
  // cpu.rs:100-111: a running minimum whose window shrinks to the best so far; the answer is the
  // minimum of the window and every distance, which a parallel min computes identically.
  float best = window;
#pragma omp parallel
  {
    float mine = window;
#pragma omp for schedule(dynamic, 256)
    for (long i = 0; i < (long)n; ++i) {
      float d;
      if (oracle::bounded(s, oracle::moved_f32(T, pts + 3 * i), mine, &d)) mine = std::min(mine, d);
    }
#pragma omp critical
    best = std::min(best, mine);
  }
  out[0] = (double)best;
}

void oracle_exec_inside(const oracle_scene *s, const double *pts, size_t n, const double T[16],
                        int retest, float *depth, uint32_t *inside) {
#pragma omp parallel for schedule(dynamic, 128)
  for (long i = 0; i < (long)n; ++i) {
    oracle::V3 p = oracle::moved_f32(T, pts + 3 * i);
    bool out_box = false;
    for (int k = 0; k < 3; ++k) out_box |= (p[k] < s->lo[k] || p[k] > s->hi[k]);
    if (out_box || !oracle::inside(s, p, retest)) {
      depth[i] = 0.0f;
      inside[i] = 0;
      continue;
    }
    depth[i] = std::sqrt(min_d2(s, p, std::numeric_limits<float>::infinity(), 0));
    inside[i] = 1;
  }
}

// segment::classify_faces (segment.rs:318-381), `cone_basis` (:383-396), constants (:53-62).
void oracle_cone_votes(const oracle_scene *s, const double *V, size_t nv, const uint32_t *F,
                       size_t nf, const int64_t *faces, const double *ns, size_t n_faces,
                       double thick, int retest, uint8_t *good) {
  (void)nv;
  const double PI = 3.14159265358979323846;
  double angle = 15.0 * (PI / 180.0);
  double cos_a = std::cos(angle), sin_a = std::sin(angle);
  float min_hit = (float)(0.1 * thick);
  float thick32 = (float)thick;
  auto face_normal = [&](size_t f, double out[3]) {
    const double *a = V + 3 * F[3 * f], *b = V + 3 * F[3 * f + 1], *c = V + 3 * F[3 * f + 2];
    double e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
    double e2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
    double n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
    double len = std::sqrt((n[0] * n[0] + n[1] * n[1]) + n[2] * n[2]);
    double den = std::max(len, 1e-12);
    out[0] = n[0] / den;
    out[1] = n[1] / den;
    out[2] = n[2] / den;
  };
#pragma omp parallel for schedule(dynamic, 64)
  for (long q = 0; q < (long)n_faces; ++q) {
    size_t i = (size_t)faces[q];
    const double *n = ns + 3 * q;
    good[q] = 0;
    if (!(std::isfinite(n[0]) && std::isfinite(n[1]) && std::isfinite(n[2]))) continue;
    double a[3] = {0, 0, 0};
    if (std::fabs(n[0]) < 0.9) a[0] = 1.0; else a[1] = 1.0;
    double e1[3] = {n[1] * a[2] - n[2] * a[1], n[2] * a[0] - n[0] * a[2], n[0] * a[1] - n[1] * a[0]};
    double len = std::sqrt((e1[0] * e1[0] + e1[1] * e1[1]) + e1[2] * e1[2]);
    if (!(len > 0.0)) continue;
    e1[0] /= len; e1[1] /= len; e1[2] /= len;
    double e2[3] = {n[1] * e1[2] - n[2] * e1[1], n[2] * e1[0] - n[0] * e1[2], n[0] * e1[1] - n[1] * e1[0]};
    double fnv[3];
    face_normal(i, fnv);
    const double *pa = V + 3 * F[3 * i], *pb = V + 3 * F[3 * i + 1], *pc = V + 3 * F[3 * i + 2];
    double c[3] = {((pa[0] + pb[0]) + pc[0]) / 3.0, ((pa[1] + pb[1]) + pc[1]) / 3.0, ((pa[2] + pb[2]) + pc[2]) / 3.0};
    float o[3] = {(float)(c[0] - fnv[0] * 1e-3), (float)(c[1] - fnv[1] * 1e-3), (float)(c[2] - fnv[2] * 1e-3)};
    uint8_t votes = 0;
    for (int k = 0; k < 7; ++k) {
      double d[3];
      if (k == 0) {
        d[0] = -n[0]; d[1] = -n[1]; d[2] = -n[2];
      } else {
        double phi = 2.0 * PI * (double)(k - 1) / 6.0;
        double cp = std::cos(phi), sp = std::sin(phi);
        for (int j = 0; j < 3; ++j) d[j] = -(cos_a * n[j]) + sin_a * (cp * e1[j] + sp * e2[j]);
      }
      float dir[3] = {(float)d[0], (float)d[1], (float)d[2]};
      float t;
      uint32_t prim;
      oracle_first_hit(s, o, dir, 1, retest, 0, &t, &prim);
      if (prim == 0xFFFFFFFFu || !std::isfinite(t) || prim >= nf || t <= min_hit) continue;
      float ratio = t / thick32;
      if (ratio <= 0.5f || ratio >= 1.8f) continue;
      double h[3];
      face_normal(prim, h);
      double al = (h[0] * d[0] + h[1] * d[1]) + h[2] * d[2];
      if (al > 0.7) votes += 1;
    }
    good[q] = votes;
  }
}

}  // extern "C"
