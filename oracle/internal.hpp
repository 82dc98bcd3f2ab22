// Oracle internals shared by its translation units. TEST INFRASTRUCTURE ONLY.
#pragma once
#include <cstddef>
#include <cstdint>

struct oracle_scene;

namespace oracle {

// An f32 triple with glam's Vec3 arithmetic (parry3d's `Vector`): cross as
// (y*z' - y'*z, z*x' - z'*x, x*y' - x'*y) and dot as (x*x' + y*y') + z*z'.
struct V3 {
  float x, y, z;
  float operator[](int k) const { return k == 0 ? x : (k == 1 ? y : z); }
};

inline V3 sub(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 scale(V3 a, float s) { return V3{a.x * s, a.y * s, a.z * s}; }
inline float dot(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
inline V3 cross(V3 a, V3 b) {
  return V3{a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}

float tri_d2(V3 p, V3 a, V3 b, V3 c);
V3 moved_f32(const double T[16], const double *p);
bool bounded(const oracle_scene *s, V3 p, float max_dist, float *d);
bool inside(const oracle_scene *s, V3 p, int retest);

// f64 nearest neighbour with kiddo's semantics (kdtree.rs): squared distance ((dx^2 + dy^2) + dz^2),
// the global nearest with ties to the lowest index. Built once over a cloud.
class Cloud {
 public:
  Cloud(const double *pts, size_t n, double cell);
  ~Cloud();
  // The nearest point and its squared distance, among points within `reach` of q; false if none.
  bool nearest(const double q[3], double reach, int64_t *idx, double *d2) const;
  size_t size() const { return n_; }

 private:
  struct Impl;
  Impl *impl_;
  size_t n_;
};

// icp.rs:661-674 eigen_pivots and icp.rs:692-753 solve_ldlt: Eigen's pivoted LDLT of a symmetric
// 6 x 6 with the pseudo-inverse of D; false on a non-finite solution.
void eigen_pivots(const double a[36], int perm[6]);
bool solve_ldlt(const double a[36], const double b[6], double x[6]);

// kdtree.rs:182-192 nearest_below over a Cloud whose cell is at least `bound`.
bool nearest_below(const Cloud &c, const double q[3], double bound, int64_t *idx, double *dist);

}  // namespace oracle
