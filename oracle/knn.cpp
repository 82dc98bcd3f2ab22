// Oracle, nearest neighbours. TEST INFRASTRUCTURE ONLY (see oracle.h).
//
// kiddo's `ImmutableKdTree<f64, 3>` is not vendored, so its answer is restated by its contract as
// crates/sherd-core/src/spatial/kdtree.rs documents and tests it: the squared distance
// `((dx*dx + dy*dy) + dz*dz)` in f64, the global nearest, ties to the lowest index
// (kdtree.rs:338-376), a strict `d < bound` taken after the square root (kdtree.rs:182-192).
// A uniform grid whose cell is at least the search reach finds that answer exactly: every point
// within the reach of a query lies in the 27 cells around it.
#include <algorithm>
#include <cmath>
#include <limits>
#include <unordered_map>
#include <vector>

#include "internal.hpp"
#include "oracle.h"

namespace oracle {

struct Cloud::Impl {
  std::vector<double> p;
  double origin[3];
  double cell;
  std::unordered_map<uint64_t, std::vector<uint32_t>> cells;
  static uint64_t key(int64_t x, int64_t y, int64_t z) {
    const uint64_t m = (1ull << 21) - 1;
    return ((uint64_t)(x + (1 << 20)) & m) | (((uint64_t)(y + (1 << 20)) & m) << 21) |
           (((uint64_t)(z + (1 << 20)) & m) << 42);
  }
  int64_t coord(double v, int k) const { return (int64_t)std::floor((v - origin[k]) / cell); }
};

Cloud::Cloud(const double *pts, size_t n, double cell) : impl_(new Impl), n_(n) {
  impl_->p.assign(pts, pts + 3 * n);
  double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
  for (int k = 0; k < 3; ++k) {
    lo[k] = std::numeric_limits<double>::infinity();
    hi[k] = -std::numeric_limits<double>::infinity();
  }
  for (size_t i = 0; i < n; ++i)
    for (int k = 0; k < 3; ++k) {
      lo[k] = std::min(lo[k], pts[3 * i + k]);
      hi[k] = std::max(hi[k], pts[3 * i + k]);
    }
  double ext = 0.0;
  for (int k = 0; k < 3; ++k) {
    impl_->origin[k] = n ? lo[k] : 0.0;
    if (n) ext = std::max(ext, hi[k] - lo[k]);
  }
  // At most 2^20 cells per axis, so the packed key never wraps; the slack keeps a point and a
  // query one reach apart at most one cell apart after rounding.
  impl_->cell = std::max(cell, ext / (double)(1 << 19)) * (1.0 + 1e-9);
  if (!(impl_->cell > 0.0) || !std::isfinite(impl_->cell)) impl_->cell = 1.0;
  for (size_t i = 0; i < n; ++i) {
    const double *q = pts + 3 * i;
    impl_->cells[Impl::key(impl_->coord(q[0], 0), impl_->coord(q[1], 1), impl_->coord(q[2], 2))]
        .push_back((uint32_t)i);
  }
}

Cloud::~Cloud() { delete impl_; }

bool Cloud::nearest(const double q[3], double reach, int64_t *idx, double *d2) const {
  (void)reach;
  if (n_ == 0) return false;
  int64_t c[3];
  for (int k = 0; k < 3; ++k) {
    double v = (q[k] - impl_->origin[k]) / impl_->cell;
    if (!(v > -1e15 && v < 1e15)) return false;  // far outside: nothing within one cell
    c[k] = (int64_t)std::floor(v);
  }
  double best = std::numeric_limits<double>::infinity();
  int64_t bi = -1;
  for (int dx = -1; dx <= 1; ++dx)
    for (int dy = -1; dy <= 1; ++dy)
      for (int dz = -1; dz <= 1; ++dz) {
        int64_t x = c[0] + dx, y = c[1] + dy, z = c[2] + dz;
        if (x < -(1 << 20) || y < -(1 << 20) || z < -(1 << 20) || x >= (1 << 20) || y >= (1 << 20) ||
            z >= (1 << 20))
          continue;
        auto it = impl_->cells.find(Impl::key(x, y, z));
        if (it == impl_->cells.end()) continue;
        for (uint32_t i : it->second) {
          const double *p = &impl_->p[3 * i];
          double ex = q[0] - p[0], ey = q[1] - p[1], ez = q[2] - p[2];
          double d = (ex * ex + ey * ey) + ez * ez;
          if (d < best || (d == best && (int64_t)i < bi)) {
            best = d;
            bi = i;
          }
        }
      }
  if (bi < 0) return false;
  *idx = bi;
  *d2 = best;
  return true;
}

bool nearest_below(const Cloud &c, const double q[3], double bound, int64_t *idx, double *dist) {
  if (!(bound > 0.0)) return false;
  int64_t i;
  double d2;
  if (!c.nearest(q, bound, &i, &d2)) return false;
  double d = std::sqrt(d2);
  if (!(d < bound)) return false;
  *idx = i;
  *dist = d;
  return true;
}

}  // namespace oracle

extern "C" {

void oracle_nearest_below(const double *cloud, size_t n, const double *q, size_t m, double bound,
                          int64_t *idx, double *dist) {
  oracle::Cloud c(cloud, n, bound > 0.0 ? bound : 1.0);
#pragma omp parallel for schedule(dynamic, 256)
  for (long i = 0; i < (long)m; ++i) {
    int64_t j;
    double d;
    if (oracle::nearest_below(c, q + 3 * i, bound, &j, &d)) {
      idx[i] = j;
      dist[i] = d;
    } else {
      idx[i] = -1;
      dist[i] = std::numeric_limits<double>::infinity();
    }
  }
}

void oracle_nearest_unbounded(const double *cloud, size_t n, const double *q, size_t m, int64_t *idx,
                              double *dist) {
#pragma omp parallel for schedule(dynamic, 64)
  for (long k = 0; k < (long)m; ++k) {
    const double *qq = q + 3 * k;
    double best = std::numeric_limits<double>::infinity();
    int64_t bi = -1;
    for (size_t i = 0; i < n; ++i) {
      const double *p = cloud + 3 * i;
      double ex = qq[0] - p[0], ey = qq[1] - p[1], ez = qq[2] - p[2];
      double d = (ex * ex + ey * ey) + ez * ez;
      if (d < best) {
        best = d;
        bi = (int64_t)i;
      }
    }
    idx[k] = bi;
    dist[k] = bi < 0 ? std::numeric_limits<double>::infinity() : std::sqrt(best);
  }
}

void oracle_place(const double *p, size_t n, const double T[16], int direction, double *out) {
  for (size_t k = 0; k < n; ++k) {
    const double *x = p + 3 * k;
    for (int r = 0; r < 3; ++r) {
      double v = std::fma(T[4 * r + 2], x[2], std::fma(T[4 * r + 1], x[1], T[4 * r] * x[0]));
      out[3 * k + r] = direction ? v : v + T[4 * r + 3];
    }
  }
}

// grid.rs's own brute force (spatial/grid.rs:299-315), which its HashGrid is tested against.
void oracle_nearest_f32(const float *cloud, size_t n, const float *q, size_t m, float radius,
                        int strict, int64_t *idx, float *dist) {
  float r2 = radius * radius;
#pragma omp parallel for schedule(dynamic, 64)
  for (long k = 0; k < (long)m; ++k) {
    uint32_t best = 0xFFFFFFFFu;
    float best_d2 = std::numeric_limits<float>::infinity();
    const float *qq = q + 3 * k;
    for (size_t i = 0; i < n; ++i) {
      const float *p = cloud + 3 * i;
      float dx = p[0] - qq[0], dy = p[1] - qq[1], dz = p[2] - qq[2];
      float d2 = dx * dx + dy * dy + dz * dz;
      bool in = strict ? (d2 < r2) : (d2 <= r2);
      if (in && (d2 < best_d2 || (d2 == best_d2 && (uint32_t)i < best))) {
        best = (uint32_t)i;
        best_d2 = d2;
      }
    }
    idx[k] = best == 0xFFFFFFFFu ? -1 : (int64_t)best;
    dist[k] = best == 0xFFFFFFFFu ? std::numeric_limits<float>::infinity() : std::sqrt(best_d2);
  }
}

}  // extern "C"
