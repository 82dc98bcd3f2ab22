// Oracle, matching half: the coarse score, the ICP rung and the R §6 verification.
// TEST INFRASTRUCTURE ONLY (see oracle.h). Each function cites the reference lines it restates;
// sums run in the reference's order (source order, left to right) and in f64.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <limits>
#include <memory>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "internal.hpp"
#include "oracle.h"

namespace {

const double INF = std::numeric_limits<double>::infinity();

// Params::default (crates/sherd-core/src/params.rs:143-188): the R §1.1 constants R §6 reads.
const double P_TIGHT_DELTA = 0.01, P_TIGHT_RES = 0.15;
const double P_COARSE_DELTA = 0.15, P_COARSE_RES = 2.3;
const double P_STAGE1_DELTA = 0.06, P_STAGE1_RES = 0.9;
const double P_FACING_DELTA = 0.3, P_FACING_RES = 1.0;
const double P_MAX_GAP = 0.03, P_GAP_RES = 0.45;
const double P_SEAM_DELTA = 0.12, P_SEAM_RES = 1.8;
const double P_NEAR_DELTA = 0.5, P_NEAR_RES = 4.0;
const double P_PEN_DELTA = 0.06, P_PEN_RES = 0.9;
const double P_NMS_DELTA = 0.5;
const double P_ICP_DELTA = 0.04, P_ICP_RES = 0.6;
const double P_MIN_TIGHT = 0.25, P_MAX_PEN = 0.005, P_MIN_SEAM = 3.0, P_MIN_CONT_N = 0.8;
const double NORMAL_AGREE = 0.7;  // matching/coarse.rs:43
const size_t MIN_FACING = 20, MIN_NEAR = 20;  // verify.rs:61-63
const double SEAM_VOXEL = 3.0;                // verify.rs:65

enum { S_T, S_RES, S_COARSE, S_STAGE1, S_TIGHT, S_FACING, S_GAP, S_SEAM, S_NEAR, S_PEN, S_NMS, S_ICP };

// Rust's f64::max / f64::min: the non-NaN operand wins.
inline double rmax(double a, double b) { return std::fmax(a, b); }
inline double rmin(double a, double b) { return std::fmin(a, b); }

// Rust's f64::midpoint.
double midpoint(double a, double b) {
  const double LO = std::numeric_limits<double>::min() * 2.0;
  const double HI = std::numeric_limits<double>::max() / 2.0;
  double aa = std::fabs(a), ab = std::fabs(b);
  if (aa <= HI && ab <= HI) return (a + b) / 2.0;
  if (aa < LO) return a + (b / 2.0);
  if (ab < LO) return (a / 2.0) + b;
  return (a / 2.0) + (b / 2.0);
}

// f64::total_cmp as an unsigned key.
inline uint64_t total_key(double v) {
  uint64_t u;
  std::memcpy(&u, &v, 8);
  return (u & 0x8000000000000000ull) ? ~u : (u | 0x8000000000000000ull);
}

inline void apply3(const double T[16], const double *p, double out[3]) {
  for (int i = 0; i < 3; ++i) out[i] = T[4 * i] * p[0] + T[4 * i + 1] * p[1] + T[4 * i + 2] * p[2] + T[4 * i + 3];
}
inline void rotate3(const double T[16], const double *p, double out[3]) {
  for (int i = 0; i < 3; ++i) out[i] = T[4 * i] * p[0] + T[4 * i + 1] * p[1] + T[4 * i + 2] * p[2];
}
inline double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

// nalgebra's n x n product: each entry summed over k left to right.
void matmul(const double *a, const double *b, double *c, int n) {
  std::vector<double> out(n * n);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      double acc = a[i * n] * b[j];
      for (int k = 1; k < n; ++k) acc = acc + a[i * n + k] * b[k * n + j];
      out[i * n + j] = acc;
    }
  std::memcpy(c, out.data(), sizeof(double) * n * n);
}

double det3(const double m[9]) {
  return m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) +
         m[2] * (m[3] * m[7] - m[4] * m[6]);
}

// One-sided Jacobi SVD of a 3x3 (nalgebra's `svd` is not vendored): M = U diag(s) V^T with the
// singular values descending and U completed by a cross product where a value vanishes.
void svd3(const double m[9], double u[9], double s[3], double v[9]) {
  double a[9];
  std::memcpy(a, m, sizeof a);
  double vv[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  for (int sweep = 0; sweep < 60; ++sweep) {
    double off = 0.0;
    for (int p = 0; p < 2; ++p)
      for (int q = p + 1; q < 3; ++q) {
        double al = 0, be = 0, ga = 0;
        for (int i = 0; i < 3; ++i) {
          al += a[3 * i + p] * a[3 * i + p];
          be += a[3 * i + q] * a[3 * i + q];
          ga += a[3 * i + p] * a[3 * i + q];
        }
        if (ga == 0.0 || std::fabs(ga) <= 1e-300) continue;
        off = std::max(off, std::fabs(ga) / std::sqrt(al * be > 0 ? al * be : 1e-300));
        double zeta = (be - al) / (2.0 * ga);
        double t = (zeta >= 0 ? 1.0 : -1.0) / (std::fabs(zeta) + std::sqrt(1.0 + zeta * zeta));
        double c = 1.0 / std::sqrt(1.0 + t * t), sn = c * t;
        for (int i = 0; i < 3; ++i) {
          double ap = a[3 * i + p], aq = a[3 * i + q];
          a[3 * i + p] = c * ap - sn * aq;
          a[3 * i + q] = sn * ap + c * aq;
          double vp = vv[3 * i + p], vq = vv[3 * i + q];
          vv[3 * i + p] = c * vp - sn * vq;
          vv[3 * i + q] = sn * vp + c * vq;
        }
      }
    if (off < 1e-15) break;
  }
  double sig[3];
  for (int j = 0; j < 3; ++j) sig[j] = std::sqrt(a[j] * a[j] + a[3 + j] * a[3 + j] + a[6 + j] * a[6 + j]);
  int ord[3] = {0, 1, 2};
  std::sort(ord, ord + 3, [&](int x, int y) { return sig[x] > sig[y]; });
  double smax = sig[ord[0]];
  for (int c = 0; c < 3; ++c) {
    int j = ord[c];
    s[c] = sig[j];
    for (int i = 0; i < 3; ++i) v[3 * i + c] = vv[3 * i + j];
    for (int i = 0; i < 3; ++i) u[3 * i + c] = sig[j] > 1e-14 * smax && sig[j] > 0 ? a[3 * i + j] / sig[j] : 0.0;
  }
  auto col_norm = [&](int c) { return std::sqrt(u[c] * u[c] + u[3 + c] * u[3 + c] + u[6 + c] * u[6 + c]); };
  if (col_norm(0) == 0.0) { u[0] = 1; u[3] = 0; u[6] = 0; }
  if (col_norm(1) == 0.0) {
    // Any unit vector orthogonal to the first column.
    double x = u[0], y = u[3], z = u[6];
    double w[3] = {std::fabs(x) < 0.9 ? 1.0 : 0.0, std::fabs(x) < 0.9 ? 0.0 : 1.0, 0.0};
    double c[3] = {y * w[2] - z * w[1], z * w[0] - x * w[2], x * w[1] - y * w[0]};
    double l = std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
    u[1] = c[0] / l; u[4] = c[1] / l; u[7] = c[2] / l;
  }
  if (col_norm(2) == 0.0) {
    u[2] = u[3] * u[7] - u[6] * u[4];
    u[5] = u[6] * u[1] - u[0] * u[7];
    u[8] = u[0] * u[4] - u[3] * u[1];
  }
}

// --- ICP (crates/sherd-core/src/matching/icp.rs) ---------------------------------------------

struct Corr {
  std::vector<std::pair<uint32_t, uint32_t>> pairs;
  double fitness = 0, rmse = 0;
};

// icp.rs:361-392 correspondences: the nearest target strictly inside `distance`, source order.
void correspondences(const std::vector<double> &moved, const oracle::Cloud &tree, double distance,
                     Corr &c) {
  c.pairs.clear();
  c.fitness = c.rmse = 0.0;
  if (tree.size() == 0 || !(distance > 0.0)) return;
  double r2 = distance * distance;
  double err2 = 0.0;
  size_t n = moved.size() / 3;
  for (size_t i = 0; i < n; ++i) {
    int64_t j;
    double d2;
    if (!tree.nearest(&moved[3 * i], distance, &j, &d2)) continue;
    if (d2 >= r2) continue;
    err2 += d2;
    c.pairs.push_back({(uint32_t)i, (uint32_t)j});
  }
  if (c.pairs.empty()) return;
  double nc = (double)c.pairs.size(), ns = (double)n;
  c.fitness = nc / ns;
  c.rmse = std::sqrt(err2 / nc);
}

void identity4(double m[16]) {
  for (int i = 0; i < 16; ++i) m[i] = (i % 5 == 0) ? 1.0 : 0.0;
}

void homogeneous(const double r[9], const double t[3], double m[16]) {
  identity4(m);
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 3; ++j) m[4 * i + j] = r[3 * i + j];
    m[4 * i + 3] = t[i];
  }
}

// icp.rs:539-592 umeyama<f64>.
void umeyama(const std::vector<double> &moved, const double *tp, const Corr &c, double out[16]) {
  if (c.pairs.empty()) {
    identity4(out);
    return;
  }
  double one_over_n = 1.0 / (double)c.pairs.size();
  double sp[3] = {0, 0, 0}, sq[3] = {0, 0, 0};
  for (auto &ij : c.pairs)
    for (int k = 0; k < 3; ++k) {
      sp[k] = sp[k] + moved[3 * ij.first + k];
      sq[k] = sq[k] + tp[3 * ij.second + k];
    }
  double mp[3], mq[3];
  for (int k = 0; k < 3; ++k) {
    mp[k] = sp[k] * one_over_n;
    mq[k] = sq[k] * one_over_n;
  }
  double sig[9] = {0};
  for (auto &ij : c.pairs) {
    double a[3], b[3];
    for (int k = 0; k < 3; ++k) {
      a[k] = moved[3 * ij.first + k] - mp[k];
      b[k] = (tp[3 * ij.second + k] - mq[k]) * one_over_n;
    }
    for (int r = 0; r < 3; ++r)
      for (int cc = 0; cc < 3; ++cc) sig[3 * r + cc] = sig[3 * r + cc] + b[r] * a[cc];
  }
  double u[9], s[3], v[9];
  svd3(sig, u, s, v);
  double vt[9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) vt[3 * i + j] = v[3 * j + i];
  double d[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  if (det3(u) * det3(vt) < 0.0) d[8] = -1.0;
  double ud[9], rot[9];
  matmul(u, d, ud, 3);
  matmul(ud, vt, rot, 3);
  double t[3];
  for (int i = 0; i < 3; ++i) t[i] = mq[i] - (rot[3 * i] * mp[0] + rot[3 * i + 1] * mp[1] + rot[3 * i + 2] * mp[2]);
  homogeneous(rot, t, out);
}

}  // namespace

namespace oracle {

// icp.rs:661-674 eigen_pivots.
void eigen_pivots(const double a[36], int perm[6]) {
  for (int i = 0; i < 6; ++i) perm[i] = i;
  for (int k = 0; k < 6; ++k) {
    int best = k;
    for (int i = k + 1; i < 6; ++i)
      if (std::fabs(a[6 * perm[i] + perm[i]]) > std::fabs(a[6 * perm[best] + perm[best]])) best = i;
    std::swap(perm[k], perm[best]);
  }
}

// icp.rs:692-753 solve_ldlt.
bool solve_ldlt(const double a[36], const double b[6], double x[6]) {
  int perm[6];
  eigen_pivots(a, perm);
  double m[6][6];
  for (int k = 0; k < 6; ++k)
    for (int l = 0; l < 6; ++l) m[k][l] = a[6 * perm[k] + perm[l]];
  for (int k = 0; k < 6; ++k) {
    if (k > 0) {
      double temp[6] = {0};
      for (int j = 0; j < k; ++j) temp[j] = m[j][j] * m[k][j];
      double diagonal = m[k][k];
      for (int j = 0; j < k; ++j) diagonal -= m[k][j] * temp[j];
      m[k][k] = diagonal;
      for (int r = k + 1; r < 6; ++r) {
        double below = m[r][k];
        for (int j = 0; j < k; ++j) below -= m[r][j] * temp[j];
        m[r][k] = below;
      }
    }
    double pivot = m[k][k];
    if (pivot != 0.0)
      for (int r = k + 1; r < 6; ++r) m[r][k] /= pivot;
  }
  double y[6];
  for (int k = 0; k < 6; ++k) y[k] = b[perm[k]];
  for (int i = 0; i < 6; ++i) {
    double value = y[i];
    for (int j = 0; j < i; ++j) value -= m[i][j] * y[j];
    y[i] = value;
  }
  for (int i = 0; i < 6; ++i) {
    double pivot = m[i][i];
    y[i] = std::fabs(pivot) > std::numeric_limits<double>::min() ? y[i] / pivot : 0.0;
  }
  for (int i = 5; i >= 0; --i) {
    double value = y[i];
    for (int j = i + 1; j < 6; ++j) value -= m[j][i] * y[j];
    y[i] = value;
  }
  for (int k = 0; k < 6; ++k) x[perm[k]] = y[k];
  for (int k = 0; k < 6; ++k)
    if (!std::isfinite(x[k])) return false;
  return true;
}

}  // namespace oracle

namespace {

using oracle::eigen_pivots;
using oracle::solve_ldlt;

// icp.rs:770-778 euler_zyx: Rz(gamma) Ry(beta) Rx(alpha).
void euler_zyx(double al, double be, double ga, double r[9]) {
  double sa = std::sin(al), ca = std::cos(al), sb = std::sin(be), cb = std::cos(be), sg = std::sin(ga),
         cg = std::cos(ga);
  double rx[9] = {1, 0, 0, 0, ca, -sa, 0, sa, ca};
  double ry[9] = {cb, 0, sb, 0, 1, 0, -sb, 0, cb};
  double rz[9] = {cg, -sg, 0, sg, cg, 0, 0, 0, 1};
  double t[9];
  matmul(rz, ry, t, 3);
  matmul(t, rx, r, 3);
}

// icp.rs:599-640 normal_equations<f64> about the origin (Assembly::World), then icp.rs:454-483.
void point_to_plane(const std::vector<double> &moved, const double *tp, const double *tn, const Corr &c,
                    double out[16]) {
  identity4(out);
  if (c.pairs.empty() || tn == nullptr) return;
  double jtj[6][6] = {{0}}, jtr[6] = {0};
  for (auto &ij : c.pairs) {
    const double *p = &moved[3 * ij.first], *q = tp + 3 * ij.second, *n = tn + 3 * ij.second;
    double pc[3] = {p[0] - 0.0, p[1] - 0.0, p[2] - 0.0};
    double row[6] = {pc[1] * n[2] - pc[2] * n[1], pc[2] * n[0] - pc[0] * n[2], pc[0] * n[1] - pc[1] * n[0],
                     n[0], n[1], n[2]};
    double r = (p[0] - q[0]) * n[0] + (p[1] - q[1]) * n[1] + (p[2] - q[2]) * n[2];
    for (int k = 0; k < 6; ++k) {
      for (int l = 0; l <= k; ++l) jtj[k][l] = jtj[k][l] + row[k] * row[l];
      jtr[k] = jtr[k] + row[k] * r;
    }
  }
  double a[36], rhs[6], x[6];
  for (int k = 0; k < 6; ++k)
    for (int l = 0; l <= k; ++l) {
      a[6 * k + l] = jtj[k][l];
      a[6 * l + k] = jtj[k][l];
    }
  for (int k = 0; k < 6; ++k) rhs[k] = -jtr[k];
  if (!solve_ldlt(a, rhs, x)) return;
  double rot[9];
  euler_zyx(x[0], x[1], x[2], rot);
  homogeneous(rot, x + 3, out);
}

// icp.rs:308-356 register.
void register_one(const double *src, size_t ns, const double *tp, const double *tn, size_t nt,
                  const oracle::Cloud &tree, const double init[16], int estimation, double distance,
                  int max_iter, double out[21]) {
  double T[16];
  std::memcpy(T, init, sizeof T);
  if (ns == 0 || nt == 0) {
    std::memcpy(out, T, sizeof T);
    out[16] = out[17] = out[18] = out[19] = out[20] = 0.0;
    return;
  }
  std::vector<double> moved(src, src + 3 * ns), tmp(3);
  auto transform_points = [&](const double M[16]) {
    for (size_t i = 0; i < ns; ++i) {
      double o[3];
      apply3(M, &moved[3 * i], o);
      moved[3 * i] = o[0];
      moved[3 * i + 1] = o[1];
      moved[3 * i + 2] = o[2];
    }
  };
  transform_points(T);
  Corr c;
  correspondences(moved, tree, distance, c);
  double fitness = c.fitness, rmse = c.rmse;
  int iters = 0;
  bool converged = false;
  for (int it = 0; it < max_iter; ++it) {
    double U[16];
    if (estimation == 0) umeyama(moved, tp, c, U);
    else point_to_plane(moved, tp, tn, c, U);
    matmul(U, T, T, 4);
    transform_points(U);
    iters += 1;
    correspondences(moved, tree, distance, c);
    converged = std::fabs(c.fitness - fitness) < 1e-6 && std::fabs(c.rmse - rmse) < 1e-6;
    fitness = c.fitness;
    rmse = c.rmse;
    if (converged) break;
  }
  std::memcpy(out, T, sizeof T);
  out[16] = fitness;
  out[17] = rmse;
  out[18] = (double)c.pairs.size();
  out[19] = (double)iters;
  out[20] = converged ? 1.0 : 0.0;
}

// --- verification (crates/sherd-core/src/matching/verify.rs) -----------------------------------

double median_inplace(std::vector<double> &v) {
  size_t n = v.size();
  if (n == 0) return std::numeric_limits<double>::quiet_NaN();
  std::sort(v.begin(), v.end(), [](double a, double b) { return total_key(a) < total_key(b); });
  return (n % 2 == 1) ? v[n / 2] : midpoint(v[n / 2 - 1], v[n / 2]);
}

// verify.rs:338-349 one_side.
void one_side(const std::vector<double> &d, double area, const double *sc, double out[3]) {
  std::vector<double> facing;
  for (double x : d)
    if (x < sc[S_FACING]) facing.push_back(x);
  if (facing.size() < MIN_FACING) {
    out[0] = 0.0;
    out[1] = 1.0;
    out[2] = 0.0;
    return;
  }
  size_t tight = 0;
  for (double x : facing) tight += (x < sc[S_TIGHT]);
  out[0] = (double)tight / (double)facing.size();
  out[1] = median_inplace(facing) / sc[S_T];
  size_t touching = 0;
  for (double x : d) touching += (x < 2.0 * sc[S_TIGHT]);
  double share = (double)touching / (double)d.size();
  out[2] = share * area / (sc[S_T] * sc[S_T]);
}

std::vector<double> surface_distances(const double *pts, size_t n, const double T[16], const oracle_scene *scene,
                                      double max_dist) {
  std::vector<double> d(n);
  if (n) oracle_exec_bounded_distance(scene, pts, n, T, max_dist, 0, d.data());
  return d;
}

// verify.rs:541-580 one_penetration.
void one_penetration(const double *pts, size_t n, const double T[16], const oracle_scene *scene, double pen,
                     int retest, double *fraction, double *min_sd) {
  if (n == 0) {
    *fraction = 0.0;
    *min_sd = 0.0;
    return;
  }
  std::vector<float> depth(n);
  std::vector<uint32_t> in(n);
  oracle_exec_inside(scene, pts, n, T, retest, depth.data(), in.data());
  double deepest = -INF;
  size_t deeper = 0;
  for (size_t i = 0; i < n; ++i) {
    if (!in[i]) continue;
    double dist = (double)depth[i];
    deepest = rmax(deepest, dist);
    if (dist > pen) deeper += 1;
  }
  *fraction = (double)deeper / (double)n;
  if (deepest > -INF) {
    *min_sd = -deepest;
    return;
  }
  double best;
  oracle_exec_bounded_distance(scene, pts, n, T, (double)std::numeric_limits<float>::max(), 1, &best);
  *min_sd = best;
}

}  // namespace

extern "C" {

int oracle_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
  else omp_set_num_threads(omp_get_num_procs());
  return omp_get_max_threads();
#else
  (void)n;
  return 1;
#endif
}

void oracle_coarse_scores(const double *target_p, const double *target_n, size_t n_target,
                          const double *probe_p, const double *probe_n, size_t n_probe,
                          const double *poses, size_t k, double radius, double normal_agree,
                          uint32_t *counts, double *scores) {
  if (n_probe == 0 || n_target == 0) {
    for (size_t h = 0; h < k; ++h) {
      scores[h] = 0.0;
      if (counts) counts[h] = 0;
    }
    return;
  }
  oracle::Cloud tree(target_p, n_target, radius > 0 ? radius : 1.0);
  double points = (double)n_probe;
#pragma omp parallel for schedule(dynamic, 64)
  for (long h = 0; h < (long)k; ++h) {
    const double *r = poses + 12 * h, *tau = r + 9;
    uint32_t agree = 0;
    for (size_t i = 0; i < n_probe; ++i) {
      const double *p = probe_p + 3 * i, *n = probe_n + 3 * i;
      double moved[3];
      for (int a = 0; a < 3; ++a) moved[a] = r[3 * a] * p[0] + r[3 * a + 1] * p[1] + r[3 * a + 2] * p[2] + tau[a];
      int64_t near;
      double dist;
      if (!oracle::nearest_below(tree, moved, radius, &near, &dist)) continue;
      double turned[3];
      for (int a = 0; a < 3; ++a) turned[a] = r[3 * a] * n[0] + r[3 * a + 1] * n[1] + r[3 * a + 2] * n[2];
      const double *theirs = target_n + 3 * near;
      double d = theirs[0] * turned[0] + theirs[1] * turned[1] + theirs[2] * turned[2];
      if (d > normal_agree) agree += 1;
    }
    if (counts) counts[h] = agree;
    scores[h] = (double)agree / points;
  }
}

void oracle_icp_rung(const double *src, size_t n_src, const double *tgt_p, const double *tgt_n,
                     size_t n_tgt, const double *inits, size_t k, int estimation, double distance,
                     int max_iteration, double *out) {
  oracle::Cloud tree(tgt_p, n_tgt, distance > 0 ? distance : 1.0);
#pragma omp parallel for schedule(dynamic, 1)
  for (long c = 0; c < (long)k; ++c)
    register_one(src, n_src, tgt_p, tgt_n, n_tgt, tree, inits + 16 * c, estimation, distance, max_iteration,
                 out + 21 * c);
}

void oracle_icp_correspondences(const double *moved, size_t n, const double *tgt_p, size_t n_tgt, double distance,
                                 int64_t *corr) {
  oracle::Cloud tree(tgt_p, n_tgt, distance > 0 ? distance : 1.0);
  std::vector<double> m(moved, moved + 3 * n);
  Corr c;
  correspondences(m, tree, distance, c);
  for (size_t i = 0; i < n; ++i) corr[i] = -1;
  for (auto &ij : c.pairs) corr[ij.first] = (int64_t)ij.second;
}

void oracle_normal_equations(const double *moved, size_t n, const double *tgt_p, const double *tgt_n,
                             const int64_t *corr, double jtj[36], double jtr[6]) {
  // icp.rs:599-640 over the pairs in source order, rows [p x n, n], residual (p - q) . n.
  double a[6][6] = {{0}}, b[6] = {0};
  for (size_t i = 0; i < n; ++i) {
    if (corr[i] < 0) continue;
    const double *p = moved + 3 * i, *q = tgt_p + 3 * corr[i], *nn = tgt_n + 3 * corr[i];
    double row[6] = {p[1] * nn[2] - p[2] * nn[1], p[2] * nn[0] - p[0] * nn[2], p[0] * nn[1] - p[1] * nn[0],
                     nn[0], nn[1], nn[2]};
    double r = (p[0] - q[0]) * nn[0] + (p[1] - q[1]) * nn[1] + (p[2] - q[2]) * nn[2];
    for (int k = 0; k < 6; ++k) {
      for (int l = 0; l <= k; ++l) a[k][l] = a[k][l] + row[k] * row[l];
      b[k] = b[k] + row[k] * r;
    }
  }
  for (int k = 0; k < 6; ++k) {
    for (int l = 0; l <= k; ++l) jtj[6 * k + l] = jtj[6 * l + k] = a[k][l];
    jtr[k] = b[k];
  }
}

void oracle_scales_for_pair(double t, double res, double out[12]) {
  auto f = [&](double k, double m) { return rmax(k * t, m * res); };
  out[S_T] = t;
  out[S_RES] = res;
  out[S_COARSE] = f(P_COARSE_DELTA, P_COARSE_RES);
  out[S_STAGE1] = f(P_STAGE1_DELTA, P_STAGE1_RES);
  out[S_TIGHT] = f(P_TIGHT_DELTA, P_TIGHT_RES);
  out[S_FACING] = f(P_FACING_DELTA, P_FACING_RES);
  out[S_GAP] = f(P_MAX_GAP, P_GAP_RES);
  out[S_SEAM] = f(P_SEAM_DELTA, P_SEAM_RES);
  out[S_NEAR] = f(P_NEAR_DELTA, P_NEAR_RES);
  out[S_PEN] = f(P_PEN_DELTA, P_PEN_RES);
  out[S_NMS] = P_NMS_DELTA * t;
  out[S_ICP] = f(P_ICP_DELTA, P_ICP_RES) / (P_ICP_DELTA * t);
}

void oracle_pose_inverse(const double T[16], double out[16]) {
  double m[4][4];
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) m[i][j] = T[4 * i + j];
  int piv[4] = {0, 1, 2, 3};
  for (int k = 0; k < 4; ++k) {
    int best = k;
    for (int i = k + 1; i < 4; ++i)
      if (std::fabs(m[i][k]) > std::fabs(m[best][k])) best = i;
    if (best != k) {
      for (int j = 0; j < 4; ++j) std::swap(m[k][j], m[best][j]);
      std::swap(piv[k], piv[best]);
    }
    double pivot = m[k][k];
    if (pivot == 0.0) continue;
    for (int r = k + 1; r < 4; ++r) {
      m[r][k] /= pivot;
      double factor = m[r][k];
      for (int c = k + 1; c < 4; ++c) m[r][c] -= factor * m[k][c];
    }
  }
  for (int col = 0; col < 4; ++col) {
    double y[4], x[4];
    for (int i = 0; i < 4; ++i) {
      double acc = 0.0;
      for (int j = 0; j < i; ++j) acc += m[i][j] * y[j];
      y[i] = (piv[i] == col ? 1.0 : 0.0) - acc;
    }
    for (int i = 3; i >= 0; --i) {
      double acc = 0.0;
      for (int j = i + 1; j < 4; ++j) acc += m[i][j] * x[j];
      x[i] = (y[i] - acc) / m[i][i];
    }
    for (int r = 0; r < 4; ++r) out[4 * r + col] = x[r];
  }
}

double oracle_median(double *values, size_t n) {
  std::vector<double> v(values, values + n);
  double m = median_inplace(v);
  std::copy(v.begin(), v.end(), values);
  return m;
}

void oracle_verify(const oracle_side *a, const oracle_side *b, const double T[16],
                   const double sc[12], int full, int retest, double out[OS_COUNT]) {
  double inv[16];
  oracle_pose_inverse(T, inv);
  // verify.rs:318-335 fracture_scores.
  std::vector<double> d1 = surface_distances(b->pf, b->n_pf, T, a->fracture, sc[S_FACING]);
  std::vector<double> d2 = surface_distances(a->pf, a->n_pf, inv, b->fracture, sc[S_FACING]);
  double sa[3], sb[3];
  one_side(d2, a->frac_area, sc, sa);
  one_side(d1, b->frac_area, sc, sb);
  for (int i = 0; i < OS_COUNT; ++i) out[i] = 0.0;
  out[OS_TIGHT_A] = sa[0];
  out[OS_TIGHT_B] = sb[0];
  out[OS_TIGHT] = rmin(sa[0], sb[0]);
  out[OS_GAP_A] = sa[1];
  out[OS_GAP_B] = sb[1];
  out[OS_GAP] = rmax(sa[1], sb[1]);
  out[OS_CONTACT_A] = sa[2];
  out[OS_CONTACT_B] = sb[2];
  out[OS_CONTACT] = rmin(sa[2], sb[2]);
  // verify.rs:378-442 seam_score: A's breakline against the moved B breakline.
  {
    std::vector<double> moved(3 * b->n_brk), normals(3 * b->n_brk);
    for (size_t i = 0; i < b->n_brk; ++i) {
      apply3(T, b->brk_p + 3 * i, &moved[3 * i]);
      rotate3(T, b->brk_ns + 3 * i, &normals[3 * i]);
    }
    std::vector<std::array<int64_t, 3>> cells;
    if (b->n_brk > 0) {
      oracle::Cloud tree(moved.data(), b->n_brk, sc[S_SEAM] > 0 ? sc[S_SEAM] : 1.0);
      double voxel = sc[S_T] / SEAM_VOXEL;
      for (size_t i = 0; i < a->n_brk; ++i) {
        int64_t j;
        double dist;
        if (!oracle::nearest_below(tree, a->brk_p + 3 * i, sc[S_SEAM], &j, &dist)) continue;
        if (dot3(a->brk_ns + 3 * i, &normals[3 * j]) > NORMAL_AGREE) {
          const double *p = a->brk_p + 3 * i;
          cells.push_back({(int64_t)std::floor(p[0] / voxel), (int64_t)std::floor(p[1] / voxel),
                           (int64_t)std::floor(p[2] / voxel)});
        }
      }
    }
    std::sort(cells.begin(), cells.end());
    cells.erase(std::unique(cells.begin(), cells.end()), cells.end());
    out[OS_SEAM] = cells.empty() ? 0.0 : (double)cells.size() / SEAM_VOXEL;
  }
  out[OS_GAP_LIMIT] = sc[S_GAP] / sc[S_T];
  out[OS_TIGHT_DELTA] = sc[S_TIGHT] / sc[S_T];
  out[OS_CONT] = 1.0;
  out[OS_CONT_N] = -1.0;
  if (!full) {
    out[OS_PARTIAL] = 1.0;
    return;
  }
  // verify.rs:450-474 continuity_scores.
  if (a->n_margin > 0 && b->n_margin > 0) {
    oracle::Cloud tree(a->margin_p, a->n_margin, sc[S_NEAR] > 0 ? sc[S_NEAR] : 1.0);
    std::vector<double> steps, agreements;
    for (size_t i = 0; i < b->n_margin; ++i) {
      double moved[3];
      apply3(T, b->margin_p + 3 * i, moved);
      int64_t j;
      double dist;
      if (!oracle::nearest_below(tree, moved, sc[S_NEAR], &j, &dist)) continue;
      const double *np = a->margin_p + 3 * j, *nn = a->margin_n + 3 * j;
      double delta[3] = {moved[0] - np[0], moved[1] - np[1], moved[2] - np[2]};
      steps.push_back(std::fabs(dot3(delta, nn)));
      double turned[3];
      rotate3(T, b->margin_n + 3 * i, turned);
      agreements.push_back(dot3(turned, nn));
    }
    if (steps.size() > MIN_NEAR) {
      out[OS_CONT] = median_inplace(steps) / sc[S_T];
      out[OS_CONT_N] = median_inplace(agreements);
    }
  }
  // verify.rs:483-498 penetration_scores.
  if (!(a->mesh && b->mesh) || !(a->watertight && b->watertight)) {
    out[OS_PEN] = 0.0;
    out[OS_PEN_DEPTH] = 0.0;
    out[OS_PEN_UNAVAILABLE] = 1.0;
    return;
  }
  double pen_b, min_b, pen_a, min_a;
  one_penetration(b->s, b->n_s, T, a->mesh, sc[S_PEN], retest, &pen_b, &min_b);
  one_penetration(a->s, a->n_s, inv, b->mesh, sc[S_PEN], retest, &pen_a, &min_a);
  out[OS_PEN] = rmax(pen_a, pen_b);
  out[OS_PEN_DEPTH] = rmax(-min_a, -min_b) / sc[S_T];
}

int oracle_accept(const double s[OS_COUNT], const double sc[12]) {
  return s[OS_TIGHT] >= P_MIN_TIGHT && s[OS_GAP] * sc[S_T] <= sc[S_GAP] && s[OS_PEN] <= P_MAX_PEN &&
         s[OS_SEAM] >= P_MIN_SEAM && s[OS_CONT_N] >= P_MIN_CONT_N;
}

}  // extern "C"
