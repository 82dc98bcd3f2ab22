// FP64 restatement of the batched Levenberg-Marquardt refinement on the robust (Huber)
// point-to-plane seam energy. TEST INFRASTRUCTURE ONLY: the tests, the smoke check and the
// benchmark's CPU baseline use it as the checker; no product path links it.
//
// The reference has no such refinement. Its pieces follow the reference where one exists:
// correspondences are icp.rs:361-392 (the nearest target strictly inside the radius, ties to the
// lowest index, points moved row by row as ((a x + b y) + c z) + t); the rows [p x n, n] and
// residual (p - q) . n are normal_equations' (icp.rs:599-640); the 6 x 6 solve is Eigen's pivoted
// LDLT (icp.rs:661-753). The rest is the refinement's own definition, step for step:
//
//   per iteration, while the pose runs:
//     C   = correspondences at the current pose T
//     E   = sum rho(r_i), H = sum w_i row_i row_i^T, g = -sum w_i r_i row_i  over C, with w = psi(r)/r
//           (1 at r = 0), rows and residuals about the target centroid c_t (p' = T s - c_t)
//     H_kk += lambda (max(H_kk, 0) + 1e-9 trace(H) / 6 + 1e-300)        (Marquardt, floored)
//     x    = solve(H, g); stop when E = 0 or x is not finite
//     x   *= min(1, 0.05 / |x_omega|, r / |x_v|)                         (trust region)
//     T'   = [exp(x_omega^) | c_t - exp(x_omega^) c_t + x_v] T          (rotation about c_t)
//     dr_i = ((A' - A)(s_i - c_s) + (b' - b)) . n_i     (the step as a change of the shifted map)
//     E'   = sum rho(r_i + dr_i),  dE = sum [rho(r_i + dr_i) - rho(r_i)], each term formed from dr_i
//           (dr (r + dr / 2) inside the Huber kink, +-delta dr outside it, else the difference)
//     dE < 0: keep T', lambda = max(lambda / 10, 1e-12), stop when -dE <= 1e-12 E
//     else:   lambda *= 10, stop when lambda > 1e12
//
// The sums here run in source order; the device sums in its fixed tree, so the two agree to the
// rounding of the sums, not bit for bit.
#include <omp.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "internal.hpp"

namespace {

void rodrigues(const double w[3], double R[9]) {
  double th = std::sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
  double K[9] = {0.0, -w[2], w[1], w[2], 0.0, -w[0], -w[1], w[0], 0.0};
  double K2[9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) K2[3 * i + j] = K[3 * i] * K[j] + K[3 * i + 1] * K[3 + j] + K[3 * i + 2] * K[6 + j];
  double a = 1.0, b = 0.5;
  if (!(th < 1e-8)) {
    a = std::sin(th) / th;
    b = (1.0 - std::cos(th)) / (th * th);
  }
  for (int k = 0; k < 9; ++k) R[k] = (k % 4 == 0 ? 1.0 : 0.0) + a * K[k] + b * K2[k];
}

void huber(double r, double delta, double *rho, double *psi) {
  double a = std::fabs(r);
  if (a <= delta) {
    *rho = 0.5 * r * r;
    *psi = r;
  } else {
    *rho = delta * (a - 0.5 * delta);
    *psi = r < 0.0 ? -delta : delta;
  }
}

struct Pair {
  const double *src, *tp, *tn;
  size_t ns;
  double cs[3], ct[3];
  double radius;
  const oracle::Cloud *tree;
};

// p' = T s - c_t, written as the shifted map A (s - c_s) + b the device uses.
void shifted_map(const double *T, const Pair &P, double A[9], double b[3]) {
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 3; ++j) A[3 * i + j] = T[4 * i + j];
    b[i] = ((T[4 * i] * P.cs[0] + T[4 * i + 1] * P.cs[1]) + T[4 * i + 2] * P.cs[2]) + T[4 * i + 3] - P.ct[i];
  }
}

// Residual of source point i against target j at the shifted map, with p' and n.
double residual(const Pair &P, const double A[9], const double b[3], size_t i, int64_t j, double p[3],
                const double **n) {
  const double *s = P.src + 3 * i, *q = P.tp + 3 * j;
  *n = P.tn + 3 * j;
  double s0 = s[0] - P.cs[0], s1 = s[1] - P.cs[1], s2 = s[2] - P.cs[2];
  for (int k = 0; k < 3; ++k) p[k] = ((A[3 * k] * s0 + A[3 * k + 1] * s1) + A[3 * k + 2] * s2) + b[k];
  double qq[3] = {q[0] - P.ct[0], q[1] - P.ct[1], q[2] - P.ct[2]};
  const double *nn = *n;
  return ((p[0] - qq[0]) * nn[0] + (p[1] - qq[1]) * nn[1]) + (p[2] - qq[2]) * nn[2];
}

void correspond(const Pair &P, const double *T, std::vector<int64_t> &corr) {
  corr.assign(P.ns, -1);
  if (!(P.radius > 0.0) || P.tree->size() == 0) return;
  double r2 = P.radius * P.radius;
  for (size_t i = 0; i < P.ns; ++i) {
    const double *s = P.src + 3 * i;
    double q[3];
    for (int k = 0; k < 3; ++k) q[k] = ((T[4 * k] * s[0] + T[4 * k + 1] * s[1]) + T[4 * k + 2] * s[2]) + T[4 * k + 3];
    int64_t j;
    double d2;
    if (P.tree->nearest(q, P.radius, &j, &d2) && d2 < r2) corr[i] = j;
  }
}

// rho(r + dr) and its change from rho(r), the change formed from dr within one branch.
void huber_change(double r, double dr, double delta, double *rho1, double *d) {
  double r1 = r + dr, psi, rho0;
  huber(r1, delta, rho1, &psi);
  double a0 = std::fabs(r), a1 = std::fabs(r1);
  if (a0 <= delta && a1 <= delta) {
    *d = dr * (r + 0.5 * dr);
  } else if (a0 > delta && a1 > delta && ((r < 0.0) == (r1 < 0.0))) {
    *d = (r < 0.0 ? -delta : delta) * dr;
  } else {
    huber(r, delta, &rho0, &psi);
    *d = *rho1 - rho0;
  }
}

// The energy at Tn over corr, and its change from T summed point by point.
void trial(const Pair &P, const double *T, const double *Tn, const std::vector<int64_t> &corr, double delta,
           double *En, double *dE) {
  double A[9], b[3], An[9], bn[3], dA[9], db[3], p[3], rho1, d;
  const double *n;
  shifted_map(T, P, A, b);
  shifted_map(Tn, P, An, bn);
  for (int k = 0; k < 9; ++k) dA[k] = An[k] - A[k];
  for (int k = 0; k < 3; ++k) db[k] = bn[k] - b[k];
  double e = 0.0, de = 0.0;
  for (size_t i = 0; i < P.ns; ++i) {
    if (corr[i] < 0) continue;
    double r = residual(P, A, b, i, corr[i], p, &n);
    const double *s = P.src + 3 * i;
    double s0 = s[0] - P.cs[0], s1 = s[1] - P.cs[1], s2 = s[2] - P.cs[2], dp[3];
    for (int k = 0; k < 3; ++k) dp[k] = ((dA[3 * k] * s0 + dA[3 * k + 1] * s1) + dA[3 * k + 2] * s2) + db[k];
    double dr = (dp[0] * n[0] + dp[1] * n[1]) + dp[2] * n[2];
    huber_change(r, dr, delta, &rho1, &d);
    e += rho1;
    de += d;
  }
  *En = e;
  *dE = de;
}

enum { RAN_OUT = 0, CONVERGED = 1, DAMPED_OUT = 2, NO_STEP = 3 };

void lm_one(const Pair &P, const double *T0, double delta, int iterations, double *T_out, double *E_out,
            uint32_t *steps_out, int32_t *status) {
  double T[16];
  std::memcpy(T, T0, sizeof T);
  double lambda = 1e-3, Ef = 0.0;
  uint32_t steps = 0;
  int st = RAN_OUT;
  std::vector<int64_t> corr;
  for (int it = 0; it < iterations && st == RAN_OUT; ++it) {
    correspond(P, T, corr);
    double A[9], b[3], p[3], rho, psi;
    const double *n;
    shifted_map(T, P, A, b);
    double E = 0.0, H[36] = {0}, g[6] = {0};
    for (size_t i = 0; i < P.ns; ++i) {
      if (corr[i] < 0) continue;
      double r = residual(P, A, b, i, corr[i], p, &n);
      huber(r, delta, &rho, &psi);
      double w = r == 0.0 ? 1.0 : psi / r;
      double row[6] = {p[1] * n[2] - p[2] * n[1], p[2] * n[0] - p[0] * n[2], p[0] * n[1] - p[1] * n[0],
                       n[0], n[1], n[2]};
      E += rho;
      for (int k = 0; k < 6; ++k) {
        for (int l = 0; l <= k; ++l) H[6 * k + l] += w * row[k] * row[l];
        g[k] -= w * r * row[k];
      }
    }
    for (int k = 0; k < 6; ++k)
      for (int l = 0; l < k; ++l) H[6 * l + k] = H[6 * k + l];
    Ef = E;
    double tr = 0.0;
    for (int k = 0; k < 6; ++k) tr += H[7 * k];
    for (int k = 0; k < 6; ++k) H[7 * k] += lambda * (std::fmax(H[7 * k], 0.0) + 1e-9 * tr / 6.0 + 1e-300);
    double x[6];
    if (!(E > 0.0) || !oracle::solve_ldlt(H, g, x)) {
      st = NO_STEP;
      break;
    }
    double wn = std::sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
    double vn = std::sqrt(x[3] * x[3] + x[4] * x[4] + x[5] * x[5]);
    double f = std::fmin(1.0, std::fmin(0.05 / std::fmax(wn, 1e-300), P.radius / std::fmax(vn, 1e-300)));
    for (int k = 0; k < 6; ++k) x[k] *= f;
    double Rw[9], tw[3], Tn[16];
    rodrigues(x, Rw);
    for (int i = 0; i < 3; ++i)
      tw[i] = P.ct[i] - ((Rw[3 * i] * P.ct[0] + Rw[3 * i + 1] * P.ct[1]) + Rw[3 * i + 2] * P.ct[2]) + x[3 + i];
    for (int i = 0; i < 3; ++i) {
      for (int j = 0; j < 4; ++j) Tn[4 * i + j] = (Rw[3 * i] * T[j] + Rw[3 * i + 1] * T[4 + j]) + Rw[3 * i + 2] * T[8 + j];
      Tn[4 * i + 3] += tw[i];
    }
    Tn[12] = Tn[13] = Tn[14] = 0.0;
    Tn[15] = 1.0;
    double En, dE;
    trial(P, T, Tn, corr, delta, &En, &dE);
    if (dE < 0.0) {
      std::memcpy(T, Tn, sizeof T);
      Ef = En;
      lambda = std::fmax(lambda * 0.1, 1e-12);
      steps = (uint32_t)(it + 1);
      if (-dE <= 1e-12 * E) st = CONVERGED;
    } else {
      lambda *= 10.0;
      if (lambda > 1e12) st = DAMPED_OUT;
    }
  }
  std::memcpy(T_out, T, sizeof T);
  *E_out = Ef;
  *steps_out = steps;
  *status = st;
}

}  // namespace

extern "C" {

// The refinement of k poses of one pair (source n_src x 3; target n_tgt x 3 with unit normals;
// correspondence radius; base poses T0 k x 16 row-major). Outputs: T_out k x 16, E_out the energy at
// the final pose over its last correspondences, steps the last accepted iteration (1-based, 0 for
// none), status 0 = ran to the iteration cap, 1 = the energy stopped dropping (relative 1e-12),
// 2 = the damping passed 1e12, 3 = no step (zero energy or a non-finite solve). OpenMP over poses.
void oracle_huber_lm(const double *src, size_t n_src, const double *tp, const double *tn, size_t n_tgt, double radius,
                     const double *T0, size_t k, double delta, int iterations, int threads, double *T_out,
                     double *E_out, uint32_t *steps, int32_t *status) {
  oracle::Cloud tree(tp, n_tgt, radius > 0.0 ? radius : 1.0);
  Pair P{src, tp, tn, n_src, {0, 0, 0}, {0, 0, 0}, radius, &tree};
  for (int a = 0; a < 3; ++a) {
    double m = 0.0;
    for (size_t i = 0; i < n_src; ++i) m += src[3 * i + a];
    P.cs[a] = n_src ? m / (double)n_src : 0.0;
    m = 0.0;
    for (size_t i = 0; i < n_tgt; ++i) m += tp[3 * i + a];
    P.ct[a] = n_tgt ? m / (double)n_tgt : 0.0;
  }
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads > 0 ? threads : omp_get_max_threads())
  for (long c = 0; c < (long)k; ++c)
    lm_one(P, T0 + 16 * c, delta, iterations, T_out + 16 * c, E_out + c, steps + c, status + c);
}

}  // extern "C"
