// Small f64 linear algebra for one lane: the point-to-plane solve (Eigen-pivoted LDLT,
// crates/sherd-core/src/matching/icp.rs:661-753), the Euler update (icp.rs:770-778), the
// point-to-point rotation (a 3x3 SVD inside Umeyama, icp.rs:539-592) and the 4x4 product and
// inverse the poses need (verify.rs:739-798).
#pragma once

#include "common.hpp"

namespace sherd {

SH_HD void eye4(double *m) {
  for (int i = 0; i < 16; ++i) m[i] = (i % 5 == 0) ? 1.0 : 0.0;
}

// nalgebra's product: each entry summed over k left to right (c may alias neither a nor b).
SH_HD void mul_nn(const double *a, const double *b, double *c, int n) {
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      double acc = a[i * n] * b[j];
      for (int k = 1; k < n; ++k) acc = acc + a[i * n + k] * b[k * n + j];
      c[i * n + j] = acc;
    }
}

SH_HD double det3(const double *m) {
  return m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) + m[2] * (m[3] * m[7] - m[4] * m[6]);
}

// icp.rs:661-674: Eigen's transposition pivot order on the original diagonal.
SH_HD void eigen_pivots(const double *a, int *perm) {
  for (int i = 0; i < 6; ++i) perm[i] = i;
  for (int k = 0; k < 6; ++k) {
    int best = k;
    for (int i = k + 1; i < 6; ++i)
      if (fabs(a[6 * perm[i] + perm[i]]) > fabs(a[6 * perm[best] + perm[best]])) best = i;
    int t = perm[k];
    perm[k] = perm[best];
    perm[best] = t;
  }
}

// icp.rs:692-753: LDL^T on the permuted matrix, pseudo-inverse of D below the smallest normal. The
// working arrays come from the caller (m 36, y 6, temp 6, perm 6): a kernel passes LDS, where the
// pivot-indexed accesses cost no private memory and the values hold no registers between steps.
SH_HD bool solve_ldlt_ws(const double *a, const double *b, double *x, double *m, double *y, double *temp, int *perm) {
  eigen_pivots(a, perm);
  for (int k = 0; k < 6; ++k)
    for (int l = 0; l < 6; ++l) m[6 * k + l] = a[6 * perm[k] + perm[l]];
  for (int k = 0; k < 6; ++k) {
    if (k > 0) {
      for (int j = 0; j < k; ++j) temp[j] = m[6 * j + j] * m[6 * k + j];
      double diagonal = m[6 * k + k];
      for (int j = 0; j < k; ++j) diagonal -= m[6 * k + j] * temp[j];
      m[6 * k + k] = diagonal;
      for (int r = k + 1; r < 6; ++r) {
        double below = m[6 * r + k];
        for (int j = 0; j < k; ++j) below -= m[6 * r + j] * temp[j];
        m[6 * r + k] = below;
      }
    }
    double pivot = m[6 * k + k];
    if (pivot != 0.0)
      for (int r = k + 1; r < 6; ++r) m[6 * r + k] /= pivot;
  }
  for (int k = 0; k < 6; ++k) y[k] = b[perm[k]];
  for (int i = 0; i < 6; ++i) {
    double v = y[i];
    for (int j = 0; j < i; ++j) v -= m[6 * i + j] * y[j];
    y[i] = v;
  }
  for (int i = 0; i < 6; ++i) {
    double pivot = m[6 * i + i];
    y[i] = fabs(pivot) > 2.2250738585072014e-308 ? y[i] / pivot : 0.0;
  }
  for (int i = 5; i >= 0; --i) {
    double v = y[i];
    for (int j = i + 1; j < 6; ++j) v -= m[6 * j + i] * y[j];
    y[i] = v;
  }
  for (int k = 0; k < 6; ++k) x[perm[k]] = y[k];
  for (int k = 0; k < 6; ++k)
    if (!(x[k] - x[k] == 0.0)) return false;
  return true;
}
SH_HD bool solve_ldlt(const double *a, const double *b, double *x) {
  double m[36], y[6], temp[6];
  int perm[6];
  return solve_ldlt_ws(a, b, x, m, y, temp, perm);
}

// icp.rs:770-778: Rz(gamma) Ry(beta) Rx(alpha).
SH_HD void euler_zyx_ws(double al, double be, double ga, double *r, double *ws /* 36 */) {
  double sa = sin(al), ca = cos(al), sb = sin(be), cb = cos(be), sg = sin(ga), cg = cos(ga);
  double *rx = ws, *ry = ws + 9, *rz = ws + 18, *t = ws + 27;
  const double x[9] = {1, 0, 0, 0, ca, -sa, 0, sa, ca}, y[9] = {cb, 0, sb, 0, 1, 0, -sb, 0, cb},
               z[9] = {cg, -sg, 0, sg, cg, 0, 0, 0, 1};
  for (int i = 0; i < 9; ++i) {
    rx[i] = x[i];
    ry[i] = y[i];
    rz[i] = z[i];
  }
  mul_nn(rz, ry, t, 3);
  mul_nn(t, rx, r, 3);
}
SH_HD void euler_zyx(double al, double be, double ga, double *r) {
  double ws[36];
  euler_zyx_ws(al, be, ga, r, ws);
}

// One-sided Jacobi SVD of a 3x3: M = U diag(s) V^T, singular values descending, U completed by a
// cross product where a singular value vanishes.
SH_HD void svd3_ws(const double *m, double *u, double *s, double *v, double *a, double *vv, double *sig, int *ord) {
  for (int i = 0; i < 9; ++i) {
    a[i] = m[i];
    vv[i] = (i % 4 == 0) ? 1.0 : 0.0;
  }
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
        if (fabs(ga) <= 1e-300) continue;
        double o = fabs(ga) / sqrt(al * be > 0 ? al * be : 1e-300);
        off = off < o ? o : off;
        double zeta = (be - al) / (2.0 * ga);
        double t = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
        double c = 1.0 / sqrt(1.0 + t * t), sn = c * t;
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
  for (int j = 0; j < 3; ++j) sig[j] = sqrt(a[j] * a[j] + a[3 + j] * a[3 + j] + a[6 + j] * a[6 + j]);
  for (int j = 0; j < 3; ++j) ord[j] = j;
  for (int x = 0; x < 3; ++x)
    for (int y = x + 1; y < 3; ++y)
      if (sig[ord[y]] > sig[ord[x]]) {
        int t = ord[x];
        ord[x] = ord[y];
        ord[y] = t;
      }
  double smax = sig[ord[0]];
  for (int c = 0; c < 3; ++c) {
    int j = ord[c];
    s[c] = sig[j];
    for (int i = 0; i < 3; ++i) v[3 * i + c] = vv[3 * i + j];
    for (int i = 0; i < 3; ++i) u[3 * i + c] = (sig[j] > 1e-14 * smax && sig[j] > 0) ? a[3 * i + j] / sig[j] : 0.0;
  }
  double n0 = sqrt(u[0] * u[0] + u[3] * u[3] + u[6] * u[6]);
  if (n0 == 0.0) {
    u[0] = 1;
    u[3] = 0;
    u[6] = 0;
  }
  double n1 = sqrt(u[1] * u[1] + u[4] * u[4] + u[7] * u[7]);
  if (n1 == 0.0) {
    double x = u[0], y = u[3], z = u[6];
    double w0 = fabs(x) < 0.9 ? 1.0 : 0.0, w1 = fabs(x) < 0.9 ? 0.0 : 1.0, w2 = 0.0;
    double c0 = y * w2 - z * w1, c1 = z * w0 - x * w2, c2 = x * w1 - y * w0;
    double l = sqrt(c0 * c0 + c1 * c1 + c2 * c2);
    u[1] = c0 / l;
    u[4] = c1 / l;
    u[7] = c2 / l;
  }
  double n2 = sqrt(u[2] * u[2] + u[5] * u[5] + u[8] * u[8]);
  if (n2 == 0.0) {
    u[2] = u[3] * u[7] - u[6] * u[4];
    u[5] = u[6] * u[1] - u[0] * u[7];
    u[8] = u[0] * u[4] - u[3] * u[1];
  }
}

SH_HD void svd3(const double *m, double *u, double *s, double *v) {
  double a[9], vv[9], sig[3];
  int ord[3];
  svd3_ws(m, u, s, v, a, vv, sig, ord);
}

// Umeyama's rotation and translation from the means and the scaled covariance (icp.rs:578-591).
// ws: 78 doubles and 3 ints of working space (LDS in a kernel, as solve_ldlt_ws's).
SH_HD void umeyama_pose_ws(const double *sigma, const double *mp, const double *mq, double *out, double *ws, int *iws) {
  double *u = ws, *s = ws + 9, *v = ws + 12, *vt = ws + 21, *d = ws + 30, *ud = ws + 39, *rot = ws + 48;
  svd3_ws(sigma, u, s, v, ws + 57, ws + 66, ws + 75, iws);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) vt[3 * i + j] = v[3 * j + i];
  for (int i = 0; i < 9; ++i) d[i] = (i % 4 == 0) ? 1.0 : 0.0;
  if (det3(u) * det3(vt) < 0.0) d[8] = -1.0;
  mul_nn(u, d, ud, 3);
  mul_nn(ud, vt, rot, 3);
  eye4(out);
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 3; ++j) out[4 * i + j] = rot[3 * i + j];
    out[4 * i + 3] = mq[i] - (rot[3 * i] * mp[0] + rot[3 * i + 1] * mp[1] + rot[3 * i + 2] * mp[2]);
  }
}

SH_HD void umeyama_pose(const double *sigma, const double *mp, const double *mq, double *out) {
  double ws[78];
  int iws[3];
  umeyama_pose_ws(sigma, mp, mq, out, ws, iws);
}

// The same Umeyama step for one lane with every array index fixed at compile time (the loops
// unrolled, the singular values put in order by swapping whole columns rather than through an
// index array), so it stays in registers: the operations of svd3_ws and umeyama_pose_ws in the same
// order, hence the same bits, without the LDS round trip per access of the working-space form.
__device__ __forceinline__ void mul3_reg(const double *a, const double *b, double *c) {
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      double acc = a[i * 3] * b[j];
      acc = acc + a[i * 3 + 1] * b[3 + j];
      acc = acc + a[i * 3 + 2] * b[6 + j];
      c[i * 3 + j] = acc;
    }
}

__device__ __forceinline__ void svd3_reg(const double *m, double *u, double *s, double *v) {
  double a[9], vv[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    a[i] = m[i];
    vv[i] = (i % 4 == 0) ? 1.0 : 0.0;
  }
  for (int sweep = 0; sweep < 60; ++sweep) {
    double off = 0.0;
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
      for (int q = p + 1; q < 3; ++q) {
        double al = 0, be = 0, ga = 0;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          al += a[3 * i + p] * a[3 * i + p];
          be += a[3 * i + q] * a[3 * i + q];
          ga += a[3 * i + p] * a[3 * i + q];
        }
        if (fabs(ga) <= 1e-300) continue;
        double o = fabs(ga) / sqrt(al * be > 0 ? al * be : 1e-300);
        off = off < o ? o : off;
        double zeta = (be - al) / (2.0 * ga);
        double t = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
        double c = 1.0 / sqrt(1.0 + t * t), sn = c * t;
#pragma unroll
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
  // Columns with their singular values, ordered by the same exchanges as svd3_ws's index sort.
  double sg[3], ac[3][3], vc[3][3];
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    sg[j] = sqrt(a[j] * a[j] + a[3 + j] * a[3 + j] + a[6 + j] * a[6 + j]);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      ac[j][i] = a[3 * i + j];
      vc[j][i] = vv[3 * i + j];
    }
  }
#pragma unroll
  for (int x = 0; x < 3; ++x)
#pragma unroll
    for (int y = x + 1; y < 3; ++y)
      if (sg[y] > sg[x]) {
        double t = sg[x];
        sg[x] = sg[y];
        sg[y] = t;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          t = ac[x][i];
          ac[x][i] = ac[y][i];
          ac[y][i] = t;
          t = vc[x][i];
          vc[x][i] = vc[y][i];
          vc[y][i] = t;
        }
      }
  const double smax = sg[0];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    s[c] = sg[c];
#pragma unroll
    for (int i = 0; i < 3; ++i) v[3 * i + c] = vc[c][i];
#pragma unroll
    for (int i = 0; i < 3; ++i) u[3 * i + c] = (sg[c] > 1e-14 * smax && sg[c] > 0) ? ac[c][i] / sg[c] : 0.0;
  }
  double n0 = sqrt(u[0] * u[0] + u[3] * u[3] + u[6] * u[6]);
  if (n0 == 0.0) {
    u[0] = 1;
    u[3] = 0;
    u[6] = 0;
  }
  double n1 = sqrt(u[1] * u[1] + u[4] * u[4] + u[7] * u[7]);
  if (n1 == 0.0) {
    double x = u[0], y = u[3], z = u[6];
    double w0 = fabs(x) < 0.9 ? 1.0 : 0.0, w1 = fabs(x) < 0.9 ? 0.0 : 1.0, w2 = 0.0;
    double c0 = y * w2 - z * w1, c1 = z * w0 - x * w2, c2 = x * w1 - y * w0;
    double l = sqrt(c0 * c0 + c1 * c1 + c2 * c2);
    u[1] = c0 / l;
    u[4] = c1 / l;
    u[7] = c2 / l;
  }
  double n2 = sqrt(u[2] * u[2] + u[5] * u[5] + u[8] * u[8]);
  if (n2 == 0.0) {
    u[2] = u[3] * u[7] - u[6] * u[4];
    u[5] = u[6] * u[1] - u[0] * u[7];
    u[8] = u[0] * u[4] - u[3] * u[1];
  }
}

// umeyama_pose_ws in registers (sigma, mp, mq may be in LDS; out gets the 4x4).
__device__ __forceinline__ void umeyama_pose_reg(const double *sigma, const double *mp, const double *mq,
                                                 double *out) {
  double u[9], s[3], v[9], vt[9], d[9], ud[9], rot[9];
  svd3_reg(sigma, u, s, v);
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) vt[3 * i + j] = v[3 * j + i];
#pragma unroll
  for (int i = 0; i < 9; ++i) d[i] = (i % 4 == 0) ? 1.0 : 0.0;
  if (det3(u) * det3(vt) < 0.0) d[8] = -1.0;
  mul3_reg(u, d, ud);
  mul3_reg(ud, vt, rot);
#pragma unroll
  for (int i = 0; i < 16; ++i) out[i] = (i % 5 == 0) ? 1.0 : 0.0;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
#pragma unroll
    for (int j = 0; j < 3; ++j) out[4 * i + j] = rot[3 * i + j];
    out[4 * i + 3] = mq[i] - (rot[3 * i] * mp[0] + rot[3 * i + 1] * mp[1] + rot[3 * i + 2] * mp[2]);
  }
}

// verify.rs:739-798 pose_inverse: LU with partial pivoting (first index on a tie), then one
// forward and one back substitution per column.
SH_HD void pose_inverse(const double *T, double *out) {
  double m[4][4];
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) m[i][j] = T[4 * i + j];
  int piv[4] = {0, 1, 2, 3};
  for (int k = 0; k < 4; ++k) {
    int best = k;
    for (int i = k + 1; i < 4; ++i)
      if (fabs(m[i][k]) > fabs(m[best][k])) best = i;
    if (best != k) {
      for (int j = 0; j < 4; ++j) {
        double t = m[k][j];
        m[k][j] = m[best][j];
        m[best][j] = t;
      }
      int t = piv[k];
      piv[k] = piv[best];
      piv[best] = t;
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

}  // namespace sherd
