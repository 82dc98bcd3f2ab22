"""FP64 restatement of the robust point-to-plane seam energy and its se(3) gradient.

Test infrastructure (the oracle): only the tests, the smoke check and the benchmark's CPU baseline
use it, and only as the checker.

The reference has no function for this energy; its nearest relative is ICP's point-to-plane
normal equations (crates/sherd-core/src/matching/icp.rs:599-640), whose rows are [p x n, n] and
residuals r = (p - q) . n. Here, for a pose with base transform T0 and frozen correspondences
(source point s_i -> target point q_i with normal n_i):

    p_i(xi) = exp(omega^) (T0 s_i) + v,        xi = (omega, v)   (rotation first, icp.rs:432-433)
    r_i     = (p_i - q_i) . n_i
    E(xi)   = sum_i rho(r_i),   rho(r) = r^2 / 2 if |r| <= delta else delta (|r| - delta / 2)
    dE/dv     = sum_i psi(r_i) n_i
    dE/domega = J_l(omega)^T sum_i psi(r_i) (exp(omega^) T0 s_i) x n_i,    psi = rho'

J_l is SO(3)'s left Jacobian. At xi = 0 the gradient is sum psi(r) [p x n, n], and as delta grows it
becomes normal_equations' JTr over the same pairs.

Parity: pinned by FP64 central finite differences (h = 1e-4) of energy(), and by the delta -> oo,
xi = 0 identity against the C++ oracle's normal_equations (a restatement of icp.rs:599-640 itself
pinned by the reference's ICP ladder dumps). Bit level: parity unpinned (no reference values exist).
"""

import numpy as np


def hat(w):
    return np.array([[0.0, -w[2], w[1]], [w[2], 0.0, -w[0]], [-w[1], w[0], 0.0]])


def rodrigues(w):
    """exp(w^) in f64."""
    th = np.linalg.norm(w)
    K = hat(w)
    if th < 1e-8:
        return np.eye(3) + K + 0.5 * K @ K
    return np.eye(3) + np.sin(th) / th * K + (1.0 - np.cos(th)) / (th * th) * K @ K


def left_jacobian(w):
    th = np.linalg.norm(w)
    K = hat(w)
    if th < 1e-6:
        return np.eye(3) + 0.5 * K + K @ K / 6.0
    return np.eye(3) + (1.0 - np.cos(th)) / (th * th) * K + (th - np.sin(th)) / th ** 3 * K @ K


def base_points(T0, s):
    """T0 s_i, row by row: (a x + b y) + c z + t."""
    T0 = np.asarray(T0, np.float64).reshape(4, 4)
    s = np.asarray(s, np.float64)
    return np.stack([((T0[k, 0] * s[:, 0] + T0[k, 1] * s[:, 1]) + T0[k, 2] * s[:, 2]) + T0[k, 3] for k in range(3)],
                    axis=1)


def residuals(p0, q, n, xi):
    R = rodrigues(np.asarray(xi[:3], np.float64))
    a = p0 @ R.T
    p = a + np.asarray(xi[3:], np.float64)
    return np.einsum("ij,ij->i", p - q, n), a


def rho_psi(r, delta):
    inner = np.abs(r) <= delta
    rho = np.where(inner, 0.5 * r * r, delta * (np.abs(r) - 0.5 * delta))
    psi = np.where(inner, r, delta * np.sign(r))
    return rho, psi


def energy(p0, q, n, xi, delta):
    """E(xi) over the corresponding pairs (p0_i = T0 s_i, q_i, n_i)."""
    r, _ = residuals(p0, q, n, xi)
    return float(rho_psi(r, delta)[0].sum())


def gradient(p0, q, n, xi, delta):
    """dE/dxi = (dE/domega, dE/dv)."""
    r, a = residuals(p0, q, n, xi)
    _, psi = rho_psi(r, delta)
    gv = (psi[:, None] * n).sum(0)
    gw = left_jacobian(np.asarray(xi[:3], np.float64)).T @ (psi[:, None] * np.cross(a, n)).sum(0)
    return np.concatenate([gw, gv])


def finite_difference(p0, q, n, xi, delta, h=1e-4):
    """Central differences of energy() in each twist coordinate."""
    xi = np.asarray(xi, np.float64)
    g = np.empty(6)
    for k in range(6):
        e = np.zeros(6)
        e[k] = h
        g[k] = (energy(p0, q, n, xi + e, delta) - energy(p0, q, n, xi - e, delta)) / (2 * h)
    return g
