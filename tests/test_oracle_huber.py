"""The FP64 restatement of the robust seam energy (oracle/huber.py), pinned without a reference
value: central finite differences, the delta -> infinity identity with ICP's normal equations
(icp.rs:599-640, through the C++ oracle), and the exponential map's own identities."""

import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import huber as H  # noqa: E402
import oracle_lib as O  # noqa: E402
from test_oracle_known import patch, rigid  # noqa: E402


def problem(seed=0, noise=0.02):
    rng = np.random.default_rng(seed)
    pts, nrm = patch(30, 0.3, -4.5)
    truth = rigid(rng.normal(size=3), 0.04, rng.normal(scale=0.05, size=3))
    src = pts[rng.random(len(pts)) < 0.8] + rng.normal(scale=noise, size=(0, 3)).sum(0)
    src = src + rng.normal(scale=noise, size=src.shape)
    T0 = np.linalg.inv(truth)
    p0 = H.base_points(T0, src)
    corr = O.icp_correspondences(p0, pts, 0.5)
    m = corr >= 0
    return p0, pts, nrm, corr, m


@pytest.mark.parametrize("delta", [1e30, 0.5, 0.05])
@pytest.mark.parametrize("twist", [0, 1])
def test_gradient_against_central_differences(delta, twist):
    p0, pts, nrm, corr, m = problem()
    xi = np.zeros(6) if twist == 0 else np.array([0.01, -0.02, 0.015, 0.03, -0.01, 0.02])
    q, n = pts[corr[m]], nrm[corr[m]]
    g = H.gradient(p0[m], q, n, xi, delta)
    fd = H.finite_difference(p0[m], q, n, xi, delta, h=1e-4)
    # Smooth energies agree to the O(h^2) truncation; a small delta puts a few residuals across the
    # Huber kink within h, which central differences blur.
    tol = 1e-6 if delta > 1.0 else 5e-4
    assert np.abs(g - fd).max() <= tol * np.abs(fd).max()


def test_infinite_delta_at_zero_twist_is_the_normal_equations_jtr():
    p0, pts, nrm, corr, m = problem(1)
    _, jtr = O.normal_equations(p0, pts, nrm, corr)
    g = H.gradient(p0[m], pts[corr[m]], nrm[corr[m]], np.zeros(6), 1e30)
    assert np.abs(g - jtr).max() <= 1e-13 * np.abs(jtr).max()


def test_exponential_map_identities():
    rng = np.random.default_rng(3)
    for _ in range(20):
        w = rng.normal(size=3) * rng.uniform(0, 2)
        R = H.rodrigues(w)
        assert np.allclose(R @ R.T, np.eye(3), atol=1e-14) and abs(np.linalg.det(R) - 1) < 1e-14
        # exp((w + d)^) = exp((J_l(w) d)^) exp(w^) to first order in d.
        d = rng.normal(size=3) * 1e-6
        lhs = H.rodrigues(w + d)
        rhs = H.rodrigues(H.left_jacobian(w) @ d) @ R
        assert np.abs(lhs - rhs).max() < 1e-10
    assert np.allclose(H.rodrigues(np.array([0, 0, np.pi / 2])) @ [1, 0, 0], [0, 1, 0], atol=1e-15)


def test_huber_is_continuous_at_delta_and_quadratic_inside():
    r = np.array([-2.0, -0.5, 0.0, 0.3, 0.5, 0.5 + 1e-12, 3.0])
    rho, psi = H.rho_psi(r, 0.5)
    assert np.allclose(rho[[1, 3]], 0.5 * r[[1, 3]] ** 2)
    assert abs(rho[4] - rho[5]) < 1e-11 and psi[5] == 0.5 and psi[0] == -0.5
    assert rho[6] == 0.5 * (3.0 - 0.25)


def test_energy_is_invariant_under_a_common_rigid_motion():
    p0, pts, nrm, corr, m = problem(4)
    M = rigid([0.3, 0.1, -0.9], 0.7, [1.0, -2.0, 0.5])
    move = lambda x: x @ M[:3, :3].T + M[:3, 3]
    e1 = H.energy(p0[m], pts[corr[m]], nrm[corr[m]], np.zeros(6), 0.05)
    e2 = H.energy(move(p0[m]), move(pts[corr[m]]), nrm[corr[m]] @ M[:3, :3].T, np.zeros(6), 0.05)
    assert abs(e1 - e2) <= 1e-12 * e1
