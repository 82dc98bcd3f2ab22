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


# ---- the batched LM's FP64 restatement (oracle/huber_lm.cpp) --------------------------------------


def _seams(n_poses=12, seed=30):
    """The seam workload on a small vessel (4 fragments, 40 t high): clouds of 10k-40k points per
    side whose normals span 3-D, base poses within 1 degree and 0.1 t of the truth."""
    from sherd_hip import synth

    col = synth.vessel(t=1.0, edges_per_t=2.0, n_frag=4, seed=11, height_t=40.0, rmin_t=15.0, rmax_t=20.0)
    return synth.seam_workload(col, poses=n_poses, seed=seed)


def _gap(T, U):
    """(rotation angle in rad, translation distance) between poses."""
    dr = T[..., :3, :3] @ np.swapaxes(U[..., :3, :3], -1, -2)
    # |R - I|_F = 2 sqrt(2) sin(angle / 2): no loss of precision at small angles (unlike arccos).
    ang = 2.0 * np.arcsin(np.minimum(np.linalg.norm(dr - np.eye(3), axis=(-2, -1)) / (2.0 * np.sqrt(2.0)), 1.0))
    return ang, np.linalg.norm(T[..., :3, 3] - U[..., :3, 3], axis=-1)


@pytest.mark.parametrize("delta", [1e30, 0.05])
def test_lm_recovers_an_exact_copy(delta):
    """Source = the target's own points under the inverse truth: E = 0 exactly at the truth, so the
    refinement must land on it from 1 degree / 0.1 t away."""
    w = _seams()
    p = w["pairs"][0]
    truth = p["truth"]
    inv = np.linalg.inv(truth)
    src = p["target"] @ inv[:3, :3].T + inv[:3, 3]
    T0 = w["T0"][w["cand_pair"] == 0]
    T, E, steps, status = O.huber_lm(src, p["target"], p["target_n"], w["radius"], T0, delta, 60)
    ang, tr = _gap(T, np.broadcast_to(truth, T.shape))
    assert ang.max() < 1e-9 and tr.max() < 1e-9, (ang.max(), tr.max())
    assert np.all(steps > 0) and np.all(status != O.LM_RAN_OUT)
    assert np.all(E <= 1e-16 * len(src))


def _rodrigues_scipy(w):
    from scipy.spatial.transform import Rotation

    return Rotation.from_rotvec(w).as_matrix()


def test_lm_first_step_equals_an_independent_damped_gauss_newton_step():
    """One iteration, recomputed with numpy/scipy from the definition: IRLS weights psi(r)/r, rows
    [p' x n, n] about the target centroid, Marquardt damping with its floor, numpy's solve, the trust
    region, scipy's rotation vector, the rotation about the target centroid, the accept test."""
    w = _seams()
    delta, lam = w["delta"], 1e-3
    for pi in range(3):
        p = w["pairs"][pi]
        T0 = w["T0"][w["cand_pair"] == pi][:3]
        T1, _, steps, _ = O.huber_lm(p["source"], p["target"], p["target_n"], w["radius"], T0, delta, 1)
        ct = p["target"].sum(0) / len(p["target"])
        for i in range(len(T0)):
            moved = H.base_points(T0[i], p["source"])
            c = O.icp_correspondences(moved, p["target"], w["radius"])
            m = c >= 0
            q, n = p["target"][c[m]], p["target_n"][c[m]]
            pp = moved[m] - ct
            r = np.einsum("ij,ij->i", moved[m] - q, n)
            rho, psi = H.rho_psi(r, delta)
            wt = np.where(r == 0, 1.0, psi / np.where(r == 0, 1.0, r))
            J = np.concatenate([np.cross(pp, n), n], axis=1)
            Hm = (J * wt[:, None]).T @ J
            g = -(J * (wt * r)[:, None]).sum(0)
            d = np.diag(Hm).copy()
            Hd = Hm + np.diag(lam * (np.maximum(d, 0) + 1e-9 * d.sum() / 6 + 1e-300))
            x = np.linalg.solve(Hd, g)
            x *= min(1.0, 0.05 / np.linalg.norm(x[:3]), w["radius"] / np.linalg.norm(x[3:]))
            Rw = _rodrigues_scipy(x[:3])
            S = np.eye(4)
            S[:3, :3] = Rw
            S[:3, 3] = ct - Rw @ ct + x[3:]
            Tn = S @ T0[i]
            rho1 = H.rho_psi(np.einsum("ij,ij->i", H.base_points(Tn, p["source"])[m] - q, n), delta)[0]
            if (rho1 - rho).sum() < 0.0:
                assert steps[i] == 1
                ang, tr = _gap(T1[i], Tn)
                assert ang < 1e-12 and tr < 1e-11, (pi, i, ang, tr)
            else:
                assert steps[i] == 0 and np.array_equal(T1[i], T0[i])


def test_lm_converged_poses_are_stationary():
    """Where the refinement stops because the energy no longer drops, the gradient of the energy
    (oracle/huber.py, over the correspondences at the final pose) has all but vanished."""
    w = _seams(n_poses=18)
    for pi, p in enumerate(w["pairs"]):
        T0 = w["T0"][w["cand_pair"] == pi]
        T, _, _, status = O.huber_lm(p["source"], p["target"], p["target_n"], w["radius"], T0, w["delta"], 40)
        for i in range(len(T0)):
            g = []
            for M in (T0[i], T[i]):
                moved = H.base_points(M, p["source"])
                c = O.icp_correspondences(moved, p["target"], w["radius"])
                m = c >= 0
                g.append(H.gradient(moved[m], p["target"][c[m]], p["target_n"][c[m]], np.zeros(6), w["delta"]))
            if status[i] == O.LM_CONVERGED:
                assert np.linalg.norm(g[1]) <= 1e-5 * np.linalg.norm(g[0]), (pi, i)


def test_lm_is_equivariant_under_a_rigid_motion_of_the_target():
    w = _seams()
    p = w["pairs"][1]
    T0 = w["T0"][w["cand_pair"] == 1]
    G = rigid([0.2, -0.5, 0.8], 0.9, [12.0, -7.0, 3.0])
    tgt = p["target"] @ G[:3, :3].T + G[:3, 3]
    tn = p["target_n"] @ G[:3, :3].T
    T, _, s1, _ = O.huber_lm(p["source"], p["target"], p["target_n"], w["radius"], T0, w["delta"], 30)
    U, _, s2, _ = O.huber_lm(p["source"], tgt, tn, w["radius"], G @ T0, w["delta"], 30)
    ang, tr = _gap(U, G @ T)
    # The last accepted step can differ by one where the energy's drop is at its rounding.
    assert ang.max() < 1e-9 and tr.max() < 1e-8 and np.abs(s1.astype(int) - s2.astype(int)).max() <= 1
