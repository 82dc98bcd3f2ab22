"""The R §9 oracle (oracle/refine.py) against what pins it: the reference's own rng and refine
tests (crates/sherd-core/src/rng.rs:183-289, crates/sherd-core/src/refine.rs:369-497) and the slab
dump's refine/ row (selection, radii, both rungs, fitness, rmse, final poses)."""

import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import refine as R  # noqa: E402

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "slab")
PIECES = ("pieceA", "pieceB")


def rot_deg(A, B):
    R3 = A[:3, :3] @ B[:3, :3].T
    return np.degrees(np.arccos(np.clip((np.trace(R3) - 1.0) / 2.0, -1.0, 1.0)))


# ---- rng.rs ----------------------------------------------------------------------------------------


def test_chacha8_stream_is_the_pinned_one():
    # rng.rs:183-189: ChaCha8Rng::seed_from_u64(0), first four u32 words.
    g = R.ChaCha8(0)
    assert [g.next_u32() for _ in range(4)] == [0xA79A3B6C, 0xB585F767, 0xBAD8C037, 0x7746A55F]
    a, b, c = R.ChaCha8(0), R.ChaCha8(0), R.ChaCha8(1)
    wa = [a.next_u32() for _ in range(40)]  # across a block boundary
    assert wa == [b.next_u32() for _ in range(40)]
    assert wa[:4] != [c.next_u32() for _ in range(4)]


def test_without_replacement_is_distinct_capped_and_a_prefix():
    # rng.rs:258-277.
    a = R.without_replacement(1000, 20_000, R.ChaCha8(0))
    assert len(a) == 1000 and len(set(a)) == 1000
    b = R.without_replacement(1000, 100, R.ChaCha8(0))
    assert b == a[:100]
    assert R.without_replacement(1000, 100, R.ChaCha8(1)) != b
    assert R.without_replacement(0, 10, R.ChaCha8(0)) == []
    assert R.without_replacement(10, 0, R.ChaCha8(0)) == []


def test_uniform_below_is_uniform_and_in_range():
    # rng.rs:279-289.
    g = R.ChaCha8(7)
    counts = np.zeros(5, int)
    for _ in range(50_000):
        k = R.uniform_below(g, 5)
        assert 0 <= k < 5
        counts[k] += 1
    assert np.all((counts >= 9_000) & (counts < 11_000)), counts
    assert R.uniform_below(g, 1) == 0


def test_uniform_below_rejects_the_biased_low_words():
    # Lemire's threshold: for n = 3, a low word below 2^32 mod 3 = 1 is redrawn. A generator fed
    # known words shows which draw survives.
    class Fixed:
        def __init__(self, words):
            self.w = list(words)

        def next_u32(self):
            return self.w.pop(0)

    # 0 * 3 has low word 0 < 1: rejected; 0x80000000 * 3 = 0x1_8000_0000: high word 1, low kept.
    assert R.uniform_below(Fixed([0, 0x80000000]), 3) == 1
    assert R.uniform_below(Fixed([0xFFFFFFFF]), 3) == 2


# ---- refine.rs --------------------------------------------------------------------------------------


def test_vertex_normals_are_open3ds_area_weighted_sum():
    # refine.rs:369-400: (0, 10, 1)/sqrt(101) for the shared vertex, a lone face's own normal, and
    # (0, 0, 1) for a vertex no face references.
    V = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [10, 0, 0], [9, 9, 9]], dtype=np.float64)
    F = np.array([[0, 1, 2], [0, 3, 4]])
    n = R.vertex_normals(V, F)
    s = np.sqrt(101.0)
    assert np.abs(n[0] - [0.0, 10.0 / s, 1.0 / s]).max() < 1e-15
    assert n[1].tobytes() == np.array([0.0, 0.0, 1.0]).tobytes()
    assert n[5].tobytes() == np.array([0.0, 0.0, 1.0]).tobytes()


def test_vertex_normals_add_in_face_order():
    # The vectorised accumulation equals the sequential face loop bit for bit.
    rng = np.random.default_rng(3)
    V = rng.normal(size=(40, 3)) * np.array([1.0, 1e-3, 1e3])
    F = rng.integers(0, 40, size=(300, 3))
    got = R.vertex_normals(V, F)
    acc = np.zeros((40, 3))
    for f in F:
        o = V[f[0]]
        e1, e2 = V[f[1]] - o, V[f[2]] - o
        face = np.array([e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]])
        for v in f:
            for a in range(3):
                acc[v, a] = acc[v, a] + face[a]
    for v in range(40):
        L = np.sqrt((acc[v, 0] * acc[v, 0] + acc[v, 1] * acc[v, 1]) + acc[v, 2] * acc[v, 2])
        want = acc[v] / L if L > 0 else np.array([0.0, 0.0, 1.0])
        assert got[v].tobytes() == want.tobytes(), v


def test_the_selection_is_the_fracture_faces_own_vertices():
    # refine.rs:404-419.
    V = np.array([[0.0, 0, 0], [1.0, 0, 0], [10.0, 0, 0]])
    C = np.array([[0.1, 0, 0], [1.2, 0, 0], [9.5, 0, 0]])
    frac = np.array([True, False, True])
    assert list(R.select_candidates(V, C, frac, 2.0, 0.1)) == [0]
    assert list(R.select_candidates(V, C, frac, 2.0, 1.0)) == [0, 2]
    capped = R.cap_selection(R.select_candidates(V, C, frac, 2.0, 1.0), 1, 0)
    assert len(capped) == 1 and capped[0] in (0, 2)


def test_the_selection_is_strict_at_the_radius():
    # d < radius: a vertex exactly one radius from its centroid is out; a hair closer is in.
    C = np.array([[0.0, 0.0, 0.0]])
    r = max(0.15 * 2.0, 1.5 * 0.1)
    V = np.array([[r, 0.0, 0.0], [np.nextafter(r, 0.0), 0.0, 0.0]])
    assert list(R.select_candidates(V, C, [True], 2.0, 0.1)) == [1]


def saddle():
    pts, nrm = [], []
    for i in range(20):
        for j in range(20):
            x, y = i * 0.1, j * 0.1
            z = 0.4 * x * x - 0.25 * y * y + 0.1 * x * y
            g = np.array([-(0.8 * x + 0.1 * y), -(0.1 * x - 0.5 * y), 1.0])
            pts.append([x, y, z])
            nrm.append(g / np.linalg.norm(g))
    return np.array(pts), np.array(nrm)


def test_the_walk_starts_at_the_groups_first_fragment_and_reaches_each_once():
    # refine.rs:423-470: used order (1,2),(0,1); the walk from 0 takes (0,1) then (1,2), and both
    # moved fragments come back onto fragment 0's plane; an empty group adds nothing.
    cl = saddle()
    pieces = [(1.0, 0.05, cl)] * 3
    poses = np.stack([np.eye(4)] * 3)
    poses[1][2, 3] = 0.004
    poses[2][2, 3] = 0.006
    out, joins = R.refine_joins(pieces, poses, [[0, 1, 2], []], [(1, 2), (0, 1)])
    assert [(j["fixed"], j["moving"]) for j in joins] == [(0, 1), (1, 2)]
    for n in (1, 2):
        assert abs(out[n][2, 3]) < 1e-9, out[n]
    assert len(joins) == 2


def test_a_join_without_a_cloud_is_skipped():
    # refine.rs:474-497.
    pieces = [(1.0, 0.1, (np.zeros((1, 3)), np.array([[0.0, 0.0, 1.0]]))), (1.0, 0.1, None)]
    poses = np.stack([np.eye(4)] * 2)
    out, joins = R.refine_joins(pieces, poses, [[0, 1]], [(0, 1)])
    assert joins == [] and np.array_equal(out, poses)


@pytest.fixture(scope="module")
def slab():
    z = np.load(os.path.join(GOLD, "refine.npz"))
    frags = [np.load(os.path.join(GOLD, p + ".npz")) for p in PIECES]
    return z, frags


def test_slab_selection_is_the_dumps_own(slab):
    z, frags = slab
    for i, (p, f) in enumerate(zip(PIECES, frags)):
        C = R.face_centroids(f["V"], f["F"])
        idx = R.select_candidates(f["V0"], C, f["frac"], z["thick"][i], z["res"][i])
        assert np.array_equal(idx, z["idx_" + p]), p


def test_slab_refinement_is_the_dumps_own(slab):
    # crates/sherd-parity/src/stages/refine.rs:56-79: radii exact, each rung within 0.2 deg and
    # 0.02 t, fitness within 1e-4, and the final poses.
    z, frags = slab
    pieces = []
    for i, (p, f) in enumerate(zip(PIECES, frags)):
        idx = z["idx_" + p]
        nrm = R.vertex_normals(f["V0"], f["F"])
        pieces.append((z["thick"][i], z["res"][i], (f["V0"][idx], nrm[idx])))
    out, joins = R.refine_joins(pieces, z["poses_before"], [list(z["group_members"])], [tuple(u) for u in z["used"]])
    assert len(joins) == 1
    j = joins[0]
    assert (j["fixed"], j["moving"]) == (z["join_fixed"][0], z["join_moving"][0])
    assert j["dist"].tobytes() == z["join_dist"][0].tobytes()
    t = min(z["thick"])
    for r in range(2):
        want = z["join_T"][0, r]
        assert rot_deg(j["rungs"][r], want) < 0.2
        assert np.abs(j["rungs"][r][:3, 3] - want[:3, 3]).max() < 0.02 * t
    assert abs(j["fitness"] - z["join_fitness"][0]) < 1e-4
    assert j["rmse_t"] < 1e-12 and z["join_rmse_t"][0] < 1e-12
    for n in range(2):
        assert rot_deg(out[n], z["poses_final"][n]) < 0.2
        assert np.abs(out[n][:3, 3] - z["poses_final"][n][:3, 3]).max() < 0.02 * t
