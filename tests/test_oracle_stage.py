"""The R §5.1-5.5 oracle (oracle/stage.py) against what pins it: nms.rs's own tests and the slab
dump's hypotheses and kept lists (crates/sherd-parity/src/stages: the coarse and stage-1 rows)."""

import os
import sys

import numpy as np
import pytest

import slab_data as S

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import stage as G  # noqa: E402

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "slab")


def rot_z(deg):
    c, s = np.cos(np.radians(deg)), np.sin(np.radians(deg))
    return np.array([[c, -s, 0.0], [s, c, 0.0], [0.0, 0.0, 1.0]])


def table(rots, xs):
    return np.array([np.concatenate([r.reshape(9), [x, 0.0, 0.0]]) for r, x in zip(rots, xs)])


def test_the_order_is_descending_with_the_index_breaking_ties():
    # nms.rs:121-127.
    score = [0.1, 0.5, 0.5, 0.2, 0.5]
    assert list(G.order_by_score(score)) == [1, 2, 4, 3, 0]
    assert list(G.order_by_score(score, 3)) == [1, 2, 4]
    assert list(G.order_by_score([], 10)) == []


def test_one_pose_survives_a_cluster_and_a_distinct_pose_beside_it():
    # nms.rs:131-143.
    P = table([rot_z(0), rot_z(2), rot_z(40), rot_z(0)], [0.0, 0.1, 0.0, 5.0])
    score = np.array([0.9, 0.8, 0.7, 0.6])
    order = G.order_by_score(score)
    assert G.nms(order, P, score, 1.0, 10, 0.0) == [0, 2, 3]
    assert G.nms(order, P, score, 9.0, 10, 0.0) == [0, 2]


def test_the_walk_stops_at_topk_and_breaks_at_the_floor():
    # nms.rs:146-157.
    P = table([rot_z(0)] * 4, [0.0, 10.0, 20.0, 30.0])
    score = np.array([0.9, 0.5, 0.5, 0.1])
    order = G.order_by_score(score)
    assert G.nms(order, P, score, 1.0, 2, 0.0) == [0, 1]
    assert G.nms(order, P, score, 1.0, 10, 0.5) == [0, 1, 2]
    assert G.nms(order, P, score, 1.0, 10, 0.6) == [0]
    assert G.nms([], P, score, 1.0, 10, 0.0) == []


def test_topk_zero_keeps_one_pose():
    # nms.rs:168-178.
    P = table([rot_z(0)] * 3, [0.0, 10.0, 20.0])
    score = np.array([0.9, 0.5, 0.1])
    order = G.order_by_score(score)
    assert G.nms(order, P, score, 1.0, 0, 0.0) == [0]
    assert G.nms(order, P, score, 1.0, 0, 1.0) == []
    assert G.nms([], P, score, 1.0, 0, 0.0) == []


def test_the_duplicate_test_is_a_conjunction_and_the_turn_is_eighteen_degrees():
    # nms.rs:183-201.
    P = table([rot_z(0), rot_z(90), rot_z(0)], [0.0, 0.0, 100.0])
    score = np.array([0.9, 0.8, 0.7])
    assert G.nms(G.order_by_score(score), P, score, 1.0, 10, 0.0) == [0, 1, 2]
    I = table([rot_z(0)], [0.0])[0]
    assert G.same_candidate(table([rot_z(18.1)], [0.0])[0], I, 1.0)
    assert not G.same_candidate(table([rot_z(18.3)], [0.0])[0], I, 1.0)
    assert G.same_candidate(table([rot_z(-18.1)], [0.0])[0], I, 1.0)


@pytest.fixture(scope="module")
def slab():
    a, b, p = S.piece("pieceA"), S.piece("pieceB"), S.pair()
    st = np.load(os.path.join(GOLD, "stage.npz"))
    fa = dict(p=a["brk_P"], ns=a["brk_ns"], f=a["brk_f"], t=a["brk_t"], dih=a["brk_dih"], sub=a["brk_sub"])
    fb = dict(p=b["brk_P"], ns=b["brk_ns"], f=b["brk_f"], t=b["brk_t"], dih=b["brk_dih"], sub=b["brk_sub"])
    return fa, fb, p, st


def test_slab_hypotheses_are_the_dumps(slab):
    fa, fb, p, _ = slab
    pa, pb = G.build_with(fa["dih"], fb["dih"], p["hyp_ia"], p["hyp_ib"])
    assert np.array_equal(pa, p["hyp_pa"]) and np.array_equal(pb, p["hyp_pb"])
    hyp = G.poses(fa, fb, p["hyp_ia"], p["hyp_ib"], pa, pb)
    want = S.hypothesis_poses(S.piece("pieceA"), S.piece("pieceB"), p["hyp_ia"], p["hyp_ib"], pa, pb)
    assert hyp.tobytes() == want.tobytes()


def test_slab_suppressions_on_the_dumps_own_walks(slab):
    # R §5.3 over the dump's nms1.order and R §5.5 over its nms2.order keep exactly its lists.
    fa, fb, p, st = slab
    sc = p["scales"]
    hyp = G.poses(fa, fb, p["hyp_ia"], p["hyp_ib"], p["hyp_pa"], p["hyp_pb"])
    kept1 = G.nms(st["nms1_order"], hyp, p["coarse_cs"], sc[10], 250, G.COARSE_FLOOR)
    assert kept1 == list(p["nms1_kept"])
    kept2 = G.nms(st["nms2_order"], G.pose12_of(p["s1_T"]), p["s1_score"], sc[10], 10, G.STAGE1_FLOOR)
    assert kept2 == list(p["nms2_kept"])
    # The walks are descending by score (their ties are numpy's, PMC-6).
    assert np.all(np.diff(p["coarse_cs"][st["nms1_order"]]) <= 0)
    assert np.all(np.diff(p["s1_score"][st["nms2_order"]]) <= 0)
