"""R §5.1-5.5 on the device (crates/sherd-core/src/matching/pair.rs:188-264): hypotheses,
suppression and the batched stage-1 hand-off, against the slab dump and the oracle (oracle/stage.py).

Integer and index results (hypothesis pairs, kept lists) must be identical; hypothesis poses are
the reference's f64 arithmetic and must be bit-identical; the stage-1 transforms come from the
device ICP and are held to the reference's ladder tolerance.
"""

import os
import sys

import numpy as np
import pytest

import oracle_lib as O
import slab_data as S
from test_gpu_match import pose_gap

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import stage as G  # noqa: E402

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "slab")


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def ex():
    import sherd_hip

    e = sherd_hip.HipExecutor(0)
    yield e
    e.close()


@pytest.fixture(scope="module")
def slab(ex):
    a, b, p = S.piece("pieceA"), S.piece("pieceB"), S.pair()
    st = np.load(os.path.join(GOLD, "stage.npz"))
    fa = dict(p=a["brk_P"], ns=a["brk_ns"], f=a["brk_f"], t=a["brk_t"], dih=a["brk_dih"], sub=a["brk_sub"])
    fb = dict(p=b["brk_P"], ns=b["brk_ns"], f=b["brk_f"], t=b["brk_t"], dih=b["brk_dih"], sub=b["brk_sub"])
    da = ex.frames(fa["p"], fa["ns"], fa["f"], fa["t"], fa["dih"], fa["sub"])
    db = ex.frames(fb["p"], fb["ns"], fb["f"], fb["t"], fb["dih"], fb["sub"])
    return dict(a=a, b=b, p=p, st=st, fa=fa, fb=fb, da=da, db=db)


def test_hypotheses_are_the_dumps(ex, slab):
    say("case: slab hypotheses on the device == hyp.pa / hyp.pb, poses bit-exact vs oracle")
    p = slab["p"]
    pa, pb, poses = ex.hypotheses(slab["da"], slab["db"], p["hyp_ia"], p["hyp_ib"])
    assert len(pa) == len(p["hyp_pa"]) == 16337
    assert np.array_equal(pa, p["hyp_pa"]) and np.array_equal(pb, p["hyp_pb"])
    want = G.poses(slab["fa"], slab["fb"], p["hyp_ia"], p["hyp_ib"], p["hyp_pa"], p["hyp_pb"])
    assert poses.tobytes() == want.tobytes()
    # Empty subsets give no hypothesis.
    assert len(ex.hypotheses(slab["da"], slab["db"], [], p["hyp_ib"])[0]) == 0


def test_suppression_on_the_dumps_own_walks(ex, slab):
    say("case: device NMS over nms1.order / nms2.order == nms1.kept / nms2.kept")
    p, st = slab["p"], slab["st"]
    sc = p["scales"]
    hyp = G.poses(slab["fa"], slab["fb"], p["hyp_ia"], p["hyp_ib"], p["hyp_pa"], p["hyp_pb"])
    k1 = ex.nms(st["nms1_order"], hyp, p["coarse_cs"], sc[10], 250, G.COARSE_FLOOR)
    assert np.array_equal(k1, p["nms1_kept"])
    k2 = ex.nms(st["nms2_order"], G.pose12_of(p["s1_T"]), p["s1_score"], sc[10], 10, G.STAGE1_FLOOR)
    assert np.array_equal(k2, p["nms2_kept"])


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_suppression_against_the_oracle_on_random_walks(ex, seed):
    say(f"case: device NMS vs oracle, random clusters, seed {seed}")
    rng = np.random.default_rng(seed)
    n = 700
    centers = rng.normal(scale=3.0, size=(12, 3))
    P = []
    for k in range(n):
        c = centers[rng.integers(12)] + rng.normal(scale=0.3, size=3)
        ax = rng.normal(size=3)
        ax /= np.linalg.norm(ax)
        ang = rng.uniform(0, np.radians(40))
        K = np.array([[0, -ax[2], ax[1]], [ax[2], 0, -ax[0]], [-ax[1], ax[0], 0]])
        R = np.eye(3) + np.sin(ang) * K + (1 - np.cos(ang)) * K @ K
        P.append(np.concatenate([R.reshape(9), c]))
    P = np.array(P)
    score = rng.integers(0, 60, size=n) / 60.0
    order = G.order_by_score(score)
    for tol, topk, floor in [(1.0, 250, 0.1), (0.5, 10, 0.05), (2.0, 0, 0.0), (1.0, 1000, 0.0), (1.0, 5, 2.0)]:
        want = G.nms(order, P, score, tol, topk, floor)
        got = ex.nms(order, P, score, tol, topk, floor)
        assert list(got) == want, (tol, topk, floor)


def test_stage1_hand_off_on_the_slab_pair(ex, slab):
    say("case: batched stage 1 on the slab pair (the dump's subsets and probe) vs the oracle and the dump")
    a, b, p = slab["a"], slab["b"], slab["p"]
    sc = p["scales"]
    idx = p["coarse_idx"]
    src, tgt = ex.cloud(b["brk_P"][b["brk_sub"]]), ex.cloud(a["brk_P"], a["brk_ns"])
    pair = dict(a=slab["da"], b=slab["db"], ia=p["hyp_ia"], ib=p["hyp_ib"], probe_p=b["brk_P"][idx],
                probe_n=b["brk_ns"][idx], s1_source=src, s1_target=tgt, scales=sc)
    r = ex.match_stage1([pair, pair])
    for got in r:
        assert got["n_hyp"] == 16337
        # R §5.3 with this port's walk (ties by index) over the device's bit-exact coarse scores.
        hyp = G.poses(slab["fa"], slab["fb"], p["hyp_ia"], p["hyp_ib"], p["hyp_pa"], p["hyp_pb"])
        want1 = G.nms(G.order_by_score(p["coarse_cs"], G.COARSE_ORDER_LIMIT), hyp, p["coarse_cs"], sc[10], 250,
                      G.COARSE_FLOOR)
        assert list(got["kept1"]) == want1
        # The stage-1 ladder from those hypotheses, vs the oracle's.
        T = G.s1_ladder(G.homogeneous(hyp[want1]), b["brk_P"][b["brk_sub"]], a["brk_P"], sc)
        ang, tr = pose_gap(got["s1_T"], T, sc[0])
        say(f"  stage-1 vs oracle: worst {ang.max():.2e} deg {tr.max():.2e} t over {len(T)}")
        assert np.mean((ang < 0.05) & (tr < 0.01)) > 0.99
        # The re-score of the device's own transforms, and the second suppression over it.
        s1, _ = O.coarse_scores(a["brk_P"], a["brk_ns"], b["brk_P"][b["brk_sub"]], b["brk_ns"][b["brk_sub"]],
                                G.pose12_of(got["s1_T"]), sc[3])
        assert np.array_equal(got["s1_score"], s1)
        want2 = G.nms(G.order_by_score(s1), G.pose12_of(got["s1_T"]), s1, sc[10], 10, G.STAGE1_FLOOR)
        assert list(got["kept2"]) == want2
        assert got["best1"] == s1.max()
        # Where the kept hypotheses are the dump's, so are the stage-1 poses (the reference's
        # ladder tolerance, stage1.rs:55-63).
        common = [k for k, h in enumerate(got["kept1"]) if h in set(p["nms1_kept"])]
        pos = {h: i for i, h in enumerate(p["nms1_kept"])}
        ang, tr = pose_gap(got["s1_T"][common], p["s1_T"][[pos[got["kept1"][k]] for k in common]], sc[0])
        say(f"  {len(common)} kept hypotheses shared with the dump; stage-1 vs s1.T: worst {ang.max():.2e} deg")
        assert np.mean((ang < 0.05) & (tr < 0.01)) > 0.95
    assert np.array_equal(r[0]["kept2"], r[1]["kept2"]) and np.array_equal(r[0]["s1_T"], r[1]["s1_T"])
    src.close()
    tgt.close()
