"""Batched ICP on the device: rungs and ladders over many cloud pairs (ladder.rs:105-128).

A batch's candidates must come out exactly as the per-call Executor::icp_rung gives them one pair
at a time, whatever the workgroup size and however the batch is composed; the ladder must equal its
rungs chained by hand, replay identically from its captured graph, and reach the reference's dumped
stage-1 and stage-2 poses.
"""

import numpy as np
import pytest

import icp_chaos
import oracle_lib as O
import slab_data as S
from test_gpu_match import pose_gap
from test_oracle_known import patch, rigid

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def ex():
    import sherd_hip

    e = sherd_hip.HipExecutor(0)
    yield e
    e.close()


def synthetic_pair(n_side, seed, normals=True):
    """A wavy patch as target, a rigidly displaced and thinned copy as source."""
    rng = np.random.default_rng(seed)
    pts, nrm = patch(n_side, 10.0 / n_side, -5.0)
    truth = rigid(rng.normal(size=3), 0.05, rng.normal(scale=0.1, size=3))
    inv = np.linalg.inv(truth)
    keep = rng.random(len(pts)) < 0.7
    src = (inv[:3, :3] @ pts[keep].T).T + inv[:3, 3] + rng.normal(scale=1e-3, size=(keep.sum(), 3))
    return src, pts, (nrm if normals else None)


def perturbed(k, seed, ang=0.05, tr=0.1):
    rng = np.random.default_rng(seed)
    return np.stack([rigid(rng.normal(size=3), ang * rng.random(), rng.normal(scale=tr, size=3)) for _ in range(k)])


def stage1_setup():
    a, b, p = S.piece("pieceA"), S.piece("pieceB"), S.pair()
    poses = S.hypothesis_poses(a, b, p["hyp_ia"], p["hyp_ib"], p["hyp_pa"], p["hyp_pb"])
    inits = np.tile(np.eye(4), (len(p["nms1_kept"]), 1, 1))
    inits[:, :3, :3] = poses[p["nms1_kept"], :9].reshape(-1, 3, 3)
    inits[:, :3, 3] = poses[p["nms1_kept"], 9:]
    return a, b, p, inits


@pytest.mark.parametrize("estimation", ["point_to_point", "point_to_plane"])
def test_batch_rung_equals_the_per_call_rung(ex, estimation):
    say(f"case: batch rung == per-call rung ({estimation}), 4 pairs of 300-9000 points")
    specs = [(14, 1), (40, 2), (75, 3), (110, 4)]
    raw = [synthetic_pair(n, s) for n, s in specs]
    clouds = [(ex.cloud(src), ex.cloud(tp, tn), 0.6) for src, tp, tn in raw]
    cand_pair = np.array([k % 4 for k in range(24)], dtype=np.uint32)
    inits = perturbed(24, 7)
    got = ex.icp_batch_rung(clouds, cand_pair, inits, estimation, 25)
    for p, (src, tp, tn) in enumerate(raw):
        sel = cand_pair == p
        want = ex.icp_rung(src, tp, tn, inits[sel], estimation, 0.6, 25)
        assert np.array_equal(got[sel], want), p
    # Against the oracle (sequential sums, so rounding differs): the transform bar of 1e-6 rad and
    # 1e-5 t on every candidate whose rung is a function of its input (the one-ULP probe of
    # crosscheck.rs:173-215), converged or not; the others are counted and capped.
    est = {"point_to_point": 0, "point_to_plane": 1}[estimation]
    for p, (src, tp, tn) in enumerate(raw):
        sel = cand_pair == p
        det, To = icp_chaos.determined([(src, tp, tn, est, 0.6, 25)], inits[sel])
        ang, tr = pose_gap(got[sel][:, :16].reshape(-1, 4, 4), To, 1.0)
        say(f"  pair {p}: {(~det).sum()} of {len(det)} chaotic; worst over determined {ang[det].max():.2e} deg "
            f"{tr[det].max():.2e} t")
        assert (~det).sum() <= 1
        assert np.radians(ang[det]).max() < 1e-6 and tr[det].max() < 1e-5


def test_workgroup_sizes_give_identical_results(ex):
    say("case: ICP workgroup sizes 64/128/256/512 identical")
    raw = [synthetic_pair(n, s) for n, s in [(30, 11), (90, 12)]]
    clouds = [(ex.cloud(src), ex.cloud(tp, tn), 0.5) for src, tp, tn in raw]
    cand_pair = np.array([0, 1, 1, 0, 1], dtype=np.uint32)
    inits = perturbed(5, 13)
    outs = []
    try:
        for b in (64, 128, 256, 512):
            ex.icp_block(b)
            outs.append(ex.icp_batch_rung(clouds, cand_pair, inits, "point_to_plane", 30))
    finally:
        ex.icp_block(0)
    for o in outs[1:]:
        assert np.array_equal(o, outs[0])


def test_wide_form_chains_equal_the_per_call_rung(ex):
    say("case: wide form split into chains (72 candidates x 31 spans) == one candidate per call, and on replay")
    # Over 2048 spans of 512 points the wide form runs the queue's halves as two chains on two
    # streams; each candidate's arithmetic must not change (one candidate per call never splits).
    raw = [synthetic_pair(150, 21), synthetic_pair(150, 22)]
    assert min(len(r[0]) for r in raw) > 30 * 512
    clouds = [(ex.cloud(src), ex.cloud(tp, tn), 0.5) for src, tp, tn in raw]
    cand_pair = np.array([k % 2 for k in range(72)], dtype=np.uint32)
    inits = perturbed(72, 23)
    try:
        ex.icp_block(1)
        outs = [ex.icp_batch_rung(clouds, cand_pair, inits, "point_to_point", 20) for _ in range(3)]
        for p, (src, tp, tn) in enumerate(raw):
            sel = np.nonzero(cand_pair == p)[0]
            for i in sel[:6]:
                want = ex.icp_rung(src, tp, tn, inits[i:i + 1], "point_to_point", 0.5, 20)
                assert np.array_equal(outs[0][i:i + 1], want), i
    finally:
        ex.icp_block(0)
    for o in outs[1:]:
        assert np.array_equal(o, outs[0])


def test_stage1_ladder_matches_chained_rungs_and_the_golden(ex):
    say("case: stage-1 ladder, 250 candidates x 2 rungs, graph replay")
    a, b, p, inits = stage1_setup()
    sc = p["scales"]
    src, tgt = ex.cloud(b["brk_P"][b["brk_sub"]]), ex.cloud(a["brk_P"], a["brk_ns"])
    rungs = [("point_to_point", 20), ("point_to_point", 20)]
    pairs = [[(src, tgt, k * sc[0] * sc[11])] for k in (0.2, 0.08)]
    cp = np.zeros(len(inits), dtype=np.uint32)
    first = ex.icp_batch_ladder(pairs, rungs, cp, inits)
    # Eager (tuning), eager, then captured on the repeat, then replayed.
    for _ in range(3):
        again = ex.icp_batch_ladder(pairs, rungs, cp, inits)
        assert np.array_equal(first, again)
    T = inits
    for r, k in enumerate((0.2, 0.08)):
        step = ex.icp_rung(b["brk_P"][b["brk_sub"]], a["brk_P"], a["brk_ns"], T, 0, k * sc[0] * sc[11], 20)
        assert np.array_equal(first[:, r], step), r
        T = step[:, :16].reshape(-1, 4, 4)
    ang, tr = pose_gap(first[:, 1, :16].reshape(-1, 4, 4), p["s1_T"], sc[0])
    assert np.mean((ang < 0.05) & (tr < 0.01)) > 0.95


def test_stage2_ladder_with_its_own_clouds_per_rung(ex):
    say("case: stage-2 ladder, 10 candidates x 4 rungs over pc_reg then pc_frac")
    from test_oracle_slab import stage2_rungs

    p = S.pair()
    sc = p["scales"]
    T0 = p["s1_T"][p["nms2_kept"]]
    steps = stage2_rungs()
    keep = []
    pairs = []
    for name, src, tp, tn, k in steps:
        s, t = ex.cloud(src), ex.cloud(tp, tn)
        keep += [s, t]
        pairs.append([(s, t, k * sc[0] * sc[11])])
    out = ex.icp_batch_ladder(pairs, [("point_to_plane", 30)] * 4, np.zeros(len(T0), np.uint32), T0)
    T = T0
    for r, (name, src, tp, tn, k) in enumerate(steps):
        step = ex.icp_rung(src, tp, tn, T, 1, k * sc[0] * sc[11], 30)
        assert np.array_equal(out[:, r], step), name
        T = step[:, :16].reshape(-1, 4, 4)
        # Against the reference's dump, its own tolerance (sherd-parity/src/stages/stage2.rs:42-48).
        ang, tr = pose_gap(T, p[name], sc[0])
        assert ang.max() < 0.05 and tr.max() < 0.01, name
    # Against the oracle's ladder: the transform bar on every determined candidate.
    det, To = icp_chaos.determined([(src, tp, tn, 1, k * sc[0] * sc[11], 30) for _, src, tp, tn, k in steps], T0)
    ang, tr = pose_gap(T, To, sc[0])
    say(f"  vs oracle: {(~det).sum()} of {len(det)} chaotic; worst over determined {ang[det].max():.2e} deg "
        f"{tr[det].max():.2e} t")
    assert (~det).sum() <= 1 and np.radians(ang[det]).max() < 1e-6 and tr[det].max() < 1e-5


def test_many_pairs_in_one_ladder_equal_their_own_ladders(ex):
    say("case: 6 pairs x 8 candidates in one ladder == each pair alone")
    raw = [synthetic_pair(n, 20 + i) for i, n in enumerate([12, 25, 50, 70, 20, 100])]
    clouds = [(ex.cloud(src), ex.cloud(tp, tn)) for src, tp, tn in raw]
    rungs = [("point_to_point", 15), ("point_to_plane", 15), ("point_to_plane", 15)]
    dists = [1.0, 0.3, 0.1]
    pairs = [[(s, t, d) for s, t in clouds] for d in dists]
    cp = np.repeat(np.arange(6, dtype=np.uint32), 8)
    inits = perturbed(48, 21)
    all_ = ex.icp_batch_ladder(pairs, rungs, cp, inits)
    for p in range(6):
        sel = cp == p
        alone = ex.icp_batch_ladder([[pr[p]] for pr in pairs], rungs, np.zeros(sel.sum(), np.uint32), inits[sel])
        assert np.array_equal(all_[sel], alone), p


def test_empty_clouds_keep_the_pose(ex):
    say("case: ICP with an empty cloud")
    src, tp, tn = synthetic_pair(20, 30)
    empty = ex.cloud(np.zeros((0, 3)))
    full_s, full_t = ex.cloud(src), ex.cloud(tp, tn)
    inits = perturbed(3, 31)
    pairs = [[(empty, full_t, 0.5), (full_s, empty, 0.5), (full_s, full_t, 0.5)]] * 2
    out = ex.icp_batch_ladder(pairs, [("point_to_point", 10)] * 2, np.array([0, 1, 2], np.uint32), inits)
    for i in (0, 1):
        for r in (0, 1):
            assert np.array_equal(out[i, r, :16], inits[i].reshape(16)) and out[i, r, 16] == 0.0
    assert out[2, 1, 16] > 0.5
