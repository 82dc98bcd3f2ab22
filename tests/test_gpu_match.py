"""HIP nearest neighbours, coarse scores and ICP rungs against the oracle and the slab golden
vectors (GPU).

Bit-exact rows: neighbour indices and distances (the f64 kiddo arithmetic), coarse agree counts
and scores (also against the reference's own coarse.cs.npy). Tolerance rows: ICP poses, whose f64
sums run in a different (fixed) order on the device than the CPU's sequential source order.
"""

import sys

import numpy as np
import pytest

import icp_chaos
import oracle_lib as O
import slab_data as S

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ex():
    import sherd_hip

    e = sherd_hip.HipExecutor(0)
    yield e
    e.close()


@pytest.mark.parametrize("seed", [0, 1])
def test_nearest_below_matches_the_oracle(ex, seed):
    say(f"case: nearest_below seed={seed}")
    rng = np.random.default_rng(seed)
    pts = rng.uniform(-4, 4, size=(3000, 3))
    pts = np.concatenate([pts, pts * [-1, 1, 1]])  # mirrored: exact ties on x = 0
    q = rng.uniform(-5, 5, size=(20000, 3))
    q[:5000, 0] = 0.0
    for bound in (0.05, 0.3, 1.0, 3.0):
        gi, gd = ex.nearest_below(pts, q, bound)
        oi, od = O.nearest_below(pts, q, bound)
        assert np.array_equal(gi, oi) and np.array_equal(gd, od), bound


def test_nearest_below_boundary_and_known_answers(ex):
    say("case: nearest_below known answers")
    grid = np.array([[i, j, 0.0] for i in range(5) for j in range(5)])
    idx, d = ex.nearest_below(grid, [[0.1, 0.2, 0], [2.9, 3.1, 0], [100, 100, 0]], 2.0)
    assert list(idx) == [0, 18, -1] and abs(d[0] - np.hypot(0.2, 0.1)) < 1e-15
    one = np.array([[7.0, 11.0, 13.0]])
    dd = np.sqrt(339.0)
    assert ex.nearest_below(one, [[0, 0, 0]], dd)[0][0] == -1
    assert ex.nearest_below(one, [[0, 0, 0]], np.nextafter(dd, np.inf))[0][0] == 0


def test_coarse_scores_bit_exact_against_the_golden(ex):
    say("case: slab coarse 16 337 hypotheses")
    ci = S.coarse_inputs()
    s, c = ex.coarse_scores(ci["target_p"], ci["target_n"], ci["probe_p"], ci["probe_n"], ci["poses"],
                            ci["radius"])
    assert np.array_equal(s.view(np.uint64), ci["expected"].view(np.uint64))
    os_, oc = O.coarse_scores(ci["target_p"], ci["target_n"], ci["probe_p"], ci["probe_n"], ci["poses"],
                              ci["radius"])
    assert np.array_equal(c, oc)


def test_stage1_rescore_bit_exact(ex):
    say("case: stage-1 re-score")
    a, b, p = S.piece("pieceA"), S.piece("pieceB"), S.pair()
    s, _ = ex.coarse_scores(a["brk_P"], a["brk_ns"], b["brk_P"][b["brk_sub"]], b["brk_ns"][b["brk_sub"]],
                            p["s1_T"], p["scales"][3])
    assert np.array_equal(s, p["s1_score"])


def test_coarse_edge_cases(ex):
    say("case: coarse edge cases")
    line = np.array([[k, 0.0, 0.0] for k in range(4)])
    up = np.tile([0.0, 0, 1], (4, 1))
    ident = np.concatenate([np.eye(3).reshape(9), np.zeros(3)])[None]
    tilt = lambda c: [[np.sqrt(1 - c * c), 0.0, c]]
    assert ex.coarse_scores(line, up, [[0, 0, 0]], tilt(0.71), ident, 0.5)[0][0] == 1.0
    assert ex.coarse_scores(line, up, [[0, 0, 0]], tilt(0.7), ident, 0.5)[0][0] == 0.0
    assert ex.coarse_scores(line[:1], up[:1], [[0, 0.5, 0]], [[0, 0, 1]], ident, 0.5)[0][0] == 0.0
    assert ex.coarse_scores(line, up, np.zeros((0, 3)), np.zeros((0, 3)), ident, 0.5)[0][0] == 0.0


def pose_gap(a, b, t):
    """Rotation gap in degrees (from |Ra Rb^T - I|_F / sqrt 2, exact to rounding for small angles,
    where an arccos of the trace bottoms out near 5e-6 degrees) and translation gap in t."""
    dr = a[:, :3, :3] @ np.transpose(b[:, :3, :3], (0, 2, 1)) - np.eye(3)
    ang = 2.0 * np.arcsin(np.clip(np.linalg.norm(dr, axis=(1, 2)) / (2.0 * np.sqrt(2.0)), 0, 1))
    return np.degrees(ang), np.linalg.norm(a[:, :3, 3] - b[:, :3, 3], axis=1) / t


def test_stage1_ladder_against_the_oracle_and_the_golden(ex):
    say("case: stage-1 ladder, 250 candidates x 2 rungs")
    a, b, p = S.piece("pieceA"), S.piece("pieceB"), S.pair()
    sc = p["scales"]
    poses = S.hypothesis_poses(a, b, p["hyp_ia"], p["hyp_ib"], p["hyp_pa"], p["hyp_pb"])
    inits = np.tile(np.eye(4), (len(p["nms1_kept"]), 1, 1))
    inits[:, :3, :3] = poses[p["nms1_kept"], :9].reshape(-1, 3, 3)
    inits[:, :3, 3] = poses[p["nms1_kept"], 9:]
    src = b["brk_P"][b["brk_sub"]]
    Tg = inits
    for k in (0.2, 0.08):
        d = k * sc[0] * sc[11]
        Tg = ex.icp_rung(src, a["brk_P"], a["brk_ns"], Tg, 0, d, 20)[:, :16].reshape(-1, 4, 4)
    # The north star's transform bar, 1e-6 rad and 1e-5 t, on every candidate whose ladder is a
    # function of its input (the one-ULP probe of crosscheck.rs:173-215); the others are counted.
    rungs = [(src, a["brk_P"], a["brk_ns"], 0, k * sc[0] * sc[11], 20) for k in (0.2, 0.08)]
    det, To = icp_chaos.determined(rungs, inits)
    ang, tr = pose_gap(Tg, To, sc[0])
    say(f"  stage 1 vs oracle: {(~det).sum()} of {len(det)} chaotic (excused); worst over determined "
        f"{ang[det].max():.2e} deg {tr[det].max():.2e} t, over all {ang.max():.2e} deg {tr.max():.2e} t")
    assert (~det).sum() <= len(det) // 20
    assert np.radians(ang[det]).max() < 1e-6 and tr[det].max() < 1e-5
    ang2, tr2 = pose_gap(Tg, p["s1_T"], sc[0])
    # Against the reference's own dump (Open3D's sums): its tolerance, 0.05 degrees / 0.01 t
    # (sherd-parity/src/stages/stage1.rs:55-63).
    assert np.mean((ang2 < 0.05) & (tr2 < 0.01)) > 0.95


def test_stage2_ladder_against_the_golden(ex):
    say("case: stage-2 ladder, 10 candidates x 4 rungs")
    from test_oracle_slab import stage2_rungs

    p = S.pair()
    sc = p["scales"]
    T0 = T = p["s1_T"][p["nms2_kept"]]
    rungs = []
    for name, src, tp, tn, k in stage2_rungs():
        d = k * sc[0] * sc[11]
        T = ex.icp_rung(src, tp, tn, T, 1, d, 30)[:, :16].reshape(-1, 4, 4)
        # Against the reference's dump: its tolerance (sherd-parity/src/stages/stage2.rs:42-48).
        ang, tr = pose_gap(T, p[name], sc[0])
        assert ang.max() < 0.05 and tr.max() < 0.01, name
        # Against the oracle: the transform bar on every determined candidate after this rung.
        rungs.append((src, tp, tn, 1, d, 30))
        det, To = icp_chaos.determined(rungs, T0)
        ang, tr = pose_gap(T, To, sc[0])
        say(f"  {name} vs oracle: {(~det).sum()} of {len(det)} chaotic (excused); worst over determined "
            f"{ang[det].max():.2e} deg {tr[det].max():.2e} t")
        assert (~det).sum() <= 1
        assert np.radians(ang[det]).max() < 1e-6 and tr[det].max() < 1e-5, name


def test_icp_known_answers(ex):
    say("case: ICP known answers")
    from test_oracle_known import patch, rigid

    pts, _ = patch(20, 0.5, -5.0)
    truth = rigid([0.3, -0.7, 0.6], 0.04, [0.11, -0.07, 0.05])
    inv = np.linalg.inv(truth)
    src = (inv[:3, :3] @ pts.T).T + inv[:3, 3]
    out = ex.icp_rung(src, pts, None, np.eye(4)[None], "point_to_point", 1.0, 20)[0]
    T = out[:16].reshape(4, 4)
    assert np.abs((T[:3, :3] @ src.T).T + T[:3, 3] - pts).max() < 1e-9 and out[20] == 1.0
    plane = np.array([[-5.0 + i * 0.25, -5.0 + j * 0.25, 0.0] for i in range(40) for j in range(40)])
    nrm = np.tile([0.0, 0.0, 1.0], (len(plane), 1))
    place = rigid([1.0, 0.0, 0.0], 0.05, [0.0, 0.0, 0.35])
    src = (place[:3, :3] @ (plane + [0.13, -0.07, 0.0]).T).T + place[:3, 3]
    out = ex.icp_rung(src, plane, nrm, np.eye(4)[None], "point_to_plane", 2.0, 30)[0]
    T = out[:16].reshape(4, 4)
    assert np.abs(((T[:3, :3] @ src.T).T + T[:3, 3])[:, 2]).max() < 1e-9
    tgt = np.array([[0.0, 0, 0], [100.0, 0, 0]])
    fit = lambda x: ex.icp_rung([[x, 0, 0]], tgt, None, np.eye(4)[None], 0, 1.3, 0)[0][16]
    assert fit(np.nextafter(1.3, 0)) == 1.0 and fit(1.3) == 0.0
