"""R §9's full-resolution refinement on the device (crates/sherd-core/src/refine.rs), and the wide
ICP form it registers with.

The wide form (every span of a candidate its own workgroup) must give the batched form's
registrations to the bit; the fracture cloud (selection, cap, vertex normals) and the placement are
integer and fused-f64 work and must equal the oracle's bit for bit; the walk must take the
reference's joins in the reference's order and land on the slab dump's refined poses.
"""

import os
import sys

import numpy as np
import pytest

import oracle_lib as O
from test_gpu_icp_batch import perturbed, synthetic_pair
from test_gpu_match import pose_gap

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import refine as R  # noqa: E402

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "slab")
PIECES = ("pieceA", "pieceB")


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def ex():
    import sherd_hip

    e = sherd_hip.HipExecutor(0)
    yield e
    e.icp_block(0)
    e.close()


@pytest.fixture(scope="module")
def slab():
    z = np.load(os.path.join(GOLD, "refine.npz"))
    frags = [np.load(os.path.join(GOLD, p + ".npz")) for p in PIECES]
    return z, frags


@pytest.mark.parametrize("estimation", ["point_to_point", "point_to_plane"])
def test_wide_form_equals_the_batched_form(ex, estimation):
    say(f"case: wide ICP == batched ICP bit for bit ({estimation}), ladder of 3 rungs, 5 pairs incl. one empty")
    specs = [(14, 1), (40, 2), (75, 3), (130, 4)]
    raw = [synthetic_pair(n, s) for n, s in specs]
    clouds = [(ex.cloud(src), ex.cloud(tp, tn)) for src, tp, tn in raw]
    empty = ex.cloud(np.zeros((0, 3)))
    rungs = [(0.9, 30), (0.5, 30), (0.2, 30)]
    pairs = [[(s, t, d) for s, t in clouds] + [(empty, clouds[0][1], d)] for d, _ in rungs]
    specs_r = [(estimation, it) for _, it in rungs]
    cand_pair = np.array([k % 5 for k in range(23)], dtype=np.uint32)
    inits = perturbed(23, 11)
    out = {}
    for mode in (256, 1, 64):
        ex.icp_block(mode)
        out[mode] = ex.icp_batch_ladder(pairs, specs_r, cand_pair, inits)
        out[mode, "captured"] = ex.icp_batch_ladder(pairs, specs_r, cand_pair, inits)  # captured on the repeat
        out[mode, "replayed"] = ex.icp_batch_ladder(pairs, specs_r, cand_pair, inits)
    ex.icp_block(0)
    for key, v in out.items():
        assert np.array_equal(v, out[256]), key
    assert np.array_equal(out[256][cand_pair == 4][:, :, :16].reshape(-1, 16),
                          np.repeat(inits[cand_pair == 4].reshape(-1, 16), 3, axis=0))


def test_wide_form_on_one_large_cloud(ex):
    say("case: one candidate, 60 000-point source: wide == batched, both near the oracle")
    src, tp, tn = synthetic_pair(300, 5)
    s, t = ex.cloud(src), ex.cloud(tp, tn)
    init = perturbed(1, 12)
    ex.icp_block(1)
    wide = ex.icp_batch_rung([(s, t, 0.5)], np.zeros(1, np.uint32), init, "point_to_plane", 40)
    ex.icp_block(512)
    batched = ex.icp_batch_rung([(s, t, 0.5)], np.zeros(1, np.uint32), init, "point_to_plane", 40)
    ex.icp_block(0)
    assert np.array_equal(wide, batched)
    o = O.icp_rung(src, tp, tn, init, "point_to_plane", 0.5, 40)
    ang, tr = pose_gap(wide[:, :16].reshape(-1, 4, 4), o[:, :16].reshape(-1, 4, 4), 1.0)
    say(f"  vs oracle: {ang.max():.2e} deg {tr.max():.2e}; iterations {wide[0, 19]:.0f} / {o[0, 19]:.0f}")
    assert ang.max() < 1e-4 and tr.max() < 1e-6


def test_fracture_cloud_is_the_oracles_and_the_dumps(ex, slab):
    say("case: slab fracture clouds: selection == refine/<piece>.idx, points and normals == oracle")
    z, frags = slab
    for i, (p, f) in enumerate(zip(PIECES, frags)):
        from sherd_hip import face_centroids

        C = face_centroids(f["V"], f["F"])
        assert C.tobytes() == R.face_centroids(f["V"], f["F"]).tobytes()
        cl = ex.fracture_cloud(f["V0"], f["F"], C, f["frac"], z["thick"][i], z["res"][i])
        pts, nrm, idx = cl.read(idx=True)
        assert np.array_equal(idx, z["idx_" + p]), p
        oi, op, on = R.fracture_cloud(f["V0"], f["F"], C, f["frac"], z["thick"][i], z["res"][i])
        assert pts.tobytes() == op.tobytes() and nrm.tobytes() == on.tobytes(), p
        cl.close()


@pytest.mark.parametrize("cap,seed", [(1000, 0), (4752, 5)])
def test_capped_selection_is_the_oracles_draw(ex, slab, cap, seed):
    say(f"case: fracture cloud capped at {cap} (seed {seed}): the ChaCha8 draw == oracle")
    z, frags = slab
    f = frags[1]
    C = R.face_centroids(f["V"], f["F"])
    cl = ex.fracture_cloud(f["V0"], f["F"], C, f["frac"], z["thick"][1], z["res"][1], max_points=cap, seed=seed)
    oi, op, on = R.fracture_cloud(f["V0"], f["F"], C, f["frac"], z["thick"][1], z["res"][1], seed=seed,
                                  max_points=cap)
    pts, nrm, idx = cl.read(idx=True)
    assert len(idx) == cap and np.array_equal(idx, oi)
    assert pts.tobytes() == op.tobytes() and nrm.tobytes() == on.tobytes()


def test_cloud_from_indices_and_placement_are_bit_exact(ex, slab):
    say("case: cloud_from_indices + placed: fused placement == oracle bit for bit")
    z, frags = slab
    f = frags[0]
    idx = z["idx_pieceA"][::-3]  # an unsorted selection keeps its order
    cl = ex.cloud_from_indices(f["V0"], f["F"], idx)
    pts, nrm, ix = cl.read(idx=True)
    assert np.array_equal(ix, idx)
    on = R.vertex_normals(f["V0"], f["F"])[idx]
    assert pts.tobytes() == f["V0"][idx].tobytes() and nrm.tobytes() == on.tobytes()
    T = z["poses_before"][1]
    pl = cl.placed(T)
    pp, pn, pi = pl.read(idx=True)
    wp, wn = R.placed(f["V0"][idx], on, T)
    assert pp.tobytes() == wp.tobytes() and pn.tobytes() == wn.tobytes() and np.array_equal(pi, idx)
    pl.close()
    cl.close()


def slab_pieces(ex, z, frags):
    clouds = [ex.cloud_from_indices(f["V0"], f["F"], z["idx_" + p]) for p, f in zip(PIECES, frags)]
    return [(z["thick"][i], z["res"][i], clouds[i]) for i in range(2)], clouds


def test_slab_refinement_on_the_device(ex, slab):
    say("case: slab R §9 on the device vs refine/joins.json and poses_final.json, and vs the oracle")
    z, frags = slab
    pieces, clouds = slab_pieces(ex, z, frags)
    groups, used = [list(z["group_members"])], [tuple(u) for u in z["used"]]
    out, joins = ex.refine_joins(pieces, z["poses_before"], groups, used)
    assert len(joins) == 1
    j = joins[0]
    assert (j["fixed"], j["moving"]) == (z["join_fixed"][0], z["join_moving"][0])
    assert j["dist"].tobytes() == z["join_dist"][0].tobytes()
    t = min(z["thick"])
    for r in range(2):
        ang, tr = pose_gap(j["rungs"][r][None], z["join_T"][0, r][None], t)
        say(f"  rung {r + 1} vs dump: {ang[0]:.2e} deg {tr[0]:.2e} t")
        assert ang[0] < 0.2 and tr[0] < 0.02
    assert abs(j["fitness"] - z["join_fitness"][0]) < 1e-4
    assert j["rmse_t"] < 1e-12
    ang, tr = pose_gap(out, z["poses_final"], t)
    assert ang.max() < 0.2 and tr.max() < 0.02
    # The oracle from the same clouds: the same fixed point, to the rounding of the sums.
    opieces = []
    for i, (p, f) in enumerate(zip(PIECES, frags)):
        idx = z["idx_" + p]
        opieces.append((z["thick"][i], z["res"][i], (f["V0"][idx], R.vertex_normals(f["V0"], f["F"])[idx])))
    oout, ojoins = R.refine_joins(opieces, z["poses_before"], groups, used)
    ang, tr = pose_gap(out, oout, t)
    say(f"  vs oracle: {ang.max():.2e} deg {tr.max():.2e} t; fitness {j['fitness']} / {ojoins[0]['fitness']}")
    assert ang.max() < 1e-6 and tr.max() < 1e-8
    assert abs(j["fitness"] - ojoins[0]["fitness"]) < 1e-4
    # Both forms of the ICP give the same refinement.
    ex.icp_block(512)
    out2, joins2 = ex.refine_joins(pieces, z["poses_before"], groups, used)
    ex.icp_block(0)
    assert np.array_equal(out2, out) and np.array_equal(joins2[0]["rungs"], j["rungs"])
    for c in clouds:
        c.close()


def test_the_walk_on_the_device(ex):
    say("case: the walk (refine.rs:423-470 and 474-497): order, pull-back, skipped join, empty group")
    pts, nrm = [], []
    for i in range(20):
        for j in range(20):
            x, y = i * 0.1, j * 0.1
            g = np.array([-(0.8 * x + 0.1 * y), -(0.1 * x - 0.5 * y), 1.0])
            pts.append([x, y, 0.4 * x * x - 0.25 * y * y + 0.1 * x * y])
            nrm.append(g / np.linalg.norm(g))
    V = np.array(pts)
    cl = ex.cloud(V, np.array(nrm))
    pieces = [(1.0, 0.05, cl)] * 3
    poses = np.stack([np.eye(4)] * 3)
    poses[1][2, 3] = 0.004
    poses[2][2, 3] = 0.006
    out, joins = ex.refine_joins(pieces, poses, [[0, 1, 2], []], [(1, 2), (0, 1)])
    assert [(j["fixed"], j["moving"]) for j in joins] == [(0, 1), (1, 2)]
    for n in (1, 2):
        assert abs(out[n][2, 3]) < 1e-9, out[n]
    out, joins = ex.refine_joins([(1.0, 0.1, cl), (1.0, 0.1, None)], poses[:2], [[0, 1]], [(0, 1)])
    assert joins == [] and np.array_equal(out, poses[:2])
    cl.close()
