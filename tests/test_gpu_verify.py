"""HIP R §6 verification against the oracle (bit for bit) and the reference's slab dump (its own
tolerances, `accepted` identical) (GPU).

Cases: the slab's ten stage-2 candidates one at a time and as one batch, the same candidates
replicated over several pairs in one batch, the cheap half (full = 0), the FP64 retest, and the
reference's wall/box known answers (verify.rs:889-1120).
"""

import sys

import numpy as np
import pytest

import oracle_lib as O
import slab_data as S

pytestmark = pytest.mark.gpu

TOL = dict(tight=0.01, gap=0.002, seam=0.34, cont=0.005, cont_n=0.01, pen=0.0005)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ex():
    import sherd_hip

    e = sherd_hip.HipExecutor(0)
    yield e
    e.close()


def sides(ex):
    dev, ora, keep = [], [], []
    for name in ("pieceA", "pieceB"):
        pc = S.piece(name)
        v32 = pc["V"].astype(np.float32)
        fr, me = ex.scene(v32, pc["F"], pc["frac"]), ex.scene(v32, pc["F"])
        ofr, ome = O.Scene(v32, pc["F"], pc["frac"]), O.Scene(v32, pc["F"])
        keep += [fr, me, ofr, ome]
        args = (pc["watertight"], pc["frac_area"], pc["Pf"], pc["S"], pc["brk_P"], pc["brk_ns"], pc["margin_p"],
                pc["margin_n"])
        dev.append(ex.fragment(fr, me, *args))
        ora.append(O.SideData(ofr, ome, *args))
    return dev, ora, keep


@pytest.fixture(scope="module")
def slab(ex):
    return sides(ex)


def test_slab_candidates_bit_exact_and_within_the_reference(ex, slab):
    dev, ora, _ = slab
    p = S.pair()
    sc = p["scales"]
    T = p["s2_T_frac2"]
    say("case: slab verify batch of 10")
    got, acc = ex.verify_batch([(dev[0], dev[1], sc)], np.zeros(len(T)), T)
    for i in range(len(T)):
        want = O.verify(ora[0], ora[1], T[i], sc)
        assert np.array_equal(got[i].view(np.uint64), want.view(np.uint64)), (i, got[i] - want)
        assert acc[i] == O.accept(want, sc) == bool(p["s2_accepted"][i])
        theirs = p["s2_scores"][i]
        for k, tl in TOL.items():
            assert abs(got[i][O.OS_KEYS.index(k)] - theirs[k]) <= tl, (i, k)
    say("case: slab verify one candidate per call")
    for i in range(len(T)):
        one, a1 = ex.verify_batch([(dev[0], dev[1], sc)], [0], T[i:i + 1])
        assert np.array_equal(one[0].view(np.uint64), got[i].view(np.uint64)) and a1[0] == acc[i]


def test_many_pairs_in_one_batch(ex, slab):
    """The same candidates under several pairs (both orders, other scales) in one batch answer
    what each answers alone: a batch is grouping, not arithmetic."""
    dev, ora, _ = slab
    p = S.pair()
    sc = p["scales"]
    sc2 = O.scales_for_pair(28.0, 2.5)
    T = p["s2_T_frac2"]
    Tinv = np.array([O.pose_inverse(t) for t in T])
    pairs = [(dev[0], dev[1], sc), (dev[1], dev[0], sc), (dev[0], dev[1], sc2)]
    cp = np.repeat([0, 1, 2], len(T))
    cT = np.concatenate([T, Tinv, T])
    say("case: 30 candidates over 3 pairs")
    got, acc = ex.verify_batch(pairs, cp, cT)
    for i in range(len(cT)):
        a, b, s_ = [(ora[0], ora[1], sc), (ora[1], ora[0], sc), (ora[0], ora[1], sc2)][cp[i]]
        want = O.verify(a, b, cT[i], s_)
        assert np.array_equal(got[i].view(np.uint64), want.view(np.uint64)), i
        assert acc[i] == O.accept(want, s_)


@pytest.mark.parametrize("full,retest", [(False, False), (True, True)])
def test_cheap_half_and_retest(ex, slab, full, retest):
    dev, ora, _ = slab
    p = S.pair()
    sc = p["scales"]
    T = p["s2_T_frac2"]
    say(f"case: full={full} retest={retest}")
    got, acc = ex.verify_batch([(dev[0], dev[1], sc)], np.zeros(len(T)), T, full=full, retest=retest)
    for i in range(len(T)):
        want = O.verify(ora[0], ora[1], T[i], sc, full=full, retest=retest)
        assert np.array_equal(got[i].view(np.uint64), want.view(np.uint64)), i
    if not full:
        assert not acc.any()


def box_mesh(lo, hi):
    v = np.array([
        [lo[0], lo[1], lo[2]], [hi[0], lo[1], lo[2]], [hi[0], hi[1], lo[2]], [lo[0], hi[1], lo[2]],
        [lo[0], lo[1], hi[2]], [hi[0], lo[1], hi[2]], [hi[0], hi[1], hi[2]], [lo[0], hi[1], hi[2]],
    ], dtype=np.float32)
    f = np.array([[0, 2, 1], [0, 3, 2], [4, 5, 6], [4, 6, 7], [0, 1, 5], [0, 5, 4], [3, 7, 6], [3, 6, 2],
                  [0, 4, 7], [0, 7, 3], [1, 2, 6], [1, 6, 5]], dtype=np.uint32)
    return v, f


def test_known_answers_through_the_device(ex):
    """verify.rs:889-1060: walls, seams, continuity and penetration with exact answers."""
    say("case: wall/box known answers")
    wall_v = np.array([[0, 0, 0], [0, 2, 0], [0, 2, 2], [0, 0, 2]], np.float32)
    wall_f = np.array([[0, 1, 2], [0, 2, 3]], np.uint32)
    wall = ex.scene(wall_v, wall_f)
    samples = np.array([[0.0, 0.2 + i * 0.2, 0.2 + j * 0.2] for i in range(10) for j in range(10)])
    curve = np.array([[0.0, k * 0.1, 0.0] for k in range(30)])
    up = np.tile([0.0, 0.0, 1.0], (30, 1))
    grid = np.array([[i * 0.1, j * 0.1, 0.0] for i in range(5) for j in range(5)])
    gup = np.tile([0.0, 0.0, 1.0], (25, 1))
    va, fa = box_mesh([0, 0, 0], [10, 10, 10])
    vb, fb = box_mesh([20, 20, 20], [30, 30, 30])
    ma, mb = ex.scene(va, fa), ex.scene(vb, fb)
    e = np.zeros((0, 3))
    a = ex.fragment(wall, ma, True, 4.0, samples, [[4, 5, 6], [1, 2, 3]], curve, up, grid, gup)
    b = ex.fragment(wall, mb, True, 4.0, samples, [[4, 5, 6], [1, 5.5, 6.5], [50, 5, 5], [-5, 5, 5]], curve, up,
                    grid, gup)
    sc1 = O.scales_for_pair(1.0, 0.0)
    sc3 = O.scales_for_pair(3.0, 0.0)
    shifted = lambda dx: np.array([[1, 0, 0, dx], [0, 1, 0, 0], [0, 0, 1, 0], [0, 0, 0, 1.0]])
    got, _ = ex.verify_batch([(a, b, sc1), (a, b, sc3)], [0, 0, 0, 1],
                             np.stack([np.eye(4), shifted(0.005), shifted(0.5), np.eye(4)]))
    k = O.OS_KEYS.index
    assert abs(got[0][k("tight")] - 1.0) < 1e-12 and abs(got[0][k("contact")] - 4.0) < 1e-9
    assert abs(got[1][k("gap")] - 0.005) < 1e-6
    assert (got[2][k("tight")], got[2][k("gap")], got[2][k("contact")]) == (0.0, 1.0, 0.0)
    assert abs(got[3][k("seam")] - 1.0) < 1e-12
    assert abs(got[0][k("cont")]) < 1e-12 and abs(got[0][k("cont_n")] - 1.0) < 1e-12
    assert abs(got[0][k("pen")] - 0.5) < 1e-12 and abs(got[0][k("pen_depth")] - 4.0) < 1e-5
    holed = ex.fragment(wall, mb, False, 4.0, samples, [[5, 5, 5]], curve, up, e, e)
    got, _ = ex.verify_batch([(a, holed, sc1)], [0], np.eye(4)[None])
    assert (got[0][k("pen")], got[0][k("pen_unavailable")]) == (0.0, 1.0)
    assert (got[0][k("cont")], got[0][k("cont_n")]) == (1.0, -1.0)


def moved_f32(T, P):
    """T.p in f64, left to right, narrowed once to f32 (executor/cpu.rs:192-200)."""
    return np.stack([((T[r, 0] * P[:, 0] + T[r, 1] * P[:, 1]) + T[r, 2] * P[:, 2]) + T[r, 3] for r in range(3)],
                    axis=1).astype(np.float32)


def parity_rays(frags, mesh_scenes, pairs, cp, cT):
    """The unit of SHERD_M_VERIFY_RAYS on the host: 3 rays per surface sample whose moved position
    lies in the other mesh's box (RayScene::inside's reject, bvh.rs:146-149), both sides of every
    candidate whose meshes are watertight (verify.rs:490-493, 541)."""
    boxes = [sc.info()[1:] for sc in mesh_scenes]
    n = 0
    for i in range(len(cT)):
        a, b = pairs[cp[i]]
        if not (frags[a]["watertight"] and frags[b]["watertight"]):
            continue
        T = np.asarray(cT[i], np.float64).reshape(4, 4)
        for own, other, M in ((a, b, O.pose_inverse(T).reshape(4, 4)), (b, a, T)):
            q = moved_f32(M, np.asarray(frags[own]["S"], np.float64).reshape(-1, 3))
            lo, hi = boxes[other]
            n += int(np.all((q >= lo) & (q <= hi), axis=1).sum())
    return 3 * n


def test_small_vessel_all_pairs_bit_exact(ex):
    """Six fragments of a small vessel, all fifteen pairs, the true relative poses and perturbations
    up to 90 degrees and 10 t: adjacent pairs touch, others stand apart, so the closest approach runs
    through every path (the per-lane search, the whole-wave queue for long queries)."""
    from sherd_hip import synth

    say("case: small vessel, 6 fragments, 15 pairs x 10 candidates, bit-exact vs oracle")
    col = synth.vessel(t=1.0, edges_per_t=1.5, n_frag=6, seed=3, height_t=60.0, rmin_t=14.0, rmax_t=18.0)
    frags = col["fragments"]
    scales = synth.scales_for_pair(col["t"], col["res"])
    pairs, cp, cT = synth.candidates(col, per_pair=10, seed=5)
    dev, ora, keep = [], [], []
    for fr in frags:
        fs, ms = ex.scene(fr["V"], fr["F"], fr["frac"]), ex.scene(fr["V"], fr["F"])
        ofs, oms = O.Scene(fr["V"], fr["F"], fr["frac"]), O.Scene(fr["V"], fr["F"])
        keep += [fs, ms, ofs, oms]
        args = (fr["watertight"], fr["frac_area"], fr["Pf"], fr["S"], fr["brk_P"], fr["brk_ns"], fr["margin_p"],
                fr["margin_n"])
        dev.append(ex.fragment(fs, ms, *args))
        ora.append(O.SideData(ofs, oms, *args))
    for retest in (False, True):
        before = ex.stats("verify_rays")
        scores, accepted = ex.verify_batch([(dev[a], dev[b], scales) for a, b in pairs], cp, cT, retest=retest)
        after = ex.stats("verify_rays")
        want_rays = parity_rays(frags, [ms for ms in keep[1::4]], pairs, cp, cT)
        say(f"  parity rays (3 per sample past the mesh-box reject): {after['items'] - before['items']}, "
            f"host count {want_rays}")
        assert after["calls"] - before["calls"] == 1
        assert after["items"] - before["items"] == want_rays > 0
        for i in range(len(cT)):
            a, b = pairs[cp[i]]
            want = O.verify(ora[a], ora[b], cT[i], scales, retest=retest)
            assert np.array_equal(scores[i].view(np.uint64), want.view(np.uint64)), (i, retest)
            assert bool(accepted[i]) == O.accept(want, scales)
    far = scores[:, 15] < -5.0
    say(f"  closest approach beyond 5 t on {int(far.sum())} of {len(cT)} candidates; accepted {int(accepted.sum())}")
    assert far.any() and accepted.any()


def test_closest_approach_at_every_separation(ex):
    """Two fragments of a 3 edges/t vessel (tens of thousands of triangles each) set apart at
    0.05 t to 100 t under random rotations: nothing is inside, so R §6.4 reads the closest approach
    (verify.rs:567-578), which the pair search over sample boxes and mesh nodes must find bit for bit
    as the oracle's per-sample minimum does."""
    from sherd_hip import synth

    say("case: closest approach, 2 fragments, separations 0.05-100 t, bit-exact vs oracle")
    col = synth.vessel(t=1.0, edges_per_t=3.0, n_frag=2, seed=21, height_t=40.0, rmin_t=15.0, rmax_t=20.0)
    frags = col["fragments"]
    scales = synth.scales_for_pair(col["t"], col["res"])
    rng = np.random.default_rng(8)
    truth = col["poses"][0] @ np.linalg.inv(col["poses"][1])
    cand = []
    for gap in (0.05, 0.3, 1.0, 3.0, 10.0, 30.0, 100.0):
        for _ in range(3):
            d = np.eye(4)
            d[:3, :3] = synth._random_rotation(rng, np.pi)
            v = rng.normal(size=3)
            d[:3, 3] = v / np.linalg.norm(v) * (gap + 60.0 * (gap > 1.0))
            cand.append(d @ truth)
    cT = np.array(cand)
    dev, ora, keep = [], [], []
    for fr in frags:
        fs, ms = ex.scene(fr["V"], fr["F"], fr["frac"]), ex.scene(fr["V"], fr["F"])
        ofs, oms = O.Scene(fr["V"], fr["F"], fr["frac"]), O.Scene(fr["V"], fr["F"])
        keep += [fs, ms, ofs, oms]
        args = (fr["watertight"], fr["frac_area"], fr["Pf"], fr["S"], fr["brk_P"], fr["brk_ns"], fr["margin_p"],
                fr["margin_n"])
        dev.append(ex.fragment(fs, ms, *args))
        ora.append(O.SideData(ofs, oms, *args))
    cp = np.zeros(len(cT), np.uint32)
    for retest in (False, True):
        scores, _ = ex.verify_batch([(dev[0], dev[1], scales)], cp, cT, retest=retest)
        for i in range(len(cT)):
            want = O.verify(ora[0], ora[1], cT[i], scales, retest=retest)
            assert np.array_equal(scores[i].view(np.uint64), want.view(np.uint64)), (i, retest, scores[i] - want)
    apart = scores[:, 15] < 0.0
    say(f"  {int(apart.sum())} of {len(cT)} candidates with nothing inside; closest approach "
        f"{-scores[apart, 15].min():.3f}-{-scores[apart, 15].max():.3f} t")
    assert apart.sum() >= len(cT) // 2
