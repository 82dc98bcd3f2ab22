"""R §3.4.4-3.4.8 on the device (crates/sherd-core/src/fragment/segment.rs:544-635): every mask and
the shell reference normals against the oracle (oracle/segment.py) bit for bit, and the final
labels against the slab dump's seg.frac_final from its own seg.frac_raw, seg.rep and seg.near."""

import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import segment as SG  # noqa: E402

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
def seg():
    return np.load(os.path.join(GOLD, "seg.npz"))


def check_against_oracle(got, trace, want):
    assert np.array_equal(trace["majority"], want["majority"])
    assert np.array_equal(trace["islands"], want["islands"])
    assert np.array_equal(trace["has_ref"], want["has_ref"])
    h = want["has_ref"]
    assert trace["reference"][h].tobytes() == want["reference"][h].tobytes()
    assert trace["passes"] == want["passes"]
    assert np.array_equal(got, want["final"])


@pytest.mark.parametrize("name", ["pieceA", "pieceB"])
def test_the_slab_dumps_labels(ex, seg, name):
    say(f"case: segmentation clean-up on the dump's inputs, {name}")
    g = np.load(os.path.join(GOLD, name + ".npz"))
    V, F, t = g["V"], g["F"], float(g["thick_t"])
    raw, rep, near = seg[name + "_frac_raw"], seg[name + "_rep"], seg[name + "_near"]
    got, trace = ex.segment_cleanup(V, F, t, raw, rep, near, trace=True)
    want = SG.cleanup(V, F, t, raw, rep, near)
    check_against_oracle(got, trace, want)
    assert np.array_equal(got, seg[name + "_frac_final"])
    assert np.array_equal(trace["majority"], seg[name + "_frac_majority"])
    assert np.array_equal(trace["islands"], seg[name + "_frac_islands"])


@pytest.mark.parametrize("seed", [1, 2])
def test_a_noisy_vote_on_the_slab(ex, seg, seed):
    say(f"case: segmentation clean-up, slab pieceA, vote with 30 % flipped (seed {seed})")
    g = np.load(os.path.join(GOLD, "pieceA.npz"))
    V, F, t = g["V"], g["F"], float(g["thick_t"])
    raw = seg["pieceA_frac_raw"].copy()
    flip = np.random.default_rng(seed).random(len(raw)) < 0.3
    raw[flip] = ~raw[flip]
    got, trace = ex.segment_cleanup(V, F, t, raw, seg["pieceA_rep"], seg["pieceA_near"], trace=True)
    check_against_oracle(got, trace, SG.cleanup(V, F, t, raw, seg["pieceA_rep"], seg["pieceA_near"]))


def test_a_vessel_fragment(ex):
    say("case: segmentation clean-up of a vessel fragment (its own grid, a vote with 20 % flipped)")
    from sherd_hip import synth

    col = synth.vessel(t=1.0, edges_per_t=5.5, n_frag=4, seed=5, height_t=20.0, rmin_t=10.0, rmax_t=16.0)
    fr = col["fragments"][0]
    V, F, t = fr["V"].astype(np.float64), fr["F"], col["t"]
    _, _, cen = SG.face_geometry(V, F)
    rep, near = synth.coarse_grid(cen, t / 8.0)
    raw = fr["frac"].copy()
    flip = np.random.default_rng(7).random(len(raw)) < 0.2
    raw[flip] = ~raw[flip]
    say(f"  {len(F)} faces, {len(rep)} representatives")
    got, trace = ex.segment_cleanup(V, F, t, raw, rep, near, trace=True)
    want = SG.cleanup(V, F, t, raw, rep, near)
    check_against_oracle(got, trace, want)
    areas = SG.face_geometry(V, F)[1]
    before, after = SG.agreement(raw, fr["frac"], areas), SG.agreement(got, fr["frac"], areas)
    say(f"  agreement with the true labels: vote {before:.4f}, cleaned {after:.4f}")
    assert after > 0.99 and after > before


def test_errors_and_no_wall(ex, seg):
    say("case: segmentation clean-up error paths")
    import sherd_hip

    g = np.load(os.path.join(GOLD, "pieceB.npz"))
    V, F = g["V"], g["F"]
    raw, rep, near = seg["pieceB_frac_raw"], seg["pieceB_rep"], seg["pieceB_near"].copy()
    assert not ex.segment_cleanup(V, F, 0.0, raw, rep, near).any()
    near[5] = len(rep)
    with pytest.raises(sherd_hip.HipError, match="representative"):
        ex.segment_cleanup(V, F, float(g["thick_t"]), raw, rep, near)
    # The context stays usable.
    got = ex.segment_cleanup(V, F, float(g["thick_t"]), raw, rep, seg["pieceB_near"])
    assert np.array_equal(got, seg["pieceB_frac_final"])


@pytest.mark.parametrize("name", ["pieceA", "pieceB"])
def test_smoothed_normals(ex, seg, name):
    say(f"case: R 3.4.2 smoothed normals, {name}")
    g = np.load(os.path.join(GOLD, name + ".npz"))
    V, F, t = g["V"], g["F"], float(g["thick_t"])
    rep, near = seg[name + "_rep"], seg[name + "_near"]
    got = ex.smoothed_normals(V, F, rep, near, t / 3.0)
    normals, areas, cen = SG.face_geometry(V, F)
    ref_g, _ = SG.ball_normals(cen, normals, areas, np.ones(len(F), bool), cen[rep.astype(np.int64)], t / 3.0)
    want = ref_g[near.astype(np.int64)]
    assert got.tobytes() == want.tobytes()
    # The dump's own NS (every fourth face), whose Python balls sum in another order.
    faces = g["seg_faces"]
    cosang = np.sum(got[faces] * g["seg_NS"], axis=1)
    assert cosang.min() > np.cos(np.radians(1e-5))
