"""HIP LBVH queries against the oracle on the same inputs (GPU): first hits, the inside test and
the bounded distance, bit for bit.

Inputs: the slab fixture's own meshes, samples and stage-2 poses (tests/golden/slab), the
reference's box/wall known answers, and seeded random triangle soups that stress the tree
(non-manifold, overlapping, degenerate faces).
"""

import sys

import numpy as np
import pytest

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
    say("device:", e.device_name())
    yield e
    e.close()


def box_mesh(lo, hi):
    v = np.array([
        [lo[0], lo[1], lo[2]], [hi[0], lo[1], lo[2]], [hi[0], hi[1], lo[2]], [lo[0], hi[1], lo[2]],
        [lo[0], lo[1], hi[2]], [hi[0], lo[1], hi[2]], [hi[0], hi[1], hi[2]], [lo[0], hi[1], hi[2]],
    ], dtype=np.float32)
    f = np.array([[0, 2, 1], [0, 3, 2], [4, 5, 6], [4, 6, 7], [0, 1, 5], [0, 5, 4], [3, 7, 6], [3, 6, 2],
                  [0, 4, 7], [0, 7, 3], [1, 2, 6], [1, 6, 5]], dtype=np.uint32)
    return v, f


def test_box_known_answers(ex):
    say("case: box first hits / inside")
    v, f = box_mesh([-1, -1, -1], [1, 1, 1])
    s = ex.scene(v, f)
    n, lo, hi = s.info()
    assert n == 12 and list(lo) == [-1, -1, -1] and list(hi) == [1, 1, 1]
    t, prim = ex.first_hit(s, [[0, 0, 5], [0, 0, 0], [0, 0, 5]], [[0, 0, -1], [0, 0, -1], [0, 0, 1]])
    assert abs(t[0] - 4.0) < 1e-5 and abs(t[1] - 1.0) < 1e-5 and not np.isfinite(t[2])
    assert prim[2] == 0xFFFFFFFF and prim[0] == 2
    v10, f10 = box_mesh([0, 0, 0], [10, 10, 10])
    s10 = ex.scene(v10, f10)
    depth, ins = ex.inside(s10, [[4, 5, 6], [1, 5.5, 6.5], [50, 5, 5], [-5, 5, 5], [5, 5, 9]], np.eye(4))
    assert list(ins) == [1, 1, 0, 0, 1]
    assert abs(depth[0] - 4.0) < 1e-5 and abs(depth[4] - 1.0) < 1e-5


def test_single_triangle_scene(ex):
    say("case: one-triangle scene")
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    s = ex.scene(v, np.array([[0, 1, 2]], np.uint32))
    t, p = ex.first_hit(s, [[0.25, 0.25, 1.0]], [[0, 0, -1]])
    assert abs(t[0] - 1.0) < 1e-6 and p[0] == 0
    d = ex.bounded_distance(s, [[0.25, 0.25, 0.5], [-3, -4, 0]], np.eye(4), 10.0)
    assert abs(d[0] - 0.5) < 1e-6 and abs(d[1] - 5.0) < 1e-6


@pytest.mark.parametrize("name", ["pieceA", "pieceB"])
@pytest.mark.parametrize("retest", [False, True])
def test_thickness_rays_bit_exact(ex, name, retest):
    say(f"case: thickness rays {name} retest={retest}")
    pc = S.piece(name)
    v0 = pc["V0"].astype(np.float32)
    o, d = S.thickness_rays(name)
    dev = ex.scene(v0, pc["F"])
    t, prim = ex.first_hit(dev, o, d, retest=retest)
    ot, op = O.Scene(v0, pc["F"]).first_hit(o, d, retest=retest)
    assert np.array_equal(t.view(np.uint32), ot.view(np.uint32))
    assert np.array_equal(prim, op)


def slab_scenes(ex):
    out = {}
    for name in ("pieceA", "pieceB"):
        pc = S.piece(name)
        v32 = pc["V"].astype(np.float32)
        out[name] = dict(
            frac=(ex.scene(v32, pc["F"], pc["frac"]), O.Scene(v32, pc["F"], pc["frac"])),
            mesh=(ex.scene(v32, pc["F"]), O.Scene(v32, pc["F"])),
        )
    return out


@pytest.fixture(scope="module")
def scenes(ex):
    return slab_scenes(ex)


def test_bounded_distance_bit_exact_on_the_stage2_poses(ex, scenes):
    p = S.pair()
    sc = p["scales"]
    b = S.piece("pieceB")
    dev, ora = scenes["pieceA"]["frac"]
    for i, T in enumerate(p["s2_T_frac2"]):
        say(f"case: bounded distance candidate {i}")
        got = ex.bounded_distance(dev, b["Pf"], T, sc[5])
        want = ora.bounded_distance(b["Pf"], T, sc[5])
        assert np.array_equal(got, want), (i, np.sum(got != want))
        gmin = ex.bounded_distance(dev, b["S"], T, np.finfo(np.float32).max, reduce="min")
        wmin = ora.bounded_distance(b["S"], T, np.finfo(np.float32).max, reduce_min=True)
        assert gmin[0] == wmin[0]


@pytest.mark.parametrize("retest", [False, True])
def test_inside_bit_exact_on_the_stage2_poses(ex, scenes, retest):
    p = S.pair()
    b = S.piece("pieceB")
    dev, ora = scenes["pieceA"]["mesh"]
    for i, T in enumerate(p["s2_T_frac2"]):
        say(f"case: inside candidate {i} retest={retest}")
        gd, gi = ex.inside(dev, b["S"], T, retest=retest)
        od, oi = ora.inside(b["S"], T, retest=retest)
        assert np.array_equal(gi, oi), (i, np.sum(gi != oi))
        assert np.array_equal(gd.view(np.uint32), od.view(np.uint32))


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_random_soups_against_the_oracle(ex, seed):
    """Overlapping, non-manifold and degenerate triangles with many exact ties."""
    say(f"case: random soup seed={seed}")
    rng = np.random.default_rng(seed)
    nv, nf = 2000, 5000
    v = rng.uniform(-10, 10, size=(nv, 3)).astype(np.float32)
    v[: nv // 4] = np.round(v[: nv // 4])  # lattice vertices: coplanar faces and exact ties
    f = rng.integers(0, nv, size=(nf, 3)).astype(np.uint32)
    f[:50, 2] = f[:50, 1]  # degenerate faces
    dev = ex.scene(v, f)
    ora = O.Scene(v, f)
    o = rng.uniform(-12, 12, size=(4000, 3)).astype(np.float32)
    o[:500] = np.round(o[:500])
    d = rng.normal(size=(4000, 3))
    d[:300] = [[1, 0, 0], [0, -1, 0], [0, 0, 1]] * 100
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    for retest in (False, True):
        t, p = ex.first_hit(dev, o, d, retest=retest)
        ot, op = ora.first_hit(o, d, retest=retest)
        assert np.array_equal(t.view(np.uint32), ot.view(np.uint32)) and np.array_equal(p, op)
    pts = rng.uniform(-12, 12, size=(3000, 3))
    T = np.eye(4)
    T[:3, 3] = [0.25, -0.5, 0.125]
    gd, gi = ex.inside(dev, pts, T)
    od, oi = ora.inside(pts, T)
    assert np.array_equal(gi, oi) and np.array_equal(gd.view(np.uint32), od.view(np.uint32))
    for w in (0.5, 3.0, np.inf):
        g = ex.bounded_distance(dev, pts, T, w)
        o_ = ora.bounded_distance(pts, T, w)
        assert np.array_equal(g, o_)


@pytest.mark.parametrize("name", ["pieceA", "pieceB"])
def test_cone_votes_bit_exact_and_against_the_dump(ex, name):
    say(f"case: cone votes {name}")
    pc = S.piece(name)
    v32 = pc["V"].astype(np.float32)
    dev = ex.scene(v32, pc["F"])
    good = ex.cone_votes(dev, pc["V"], pc["F"], pc["seg_faces"], pc["seg_NS"], pc["thick_t"])
    want = O.Scene(v32, pc["F"]).cone_votes(pc["V"], pc["F"], pc["seg_faces"], pc["seg_NS"], pc["thick_t"])
    assert np.array_equal(good, want)
    assert np.mean(good == pc["seg_good"]) > 0.999
