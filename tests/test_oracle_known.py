"""The oracle against the reference's own known-answer unit tests (CPU).

Each test names the reference test it transcribes. These pin the parts of the oracle the slab
fixture does not reach on its own: boundary conventions (strict vs inclusive radii), tie rules,
closed-form distances, and the pose algebra.
"""

import numpy as np
import pytest

import oracle_lib as O

I4 = np.eye(4)


def box_mesh(lo, hi):
    """bvh.rs:222-252 / verify.rs:838-868: an axis-aligned closed box, outward normals."""
    v = np.array([
        [lo[0], lo[1], lo[2]], [hi[0], lo[1], lo[2]], [hi[0], hi[1], lo[2]], [lo[0], hi[1], lo[2]],
        [lo[0], lo[1], hi[2]], [hi[0], lo[1], hi[2]], [hi[0], hi[1], hi[2]], [lo[0], hi[1], hi[2]],
    ], dtype=np.float64)
    f = np.array([[0, 2, 1], [0, 3, 2], [4, 5, 6], [4, 6, 7], [0, 1, 5], [0, 5, 4], [3, 7, 6], [3, 6, 2],
                  [0, 4, 7], [0, 7, 3], [1, 2, 6], [1, 6, 5]], dtype=np.uint32)
    return v, f


def wall():
    """verify.rs:826-835: the square x = 0, [0, 2]^2 in y and z, two triangles."""
    v = np.array([[0, 0, 0], [0, 2, 0], [0, 2, 2], [0, 0, 2]], dtype=np.float64)
    return v, np.array([[0, 1, 2], [0, 2, 3]], dtype=np.uint32)


def wall_samples():
    return np.array([[0.0, 0.2 + i * 0.2, 0.2 + j * 0.2] for i in range(10) for j in range(10)])


def shifted(dx):
    t = np.eye(4)
    t[0, 3] = dx
    return t


# --- bvh.rs:255-280 ----------------------------------------------------------------------------

def test_scene_first_hits_and_misses():
    v, f = box_mesh([-1.0] * 3, [1.0] * 3)
    s = O.Scene(v, f)
    assert s.n_faces == 12
    t, prim = s.first_hit([[0, 0, 5], [0, 0, 0], [0, 0, 5]], [[0, 0, -1], [0, 0, -1], [0, 0, 1]])
    assert abs(t[0] - 4.0) < 1e-5 and abs(t[1] - 1.0) < 1e-5
    assert not np.isfinite(t[2]) and prim[2] == 0xFFFFFFFF
    # Both lid triangles meet the ray on their shared diagonal: the lowest face wins the tie.
    assert prim[0] == 2


def test_closest_distance_from_either_side():
    v, f = box_mesh([-1.0] * 3, [1.0] * 3)
    s = O.Scene(v, f)
    assert abs(s.distance([0, 0, 3]) - 2.0) < 1e-5
    assert abs(s.distance([0, 0, 0.5]) - 0.5) < 1e-5
    # Brute force and the BVH answer the same bits on a sweep.
    rng = np.random.default_rng(5)
    for p in rng.uniform(-3, 3, size=(300, 3)):
        assert s.distance(p) == s.distance(p, brute=True)


def test_distance_closed_forms():
    """Point-triangle distance against what the geometry fixes: above the face, beyond an edge,
    beyond a vertex of the right triangle (0,0,0) (1,0,0) (0,1,0)."""
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], dtype=np.float64)
    s = O.Scene(v, np.array([[0, 1, 2]], dtype=np.uint32))
    cases = [([0.25, 0.25, 0.5], 0.5), ([0.5, -0.75, 0.0], 0.75), ([-3.0, -4.0, 0.0], 5.0),
             ([2.0, 2.0, 0.0], np.sqrt(2 * 1.5 ** 2)), ([0.2, 0.2, -2.0], 2.0), ([1.0, 1.0, 1.0], np.sqrt(0.5 + 1.0))]
    for p, want in cases:
        assert abs(s.distance(p) - want) < 1e-6, (p, s.distance(p), want)


# --- the inside test (bvh.rs:146-178, verify.rs:1000-1060) --------------------------------------

def test_ray_parity_in_a_box():
    v, f = box_mesh([0.0] * 3, [10.0] * 3)
    s = O.Scene(v, f)
    depth, ins = s.inside([[4, 5, 6], [1, 5.5, 6.5], [50, 5, 5], [-5, 5, 5], [5, 5, 9]], I4)
    assert list(ins) == [1, 1, 0, 0, 1]
    assert abs(depth[0] - 4.0) < 1e-5 and abs(depth[1] - 1.0) < 1e-5 and abs(depth[4] - 1.0) < 1e-5
    # An axis ray through a face diagonal counts twice; the majority of three still finds it.
    assert s.crossings([5, 5, 9], [0, 0, 1]) in (0, 2)


def _side(fracture=None, mesh=None, watertight=False, area=0.0, pf=(), s=(), brk_p=(), brk_ns=(),
          margin_p=(), margin_n=()):
    arr = lambda x: np.asarray(x, dtype=np.float64).reshape(-1, 3)
    return O.SideData(fracture, mesh, watertight, area, arr(pf), arr(s), arr(brk_p), arr(brk_ns),
                      arr(margin_p), arr(margin_n))


def _s(scores, key):
    return scores[O.OS_KEYS.index(key)]


# --- verify.rs:889-1120 ------------------------------------------------------------------------

def test_fracture_scores_on_two_walls():
    sc = O.scales_for_pair(1.0, 0.0)
    v, f = wall()
    scene = O.Scene(v, f)
    a = _side(scene, pf=wall_samples(), area=4.0)
    b = _side(scene, pf=wall_samples(), area=4.0)
    s = O.verify(a, b, I4, sc, full=False)
    assert all(abs(_s(s, k) - 1.0) < 1e-12 for k in ("tightA", "tightB", "tight"))
    assert all(_s(s, k) < 1e-6 for k in ("gapA", "gapB", "gap"))
    assert all(abs(_s(s, k) - 4.0) < 1e-9 for k in ("contactA", "contactB", "contact"))
    s = O.verify(a, b, shifted(0.005), sc, full=False)
    assert abs(_s(s, "tight") - 1.0) < 1e-12 and abs(_s(s, "gap") - 0.005) < 1e-6
    s = O.verify(a, b, shifted(0.05), sc, full=False)
    assert _s(s, "tight") == 0.0 and _s(s, "contact") == 0.0 and abs(_s(s, "gap") - 0.05) < 1e-6
    s = O.verify(a, b, shifted(0.5), sc, full=False)
    assert (_s(s, "tight"), _s(s, "gap"), _s(s, "contact")) == (0.0, 1.0, 0.0)


def test_the_two_sides_are_not_the_same_number():
    sc = O.scales_for_pair(1.0, 0.0)
    v, f = wall()
    scene = O.Scene(v, f)
    off = wall_samples().copy()
    off[:, 0] = 0.05
    s = O.verify(_side(scene, pf=wall_samples(), area=4.0), _side(scene, pf=off, area=4.0), I4, sc, full=False)
    assert abs(_s(s, "tightA") - 1.0) < 1e-12 and _s(s, "tightB") == 0.0 and _s(s, "tight") == 0.0
    assert abs(_s(s, "gap") - 0.05) < 1e-6


def test_seam_counts_voxels_of_the_shared_curve():
    v, f = wall()
    scene = O.Scene(v, f)
    curve = np.array([[0.0, k * 0.1, 0.0] for k in range(30)])
    up = np.tile([0.0, 0.0, 1.0], (30, 1))
    a = _side(scene, brk_p=curve, brk_ns=up)
    sc = O.scales_for_pair(3.0, 0.0)
    assert abs(_s(O.verify(a, a, I4, sc, full=False), "seam") - 1.0) < 1e-12
    assert _s(O.verify(a, a, shifted(0.5), sc, full=False), "seam") == 0.0
    flipped = _side(scene, brk_p=curve, brk_ns=-up)
    assert _s(O.verify(a, flipped, I4, sc, full=False), "seam") == 0.0
    long = np.array([[0.0, k * 0.1, 0.0] for k in range(60)])
    a2 = _side(scene, brk_p=long, brk_ns=np.tile([0.0, 0.0, 1.0], (60, 1)))
    assert abs(_s(O.verify(a2, a2, I4, sc, full=False), "seam") - 2.0) < 1e-12


def test_continuity_step_and_normal_agreement():
    v, f = wall()
    scene = O.Scene(v, f)
    grid = np.array([[i * 0.1, j * 0.1, 0.0] for i in range(5) for j in range(5)])
    up = np.tile([0.0, 0.0, 1.0], (25, 1))
    a = _side(scene, margin_p=grid, margin_n=up)
    sc = O.scales_for_pair(1.0, 0.0)
    s = O.verify(a, a, I4, sc)
    assert abs(_s(s, "cont")) < 1e-12 and abs(_s(s, "cont_n") - 1.0) < 1e-12
    step = np.eye(4)
    step[2, 3] = 0.05
    s = O.verify(a, a, step, sc)
    assert abs(_s(s, "cont") - 0.05) < 1e-12 and abs(_s(s, "cont_n") - 1.0) < 1e-12
    flip = np.eye(4)
    flip[1, 1] = flip[2, 2] = -1.0
    assert abs(_s(O.verify(a, a, flip, sc), "cont_n") + 1.0) < 1e-12
    few = _side(scene, margin_p=grid[:10], margin_n=up[:10])
    s = O.verify(a, few, I4, sc)
    assert (_s(s, "cont"), _s(s, "cont_n")) == (1.0, -1.0)
    none = _side(scene)
    assert (_s(O.verify(a, none, I4, sc), "cont"), _s(O.verify(none, a, I4, sc), "cont_n")) == (1.0, -1.0)


def test_penetration_finds_what_is_inside():
    va, fa = box_mesh([0.0] * 3, [10.0] * 3)
    vb, fb = box_mesh([20.0] * 3, [30.0] * 3)
    mesh_a, mesh_b = O.Scene(va, fa), O.Scene(vb, fb)
    vw, fw = wall()
    frac = O.Scene(vw, fw)
    sc = O.scales_for_pair(1.0, 0.0)
    a = _side(frac, mesh_a, True, s=[[4, 5, 6], [1, 2, 3]])
    b = _side(frac, mesh_b, True, s=[[4, 5, 6], [1, 5.5, 6.5], [50, 5, 5], [-5, 5, 5]])
    s = O.verify(a, b, I4, sc)
    assert _s(s, "pen_unavailable") == 0.0
    assert abs(_s(s, "pen") - 0.5) < 1e-12 and abs(_s(s, "pen_depth") - 4.0) < 1e-5
    diag = _side(frac, mesh_b, True, s=[[5, 5, 9]])
    assert abs(_s(O.verify(a, diag, I4, sc), "pen") - 1.0) < 1e-12
    far = _side(frac, mesh_b, True, s=[[-4, 5, 5], [-9, 5, 5]])
    s = O.verify(a, far, I4, sc)
    assert _s(s, "pen") == 0.0 and abs(_s(s, "pen_depth") + 4.0) < 1e-5
    holed = _side(frac, mesh_b, False, s=[[5, 5, 5]])
    s = O.verify(a, holed, I4, sc)
    assert (_s(s, "pen"), _s(s, "pen_depth"), _s(s, "pen_unavailable")) == (0.0, 0.0, 1.0)


def test_acceptance_refuses_on_any_one_limit():
    sc = O.scales_for_pair(1.0, 0.0)
    good = np.zeros(20)
    good[O.OS_KEYS.index("tight")] = 0.25
    good[O.OS_KEYS.index("gap")] = sc[6] / sc[0]
    good[O.OS_KEYS.index("pen")] = 0.005
    good[O.OS_KEYS.index("seam")] = 3.0
    good[O.OS_KEYS.index("cont_n")] = 0.8
    assert O.accept(good, sc)
    for key, step in [("tight", -1e-9), ("gap", 1e-9), ("pen", 1e-9), ("seam", -1e-9), ("cont_n", -1e-9)]:
        bad = good.copy()
        bad[O.OS_KEYS.index(key)] += step
        assert not O.accept(bad, sc), key


def test_pose_inverse_is_numpys():
    """verify.rs:1239-1288: np.linalg.inv of a real stage-2 pose, within 4 ULP per entry."""
    t = np.array([
        [-0.32261010392212647, 0.936147816701064, -0.13982126495346298, -34.03820147233827],
        [-0.8883980542540967, -0.35044568486254296, -0.2965412604659947, -297.53302685991065],
        [-0.32660621250198646, 0.028549732871869956, 0.9447292176640116, 114.44892672052433],
        [0.0, 0.0, 0.0, 1.0]])
    want = np.array([
        [-0.3226101039221262, -0.8883980542540941, -0.3266062125019848, -237.92909937188142],
        [0.9361478167010636, -0.35044568486254224, 0.028549732871869994, -75.67186365972935],
        [-0.13982126495346295, -0.29654126046599444, 0.9447292176640085, -201.11332820507005],
        [0.0, 0.0, 0.0, 1.0]])
    got = O.pose_inverse(t)
    bound = 4 * np.finfo(np.float64).eps * np.maximum(np.abs(want), 1.0)
    assert np.all(np.abs(got - want) <= bound)


def test_median_is_numpys():
    assert O.median([3.0, 1.0, 2.0]) == 2.0
    assert O.median([4.0, 1.0, 3.0, 2.0]) == 2.5
    assert O.median([7.0]) == 7.0
    assert np.isnan(O.median([]))
    rng = np.random.default_rng(1)
    for n in (5, 6, 101, 1000):
        x = rng.normal(size=n)
        assert O.median(x) == np.median(x)


def test_scales_are_the_slab_pairs_own():
    """scales.rs:134-150: twelve exact doubles."""
    sc = O.scales_for_pair(30.127979278564453, 2.2784501327205304)
    want = [30.127979278564453, 2.2784501327205304, 5.240435305257219, 2.0506051194484773,
            0.3417675199080796, 9.038393783569335, 1.0253025597242387, 4.101210238896955,
            15.063989639282227, 2.0506051194484773, 15.063989639282227, 1.1343858038007926]
    assert list(sc) == want


# --- kNN: kdtree.rs:261-500, grid.rs:341-405 ---------------------------------------------------

def lattice():
    return np.array([[i, j, 0.0] for i in range(5) for j in range(5)], dtype=np.float64)


def test_nearest_below_is_strict_and_ties_go_low():
    p = lattice()
    idx, d = O.nearest_below(p, [[0.1, 0.2, 0.0], [2.9, 3.1, 0.0], [100, 100, 0]], 2.0)
    assert idx[0] == 0 and abs(d[0] - np.hypot(0.2, 0.1)) < 1e-15
    assert idx[1] == 18 and idx[2] == -1
    # kdtree.rs:403-411: 7^2 + 11^2 + 13^2 = 339; the bound at the distance itself is a miss.
    one = np.array([[7.0, 11.0, 13.0]])
    dd = np.sqrt(339.0)
    assert O.nearest_below(one, [[0, 0, 0]], dd)[0][0] == -1
    above = np.nextafter(dd, np.inf)
    assert O.nearest_below(one, [[0, 0, 0]], above)[0][0] == 0


def test_mirrored_cloud_ties_resolve_to_the_lower_index():
    """kdtree.rs:309-353: every point mirrored through x = 0; queries on the plane tie."""
    rng = np.random.default_rng(0)
    pts = []
    for p in rng.uniform(-4, 4, size=(300, 3)):
        pts.append(p)
        pts.append([-p[0], p[1], p[2]])
    pts = np.array(pts)
    q = np.column_stack([np.zeros(400), rng.uniform(-4, 4, 400), rng.uniform(-4, 4, 400)])
    idx, d = O.nearest_below(pts, q, 100.0)
    d2 = ((pts[None, :, :] - q[:, None, :]) ** 2)
    d2 = (d2[..., 0] + d2[..., 1]) + d2[..., 2]
    want = np.argmin(d2, axis=1)  # numpy's argmin takes the first minimum
    assert np.array_equal(idx, want)
    ties = np.sum(np.sum(d2 == d2.min(axis=1, keepdims=True), axis=1) > 1)
    assert ties > 100


def test_f32_mirror_tie_and_boundary():
    """grid.rs:379-405: ties to the lowest original index, strict vs inclusive radius."""
    pts = np.array([[1.0, 2.0, 3.0]] * 12 + [[1.0, 2.0, 3.5]], dtype=np.float32)
    idx, d = O.nearest_f32(pts, [[1.0, 2.0, 3.0], [1.0, 2.0, 3.6]], 1.0, strict=False)
    assert idx[0] == 0 and d[0] == 0.0
    assert idx[1] == 12 and d[1] == np.float32(0.099999905)
    one = np.array([[0.0, 0.0, 0.0]], dtype=np.float32)
    assert O.nearest_f32(one, [[1.0, 0, 0]], 1.0, strict=False)[0][0] == 0
    assert O.nearest_f32(one, [[1.0, 0, 0]], 1.0, strict=True)[0][0] == -1
    inside = np.nextafter(np.float32(1.0), np.float32(0.0))
    assert O.nearest_f32(one, [[inside, 0, 0]], 1.0, strict=True)[0][0] == 0


# --- coarse.rs:207-320 -------------------------------------------------------------------------

def line(n, x0=0.0, spacing=1.0):
    return np.array([[x0 + spacing * k, 0.0, 0.0] for k in range(n)]), np.tile([0.0, 0.0, 1.0], (n, 1))


def ident_pose():
    return np.concatenate([np.eye(3).reshape(9), np.zeros(3)])[None, :]


def test_coarse_fraction_boundary_and_normals():
    ap, an = line(10)
    s, c = O.coarse_scores(ap, an, ap[[0, 3, 7, 9]], an[[0, 3, 7, 9]], ident_pose(), 0.5)
    assert s[0] == 1.0 and c[0] == 4
    far = ap[[0, 3, 7, 9]] + [0, 0.6, 0]
    assert O.coarse_scores(ap, an, far, an[:4], ident_pose(), 0.5)[0][0] == 0.0
    a1, n1 = line(1)
    assert O.coarse_scores(a1, n1, [[0, 0.25, 0]], [[0, 0, 1]], ident_pose(), 0.5)[0][0] == 1.0
    assert O.coarse_scores(a1, n1, [[0, 0.5, 0]], [[0, 0, 1]], ident_pose(), 0.5)[0][0] == 0.0
    a4, n4 = line(4)
    assert O.coarse_scores(a4, n4, [[0, 0, 0], [1, 0, 0]], [[0, 0, -1]] * 2, ident_pose(), 0.5)[0][0] == 0.0
    tilt = lambda c: [[np.sqrt(1 - c * c), 0.0, c]]
    assert O.coarse_scores(a4, n4, [[0, 0, 0]], tilt(0.71), ident_pose(), 0.5)[0][0] == 1.0
    assert O.coarse_scores(a4, n4, [[0, 0, 0]], tilt(0.7), ident_pose(), 0.5)[0][0] == 0.0
    probe = [[0, 0, 0], [1, 0, 0], [0, 9, 0], [0, -9, 0]]
    assert O.coarse_scores(a4, n4, probe, [[0, 0, 1]] * 4, ident_pose(), 0.5)[0][0] == 0.5
    assert O.coarse_scores(a4, n4, np.zeros((0, 3)), np.zeros((0, 3)), ident_pose(), 0.5)[0][0] == 0.0


# --- icp.rs:917-1120 ---------------------------------------------------------------------------

def patch(n, step, offset):
    pts, nrm = [], []
    for i in range(n):
        for j in range(n):
            x, y = offset + i * step, offset + j * step
            z = 0.3 * np.sin(0.8 * x) + 0.2 * np.cos(1.1 * y)
            dx, dy = 0.24 * np.cos(0.8 * x), -0.22 * np.sin(1.1 * y)
            nn = np.sqrt(dx * dx + dy * dy + 1.0)
            pts.append([x, y, z])
            nrm.append([-dx / nn, -dy / nn, 1.0 / nn])
    return np.array(pts), np.array(nrm)


def rigid(axis, angle, t):
    axis = np.asarray(axis, float) / np.linalg.norm(axis)
    k = np.array([[0, -axis[2], axis[1]], [axis[2], 0, -axis[0]], [-axis[1], axis[0], 0]])
    r = np.eye(3) + np.sin(angle) * k + (1 - np.cos(angle)) * (k @ k)
    m = np.eye(4)
    m[:3, :3] = r
    m[:3, 3] = t
    return m


def test_point_to_point_recovers_a_rigid_transform():
    pts, _ = patch(20, 0.5, -5.0)
    truth = rigid([0.3, -0.7, 0.6], 0.04, [0.11, -0.07, 0.05])
    inv = np.linalg.inv(truth)
    src = (inv[:3, :3] @ pts.T).T + inv[:3, 3]
    out = O.icp_rung(src, pts, None, np.eye(4)[None], "point_to_point", 1.0, 20)[0]
    T = out[:16].reshape(4, 4)
    moved = (T[:3, :3] @ src.T).T + T[:3, 3]
    assert np.abs(moved - pts).max() < 1e-9
    assert abs(out[16] - 1.0) < 1e-15 and out[17] < 1e-9 and out[20] == 1.0 and out[19] <= 3


def test_point_to_plane_converges_to_a_plane_pair():
    pts = np.array([[-5.0 + i * 0.25, -5.0 + j * 0.25, 0.0] for i in range(40) for j in range(40)])
    nrm = np.tile([0.0, 0.0, 1.0], (len(pts), 1))
    place = rigid([1.0, 0.0, 0.0], 0.05, [0.0, 0.0, 0.35])
    src = (place[:3, :3] @ (pts + [0.13, -0.07, 0.0]).T).T + place[:3, 3]
    out = O.icp_rung(src, pts, nrm, np.eye(4)[None], "point_to_plane", 2.0, 30)[0]
    T = out[:16].reshape(4, 4)
    moved = (T[:3, :3] @ src.T).T + T[:3, 3]
    assert np.abs(moved[:, 2]).max() < 1e-9
    assert out[20] == 1.0 and abs(out[16] - 1.0) < 1e-15


def test_correspondence_radius_is_strict():
    tgt = np.array([[0.0, 0, 0], [100.0, 0, 0]])
    d = 1.3
    fit = lambda x: O.icp_rung([[x, 0, 0]], tgt, None, np.eye(4)[None], 0, d, 0)[0][16]
    assert abs(fit(np.nextafter(d, 0)) - 1.0) < 1e-15
    assert fit(d) == 0.0
    assert fit(np.nextafter(d, 2)) == 0.0


def test_nothing_to_match_leaves_the_pose_alone():
    pts, nrm = patch(10, 0.4, -2.0)
    init = rigid([0.0, 1.0, 0.0], 0.2, [50.0, 0.0, 0.0])
    for est in (0, 1):
        out = O.icp_rung(pts, pts, nrm, init[None], est, 0.01, 30)[0]
        assert np.array_equal(out[:16].reshape(4, 4), init) and out[16] == 0.0


@pytest.mark.parametrize("brute", [True])
def test_bvh_and_brute_force_agree_on_rays(brute):
    rng = np.random.default_rng(3)
    v = rng.uniform(-5, 5, size=(300, 3))
    f = rng.integers(0, 300, size=(400, 3)).astype(np.uint32)
    s = O.Scene(v, f)
    o = rng.uniform(-6, 6, size=(500, 3))
    d = rng.normal(size=(500, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    for retest in (False, True):
        t1, p1 = s.first_hit(o, d, retest=retest)
        t2, p2 = s.first_hit(o, d, retest=retest, brute=True)
        assert np.array_equal(t1, t2) and np.array_equal(p1, p2)
    for i in range(100):
        assert s.crossings(o[i], d[i]) == s.crossings(o[i], d[i], brute=True)
