"""Oracle, R §3.4.4-3.4.8's post-processing of the cone vote. TEST INFRASTRUCTURE ONLY: tests/
import it as the checker; nothing in the product imports, links or runs it.

A restatement of crates/sherd-core/src/fragment/segment.rs:403-480 and 544-635 with the mesh helpers
it calls, in numpy and Python integers:

  face_geometry      mesh/geometry.rs:70-96: f64 unit normals (length floored at 1e-12), areas,
                     centroids, elementwise in its order
  face_adjacency     mesh/adjacency.rs:38-51, 138-158: the 3m edge rows slot-major, a stable sort by
                     key, every two consecutive rows with one key a pair
  within             spatial/kdtree.rs:219-245: every point with kiddo's squared distance
                     (dx^2 + dy^2) + dz^2 <= r^2, ascending by index
  majority           segment.rs:403-428: per representative the f64 area sums over its ball in
                     ascending index order; a face takes its nearest representative's answer
  ball_normals       segment.rs:267-305 over the shell faces only (:584-608)
  drop_small         mesh/components.rs:120-205: components of mask == target over the pairs, each
                     flipped when its area is below the bound
  grow               segment.rs:438-480: the growth passes, candidates read before any flip
  cleanup            segment.rs:544-635's steps 3.4.4 to 3.4.8 in order
  vertex_votes       the per-vertex form of R §3.4.3's rays (one ray per vertex along the inward
                     vertex normal, refine.rs:118-150's normals), with the test of the cone's first
                     ray (segment.rs:330-372), and a face fracture when fewer than two of its
                     vertices voted

One deliberate difference, shared with the device: a component's area is the exact sum of its faces'
areas, each first truncated to a multiple of 2^(E - 64) where 2^E is the least power of two above
t^2 (mesh/components.rs:188-196 sums in face order in f64). The two rules give different labels
only for a component whose area lies within the f64 sum's rounding of the bound; `drop_small`
with exact=False is the reference's rule, and the tests check the two agree.

What pins it: the slab dump's own masks after each step (seg.frac_majority, seg.frac_islands,
seg.frac_final, seg.ref / seg.has_ref) from its own seg.frac_raw, seg.rep and seg.near by
area-weighted agreement (the reference's parity gate, crates/sherd-parity/src/stages/
segmentation.rs), brute-force connected components (scipy's csgraph) for the islands, and known
answers on small meshes.
"""

import math

import numpy as np

NORMAL_FLOOR = 1e-12
MAJORITY_T = 0.25
REFERENCE_T = 0.5
MIN_FRACTURE_T2 = 0.5
MIN_SHELL_T2 = 2.0
MAX_GROWTH_PASSES = 60
BOUNDARY_ANGLE = 25.0


def face_geometry(V, F):
    V = np.asarray(V, dtype=np.float64)
    F = np.asarray(F).astype(np.int64)
    a, b, c = V[F[:, 0]], V[F[:, 1]], V[F[:, 2]]
    e1, e2 = b - a, c - a
    n = np.stack([e1[:, 1] * e2[:, 2] - e1[:, 2] * e2[:, 1],
                  e1[:, 2] * e2[:, 0] - e1[:, 0] * e2[:, 2],
                  e1[:, 0] * e2[:, 1] - e1[:, 1] * e2[:, 0]], axis=1)
    ln = np.sqrt((n[:, 0] * n[:, 0] + n[:, 1] * n[:, 1]) + n[:, 2] * n[:, 2])
    den = np.maximum(ln, NORMAL_FLOOR)
    normals = n / den[:, None]
    areas = 0.5 * ln
    cen = ((a + b) + c) / 3.0
    return normals, areas, cen


def face_adjacency(F):
    """(fa, fb): one pair per two consecutive rows of the stably key-sorted edge list."""
    F = np.asarray(F).astype(np.int64)
    n = len(F)
    rows = np.concatenate([F[:, [0, 1]], F[:, [1, 2]], F[:, [2, 0]]])
    lo, hi = np.minimum(rows[:, 0], rows[:, 1]), np.maximum(rows[:, 0], rows[:, 1])
    key = (lo.astype(np.uint64) << np.uint64(32)) | hi.astype(np.uint64)
    order = np.argsort(key, kind="stable")
    k = key[order]
    same = np.nonzero(k[1:] == k[:-1])[0]
    return (order[same] % n).astype(np.int64), (order[same + 1] % n).astype(np.int64)


def within(points, q, r, candidates=None):
    """Indices of points with (dx^2 + dy^2) + dz^2 <= r^2, ascending."""
    P = points if candidates is None else points[candidates]
    d = P - q
    d2 = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
    ids = np.nonzero(d2 <= r * r)[0]
    if candidates is not None:
        ids = np.sort(np.asarray(candidates)[ids])
    return ids


def _balls(points, queries, r):
    """Per query, the ascending indices within r (a KD-tree for candidates, the exact test here)."""
    from scipy.spatial import cKDTree

    tree = cKDTree(points)
    cands = tree.query_ball_point(queries, r * (1.0 + 1e-9) + 1e-300)
    return [within(points, q, r, np.asarray(c, dtype=np.int64)) for q, c in zip(queries, cands)]


def majority(cen, areas, frac, rep, near, t):
    r = t * MAJORITY_T
    per_rep = np.zeros(len(rep), dtype=bool)
    for i, ball in enumerate(_balls(cen, cen[np.asarray(rep, dtype=np.int64)], r)):
        wet = 0.0
        total = 0.0
        for j in ball:
            a = float(areas[j])
            total += a
            if frac[j]:
                wet += a
        per_rep[i] = wet > 0.5 * total
    return per_rep[np.asarray(near, dtype=np.int64)]


def ball_normals(cen, normals, areas, mask, queries, r):
    """Area-weighted mean normal of the faces with mask set within r of each query, and whether any."""
    sub = np.nonzero(mask)[0]
    out = np.zeros((len(queries), 3))
    has = np.zeros(len(queries), dtype=bool)
    if len(sub) == 0:
        return out, has
    for i, ball in enumerate(_balls(cen[sub], queries, r)):
        m = [0.0, 0.0, 0.0]
        w = 0.0
        for jj in ball:
            j = sub[jj]
            a = float(areas[j])
            for k in range(3):
                m[k] += float(normals[j, k]) * a
            w += a
        den = max(w, 1e-12)
        m = [v / den for v in m]
        norm = max(math.sqrt((m[0] * m[0] + m[1] * m[1]) + m[2] * m[2]), 1e-12)
        out[i] = [m[0] / norm, m[1] / norm, m[2] / norm]
        has[i] = len(ball) > 0
    return out, has


def components(mask, target, fa, fb):
    """Component root (lowest face index) of every face with mask == target, -1 elsewhere."""
    n = len(mask)
    parent = list(range(n))

    def find(x):
        while parent[x] != x:
            parent[x] = parent[parent[x]]
            x = parent[x]
        return x

    for a, b in zip(fa.tolist(), fb.tolist()):
        if mask[a] == target and mask[b] == target:
            ra, rb = find(a), find(b)
            if ra != rb:
                parent[max(ra, rb)] = min(ra, rb)
    return np.array([find(i) if mask[i] == target else -1 for i in range(n)], dtype=np.int64)


def area_exponent(t):
    """E with 2^(E-1) <= t^2 < 2^E: the exact sums count in units of 2^(E - 64)."""
    return math.frexp(t * t)[1]


def quantized(a, e):
    return int(math.floor(math.ldexp(float(a), 64 - e)))


def drop_small(mask, target, min_area, fa, fb, areas, t, exact=True):
    mask = np.array(mask, dtype=bool)
    root = components(mask, target, fa, fb)
    inset = root >= 0
    if exact:
        e = area_exponent(t)
        sums = {}
        for i in np.nonzero(inset)[0].tolist():
            sums[root[i]] = sums.get(root[i], 0) + quantized(areas[i], e)
        bound = math.ceil(math.ldexp(min_area, 64 - e))
        small = {r for r, s in sums.items() if s < bound}
    else:
        sums = {}
        for i in np.nonzero(inset)[0].tolist():
            sums[root[i]] = sums.get(root[i], 0.0) + float(areas[i])
        small = {r for r, s in sums.items() if s < min_area}
    for i in np.nonzero(inset)[0].tolist():
        if root[i] in small:
            mask[i] = not target
    return mask


def grow(frac, fa, fb, normals, reference, has_ref, angle_deg=BOUNDARY_ANGLE):
    frac = np.array(frac, dtype=bool)
    cos_grow = math.cos(math.radians(angle_deg))
    n, r = normals, reference
    dot = (n[:, 0] * r[:, 0] + n[:, 1] * r[:, 1]) + n[:, 2] * r[:, 2]
    eligible = has_ref & (dot > cos_grow)
    passes = 0
    for _ in range(MAX_GROWTH_PASSES):
        diff = frac[fa] != frac[fb]
        cand = np.zeros(len(frac), dtype=bool)
        cand[np.where(frac[fa[diff]], fa[diff], fb[diff])] = True
        if not cand.any():
            break
        passes += 1
        flip = cand & eligible
        if not flip.any():
            break
        frac[flip] = False
    return frac, passes


def cleanup(V, F, t, frac_raw, rep, near, angle_deg=BOUNDARY_ANGLE, exact=True):
    """Steps 3.4.4-3.4.8: a dict of the masks after each step, the reference normals and the labels."""
    normals, areas, cen = face_geometry(V, F)
    rep = np.asarray(rep, dtype=np.int64)
    near = np.asarray(near, dtype=np.int64)
    fa, fb = face_adjacency(F)
    frac = majority(cen, areas, np.asarray(frac_raw, dtype=bool), rep, near, t)
    out = {"majority": frac.copy()}
    frac = drop_small(frac, True, MIN_FRACTURE_T2 * t * t, fa, fb, areas, t, exact)
    frac = drop_small(frac, False, MIN_SHELL_T2 * t * t, fa, fb, areas, t, exact)
    out["islands"] = frac.copy()
    ref_g, has_g = ball_normals(cen, normals, areas, ~frac, cen[rep], t * REFERENCE_T)
    reference, has_ref = ref_g[near], has_g[near]
    out["reference"], out["has_ref"] = reference, has_ref
    frac, out["passes"] = grow(frac, fa, fb, normals, reference, has_ref, angle_deg)
    out["grown"] = frac.copy()
    frac = drop_small(frac, True, MIN_FRACTURE_T2 * t * t, fa, fb, areas, t, exact)
    out["final"] = frac
    return out


def agreement(a, b, areas):
    """Area-weighted fraction of faces whose two labels agree (segment.rs:655)."""
    return float(np.sum(areas[np.asarray(a) == np.asarray(b)]) / np.sum(areas))


def vertex_votes(scene, V, F, t, retest=False):
    """One ray per vertex from 1e-3 inside along its normal (refine.rs:118-150's area-weighted vertex
    normal), straight down, narrowed to f32 once; good when the first hit (scene: the oracle's
    scene of V narrowed to f32) lies in the wall window and its face looks back along the ray
    (segment.rs:357-372: prim valid, dh > f32(0.1 t), 0.5 < dh / f32(t) < 1.8 in f32, h . d > 0.7 in
    f64). Returns (good per vertex as u8, fracture label per face as u8)."""
    import refine as R

    V = np.asarray(V, dtype=np.float64)
    F = np.asarray(F, dtype=np.int64).reshape(-1, 3)
    n = R.vertex_normals(V, F)
    ok = np.isfinite(n).all(axis=1) & (n != 0.0).any(axis=1)
    origin = (V - n * 1e-3).astype(np.float32)
    d = -n
    hit_t, prim = scene.first_hit(origin, d.astype(np.float32), retest=retest)
    hit_t = np.asarray(hit_t, dtype=np.float32)
    prim = np.asarray(prim, dtype=np.int64)
    min_hit, th = np.float32(0.1 * t), np.float32(t)
    with np.errstate(invalid="ignore", divide="ignore"):
        ratio = hit_t / th
    valid = ok & (prim != 0xFFFFFFFF) & np.isfinite(hit_t) & (prim < len(F)) & (hit_t > min_hit)
    valid &= (ratio > np.float32(0.5)) & (ratio < np.float32(1.8))
    good = np.zeros(len(V), np.uint8)
    idx = np.nonzero(valid)[0]
    if len(idx):
        normals, _, _ = face_geometry(V, F)
        h = normals[prim[idx]]
        al = (h[:, 0] * d[idx, 0] + h[:, 1] * d[idx, 1]) + h[:, 2] * d[idx, 2]
        good[idx[al > 0.7]] = 1
    frac = ((good[F[:, 0]].astype(int) + good[F[:, 1]] + good[F[:, 2]]) < 2).astype(np.uint8)
    return good, frac
