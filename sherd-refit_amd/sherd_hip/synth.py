"""Synthetic fragment collections for the benchmark workloads (host-side input generation).

`vessel(...)` is the thin-wall vessel of the benchmark configurations: a surface of revolution
(height 200 t, outer radius a seeded cubic spline in [40 t, 80 t], wall t, rim thickened to 1.4 t)
cut into fragments by seeded Voronoi seeds on the mid-surface whose distances carry seeded smooth
noise of amplitude 0.25 t, so the fracture lines wander. Each fragment is a closed, consistently
oriented triangle mesh: its outer and inner shell, the end caps it owns, and the fracture walls it
shares with its neighbours. Neighbouring fragments are built on one shared vertex lattice, so
their fracture surfaces are exactly complementary and the relative pose that reassembles them is
known to machine precision. Every fragment is then moved by its own seeded rigid pose.

Per fragment the generator also produces what R §6 reads (verify.rs:223-302 `Surfaces`), drawn the
way R §3.6 draws them (params.rs:179-187 defaults): fracture samples Pf (150 per t^2, clamped to
[5 000, 12 000]), 20 000 surface samples S, the breakline (shell/fracture junction) thinned to
0.5 t with shell macro-normals, and 6 000 shell-margin points between 0.15 t and 0.6 t of it.
The meshes are f32; the samples are f64.
"""

import numpy as np
from scipy.interpolate import CubicSpline
from scipy.spatial import cKDTree

# Params::default (crates/sherd-core/src/params.rs:143-188) values R §3.6 draws with.
FRAC_PER_T2, MIN_FRAC, MAX_FRAC = 150.0, 5000, 12000
SURFACE_POINTS, MARGIN_POINTS = 20000, 6000
BRK_VOXEL, MACRO_INNER, MACRO_OUTER = 0.5, 0.15, 0.60
SUB_VOXEL = 2.0  # the hypothesis subset: breakline points thinned on a 2 t voxel grid


def scales_for_pair(t, res):
    """Scales::for_pair (crates/sherd-core/src/matching/scales.rs:72-88) under Params::default, in
    scales.json order: t, res, coarse, stage1, tight, facing, gap, seam, near, pen, nms, icp."""
    f = lambda k, m: max(k * t, m * res)
    return np.array([t, res, f(0.15, 2.3), f(0.06, 0.9), f(0.01, 0.15), f(0.3, 1.0), f(0.03, 0.45),
                     f(0.12, 1.8), f(0.5, 4.0), f(0.06, 0.9), 0.5 * t, f(0.04, 0.6) / (0.04 * t)])


def _value_noise(p, wavelength, seed):
    """Smooth seeded noise in [-1, 1]: trilinear interpolation of hashed lattice values."""
    q = p / wavelength
    i0 = np.floor(q).astype(np.int64)
    f = q - i0
    f = f * f * (3.0 - 2.0 * f)
    out = np.zeros(len(p))
    for dx in (0, 1):
        for dy in (0, 1):
            for dz in (0, 1):
                h = (i0[:, 0] + dx) * 73856093 ^ (i0[:, 1] + dy) * 19349663 ^ (i0[:, 2] + dz) * 83492791 ^ (seed * 2654435761)
                h = (h ^ (h >> 13)) * 1274126177
                v = ((h ^ (h >> 16)) & 0xFFFF) / 32767.5 - 1.0
                w = (f[:, 0] if dx else 1 - f[:, 0]) * (f[:, 1] if dy else 1 - f[:, 1]) * (f[:, 2] if dz else 1 - f[:, 2])
                out += w * v
    return out


def _random_rotation(rng, max_angle=np.pi):
    axis = rng.normal(size=3)
    axis /= np.linalg.norm(axis)
    ang = rng.uniform(0.0, max_angle)
    k = np.array([[0, -axis[2], axis[1]], [axis[2], 0, -axis[0]], [-axis[1], axis[0], 0]])
    return np.eye(3) + np.sin(ang) * k + (1 - np.cos(ang)) * (k @ k)


def _sample_faces(v, f, faces, n, rng):
    """n area-weighted uniform points on the listed faces: (points, face ids)."""
    a, b, c = v[f[faces, 0]], v[f[faces, 1]], v[f[faces, 2]]
    area = 0.5 * np.linalg.norm(np.cross(b - a, c - a), axis=1)
    pick = rng.choice(len(faces), size=n, p=area / area.sum())
    r1 = np.sqrt(rng.uniform(size=n))[:, None]
    r2 = rng.uniform(size=n)[:, None]
    pts = (1 - r1) * a[pick] + r1 * (1 - r2) * b[pick] + r1 * r2 * c[pick]
    return pts, faces[pick], area.sum()


def vessel(t=1.0, edges_per_t=5.5, n_frag=8, seed=10, height_t=200.0, rmin_t=40.0, rmax_t=80.0,
           noise_t=0.25, pose_seed=None):
    """The vessel collection. Returns a dict with `fragments` (list of dicts), `poses` (4x4 each,
    mapping the unposed vessel frame into the fragment's own frame), `adjacent` (set of index
    pairs sharing a fracture wall), `t` and `res`."""
    rng = np.random.default_rng(seed)
    e = t / edges_per_t
    H = height_t * t
    zc = np.linspace(0.0, H, 6)
    rc = rng.uniform(rmin_t * t, rmax_t * t, size=6)
    spline = CubicSpline(zc, rc)
    nz = max(4, int(round(H / e)))
    zs = np.linspace(0.0, H, nz + 1)
    r_out = np.clip(spline(zs), rmin_t * t, rmax_t * t)
    thick = np.full(nz + 1, t)
    rim = zs > H - 3.0 * t
    thick[rim] = t + 0.4 * t * (zs[rim] - (H - 3.0 * t)) / (3.0 * t)
    r_in = r_out - thick
    nth = max(8, int(round(2 * np.pi * np.mean(0.5 * (r_out + r_in)) / e)))
    th = np.linspace(0.0, 2 * np.pi, nth, endpoint=False)
    K = max(1, int(round(t / e)))  # lattice levels through the wall

    # Cell labels: nearest seed on the mid-surface, distances perturbed by smooth noise.
    rng_seeds = np.random.default_rng(seed + 1)
    ci, cj = np.meshgrid(np.arange(nth), np.arange(nz), indexing="ij")
    cth = (ci + 0.5) * (2 * np.pi / nth)
    czz = (cj + 0.5) * (H / nz)
    crm = 0.5 * (np.interp(czz, zs, r_out) + np.interp(czz, zs, r_in))
    cell_pts = np.stack([crm * np.cos(cth), crm * np.sin(cth), czz], axis=-1).reshape(-1, 3)
    seeds = cell_pts[rng_seeds.choice(len(cell_pts), size=n_frag, replace=False)]
    dist = np.empty((n_frag, len(cell_pts)))
    for k in range(n_frag):
        dist[k] = np.linalg.norm(cell_pts - seeds[k], axis=1) + noise_t * t * _value_noise(cell_pts, 2.0 * t, seed + 2 + k)
    label = np.argmin(dist, axis=0).reshape(nth, nz).astype(np.int32)
    del dist

    # The shared lattice: node (i, j, l) -> index i*(nz+1)*(K+1) + j*(K+1) + l.
    def node(i, j, l):
        return ((i % nth) * (nz + 1) + j) * (K + 1) + l

    L = (np.arange(K + 1) / K)[None, None, :]
    rad = r_in[None, :, None] + L * thick[None, :, None]
    X = rad * np.cos(th)[:, None, None]
    Y = rad * np.sin(th)[:, None, None]
    Z = np.broadcast_to(zs[None, :, None], X.shape)
    verts = np.stack([X, Y, Z], axis=-1).reshape(-1, 3).astype(np.float32).astype(np.float64)

    tris, owner, is_frac, partner = [], [], [], []

    def add(a, b, c, lab, frac, other=-1):
        tris.append(np.stack([a, b, c], axis=-1))
        owner.append(np.broadcast_to(lab, np.shape(a)).ravel() if np.ndim(lab) else np.full(np.size(a), lab))
        is_frac.append(np.full(np.size(a), frac, dtype=bool))
        # The fragment on the other side of a fracture face (-1 on the shell).
        partner.append(np.broadcast_to(other, np.shape(a)).ravel() if np.ndim(other) else np.full(np.size(a), other))

    I, J = ci, cj
    lab = label
    # Outer shell (level K) outward, inner shell (level 0) inward.
    a, b, c, d = node(I, J, K), node(I + 1, J, K), node(I + 1, J + 1, K), node(I, J + 1, K)
    add(a, b, c, lab, False)
    add(a, c, d, lab, False)
    a, b, c, d = node(I, J, 0), node(I + 1, J, 0), node(I + 1, J + 1, 0), node(I, J + 1, 0)
    add(a, c, b, lab, False)
    add(a, d, c, lab, False)
    for l in range(K):
        # Walls at theta_{i+1} between cell (i, j) and (i+1, j).
        nb = np.roll(label, -1, axis=0)
        m = nb != label
        ii, jj = I[m], J[m]
        p0, p1, p2, p3 = node(ii + 1, jj, l), node(ii + 1, jj + 1, l), node(ii + 1, jj + 1, l + 1), node(ii + 1, jj, l + 1)
        add(p0, p1, p2, label[m], True, nb[m])
        add(p0, p2, p3, label[m], True, nb[m])
        add(p0, p2, p1, nb[m], True, label[m])
        add(p0, p3, p2, nb[m], True, label[m])
        # Walls at z_{j+1} between cell (i, j) and (i, j+1).
        up = label[:, 1:]
        m = up != label[:, :-1]
        ii, jj = I[:, :-1][m], J[:, :-1][m]
        p0, p1, p2, p3 = node(ii, jj + 1, l), node(ii, jj + 1, l + 1), node(ii + 1, jj + 1, l + 1), node(ii + 1, jj + 1, l)
        add(p0, p1, p2, label[:, :-1][m], True, up[m])
        add(p0, p2, p3, label[:, :-1][m], True, up[m])
        add(p0, p2, p1, up[m], True, label[:, :-1][m])
        add(p0, p3, p2, up[m], True, label[:, :-1][m])
        # End caps (shell): bottom faces -z, top faces +z.
        ib = np.arange(nth)
        p0, p1, p2, p3 = node(ib, 0, l), node(ib, 0, l + 1), node(ib + 1, 0, l + 1), node(ib + 1, 0, l)
        add(p0, p2, p1, label[:, 0], False)
        add(p0, p3, p2, label[:, 0], False)
        p0, p1, p2, p3 = node(ib, nz, l), node(ib, nz, l + 1), node(ib + 1, nz, l + 1), node(ib + 1, nz, l)
        add(p0, p1, p2, label[:, -1], False)
        add(p0, p2, p3, label[:, -1], False)

    tris = np.concatenate([x.reshape(-1, 3) for x in tris])
    owner = np.concatenate(owner)
    is_frac = np.concatenate(is_frac)
    partner = np.concatenate(partner).astype(np.int32)

    prng = np.random.default_rng(seed + 3 if pose_seed is None else pose_seed)
    frags, poses, adjacent = [], [], set()
    wall = (np.roll(label, -1, axis=0) != label)
    for k, l2 in zip(label[wall], np.roll(label, -1, axis=0)[wall]):
        adjacent.add((min(k, l2), max(k, l2)))
    upm = label[:, 1:] != label[:, :-1]
    for k, l2 in zip(label[:, :-1][upm], label[:, 1:][upm]):
        adjacent.add((min(k, l2), max(k, l2)))
    res = e
    for k in range(n_frag):
        sel = owner == k
        ft = tris[sel]
        used, inv = np.unique(ft, return_inverse=True)
        v = verts[used]
        f = inv.reshape(-1, 3).astype(np.uint32)
        frac = is_frac[sel]
        R = _random_rotation(prng)
        tau = prng.uniform(-50 * t, 50 * t, size=3)
        P = np.eye(4)
        P[:3, :3] = R
        P[:3, 3] = tau
        frag = _fragment_data(v, f, frac, t, np.random.default_rng(seed + 100 + k))
        frag = _posed(frag, P)
        frag["name"] = f"frag{k}"
        frag["partner"] = partner[sel]
        frags.append(frag)
        poses.append(P)
    return dict(fragments=frags, poses=poses, adjacent=adjacent, t=t, res=res)


def _macro_normals(cen, fn, area, mask, pts, r):
    """Per point, the area-weighted mean normal of the faces of `mask` whose centroid lies within r
    (normalised; zero where none is)."""
    from scipy.sparse import csr_matrix

    sel = np.nonzero(mask)[0]
    out = np.zeros((len(pts), 3))
    if len(sel) == 0 or len(pts) == 0:
        return out
    ball = cKDTree(cen[sel]).query_ball_point(pts, r, workers=-1)
    rows = np.repeat(np.arange(len(pts)), [len(b) for b in ball])
    cols = sel[np.concatenate([np.asarray(b, dtype=np.int64) for b in ball])] if len(rows) else np.zeros(0, np.int64)
    M = csr_matrix((area[cols], (rows, cols)), shape=(len(pts), len(cen)))
    out = M @ fn
    ln = np.linalg.norm(out, axis=1, keepdims=True)
    return np.where(ln > 0, out / np.maximum(ln, 1e-300), 0.0)


def _fragment_data(v, f, frac, t, rng, brk_voxel_t=BRK_VOXEL, sub_voxel_t=SUB_VOXEL):
    """R §3.6's samples of one labelled mesh. The breakline is the shell/fracture junction's
    vertices thinned on a brk_voxel_t grid with shell ring normals; with brk_voxel_t None it is every
    junction vertex (the reference takes every junction edge, fragment/breakline.rs:230-232) with
    macro normals over 0.6 t (breakline.rs's annulus, without its inner radius), and the hypothesis
    subset on a sub_voxel_t grid."""
    fi = np.arange(len(f))
    a, b, c = v[f[:, 0]], v[f[:, 1]], v[f[:, 2]]
    n = np.cross(b - a, c - a)
    ln = np.linalg.norm(n, axis=1)
    fn = n / np.maximum(ln, 1e-12)[:, None]
    frac_area = 0.5 * ln[frac].sum()
    n_pf = int(np.clip(round(FRAC_PER_T2 * frac_area / (t * t)), MIN_FRAC, MAX_FRAC))
    pf, pf_face, _ = _sample_faces(v, f, fi[frac], n_pf, rng)
    s, s_face, _ = _sample_faces(v, f, fi, SURFACE_POINTS, rng)
    # Breakline: shell-face vertices that also belong to a fracture face, thinned on a 0.5 t voxel
    # grid; the shell macro-normal is the mean shell face normal around the vertex.
    shell = ~frac
    on_frac = np.zeros(len(v), bool)
    on_frac[f[frac].ravel()] = True
    on_shell = np.zeros(len(v), bool)
    on_shell[f[shell].ravel()] = True
    vn = np.zeros_like(v)
    np.add.at(vn, f[shell].ravel(), np.repeat(fn[shell], 3, axis=0))
    brk = np.nonzero(on_frac & on_shell & (np.linalg.norm(vn, axis=1) > 0))[0]
    if brk_voxel_t is not None:
        key = np.floor(v[brk] / (brk_voxel_t * t)).astype(np.int64)
        _, first = np.unique(key, axis=0, return_index=True)
        brk = brk[np.sort(first)]
    brk_p = v[brk]
    # Frames (R §3.5.4 / R §3.6): the fracture macro-normal around the vertex, the in-plane axis
    # pointing away from the fracture (minus the fracture normal's part across the shell normal),
    # the tangent ns x f, and the dihedral angle between the two normals. Across a seam the shell
    # normals agree and the fracture normals oppose, so the dihedral angles add to 180 degrees.
    if brk_voxel_t is not None:
        brk_ns = vn[brk] / np.linalg.norm(vn[brk], axis=1, keepdims=True)
        vf = np.zeros_like(v)
        np.add.at(vf, f[frac].ravel(), np.repeat(fn[frac], 3, axis=0))
        nf = vf[brk] / np.maximum(np.linalg.norm(vf[brk], axis=1, keepdims=True), 1e-300)
    else:
        cen0 = (a + b + c) / 3.0
        area0 = 0.5 * ln
        brk_ns = _macro_normals(cen0, fn, area0, shell, brk_p, MACRO_OUTER * t)
        nf = _macro_normals(cen0, fn, area0, frac, brk_p, MACRO_OUTER * t)
        ok = (np.linalg.norm(brk_ns, axis=1) > 0.5) & (np.linalg.norm(nf, axis=1) > 0.5)
        brk, brk_p, brk_ns, nf = brk[ok], brk_p[ok], brk_ns[ok], nf[ok]
    inplane = nf - (nf * brk_ns).sum(axis=1, keepdims=True) * brk_ns
    brk_f = -inplane / np.maximum(np.linalg.norm(inplane, axis=1, keepdims=True), 1e-300)
    brk_t = np.cross(brk_ns, brk_f)
    brk_dih = np.degrees(np.arccos(np.clip((brk_ns * nf).sum(axis=1), -1.0, 1.0)))
    skey = np.floor(brk_p / (sub_voxel_t * t)).astype(np.int64)
    _, sfirst = np.unique(skey, axis=0, return_index=True)
    brk_sub = np.sort(sfirst).astype(np.int64)
    # Shell margin: shell faces whose centroid lies 0.15 t .. 0.6 t from the breakline.
    cen = (a + b + c) / 3.0
    shell_faces = fi[shell]
    dist, _ = cKDTree(v[np.nonzero(on_frac & on_shell)[0]]).query(cen[shell_faces], distance_upper_bound=1.01 * MACRO_OUTER * t, workers=-1)
    band = shell_faces[(dist >= MACRO_INNER * t) & (dist <= MACRO_OUTER * t)]
    margin_p, margin_face, _ = _sample_faces(v, f, band, MARGIN_POINTS, rng)
    return dict(V=v.astype(np.float32), F=f, frac=frac, frac_area=float(frac_area), watertight=True,
                Pf=pf, S=s, brk_P=brk_p, brk_ns=brk_ns, brk_f=brk_f, brk_t=brk_t, brk_dih=brk_dih, brk_sub=brk_sub,
                margin_p=margin_p, margin_n=fn[margin_face], face_normals=fn)


def _perlin1(x, seed):
    """Seeded 1-D gradient (Perlin) noise in about [-1, 1]: a random slope at every integer, the
    two neighbouring ramps blended by the quintic fade."""
    i0 = np.floor(x).astype(np.int64)
    f = x - i0
    lo, hi = int(i0.min()), int(i0.max()) + 1
    g = np.random.default_rng(seed).uniform(-1.0, 1.0, size=hi - lo + 1)
    a = g[i0 - lo] * f
    b = g[i0 + 1 - lo] * (f - 1.0)
    w = f * f * f * (f * (6.0 * f - 15.0) + 10.0)
    return 2.0 * (a + w * (b - a))


def _quads(G):
    """Two triangles per cell of an index grid G (m+1, n+1), wound so that the normal is
    d/d(first index) x d/d(second index)."""
    a, b, c, d = G[:-1, :-1], G[1:, :-1], G[1:, 1:], G[:-1, 1:]
    return np.concatenate([np.stack([a, b, c], -1).reshape(-1, 3), np.stack([a, c, d], -1).reshape(-1, 3)])


def slab_plate(t=1.0, edge_t=0.06, length_t=60.0, width_t=40.0, amp_t=0.3, wavelength_t=2.0, jitter_t=0.01,
               seed=0, jitter_seed=1, pose_seed=2, max_deg=30.0, max_shift_t=5.0, rounded_rims=True):
    """a plate length_t x width_t x t (t thick) split in two across its
    width by a z-extruded polyline x = length/2 + amp * perlin(y / wavelength) (seed), every vertex
    jittered by N(0, jitter) (jitter_seed; the shared fracture lattice identically on both sides),
    meshed closed at edge_t per edge (both pieces on one fracture lattice, so the fracture surfaces
    are exactly complementary), fragment B moved by a seeded rigid transform (axis uniform on the
    sphere, angle uniform in (0, max_deg], translation uniform in a ball of max_shift_t; pose_seed).
    With rounded_rims the plate's three outer sides of each piece are half-round (every rim vertex
    pushed outward by sqrt((t/2)^2 - z^2), the cut's end columns too, identically on both pieces):
    two straight square rims would fit each other face to face as exactly as the cut does, and a
    half-round rim touches another only along a line. Returns the collection dict of `vessel` (fragments with V, F, frac, samples, frames; poses;
    adjacent; t; res)."""
    L, W = length_t * t, width_t * t
    nu = max(4, int(round(0.5 * L / (edge_t * t))))
    nv = max(4, int(round(W / (edge_t * t))))
    nw = max(2, int(round(t / (edge_t * t))))
    vv = np.linspace(0.0, 1.0, nv + 1)
    ww = np.linspace(0.0, 1.0, nw + 1)
    y = -0.5 * W + W * vv
    z = -0.5 * t + t * ww
    xc = 0.5 * L + amp_t * t * _perlin1(y / (wavelength_t * t), seed)
    jrng = np.random.default_rng(jitter_seed)
    cut_jitter = jrng.normal(scale=jitter_t * t, size=(nv + 1, nw + 1, 3))
    frags = []
    for side in (0, 1):
        uu = np.linspace(0.0, 1.0, nu + 1)
        U, Vg, Wg = np.meshgrid(uu, np.arange(nv + 1), np.arange(nw + 1), indexing="ij")
        X = uu[:, None, None] * xc[None, :, None] if side == 0 else xc[None, :, None] + uu[:, None, None] * (L - xc[None, :, None])
        X = np.broadcast_to(X, U.shape)
        P = np.stack([X, np.broadcast_to(y[None, :, None], U.shape), np.broadcast_to(z[None, None, :], U.shape)], -1)
        P = P + jrng.normal(scale=jitter_t * t, size=P.shape)
        cut = nu if side == 0 else 0
        P[cut] = np.stack([np.broadcast_to(xc[:, None], (nv + 1, nw + 1)), np.broadcast_to(y[:, None], (nv + 1, nw + 1)),
                           np.broadcast_to(z[None, :], (nv + 1, nw + 1))], -1) + cut_jitter
        if rounded_rims:
            bulge = np.sqrt(np.clip(0.25 * t * t - z * z, 0.0, None))
            outer = 0 if side == 0 else nu
            P[outer, :, :, 0] += (-1.0 if side == 0 else 1.0) * bulge[None, :]
            P[:, 0, :, 1] -= bulge[None, :]
            P[:, -1, :, 1] += bulge[None, :]
        idx = np.arange(U.size).reshape(U.shape)
        parts = [_quads(idx[:, :, -1]), _quads(idx[:, :, 0].T), _quads(idx[:, -1, :].T), _quads(idx[:, 0, :]),
                 _quads(idx[-1]), _quads(idx[0].T)]
        frac_part = 4 if side == 0 else 5  # the fracture is the u = 1 face of A, the u = 0 face of B
        F = np.concatenate(parts)
        is_frac = np.concatenate([np.full(len(q), k == frac_part) for k, q in enumerate(parts)])
        used, inv = np.unique(F, return_inverse=True)
        F = inv.reshape(-1, 3).astype(np.uint32)
        V = P.reshape(-1, 3)[used]
        del P, U, Vg, Wg, X
        frag = _fragment_data(V, F, is_frac, t, np.random.default_rng(seed + 100 + side))
        frag["partner"] = np.where(is_frac, 1 - side, -1).astype(np.int32)
        frags.append(frag)
    prng = np.random.default_rng(pose_seed)
    axis = prng.normal(size=3)
    axis /= np.linalg.norm(axis)
    ang = np.radians(max_deg) * (1.0 - prng.uniform())  # in (0, max_deg]
    K = np.array([[0, -axis[2], axis[1]], [axis[2], 0, -axis[0]], [-axis[1], axis[0], 0]])
    PB = np.eye(4)
    PB[:3, :3] = np.eye(3) + np.sin(ang) * K + (1 - np.cos(ang)) * (K @ K)
    v = prng.normal(size=3)
    PB[:3, 3] = v / np.linalg.norm(v) * max_shift_t * t * prng.uniform() ** (1.0 / 3.0)
    frags[1] = dict(_posed(frags[1], PB), partner=frags[1]["partner"])
    for k, fr in enumerate(frags):
        fr["name"] = f"slab{k}"
    return dict(fragments=frags, poses=[np.eye(4), PB], adjacent={(0, 1)}, t=t, res=edge_t * t)


def _posed(frag, P):
    R, tau = P[:3, :3], P[:3, 3]
    out = dict(frag)
    out["V"] = (frag["V"].astype(np.float64) @ R.T + tau).astype(np.float32)
    for key in ("Pf", "S", "brk_P", "margin_p"):
        out[key] = frag[key] @ R.T + tau
    for key in ("brk_ns", "brk_f", "brk_t", "margin_n", "face_normals"):
        out[key] = frag[key] @ R.T
    return out


def remove_patch(frag, fraction, rng):
    """A non-watertight copy: drop `fraction` of the faces around a random shell face."""
    v, f = frag["V"].astype(np.float64), frag["F"]
    cen = v[f].mean(axis=1)
    start = cen[rng.integers(len(f))]
    order = np.argsort(np.linalg.norm(cen - start, axis=1))
    keep = np.ones(len(f), bool)
    keep[order[: int(fraction * len(f))]] = False
    out = dict(frag)
    out["F"] = f[keep]
    out["frac"] = frag["frac"][keep]
    if "partner" in frag:
        out["partner"] = frag["partner"][keep]
    out["watertight"] = False
    return out


def candidates(collection, per_pair=10, seed=20):
    """Candidate poses for every pair (a < b): the true relative pose (B's frame into A's,
    P_A . P_B^-1) and per_pair - 1 perturbations of it, from a tenth of a degree and a hundredth
    of t up to 90 degrees and 10 t. Returns (pairs [(a, b)], cand_pair, cand_T)."""
    rng = np.random.default_rng(seed)
    P = collection["poses"]
    t = collection["t"]
    n = len(P)
    angles = np.radians([0.1, 0.3, 1.0, 2.0, 5.0, 10.0, 20.0, 45.0, 90.0])
    shifts = t * np.array([0.01, 0.03, 0.1, 0.2, 0.5, 1.0, 2.0, 5.0, 10.0])
    pairs, cp, cT = [], [], []
    for a in range(n):
        for b in range(a + 1, n):
            truth = P[a] @ np.linalg.inv(P[b])
            pairs.append((a, b))
            for k in range(per_pair):
                T = truth.copy()
                if k > 0:
                    j = (k - 1) % len(angles)
                    d = np.eye(4)
                    d[:3, :3] = _random_rotation(rng, angles[j])
                    v = rng.normal(size=3)
                    d[:3, 3] = shifts[j] * v / np.linalg.norm(v)
                    T = d @ truth
                cp.append(len(pairs) - 1)
                cT.append(T)
    return pairs, np.array(cp, np.uint32), np.array(cT)


def _uniform_rotation_within(rng, max_angle):
    """A rotation uniform on SO(3) restricted to angles <= max_angle (Haar density of the angle is
    (1 - cos a) / pi; sampled by rejection)."""
    axis = rng.normal(size=3)
    axis /= np.linalg.norm(axis)
    while True:
        ang = rng.uniform(0.0, max_angle)
        if rng.uniform() * (1.0 - np.cos(max_angle)) <= 1.0 - np.cos(ang):
            break
    k = np.array([[0, -axis[2], axis[1]], [axis[2], 0, -axis[0]], [-axis[1], axis[0], 0]])
    return np.eye(3) + np.sin(ang) * k + (1 - np.cos(ang)) * (k @ k)


def _breakline(rng, n, length, t):
    """n points along a smooth space curve of the given length (a fracture rim), with the shell
    normal at each: a sum of low-frequency sines, sampled uniformly in arc parameter, with
    0.02 t of jitter across the curve."""
    freq = rng.uniform(0.5, 2.0, size=(3, 3))
    amp = rng.uniform(0.05, 0.15, size=(3, 3)) * length
    phase = rng.uniform(0.0, 2 * np.pi, size=(3, 3))
    s = np.sort(rng.uniform(0.0, 1.0, size=n))
    u = 2 * np.pi * s[:, None]
    p = np.stack([(amp[k] * np.sin(freq[k] * u + phase[k])).sum(1) for k in range(3)], axis=1)
    p[:, 0] += length * s
    dp = np.stack([(amp[k] * freq[k] * np.cos(freq[k] * u + phase[k])).sum(1) for k in range(3)], axis=1)
    dp[:, 0] += length / (2 * np.pi)
    tan = dp / np.linalg.norm(dp, axis=1, keepdims=True)
    ref = np.array([0.0, 0.0, 1.0])
    nrm = np.cross(tan, ref)
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    p += rng.normal(scale=0.02 * t, size=p.shape)
    return p, nrm


def icp_batch_workload(t=1.0, n_pairs=28, poses=64, nmin=2000, nmax=60000, length_t=60.0, seed=20, size_seed=21):
    """The batched ICP ladder's workload: n_pairs breakline pairs whose point counts per side are
    log-uniform in [nmin, nmax] (size_seed), each the same curve sampled twice with B's copy under a
    rigid pose; and `poses` initial poses per pair, the truth composed with a rotation uniform on
    SO(3) within 45 degrees and a translation in a 2 t ball (seed). Rungs: 1.0, 0.3, 0.1, 0.03 t,
    20 iterations each, point-to-point (the breakline estimator, ladder.rs:131-140)."""
    srng = np.random.default_rng(size_seed)
    rng = np.random.default_rng(seed)
    lo, hi = np.log(nmin), np.log(nmax)
    pairs = []
    for p in range(n_pairs):
        ns, nt = (int(round(np.exp(srng.uniform(lo, hi)))) for _ in range(2))
        crng = np.random.default_rng(1000 + p)
        tgt, tn = _breakline(crng, nt, length_t * t, t)
        # The source samples the same curve: re-draw parameters from the same curve seed.
        src_curve, _ = _breakline(np.random.default_rng(1000 + p), ns, length_t * t, t)
        R = _random_rotation(crng)
        tau = crng.normal(scale=5.0 * t, size=3)
        truth = np.eye(4)
        truth[:3, :3] = R
        truth[:3, 3] = tau
        inv = np.linalg.inv(truth)
        src = src_curve @ inv[:3, :3].T + inv[:3, 3]
        inits = np.empty((poses, 4, 4))
        for k in range(poses):
            d = np.eye(4)
            d[:3, :3] = _uniform_rotation_within(rng, np.radians(45.0))
            v = rng.normal(size=3)
            d[:3, 3] = v / np.linalg.norm(v) * 2.0 * t * rng.uniform() ** (1.0 / 3.0)
            inits[k] = d @ truth
        pairs.append(dict(source=src, target=tgt, target_n=tn, truth=truth, inits=inits))
    rungs = [(1.0 * t, 20), (0.3 * t, 20), (0.1 * t, 20), (0.03 * t, 20)]
    return dict(pairs=pairs, rungs=rungs, t=t)


def huber_workload(t=1.0, n_pairs=16, poses=32, nmin=10000, nmax=40000, length_t=60.0, seed=30):
    """The robust seam energy's workload: n_pairs breakline pairs with 10k-40k points per side
    (log-uniform), each the same curve sampled twice with B's copy under a rigid pose, and `poses`
    base poses per pair near the truth (rotations within 2 degrees, translations within 0.1 t), so
    that correspondences at 0.3 t are meaningful. Huber delta 0.05 t."""
    rng = np.random.default_rng(seed)
    lo, hi = np.log(nmin), np.log(nmax)
    pairs = []
    for p in range(n_pairs):
        ns, nt = (int(round(np.exp(rng.uniform(lo, hi)))) for _ in range(2))
        tgt, tn = _breakline(np.random.default_rng(5000 + p), nt, length_t * t, t)
        src_curve, _ = _breakline(np.random.default_rng(5000 + p), ns, length_t * t, t)
        crng = np.random.default_rng(7000 + p)
        truth = np.eye(4)
        truth[:3, :3] = _random_rotation(crng)
        truth[:3, 3] = crng.normal(scale=5.0 * t, size=3)
        inv = np.linalg.inv(truth)
        src = src_curve @ inv[:3, :3].T + inv[:3, 3]
        T0 = np.empty((poses, 4, 4))
        for k in range(poses):
            d = np.eye(4)
            d[:3, :3] = _uniform_rotation_within(rng, np.radians(2.0))
            v = rng.normal(size=3)
            d[:3, 3] = v / np.linalg.norm(v) * 0.1 * t * rng.uniform() ** (1.0 / 3.0)
            T0[k] = d @ truth
        pairs.append(dict(source=src, target=tgt, target_n=tn, truth=truth, T0=T0))
    return dict(pairs=pairs, delta=0.05 * t, radius=0.3 * t, t=t)


def _seam_side(frag, other, n, t, band_t, frac_share, rng):
    """n points on `frag`'s side of its seam with fragment `other`, each with its face's normal: a
    frac_share of them on the fracture faces the two share, the rest on the shell faces whose centroid
    lies within band_t of the seam's breakline (the shell/fracture junction of those faces)."""
    V = frag["V"].astype(np.float64)
    F = frag["F"]
    fn = frag["face_normals"]
    faces = np.nonzero(frag["frac"] & (frag["partner"] == other))[0]
    if len(faces) == 0:
        return None
    n_f = int(round(frac_share * n))
    pf, pff, _ = _sample_faces(V, F, faces, n_f, rng)
    on_shell = np.zeros(len(V), bool)
    on_shell[F[~frag["frac"]].ravel()] = True
    brk = np.unique(F[faces].ravel())
    brk = brk[on_shell[brk]]
    shell = np.nonzero(~frag["frac"])[0]
    cen = V[F[shell]].mean(axis=1)
    d, _ = cKDTree(V[brk]).query(cen, distance_upper_bound=1.01 * band_t * t, workers=-1)
    band = shell[d <= band_t * t]
    ps, psf, _ = _sample_faces(V, F, band, n - n_f, rng)
    return np.concatenate([pf, ps]), fn[np.concatenate([pff, psf])]


def seam_workload(col, poses=512, nmin=10000, nmax=40000, seed=30, band_t=0.6, frac_share=0.75, max_deg=1.0,
                  max_shift_t=0.1):
    """The robust seam energy's workload on a vessel collection's seams: for every adjacent pair
    (a, b), a cloud on each side of their seam (log-uniform in [nmin, nmax] points per side, a
    frac_share of them on the shared fracture faces and the rest on the shell band within band_t of
    the breakline, each point with its face's normal: fracture normals across the wall in two
    directions, shell normals radial, so together they span 3-D), B's cloud as the source and A's as
    the target. `poses` base poses spread over the pairs in turn: the true relative pose composed
    with a rotation uniform within max_deg about the target seam's centroid and a translation in a
    ball of max_shift_t. Huber delta 0.05 t, correspondence radius 0.3 t."""
    t = col["t"]
    rng = np.random.default_rng(seed)
    lo, hi = np.log(nmin), np.log(nmax)
    frags, P = col["fragments"], col["poses"]
    pairs = []
    for a, b in sorted(col["adjacent"]):
        nt, ns = (int(round(np.exp(rng.uniform(lo, hi)))) for _ in range(2))
        tgt = _seam_side(frags[a], b, nt, t, band_t, frac_share, rng)
        src = _seam_side(frags[b], a, ns, t, band_t, frac_share, rng)
        if tgt is None or src is None:
            continue
        pairs.append(dict(a=a, b=b, source=src[0], target=tgt[0], target_n=tgt[1], truth=P[a] @ np.linalg.inv(P[b])))
    cp = np.sort(np.arange(poses) % len(pairs)).astype(np.uint32)
    T0 = np.empty((poses, 4, 4))
    for i, p in enumerate(cp):
        c = pairs[p]["target"].mean(axis=0)
        R = _uniform_rotation_within(rng, np.radians(max_deg))
        v = rng.normal(size=3)
        d = np.eye(4)
        d[:3, :3] = R
        d[:3, 3] = c - R @ c + v / np.linalg.norm(v) * max_shift_t * t * rng.uniform() ** (1.0 / 3.0)
        T0[i] = d @ pairs[p]["truth"]
    return dict(pairs=pairs, cand_pair=cp, T0=T0, delta=0.05 * t, radius=0.3 * t, t=t)


def coarse_grid(centroids, spacing):
    """The voxel grid R §3.4 cleans its vote over (crates/sherd-core/src/fragment/segment.rs:
    189-260): the lowest face index of every occupied voxel of side `spacing` (voxel
    floor((c - (min - 1)) / spacing)), ascending, and each face's nearest representative (a
    KD-tree; ties may resolve otherwise than the reference's). Input for the segmentation clean-up
    on synthetic meshes."""
    from scipy.spatial import cKDTree

    c = np.asarray(centroids, dtype=np.float64)
    lo = c.min(axis=0) - 1.0
    vox = np.floor((c - lo) / spacing).astype(np.int64)
    _, first = np.unique(vox, axis=0, return_index=True)
    rep = np.sort(first)
    _, near = cKDTree(c[rep]).query(c)
    return rep.astype(np.uint32), near.astype(np.uint32)
