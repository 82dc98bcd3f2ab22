"""Oracle, R §9's full-resolution refinement. TEST INFRASTRUCTURE ONLY: tests/ import it as the
checker; nothing in the product imports, links or runs it.

A restatement of crates/sherd-core/src/refine.rs over numpy and the C oracle (tests/oracle_lib.py:
kiddo's unbounded nearest by brute force, the fused placement, icp::register):

  vertex_normals      refine.rs:118-150  Open3D's area-weighted vertex normals
  select_candidates   refine.rs:163-182  the fracture vertices of an original mesh
  ChaCha8, uniform_below, without_replacement
                      crates/sherd-core/src/rng.rs:96-156 over rand_chacha 0.10's ChaCha8Rng
                      (rand_core's seed_from_u64: PCG32 expansion of the u64 into the key)
  cap_selection       refine.rs:186-192
  fracture_cloud      refine.rs:216-227
  placed              refine.rs:99-104 (types.rs:202-214)
  refine_joins        refine.rs:270-303 with refine_one (refine.rs:306-350)

What pins each: the rng.rs tests' golden words and properties (rng.rs:183-289), refine.rs's own
tests (normals :369-400, selection :404-419, the walk :423-470 and :474-497), and the slab dump's
refine/ row (selection exact, radii exact, rungs at 0.2 deg / 0.02 t, fitness 1e-4;
crates/sherd-parity/src/stages/refine.rs:56-79). The ICP underneath is the oracle's icp_rung.
"""

import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "tests"))
import oracle_lib as O  # noqa: E402

MAX_POINTS = 150_000
SELECT_T, SELECT_RES = 0.15, 1.5
RUNGS = (0.05, 0.02)
ITERATIONS = 40

_M32 = 0xFFFFFFFF
_M64 = 0xFFFFFFFFFFFFFFFF


def vertex_normals(V, F):
    """refine.rs:118-150: per vertex the sum of its faces' (v1 - v0) x (v2 - v0), added in face
    order, then divided once by sqrt((x^2 + y^2) + z^2); (0, 0, 1) when that is not positive."""
    V = np.asarray(V, dtype=np.float64)
    F = np.asarray(F, dtype=np.int64).reshape(-1, 3)
    o, b, c = V[F[:, 0]], V[F[:, 1]], V[F[:, 2]]
    e1, e2 = b - o, c - o
    face = np.stack([e1[:, 1] * e2[:, 2] - e1[:, 2] * e2[:, 1],
                     e1[:, 2] * e2[:, 0] - e1[:, 0] * e2[:, 2],
                     e1[:, 0] * e2[:, 1] - e1[:, 1] * e2[:, 0]], axis=1)
    n = np.zeros((len(V), 3))
    # ufunc.at applies the additions one index at a time in array order: face-major, as the loop.
    np.add.at(n, F.reshape(-1), np.repeat(face, 3, axis=0))
    length = np.sqrt((n[:, 0] * n[:, 0] + n[:, 1] * n[:, 1]) + n[:, 2] * n[:, 2])
    ok = length > 0.0
    out = np.empty_like(n)
    out[ok] = n[ok] / length[ok, None]
    out[~ok] = (0.0, 0.0, 1.0)
    return out


def select_candidates(V, centroids, fracture, thick, res):
    """refine.rs:163-182: ascending indices of the vertices whose nearest centroid is a fracture
    face strictly closer than max(0.15 t, 1.5 res)."""
    if len(centroids) == 0:
        return np.zeros(0, np.int64)
    radius = max(SELECT_T * thick, SELECT_RES * res)
    j, d = O.nearest_unbounded(centroids, V)
    keep = np.asarray(fracture, dtype=bool)[j] & (d < radius)
    return np.nonzero(keep)[0]


class ChaCha8:
    """rand_chacha's ChaCha8Rng from seed_from_u64: 8 rounds, 64-bit counter, stream 0."""

    def __init__(self, seed):
        state = seed & _M64
        key = []
        for _ in range(8):
            state = (state * 6364136223846793005 + 11634580027462260723) & _M64
            xs = (((state >> 18) ^ state) >> 27) & _M32
            rot = state >> 59
            key.append(((xs >> rot) | (xs << ((32 - rot) & 31))) & _M32)
        self.key = key
        self.counter = 0
        self.buf = []

    @staticmethod
    def _rotl(x, n):
        return ((x << n) | (x >> (32 - n))) & _M32

    def _block(self):
        s = [0x61707865, 0x3320646E, 0x79622D32, 0x6B206574] + self.key + [
            self.counter & _M32, self.counter >> 32, 0, 0]
        x = list(s)

        def qr(a, b, c, d):
            x[a] = (x[a] + x[b]) & _M32
            x[d] = self._rotl(x[d] ^ x[a], 16)
            x[c] = (x[c] + x[d]) & _M32
            x[b] = self._rotl(x[b] ^ x[c], 12)
            x[a] = (x[a] + x[b]) & _M32
            x[d] = self._rotl(x[d] ^ x[a], 8)
            x[c] = (x[c] + x[d]) & _M32
            x[b] = self._rotl(x[b] ^ x[c], 7)

        for _ in range(4):
            qr(0, 4, 8, 12), qr(1, 5, 9, 13), qr(2, 6, 10, 14), qr(3, 7, 11, 15)
            qr(0, 5, 10, 15), qr(1, 6, 11, 12), qr(2, 7, 8, 13), qr(3, 4, 9, 14)
        self.counter += 1
        self.buf = [(x[i] + s[i]) & _M32 for i in range(16)][::-1]

    def next_u32(self):
        if not self.buf:
            self._block()
        return self.buf.pop()


def uniform_below(rng, n):
    """rng.rs:140-156: Lemire's multiply-shift with rejection."""
    m = rng.next_u32() * n
    low = m & _M32
    if low < n:
        threshold = ((-n) & _M32) % n
        while low < threshold:
            m = rng.next_u32() * n
            low = m & _M32
    return m >> 32


def without_replacement(n, take, rng):
    """rng.rs:127-137: a partial Fisher-Yates shuffle of 0..n, its first min(take, n) entries."""
    take = min(take, n)
    pool = list(range(n))
    for i in range(take):
        j = i + uniform_below(rng, n - i)
        pool[i], pool[j] = pool[j], pool[i]
    return pool[:take]


def cap_selection(idx, max_points, seed):
    """refine.rs:186-192."""
    idx = np.asarray(idx, dtype=np.int64)
    if len(idx) > max_points:
        take = without_replacement(len(idx), max_points, ChaCha8(seed))
        idx = idx[np.asarray(take, dtype=np.int64)]
    return idx


def fracture_cloud(V, F, centroids, fracture, thick, res, seed=0, max_points=MAX_POINTS):
    """refine.rs:216-227: (idx, points, normals)."""
    idx = cap_selection(select_candidates(V, centroids, fracture, thick, res), max_points, seed)
    V = np.asarray(V, dtype=np.float64)
    return idx, V[idx], vertex_normals(V, F)[idx]


def face_centroids(V, F):
    """geometry.rs:88-93: (a + b + c) / 3 left to right."""
    V = np.asarray(V, dtype=np.float64)
    F = np.asarray(F, dtype=np.int64).reshape(-1, 3)
    return (V[F[:, 0]] + V[F[:, 1]] + V[F[:, 2]]) / 3.0


def placed(points, normals, T):
    """refine.rs:99-104."""
    return O.place(points, T), O.place(normals, T, direction=True)


def icp_dist(t, res, k):
    """scales.rs:72-88 (Params::default icp_delta 0.04, icp_res 0.6) and :100."""
    icp = max(0.04 * t, 0.6 * res) / (0.04 * t)
    return k * t * icp


def refine_joins(pieces, poses, groups, used):
    """refine.rs:270-303. pieces: [(thick, res, (points, normals) or None)]; poses (n, 4, 4);
    groups: lists of ids; used: [(a, b)]. Returns (poses, joins in walk order)."""
    out = np.array(poses, dtype=np.float64).copy()
    joins = []
    for group in groups:
        if len(group) < 2:
            continue
        done = [group[0]]
        edges = [(a, b) for a, b in used if a in group and b in group]
        while True:
            step = next((k for k, (a, b) in enumerate(edges) if (a in done) != (b in done)), None)
            if step is None:
                break
            a, b = edges.pop(step)
            fixed, moving = (a, b) if a in done else (b, a)
            j = refine_one(pieces, out, fixed, moving)
            if j is not None:
                out[moving] = j["rungs"][1] @ out[moving]
                joins.append(j)
            done.append(moving)
    return out, joins


def refine_one(pieces, poses, fixed, moving):
    """refine.rs:306-350: two point-to-plane rungs from the identity on the placed clouds."""
    tp, sp = pieces[fixed], pieces[moving]
    if tp[2] is None or sp[2] is None:
        return None
    t, res = min(tp[0], sp[0]), max(tp[1], sp[1])
    source = O.place(sp[2][0], poses[moving])
    tgt_p, tgt_n = placed(tp[2][0], tp[2][1], poses[fixed])
    pose = np.eye(4)
    rungs, dist = [], []
    last = None
    for k in RUNGS:
        d = icp_dist(t, res, k)
        r = O.icp_rung(source, tgt_p, tgt_n, pose[None], "point_to_plane", d, ITERATIONS)[0]
        pose = r[:16].reshape(4, 4)
        rungs.append(pose)
        dist.append(d)
        last = r
    return dict(fixed=fixed, moving=moving, dist=np.array(dist), rungs=np.array(rungs), fitness=float(last[16]),
                rmse_t=float(last[17]) / t)
