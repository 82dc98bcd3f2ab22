"""Oracle, R §3.2's wall thickness. TEST INFRASTRUCTURE ONLY: tests/ import it as the checker;
nothing in the product imports, links or runs it.

A restatement of crates/sherd-core/src/fragment/thickness.rs in numpy float32:

  face_indices        thickness.rs:69-72, R §3.2's stride rule (every face below 300 000)
  percentile90        thickness.rs:267-290, np.percentile(x, 90) on a float32 array: the virtual
                      index 0.9 (n - 1) in f64, the interpolation in f32, from the upper end when
                      gamma >= 0.5
  bin_edges           thickness.rs:249-262, np.linspace(first, last, 61) in f32
  hist_mode           thickness.rs:207-246, np.histogram(x, 60, (0, p90)) with numpy's two
                      corrections at a bin edge, the centre of the first fullest bin
  thickness_from_hits thickness.rs:139-161 with looking_back (:184-203): the valid hits, the ones
                      whose face looks back along the ray (f64 normals, > 0.7), the two modes

What pins it: numpy's own np.percentile and np.histogram on random float32 arrays (the functions
above reduce to them), and the slab dump's thick.t from its own thick.t_hit / thick.prim
(bit-exact; crates/sherd-parity/src/stages/thickness.rs's injected mode).
"""

import numpy as np

RAYS = 300_000
RAY_OFFSET = 1e-3
LOOKS_BACK_COS = 0.7
MIN_HITS = 100
BINS = 60


def face_indices(n_faces, cap=RAYS):
    """The faces that cast a ray: all of them up to `cap`, else every ceil(n / cap)-th."""
    stride = max(1, -(-n_faces // cap))
    return np.arange(0, n_faces, stride, dtype=np.uint32)


def percentile90(x):
    v = np.sort(np.asarray(x, dtype=np.float32))
    n = len(v)
    if n == 1:
        return np.float32(v[0])
    vi = float(n - 1) * 0.9
    prev = np.floor(vi)
    i = int(prev)
    gamma = vi - prev
    a, b = v[i], v[min(i + 1, n - 1)]
    diff = np.float32(b - a)
    if gamma >= 0.5:
        return np.float32(b - np.float32(diff * np.float32(1.0 - gamma)))
    return np.float32(a + np.float32(diff * np.float32(gamma)))


def bin_edges(first, last):
    first, last = np.float32(first), np.float32(last)
    step = np.float32((last - first) / np.float32(BINS))
    e = np.array([np.float32(np.float32(k) * step + first) for k in range(BINS + 1)], dtype=np.float32)
    e[BINS] = last
    return e


def hist_mode(x):
    x = np.asarray(x, dtype=np.float32)
    p90 = percentile90(x)
    first, last = (np.float32(-0.5), np.float32(0.5)) if p90 == 0.0 else (np.float32(0.0), p90)
    edges = bin_edges(first, last)
    denom = np.float32(last - first)
    v = x[(x >= first) & (x <= last)]
    # Elementwise float32, the reference's per-value steps (numpy's own binning).
    i = (((v - first) / denom) * np.float32(BINS)).astype(np.int64)
    i[i == BINS] -= 1
    i -= (v < edges[i]).astype(np.int64)
    i += ((i != BINS - 1) & (v >= edges[np.minimum(i + 1, BINS)])).astype(np.int64)
    counts = np.bincount(i, minlength=BINS)
    k = int(np.argmax(counts))  # the first of the fullest
    return np.float32(np.float32(0.5) * np.float32(edges[k] + edges[k + 1]))


def looking_back(normals, idx, t_hit, prim):
    """The valid hits' distances, and those of them whose face looks back along the ray."""
    t_hit = np.asarray(t_hit, dtype=np.float32)
    prim = np.asarray(prim, dtype=np.uint32)
    ok = np.isfinite(t_hit) & (prim < len(normals))
    hit = normals[prim[ok].astype(np.int64)]
    frm = normals[np.asarray(idx)[ok].astype(np.int64)]
    dot = -((hit[:, 0] * frm[:, 0] + hit[:, 1] * frm[:, 1]) + hit[:, 2] * frm[:, 2])
    d = t_hit[ok]
    return d, d[dot > LOOKS_BACK_COS]


def thickness_from_hits(normals, idx, t_hit, prim):
    """(t, thick_mode), or None below MIN_HITS valid hits."""
    d, far = looking_back(normals, idx, t_hit, prim)
    if len(d) < MIN_HITS:
        return None
    raw = hist_mode(d)
    t = hist_mode(far) if len(far) >= MIN_HITS else raw
    return t, raw
