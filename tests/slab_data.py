"""The slab golden vectors (tests/golden/slab, written by tests/golden/make_slab.py) and the inputs
the reference derives from them, with the reference's own arithmetic.

Used by the parity tests on both sides: the oracle against the reference's dump, and the HIP
path against the oracle on the same inputs.
"""

import functools
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden", "slab")


@functools.lru_cache(maxsize=None)
def piece(name):
    with np.load(os.path.join(GOLD, name + ".npz")) as z:
        return {k: z[k] for k in z.files}


@functools.lru_cache(maxsize=None)
def pair():
    with np.load(os.path.join(GOLD, "pair.npz")) as z:
        d = {k: z[k] for k in z.files}
    with open(os.path.join(GOLD, "s2_scores.json")) as fh:
        d["s2_scores"] = json.load(fh)
    with open(os.path.join(GOLD, "ground_truth.json")) as fh:
        d["ground_truth"] = json.load(fh)
    return d


def hypothesis_poses(a, b, ia, ib, pa, pb):
    """hypotheses.rs:187-212 `pose`: R = RA.RB^T with the tangent and in-plane axis of B negated,
    tau = P_A - R.P_B, every three-term sum in j = 0, 1, 2 order. Returns (k, 12) = R then tau."""
    ka = ia[pa]
    kb = ib[pb]
    axes_a = [a["brk_t"][ka], a["brk_ns"][ka], a["brk_f"][ka]]
    axes_b = [b["brk_t"][kb], b["brk_ns"][kb], b["brk_f"][kb]]
    k = len(pa)
    rot = np.empty((k, 3, 3))
    for i in range(3):
        for c in range(3):
            rot[:, i, c] = (axes_a[0][:, i] * -axes_b[0][:, c] + axes_a[1][:, i] * axes_b[1][:, c]) + axes_a[2][
                :, i
            ] * -axes_b[2][:, c]
    pa_pts = a["brk_P"][ka]
    pb_pts = b["brk_P"][kb]
    tau = np.empty((k, 3))
    for i in range(3):
        tau[:, i] = pa_pts[:, i] - ((rot[:, i, 0] * pb_pts[:, 0] + rot[:, i, 1] * pb_pts[:, 1]) + rot[:, i, 2] * pb_pts[:, 2])
    return np.concatenate([rot.reshape(k, 9), tau], axis=1)


def coarse_inputs():
    """The injected coarse row (crates/sherd-parity/src/stages/coarse.rs:70-110)."""
    a, b, p = piece("pieceA"), piece("pieceB"), pair()
    poses = hypothesis_poses(a, b, p["hyp_ia"], p["hyp_ib"], p["hyp_pa"], p["hyp_pb"])
    idx = p["coarse_idx"]
    return dict(
        target_p=a["brk_P"], target_n=a["brk_ns"], probe_p=b["brk_P"][idx], probe_n=b["brk_ns"][idx],
        poses=poses, radius=p["scales"][2], expected=p["coarse_cs"],
    )


def face_geometry(v, f):
    """mesh/geometry.rs:70-96 in f64: unit normals (floored), areas, centroids."""
    f = f.astype(np.int64)
    a, b, c = v[f[:, 0]], v[f[:, 1]], v[f[:, 2]]
    e1 = b - a
    e2 = c - a
    n0 = e1[:, 1] * e2[:, 2] - e1[:, 2] * e2[:, 1]
    n1 = e1[:, 2] * e2[:, 0] - e1[:, 0] * e2[:, 2]
    n2 = e1[:, 0] * e2[:, 1] - e1[:, 1] * e2[:, 0]
    length = np.sqrt((n0 * n0 + n1 * n1) + n2 * n2)
    den = np.maximum(length, 1e-12)
    normals = np.stack([n0 / den, n1 / den, n2 / den], axis=1)
    centroids = ((a + b) + c) / 3.0
    return normals, 0.5 * length, centroids


def thickness_rays(name):
    """thickness.rs:115-132 cast_inward_rays over every face (face_indices stride 1 below 300 000
    faces): origin = (c - n*1e-3) narrowed to f32, direction = -n narrowed."""
    pc = piece(name)
    n, _, c = face_geometry(pc["V0"], pc["F"])
    d = -n
    origin = (c + d * 1e-3).astype(np.float32)
    return origin, d.astype(np.float32)


def stage2_candidates():
    return pair()["s2_T_frac2"]
