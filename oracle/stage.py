"""Oracle, R §5.1-5.5 for one pair. TEST INFRASTRUCTURE ONLY: tests/ import it as the checker;
nothing in the product imports, links or runs it.

A restatement over numpy and the C oracle (tests/oracle_lib.py: coarse scores, ICP) of:

  build_with, poses     crates/sherd-core/src/matching/hypotheses.rs:149-212
  order_by_score        crates/sherd-core/src/matching/nms.rs:50-58
  nms                   nms.rs:62-104
  stage1                crates/sherd-core/src/matching/pair.rs:188-264 (suppress, stage1,
                        suppress_stage1) with ladder.rs:51-61, 131-139, 177-210

Pinned by nms.rs's own tests (:121-202), the slab dump's hyp.pa / hyp.pb (the dihedral filter's
pairs, exactly), nms1.kept and nms2.kept from the dump's own walk orders (exactly), coarse.cs and
s1.score (bit-exact through the C oracle) and s1.T (at the reference's ladder tolerance).
"""

import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "tests"))
import oracle_lib as O  # noqa: E402

ROT_TOL = 2.9
COARSE_ORDER_LIMIT = 5000
COARSE_FLOOR = 0.1
STAGE1_FLOOR = 0.05
STAGE1_RUNGS = (0.2, 0.08)
STAGE1_ITERATIONS = 20
DIHEDRAL_TOL = 25.0


def build_with(dih_a, dih_b, ia, ib, tol=DIHEDRAL_TOL):
    """hypotheses.rs:149-165: (pa, pb), positions into ia / ib, row-major over (ia, ib)."""
    da = np.asarray(dih_a, dtype=np.float64)[np.asarray(ia)]
    db = np.asarray(dih_b, dtype=np.float64)[np.asarray(ib)]
    ok = np.abs(da[:, None] + db[None, :] - 180.0) < tol
    pa, pb = np.nonzero(ok)
    return pa.astype(np.int64), pb.astype(np.int64)


def poses(a, b, ia, ib, pa, pb):
    """hypotheses.rs:172-212: (k, 12) = R row-major, then tau, in the reference's sums."""
    ka, kb = np.asarray(ia)[pa], np.asarray(ib)[pb]
    A = [a["t"][ka], a["ns"][ka], a["f"][ka]]
    B = [b["t"][kb], b["ns"][kb], b["f"][kb]]
    k = len(pa)
    out = np.empty((k, 12))
    for i in range(3):
        for c in range(3):
            out[:, 3 * i + c] = (A[0][:, i] * -B[0][:, c] + A[1][:, i] * B[1][:, c]) + A[2][:, i] * -B[2][:, c]
    pA, pB = a["p"][ka], b["p"][kb]
    for i in range(3):
        out[:, 9 + i] = pA[:, i] - ((out[:, 3 * i] * pB[:, 0] + out[:, 3 * i + 1] * pB[:, 1]) + out[:, 3 * i + 2] * pB[:, 2])
    return out


def order_by_score(score, limit=None):
    """nms.rs:50-58: descending, ties by ascending index, truncated."""
    score = np.asarray(score, dtype=np.float64)
    order = np.lexsort((np.arange(len(score)), -score)).astype(np.int64)
    return order if limit is None else order[:limit]


def same_candidate(x, y, tol):
    """nms.rs:75-78 and 97-104 for two 12-double poses (x the walked one, y the kept one)."""
    d = x[9:12] - y[9:12]
    if not (np.sqrt((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]) < tol):
        return False
    tr = 0.0
    for i in range(3):
        tr += (x[i] * y[i] + x[3 + i] * y[3 + i]) + x[6 + i] * y[6 + i]
    return tr > ROT_TOL


def nms(order, pose12, score, tol, topk, floor):
    """nms.rs:62-84."""
    kept = []
    for h in order:
        if score[h] < floor:
            break
        if not any(same_candidate(pose12[h], pose12[i], tol) for i in kept):
            kept.append(int(h))
        if len(kept) >= topk:
            break
    return kept


def pose12_of(T):
    T = np.asarray(T, dtype=np.float64).reshape(-1, 4, 4)
    return np.concatenate([T[:, :3, :3].reshape(-1, 9), T[:, :3, 3]], axis=1)


def homogeneous(p12):
    T = np.tile(np.eye(4), (len(p12), 1, 1))
    T[:, :3, :3] = p12[:, :9].reshape(-1, 3, 3)
    T[:, :3, 3] = p12[:, 9:]
    return T


def stage1(a, b, ia, ib, probe_idx, s1_source, s1_target, scales, n1=250, n2=10, tol=DIHEDRAL_TOL):
    """pair.rs:188-264 up to the stage-2 inits. a / b: dicts of the frames (p, ns, f, t, dih, sub).
    Returns a dict with the hypotheses, scores, kept lists, stage-1 transforms and scores."""
    pa, pb = build_with(a["dih"], b["dih"], ia, ib, tol)
    hyp = poses(a, b, ia, ib, pa, pb)
    cs, cnt = O.coarse_scores(a["p"], a["ns"], b["p"][probe_idx], b["ns"][probe_idx], hyp, scales[2])
    kept1 = nms(order_by_score(cs, COARSE_ORDER_LIMIT), hyp, cs, scales[10], n1, COARSE_FLOOR)
    T = s1_ladder(homogeneous(hyp[kept1]), s1_source, s1_target, scales)
    sub = np.asarray(b["sub"])
    s1, _ = O.coarse_scores(a["p"], a["ns"], b["p"][sub], b["ns"][sub], pose12_of(T), scales[3])
    kept2 = nms(order_by_score(s1), pose12_of(T), s1, scales[10], n2, STAGE1_FLOOR)
    return dict(pa=pa, pb=pb, hyp=hyp, cs=cs, kept1=kept1, T=T, s1=s1, kept2=kept2)


def s1_ladder(inits, source, target, scales):
    """ladder.rs:131-139: two point-to-point rungs, 20 iterations each, the last rung's poses."""
    T = inits
    for k in STAGE1_RUNGS:
        T = O.icp_rung(source, target, None, T, "point_to_point", k * scales[0] * scales[11],
                       STAGE1_ITERATIONS)[:, :16].reshape(-1, 4, 4)
    return T
