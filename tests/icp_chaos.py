"""Which ICP candidates' ladders are a function of their input at double precision (test helper).

The reference's GPU cross-check (crates/sherd-gpu/src/crosscheck.rs:90-215, `Column` and
`determined_batch`) excludes from a tolerance row the candidates whose ladder is chaotic: it
re-climbs the ladder on the CPU from the twelve initial poses one ULP away from the candidate's
own (each entry of the 3 x 4 block moved to its next representable neighbour, one at a time), and
calls a candidate determined when all twelve answers land within the row's tolerance of the
unperturbed one. An excluded candidate is counted and reported, never silently dropped. This is
that probe over the oracle's ladder (the checker, never the thing measured).
"""

import numpy as np

import oracle_lib as O


def ladder(rungs, inits):
    """The oracle's ladder: rungs = [(src, tgt_p, tgt_n, estimation, distance, max_iteration)], each
    rung from the previous rung's poses. Returns the final poses (k, 4, 4)."""
    T = np.asarray(inits, np.float64).reshape(-1, 4, 4)
    for src, tp, tn, est, dist, iters in rungs:
        T = O.icp_rung(src, tp, tn, T, est, dist, iters)[:, :16].reshape(-1, 4, 4)
    return T


def ulp_neighbours(T):
    """The twelve poses one ULP from T: each entry of the 3 x 4 block moved to its next
    representable value toward +inf (crosscheck.rs:196-199)."""
    out = np.repeat(np.asarray(T, np.float64).reshape(1, 4, 4), 12, axis=0)
    for e in range(12):
        r, c = divmod(e, 4)
        out[e, r, c] = np.nextafter(out[e, r, c], np.inf)
    return out


def gap(a, b):
    """(rotation angle in rad from |Ra Rb^T - I|_F, translation distance) between pose stacks."""
    dr = a[..., :3, :3] @ np.swapaxes(b[..., :3, :3], -1, -2) - np.eye(3)
    ang = 2.0 * np.arcsin(np.clip(np.linalg.norm(dr, axis=(-2, -1)) / (2.0 * np.sqrt(2.0)), 0, 1))
    return ang, np.linalg.norm(a[..., :3, 3] - b[..., :3, 3], axis=-1)


def determined(rungs, inits, rad_tol=1e-6, trans_tol=1e-5):
    """(mask of determined candidates, the oracle's unperturbed final poses)."""
    inits = np.asarray(inits, np.float64).reshape(-1, 4, 4)
    base = ladder(rungs, inits)
    probe = np.concatenate([ulp_neighbours(T) for T in inits])
    moved = ladder(rungs, probe).reshape(len(inits), 12, 4, 4)
    ang, tr = gap(moved, base[:, None])
    return (ang.max(axis=1) <= rad_tol) & (tr.max(axis=1) <= trans_tol), base
