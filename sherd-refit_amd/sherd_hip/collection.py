"""The 40-fragment collection end to end on the device, batched against the plain schedule.

Workload: three synthetic thin-wall vessels (the vessel generator; walls t = 0.8, 1.0, 1.3 at 4, 5
and 6 edges per t; 13 + 13 + 14 fragments), two fragments made non-watertight by removing 5 % of
their faces around a random shell face, so their pairs take R §6.4's `pen_unavailable` path
(verify.rs:490-493). Every one of the 780 pairs gets `per_pair` candidate poses: the true relative
pose for fragments of one vessel (P_A P_B^-1), a random one across vessels, perturbed.

Pipeline per candidate (R §5.4's breakline ladder, then the robust refinement, then R §6):
  1. ICP ladder on the breaklines, point-to-point, rungs 0.2 t and 0.08 t (stretched by the pair's
     Scales::icp_dist, ladder.rs:116), 20 iterations each;
  2. Huber LM on the same breaklines (target shell normals), radius icp_dist(0.08), delta 0.04;
  3. verify::verify + accept with the FP64 retest of degenerate parity hits (pen threshold
     max(0.06 t, 0.9 res)).

Batched: one ladder call, one LM batch, one verification batch over all pairs. Plain: the
reference's schedule, one call per pair and rung (ICP), per pair (LM) and per candidate (R §6).
The two give identical results.
"""

import time

import numpy as np

from . import synth

VESSELS = [(0.8, 4.0, 13, 40), (1.0, 5.0, 13, 41), (1.3, 6.0, 14, 42)]  # (t, edges per t, fragments, seed)
OPEN_FRAGMENTS = (7, 23)


def build(height_t=200.0, vessels=VESSELS, open_fragments=OPEN_FRAGMENTS, **vessel_kw):
    """The collection: `vessels` (t, edges per t, fragments, seed) from the vessel generator (extra
    keywords go to it), the fragments of `open_fragments` made non-watertight."""
    frags, poses, vessel_of, t_of, res_of = [], [], [], [], []
    adjacent = set()
    for v, (t, e, n, seed) in enumerate(vessels):
        col = synth.vessel(t=t, edges_per_t=e, n_frag=n, seed=seed, height_t=height_t, **vessel_kw)
        base = len(frags)
        for i, fr in enumerate(col["fragments"]):
            frags.append(fr)
            poses.append(col["poses"][i])
            vessel_of.append(v)
            t_of.append(col["t"])
            res_of.append(col["res"])
        for a, b in col["adjacent"]:
            adjacent.add((base + a, base + b))
    rng = np.random.default_rng(43)
    for k in open_fragments:
        frags[k] = synth.remove_patch(frags[k], 0.05, rng)
    return dict(fragments=frags, poses=poses, vessel_of=vessel_of, t_of=t_of, res_of=res_of, adjacent=adjacent)


def candidates(coll, per_pair=4, seed=44):
    rng = np.random.default_rng(seed)
    P, V = coll["poses"], coll["vessel_of"]
    n = len(P)
    angles = np.radians([1.0, 5.0, 20.0])
    shifts = np.array([0.1, 0.5, 2.0])
    pairs, cp, cT = [], [], []
    for a in range(n):
        for b in range(a + 1, n):
            if V[a] == V[b]:
                base = P[a] @ np.linalg.inv(P[b])
            else:
                base = np.eye(4)
                base[:3, :3] = synth._random_rotation(rng)
                base[:3, 3] = rng.normal(scale=20.0, size=3)
            pairs.append((a, b))
            for k in range(per_pair):
                T = base.copy()
                if k > 0:
                    j = (k - 1) % len(angles)
                    d = np.eye(4)
                    d[:3, :3] = synth._random_rotation(rng, angles[j])
                    v = rng.normal(size=3)
                    d[:3, 3] = shifts[j] * coll["t_of"][a] * v / np.linalg.norm(v)
                    T = d @ base
                cp.append(len(pairs) - 1)
                cT.append(T)
    return pairs, np.array(cp, np.uint32), np.array(cT)


def pair_scales(coll, pairs):
    return [synth.scales_for_pair(min(coll["t_of"][a], coll["t_of"][b]), max(coll["res_of"][a], coll["res_of"][b]))
            for a, b in pairs]


class Device:
    """The collection resident on the device: scenes, fragments, breakline clouds, and the pair
    tables of the batch calls (built once per pair list, as a host keeps its pair set)."""

    def __init__(self, ex, coll):
        self.ex = ex
        self._tables = {}
        self.scenes, self.frags, self.brk = [], [], []
        for fr in coll["fragments"]:
            fs, ms = ex.scene(fr["V"], fr["F"], fr["frac"]), ex.scene(fr["V"], fr["F"])
            self.scenes.append((fs, ms))
            self.frags.append(ex.fragment(fs, ms, fr["watertight"], fr["frac_area"], fr["Pf"], fr["S"], fr["brk_P"],
                                          fr["brk_ns"], fr["margin_p"], fr["margin_n"]))
            self.brk.append(ex.cloud(fr["brk_P"], fr["brk_ns"]))

    def tables(self, pairs, scales):
        """The ladder's pair table per rung, the LM's and verification's, for these pairs and scales
        (kept with the lists they were built from)."""
        key = (id(pairs), id(scales))
        if key not in self._tables:
            import sherd_hip as H

            ladder = [H.IcpPairs([(self.brk[b], self.brk[a], k * sc[0] * sc[11]) for (a, b), sc in zip(pairs, scales)])
                      for k in RUNGS]
            lm = H.IcpPairs([(self.brk[b], self.brk[a], RUNGS[-1] * sc[0] * sc[11]) for (a, b), sc in zip(pairs, scales)])
            ver = H.VerifyPairs([(self.frags[a], self.frags[b], sc) for (a, b), sc in zip(pairs, scales)])
            self._tables[key] = (pairs, scales, ladder, lm, ver)
        return self._tables[key][2:]

    def close(self):
        self._tables.clear()
        for c in self.brk:
            c.close()
        for f in self.frags:
            f.close()
        for a, b in self.scenes:
            a.close()
            b.close()


RUNGS = (0.2, 0.08)
ITERATIONS = 20
LM_DELTA, LM_ITERATIONS = 0.04, 10


def run_batched(dev, coll, pairs, cand_pair, cand_T, scales):
    ex = dev.ex
    t0 = time.perf_counter()
    ladder_pairs, lm_pairs, vpairs = dev.tables(pairs, scales)
    reg = ex.icp_batch_ladder(ladder_pairs, [("point_to_point", ITERATIONS)] * len(RUNGS), cand_pair, cand_T)
    T1 = reg[:, -1, :16].reshape(-1, 4, 4)
    t_icp = time.perf_counter() - t0
    hb = ex.huber(lm_pairs, cand_pair, T1)
    T2, E, steps = hb.lm(LM_DELTA, LM_ITERATIONS, 32)
    hb.close()
    t_lm = time.perf_counter() - t0 - t_icp
    scores, accepted = ex.verify_batch(vpairs, cand_pair, T2, full=True, retest=True)
    el = time.perf_counter() - t0
    return dict(reg=reg, T_lm=T2, E=E, scores=scores, accepted=accepted, seconds=el, icp_s=t_icp, lm_s=t_lm,
                verify_s=el - t_icp - t_lm)


def run_plain(dev, coll, pairs, cand_pair, cand_T, scales):
    ex = dev.ex
    fr = coll["fragments"]
    t0 = time.perf_counter()
    n = len(cand_T)
    reg = np.empty((n, len(RUNGS), 21))
    for p, ((a, b), sc) in enumerate(zip(pairs, scales)):
        sel = np.nonzero(cand_pair == p)[0]
        T = cand_T[sel]
        for r, k in enumerate(RUNGS):
            o = ex.icp_rung(fr[b]["brk_P"], fr[a]["brk_P"], fr[a]["brk_ns"], T, 0, k * sc[0] * sc[11], ITERATIONS)
            reg[sel, r] = o
            T = o[:, :16].reshape(-1, 4, 4)
    t_icp = time.perf_counter() - t0
    T2 = np.empty((n, 4, 4))
    E = np.empty(n)
    for p, ((a, b), sc) in enumerate(zip(pairs, scales)):
        sel = np.nonzero(cand_pair == p)[0]
        hb = ex.huber([(dev.brk[b], dev.brk[a], RUNGS[-1] * sc[0] * sc[11])], np.zeros(len(sel), np.uint32),
                      reg[sel, -1, :16].reshape(-1, 4, 4))
        T2[sel], E[sel], _ = hb.lm(LM_DELTA, LM_ITERATIONS, 32)
        hb.close()
    t_lm = time.perf_counter() - t0 - t_icp
    vpairs = dev.tables(pairs, scales)[2]
    scores = np.empty((n, 20))
    accepted = np.empty(n, bool)
    for i in range(n):
        s1, a1 = ex.verify_batch(vpairs, cand_pair[i:i + 1], T2[i:i + 1], full=True, retest=True)
        scores[i], accepted[i] = s1[0], a1[0]
    el = time.perf_counter() - t0
    return dict(reg=reg, T_lm=T2, E=E, scores=scores, accepted=accepted, seconds=el, icp_s=t_icp, lm_s=t_lm,
                verify_s=el - t_icp - t_lm)
