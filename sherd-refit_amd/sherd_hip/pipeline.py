"""One pair through the whole path on the device: segmentation rays, R §3.6's sampling, R §5.1-5.5's
match, a stage-2 ladder and R §6's verification (BASELINE configuration 1's "full pipeline").

Host code: the data flow between the entry points and R §3.6's sampling of the labelled meshes
(synth's numpy restatement of the samples R §6 reads). Everything per point, per ray and per
candidate runs on the device:
  1. the per-vertex inward rays and the face labels they give (sherd_vertex_votes), cleaned by
     R §3.4.4-3.4.8 (sherd_segment_cleanup);
  2. the fragments' sample sets from those labels (Pf, S, the breakline with its frames, margins);
  3. R §5.1-5.5 for the pair, A fixed and B moving (sherd_match_stage1: hypotheses, coarse scores,
     suppression, the breakline ladder, the re-score, the second suppression);
  4. a stage-2 ladder from the kept poses on the fracture samples (point-to-point, two rungs);
  5. verify::verify + accept of every stage-2 pose (sherd_verify_batch), ranked by R §5.7.
"""

import time

import numpy as np

from . import SCORE_KEYS, synth


def _segment(ex, fr, t, cleanup=True):
    """The rays' fracture mask, then (cleanup) R §3.4.4-3.4.8 over it on the t/8 voxel grid of the
    face centroids, as segment::run follows the vote (segment.rs:544-635)."""
    V64 = fr["V"].astype(np.float64)
    scene = ex.scene(fr["V"], fr["F"])
    good, frac = ex.vertex_votes(scene, V64, fr["F"], t)
    if cleanup:
        F = fr["F"]
        cen = ((V64[F[:, 0]] + V64[F[:, 1]]) + V64[F[:, 2]]) / 3.0
        rep, near = synth.coarse_grid(cen, t / 8.0)
        frac = ex.segment_cleanup(V64, F, t, frac, rep, near)
    return scene, good, frac.astype(bool)


def run_pair(ex, fa, fb, t, res, seed=0, probe=60, rungs=(0.1, 0.04), iterations=30, cleanup=True, keep=5):
    """A (fa) fixed, B (fb) moving, both meshes in their own frames (f32 V, u32 F). Returns a dict of
    the labels, the stage-1 and stage-2 poses, the scores and verdicts, `order` (the kept candidates'
    indices, best first by R §5.7's seam * tight) and per-step seconds. The
    breakline is every shell/fracture junction vertex with macro normals over 0.6 t, and the
    hypothesis subset its 0.5 t voxel representatives, as the reference builds them
    (fragment/breakline.rs:230-286)."""
    times = {}
    t0 = time.perf_counter()
    sa, good_a, frac_a = _segment(ex, fa, t, cleanup)
    sb, good_b, frac_b = _segment(ex, fb, t, cleanup)
    times["segment"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    rng = np.random.default_rng(seed)
    kw = dict(brk_voxel_t=None, sub_voxel_t=0.5)
    da = synth._fragment_data(fa["V"].astype(np.float64), fa["F"], frac_a, t, np.random.default_rng(seed + 1), **kw)
    db = synth._fragment_data(fb["V"].astype(np.float64), fb["F"], frac_b, t, np.random.default_rng(seed + 2), **kw)
    times["sample"] = time.perf_counter() - t0
    scales = synth.scales_for_pair(t, res)
    t0 = time.perf_counter()
    handles = []
    frames = [ex.frames(d["brk_P"], d["brk_ns"], d["brk_f"], d["brk_t"], d["brk_dih"], d["brk_sub"]) for d in (da, db)]
    s1_src = ex.cloud(db["brk_P"][db["brk_sub"]])
    s1_tgt = ex.cloud(da["brk_P"], da["brk_ns"])
    pick = np.sort(rng.choice(db["brk_sub"], size=min(probe, len(db["brk_sub"])), replace=False))
    m = ex.match_stage1([dict(a=frames[0], b=frames[1], probe_p=db["brk_P"][pick], probe_n=db["brk_ns"][pick],
                              s1_source=s1_src, s1_target=s1_tgt, scales=scales)])[0]
    times["match"] = time.perf_counter() - t0
    handles += frames + [s1_src, s1_tgt]
    t0 = time.perf_counter()
    inits = m["s1_T"][m["kept2"]] if len(m["kept2"]) else np.eye(4)[None]
    src, tgt = ex.cloud(db["Pf"]), ex.cloud(da["Pf"])
    handles += [src, tgt]
    reg = ex.icp_batch_ladder([[(src, tgt, k * scales[0] * scales[11])] for k in rungs],
                              [("point_to_point", iterations)] * len(rungs), np.zeros(len(inits), np.uint32), inits)
    T2 = reg[:, -1, :16].reshape(-1, 4, 4)
    times["stage2"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    fsa, fsb = ex.scene(fa["V"], fa["F"], frac_a), ex.scene(fb["V"], fb["F"], frac_b)
    ha = ex.fragment(fsa, sa, True, da["frac_area"], da["Pf"], da["S"], da["brk_P"], da["brk_ns"], da["margin_p"],
                     da["margin_n"])
    hb = ex.fragment(fsb, sb, True, db["frac_area"], db["Pf"], db["S"], db["brk_P"], db["brk_ns"], db["margin_p"],
                     db["margin_n"])
    handles += [fsa, fsb, ha, hb]
    scores, accepted = ex.verify_batch([(ha, hb, scales)], np.zeros(len(T2), np.uint32), T2, retest=True)
    times["verify"] = time.perf_counter() - t0
    # R §5.7 (matching/pair.rs:441-456): seam * tight descending, stable, the first `keep` kept.
    key = scores[:, SCORE_KEYS.index("seam")] * scores[:, SCORE_KEYS.index("tight")]
    order = np.argsort(-key, kind="stable")[:keep]
    out = dict(good=(good_a, good_b), frac=(frac_a, frac_b), data=(da, db), scales=scales, match=m, stage2=reg,
               T=T2, scores=scores, accepted=accepted, order=order, seconds=times)
    for h in handles[::-1]:
        h.close()
    sa.close()
    sb.close()
    return out
