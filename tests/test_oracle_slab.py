"""The oracle against the reference's slab fixture dump (CPU).

The rows and tolerances are the reference's own parity harness rows (crates/sherd-parity/src/
stages/*.rs, D §10.2): `coarse` bit-exact, `stage1`/`stage2` within 0.05 deg / 0.01 t, `verify`
within tight 0.01, gap 0.002 t, seam 0.34 t, cont 0.005 t, cont_n 0.01, pen 0.0005 with
`accepted` identical. The thickness rays are checked against the Embree first hits the dump
holds (the reference's own parry path was measured at |dt| <= 1.2e-3 against them, bvh.rs:13-15).
"""

import numpy as np
import pytest

import oracle_lib as O
import slab_data as S


def sides():
    out = []
    keep = []
    for name in ("pieceA", "pieceB"):
        pc = S.piece(name)
        v32 = pc["V"].astype(np.float32)
        frac = O.Scene(v32, pc["F"], pc["frac"])
        mesh = O.Scene(v32, pc["F"])
        keep += [frac, mesh]
        out.append(O.SideData(frac, mesh, pc["watertight"], pc["frac_area"], pc["Pf"], pc["S"],
                              pc["brk_P"], pc["brk_ns"], pc["margin_p"], pc["margin_n"]))
    return out[0], out[1], keep


def pose_gap(a, b, t):
    """Rotation gap in degrees (from |Ra Rb^T - I|_F / sqrt 2, exact to rounding for small angles,
    where an arccos of the trace bottoms out near 5e-6 degrees) and translation gap in t."""
    dr = a[:, :3, :3] @ np.transpose(b[:, :3, :3], (0, 2, 1)) - np.eye(3)
    ang = 2.0 * np.arcsin(np.clip(np.linalg.norm(dr, axis=(1, 2)) / (2.0 * np.sqrt(2.0)), 0, 1))
    return np.degrees(ang), np.linalg.norm(a[:, :3, 3] - b[:, :3, 3], axis=1) / t


def test_coarse_scores_are_bit_exact():
    ci = S.coarse_inputs()
    s, _ = O.coarse_scores(ci["target_p"], ci["target_n"], ci["probe_p"], ci["probe_n"], ci["poses"],
                           ci["radius"])
    assert len(s) == 16337
    assert np.array_equal(s.view(np.uint64), ci["expected"].view(np.uint64))


def test_verify_scores_within_the_reference_tolerances():
    a, b, _keep = sides()
    p = S.pair()
    sc = p["scales"]
    tol = dict(tight=0.01, gap=0.002, seam=0.34, cont=0.005, cont_n=0.01, pen=0.0005)
    for i, T in enumerate(p["s2_T_frac2"]):
        s = O.verify(a, b, T, sc)
        theirs = p["s2_scores"][i]
        for k, tl in tol.items():
            assert abs(s[O.OS_KEYS.index(k)] - theirs[k]) <= tl, (i, k, s[O.OS_KEYS.index(k)], theirs[k])
        assert O.accept(s, sc) == bool(p["s2_accepted"][i]), i
    assert int(np.sum(p["s2_accepted"])) == 2


def test_stage1_ladder_reaches_the_references_poses():
    a, b, p = S.piece("pieceA"), S.piece("pieceB"), S.pair()
    sc = p["scales"]
    poses = S.hypothesis_poses(a, b, p["hyp_ia"], p["hyp_ib"], p["hyp_pa"], p["hyp_pb"])
    inits = np.tile(np.eye(4), (len(p["nms1_kept"]), 1, 1))
    inits[:, :3, :3] = poses[p["nms1_kept"], :9].reshape(-1, 3, 3)
    inits[:, :3, 3] = poses[p["nms1_kept"], 9:]
    T = inits
    for k in (0.2, 0.08):  # ladder.rs:51-53
        T = O.icp_rung(b["brk_P"][b["brk_sub"]], a["brk_P"], a["brk_ns"], T, 0, k * sc[0] * sc[11], 20)[:, :16]
        T = T.reshape(-1, 4, 4)
    ang, tr = pose_gap(T, p["s1_T"], sc[0])
    assert ang.max() < 0.05 and tr.max() < 0.01
    assert np.median(ang) < 1e-6 and np.median(tr) < 1e-9
    # R §5.4's re-score of the reference's own stage-1 poses is bit-exact.
    s, _ = O.coarse_scores(a["brk_P"], a["brk_ns"], b["brk_P"][b["brk_sub"]], b["brk_ns"][b["brk_sub"]],
                           p["s1_T"], sc[3])
    assert np.array_equal(s, p["s1_score"])


def reg_cloud(pc, reg_points=6000):
    """samples.rs:601-615 registration_split, then pc_reg = Pf prefix + margin prefix."""
    nf, nm = len(pc["Pf"]), len(pc["margin_p"])
    if nf + nm <= reg_points:
        tf, tm = nf, nm
    else:
        tf = min(max(int(np.round(nf * reg_points / (nf + nm))), 1), nf)
        tm = min(max(reg_points - tf, 0), nm)
    return (np.concatenate([pc["Pf"][:tf], pc["margin_p"][:tm]]),
            np.concatenate([pc["nf"][:tf], pc["margin_n"][:tm]]))


def stage2_rungs():
    a, b = S.piece("pieceA"), S.piece("pieceB")
    ra, rna = reg_cloud(a)
    rb, _ = reg_cloud(b)
    # ladder.rs:142-160: two pc_reg rungs then two pc_frac rungs, point-to-plane, 30 iterations.
    return [("s2_T_reg1", rb, ra, rna, 0.2), ("s2_T_reg2", rb, ra, rna, 0.08),
            ("s2_T_frac1", b["Pf"], a["Pf"], a["nf"], 0.08), ("s2_T_frac2", b["Pf"], a["Pf"], a["nf"], 0.04)]


def test_stage2_ladder_reaches_the_references_poses():
    p = S.pair()
    sc = p["scales"]
    T = p["s1_T"][p["nms2_kept"]]
    for name, src, tp, tn, k in stage2_rungs():
        T = O.icp_rung(src, tp, tn, T, 1, k * sc[0] * sc[11], 30)[:, :16].reshape(-1, 4, 4)
        ang, tr = pose_gap(T, p[name], sc[0])
        assert ang.max() < 0.05 and tr.max() < 0.01, name


@pytest.mark.parametrize("name", ["pieceA", "pieceB"])
def test_thickness_rays_against_the_dumps_first_hits(name):
    pc = S.piece(name)
    scene = O.Scene(pc["V0"].astype(np.float32), pc["F"])
    o, d = S.thickness_rays(name)
    t, prim = scene.first_hit(o, d)
    gt, gp = pc["thick_t_hit"], pc["thick_prim"]
    assert np.array_equal(np.isfinite(gt), np.isfinite(t))
    both = np.isfinite(gt)
    assert np.abs(gt[both] - t[both]).max() <= 2e-3
    assert np.mean(gp[both] == prim[both]) > 0.999
    # The FP64 retest changes nothing on a well-conditioned mesh.
    t2, p2 = scene.first_hit(o, d, retest=True)
    assert np.mean(t2 == t) > 0.999


@pytest.mark.parametrize("name", ["pieceA", "pieceB"])
def test_cone_votes_against_the_dump(name):
    pc = S.piece(name)
    scene = O.Scene(pc["V"].astype(np.float32), pc["F"])
    good = scene.cone_votes(pc["V"], pc["F"], pc["seg_faces"], pc["seg_NS"], pc["thick_t"])
    assert np.mean(good == pc["seg_good"]) > 0.999
    votes = pc["seg_votes"]
    assert np.mean((good < votes) == (pc["seg_good"] < votes)) > 0.999
