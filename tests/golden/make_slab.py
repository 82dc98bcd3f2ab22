"""Writes the slab golden vectors under tests/golden/slab/ from the reference's fixture dump.

Run in a container that has the reference checked out:

    python tests/golden/make_slab.py /root/reference/fixtures/slab/dump

Everything written is data the reference's own fixture dump holds (fixtures/slab/dump, produced by
tools/dump_fixtures.py of the reference), plus a few arrays derived from it with the reference's
own arithmetic, spelled out below with the file:line each follows:

* face normals and areas: crates/sherd-core/src/mesh/geometry.rs:70-96 (f64, `NORMAL_FLOOR` 1e-12);
* `frac_area = A[frac].sum()` with numpy's pairwise sum, as crates/sherd-parity/src/stages/pairs.rs:343-351
  does through `masked_area` (crates/sherd-core/src/fragment/segment.rs:640);
* the shell margin `Pm = S[margin_idx]`, `Nm = FN[sp[margin_idx]]`
  (crates/sherd-parity/src/stages/pairs.rs:414-416);
* the fracture-sample normals `FN[fp]` (crates/sherd-parity/src/stages/pairs.rs:286).

R §3.4.4-3.4.8's row (seg.npz) is the dump's own seg.rep, seg.near, seg.frac_raw, seg.frac_majority,
seg.frac_islands, seg.frac_final, seg.ref and seg.has_ref of each piece.

R §9's row (refine.npz) is the dump's own: refine/<piece>.idx, refine/joins.json and
refine/poses_final.json, the assembly's poses.json, groups.json and used.json it starts from, and
each fragment's mesh.stats.json thickness and resolution (crates/sherd-parity/src/stages/refine.rs:
139-205 reads the same files).

Only numpy is used, and only `numpy.load` with its default `allow_pickle=False`.
"""

import json
import os
import sys

import numpy as np

PIECES = ("pieceA", "pieceB")
PAIR = "pieceA__pieceB"
HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "slab")


def face_geometry(v, f):
    """geometry.rs:70-96: f64 normals (floored length) and areas, elementwise in its order."""
    a, b, c = v[f[:, 0]], v[f[:, 1]], v[f[:, 2]]
    e1 = b - a
    e2 = c - a
    n0 = e1[:, 1] * e2[:, 2] - e1[:, 2] * e2[:, 1]
    n1 = e1[:, 2] * e2[:, 0] - e1[:, 0] * e2[:, 2]
    n2 = e1[:, 0] * e2[:, 1] - e1[:, 1] * e2[:, 0]
    length = np.sqrt((n0 * n0 + n1 * n1) + n2 * n2)
    den = np.maximum(length, 1e-12)
    normals = np.stack([n0 / den, n1 / den, n2 / den], axis=1)
    return normals, 0.5 * length


def arrays_dir(frag_dir, t):
    """pairs.rs:480-505: the directory whose md.params.json carries the pair's t bit for bit."""
    cands = [frag_dir]
    md_t = os.path.join(frag_dir, "md_t")
    if os.path.isdir(md_t):
        cands += [os.path.join(md_t, d) for d in sorted(os.listdir(md_t))]
    for d in cands:
        p = os.path.join(d, "md.params.json")
        if os.path.isfile(p):
            with open(p) as fh:
                params = json.load(fh)
            if np.float64(params["t"]).tobytes() == np.float64(t).tobytes():
                return d
    raise SystemExit(f"no arrays at t={t} under {frag_dir}")


def main(dump):
    os.makedirs(OUT, exist_ok=True)
    pair_dir = os.path.join(dump, "pairs", PAIR)
    with open(os.path.join(pair_dir, "md_used.json")) as fh:
        used = json.load(fh)
    t_pair = used["t"]
    for piece in PIECES:
        d = os.path.join(dump, "fragments", piece)
        ld = lambda name, base=d: np.load(os.path.join(base, name + ".npy"))
        v = ld("mesh.V")
        f = ld("mesh.F").astype(np.uint32)
        frac = ld("seg.frac_final").astype(bool)
        normals, areas = face_geometry(v, f.astype(np.int64))
        frac_area = areas[frac].sum()
        with open(os.path.join(d, "mesh.watertight.json")) as fh:
            wt = json.load(fh)
        md = arrays_dir(d, t_pair)
        mdl = lambda name: ld(name, md)
        s = mdl("md.S")
        sp = mdl("md.sp").astype(np.int64)
        midx = mdl("md.margin_idx").astype(np.int64)
        fp = mdl("md.fp").astype(np.int64)
        # The thickness rays of R §3.2 run on the loaded mesh (thickness.rs:99-132).
        v0 = ld("load.V0")
        np.savez_compressed(
            os.path.join(OUT, piece + ".npz"),
            V=v,
            F=f,
            frac=frac,
            watertight=np.bool_(wt["watertight"]),
            n_boundary=np.int64(wt["n_boundary"]),
            frac_area=np.float64(frac_area),
            Pf=mdl("md.Pf"),
            nf=normals[fp],
            S=s,
            margin_p=s[midx],
            margin_n=normals[sp[midx]],
            brk_P=mdl("md.brk_P"),
            brk_ns=mdl("md.brk_ns"),
            brk_f=mdl("md.brk_f"),
            brk_t=mdl("md.brk_t"),
            brk_dih=mdl("md.brk_dih"),
            brk_sub=mdl("md.brk_sub").astype(np.int64),
            V0=v0,
            thick_t_hit=ld("thick.t_hit").astype(np.float32),
            thick_prim=ld("thick.prim").astype(np.uint32),
            thick_t=np.float64(json.load(open(os.path.join(d, "thick.t.json")))),
            # Every fourth face of R §3.4.3's cone vote: enough to pin the vote, a quarter the size.
            seg_faces=np.arange(0, len(f), 4, dtype=np.int64),
            seg_NS=ld("seg.NS")[::4],
            seg_good=ld("seg.good").astype(np.uint8)[::4],
            seg_votes=np.int64(json.load(open(os.path.join(d, "seg.info.json")))["votes"]),
        )
    pl = lambda name: np.load(os.path.join(pair_dir, name + ".npy"))
    with open(os.path.join(pair_dir, "scales.json")) as fh:
        scales = json.load(fh)
    keys = ["t", "res", "coarse", "stage1", "tight", "facing", "gap", "seam", "near", "pen", "nms", "icp"]
    np.savez_compressed(
        os.path.join(OUT, "pair.npz"),
        scales=np.array([scales[k] for k in keys], dtype=np.float64),
        t_a=np.float64(used["t_a"]),
        t_b=np.float64(used["t_b"]),
        res_a=np.float64(used["res_a"]),
        res_b=np.float64(used["res_b"]),
        hyp_ia=pl("hyp.ia").astype(np.int64),
        hyp_ib=pl("hyp.ib").astype(np.int64),
        hyp_pa=pl("hyp.pa").astype(np.int64),
        hyp_pb=pl("hyp.pb").astype(np.int64),
        coarse_idx=pl("coarse.idx").astype(np.int64),
        coarse_cs=pl("coarse.cs"),
        nms1_kept=pl("nms1.kept").astype(np.int64),
        nms2_kept=pl("nms2.kept").astype(np.int64),
        s1_T=pl("s1.T"),
        s1_score=pl("s1.score"),
        s2_T_reg1=pl("s2.T_reg1"),
        s2_T_reg2=pl("s2.T_reg2"),
        s2_T_frac1=pl("s2.T_frac1"),
        s2_T_frac2=pl("s2.T_frac2"),
        s2_T=pl("s2.T"),
        s2_accepted=pl("s2.accepted").astype(bool),
    )
    with open(os.path.join(pair_dir, "s2.scores.json")) as fh:
        scores = json.load(fh)
    with open(os.path.join(OUT, "s2_scores.json"), "w") as fh:
        json.dump(scores, fh, indent=1)
    with open(os.path.join(dump, "..", "input", "ground_truth.json")) as fh:
        gt = json.load(fh)
    with open(os.path.join(OUT, "ground_truth.json"), "w") as fh:
        json.dump(gt, fh, indent=1)
    write_refine(dump)
    write_stage(dump)


def write_stage(dump):
    """stage.npz: R §5.3 and R §5.5's own walk orders (PMC-6: numpy's unstable argsort), which the
    parity of the greedy suppression is measured on."""
    pair_dir = os.path.join(dump, "pairs", PAIR)
    ld = lambda name: np.load(os.path.join(pair_dir, name + ".npy")).astype(np.int64)
    np.savez_compressed(os.path.join(OUT, "stage.npz"), nms1_order=ld("nms1.order"), nms2_order=ld("nms2.order"))


def write_refine(dump):
    """refine.npz: R §9's inputs and outputs for the slab's one join."""
    rd, ad = os.path.join(dump, "refine"), os.path.join(dump, "assembly")
    jl = lambda path: json.load(open(path))
    index = {p: i for i, p in enumerate(PIECES)}
    poses = lambda path: np.array([jl(path)[p] for p in PIECES], dtype=np.float64)
    stats = [jl(os.path.join(dump, "fragments", p, "mesh.stats.json")) for p in PIECES]
    joins = jl(os.path.join(rd, "joins.json"))
    groups = jl(os.path.join(ad, "groups.json"))
    used = jl(os.path.join(ad, "used.json"))
    out = dict(
        poses_before=poses(os.path.join(ad, "poses.json")),
        poses_final=poses(os.path.join(rd, "poses_final.json")),
        thick=np.array([s["thick"] for s in stats], dtype=np.float64),
        res=np.array([s["res"] for s in stats], dtype=np.float64),
        group_members=np.array([index[n] for g in groups for n in g], dtype=np.int64),
        group_sizes=np.array([len(g) for g in groups], dtype=np.int64),
        used=np.array([[index[u["a"]], index[u["b"]]] for u in used], dtype=np.int64).reshape(-1, 2),
        join_fixed=np.array([index[j["fixed"]] for j in joins], dtype=np.int64),
        join_moving=np.array([index[j["moving"]] for j in joins], dtype=np.int64),
        join_dist=np.array([j["dist"] for j in joins], dtype=np.float64),
        join_T=np.array([j["T_rung"] for j in joins], dtype=np.float64),
        join_fitness=np.array([j["fitness"] for j in joins], dtype=np.float64),
        join_rmse_t=np.array([j["rmse_t"] for j in joins], dtype=np.float64),
    )
    for p in PIECES:
        out["idx_" + p] = np.load(os.path.join(rd, p + ".idx.npy")).astype(np.int64)
    np.savez_compressed(os.path.join(OUT, "refine.npz"), **out)

    # R §3.4.4-3.4.8's inputs and intermediate masks for the segmentation post-processing row: the
    # voxel representatives and each face's nearest one, the cone vote's mask, and what the majority
    # filter, the island passes and the boundary growth left (crates/sherd-parity/src/stages/
    # segmentation.rs compares the same files).
    seg = {}
    for p in PIECES:
        d = os.path.join(dump, "fragments", p)
        for k in ("rep", "near", "frac_raw", "frac_majority", "frac_islands", "frac_final", "ref", "has_ref"):
            a = np.load(os.path.join(d, "seg." + k + ".npy"))
            if a.dtype == np.int64:
                a = a.astype(np.uint32)
            seg[p + "_" + k] = a
    np.savez_compressed(os.path.join(OUT, "seg.npz"), **seg)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/fixtures/slab/dump")
