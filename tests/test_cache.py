"""The fragment cache reader and MatchData::own (CPU; the library's host code, no device).

The reader against cache.rs's own tests (crates/sherd-core/src/fragment/cache.rs:746-1030: the
sample fragment's round trip bit for bit, every refusal with its message, empty breaklines and
samples read back as empty); MatchData against samples.rs (the registration split's known answers,
:811-840, the half-to-even rounding, :833-840) and against the slab golden vectors: a cache written
from the reference's dump arrays (narrowed to the cache's f32, as the format stores them) yields
the slab's inputs, and from them the oracle reproduces the slab rows within the reference's own
parity tolerances."""

import json
import os

import numpy as np
import pytest

import oracle_lib as O
import sherd_cache as W
import slab_data as SD

sherd_hip = pytest.importorskip("sherd_hip")
LIB = sherd_hip.LIB_PATH
pytestmark = pytest.mark.skipif(not os.path.exists(LIB), reason="libsherd_hip.so not built")


def load(arrays, meta):
    return sherd_hip.FragmentCache(data=W.to_bytes(arrays, meta))


def refused(arrays, meta, needle):
    with pytest.raises(sherd_hip.HipError) as e:
        load(arrays, meta)
    assert e.value.status == sherd_hip.E_FORMAT and needle in str(e.value), str(e.value)


def test_the_sample_fragment_round_trips_bit_for_bit():
    arrays, meta = W.tetrahedron()
    c = load(arrays, meta)
    got = c.arrays()
    for k, v in arrays.items():
        want = np.asarray(v, dtype=got[k].dtype).reshape(got[k].shape)
        assert np.array_equal(got[k], want), k
    i = c.info
    assert i["name"] == "pieceA" and i["cache_version"] == 7 and i["watertight"] == 1
    for k in ("thick", "thick_mode", "res", "frac_area", "area"):
        assert i[k] == meta[k], k  # shortest round-trip text reads back to the same bits
    assert (i["nv"], i["nf"], i["n_brk"], i["n_sub"], i["n_s"], i["n_pf"], i["n_margin"]) == (4, 4, 2, 1, 3, 1, 2)


@pytest.mark.parametrize("field,value,needle", [("cache_version", 8, "cache_version"),
                                                ("format", "other", "format `other`"),
                                                ("algo_ref", "x", "algo_ref `x`")])
def test_a_file_from_another_build_is_refused_with_its_reason(field, value, needle):
    arrays, meta = W.tetrahedron()
    meta[field] = value
    refused(arrays, meta, needle)


@pytest.mark.parametrize("edit,needle", [
    (lambda a: a.update(brk_sub=[7]), "brk_sub 7 is outside brk_P (2)"),
    (lambda a: a.update(brk_ns=a["brk_ns"][:1]), "frames do not describe brk_P"),
    (lambda a: a.update(brk_valid=[True]), "brk_valid does not describe brk_P"),
    (lambda a: a.update(sp=a["sp"][:2]), "sp/fp do not describe S/Pf"),
    (lambda a: a.update(fp=[9]), "sample face 9 is outside F (4)"),
    (lambda a: a.update(margin_idx=[11]), "margin_idx 11 is outside S (3)"),
    (lambda a: a.update(labels=a["labels"][:3]), "tensor labels"),
    (lambda a: a.update(labels=[0, 1, 3, 0]), "label 3"),
    (lambda a: a.update(F=[[0, 1, 99]] + a["F"][1:]), "triangle index 99 is outside V (4)"),
])
def test_a_cache_that_does_not_hang_together_is_refused(edit, needle):
    arrays, meta = W.tetrahedron()
    edit(arrays)
    refused(arrays, meta, needle)


def test_malformed_files_are_refused():
    arrays, meta = W.tetrahedron()
    good = W.to_bytes(arrays, meta)
    for bad, needle in [(good[:5], "shorter than a safetensors header"),
                        (good[:-4], "its bytes lie outside the file"),
                        (good + b"junk", "do not cover the data section"),
                        ((10**9).to_bytes(8, "little") + good[8:], "runs past the file"),
                        (good[:8] + b"[" + good[9:], "not a JSON object")]:
        with pytest.raises(sherd_hip.HipError) as e:
            sherd_hip.FragmentCache(data=bad)
        assert needle in str(e.value), (needle, str(e.value))
    # A tensor of the wrong dtype, written by hand into an otherwise valid file.
    from safetensors.numpy import save

    t = {k: np.asarray(v) for k, v in arrays.items()}
    t["V"] = np.asarray(arrays["V"], np.float64)
    for k, dt in (("F", np.uint32), ("labels", np.uint8), ("brk_sub", np.uint32), ("sp", np.uint32),
                  ("fp", np.uint32), ("margin_idx", np.uint32)):
        t[k] = np.asarray(arrays[k], dt)
    for k in ("brk_P", "brk_ns", "brk_nf", "brk_f", "S", "Pf"):
        t[k] = np.asarray(arrays[k], np.float32)
    t["brk_valid"] = np.asarray(arrays["brk_valid"], bool)
    raw = save(t, metadata={"sherd": json.dumps(meta)})
    with pytest.raises(sherd_hip.HipError) as e:
        sherd_hip.FragmentCache(data=raw)
    assert "tensor V is F64 [4, 3], expected F32 [n, 3]" in str(e.value)
    del t["Pf"]
    with pytest.raises(sherd_hip.HipError):
        sherd_hip.FragmentCache(data=save(t, metadata={"sherd": json.dumps(meta)}))
    with pytest.raises(sherd_hip.HipError) as e:
        sherd_hip.FragmentCache(data=save({"V": np.zeros((1, 3), np.float32)}))
    assert "no `sherd` metadata key" in str(e.value)
    with pytest.raises(sherd_hip.HipError) as e:
        sherd_hip.FragmentCache(path="/nonexistent/x.sherd")
    assert e.value.status == sherd_hip.E_FORMAT


def test_empty_breaklines_and_samples_read_back_empty():
    arrays, meta = W.tetrahedron()
    arrays.update(brk_P=np.zeros((0, 3)), brk_ns=np.zeros((0, 3)), brk_nf=np.zeros((0, 3)), brk_f=np.zeros((0, 3)),
                  brk_sub=[], brk_valid=np.zeros(0, bool), S=np.zeros((0, 3)), sp=[], Pf=np.zeros((0, 3)), fp=[],
                  margin_idx=[])
    c = load(arrays, meta)
    i = c.info
    assert (i["n_brk"], i["n_sub"], i["n_s"], i["n_pf"], i["n_margin"]) == (0, 0, 0, 0, 0)
    m = c.match_data()
    assert m["n_reg"] == 0 and m["reg_p"].shape == (0, 3)


def _normals32(V, F):
    n, _, _ = SD.face_geometry(np.asarray(V, np.float32).astype(np.float64), np.asarray(F))
    return n.astype(np.float32).astype(np.float64)


def test_match_data_of_the_sample_fragment():
    arrays, meta = W.tetrahedron()
    m = load(arrays, meta).match_data()
    FN = _normals32(arrays["V"], arrays["F"])
    S32 = np.asarray(arrays["S"], np.float32).astype(np.float64)
    assert np.array_equal(m["face_normals"], FN)
    assert np.array_equal(m["s"], S32) and np.array_equal(m["s_n"], FN[[0, 1, 3]])
    assert np.array_equal(m["s_frac"], [False, True, False])
    assert np.array_equal(m["margin_p"], S32[[0, 2]]) and np.array_equal(m["margin_n"], FN[[0, 3]])
    assert np.array_equal(m["pf_n"], FN[[1]])
    # Everything fits under 6000: pc_reg is Pf then the margin.
    assert (m["n_reg_frac"], m["n_reg_margin"]) == (1, 2)
    assert np.array_equal(m["reg_p"], np.concatenate([m["pf"], m["margin_p"]]))
    # ns x f and the dihedral of orthogonal normals.
    assert np.array_equal(m["brk_t"], [[0.0, -1.0, 0.0], [-1.0, 0.0, 0.0]])
    assert np.allclose(m["brk_dih"], 90.0, atol=0, rtol=1e-15)
    assert np.array_equal(m["brk_sub"], [1]) and m["t"] == meta["thick"] and m["frac_area"] == meta["frac_area"]
    # Arrays built at another t than the fragment's own are not MatchData::own's.
    meta2 = dict(meta, md_params=W.md_params(1.0))
    with pytest.raises(sherd_hip.HipError) as e:
        load(arrays, meta2).match_data()
    assert e.value.status == sherd_hip.E_LIMIT


def _sized(n_pf, n_margin):
    arrays, meta = W.tetrahedron()
    n_s = max(n_margin, 1)
    arrays.update(S=np.zeros((n_s, 3)), sp=np.zeros(n_s), Pf=np.zeros((n_pf, 3)), fp=np.zeros(n_pf),
                  margin_idx=np.arange(n_margin))
    return load(arrays, meta)


@pytest.mark.parametrize("n_pf,n_margin,reg,want", [
    (4000, 6000, 6000, (2400, 3600)),   # samples.rs:816-820
    (100, 200, 6000, (100, 200)),       # below the budget nothing is dropped
    (4000, 6000, 0, (4000, 6000)),      # reg_points 0 keeps everything
    (1, 1_000_000, 6000, (1, 5999)),    # one fracture point survives
    (5, 3, 4, (2, 2)),                  # share 2.5 rounds half to even
    (7, 1, 4, (4, 0)),                  # share 3.5 rounds to 4, leaving 0 for the margin
])
def test_the_registration_split_keeps_the_proportion_and_the_count(n_pf, n_margin, reg, want):
    m = _sized(n_pf, n_margin).match_data(reg)
    assert (m["n_reg_frac"], m["n_reg_margin"]) == want and m["n_reg"] == sum(want)


# ---- the slab: a cache written from the reference's dump arrays ------------------------------

def slab_cache(name):
    pc = SD.piece(name)
    ca = np.load(os.path.join(SD.GOLD, "cache_arrays.npz"))
    t = SD.pair()["scales"][0]
    arrays = dict(V=pc["V"], F=pc["F"], labels=pc["frac"].astype(np.uint8), brk_P=pc["brk_P"], brk_ns=pc["brk_ns"],
                  brk_nf=ca[name + "_brk_nf"], brk_f=pc["brk_f"], brk_sub=pc["brk_sub"],
                  brk_valid=ca[name + "_brk_valid"], S=pc["S"], sp=ca[name + "_sp"], Pf=pc["Pf"], fp=ca[name + "_fp"],
                  margin_idx=ca[name + "_margin_idx"])
    stats = json.loads(str(ca[name + "_stats"]))
    md = json.loads(str(ca[name + "_md_params"]))
    # The arrays are the pair's (piece A's at the pair's t, as the pair reads them), so the cache
    # states that t as the fragment's own.
    meta = W.meta(name, md["t"], stats["res"], float(pc["frac_area"]), watertight=bool(pc["watertight"]),
                  n_boundary=int(pc["n_boundary"]), area=stats["area"], thick_mode=stats["thick_mode"],
                  seed=md["seed"])
    return load(arrays, meta), pc, t


@pytest.fixture(scope="module")
def slab():
    return {n: slab_cache(n) for n in ("pieceA", "pieceB")}


@pytest.mark.parametrize("name", ["pieceA", "pieceB"])
def test_the_slab_caches_match_data_is_the_pairs_inputs(slab, name):
    c, pc, _ = slab[name]
    m = c.match_data()
    n32 = lambda a: np.asarray(a, np.float32).astype(np.float64)
    assert np.array_equal(m["pf"], n32(pc["Pf"])) and np.array_equal(m["margin_p"], n32(pc["margin_p"]))
    assert np.array_equal(m["brk_p"], n32(pc["brk_P"])) and np.array_equal(m["brk_sub"], pc["brk_sub"])
    # Normals from the f32 vertices against the dump's f64 ones; the tangent and dihedral from the
    # f32 frames against the dump's f64 ones.
    assert np.abs(m["pf_n"] - pc["nf"]).max() < 1e-5 and np.abs(m["margin_n"] - pc["margin_n"]).max() < 1e-5
    assert np.abs(m["brk_t"] - pc["brk_t"]).max() < 1e-6
    assert np.abs(m["brk_dih"] - pc["brk_dih"]).max() < 1e-3
    tf = m["n_reg_frac"]
    assert m["n_reg"] == 6000 and np.array_equal(m["reg_p"][:tf], m["pf"][:tf])
    assert np.array_equal(m["reg_p"][tf:], m["margin_p"][: 6000 - tf])


def _inputs(slab):
    out = {}
    for name in ("pieceA", "pieceB"):
        c = slab[name][0]
        m = c.match_data()
        arr = c.arrays()
        out[name] = dict(m, V=arr["V"], F=arr["F"], frac=arr["labels"] == 1, watertight=bool(m["watertight"]))
    return out


def test_the_slab_rows_from_the_cache(slab):
    """coarse, stage 1 and verify from the cache's arrays (narrowed to f32) against the dump's
    rows: coarse bit-exact (no probe crosses the radius by the narrowing); stage 1 within
    0.05 deg / 0.01 t; verify within tight 0.01, gap 0.002 t, seam 0.34 t, cont 0.005 t,
    cont_n 0.01, pen 0.0005 with accepted identical (crates/sherd-parity/src/stages/*.rs)."""
    x = _inputs(slab)
    a, b, p = x["pieceA"], x["pieceB"], SD.pair()
    sc = p["scales"]
    fa = dict(brk_t=a["brk_t"], brk_ns=a["brk_ns"], brk_f=a["brk_f"], brk_P=a["brk_p"])
    fb = dict(brk_t=b["brk_t"], brk_ns=b["brk_ns"], brk_f=b["brk_f"], brk_P=b["brk_p"])
    poses = SD.hypothesis_poses(fa, fb, p["hyp_ia"], p["hyp_ib"], p["hyp_pa"], p["hyp_pb"])
    idx = p["coarse_idx"]
    s, _ = O.coarse_scores(a["brk_p"], a["brk_ns"], b["brk_p"][idx], b["brk_ns"][idx], poses, sc[2])
    assert np.array_equal(s.view(np.uint64), p["coarse_cs"].view(np.uint64))
    # Stage 1 from the reference's kept hypotheses.
    inits = np.tile(np.eye(4), (len(p["nms1_kept"]), 1, 1))
    inits[:, :3, :3] = poses[p["nms1_kept"], :9].reshape(-1, 3, 3)
    inits[:, :3, 3] = poses[p["nms1_kept"], 9:]
    T = inits
    for k in (0.2, 0.08):
        T = O.icp_rung(b["brk_p"][b["brk_sub"]], a["brk_p"], a["brk_ns"], T, 0, k * sc[0] * sc[11], 20)[:, :16]
        T = T.reshape(-1, 4, 4)
    from test_oracle_slab import pose_gap

    ang, tr = pose_gap(T, p["s1_T"], sc[0])
    assert ang.max() < 0.05 and tr.max() < 0.01
    # Verification at the reference's stage-2 poses.
    keep, sides = [], []
    for side in (a, b):
        v32 = np.asarray(side["V"], np.float32)
        fs, ms = O.Scene(v32, side["F"], side["frac"]), O.Scene(v32, side["F"])
        keep += [fs, ms]
        sides.append(O.SideData(fs, ms, side["watertight"], side["frac_area"], side["pf"], side["s"], side["brk_p"],
                                side["brk_ns"], side["margin_p"], side["margin_n"]))
    tol = dict(tight=0.01, gap=0.002, seam=0.34, cont=0.005, cont_n=0.01, pen=0.0005)
    for i, T in enumerate(p["s2_T_frac2"]):
        s = O.verify(sides[0], sides[1], T, sc)
        theirs = p["s2_scores"][i]
        for k, tl in tol.items():
            assert abs(s[O.OS_KEYS.index(k)] - theirs[k]) <= tl, (i, k)
        assert O.accept(s, sc) == bool(p["s2_accepted"][i]), i


def test_the_narrowing_moves_one_stage2_candidate():
    """The stage-2 rungs on the cache's (f32) arrays from the dump's own previous-rung poses: every
    candidate stays within 4e-4 deg of the dump's pose except candidate 8 at the first pc_frac
    rung, which the narrowing of the fracture samples moves 0.23 deg (a correspondence crosses the
    rung's radius). With the dump's f64 arrays the same rung is within 1e-4 deg."""
    from test_oracle_slab import pose_gap, stage2_rungs

    x = _inputs({n: slab_cache(n) for n in ("pieceA", "pieceB")})
    a, b, p = x["pieceA"], x["pieceB"], SD.pair()
    sc = p["scales"]
    prev = p["s1_T"][p["nms2_kept"]]
    cache_rungs = [("s2_T_reg1", b["reg_p"], a["reg_p"], a["reg_n"], 0.2),
                   ("s2_T_reg2", b["reg_p"], a["reg_p"], a["reg_n"], 0.08),
                   ("s2_T_frac1", b["pf"], a["pf"], a["pf_n"], 0.08), ("s2_T_frac2", b["pf"], a["pf"], a["pf_n"], 0.04)]
    for (name, src, tp, tn, k), (_, s64, t64, n64, _) in zip(cache_rungs, stage2_rungs()):
        T = O.icp_rung(src, tp, tn, prev, 1, k * sc[0] * sc[11], 30)[:, :16].reshape(-1, 4, 4)
        ang, tr = pose_gap(T, p[name], sc[0])
        T64 = O.icp_rung(s64, t64, n64, prev, 1, k * sc[0] * sc[11], 30)[:, :16].reshape(-1, 4, 4)
        ang64, _ = pose_gap(T64, p[name], sc[0])
        assert ang64.max() < 1e-4, name
        if name == "s2_T_frac1":
            assert np.nonzero(ang > 4e-4)[0].tolist() == [8] and 0.2 < ang[8] < 0.25 and tr[8] < 0.02
        else:
            assert ang.max() < 4e-4 and tr.max() < 1e-4, name
        prev = p[name]

