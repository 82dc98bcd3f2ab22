"""Benchmark of the MI355X sherd-refit backend on one GPU.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Headline (the full collection): 40 fragments of three synthetic thin-wall
vessels (walls t = 0.8 / 1.0 / 1.3 at 4 / 5 / 6 edges per t, 25.3 M triangles, two fragments made
non-watertight), all 780 pairs x 4 candidate poses. One step = the batched pipeline over every
candidate with the inputs resident on the device: the breakline ICP ladder (sherd_icp_batch_ladder),
the Huber LM (sherd_huber_lm, f32 sums, FP64 6x6 solve) and R §6's verification with the FP64
retest (sherd_verify_batch). `value` is pair-verifications per second over the timed steps (3120
candidates per step); setup (scenes, parity grids, distance bounds, clouds) is timed apart. Beside it:
  * the plain (per-pair / per-candidate) HIP schedule over its own timed steps, the batched/plain
    ratio, and a check that every result is identical;
  * the roofline of R §6.4's inside pass (k_pen's box reject, k_pen_groups, k_pen_rays' parity rays), its launch
    times from HIP events on its stream over the timed steps, with the measured HBM traffic of the committed
    counter profile beside it;
  * the CPU baseline: the oracle's end-to-end pipeline (C++, `oracle/`) on a bounded sample of the
    collection's candidates, on all host threads over candidates and on one, median of five;
  * sections, each with its own oracle sample: configuration 2's vessel (`vessel`: verification,
    cone-vote rays/s, kNN, LBVH build), configuration 1's slab plate (`slab`), configuration 3's
    breakline ladder (`icp`), configuration 4's robust energy and LM (`huber`), R §9's refinement
    (`refine`), R §5.1-5.5 (`match`), R §3.2's thickness (`thickness`) and R §3.4.4-3.4.8's
    segmentation clean-up (`segment`).
"""

import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "sherd-refit_amd"))

METRIC = "pair-verifications/s + rays/s, batched vs plain HIP, 1 MI355X, same accepted seams"
HBM_PEAK_GBS = 8000.0
# The three sections that run concurrently on side streams after R §6.4's inside pass.
OVERLAPPED_KERNELS = ("k_frac", "k_seam", "k_cont")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def algorithmic_bytes(kernel, frags, pairs, cand_pair):
    """Bytes of one launch of a verification kernel over the batch in its per-point unit: every
    point of a (candidate, side) task read once (f64 samples, 24 B) and every per-point output
    written once. The meshes and fracture scenes the points are tested against are not counted: a
    launch reads only the parts near its points, and the bytes it does move are the measured
    traffic reported beside it. A margin cloud an R §6.3 task searches is counted once per launch."""
    used = set()
    total = 0
    for c in cand_pair:
        a, b = pairs[c]
        fa, fb = frags[a], frags[b]
        if kernel == "k_pen":
            total += 24 * (len(fa["S"]) + len(fb["S"]))
        elif kernel == "k_frac":
            total += (24 + 8) * (len(fa["Pf"]) + len(fb["Pf"]))  # the point, its distance key
        elif kernel == "k_seam":
            total += 48 * (len(fa["brk_P"]) + len(fb["brk_P"]))
        elif kernel == "k_cont":
            total += (48 + 16) * len(fb["margin_p"])  # point and normal, two keys
            used.add(a)
    return total + sum(48 * len(frags[i]["margin_p"]) for i in used)


def cpu_baseline(col, pairs, scales, cand_pair, cand_T, limit_s=20.0):
    """The oracle on a bounded sample of the candidates: pair-verifications/s on this host."""
    sys.path.insert(0, os.path.join(HERE, "tests"))
    import oracle_lib as O

    cores = min(os.cpu_count() or 1, 16)
    O.set_threads(cores)
    frags = col["fragments"]
    sample_pairs = sorted(set(int(c) for c in cand_pair[:30]))
    need = sorted(set(i for p in sample_pairs for i in pairs[p]))
    sides = {}
    keep = []
    for i in need:
        fr = frags[i]
        fs, ms = O.Scene(fr["V"], fr["F"], fr["frac"]), O.Scene(fr["V"], fr["F"])
        keep += [fs, ms]
        sides[i] = O.SideData(fs, ms, fr["watertight"], fr["frac_area"], fr["Pf"], fr["S"], fr["brk_P"],
                              fr["brk_ns"], fr["margin_p"], fr["margin_n"])
    t0 = time.perf_counter()
    done = 0
    out = []
    for k in range(min(30, len(cand_pair))):
        a, b = pairs[int(cand_pair[k])]
        out.append(O.verify(sides[a], sides[b], cand_T[k], scales))
        done += 1
        if time.perf_counter() - t0 > limit_s:
            break
    el = time.perf_counter() - t0
    return dict(value=done / el, unit="pair-verifications/s", cores=cores, kind="port",
                sample=f"oracle verify of the first {done} candidates (pairs {sample_pairs[0]}-{sample_pairs[-1]}) "
                       f"of the same workload, {cores} OpenMP threads, scene builds excluded"), out


def icp_ladder(ex, synth, args):
    """The batched ICP ladder: pairs x poses candidates climb four point-to-point rungs (1.0, 0.3,
    0.1, 0.03 t, 20 iterations each) in one call, against the plain schedule (one per-call
    icp_rung per pair per rung, as ladder::climb_all issues them) and the oracle on a sample."""
    w = synth.icp_batch_workload(n_pairs=args.icp_pairs, poses=args.icp_poses)
    P = w["pairs"]
    clouds = [(ex.cloud(p["source"]), ex.cloud(p["target"], p["target_n"])) for p in P]
    rungs = [("point_to_point", it) for _, it in w["rungs"]]
    per_rung = [[(s, t, r) for s, t in clouds] for r, _ in w["rungs"]]
    cp = np.repeat(np.arange(len(P), dtype=np.uint32), args.icp_poses)
    inits = np.concatenate([p["inits"] for p in P])
    ns = np.array([len(p["source"]) for p in P])[cp]
    t0 = time.perf_counter()
    ex.icp_batch_ladder(per_rung, rungs, cp, inits)  # tunes the workgroup size per size bucket
    t_tune = time.perf_counter() - t0
    ex.icp_batch_ladder(per_rung, rungs, cp, inits)  # captures the graph
    t0 = time.perf_counter()
    out = ex.icp_batch_ladder(per_rung, rungs, cp, inits)
    el = time.perf_counter() - t0
    iters = out[:, :, 19]
    cand_iter = float(iters.sum())
    queries = float((ns[:, None] * (iters + 1)).sum())
    log(f"icp ladder: {len(cp)} candidates x {len(rungs)} rungs in {el:.3f}s ({cand_iter:.0f} cand-iter, "
        f"{queries / 1e9:.2f} G point-queries); first call with tuning {t_tune:.2f}s")
    # Plain: the reference's schedule, one Executor::icp_rung per pair and rung (over the first
    # --icp-plain-pairs pairs; the ratio is against the batched time scaled to the same pairs).
    n_plain = len(P) if args.icp_plain_pairs < 0 else min(args.icp_plain_pairs, len(P))
    t0 = time.perf_counter()
    same = True
    for p, pr in enumerate(P[:n_plain]):
        T = pr["inits"]
        for r, (dist, it) in enumerate(w["rungs"]):
            o = ex.icp_rung(pr["source"], pr["target"], pr["target_n"], T, 0, dist, it)
            same &= bool(np.array_equal(o, out[cp == p][:, r]))
            T = o[:, :16].reshape(-1, 4, 4)
    el_plain = time.perf_counter() - t0
    if n_plain:
        log(f"icp plain: {n_plain} pairs in {el_plain:.3f}s; identical to batched: {same}")
    plain_frac = float(np.sum(cp < n_plain)) / len(cp)
    res = dict(candidates=int(len(cp)), rungs=len(rungs), cand_iter=int(cand_iter),
               point_queries=int(queries), seconds=round(el, 4),
               cand_iter_per_s=round(cand_iter / el, 1), point_queries_per_s=round(queries / el, 1),
               plain_pairs=n_plain, plain_seconds=round(el_plain, 4) if n_plain else None,
               batched_vs_plain=round(el_plain / (el * plain_frac), 3) if n_plain else None,
               identical_plain_batched=same if n_plain else None,
               converged_last_rung=int(out[:, -1, 20].sum()),
               workload=f"{len(P)} breakline pairs x {args.icp_poses} poses, {int(ns.min())}-{int(ns.max())} "
                        f"source points (log-uniform), rungs 1.0/0.3/0.1/0.03 t x 20 iterations")
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(HERE, "tests"))
        import oracle_lib as O

        O.set_threads(min(os.cpu_count() or 1, 16))
        pr = P[0]
        k = 8
        t0 = time.perf_counter()
        T = pr["inits"][:k]
        c_iters = 0.0
        for r, (dist, it) in enumerate(w["rungs"]):
            o = O.icp_rung(pr["source"], pr["target"], pr["target_n"], T, 0, dist, it)
            c_iters += o[:, 19].sum()
            T = o[:, :16].reshape(-1, 4, 4)
        el_cpu = time.perf_counter() - t0
        res["cpu_cand_iter_per_s"] = round(c_iters / el_cpu, 1)
        res["cpu_sample"] = (f"oracle icp_rung ladder of pair 0's first {k} candidates ({len(pr['source'])} source "
                             f"points), {min(os.cpu_count() or 1, 16)} OpenMP threads")
    for s_, t_ in clouds:
        s_.close()
        t_.close()
    return res


def slab_section(ex, synth, args):
    """the slab plate (60 t x 40 t x t, Perlin-cut, jittered, edge 0.06 t)
    through the whole path on the device (sherd_hip/pipeline.py): one inward ray per vertex for the
    segmentation, R §3.6's sampling of the labelled meshes, R §5.1-5.5's match, a stage-2 ladder and
    verification; rays/s of the per-vertex rays, and the oracle's rays on fragment A as the CPU
    baseline of this row (checked bit for bit)."""
    from sherd_hip import SCORE_KEYS, pipeline

    t0 = time.perf_counter()
    col = synth.slab_plate()
    gen_s = time.perf_counter() - t0
    fa, fb = col["fragments"]
    ex.profile(True)
    out = pipeline.run_pair(ex, fa, fb, col["t"], col["res"])
    prof = ex.profile_read()
    ex.profile(False)
    truth = col["poses"][0] @ np.linalg.inv(col["poses"][1])
    ang, tr = pose_gap_deg_t(out["T"], np.broadcast_to(truth, out["T"].shape))
    acc = out["accepted"]
    rays = len(fa["V"]) + len(fb["V"])
    ray_ms = prof.get("k_vertex_votes", {}).get("ms", float("nan"))
    agree = []
    for fr, fp, P in zip(col["fragments"], out["frac"], col["poses"]):
        shell = (~fr["frac"]) & (np.abs(fr["face_normals"] @ P[:3, 2]) > 0.9)
        agree.append([round(float(np.mean(~fp[shell])), 4), round(float(np.mean(fp[fr["frac"]])), 4)])
    res = dict(workload=f"slab plate 60 t x 40 t x t, edge 0.06 t, {len(fa['F'])} + {len(fb['F'])} triangles, "
                        f"Perlin cut (0.3 t at 2 t), jitter 0.01 t, B posed {np.degrees(np.arccos(np.clip((np.trace(truth[:3, :3]) - 1) / 2, -1, 1))):.1f} deg",
               generation_seconds=round(gen_s, 2), triangles=[int(len(fa["F"])), int(len(fb["F"]))], rays=int(rays),
               vertex_ray_kernel_ms=round(ray_ms, 3), rays_per_s=round(rays / (ray_ms * 1e-3), 1),
               labels_agree_shell_fracture=agree, seconds={k: round(v, 4) for k, v in out["seconds"].items()},
               pipeline_seconds=round(sum(out["seconds"].values()), 4), hypotheses=int(out["match"]["n_hyp"]),
               stage2_candidates=int(len(out["T"])), accepted=int(acc.sum()))
    best = int(out["order"][0])
    res.update(best=dict(accepted=bool(acc[best]), gap_to_truth_deg_t=[round(float(ang[best]), 4), round(float(tr[best]), 4)],
                         seam_x_tight=round(float(out["scores"][best, SCORE_KEYS.index("seam")]
                                                  * out["scores"][best, SCORE_KEYS.index("tight")]), 3),
                         runner_up_seam_x_tight=(round(float(out["scores"][out["order"][1], SCORE_KEYS.index("seam")]
                                                             * out["scores"][out["order"][1], SCORE_KEYS.index("tight")]), 3)
                                                 if len(out["order"]) > 1 else None)))
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(HERE, "oracle"))
        sys.path.insert(0, os.path.join(HERE, "tests"))
        import oracle_lib as O
        import segment as SG

        cores = min(os.cpu_count() or 1, 16)
        O.set_threads(cores)
        osc = O.Scene(fa["V"], fa["F"])
        t0 = time.perf_counter()
        og, _ = SG.vertex_votes(osc, fa["V"].astype(np.float64), fa["F"], col["t"])
        el = time.perf_counter() - t0
        res.update(cpu_rays_per_s=round(len(fa["V"]) / el, 1), cpu_identical_on_fragment_a=bool(np.array_equal(og, out["good"][0])),
                   cpu_sample=f"oracle vertex_votes (normals in numpy, first hits by the C++ oracle on {cores} threads) "
                              f"over fragment A's {len(fa['V'])} vertices")
    log(f"slab: {rays} vertex rays in {ray_ms:.2f} ms; pipeline {res['pipeline_seconds']:.3f}s; accepted {int(acc.sum())} of "
        f"{len(acc)}; best {res['best']}"
        + (f"; oracle {res['cpu_rays_per_s']:.0f} rays/s, identical {res['cpu_identical_on_fragment_a']}" if not args.no_cpu else ""))
    return res


def huber_section(ex, synth, col, args):
    """the robust seam energy over 512 batched poses on the vessel's seams
    (configuration 2's generator; per adjacent pair a cloud on each side of the seam, 10k-40k points
    per side, fracture and shell normals that span 3-D; base poses within 1 degree and 0.1 t of the
    truth): energy+gradient evaluations/s in f32 and f64, the f32 error against the FP64 oracle, the
    device gradient against FP64 central differences (h = 1e-4), and the batched LM in f64 and f32
    with its transform error against the FP64 oracle's LM on a bounded sample of the poses."""
    sys.path.insert(0, os.path.join(HERE, "oracle"))
    sys.path.insert(0, os.path.join(HERE, "tests"))
    import huber as H
    import oracle_lib as O

    w = synth.seam_workload(col, poses=512)
    P, cp, T0 = w["pairs"], w["cand_pair"], w["T0"]
    clouds = [(ex.cloud(p["source"]), ex.cloud(p["target"], p["target_n"]), w["radius"]) for p in P]
    b = ex.huber(clouds, cp, T0)
    k = len(cp)
    xi = np.random.default_rng(31).normal(scale=1e-3, size=(k, 6))
    corr = b.correspondences()
    pairs_used = int((corr >= 0).sum())
    sizes = [len(p["source"]) for p in P] + [len(p["target"]) for p in P]
    res = dict(poses=k, seams=len(P), points_per_side=[int(min(sizes)), int(max(sizes))],
               corresponding_points=pairs_used, delta=w["delta"], radius=w["radius"])
    for prec in (32, 64):
        b.energy(xi, w["delta"], prec, grad=True)
        reps = 5
        t0 = time.perf_counter()
        for _ in range(reps):
            E, G = b.energy(xi, w["delta"], prec, grad=True)
        el = (time.perf_counter() - t0) / reps
        res[f"f{prec}_eval_ms"] = round(1e3 * el, 3)
        res[f"f{prec}_evals_per_s"] = round(k / el, 1)
        res[f"f{prec}_point_terms_per_s"] = round(pairs_used / el, 1)
        if prec == 32:
            E32, G32 = E, G
        else:
            E64, G64 = E, G
    # Error against the FP64 oracle on one pose per seam, and the oracle's central differences.
    offs = np.concatenate([[0], np.cumsum([len(P[p]["source"]) for p in cp])])
    errs_e, errs_g, errs_fd = [], [], []
    for p in range(len(P)):
        i = int(np.nonzero(cp == p)[0][0])
        c = corr[offs[i]:offs[i + 1]]
        m = c >= 0
        p0 = H.base_points(T0[i], P[p]["source"])[m]
        q, n = P[p]["target"][c[m]], P[p]["target_n"][c[m]]
        e = H.energy(p0, q, n, xi[i], w["delta"])
        g = H.gradient(p0, q, n, xi[i], w["delta"])
        errs_e.append(abs(E32[i] - e) / abs(e))
        errs_g.append(np.abs(G32[i] - g).max() / np.abs(g).max())
        if p < 4:
            fd = H.finite_difference(p0, q, n, xi[i], w["delta"], h=1e-4)
            errs_fd.append(np.abs(G64[i] - fd).max() / np.abs(fd).max())
    res["f32_energy_rel_err_max"] = float(max(errs_e))
    res["f32_grad_rel_err_max"] = float(max(errs_g))
    res["f64_grad_vs_fp64_fd_rel_max"] = float(max(errs_fd))
    iters = 20
    out = {}
    for prec in (64, 32):
        b.lm(w["delta"], iters, prec)
        t0 = time.perf_counter()
        out[prec] = b.lm(w["delta"], iters, prec)
        res[f"lm_f{prec}_ms"] = round(1e3 * (time.perf_counter() - t0), 3)
    res["lm_iterations"] = iters
    res["lm_f64_mean_accepted_steps"] = round(float(out[64][2].mean()), 2)
    truths = np.stack([P[p]["truth"] for p in cp])
    ang, tr = pose_gap_deg_t(out[64][0], truths)
    res["lm_f64_to_truth_deg_t_q50_q90_max"] = [[float(x) for x in np.quantile(ang, [0.5, 0.9, 1.0])],
                                                [float(x) for x in np.quantile(tr, [0.5, 0.9, 1.0])]]
    ang0, tr0 = pose_gap_deg_t(T0, truths)
    res["base_to_truth_deg_t_max"] = [float(ang0.max()), float(tr0.max())]
    if not args.no_cpu:
        # The FP64 oracle's LM on the first 4 poses of every seam (all host threads over poses).
        sel, To, status = [], [], []
        t0 = time.perf_counter()
        for p in range(len(P)):
            ii = np.nonzero(cp == p)[0][:4]
            T, _, _, st = O.huber_lm(P[p]["source"], P[p]["target"], P[p]["target_n"], w["radius"], T0[ii],
                                     w["delta"], iters)
            sel += list(ii)
            To.append(T)
            status.append(st)
        el_cpu = time.perf_counter() - t0
        To, status, sel = np.concatenate(To), np.concatenate(status), np.array(sel)
        conv = (status == O.LM_CONVERGED) | (status == O.LM_DAMPED_OUT)
        for prec in (64, 32):
            a, d = pose_gap_deg_t(out[prec][0][sel], To)
            a = np.radians(a)
            res[f"lm_f{prec}_vs_oracle_rad_t_max_converged"] = [float(a[conv].max()), float(d[conv].max())]
            res[f"lm_f{prec}_vs_oracle_rad_t_max_all"] = [float(a.max()), float(d.max())]
        res["oracle_lm_poses"] = int(len(sel))
        res["oracle_lm_not_converged_share"] = round(float(1.0 - conv.mean()), 4)
        res["oracle_lm_poses_per_s"] = round(len(sel) / el_cpu, 2)
        res["oracle_lm_sample"] = (f"oracle/huber_lm.cpp (FP64) on the first 4 base poses of each of the {len(P)} "
                                   f"seams, {iters} iterations, OpenMP over poses")
    log(f"huber: {k} poses on {len(P)} seams, {pairs_used} corresponding points; f32 eval {res['f32_eval_ms']} ms, "
        f"f64 eval {res['f64_eval_ms']} ms; LM f64 {res['lm_f64_ms']} ms, f32 {res['lm_f32_ms']} ms"
        + (f"; f32 LM vs oracle {res['lm_f32_vs_oracle_rad_t_max_converged']}" if not args.no_cpu else ""))
    b.close()
    for s_, t_, _ in clouds:
        s_.close()
        t_.close()
    res["workload"] = ("vessel (configuration 2) seams: per adjacent pair, fracture-face and shell-band samples "
                       "on each side, 10k-40k points per side log-uniform, 512 base poses within 1 deg / 0.1 t, "
                       "delta 0.05 t, radius 0.3 t")
    return res


def _surface_join(n, seed=51):
    """One R §9-shaped join of n points a side: a smooth height field sampled every 0.05 (wall 1),
    the moving copy displaced by 0.05 degrees and 0.01."""
    rng = np.random.default_rng(seed)
    side = int(np.ceil(np.sqrt(n)))
    h = 0.05
    x, y = np.meshgrid(np.arange(side) * h, np.arange(side) * h, indexing="ij")
    x, y = x.ravel()[:n], y.ravel()[:n]
    z = 0.4 * np.sin(x / 1.7) * np.cos(y / 2.3) + 0.1 * np.sin(x * y / 9.0)
    gx = 0.4 / 1.7 * np.cos(x / 1.7) * np.cos(y / 2.3) + 0.1 * y / 9.0 * np.cos(x * y / 9.0)
    gy = -0.4 / 2.3 * np.sin(x / 1.7) * np.sin(y / 2.3) + 0.1 * x / 9.0 * np.cos(x * y / 9.0)
    pts = np.stack([x, y, z], axis=1)
    nrm = np.stack([-gx, -gy, np.ones_like(x)], axis=1)
    nrm /= np.linalg.norm(nrm, axis=1)[:, None]
    a = rng.normal(size=3)
    a /= np.linalg.norm(a)
    ang = np.radians(0.05)
    K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    D = np.eye(4)
    D[:3, :3] = np.eye(3) + np.sin(ang) * K + (1 - np.cos(ang)) * K @ K
    D[:3, 3] = rng.normal(size=3) * 0.01 / np.sqrt(3)
    return pts, nrm, D, h


def segment_section(ex, col, args):
    """R §3.4.4-3.4.8 (segment.rs:544-635) on each of the vessel's fragments: the cone vote's mask
    (good < 5 of seven rays, the face normal standing in for the smoothed one) cleaned on the device
    -- majority filter, islands, reference normals, growth, islands -- one call per fragment (the call
    uploads the f64 mesh, the vote and the voxel grid); and, as the CPU baseline of this row, the
    oracle and the device on a bounded sample (a 20 t high vessel fragment, same resolution)."""
    from sherd_hip import synth

    frags, t = col["fragments"], col["t"]
    inputs = []
    for fr in frags:
        V = fr["V"].astype(np.float64)
        F = fr["F"]
        cen = ((V[F[:, 0]] + V[F[:, 1]]) + V[F[:, 2]]) / 3.0
        rep, near = synth.coarse_grid(cen, t / 8.0)
        sc = ex.scene(fr["V"].astype(np.float32), F)
        good = ex.cone_votes(sc, V, F, None, fr["face_normals"], t)
        sc.close()
        inputs.append((V, F, good < 5, rep, near))
    V, F, raw, rep, near = inputs[0]
    ex.segment_cleanup(V, F, t, raw, rep, near)
    ex.profile(True)
    t0 = time.perf_counter()
    out = [ex.segment_cleanup(V, F, t, raw, rep, near) for (V, F, raw, rep, near) in inputs]
    el = time.perf_counter() - t0
    prof = ex.profile_read()
    ex.profile(False)
    faces = sum(len(i[1]) for i in inputs)
    res = dict(fragments=len(frags), faces=int(faces), representatives=int(sum(len(i[3]) for i in inputs)),
               seconds=round(el, 4), faces_per_s=round(faces / el, 1),
               phase_ms={k: round(v["ms"], 3) for k, v in prof.items() if k.startswith("k_seg") and v["ms"] > 0},
               fracture_fraction_vote_cleaned=[[round(float(np.mean(i[2])), 4), round(float(np.mean(o)), 4)]
                                               for i, o in zip(inputs, out)],
               agreement_with_truth=[round(float(np.mean(o == fr["frac"])), 4) for o, fr in zip(out, frags)])
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(HERE, "oracle"))
        import segment as SG

        small = synth.vessel(t=t, edges_per_t=args.edges_per_t, n_frag=4, seed=5, height_t=20.0, rmin_t=10.0,
                             rmax_t=16.0)["fragments"][0]
        V = small["V"].astype(np.float64)
        F = small["F"]
        _, _, cen = SG.face_geometry(V, F)
        rep, near = synth.coarse_grid(cen, t / 8.0)
        raw = small["frac"].copy()
        flip = np.random.default_rng(7).random(len(raw)) < 0.2
        raw[flip] = ~raw[flip]
        ex.segment_cleanup(V, F, t, raw, rep, near)
        t0 = time.perf_counter()
        got = ex.segment_cleanup(V, F, t, raw, rep, near)
        dev = time.perf_counter() - t0
        t0 = time.perf_counter()
        want = SG.cleanup(V, F, t, raw, rep, near)["final"]
        cpu_s = time.perf_counter() - t0
        res.update(cpu_sample=f"oracle cleanup (numpy + Python loops, 1 thread) of a {len(F)}-face vessel fragment "
                              f"(20 t high, 20 % of the vote flipped)",
                   cpu_seconds=round(cpu_s, 3), device_seconds_on_sample=round(dev, 4),
                   identical_to_oracle_on_sample=bool(np.array_equal(got, want)))
    log(f"segment: {len(frags)} fragments, {faces} faces in {1e3 * el:.1f} ms"
        + (f"; sample {res['device_seconds_on_sample'] * 1e3:.1f} ms vs oracle {res['cpu_seconds']:.1f} s, identical "
           f"{res['identical_to_oracle_on_sample']}" if "cpu_seconds" in res else ""))
    return res


def thickness_section(ex, col, args):
    """R §3.2 (thickness.rs:100-161) on each of the vessel's fragments: one inward ray from every
    face of the stride (every ceil(F / 300 000)-th face), the look-back filter and the two
    histogram modes, one call per fragment (the call uploads the f64 mesh); and, as the CPU
    baseline of this row, the oracle's filter and modes over the same rays on the host."""
    frags = col["fragments"]
    scenes = [ex.scene(f["V"].astype(np.float32), f["F"]) for f in frags]
    for sc, f in zip(scenes, frags):
        ex.thickness(sc, f["V"], f["F"])
    ex.profile(True)
    t0 = time.perf_counter()
    out = [ex.thickness(sc, f["V"], f["F"]) for sc, f in zip(scenes, frags)]
    el = time.perf_counter() - t0
    prof = ex.profile_read()
    ex.profile(False)
    strides = [-(-len(f["F"]) // 300_000) for f in frags]
    rays = sum(-(-len(f["F"]) // s_) for f, s_ in zip(frags, strides))
    ray_ms = prof.get("k_thick_rays", {}).get("ms", float("nan"))
    modes_ms = prof.get("k_thick_modes", {}).get("ms", float("nan"))
    res = dict(fragments=len(frags), rays=int(rays), seconds=round(el, 4),
               fragments_per_s=round(len(frags) / el, 2), ray_kernel_ms=round(ray_ms, 3),
               modes_ms=round(modes_ms, 3),
               rays_per_s=round(rays / (ray_ms * 1e-3), 1),
               t_over_wall=[round(float(o[0]) / col["t"], 4) if o else None for o in out])
    if not args.no_cpu:
        # The host's share (filter and modes) over fragment 0's rays, cast again through the
        # device's first hit so that both sides see the same hits.
        sys.path.insert(0, os.path.join(HERE, "oracle"))
        import thickness as TH

        f = frags[0]
        idx = TH.face_indices(len(f["F"]))
        V, F = np.asarray(f["V"], dtype=np.float64), np.asarray(f["F"]).astype(np.int64)  # (the device's f64 too)
        a, b, c = V[F[:, 0]], V[F[:, 1]], V[F[:, 2]]
        e1, e2 = b - a, c - a
        n = np.stack([e1[:, 1] * e2[:, 2] - e1[:, 2] * e2[:, 1], e1[:, 2] * e2[:, 0] - e1[:, 0] * e2[:, 2],
                      e1[:, 0] * e2[:, 1] - e1[:, 1] * e2[:, 0]], axis=1)
        ln = np.sqrt((n[:, 0] * n[:, 0] + n[:, 1] * n[:, 1]) + n[:, 2] * n[:, 2])
        n = n / np.maximum(ln, 1e-12)[:, None]
        cen = ((a + b) + c) / 3.0
        d = -n[idx]
        t_hit, prim = ex.first_hit(scenes[0], (cen[idx] + d * 1e-3).astype(np.float32), d.astype(np.float32))
        t0 = time.perf_counter()
        want = TH.thickness_from_hits(n, idx, t_hit, prim)
        res["cpu_modes_seconds"] = round(time.perf_counter() - t0, 4)
        res["identical_to_oracle_on_fragment_0"] = bool(want is not None and out[0] is not None and
                                                        want[0].tobytes() == out[0][0].tobytes() and
                                                        want[1].tobytes() == out[0][1].tobytes())
    res["workload"] = "vessel, 8 fragments: one inward ray per face of the 300 000-ray stride, two f32 histogram modes"
    log(f"thickness: {len(frags)} fragments, {rays} rays in {el * 1e3:.1f} ms ({ray_ms:.2f} ms of ray kernel); "
        f"t/wall {res['t_over_wall'][0]}; identical to the oracle {res.get('identical_to_oracle_on_fragment_0')}")
    for sc in scenes:
        sc.close()
    return res


def refine_section(ex, col, args):
    """R §9 (refine.rs:270-350) on the vessel: each fragment's fracture cloud (its mesh serving as
    its own original), assembly poses off the truth by 0.02 degrees about the fragment's centroid
    and 0.02 t (a pose right to a fraction of a wall, as R §8 leaves it), the walk over the
    adjacent joins; the device as it runs by itself (the wide ICP form for these batches of one),
    the one-workgroup-per-candidate form, and the oracle on the host. Then one join of 150 000
    points a side, the size R §9 caps a cloud at."""
    from sherd_hip import face_centroids

    frags = col["fragments"]
    t, res = col["t"], col["res"]
    t0 = time.perf_counter()
    clouds = [ex.fracture_cloud(f["V"], f["F"], face_centroids(f["V"], f["F"]), f["frac"], t, res) for f in frags]
    el_clouds = time.perf_counter() - t0
    rng = np.random.default_rng(50)
    poses = []
    for f, P in zip(frags, col["poses"]):
        W = np.linalg.inv(P)
        c = W[:3, :3] @ f["V"].mean(axis=0) + W[:3, 3]
        a = rng.normal(size=3)
        a /= np.linalg.norm(a)
        K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
        ang = np.radians(0.02)
        D = np.eye(4)
        D[:3, :3] = np.eye(3) + np.sin(ang) * K + (1 - np.cos(ang)) * K @ K
        v = rng.normal(size=3)
        D[:3, 3] = c - D[:3, :3] @ c + 0.02 * t * v / np.linalg.norm(v)
        poses.append(D @ W)
    poses = np.array(poses)
    pieces = [(t, res, c) for c in clouds]
    groups = [list(range(len(frags)))]
    used = sorted(tuple(sorted(p)) for p in col["adjacent"])

    def gap(W):
        # Relative pose error of every adjacent pair against the truth, degrees and t.
        worst = [0.0, 0.0]
        for a, b in used:
            E = (W[a] @ col["poses"][a]) @ np.linalg.inv(W[b] @ col["poses"][b])
            c = np.clip((np.trace(E[:3, :3]) - 1.0) / 2.0, -1.0, 1.0)
            worst[0] = max(worst[0], float(np.degrees(np.arccos(c))))
            worst[1] = max(worst[1], float(np.abs(E[:3, 3]).max() / t))
        return worst

    ex.refine_joins(pieces, poses, groups, used)
    t0 = time.perf_counter()
    out, joins = ex.refine_joins(pieces, poses, groups, used)
    el = time.perf_counter() - t0
    ex.icp_block(512)
    ex.refine_joins(pieces, poses, groups, used)
    t0 = time.perf_counter()
    out1, joins1 = ex.refine_joins(pieces, poses, groups, used)
    el1 = time.perf_counter() - t0
    ex.icp_block(0)
    same = bool(np.array_equal(out, out1))
    sizes = [c.n for c in clouds]
    log(f"refine: {len(joins)} joins, fracture clouds {min(sizes)}-{max(sizes)} points (built in {el_clouds:.3f}s); "
        f"device {el * 1e3:.1f} ms, one workgroup per join {el1 * 1e3:.1f} ms, identical {same}")
    res_ = dict(joins=len(joins), cloud_points=[int(min(sizes)), int(max(sizes))],
                fracture_clouds_seconds=round(el_clouds, 4), seconds=round(el, 4),
                joins_per_s=round(len(joins) / el, 2), one_workgroup_seconds=round(el1, 4),
                wide_vs_one_workgroup=round(el1 / el, 2), identical_wide_one_workgroup=same,
                adjacent_gap_before_deg_t=gap(poses), adjacent_gap_after_deg_t=gap(out),
                workload=f"vessel, {len(frags)} fragments, walk over {len(used)} adjacent joins, "
                         "2 point-to-plane rungs x 40 iterations")
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(HERE, "oracle"))
        import refine as R

        host = [(t, res, c.read()[:2]) for c in clouds]
        t0 = time.perf_counter()
        oout, ojoins = R.refine_joins(host, poses, groups, used)
        el_cpu = time.perf_counter() - t0
        dev = [np.degrees(np.arccos(np.clip((np.trace(out[i][:3, :3] @ oout[i][:3, :3].T) - 1) / 2, -1, 1)))
               for i in range(len(out))]
        res_["cpu_seconds"] = round(el_cpu, 3)
        res_["cpu_joins_per_s"] = round(len(ojoins) / el_cpu, 3)
        res_["cpu_vs_device_pose_deg_max"] = float(max(dev))
        res_["cpu_sample"] = "the oracle's refine_joins (icp_rung, one thread per candidate) over the same walk"
        log(f"refine oracle: {el_cpu:.2f}s; worst pose difference to the device {max(dev):.2e} deg")
    for c in clouds:
        c.close()
    # One join at R §9's cap: 150 000 points a side, a batch of one.
    pts, nrm, D, h = _surface_join(150_000)
    src = ex.cloud(pts, nrm)
    big = [(1.0, h, src), (1.0, h, src)]
    bp = np.stack([np.eye(4), D])
    ex.refine_joins(big, bp, [[0, 1]], [(0, 1)])
    t0 = time.perf_counter()
    bout, bj = ex.refine_joins(big, bp, [[0, 1]], [(0, 1)])
    el_big = time.perf_counter() - t0
    ex.icp_block(512)
    t0 = time.perf_counter()
    bout1, _ = ex.refine_joins(big, bp, [[0, 1]], [(0, 1)])
    el_big1 = time.perf_counter() - t0
    ex.icp_block(0)
    err = np.abs(bout[1] - np.eye(4)).max()
    log(f"refine 150k join: device {el_big * 1e3:.1f} ms, one workgroup {el_big1 * 1e3:.1f} ms, identical "
        f"{bool(np.array_equal(bout, bout1))}; pose back to identity within {err:.1e}")
    res_["join_150k"] = dict(seconds=round(el_big, 4), one_workgroup_seconds=round(el_big1, 4),
                             wide_vs_one_workgroup=round(el_big1 / el_big, 2),
                             identical=bool(np.array_equal(bout, bout1)), fitness=bj[0]["fitness"],
                             pose_error_max=float(err))
    src.close()
    return res_


def match_section(ex, col, args):
    """R §5.1-5.5 (pair.rs:188-264) over the vessel's 28 pairs: hypotheses from the breakline frames,
    coarse scores on a 60-point probe, the walk and suppression (250 kept), the breakline ladder
    from every kept hypothesis, its re-score and the second suppression (10 kept): one batched call
    against one call per pair (the reference's schedule), and the oracle on one pair."""
    frags = col["fragments"]
    scales = synth_scales(col)
    frames = [ex.frames(f["brk_P"], f["brk_ns"], f["brk_f"], f["brk_t"], f["brk_dih"], f["brk_sub"]) for f in frags]
    srcs = [ex.cloud(f["brk_P"][f["brk_sub"]]) for f in frags]
    tgts = [ex.cloud(f["brk_P"], f["brk_ns"]) for f in frags]
    rng = np.random.default_rng(60)
    probes = [np.sort(rng.choice(f["brk_sub"], size=min(60, len(f["brk_sub"])), replace=False)) for f in frags]
    pairs = [(a, b) for a in range(len(frags)) for b in range(a + 1, len(frags))]
    mp = [dict(a=frames[a], b=frames[b], probe_p=frags[b]["brk_P"][probes[b]], probe_n=frags[b]["brk_ns"][probes[b]],
               s1_source=srcs[b], s1_target=tgts[a], scales=scales) for a, b in pairs]
    ex.match_stage1(mp)
    t0 = time.perf_counter()
    out = ex.match_stage1(mp)
    el = time.perf_counter() - t0
    t0 = time.perf_counter()
    plain = [ex.match_stage1([m])[0] for m in mp]
    el_plain = time.perf_counter() - t0
    same = all(np.array_equal(x["kept1"], y["kept1"]) and np.array_equal(x["kept2"], y["kept2"]) and
               np.array_equal(x["s1_T"], y["s1_T"]) for x, y in zip(out, plain))
    n_hyp = sum(o["n_hyp"] for o in out)
    # How close the best of the ten stage-2 inits comes to the true relative pose, on adjacent pairs.
    gaps = []
    for (a, b), o in zip(pairs, out):
        if (a, b) not in col["adjacent"] or len(o["kept2"]) == 0:
            continue
        truth = col["poses"][a] @ np.linalg.inv(col["poses"][b])
        best = np.inf
        for k in o["kept2"]:
            T = o["s1_T"][k]
            E = T @ np.linalg.inv(truth)
            ang = np.degrees(np.arccos(np.clip((np.trace(E[:3, :3]) - 1) / 2, -1, 1)))
            best = min(best, max(ang / 5.0, np.abs(E[:3, 3]).max() / col["t"]))
        gaps.append(best)
    log(f"match stage 1: {len(pairs)} pairs, {n_hyp} hypotheses in {el * 1e3:.1f} ms batched, "
        f"{el_plain * 1e3:.1f} ms one pair per call; identical {same}")
    res = dict(pairs=len(pairs), hypotheses=int(n_hyp), seconds=round(el, 4),
               hypotheses_per_s=round(n_hyp / el, 1), plain_seconds=round(el_plain, 4),
               batched_vs_plain=round(el_plain / el, 3), identical_plain_batched=bool(same),
               kept1=int(sum(len(o["kept1"]) for o in out)), kept2=int(sum(len(o["kept2"]) for o in out)),
               adjacent_pairs_with_an_init_within_5deg_1t=int(sum(g <= 1.0 for g in gaps)),
               adjacent_pairs=len(gaps),
               workload="vessel breaklines (frames from the meshes, 2 t subset), 60-point probes, "
                        "250 / 10 kept, rungs 0.2 / 0.08 t x 20")
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(HERE, "oracle"))
        import stage as G

        a, b = pairs[0]
        fa, fb = frags[a], frags[b]
        A = dict(p=fa["brk_P"], ns=fa["brk_ns"], f=fa["brk_f"], t=fa["brk_t"], dih=fa["brk_dih"], sub=fa["brk_sub"])
        B = dict(p=fb["brk_P"], ns=fb["brk_ns"], f=fb["brk_f"], t=fb["brk_t"], dih=fb["brk_dih"], sub=fb["brk_sub"])
        t0 = time.perf_counter()
        o = G.stage1(A, B, A["sub"], B["sub"], probes[b], fb["brk_P"][fb["brk_sub"]], fa["brk_P"], scales)
        el_cpu = time.perf_counter() - t0
        agree = bool(list(o["kept1"]) == list(out[0]["kept1"]) and list(o["kept2"]) == list(out[0]["kept2"]))
        res["cpu_seconds_one_pair"] = round(el_cpu, 3)
        res["cpu_hypotheses_per_s"] = round(len(o["pa"]) / el_cpu, 1)
        res["cpu_kept_lists_identical"] = agree
        res["cpu_sample"] = "oracle stage1 (numpy + the C oracle's coarse and ICP) on the first pair"
        log(f"match oracle: {el_cpu:.2f}s for one pair; kept lists identical {agree}")
    for c in srcs + tgts:
        c.close()
    for f in frames:
        f.close()
    return res


def synth_scales(col):
    from sherd_hip import synth

    return synth.scales_for_pair(col["t"], col["res"])


def pose_gap_deg_t(a, b):
    """Rotation gap in degrees (from |Ra Rb^T - I|_F, exact to rounding at small angles) and
    translation gap between two stacks of 4 x 4 poses."""
    dr = a[:, :3, :3] @ np.transpose(b[:, :3, :3], (0, 2, 1)) - np.eye(3)
    ang = 2.0 * np.arcsin(np.clip(np.linalg.norm(dr, axis=(1, 2)) / (2.0 * np.sqrt(2.0)), 0, 1))
    return np.degrees(ang), np.linalg.norm(a[:, :3, 3] - b[:, :3, 3], axis=1)


def host_info():
    """This host's CPU model (lscpu's "Model name") and core count (nproc)."""
    import subprocess

    model = None
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
                break
    except Exception:  # noqa: BLE001 (lscpu missing: the model stays unknown)
        pass
    return model, os.cpu_count() or 1


def collection_oracle(coll, pairs, cp, cT, scales, sample, threads, repeats=5):
    """The oracle's end-to-end pipeline per candidate (the breakline ladder, the FP64 LM, verify with
    the FP64 retest), threaded over candidates as Pair::stage2_batch's rayon loop is
    (crates/sherd-core/src/matching/pair.rs:350-359): `threads` host threads, each running whole
    candidates with the oracle's own OpenMP at one thread. Returns (candidates/s median of `repeats`,
    the per-run rates, the last run's (T_icp, T_lm, scores) per sampled candidate)."""
    from concurrent.futures import ThreadPoolExecutor

    sys.path.insert(0, os.path.join(HERE, "tests"))
    import oracle_lib as O
    from sherd_hip import collection as CL

    frags = coll["fragments"]
    need = sorted(set(i for c in sample for i in pairs[cp[c]]))
    keep, side = [], {}
    for i in need:
        fr = frags[i]
        fs, ms = O.Scene(fr["V"], fr["F"], fr["frac"]), O.Scene(fr["V"], fr["F"])
        keep += [fs, ms]
        side[i] = O.SideData(fs, ms, fr["watertight"], fr["frac_area"], fr["Pf"], fr["S"], fr["brk_P"],
                              fr["brk_ns"], fr["margin_p"], fr["margin_n"])

    def one(c):
        a, b = pairs[cp[c]]
        sc = scales[cp[c]]
        fa, fb = frags[a], frags[b]
        T = cT[c][None]
        for k in CL.RUNGS:
            T = O.icp_rung(fb["brk_P"], fa["brk_P"], fa["brk_ns"], T, 0, k * sc[0] * sc[11], CL.ITERATIONS)[:, :16]
            T = T.reshape(-1, 4, 4)
        T1 = T[0]
        T2 = O.huber_lm(fb["brk_P"], fa["brk_P"], fa["brk_ns"], CL.RUNGS[-1] * sc[0] * sc[11], T, CL.LM_DELTA,
                        CL.LM_ITERATIONS, threads=1)[0][0]
        return T1, T2, O.verify(side[a], side[b], T2, sc, retest=True)

    rates, out = [], None
    with ThreadPoolExecutor(max_workers=threads, initializer=O.set_threads, initargs=(1,)) as pool:
        for _ in range(repeats):
            t0 = time.perf_counter()
            out = list(pool.map(one, sample))
            rates.append(len(sample) / (time.perf_counter() - t0))
    return float(np.median(rates)), rates, out


def collection_headline(ex, args):
    """BASELINE configuration 5 end to end, the headline: the 40-fragment, three-vessel collection
    (walls t = 0.8 / 1.0 / 1.3 at 4-6 edges per t, two fragments non-watertight), all 780 pairs x 4
    candidate poses. One step = the batched pipeline over every candidate: the breakline ICP
    ladder (one launch per rung, graph-replayed), the Huber LM (f32 sums, FP64 6x6 solve), and R §6's
    verification with the FP64 retest (sherd_hip/collection.py). The plain HIP schedule (one call per
    pair and rung, per pair, per candidate) is timed over its own steps on the same candidates, with
    every result checked identical. Setup (scenes, parity grids, distance bounds, clouds) is timed
    apart and stated."""
    from sherd_hip import collection as CL

    t0 = time.perf_counter()
    coll = CL.build()
    pairs, cp, cT = CL.candidates(coll)
    scales = CL.pair_scales(coll, pairs)
    gen_s = time.perf_counter() - t0
    n_tri = int(sum(len(f["F"]) for f in coll["fragments"]))
    t0 = time.perf_counter()
    dev = CL.Device(ex, coll)
    setup_s = time.perf_counter() - t0
    log(f"collection: {len(coll['fragments'])} fragments, {n_tri} triangles, {len(pairs)} pairs, {len(cT)} "
        f"candidates; generated in {gen_s:.1f}s, setup on the device {setup_s:.2f}s")
    for _ in range(max(args.warmup, 2)):  # the first pass tunes the ICP workgroups, the second captures
        CL.run_batched(dev, coll, pairs, cp, cT, scales)
    ex.profile(True)
    r0 = ex.stats("verify_rays")["items"]
    t0 = time.perf_counter()
    parts = np.zeros(3)
    for _ in range(args.steps):
        b = CL.run_batched(dev, coll, pairs, cp, cT, scales)
        parts += [b["icp_s"], b["lm_s"], b["verify_s"]]
    el = time.perf_counter() - t0
    rays = (ex.stats("verify_rays")["items"] - r0) / args.steps  # parity rays per step
    prof = ex.profile_read()
    ex.profile(False)
    n_cand = len(cT)
    value = n_cand * args.steps / el
    log(f"collection batched: {args.steps} steps in {el:.3f}s ({1e3 * el / args.steps:.2f} ms/step; icp "
        f"{1e3 * parts[0] / args.steps:.2f}, lm {1e3 * parts[1] / args.steps:.2f}, verify "
        f"{1e3 * parts[2] / args.steps:.2f} ms)")
    # The plain schedule over its own timed steps, after one untimed pass (none with --plain-steps 0,
    # for counter collection).
    if args.plain_steps <= 0:
        p, el_plain, plain, plain_rays = b, float("nan"), float("nan"), float("nan")
    else:
        CL.run_plain(dev, coll, pairs, cp, cT, scales)
        r0 = ex.stats("verify_rays")["items"]
        t0 = time.perf_counter()
        for _ in range(args.plain_steps):
            p = CL.run_plain(dev, coll, pairs, cp, cT, scales)
        el_plain = time.perf_counter() - t0
        plain = n_cand * args.plain_steps / el_plain
        plain_rays = (ex.stats("verify_rays")["items"] - r0) / args.plain_steps
    same = dict(icp=bool(np.array_equal(b["reg"], p["reg"])), lm=bool(np.array_equal(b["T_lm"], p["T_lm"])),
                scores=bool(np.array_equal(b["scores"].view(np.uint64), p["scores"].view(np.uint64))),
                accepted=bool(np.array_equal(b["accepted"], p["accepted"])))
    log(f"collection plain: {args.plain_steps} steps in {el_plain:.2f}s; batched/plain {value / plain:.1f}x; "
        f"identical {same}")
    log(f"collection rays: {rays:.0f} parity rays per step; {rays / (el / args.steps) / 1e6:.1f} M/s batched, "
        f"{plain_rays / (el_plain / args.plain_steps) / 1e6 if args.plain_steps > 0 else float('nan'):.2f} M/s plain")
    adj = np.array([tuple(pr) in coll["adjacent"] for pr in pairs])[cp]
    res = dict(
        value=value, ms_per_step=1e3 * el / args.steps, candidates=n_cand, prof=prof, coll=coll, pairs=pairs, cp=cp,
        reg=b["reg"], scores=b["scores"],
        section=dict(
            fragments=len(coll["fragments"]), triangles=n_tri, pairs=len(pairs), candidates=int(n_cand),
            non_watertight=list(CL.OPEN_FRAGMENTS), generation_seconds=round(gen_s, 2),
            setup_seconds=round(setup_s, 4),
            setup="scenes (LBVH) of every fragment's mesh and fracture subset, parity grids, distance bounds, "
                  "sample sets, breakline clouds; once per collection, outside the steps",
            step_ms=round(1e3 * el / args.steps, 4),
            step_icp_lm_verify_ms=[round(1e3 * x / args.steps, 4) for x in parts],
            end_to_end_with_setup_seconds=round(setup_s + el / args.steps, 4),
            plain_steps=args.plain_steps,
            plain_step_seconds=round(el_plain / args.plain_steps, 4) if args.plain_steps > 0 else None,
            plain_icp_lm_verify_s=[round(p["icp_s"], 4), round(p["lm_s"], 4), round(p["verify_s"], 4)],
            plain_pair_verifications_per_s=round(plain, 3) if args.plain_steps > 0 else None,
            batched_vs_plain=round(value / plain, 3) if args.plain_steps > 0 else None,
            batched_vs_plain_with_setup=(round((el_plain / args.plain_steps + setup_s) / (el / args.steps + setup_s), 3)
                                         if args.plain_steps > 0 else None),
            target_3x_met=bool(value / plain >= 3.0) if args.plain_steps > 0 else None, identical_plain_batched=same,
            rays=dict(unit="parity rays: 3 per surface sample past the mesh-box reject of R 6.4's inside test "
                           "(the step casts no first-hit ray)", per_step=int(rays),
                      batched_per_s=round(rays / (el / args.steps), 1),
                      plain_per_s=round(plain_rays / (el_plain / args.plain_steps), 1) if args.plain_steps > 0 else None,
                      plain_per_step=int(plain_rays) if args.plain_steps > 0 else None),
            accepted=int(b["accepted"].sum()), accepted_on_adjacent_pairs=int((b["accepted"] & adj).sum()),
            adjacent_candidates=int(adj.sum()), pen_unavailable=int(b["scores"][:, 16].sum())))
    cpu = None
    if not args.no_cpu:
        # A bounded sample: every candidate of the pairs among six fragments (the first four of the
        # first vessel, the first two of the second), so adjacent, same-vessel and cross-vessel pairs
        # all appear; all host cores threaded over candidates, and one thread, median of five each.
        S = {0, 1, 2, 3, 13, 14}
        sample = [c for c in range(n_cand) if set(pairs[cp[c]]) <= S]
        model, nproc = host_info()
        cores = min(nproc, 16)
        v_all, runs_all, out = collection_oracle(coll, pairs, cp, cT, scales, sample, cores)
        v_one, runs_one, _ = collection_oracle(coll, pairs, cp, cT, scales, sample[:8], 1)
        # Parity on the sample: the oracle's ladder against the device's (the transform bar), and
        # verify at the device's final poses bit for bit (checked by the tests; here the oracle's own
        # pipeline's accept verdicts against the device's).
        import oracle_lib as O

        T1o = np.stack([o[0] for o in out])
        T1d = b["reg"][sample, -1, :16].reshape(-1, 4, 4)
        ang, tr = pose_gap_deg_t(T1o, T1d)
        acc_o = np.array([bool(O.accept(o[2], scales[cp[c]])) for o, c in zip(out, sample)])
        cpu = dict(value=round(v_all, 3), unit="pair-verifications/s", cores=cores, kind="port",
                   sample=f"the oracle's end-to-end pipeline (breakline ladder, FP64 LM, verify with the FP64 retest) "
                          f"on all {len(sample)} candidates of the 15 pairs among fragments {sorted(S)} of the same "
                          f"collection, {cores} host threads over candidates (pair.rs:350-359), median of 5; "
                          f"scene builds excluded",
                   one_thread_value=round(v_one, 3), one_thread_sample=f"the first 8 of those candidates, 1 thread, "
                                                                         f"median of 5",
                   runs=[round(x, 3) for x in runs_all], one_thread_runs=[round(x, 3) for x in runs_one],
                   cpu_model=model, nproc=nproc,
                   icp_vs_device_worst_deg_t=[float(ang.max()), float(tr.max())],
                   accept_identical_to_device=bool(np.array_equal(acc_o, b["accepted"][sample])))
        log(f"collection oracle: {v_all:.2f} candidates/s on {cores} threads, {v_one:.2f} on 1 ({model}); ladder "
            f"vs device worst {ang.max():.1e} deg {tr.max():.1e} t; accept identical {cpu['accept_identical_to_device']}")
    res["cpu"] = cpu
    res["dev"] = dev
    return res


def vessel_section(ex, synth, args):
    """BASELINE configuration 2 (plain method on GPU vs CPU), the vessel: 8 fragments at
    --edges-per-t edges per t (5.5: 0.9-1.6 M triangles each, the per-fragment size the configuration
    states; its 17 edges per t would be about 87 M triangles in all), all 28 pairs x 10 candidate
    poses through R §6 in one batch per step, timed over the same steps; beside it the plain
    schedule, the cone-vote rays, kNN queries, the LBVH build rate and the roofline of the inside
    pass (k_pen) with the side-stream sections listed as overlapped."""
    t0 = time.perf_counter()
    col = synth.vessel(edges_per_t=args.edges_per_t)
    frags = col["fragments"]
    n_tri = int(sum(len(f["F"]) for f in frags))
    log(f"vessel: {len(frags)} fragments, {n_tri} triangles generated in {time.perf_counter() - t0:.1f}s")
    scales = synth.scales_for_pair(col["t"], col["res"])
    pairs, cand_pair, cand_T = synth.candidates(col)
    b_sb = ex.stats("scene_build")
    scenes, handles = [], []
    for fr in frags:
        fs = ex.scene(fr["V"], fr["F"], fr["frac"])
        ms = ex.scene(fr["V"], fr["F"])
        scenes.append((fs, ms))
        handles.append(ex.fragment(fs, ms, fr["watertight"], fr["frac_area"], fr["Pf"], fr["S"], fr["brk_P"],
                                   fr["brk_ns"], fr["margin_p"], fr["margin_n"]))
    st = ex.stats("scene_build")
    lbvh_mtri = (st["items"] - b_sb["items"]) / ((st["device_ns"] - b_sb["device_ns"]) * 1e-9) / 1e6
    vpairs = [(handles[a], handles[b], scales) for (a, b) in pairs]
    for _ in range(args.warmup):
        ex.verify_batch(vpairs, cand_pair, cand_T)
    ex.profile(True)
    r0 = ex.stats("verify_rays")["items"]
    t0 = time.perf_counter()
    for _ in range(args.steps):
        scores, accepted = ex.verify_batch(vpairs, cand_pair, cand_T)
    el = time.perf_counter() - t0
    prays = (ex.stats("verify_rays")["items"] - r0) / args.steps  # parity rays per step
    prof = ex.profile_read()
    ex.profile(False)
    n_cand = len(cand_T)
    value = n_cand * args.steps / el
    n_plain = max(1, min(args.plain_candidates, n_cand))
    r0 = ex.stats("verify_rays")["items"]
    t0 = time.perf_counter()
    plain_acc = np.empty(n_plain, bool)
    plain_scores = np.empty((n_plain, 20))
    for i in range(n_plain):
        s1, a1 = ex.verify_batch(vpairs, cand_pair[i:i + 1], cand_T[i:i + 1])
        plain_scores[i], plain_acc[i] = s1[0], a1[0]
    el_plain = time.perf_counter() - t0
    plain = n_plain / el_plain
    plain_prays = ex.stats("verify_rays")["items"] - r0
    same = bool(np.array_equal(plain_acc, accepted[:n_plain]) and
                np.array_equal(plain_scores.view(np.uint64), scores[:n_plain].view(np.uint64)))
    log(f"vessel verify: {value:.0f} pair-verifications/s batched, {plain:.0f} plain; identical {same}")
    # Rays: the seven-ray cone vote over every face.
    before = ex.stats("cone")
    for fr, (fs, ms) in zip(frags, scenes):
        ex.cone_votes(ms, fr["V"].astype(np.float64), fr["F"], None, fr["face_normals"], col["t"])
    after = ex.stats("cone")
    rays = after["items"] - before["items"]
    rays_per_s = rays / ((after["device_ns"] - before["device_ns"]) * 1e-9)
    # kNN queries/s: R §6.3's margin query shape, all margins against one.
    q = np.concatenate([fr["margin_p"] for fr in frags])
    b0 = ex.stats("nearest")
    for _ in range(3):
        ex.nearest_below(frags[0]["margin_p"], q, scales[8])
    b1 = ex.stats("nearest")
    knn = (b1["items"] - b0["items"]) / ((b1["device_ns"] - b0["device_ns"]) * 1e-9)
    log(f"vessel rays: {rays} cone rays at {rays_per_s / 1e6:.1f} M/s; {prays:.0f} parity rays per verify step, "
        f"{prays / (el / args.steps) / 1e6:.1f} M/s batched, {plain_prays / el_plain / 1e6:.2f} M/s plain; "
        f"kNN {knn / 1e6:.1f} M queries/s")
    nbytes = algorithmic_bytes("k_pen", frags, pairs, cand_pair)
    overlapped = {}
    for k in OVERLAPPED_KERNELS:
        if k in prof:
            ms_k = prof[k]["ms"] / prof[k]["launches"]
            b_k = algorithmic_bytes(k, frags, pairs, cand_pair)
            overlapped[k] = dict(avg_launch_ms=round(ms_k, 4), algorithmic_bytes_per_launch=b_k,
                                 achieved_gbs=round(b_k / (ms_k * 1e-3) / 1e9, 3))
    roof = kernel_roofline(INSIDE_PASS, nbytes, prof, "verify_vessel_pmc.json")
    roof["overlapped"] = overlapped
    res = dict(workload=f"vessel: 8 fragments, {args.edges_per_t} edges/t, {n_tri} triangles, 28 pairs x 10 "
                        f"candidates", pair_verifications_per_s=round(value, 3), steps=args.steps,
               ms_per_step=round(1e3 * el / args.steps, 4), plain_pair_verifications_per_s=round(plain, 3),
               plain_candidates=n_plain, batched_vs_plain=round(value / plain, 3), accepted=int(accepted.sum()),
               identical_plain_batched=same, rays_per_s=round(rays_per_s, 1), rays=int(rays),
               parity_rays=dict(unit="3 per surface sample past the mesh-box reject of R 6.4's inside test",
                                per_step=int(prays), batched_per_s=round(prays / (el / args.steps), 1),
                                plain_candidates=n_plain, plain_rays=int(plain_prays),
                                plain_per_s=round(plain_prays / el_plain, 1)),
               knn_queries_per_s=round(knn, 1), lbvh_mtri_per_s=round(lbvh_mtri, 2), kernel_ms=prof, roofline=roof)
    if not args.no_cpu:
        cpu, oracle_scores = cpu_baseline(col, pairs, scales, cand_pair, cand_T)
        k = len(oracle_scores)
        cpu["scores_identical_to_hip"] = bool(np.array_equal(np.array(oracle_scores).view(np.uint64),
                                                             scores[:k].view(np.uint64)))
        res["cpu"] = cpu
    for h in handles:
        h.close()
    for fs, ms in scenes:
        fs.close()
        ms.close()
    return res, col


def kernel_roofline(kernels, nbytes, prof, pmc_file):
    """The roofline object of one pass of consecutive launches on one stream (a tuple of kernel
    names): algorithmic bytes per pass over the sum of the launches' average times (HIP events on
    the library's stream over the timed steps) against the HBM peak; beside it the HBM bytes per
    pass read from a committed rocprofv3 counter summary (FETCH_SIZE, WRITE_SIZE, scaled by the
    copy-kernel calibration) when it holds every launch of the pass, and the bandwidth those
    measured bytes give."""
    launch_ms = {k: prof[k]["ms"] / prof[k]["launches"] for k in kernels}
    avg_ms = sum(launch_ms.values())
    achieved = nbytes / (avg_ms * 1e-3) / 1e9
    traffic, src, per = None, None, {}
    path = os.path.join(HERE, "profiles", pmc_file)
    if os.path.exists(path):
        with open(path) as fh:
            table = json.load(fh).get("kernels", {})
        per = {k: table[k]["bytes_per_launch"] for k in kernels if k in table}
        if len(per) == len(kernels):
            traffic = int(sum(per.values()))
            src = f"profiles/{pmc_file} (rocprofv3 --pmc of the same workload, committed)"
    out = dict(bound="hbm", achieved=round(achieved, 3), peak=HBM_PEAK_GBS, unit="GB/s",
               frac=round(achieved / HBM_PEAK_GBS, 6), traffic=traffic, kernel=" + ".join(kernels),
               algorithmic_bytes_per_launch=int(nbytes), avg_launch_ms=round(avg_ms, 4),
               launch_ms={k: round(v, 4) for k, v in launch_ms.items()})
    if traffic:
        out["traffic_source"] = src
        out["traffic_per_launch"] = per
        out["measured_gbs"] = round(traffic / (avg_ms * 1e-3) / 1e9, 3)
        out["measured_frac"] = round(traffic / (avg_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 6)
    return out


# The roofline of both verification sections is R §6.4's inside pass (the unit of the parity rays):
# k_pen's box reject, k_pen_groups' prefix sum over its lists and k_pen_rays' parity rays, three
# consecutive launches on the main stream, in bytes per point processed (the f64 sample, 24 B) over
# every (candidate, side) task of a watertight pair. k_pen shares the device with the apart
# candidates' closest approach (a side stream), so its launch time is the time it takes beside it.
INSIDE_PASS = ("k_pen", "k_pen_groups", "k_pen_rays")


def collection_roofline(head):
    """The roofline of the collection step's inside pass (INSIDE_PASS); every other kernel of the
    step is listed beside it with its time per launch (HIP events on its own stream)."""
    prof, coll, pairs, cp, scores = head["prof"], head["coll"], head["pairs"], head["cp"], head["scores"]
    frags = coll["fragments"]
    n_s = np.array([len(frags[a]["S"]) + len(frags[b]["S"]) for a, b in pairs])[cp]
    pen_ok = scores[:, 16] == 0.0
    points = int(n_s[pen_ok].sum())
    roof = kernel_roofline(INSIDE_PASS, 24 * points, prof, "collection_pmc.json")
    roof["points_per_launch"] = points
    roof["unit_bytes"] = ("24 B per point processed (the f64 sample) over the inside pass's three launches; "
                          "the meshes the in-box samples' rays scan are in the measured traffic, not the unit")
    main = [k for k in prof if k not in OVERLAPPED_KERNELS]
    roof["share_of_main_stream_ms"] = round(sum(prof[k]["ms"] for k in INSIDE_PASS) /
                                            sum(prof[k]["ms"] for k in main), 4)
    roof["others_ms_per_launch"] = {k: round(v["ms"] / v["launches"], 4) for k, v in prof.items()
                                    if k not in INSIDE_PASS and v["launches"]}
    return roof


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--plain-steps", type=int, default=2,
                    help="timed steps of the plain schedule on the collection (after one untimed)")
    ap.add_argument("--edges-per-t", type=float, default=5.5)
    ap.add_argument("--plain-candidates", type=int, default=280)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-vessel", action="store_true", help="skip configuration 2 (and the sections on its vessel)")
    ap.add_argument("--icp-pairs", type=int, default=28, help="pairs of the batched ICP ladder workload (0: skip)")
    ap.add_argument("--icp-poses", type=int, default=64)
    ap.add_argument("--icp-plain-pairs", type=int, default=-1,
                    help="pairs run in the plain per-call ICP schedule (-1: all, 0: none)")
    ap.add_argument("--no-huber", action="store_true")
    ap.add_argument("--no-slab", action="store_true", help="skip configuration 1 (the slab plate, whole path)")
    ap.add_argument("--no-refine", action="store_true", help="skip R §9's full-resolution refinement")
    ap.add_argument("--no-match", action="store_true", help="skip R §5.1-5.5 over the vessel's pairs")
    ap.add_argument("--no-thickness", action="store_true", help="skip R §3.2 over the vessel's fragments")
    ap.add_argument("--no-segment", action="store_true", help="skip R §3.4.4-3.4.8 over the vessel's fragments")
    ap.add_argument("--collection-only", action="store_true",
                    help="run only the collection's steps (for counter collection); prints no bench line")
    ap.add_argument("--no-collection", action="store_true",
                    help="skip the collection (the headline) and run the sections only (for kernel traces); "
                         "prints the sections as JSON, not a bench line")
    ap.add_argument("--vessel-only", action="store_true",
                    help="run only configuration 2's verification steps (for counter collection); prints no bench line")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the pair list would shard across GPUs by pair, each GPU holding the fragments its "
              "pairs name; this benchmark times one GPU's share)", file=sys.stderr)
        sys.exit(2)

    import sherd_hip
    from sherd_hip import synth

    ex = sherd_hip.HipExecutor(0)
    log("device:", ex.device_name())
    st = ex.selftest()  # the gate backend selection reads (selftest.rs:179)
    log(f"selftest: {'passed' if st['passed'] else 'FAILED'}; " + "; ".join(f"{n}: {d}" for n, _, d in st["checks"]))
    selftest = dict(passed=st["passed"], checks={n: dict(passed=ok, detail=d) for n, ok, d in st["checks"]},
                    reduce_bits=hex(st["reduce_device_bits"]), nn_gpu_ms=round(st["gpu_ms"], 4),
                    nn_cpu_ms=round(st["cpu_ms"], 3), speedup=round(st["speedup"], 2), queries=st["queries"])
    if not st["passed"]:
        sys.exit("the self-test failed: this device is not a backend")
    if args.collection_only or args.vessel_only:
        args.no_cpu = True
    if args.vessel_only:
        args.plain_candidates = 1
        vessel, _ = vessel_section(ex, synth, args)
        for k in sorted(vessel["kernel_ms"], key=lambda k: -vessel["kernel_ms"][k]["ms"]):
            v = vessel["kernel_ms"][k]
            log(f"  {k}: {v['ms'] / v['launches']:.4f} ms/launch over {v['launches']} launches")
        ex.close()
        return
    if args.no_collection:
        vessel, col = (None, None) if args.no_vessel else vessel_section(ex, synth, args)
        out = dict(vessel=vessel, slab=slab_section(ex, synth, args) if not args.no_slab else None,
                   icp=icp_ladder(ex, synth, args) if args.icp_pairs > 0 else None,
                   huber=huber_section(ex, synth, col, args) if not args.no_huber and col is not None else None,
                   refine=refine_section(ex, col, args) if not args.no_refine and col is not None else None,
                   match=match_section(ex, col, args) if not args.no_match and col is not None else None,
                   thickness=thickness_section(ex, col, args) if not args.no_thickness and col is not None else None,
                   segment=segment_section(ex, col, args) if not args.no_segment and col is not None else None)
        print(json.dumps(dict(sections=out)), flush=True)
        ex.close()
        return
    head = collection_headline(ex, args)
    roof = collection_roofline(head)
    log(f"collection roofline: {roof['kernel']} {roof['avg_launch_ms']} ms/launch, {roof['achieved']} GB/s "
        f"({roof['frac']:.3f} of peak)")
    head["dev"].close()
    if args.collection_only:
        for k in sorted(head["prof"], key=lambda k: -head["prof"][k]["ms"]):
            v = head["prof"][k]
            log(f"  {k}: {v['ms'] / v['launches']:.4f} ms/launch over {v['launches']} launches")
        ex.close()
        return
    vessel, col = (None, None) if args.no_vessel else vessel_section(ex, synth, args)
    slab = slab_section(ex, synth, args) if not args.no_slab else None
    icp = icp_ladder(ex, synth, args) if args.icp_pairs > 0 else None
    huber = huber_section(ex, synth, col, args) if not args.no_huber and col is not None else None
    refine = refine_section(ex, col, args) if not args.no_refine and col is not None else None
    match = match_section(ex, col, args) if not args.no_match and col is not None else None
    thick = thickness_section(ex, col, args) if not args.no_thickness and col is not None else None
    segm = segment_section(ex, col, args) if not args.no_segment and col is not None else None

    sec = head["section"]
    line = dict(
        metric=METRIC, value=round(head["value"], 3), unit="pair-verifications/s", n_gpus=1, steps=args.steps,
        warmup=args.warmup, ms_per_step=round(head["ms_per_step"], 3), higher_is_better=True, scaling="weak",
        vs_baseline=None, dtype="f32+f64", data="synthetic",
        config=dict(workload=f"collection (configuration 5): 40 fragments of 3 vessels (t 0.8/1.0/1.3 at 4/5/6 "
                             f"edges per t), {sec['triangles']} triangles, fragments 7 and 23 non-watertight; all "
                             f"780 pairs x 4 candidates per step through the batched breakline ICP ladder, the "
                             f"Huber LM and verification with the FP64 retest",
                    candidates_per_step=sec["candidates"], pairs=sec["pairs"], parallelism="single GPU"),
        plain_pair_verifications_per_s=sec["plain_pair_verifications_per_s"], batched_vs_plain=sec["batched_vs_plain"],
        accepted=sec["accepted"], accepted_identical_plain_batched=sec["identical_plain_batched"]["accepted"],
        collection=sec, vessel=vessel, slab=slab, icp=icp, huber=huber, refine=refine, match=match, thickness=thick,
        segment=segm, kernel_ms=head["prof"], roofline=roof, cpu_baseline=head["cpu"], selftest=selftest,
    )
    print(json.dumps(line), flush=True)
    ex.close()


if __name__ == "__main__":
    main()
