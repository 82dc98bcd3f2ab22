"""The batching and collection configurations at reduced size on their own generators (GPU):
BASELINE configuration 3 (the breakline ladder over pairs x poses, log-uniform point counts, four
rungs 1.0 -> 0.03 t) against the oracle's ladder rung by rung, and configuration 5 (the collection
end to end: ladder, Huber LM, verification with the FP64 retest) batched == plain == oracle on every
candidate, with a non-watertight fragment among them."""

import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import icp_chaos  # noqa: E402
import oracle_lib as O  # noqa: E402
from test_gpu_match import pose_gap  # noqa: E402

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def ex():
    import sherd_hip

    e = sherd_hip.HipExecutor(0)
    yield e
    e.close()


def test_batched_ladder_workload_against_the_oracle(ex):
    """icp_batch_workload at 4 pairs x 8 poses (2k-60k points per side, log-uniform), rungs 1.0 /
    0.3 / 0.1 / 0.03 t x 20 iterations in one ladder call: every candidate the one-ULP probe finds
    determined (crosscheck.rs:173-215) within 1e-6 rad / 1e-5 t of the oracle's ladder, the chaotic
    ones counted (at most one per pair)."""
    from sherd_hip import synth

    w = synth.icp_batch_workload(n_pairs=4, poses=8)
    P = w["pairs"]
    say(f"case: batched ladder, {len(P)} pairs x 8 poses, sizes {[(len(p['source']), len(p['target'])) for p in P]}")
    clouds = [(ex.cloud(p["source"]), ex.cloud(p["target"], p["target_n"])) for p in P]
    per_rung = [[(s, t, r) for s, t in clouds] for r, _ in w["rungs"]]
    cp = np.repeat(np.arange(len(P), dtype=np.uint32), 8)
    inits = np.concatenate([p["inits"] for p in P])
    out = ex.icp_batch_ladder(per_rung, [("point_to_point", it) for _, it in w["rungs"]], cp, inits)
    T = out[:, -1, :16].reshape(-1, 4, 4)
    chaotic = 0
    for p, pr in enumerate(P):
        sel = cp == p
        rungs = [(pr["source"], pr["target"], pr["target_n"], 0, r, it) for r, it in w["rungs"]]
        det, To = icp_chaos.determined(rungs, inits[sel])
        ang, tr = pose_gap(T[sel], To, 1.0)
        chaotic += int((~det).sum())
        say(f"  pair {p}: {(~det).sum()} chaotic of {det.size}; worst over determined {ang[det].max():.2e} deg "
            f"{tr[det].max():.2e} t")
        assert (~det).sum() <= 1
        assert np.radians(ang[det]).max() < 1e-6 and tr[det].max() < 1e-5
    for s_, t_ in clouds:
        s_.close()
        t_.close()


def test_collection_end_to_end_batched_plain_oracle(ex):
    """collection.py on one small vessel (6 fragments, 3 edges per t, 40 t high) with fragment 2
    opened by a 5 % patch: all 15 pairs x 4 candidates through the ladder, the LM and verification
    with the FP64 retest, batched and plain identical, and the oracle's verification at the final
    poses identical on every candidate (pen_unavailable on the open fragment's pairs)."""
    from sherd_hip import collection as CL

    say("case: collection, 1 vessel x 6 fragments, fragment 2 open, 60 candidates")
    coll = CL.build(height_t=40.0, vessels=[(1.0, 3.0, 6, 17)], open_fragments=(2,), rmin_t=15.0, rmax_t=20.0)
    pairs, cp, cT = CL.candidates(coll)
    scales = CL.pair_scales(coll, pairs)
    dev = CL.Device(ex, coll)
    b = CL.run_batched(dev, coll, pairs, cp, cT, scales)
    p = CL.run_plain(dev, coll, pairs, cp, cT, scales)
    assert np.array_equal(b["reg"], p["reg"]) and np.array_equal(b["T_lm"], p["T_lm"])
    assert np.array_equal(b["scores"].view(np.uint64), p["scores"].view(np.uint64))
    assert np.array_equal(b["accepted"], p["accepted"])
    side = {}
    keep = []
    for i, fr in enumerate(coll["fragments"]):
        fs, ms = O.Scene(fr["V"], fr["F"], fr["frac"]), O.Scene(fr["V"], fr["F"])
        keep += [fs, ms]
        side[i] = O.SideData(fs, ms, fr["watertight"], fr["frac_area"], fr["Pf"], fr["S"], fr["brk_P"], fr["brk_ns"],
                              fr["margin_p"], fr["margin_n"])
    for c in range(len(cT)):
        a_, b_ = pairs[cp[c]]
        want = O.verify(side[a_], side[b_], b["T_lm"][c], scales[cp[c]], retest=True)
        assert np.array_equal(want.view(np.uint64), b["scores"][c].view(np.uint64)), c
        assert bool(b["accepted"][c]) == O.accept(want, scales[cp[c]]), c
    opened = np.array([2 in pr for pr in pairs])[cp]
    assert np.all(b["scores"][opened, 16] == 1.0) and np.all(b["scores"][~opened, 16] == 0.0)
    adj = np.array([tuple(pr) in coll["adjacent"] for pr in pairs])[cp]
    say(f"  accepted {int(b['accepted'].sum())} ({int((b['accepted'] & adj).sum())} on adjacent pairs); "
        f"pen_unavailable on {int(opened.sum())}")
    assert (b["accepted"] & adj).any() and not (b["accepted"] & ~adj).any()
    dev.close()
