"""BASELINE configuration 1 on the device (GPU): the slab plate's per-vertex inward rays against the
oracle bit for bit, and the whole path on one pair (segmentation rays, sampling, R §5.1-5.5's match,
a stage-2 ladder, R §6's verification) finding the seam the generator made."""

import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import oracle_lib as O  # noqa: E402
import segment as SG  # noqa: E402
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


@pytest.mark.parametrize("retest", [False, True])
def test_vertex_votes_bit_exact_against_the_oracle(ex, retest):
    from sherd_hip import synth

    say(f"case: slab plate 8 t x 6 t at 0.15 t, per-vertex rays, retest={retest}")
    col = synth.slab_plate(length_t=8.0, width_t=6.0, edge_t=0.15)
    for fr, P in zip(col["fragments"], col["poses"]):
        V64 = fr["V"].astype(np.float64)
        scene = ex.scene(fr["V"], fr["F"])
        good, frac = ex.vertex_votes(scene, V64, fr["F"], 1.0, retest=retest)
        og, of = SG.vertex_votes(O.Scene(fr["V"], fr["F"]), V64, fr["F"], 1.0, retest=retest)
        assert np.array_equal(good, og) and np.array_equal(frac, of)
        shell = (~fr["frac"]) & (np.abs(fr["face_normals"] @ P[:3, 2]) > 0.9)  # the two plate faces
        say(f"  {len(good)} vertices, {int(good.sum())} vote shell; labels agree with the generator on "
            f"{np.mean(frac[shell] == 0):.4f} of shell faces, {np.mean(frac[fr['frac']] == 1):.4f} of fracture faces")
        assert np.mean(frac[shell] == 0) > 0.9 and np.mean(frac[fr["frac"]] == 1) > 0.8
        scene.close()


def test_slab_pipeline_finds_the_seam(ex):
    """The 60 t x 40 t plate at 0.06 t (the full configuration): the labels from the rays, cleaned by
    R §3.4.4-3.4.8, drive the sampling, R §5.1-5.5 finds stage-1 poses, the stage-2 ladder refines
    them, verification scores them, and the candidate R §5.7 ranks first (seam * tight) is accepted
    within 0.5 degrees and 0.1 t of the generator's relative pose. The plate's straight outer rims
    are fracture to the rays too (they look along the plate), so rim-to-rim contacts are accepted as
    well; the ranking is what separates the seam from them."""
    from sherd_hip import SCORE_KEYS, pipeline, synth

    say("case: slab plate 60 t x 40 t at 0.06 t through the whole path")
    col = synth.slab_plate()
    fa, fb = col["fragments"]
    out = pipeline.run_pair(ex, fa, fb, col["t"], col["res"])
    truth = col["poses"][0] @ np.linalg.inv(col["poses"][1])
    acc = out["accepted"]
    ang, tr = pose_gap(out["T"], np.broadcast_to(truth, out["T"].shape), 1.0)
    key = out["scores"][:, SCORE_KEYS.index("seam")] * out["scores"][:, SCORE_KEYS.index("tight")]
    best = int(out["order"][0])
    say(f"  {len(out['T'])} stage-2 candidates, {int(acc.sum())} accepted; best {best}: {ang[best]:.3f} deg "
        f"{tr[best]:.4f} t, seam*tight {key[best]:.2f} (next {np.sort(key)[-2]:.2f}); "
        f"seconds {dict((k, round(v, 3)) for k, v in out['seconds'].items())}")
    assert np.array_equal(out["order"], np.argsort(-key, kind="stable")[:5])
    assert acc[best] and ang[best] < 0.5 and tr[best] < 0.1
