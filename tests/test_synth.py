"""The synthetic workloads' host helpers against what fixes them: the pair scales every bench and
collection number reads (matching/scales.rs:72-88) against the slab dump's scales.json and the
oracle's restatement, and the slab plate generator's mesh (closed, outward, complementary cut)."""

import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "sherd-refit_amd"))
import oracle_lib as O  # noqa: E402
from sherd_hip import synth  # noqa: E402

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "slab")


def test_scales_for_pair_is_the_slab_dumps():
    p = np.load(os.path.join(GOLD, "pair.npz"))
    t = min(float(p["t_a"]), float(p["t_b"]))
    res = max(float(p["res_a"]), float(p["res_b"]))
    assert list(synth.scales_for_pair(t, res)) == list(p["scales"])


@pytest.mark.parametrize("t", [0.8, 1.0, 1.3, 30.0])
@pytest.mark.parametrize("res", [0.0, 0.01, 0.06, 0.25, 2.0])
def test_scales_for_pair_agrees_with_the_oracle(t, res):
    # Both branches of every max(k t, m res) are taken across this grid.
    assert list(synth.scales_for_pair(t, res)) == list(O.scales_for_pair(t, res))


@pytest.mark.parametrize("rounded", [False, True])
def test_slab_plate_is_closed_outward_and_its_cut_complementary(rounded):
    col = synth.slab_plate(length_t=6.0, width_t=4.0, edge_t=0.2, rounded_rims=rounded)
    cuts = []
    for fr, P in zip(col["fragments"], col["poses"]):
        V, F = fr["V"].astype(np.float64), fr["F"].astype(np.int64)
        e = np.sort(np.concatenate([F[:, [0, 1]], F[:, [1, 2]], F[:, [2, 0]]]), axis=1)
        _, cnt = np.unique(e, axis=0, return_counts=True)
        assert np.all(cnt == 2)  # every edge between exactly two faces
        n = np.cross(V[F[:, 1]] - V[F[:, 0]], V[F[:, 2]] - V[F[:, 0]])
        vol = np.einsum("ij,ij->i", V[F[:, 0]], n).sum() / 6.0
        half = 0.5 * 6.0 * 4.0 * 1.0
        assert half * 0.9 < vol < half * (1.6 if rounded else 1.1)
        # The cut in the generator's frame: B's vertices moved back by its pose.
        Vb = (V - P[:3, 3]) @ P[:3, :3]
        cuts.append(np.unique(np.round(Vb[np.unique(F[fr["frac"]])], 5), axis=0))
    assert len(cuts[0]) == len(cuts[1]) and np.allclose(cuts[0], cuts[1], atol=2e-5)
