"""The R §3.4.4-3.4.8 oracle (oracle/segment.py) against what pins it: the slab dump's own masks
after the majority filter, the island passes and the growth (crates/sherd-parity/src/stages/
segmentation.rs compares the same files), scipy's connected components for the islands, and known
answers for the adjacency's pairing."""

import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import segment as SG  # noqa: E402

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "slab")


def test_adjacency_pairs_consecutive_rows():
    # Two triangles sharing edge (1, 2); a third on the same edge makes a chain of two pairs.
    F = np.array([[0, 1, 2], [2, 1, 3], [1, 2, 4]])
    fa, fb = SG.face_adjacency(F)
    # Rows of edge (1, 2): face 0 slot 1, face 1 slot 0, face 2 slot 0 -> positions 3, 1, 2;
    # stably sorted: 1 (face 1), 2 (face 2), 3 (face 0).
    assert list(zip(fa.tolist(), fb.tolist())) == [(1, 2), (2, 0)]
    fa, fb = SG.face_adjacency(F[:2])
    assert list(zip(fa.tolist(), fb.tolist())) == [(1, 0)]


def test_components_match_scipy():
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import connected_components

    rng = np.random.default_rng(3)
    n = 400
    F = rng.integers(0, 120, size=(n, 3))
    F = F[(F[:, 0] != F[:, 1]) & (F[:, 1] != F[:, 2]) & (F[:, 0] != F[:, 2])]
    fa, fb = SG.face_adjacency(F)
    mask = rng.random(len(F)) < 0.6
    root = SG.components(mask, True, fa, fb)
    keep = mask[fa] & mask[fb]
    g = coo_matrix((np.ones(keep.sum()), (fa[keep], fb[keep])), shape=(len(F), len(F)))
    _, lab = connected_components(g, directed=False)
    ins = np.nonzero(mask)[0]
    for i in ins:
        for j in ins:
            assert (root[i] == root[j]) == (lab[i] == lab[j])
    assert np.all(root[~mask] == -1)
    assert all(root[i] == min(ins[lab[ins] == lab[i]]) for i in ins)


def test_exact_area_sums_are_exact_integers():
    e = SG.area_exponent(3.0)
    assert 2.0 ** (e - 1) <= 9.0 < 2.0 ** e
    assert SG.quantized(9.0, e) == 9 * 2 ** (64 - e)
    assert SG.quantized(0.1, e) == int(np.floor(np.ldexp(0.1, 64 - e)))


@pytest.fixture(scope="module")
def seg():
    return np.load(os.path.join(GOLD, "seg.npz"))


@pytest.mark.parametrize("name", ["pieceA", "pieceB"])
def test_the_slab_dumps_masks_step_by_step(seg, name):
    g = np.load(os.path.join(GOLD, name + ".npz"))
    V, F, t = g["V"], g["F"], float(g["thick_t"])
    out = SG.cleanup(V, F, t, seg[name + "_frac_raw"], seg[name + "_rep"], seg[name + "_near"])
    assert np.array_equal(out["majority"], seg[name + "_frac_majority"])
    assert np.array_equal(out["islands"], seg[name + "_frac_islands"])
    assert np.array_equal(out["has_ref"], seg[name + "_has_ref"])
    h = seg[name + "_has_ref"]
    cosang = np.sum(out["reference"][h] * seg[name + "_ref"][h], axis=1)
    assert cosang.min() > np.cos(np.radians(1e-5))  # the dump sums its balls in another order
    assert np.array_equal(out["final"], seg[name + "_frac_final"])
    # The exact island sums and the reference's face-order f64 sums decide alike here.
    out2 = SG.cleanup(V, F, t, seg[name + "_frac_raw"], seg[name + "_rep"], seg[name + "_near"], exact=False)
    assert np.array_equal(out["final"], out2["final"]) and np.array_equal(out["islands"], out2["islands"])


def test_vertex_votes_known_answers_on_a_plate():
    """The per-vertex rays on a flat plate of wall 1: a shell vertex away from the rim looks straight
    through the wall (hit at 1.0, the far face looking back) and votes; a vertex on the middle of a rim
    looks along the plate (hit 8 or more away, beyond 1.8 t) and does not; a shell vertex's ray at
    0.5 t (a plate of wall 0.5 read with t = 1) falls out of the window. Faces take the vertices'
    majority."""
    import oracle_lib as O
    from sherd_hip import synth

    col = synth.slab_plate(length_t=8.0, width_t=6.0, edge_t=0.25, amp_t=0.0, jitter_t=0.0, max_deg=1e-9,
                           max_shift_t=0.0, rounded_rims=False)
    fr = col["fragments"][0]
    V, F = fr["V"].astype(np.float64), fr["F"].astype(np.int64)
    good, frac = SG.vertex_votes(O.Scene(fr["V"], fr["F"]), V, F, 1.0)
    top = np.isclose(V[:, 2], 0.5) & (V[:, 0] > 0.6) & (V[:, 0] < 3.4) & (np.abs(V[:, 1]) < 2.4)
    rim_mid = np.isclose(V[:, 1], 3.0) & np.isclose(V[:, 2], 0.0, atol=0.13) & (V[:, 0] > 1.0) & (V[:, 0] < 3.0)
    assert top.sum() > 50 and rim_mid.sum() > 3
    assert good[top].all() and not good[rim_mid].any()
    cen = V[F].mean(axis=1)
    inner_top = np.isclose(cen[:, 2], 0.5) & (cen[:, 0] > 0.6) & (cen[:, 0] < 3.4) & (np.abs(cen[:, 1]) < 2.4)
    # Fracture faces off the shell edges (no vertex on a shell, whose edge normals lean 45 degrees
    # and see the other shell at sqrt(2) t) are all labelled fracture.
    off_edge = fr["frac"] & (np.abs(V[F][:, :, 2]).max(axis=1) < 0.49)
    assert not frac[inner_top].any() and off_edge.sum() > 20 and frac[off_edge].all()
    # Read as a wall of 2 t: the hit at 1.0 is 0.5 t, on the window's edge (not inside it).
    good2, _ = SG.vertex_votes(O.Scene(fr["V"], fr["F"]), V, F, 2.0)
    assert not good2[top].any()
