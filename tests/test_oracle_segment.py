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
