"""The R §3.2 oracle (oracle/thickness.py) against what pins it: numpy's own percentile and
histogram (which the reference's functions restate in float32), and the slab dump's thick.t from
its own rays (crates/sherd-parity/src/stages/thickness.rs's injected mode: bit for bit)."""

import os
import sys

import numpy as np
import pytest

import slab_data as S

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import thickness as TH  # noqa: E402


def ulps(a, b):
    return abs(int(np.float32(a).view(np.int32)) - int(np.float32(b).view(np.int32)))


def test_the_stride_rule():
    assert np.array_equal(TH.face_indices(10), np.arange(10))
    assert np.array_equal(TH.face_indices(300_000), np.arange(300_000))
    assert np.array_equal(TH.face_indices(300_001), np.arange(0, 300_001, 2))
    assert np.array_equal(TH.face_indices(1_000_000), np.arange(0, 1_000_000, 4))


@pytest.mark.parametrize("n", [1, 2, 5, 11, 100, 101, 1001, 20_000])
def test_percentile90_is_numpys(n):
    x = np.random.default_rng(n).gamma(3.0, 2.0, size=n).astype(np.float32)
    p = TH.percentile90(x)
    assert p.dtype == np.float32
    q = np.float32(np.percentile(x, 90))
    if (n - 1) % 10 == 0:
        # An integral virtual index: no interpolation, the order statistic itself.
        assert p == q == np.sort(x)[(n - 1) * 9 // 10]
    else:
        # numpy's interpolation rounding differs between releases by an ulp or two.
        assert ulps(p, q) <= 4


@pytest.mark.parametrize("seed", range(6))
def test_hist_mode_is_numpys_histogram(seed):
    rng = np.random.default_rng(seed)
    x = np.concatenate([rng.normal(3.0, 0.2, 3000), rng.uniform(0.0, 12.0, 2000)]).astype(np.float32)
    p90 = TH.percentile90(x)
    counts, edges = np.histogram(x, bins=60, range=(np.float32(0.0), p90))
    assert edges.dtype == np.float32
    k = int(np.argmax(counts))
    assert TH.hist_mode(x) == np.float32(np.float32(0.5) * (edges[k] + edges[k + 1]))


def test_hist_mode_edge_cases():
    # Every value zero: numpy widens the empty range to (-0.5, 0.5); the zeros fill bin 30.
    z = np.zeros(10, np.float32)
    counts, edges = np.histogram(z, bins=60, range=(np.float32(0.0), np.float32(0.0)))
    assert int(np.argmax(counts)) == 30 and edges[0] == -0.5 and edges[60] == 0.5
    assert TH.hist_mode(z) == np.float32(np.float32(0.5) * (edges[30] + edges[31]))
    # Values exactly on bin edges go where numpy puts them.
    e = TH.bin_edges(np.float32(0.0), np.float32(6.0))
    x = np.concatenate([e[:60], np.full(5, e[7], np.float32), [np.float32(6.0)] * 2]).astype(np.float32)
    x = np.concatenate([x, np.full(200, np.float32(6.0))])
    counts, edges = np.histogram(x, bins=60, range=(np.float32(0.0), TH.percentile90(x)))
    k = int(np.argmax(counts))
    assert TH.hist_mode(x) == np.float32(np.float32(0.5) * (edges[k] + edges[k + 1]))


def test_too_few_hits_is_none():
    n = np.tile([[0.0, 0.0, 1.0]], (120, 1))
    idx = np.arange(120, dtype=np.uint32)
    t = np.full(120, np.inf, np.float32)
    t[:99] = 1.0
    prim = np.where(np.isfinite(t), 0, 0xFFFFFFFF).astype(np.uint32)
    assert TH.thickness_from_hits(n, idx, t, prim) is None
    t[99] = 1.0
    prim[99] = 0
    # 100 valid hits, none looking back (the hit face's normal is the ray's origin face's): the raw mode.
    r = TH.thickness_from_hits(n, idx, t, prim)
    assert r is not None and r[0] == r[1]


@pytest.mark.parametrize("name", ["pieceA", "pieceB"])
def test_the_slab_dumps_thickness_from_its_own_rays(name):
    pc = S.piece(name)
    normals, _, _ = S.face_geometry(pc["V0"], pc["F"])
    idx = TH.face_indices(len(pc["F"]))
    assert len(idx) == len(pc["thick_t_hit"])
    t, raw = TH.thickness_from_hits(normals, idx, pc["thick_t_hit"], pc["thick_prim"])
    assert t.tobytes() == np.float32(pc["thick_t"]).tobytes()
    assert raw <= t * 1.05 and raw >= t * 0.9
