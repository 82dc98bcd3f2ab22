"""The entry point that allocates nothing, sherd_bounded_distance_noalloc, against
sherd_bounded_distance on the same inputs (GPU), through the C ABI.

Two calls in a row on one context, the second with fewer points, in both reductions (All, Min):
the values must equal sherd_bounded_distance's bit for bit (the same kernels run over the fixed
device array instead of the context's grow-only buffers), and the smaller call must not see the
larger call's leftovers. The smaller call is also checked against the oracle. The size limit is
refused on the host before any kernel runs.
"""

import sys

import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu

MAX_POINTS = 1 << 18


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _pose(rng, deg, shift):
    a = rng.normal(size=3)
    a /= np.linalg.norm(a)
    th = np.radians(deg)
    k = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    T = np.eye(4)
    T[:3, :3] = np.eye(3) + np.sin(th) * k + (1 - np.cos(th)) * (k @ k)
    T[:3, 3] = rng.normal(size=3) * shift
    return T


def test_noalloc_equals_bounded_distance_twice_smaller_second():
    import sherd_hip
    from sherd_hip import synth

    col = synth.vessel(t=1.0, edges_per_t=1.5, n_frag=2, seed=7, height_t=12.0, rmin_t=6.0, rmax_t=7.0)
    fr = col["fragments"][0]
    rng = np.random.default_rng(5)
    V = fr["V"].astype(np.float64)
    lo, hi = V.min(0) - 1.0, V.max(0) + 1.0
    with sherd_hip.HipExecutor(0) as ex:
        scene = ex.scene(fr["V"], fr["F"])
        for case, n in (("first, large", 200_000), ("second, smaller", 7_000)):
            # Half the points on the surface samples (distances inside the window), half spread
            # over the box (many beyond it).
            k = n // 2
            near = fr["S"][rng.integers(0, len(fr["S"]), size=k)] + rng.normal(scale=0.2, size=(k, 3))
            far = lo + rng.uniform(size=(n - k, 3)) * (hi - lo)
            pts = np.concatenate([near, far])
            T = _pose(rng, 3.0, 0.1)
            for reduce in ("all", "min"):
                say(f"case: {case} call, n={n}, reduce={reduce}")
                want = ex.bounded_distance(scene, pts, T, 0.75, reduce=reduce)
                got = ex.bounded_distance_noalloc(scene, pts, T, 0.75, reduce=reduce)
                assert got.shape == want.shape
                assert np.array_equal(got.view(np.uint64), want.view(np.uint64)), (case, reduce)
                if reduce == "all":
                    assert np.isfinite(got).sum() > n // 4 and (~np.isfinite(got)).sum() > 0
            if n <= 10_000:
                say(f"case: {case} call against the oracle")
                oscene = O.Scene(fr["V"], fr["F"])
                o_all = oscene.bounded_distance(pts, T, 0.75)
                o_min = oscene.bounded_distance(pts, T, 0.75, reduce_min=True)
                got = ex.bounded_distance_noalloc(scene, pts, T, 0.75, reduce="all")
                assert np.array_equal(got.view(np.uint64), np.asarray(o_all, np.float64).view(np.uint64))
                got = ex.bounded_distance_noalloc(scene, pts, T, 0.75, reduce="min")
                assert got[0] == float(np.asarray(o_min).ravel()[0])
        say("case: empty input and the size limit")
        assert ex.bounded_distance_noalloc(scene, np.zeros((0, 3)), np.eye(4), 0.5, reduce="min")[0] == np.float32(0.5)
        with pytest.raises(sherd_hip.HipError) as err:
            ex.bounded_distance_noalloc(scene, np.zeros((MAX_POINTS + 1, 3)), np.eye(4), 0.5)
        assert err.value.status == -4
