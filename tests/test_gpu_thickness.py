"""R §3.2 on the device (crates/sherd-core/src/fragment/thickness.rs:100-161): the filter and the two
histogram modes against the oracle (oracle/thickness.py) bit for bit, on the slab dump's own rays
(whose thick.t it must reproduce) and on the device's own rays; and the native estimate within
D §10.2's 2 % of the dump's."""

import os
import sys

import numpy as np
import pytest

import slab_data as S

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import thickness as TH  # noqa: E402

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def ex():
    import sherd_hip

    e = sherd_hip.HipExecutor(0)
    yield e
    e.close()


@pytest.mark.parametrize("name", ["pieceA", "pieceB"])
def test_from_the_dumps_rays(ex, name):
    say(f"case: thickness from the dump's rays, {name}")
    pc = S.piece(name)
    normals, _, _ = S.face_geometry(pc["V0"], pc["F"])
    idx = TH.face_indices(len(pc["F"]))
    got = ex.thickness_from_hits(normals, idx, pc["thick_t_hit"], pc["thick_prim"])
    want = TH.thickness_from_hits(normals, idx, pc["thick_t_hit"], pc["thick_prim"])
    assert got[0].tobytes() == want[0].tobytes() == np.float32(pc["thick_t"]).tobytes()
    assert got[1].tobytes() == want[1].tobytes()


@pytest.mark.parametrize("name", ["pieceA", "pieceB"])
def test_native_on_the_slab(ex, name):
    say(f"case: native thickness, {name}")
    pc = S.piece(name)
    v0 = pc["V0"].astype(np.float32)
    scene = ex.scene(v0, pc["F"])
    got = ex.thickness(scene, pc["V0"], pc["F"])
    # The oracle over the same rays, cast through the device's own first hit.
    o, d = S.thickness_rays(name)
    t_hit, prim = ex.first_hit(scene, o, d)
    normals, _, _ = S.face_geometry(pc["V0"], pc["F"])
    want = TH.thickness_from_hits(normals, TH.face_indices(len(pc["F"])), t_hit, prim)
    assert got[0].tobytes() == want[0].tobytes() and got[1].tobytes() == want[1].tobytes()
    ref = np.float32(pc["thick_t"])
    say(f"  native t {got[0]} vs the dump's {ref}: {abs(got[0] - ref) / ref:.2%}")
    assert abs(got[0] - ref) <= 0.02 * ref
    scene.close()


def test_the_stride_on_a_large_mesh(ex):
    say("case: thickness over 360 000 faces (stride 2)")
    # A wall: two finely tessellated sheets 1 apart, the top facing +z and the bottom -z.
    n = 300
    u = np.linspace(0.0, 40.0, n + 1)
    X, Y = np.meshgrid(u, u, indexing="ij")
    top = np.stack([X.ravel(), Y.ravel(), np.full(X.size, 1.0)], axis=1)
    bot = np.stack([X.ravel(), Y.ravel(), np.zeros(X.size)], axis=1)
    V = np.concatenate([top, bot])
    m = n + 1
    i, j = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    a, b, c, d = (i * m + j).ravel(), ((i + 1) * m + j).ravel(), ((i + 1) * m + j + 1).ravel(), (i * m + j + 1).ravel()
    o = m * m
    F = np.concatenate([np.stack([a, b, c], 1), np.stack([a, c, d], 1),
                        np.stack([o + a, o + c, o + b], 1), np.stack([o + a, o + d, o + c], 1)]).astype(np.uint32)
    assert len(F) > 300_000
    scene = ex.scene(V.astype(np.float32), F)
    got = ex.thickness(scene, V, F)
    idx = TH.face_indices(len(F))
    normals, _, c = S.face_geometry(V, F)
    dirs = -normals[idx]
    o = (c[idx] + dirs * 1e-3).astype(np.float32)
    t_hit, prim = ex.first_hit(scene, o, dirs.astype(np.float32))
    want = TH.thickness_from_hits(normals, idx, t_hit, prim)
    assert got[0].tobytes() == want[0].tobytes() and got[1].tobytes() == want[1].tobytes()
    assert abs(got[0] - 1.0) < 0.05
    scene.close()


def test_too_few_hits(ex):
    say("case: thickness of a lone triangle pair (no ray hits)")
    V = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0]], np.float64)
    F = np.array([[0, 1, 2], [1, 3, 2]], np.uint32)
    scene = ex.scene(V.astype(np.float32), F)
    assert ex.thickness(scene, V, F) is None
    scene.close()


def test_a_face_index_outside_the_mesh_is_an_error(ex):
    say("case: thickness with a face index outside V (checked on the device)")
    import sherd_hip

    pc = S.piece("pieceA")
    scene = ex.scene(pc["V0"].astype(np.float32), pc["F"])
    good = ex.thickness(scene, pc["V0"], pc["F"])
    F = pc["F"].copy()
    F[len(F) // 2, 1] = len(pc["V0"])
    with pytest.raises(sherd_hip.HipError, match="out of range"):
        ex.thickness(scene, pc["V0"], F)
    # The context stays usable and the next estimate starts from a clean state.
    again = ex.thickness(scene, pc["V0"], pc["F"])
    assert again[0].tobytes() == good[0].tobytes() and again[1].tobytes() == good[1].tobytes()
    scene.close()
