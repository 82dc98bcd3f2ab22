"""Writes `.sherd` fragment caches in the reference's layout (test helper).

crates/sherd-core/src/fragment/cache.rs:11-26 (the tensors), :104-176 (the metadata block, one
JSON object under `__metadata__.sherd`, in `CacheMeta`'s field order with shortest round-trip
floats) and :258-296 (`to_bytes`). Written with the safetensors package, an implementation of the
format independent of the library's reader."""

import json

import numpy as np
from safetensors.numpy import save

CACHE_VERSION = 7
ALGO_REF = "2026-09-06/9d4b9d3"


def brk_params(t):
    return dict(t=float(t), macro_inner=0.15, macro_outer=0.6, brk_voxel=0.5)


def md_params(t, seed=0, surface_points=20000, frac_per_t2=150.0, min_frac_points=5000, max_frac_points=12000,
              margin_points=6000):
    return dict(t=float(t), seed=int(seed), surface_points=int(surface_points), frac_per_t2=float(frac_per_t2),
                min_frac_points=int(min_frac_points), max_frac_points=int(max_frac_points),
                margin_points=int(margin_points))


def meta(name, thick, res, frac_area, watertight=True, n_boundary=0, area=0.0, thick_mode=0.0, brk_t=None, md_t=None,
         format="sherd-cache", cache_version=CACHE_VERSION, algo_ref=ALGO_REF, seed=0, **md):
    return dict(format=format, cache_version=cache_version, algo_ref=algo_ref, core_version="0.1.0", name=name,
                source_path=f"/data/{name}.ply", source_size=1, source_mtime_ns=0, target_faces=200000,
                face_budget=50000, area0=float(area), thick=float(thick), thick_mode=float(thick_mode),
                res=float(res), watertight=bool(watertight), n_boundary=int(n_boundary), area=float(area),
                frac_area=float(frac_area), n_orig_vertices=0, n_orig_faces=0, backend="cpu",
                brk_params=brk_params(thick if brk_t is None else brk_t),
                md_params=md_params(thick if md_t is None else md_t, seed, **md))


def to_bytes(arrays, meta_block):
    """arrays: V, F, labels, brk_P, brk_ns, brk_nf, brk_f, brk_sub, brk_valid, S, sp, Pf, fp,
    margin_idx (any dtype; narrowed to the file's)."""
    f32 = lambda a: np.ascontiguousarray(np.asarray(a, dtype=np.float32).reshape(-1, 3))
    u32 = lambda a: np.ascontiguousarray(np.asarray(a, dtype=np.uint32).reshape(-1))
    tensors = dict(
        V=f32(arrays["V"]), F=np.ascontiguousarray(np.asarray(arrays["F"], np.uint32).reshape(-1, 3)),
        labels=np.ascontiguousarray(np.asarray(arrays["labels"], np.uint8)),
        brk_P=f32(arrays["brk_P"]), brk_ns=f32(arrays["brk_ns"]), brk_nf=f32(arrays["brk_nf"]),
        brk_f=f32(arrays["brk_f"]), brk_sub=u32(arrays["brk_sub"]),
        brk_valid=np.ascontiguousarray(np.asarray(arrays["brk_valid"], bool)),
        S=f32(arrays["S"]), sp=u32(arrays["sp"]), Pf=f32(arrays["Pf"]), fp=u32(arrays["fp"]),
        margin_idx=u32(arrays["margin_idx"]))
    for k in ("labels", "brk_valid"):
        tensors[k] = tensors[k].reshape(-1)
    return save(tensors, metadata={"sherd": json.dumps(meta_block, separators=(",", ":"))})


def tetrahedron():
    """cache.rs:642-700's sample fragment: a tetrahedron, two breakline points, three surface
    samples, one fracture sample, a margin of two."""
    arrays = dict(
        V=[[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.12345678, 0.2, 0.9876543]],
        F=[[0, 2, 1], [0, 1, 3], [1, 2, 3], [2, 0, 3]], labels=[0, 1, 1, 0],
        brk_P=[[0.5, 0.0, 0.0], [0.0, 0.5, 0.12345679]], brk_ns=[[0.0, 0.0, -1.0], [0.0, -1.0, 0.0]],
        brk_nf=[[1.0, 0.0, 0.0], [0.0, 0.0, 1.0]], brk_f=[[1.0, 0.0, 0.0], [0.0, 0.0, 1.0]], brk_sub=[1],
        brk_valid=[True, True], S=[[0.25, 0.25, 0.0], [0.3, 0.1, 0.2], [0.1, 0.4, 0.1]], sp=[0, 1, 3],
        Pf=[[0.3, 0.1, 0.2]], fp=[1], margin_idx=[0, 2])
    m = meta("pieceA", 3.531017303466797, 0.12345679, 1.01538475629131, area=2.11538475629131, thick_mode=4.0441)
    return arrays, m
