"""Writes tests/golden/slab/cache_arrays.npz: the arrays of the reference's fragment cache
(crates/sherd-core/src/fragment/cache.rs:11-26) that the slab golden vectors do not already hold,
taken from the reference's fixture dump at the pair's wall thickness (the directory
tests/golden/make_slab.py reads the slab rows' arrays from).

Run in a container that has the reference checked out:

    python tests/golden/make_slab_cache.py /root/reference/fixtures/slab/dump

Per piece: `sp`, `fp` (the face of every surface and fracture sample), `margin_idx` (the margin's
indices into S), `brk_nf` (the fracture macro normals) and `brk_valid` (the frame-validity mask),
the dump's own md.sp / md.fp / md.margin_idx / md.brk_nf / md.valid; the sampling and breakline
parameters (md.params.json) and the mesh scalars (mesh.stats.json, mesh.watertight.json,
thick.target.json) the cache's metadata block carries. Only numpy is used, and only `numpy.load`
with its default `allow_pickle=False`.
"""

import json
import os
import sys

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from make_slab import PAIR, PIECES, arrays_dir  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "slab", "cache_arrays.npz")


def main(dump):
    with open(os.path.join(dump, "pairs", PAIR, "md_used.json")) as fh:
        t_pair = json.load(fh)["t"]
    out = {}
    for piece in PIECES:
        d = os.path.join(dump, "fragments", piece)
        md = arrays_dir(d, t_pair)
        ld = lambda name, base=md: np.load(os.path.join(base, name + ".npy"))
        js = lambda name, base=d: json.load(open(os.path.join(base, name + ".json")))
        params = js("md.params", md)
        stats = js("mesh.stats")
        target = js("thick.target")
        out[piece + "_sp"] = ld("md.sp").astype(np.uint32)
        out[piece + "_fp"] = ld("md.fp").astype(np.uint32)
        out[piece + "_margin_idx"] = ld("md.margin_idx").astype(np.uint32)
        out[piece + "_brk_nf"] = ld("md.brk_nf")
        out[piece + "_brk_valid"] = ld("md.valid").astype(bool)
        out[piece + "_md_params"] = np.array(json.dumps(params, sort_keys=True))
        out[piece + "_stats"] = np.array(json.dumps(stats, sort_keys=True))
        out[piece + "_target"] = np.array(json.dumps(target, sort_keys=True))
    np.savez_compressed(OUT, **out)
    print(OUT)


if __name__ == "__main__":
    main(sys.argv[1])
