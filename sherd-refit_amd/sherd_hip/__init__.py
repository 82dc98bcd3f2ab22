"""Python face of the MI355X backend: ctypes over the C ABI in include/sherd_hip.h.

`HipExecutor` mirrors the reference's `Executor` trait (crates/sherd-core/src/executor/mod.rs:105-137)
method for method: `name`, `coarse_scores`, `icp_rung`, `bounded_distance`, `inside`, plus the
scene and ray entry points the trait leaves to preprocessing (`first_hit`, `cone_votes`) and the
batched forms. Arrays are numpy arrays in the reference's layouts (n x 3 f64 points, 4 x 4 f64
poses, f32 meshes).

There is no CPU fallback: if libsherd_hip.so is missing or no gfx950 device is present, opening an
executor raises.
"""

import ctypes as C
import os
import weakref

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# SHERD_HIP_LIB names another build of the library (for instance a variant being compared).
LIB_PATH = os.environ.get("SHERD_HIP_LIB") or os.path.join(HERE, "lib", "libsherd_hip.so")

SHERD_M = dict(first_hit=0, bounded_distance=1, inside=2, coarse=3, icp=4, nearest=5, verify=6,
               scene_build=7, cone=8, huber=9, segment=10, verify_rays=11)

_lib = None


class HipError(RuntimeError):
    """A nonzero status from the C ABI, with sherd_last_error()'s message."""

    def __init__(self, status, msg):
        super().__init__(f"sherd_hip status {status}: {msg}")
        self.status = status


class InsideOutcome(C.Structure):
    _fields_ = [("depth", C.c_float), ("inside", C.c_uint32)]


class Registration(C.Structure):
    """`Registration` (crates/sherd-core/src/matching/icp.rs:182-196)."""

    _fields_ = [("transform", C.c_double * 16), ("fitness", C.c_double), ("inlier_rmse", C.c_double),
                ("correspondences", C.c_uint64), ("iterations", C.c_uint32), ("converged", C.c_uint32)]


class IcpPair(C.Structure):
    _fields_ = [("source", C.c_void_p), ("target", C.c_void_p), ("distance", C.c_double)]


class IcpRungSpec(C.Structure):
    _fields_ = [("estimation", C.c_int), ("max_iteration", C.c_int)]


class RefinePiece(C.Structure):
    """`RefinePiece` (crates/sherd-core/src/refine.rs:237-245)."""

    _fields_ = [("thick", C.c_double), ("res", C.c_double), ("cloud", C.c_void_p)]


class RefinedJoin(C.Structure):
    """`RefinedJoin` (crates/sherd-core/src/refine.rs:248-262)."""

    _fields_ = [("fixed", C.c_uint32), ("moving", C.c_uint32), ("dist", C.c_double * 2),
                ("rungs", C.c_double * 32), ("fitness", C.c_double), ("rmse_t", C.c_double)]


class MatchPair(C.Structure):
    _fields_ = [("a", C.c_void_p), ("b", C.c_void_p), ("ia", C.POINTER(C.c_uint32)), ("nia", C.c_size_t),
                ("ib", C.POINTER(C.c_uint32)), ("nib", C.c_size_t), ("probe_p", C.POINTER(C.c_double)),
                ("probe_n", C.POINTER(C.c_double)), ("n_probe", C.c_size_t), ("s1_source", C.c_void_p),
                ("s1_target", C.c_void_p), ("scales", C.c_double * 12)]


class MatchResult(C.Structure):
    _fields_ = [("n_hyp", C.c_uint32), ("n_kept1", C.c_uint32), ("n_kept2", C.c_uint32), ("pad", C.c_uint32),
                ("best1", C.c_double)]


class SegTrace(C.Structure):
    _fields_ = [("frac_majority", C.POINTER(C.c_uint8)), ("frac_islands", C.POINTER(C.c_uint8)),
                ("reference", C.POINTER(C.c_double)), ("has_ref", C.POINTER(C.c_uint8)), ("passes", C.c_int)]


class VerifyPair(C.Structure):
    _fields_ = [("a", C.c_void_p), ("b", C.c_void_p), ("scales", C.c_double * 12)]


SCORE_KEYS = [
    "tightA", "tightB", "tight", "gapA", "gapB", "gap", "contactA", "contactB", "contact", "seam",
    "gap_limit", "tight_delta", "cont", "cont_n", "pen", "pen_depth", "pen_unavailable", "partial",
    "brk", "brk_best",
]


class MethodStats(C.Structure):
    _fields_ = [("calls", C.c_uint64), ("items", C.c_uint64), ("device_ns", C.c_uint64)]


class Check(C.Structure):
    _fields_ = [("name", C.c_char * 16), ("passed", C.c_int), ("detail", C.c_char * 240)]


class SelfTestReport(C.Structure):
    """`sherd_selftest_report` (crates/sherd-gpu/src/selftest.rs:135-150 `SelfTest`)."""

    _fields_ = [("passed", C.c_int), ("n_checks", C.c_int), ("checks", Check * 4),
                ("reduce_device_bits", C.c_uint32), ("reduce_host_bits", C.c_uint32), ("queries", C.c_uint64),
                ("gpu_ms", C.c_double), ("cpu_ms", C.c_double), ("speedup", C.c_double), ("cpu_threads", C.c_int)]


class CacheInfo(C.Structure):
    _fields_ = [("name", C.c_char * 256), ("cache_version", C.c_uint32), ("target_faces", C.c_uint32),
                ("face_budget", C.c_uint32), ("n_boundary", C.c_uint32), ("watertight", C.c_int),
                ("thick", C.c_double), ("thick_mode", C.c_double), ("res", C.c_double), ("area", C.c_double),
                ("area0", C.c_double), ("frac_area", C.c_double), ("brk_t", C.c_double), ("md_t", C.c_double),
                ("nv", C.c_size_t), ("nf", C.c_size_t), ("n_brk", C.c_size_t), ("n_sub", C.c_size_t),
                ("n_s", C.c_size_t), ("n_pf", C.c_size_t), ("n_margin", C.c_size_t)]


_CACHE_ARRAYS = [("V", C.c_float, "nv", 3), ("F", C.c_uint32, "nf", 3), ("labels", C.c_uint8, "nf", 0),
                 ("brk_P", C.c_float, "n_brk", 3), ("brk_ns", C.c_float, "n_brk", 3),
                 ("brk_nf", C.c_float, "n_brk", 3), ("brk_f", C.c_float, "n_brk", 3),
                 ("brk_sub", C.c_uint32, "n_sub", 0), ("brk_valid", C.c_uint8, "n_brk", 0),
                 ("S", C.c_float, "n_s", 3), ("sp", C.c_uint32, "n_s", 0), ("Pf", C.c_float, "n_pf", 3),
                 ("fp", C.c_uint32, "n_pf", 0), ("margin_idx", C.c_uint32, "n_margin", 0)]


class CacheArraysOut(C.Structure):
    _fields_ = [(n, C.POINTER(ct)) for n, ct, _, _ in _CACHE_ARRAYS]


class MatchInfo(C.Structure):
    _fields_ = [("n_reg", C.c_size_t), ("n_reg_frac", C.c_size_t), ("n_reg_margin", C.c_size_t),
                ("n_pf", C.c_size_t), ("n_brk", C.c_size_t), ("n_sub", C.c_size_t), ("n_margin", C.c_size_t),
                ("n_s", C.c_size_t), ("nf", C.c_size_t), ("t", C.c_double), ("frac_area", C.c_double),
                ("watertight", C.c_int)]


_MATCH_ARRAYS = [("reg_p", C.c_double, "n_reg", 3), ("reg_n", C.c_double, "n_reg", 3),
                 ("pf", C.c_double, "n_pf", 3), ("pf_n", C.c_double, "n_pf", 3),
                 ("brk_p", C.c_double, "n_brk", 3), ("brk_ns", C.c_double, "n_brk", 3),
                 ("brk_f", C.c_double, "n_brk", 3), ("brk_t", C.c_double, "n_brk", 3),
                 ("brk_dih", C.c_double, "n_brk", 0), ("brk_sub", C.c_uint32, "n_sub", 0),
                 ("margin_p", C.c_double, "n_margin", 3), ("margin_n", C.c_double, "n_margin", 3),
                 ("s", C.c_double, "n_s", 3), ("s_n", C.c_double, "n_s", 3), ("s_frac", C.c_uint8, "n_s", 0),
                 ("face_normals", C.c_double, "nf", 3)]


class MatchArrays(C.Structure):
    _fields_ = [(n, C.POINTER(ct)) for n, ct, _, _ in _MATCH_ARRAYS]


_NP = {C.c_float: np.float32, C.c_double: np.float64, C.c_uint32: np.uint32, C.c_uint8: np.uint8}


def _arrays(spec, sizes, struct):
    out, ptrs = {}, {}
    for name, ct, size, cols in spec:
        n = int(getattr(sizes, size))
        a = np.zeros((n, cols) if cols else n, _NP[ct])
        out[name] = a
        ptrs[name] = _p(a, ct)
    return out, struct(**ptrs)


class FragmentCache:
    """A `.sherd` fragment cache read by the library (sherd_cache_read; cache.rs:329-438): `info`
    (the metadata and sizes), `arrays()` (the stored tensors), `match_data(reg_points)`
    (MatchData::own's arrays, samples.rs:499-560). Host memory; no device context needed."""

    def __init__(self, path=None, data=None):
        self.ptr = C.c_void_p()
        if data is not None:
            buf = np.frombuffer(bytes(data), np.uint8)
            _check(lib().sherd_cache_from_bytes(_p(buf, C.c_uint8), len(buf), C.byref(self.ptr)))
        else:
            _check(lib().sherd_cache_read(os.fsencode(path), C.byref(self.ptr)))
        i = CacheInfo()
        _check(lib().sherd_cache_info_get(self.ptr, C.byref(i)))
        self._info = i
        self.info = {k: (getattr(i, k).decode() if k == "name" else getattr(i, k)) for k, _ in CacheInfo._fields_}

    def arrays(self):
        out, st = _arrays(_CACHE_ARRAYS, self._info, CacheArraysOut)
        _check(lib().sherd_cache_arrays(self.ptr, C.byref(st)))
        out["brk_valid"] = out["brk_valid"].astype(bool)
        return out

    def match_data(self, reg_points=6000):
        mi = MatchInfo()
        _check(lib().sherd_match_data(self.ptr, int(reg_points), C.byref(mi), None))
        out, st = _arrays(_MATCH_ARRAYS, mi, MatchArrays)
        _check(lib().sherd_match_data(self.ptr, int(reg_points), C.byref(mi), C.byref(st)))
        out["s_frac"] = out["s_frac"].astype(bool)
        out.update({k: getattr(mi, k) for k, _ in MatchInfo._fields_})
        return out

    def close(self):
        if self.ptr:
            lib().sherd_cache_free(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


HEALTH_KEYS = ["calls", "failed", "arg_errors", "host_errors", "nomem", "limit", "corrupt", "injected"]


class Health(C.Structure):
    _fields_ = [(k, C.c_uint64) for k in HEALTH_KEYS]


# Status codes (include/sherd_hip.h) and the faults sherd_fault_inject arms.
E_ARG, E_HIP, E_NOMEM, E_LIMIT, E_NODEV, E_CORRUPT, E_FORMAT = -1, -2, -3, -4, -5, -6, -7
FAULT = dict(none=0, host=1, nomem=2, corrupt=3)


def lib():
    """The loaded library; raises if it was never built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise HipError(-5, f"{LIB_PATH} is missing: build it with sherd-refit_amd/build.py")
        _lib = _bind(C.CDLL(LIB_PATH))
    return _lib


def _p(a, ct):
    return None if a is None else a.ctypes.data_as(C.POINTER(ct))


def _bind(L):
    d, f, u32, u8, sz, i64 = C.c_double, C.c_float, C.c_uint32, C.c_uint8, C.c_size_t, C.c_int64
    P, vp, ci = C.POINTER, C.c_void_p, C.c_int
    sig = {
        "sherd_open": (ci, [ci, P(vp)]),
        "sherd_close": (None, [vp]),
        "sherd_name": (C.c_char_p, []),
        "sherd_last_error": (C.c_char_p, []),
        "sherd_device_name": (ci, [vp, C.c_char_p, sz]),
        "sherd_stats": (ci, [vp, ci, P(MethodStats)]),
        "sherd_selftest": (ci, [vp, P(SelfTestReport)]),
        "sherd_health_read": (ci, [vp, P(Health)]),
        "sherd_fault_inject": (ci, [vp, ci, ci]),
        "sherd_cache_read": (ci, [C.c_char_p, P(vp)]),
        "sherd_cache_from_bytes": (ci, [P(u8), sz, P(vp)]),
        "sherd_cache_free": (None, [vp]),
        "sherd_cache_info_get": (ci, [vp, P(CacheInfo)]),
        "sherd_cache_arrays": (ci, [vp, P(CacheArraysOut)]),
        "sherd_match_data": (ci, [vp, sz, P(MatchInfo), P(MatchArrays)]),
        "sherd_scene_build": (ci, [vp, P(f), sz, P(u32), sz, P(u8), P(vp)]),
        "sherd_scene_free": (None, [vp]),
        "sherd_scene_info": (ci, [vp, P(sz), P(f), P(f)]),
        "sherd_first_hit": (ci, [vp, vp, P(f), P(f), sz, ci, P(f), P(u32)]),
        "sherd_bounded_distance": (ci, [vp, vp, P(d), sz, P(d), d, ci, P(d)]),
        "sherd_bounded_distance_noalloc": (ci, [vp, vp, P(d), sz, P(d), d, ci, P(d)]),
        "sherd_vertex_votes": (ci, [vp, vp, P(d), sz, P(u32), sz, d, ci, P(u8), P(u8)]),
        "sherd_inside": (ci, [vp, vp, P(d), sz, P(d), ci, P(InsideOutcome)]),
        "sherd_nearest_below": (ci, [vp, P(d), sz, P(d), sz, d, P(i64), P(d)]),
        "sherd_coarse_scores": (ci, [vp, P(d), P(d), sz, P(d), P(d), sz, P(d), sz, d, d, P(u32), P(d)]),
        "sherd_icp_rung": (ci, [vp, P(d), sz, P(d), P(d), sz, P(d), sz, ci, d, ci, P(Registration)]),
        "sherd_fragment_create": (ci, [vp, vp, vp, ci, d, P(d), sz, P(d), sz, P(d), P(d), sz, P(d), P(d), sz,
                                       P(vp)]),
        "sherd_fragment_free": (None, [vp]),
        "sherd_cloud_create": (ci, [vp, P(d), P(d), sz, P(vp)]),
        "sherd_cloud_free": (None, [vp]),
        "sherd_icp_block": (ci, [vp, ci]),
        "sherd_frames_create": (ci, [vp, P(d), P(d), P(d), P(d), P(d), sz, P(u32), sz, P(vp)]),
        "sherd_frames_free": (None, [vp]),
        "sherd_hypotheses": (ci, [vp, vp, vp, P(u32), sz, P(u32), sz, d, sz, P(u32), P(u32), P(d), P(sz)]),
        "sherd_nms": (ci, [vp, P(u32), sz, P(d), P(d), sz, d, sz, d, P(u32), P(sz)]),
        "sherd_match_stage1": (ci, [vp, P(MatchPair), sz, sz, sz, d, P(MatchResult), P(u32), P(d), P(d), P(u32)]),
        "sherd_fracture_cloud": (ci, [vp, P(d), sz, P(u32), sz, P(d), P(u8), sz, d, d, sz, C.c_uint64, P(vp)]),
        "sherd_cloud_from_indices": (ci, [vp, P(d), sz, P(u32), sz, P(u32), sz, P(vp)]),
        "sherd_cloud_size": (sz, [vp]),
        "sherd_cloud_read": (ci, [vp, P(u32), P(d), P(d)]),
        "sherd_cloud_placed": (ci, [vp, P(d), ci, P(vp)]),
        "sherd_refine_joins": (ci, [vp, P(RefinePiece), sz, P(d), P(u32), P(sz), sz, P(u32), sz, P(d),
                                    P(RefinedJoin), P(sz)]),
        "sherd_huber_create": (ci, [vp, P(IcpPair), sz, P(u32), P(d), sz, P(vp)]),
        "sherd_huber_free": (None, [vp]),
        "sherd_huber_correspondences": (ci, [vp, P(C.c_int32), sz]),
        "sherd_huber_eval": (ci, [vp, P(d), d, ci, P(d), P(d)]),
        "sherd_huber_lm": (ci, [vp, d, ci, ci, P(d), P(d), P(u32)]),
        "sherd_icp_batch_rung": (ci, [vp, P(IcpPair), sz, P(u32), P(d), sz, ci, ci, P(Registration)]),
        "sherd_icp_batch_ladder": (ci, [vp, P(IcpPair), sz, P(IcpRungSpec), sz, P(u32), P(d), sz,
                                        P(Registration)]),
        "sherd_profile": (ci, [vp, ci]),
        "sherd_profile_read": (ci, [vp, C.c_char_p, sz]),
        "sherd_cone_votes": (ci, [vp, vp, P(d), sz, P(u32), sz, P(i64), P(d), sz, d, ci, P(u8)]),
        "sherd_thickness": (ci, [vp, vp, P(d), sz, P(u32), sz, P(C.c_float), P(C.c_float), P(ci)]),
        "sherd_thickness_from_hits": (ci, [vp, P(d), sz, P(u32), P(C.c_float), P(u32), sz, P(C.c_float),
                                           P(C.c_float), P(ci)]),
        "sherd_verify_batch": (ci, [vp, P(VerifyPair), sz, P(u32), P(d), sz, ci, ci, P(d), P(u8)]),
        "sherd_smoothed_normals": (ci, [vp, P(d), sz, P(u32), sz, P(u32), sz, P(u32), d, P(d)]),
        "sherd_segment_cleanup": (ci, [vp, P(d), sz, P(u32), sz, d, P(u8), P(u32), sz, P(u32), d, P(u8),
                                       P(SegTrace)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name, None)
        if fn is None:
            continue
        fn.restype = res
        fn.argtypes = args
    return L


def exported_symbols():
    return [n for n in dir(lib()) if n.startswith("sherd_")]


def _check(status):
    if status != 0:
        raise HipError(status, lib().sherd_last_error().decode())


def f64(a, cols=3):
    return np.ascontiguousarray(a, dtype=np.float64).reshape(-1, cols)


def f32(a, cols=3):
    return np.ascontiguousarray(a, dtype=np.float32).reshape(-1, cols)


def pose16(T):
    T = np.ascontiguousarray(T, dtype=np.float64).reshape(16)
    return T


class Scene:
    """A device LBVH over an f32 mesh (RayScene, crates/sherd-core/src/spatial/bvh.rs:36-82)."""

    def __init__(self, ex, V, F, mask=None):
        self.ex = ex
        V = f32(V)
        F = np.ascontiguousarray(F, dtype=np.uint32).reshape(-1, 3)
        m = None if mask is None else np.ascontiguousarray(mask, dtype=np.uint8)
        out = C.c_void_p()
        _check(lib().sherd_scene_build(ex.ptr, _p(V, C.c_float), len(V), _p(F, C.c_uint32), len(F),
                                       _p(m, C.c_uint8), C.byref(out)))
        self.ptr = out
        ex._own(self)

    def close(self):
        if self.ptr:
            lib().sherd_scene_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self):
        n = C.c_size_t()
        lo = np.zeros(3, np.float32)
        hi = np.zeros(3, np.float32)
        _check(lib().sherd_scene_info(self.ptr, C.byref(n), _p(lo, C.c_float), _p(hi, C.c_float)))
        return n.value, lo, hi


class Cloud:
    """A point set for ICP on the device (IcpTarget, crates/sherd-core/src/matching/icp.rs:218-300)."""

    def __init__(self, ex, pts, normals=None, _ptr=None):
        self.ex = ex
        if _ptr is not None:
            self.ptr = _ptr
            self.n = int(lib().sherd_cloud_size(_ptr))
            ex._own(self)
            return
        p = f64(pts)
        n = None if normals is None else f64(normals)
        out = C.c_void_p()
        _check(lib().sherd_cloud_create(ex.ptr, _p(p, C.c_double), _p(n, C.c_double), len(p), C.byref(out)))
        self.ptr = out
        ex._own(self)
        self.n = len(p)

    def read(self, idx=False, normals=True):
        """(points, normals or None, vertex indices or None) copied from the device."""
        pts = np.empty((self.n, 3))
        nrm = np.empty((self.n, 3)) if normals else None
        ix = np.empty(self.n, np.uint32) if idx else None
        _check(lib().sherd_cloud_read(self.ptr, _p(ix, C.c_uint32), _p(pts, C.c_double), _p(nrm, C.c_double)))
        return pts, nrm, ix

    @property
    def idx(self):
        return self.read(idx=True, normals=False)[2]

    def placed(self, T, tree=True):
        """FractureCloud::placed (refine.rs:99-104) as a new device cloud, with its point tree."""
        T = np.ascontiguousarray(T, dtype=np.float64).reshape(16)
        out = C.c_void_p()
        _check(lib().sherd_cloud_placed(self.ptr, _p(T, C.c_double), int(bool(tree)), C.byref(out)))
        return Cloud(self.ex, None, _ptr=out)

    def close(self):
        if self.ptr:
            lib().sherd_cloud_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Frames:
    """A fragment's breakline frames on the device (hypotheses.rs:47-68): points, shell normals,
    in-plane axes, tangents, dihedral angles and the hypothesis subset."""

    def __init__(self, ex, p, ns, f, t, dih, sub):
        self.ex = ex
        self._keep = [f64(p), f64(ns), f64(f), f64(t), np.ascontiguousarray(dih, dtype=np.float64).reshape(-1)]
        sub = np.ascontiguousarray(sub, dtype=np.uint32).reshape(-1)
        out = C.c_void_p()
        _check(lib().sherd_frames_create(ex.ptr, *[_p(a, C.c_double) for a in self._keep], len(self._keep[0]),
                                         _p(sub, C.c_uint32), len(sub), C.byref(out)))
        self.ptr = out
        self.n = len(self._keep[0])
        ex._own(self)

    def close(self):
        if self.ptr:
            lib().sherd_frames_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def face_centroids(V, F):
    """face_geometry's centroids (crates/sherd-core/src/mesh/geometry.rs:88-93): (a + b + c) / 3,
    left to right, per face."""
    V = f64(V).reshape(-1, 3)
    F = np.asarray(F).reshape(-1, 3)
    return (V[F[:, 0]] + V[F[:, 1]] + V[F[:, 2]]) / 3.0


# numpy views of the ABI's arrays of structures (the same layouts as the ctypes structures), so
# that batches of hundreds of pairs and thousands of registrations cross without Python loops.
_REG_DT = np.dtype([("transform", np.float64, 16), ("fitness", np.float64), ("inlier_rmse", np.float64),
                    ("correspondences", np.uint64), ("iterations", np.uint32), ("converged", np.uint32)])
_ICP_PAIR_DT = np.dtype([("source", np.uintp), ("target", np.uintp), ("distance", np.float64)])
_VERIFY_PAIR_DT = np.dtype([("a", np.uintp), ("b", np.uintp), ("scales", np.float64, 12)])
assert _REG_DT.itemsize == C.sizeof(Registration) and _ICP_PAIR_DT.itemsize == C.sizeof(IcpPair)
assert _VERIFY_PAIR_DT.itemsize == C.sizeof(VerifyPair)


def _addr(h):
    """The address behind a handle object (its c_void_p), 0 for None."""
    return 0 if h is None else (h.ptr.value or 0)


class IcpPairs:
    """A table of (source Cloud, target Cloud, distance) in the ABI's layout (sherd_icp_pair), built
    once and passed to the batch calls in place of a list, so that a batch over hundreds of pairs
    crosses the boundary without a Python loop per call. Holds its clouds (they outlive it)."""

    def __init__(self, pairs):
        pairs = list(pairs)
        self.clouds = [(s, t) for s, t, _ in pairs]
        self.arr = np.zeros(max(len(pairs), 1), _ICP_PAIR_DT)
        if pairs:
            self.arr["source"] = [_addr(s) for s, _, _ in pairs]
            self.arr["target"] = [_addr(t) for _, t, _ in pairs]
            self.arr["distance"] = [float(d) for _, _, d in pairs]
        # Source points that take part per pair (none when the target is empty).
        self.ns = np.array([s.n if t.n else 0 for s, t in self.clouds], np.int64)

    def __len__(self):
        return len(self.clouds)


def _icp_pairs(pairs):
    """[(source, target, distance)] or an IcpPairs -> an IcpPair array (numpy, ABI layout)."""
    return (pairs if isinstance(pairs, IcpPairs) else IcpPairs(pairs)).arr


class VerifyPairs:
    """A table of (Fragment A, Fragment B, Scales as 12 f64) in the ABI's layout (sherd_verify_pair),
    built once and passed to verify_batch in place of a list. Holds its fragments."""

    def __init__(self, pairs):
        pairs = list(pairs)
        self.frags = [(a, b) for a, b, _ in pairs]
        self.arr = np.zeros(max(len(pairs), 1), _VERIFY_PAIR_DT)
        if pairs:
            self.arr["a"] = [_addr(a) for a, _, _ in pairs]
            self.arr["b"] = [_addr(b) for _, b, _ in pairs]
            self.arr["scales"] = np.asarray([sc for _, _, sc in pairs], dtype=np.float64).reshape(len(pairs), 12)

    def __len__(self):
        return len(self.frags)


def _registrations(out, k):
    """A Registration array (ctypes) -> (k, 21): the transform, fitness, rmse, correspondences,
    iterations, converged."""
    a = np.frombuffer(out, dtype=_REG_DT, count=k)
    res = np.empty((k, 21))
    res[:, :16] = a["transform"]
    res[:, 16] = a["fitness"]
    res[:, 17] = a["inlier_rmse"]
    res[:, 18] = a["correspondences"]
    res[:, 19] = a["iterations"]
    res[:, 20] = a["converged"]
    return res


_EST = {"point_to_point": 0, "point_to_plane": 1}


class HuberBatch:
    """k poses of the robust point-to-plane seam energy with correspondences frozen at their base
    transforms (the C ABI's sherd_huber_*). pairs: [(source Cloud, target Cloud with normals,
    correspondence radius)]; cand_pair: the pair of each pose; T0: (k, 4, 4)."""

    def __init__(self, ex, pairs, cand_pair, T0):
        self.ex = ex
        pairs = pairs if isinstance(pairs, IcpPairs) else IcpPairs(pairs)
        self.clouds = pairs.clouds
        self.cand_pair = np.ascontiguousarray(cand_pair, dtype=np.uint32)
        T0 = np.ascontiguousarray(T0, dtype=np.float64).reshape(-1, 16)
        self.k = len(T0)
        # (A pose naming a pair out of range is the C ABI's to refuse; clipped here only for the count.)
        self.points = int(pairs.ns[np.minimum(self.cand_pair, len(pairs) - 1)].sum()) if len(pairs) else 0
        out = C.c_void_p()
        _check(lib().sherd_huber_create(ex.ptr, _p(pairs.arr, IcpPair), len(pairs), _p(self.cand_pair, C.c_uint32),
                                        _p(T0, C.c_double), self.k, C.byref(out)))
        self.ptr = out
        ex._own(self)

    def correspondences(self):
        out = np.empty(self.points, np.int32)
        _check(lib().sherd_huber_correspondences(self.ptr, _p(out, C.c_int32), self.points))
        return out

    def energy(self, xi, delta, precision=64, grad=False):
        """E (k,) and, with grad, dE/dxi (k, 6) at the twists xi (k, 6)."""
        xi = np.ascontiguousarray(xi, dtype=np.float64).reshape(self.k, 6)
        E = np.empty(self.k)
        G = np.empty((self.k, 6)) if grad else None
        _check(lib().sherd_huber_eval(self.ptr, _p(xi, C.c_double), float(delta), int(precision), _p(E, C.c_double),
                                      _p(G, C.c_double)))
        return (E, G) if grad else E

    def lm(self, delta, iterations, precision=64):
        """Batched LM from the base transforms: (T (k, 4, 4), E (k,), accepted steps (k,))."""
        T = np.empty((self.k, 16))
        E = np.empty(self.k)
        st = np.empty(self.k, np.uint32)
        _check(lib().sherd_huber_lm(self.ptr, float(delta), int(iterations), int(precision), _p(T, C.c_double),
                                    _p(E, C.c_double), _p(st, C.c_uint32)))
        return T.reshape(-1, 4, 4), E, st

    def close(self):
        if self.ptr:
            lib().sherd_huber_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Fragment:
    """One fragment as R §6 measures it (verify.rs:223-302 `Surfaces`), resident on the device."""

    def __init__(self, ex, fracture, mesh, watertight, frac_area, pf, s, brk_p, brk_ns, margin_p, margin_n):
        self.ex, self.fracture, self.mesh = ex, fracture, mesh
        arrs = [f64(x) for x in (pf, s, brk_p, brk_ns, margin_p, margin_n)]
        pf, s, bp, bn, mp, mn = arrs
        out = C.c_void_p()
        _check(lib().sherd_fragment_create(
            ex.ptr, fracture.ptr, mesh.ptr if mesh is not None else None, int(bool(watertight)), float(frac_area),
            _p(pf, C.c_double), len(pf), _p(s, C.c_double), len(s), _p(bp, C.c_double), _p(bn, C.c_double), len(bp),
            _p(mp, C.c_double), _p(mn, C.c_double), len(mp), C.byref(out)))
        self.ptr = out
        ex._own(self)

    def close(self):
        if self.ptr:
            lib().sherd_fragment_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class HipExecutor:
    """The `Executor` trait over one MI355X context."""

    def __init__(self, device=0):
        out = C.c_void_p()
        _check(lib().sherd_open(int(device), C.byref(out)))
        self.ptr = out
        self._handles = weakref.WeakSet()

    def _own(self, handle):
        """Handles (scenes, clouds, fragments, Huber batches) are freed before the context is."""
        self._handles.add(handle)

    def close(self):
        if self.ptr:
            for h in list(self._handles):
                h.close()
            lib().sherd_close(self.ptr)
            self.ptr = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @staticmethod
    def name():
        return lib().sherd_name().decode()

    def device_name(self):
        buf = C.create_string_buffer(256)
        _check(lib().sherd_device_name(self.ptr, buf, 256))
        return buf.value.decode()

    def stats(self, method):
        s = MethodStats()
        _check(lib().sherd_stats(self.ptr, SHERD_M[method], C.byref(s)))
        return dict(calls=s.calls, items=s.items, device_ns=s.device_ns)

    def selftest(self):
        """SelfTest::run (selftest.rs:179) through sherd_selftest: dict(passed, checks [(name, passed,
        detail)], reduce bits device/host, queries, gpu_ms, cpu_ms, speedup, cpu_threads)."""
        r = SelfTestReport()
        _check(lib().sherd_selftest(self.ptr, C.byref(r)))
        return dict(passed=bool(r.passed),
                    checks=[(r.checks[k].name.decode(), bool(r.checks[k].passed), r.checks[k].detail.decode())
                            for k in range(r.n_checks)],
                    reduce_device_bits=r.reduce_device_bits, reduce_host_bits=r.reduce_host_bits, queries=r.queries,
                    gpu_ms=r.gpu_ms, cpu_ms=r.cpu_ms, speedup=r.speedup, cpu_threads=r.cpu_threads)

    def health(self):
        """The context's call and failure counters (sherd_health_read)."""
        h = Health()
        _check(lib().sherd_health_read(self.ptr, C.byref(h)))
        return {k: int(getattr(h, k)) for k in HEALTH_KEYS}

    def fault_inject(self, kind, after=0):
        """Arms one fault ('host', 'nomem', 'corrupt'; 'none' disarms) at its after-th opportunity."""
        _check(lib().sherd_fault_inject(self.ptr, FAULT[kind], int(after)))

    def profile(self, on=True):
        _check(lib().sherd_profile(self.ptr, int(on)))

    def profile_read(self):
        import json

        buf = C.create_string_buffer(16384)
        _check(lib().sherd_profile_read(self.ptr, buf, 16384))
        return json.loads(buf.value.decode())

    def scene(self, V, F, mask=None):
        return Scene(self, V, F, mask)

    # ---- Executor methods -------------------------------------------------------------------

    def first_hit(self, scene, orig, dirs, retest=False):
        o, d = f32(orig), f32(dirs)
        n = len(o)
        t = np.empty(n, np.float32)
        prim = np.empty(n, np.uint32)
        _check(lib().sherd_first_hit(self.ptr, scene.ptr, _p(o, C.c_float), _p(d, C.c_float), n, int(retest),
                                     _p(t, C.c_float), _p(prim, C.c_uint32)))
        return t, prim

    def bounded_distance(self, scene, points, T, max_dist, reduce="all"):
        pts = f64(points)
        T = pose16(T)
        rd = {"all": 0, "min": 1}[reduce]
        out = np.empty(1 if rd else len(pts), np.float64)
        _check(lib().sherd_bounded_distance(self.ptr, scene.ptr, _p(pts, C.c_double), len(pts), _p(T, C.c_double),
                                            float(max_dist), rd, _p(out, C.c_double)))
        return out

    def bounded_distance_noalloc(self, scene, points, T, max_dist, reduce="all"):
        """bounded_distance through the entry point that allocates nothing (at most 262144 points)."""
        pts = f64(points)
        T = pose16(T)
        rd = {"all": 0, "min": 1}[reduce]
        out = np.empty(1 if rd else len(pts), np.float64)
        _check(lib().sherd_bounded_distance_noalloc(self.ptr, scene.ptr, _p(pts, C.c_double), len(pts),
                                                    _p(T, C.c_double), float(max_dist), rd, _p(out, C.c_double)))
        return out

    def inside(self, scene, points, T, retest=False):
        pts = f64(points)
        T = pose16(T)
        out = (InsideOutcome * max(len(pts), 1))()
        _check(lib().sherd_inside(self.ptr, scene.ptr, _p(pts, C.c_double), len(pts), _p(T, C.c_double),
                                  int(retest), out))
        arr = np.frombuffer(out, dtype=np.dtype([("depth", np.float32), ("inside", np.uint32)]))[: len(pts)]
        return arr["depth"].copy(), arr["inside"].copy()

    def nearest_below(self, cloud, queries, bound):
        """PointTree::nearest_below (kdtree.rs:182-192): (idx or -1, distance or inf)."""
        c, q = f64(cloud), f64(queries)
        idx = np.empty(len(q), np.int64)
        dist = np.empty(len(q), np.float64)
        _check(lib().sherd_nearest_below(self.ptr, _p(c, C.c_double), len(c), _p(q, C.c_double), len(q),
                                         float(bound), _p(idx, C.c_int64), _p(dist, C.c_double)))
        return idx, dist

    def coarse_scores(self, target_p, target_n, probe_p, probe_n, poses, radius, normal_agree=0.7):
        """Executor::coarse_scores; poses are (k, 12) rotation-then-translation rows or (k, 4, 4)."""
        poses = np.asarray(poses, dtype=np.float64)
        if poses.ndim == 3:
            poses = np.concatenate([poses[:, :3, :3].reshape(-1, 9), poses[:, :3, 3]], axis=1)
        poses = np.ascontiguousarray(poses)
        tp, tn, pp, pn = (f64(x) for x in (target_p, target_n, probe_p, probe_n))
        k = len(poses)
        counts = np.empty(k, np.uint32)
        scores = np.empty(k, np.float64)
        _check(lib().sherd_coarse_scores(self.ptr, _p(tp, C.c_double), _p(tn, C.c_double), len(tp),
                                         _p(pp, C.c_double), _p(pn, C.c_double), len(pp), _p(poses, C.c_double), k,
                                         float(radius), float(normal_agree), _p(counts, C.c_uint32),
                                         _p(scores, C.c_double)))
        return scores, counts

    def icp_rung(self, source, target_p, target_n, inits, estimation, distance, max_iteration):
        """Executor::icp_rung: one Registration per init, as a (k, 21) array
        [T(16), fitness, rmse, correspondences, iterations, converged] (the oracle's layout)."""
        src, tp = f64(source), f64(target_p)
        tn = None if target_n is None else f64(target_n)
        inits = np.ascontiguousarray(inits, dtype=np.float64).reshape(-1, 16)
        est = {"point_to_point": 0, "point_to_plane": 1}.get(estimation, estimation)
        k = len(inits)
        out = (Registration * max(k, 1))()
        _check(lib().sherd_icp_rung(self.ptr, _p(src, C.c_double), len(src), _p(tp, C.c_double),
                                    _p(tn, C.c_double), len(tp), _p(inits, C.c_double), k, int(est), float(distance),
                                    int(max_iteration), out))
        return _registrations(out, k)

    def cloud(self, pts, normals=None):
        return Cloud(self, pts, normals)

    def huber(self, pairs, cand_pair, T0):
        return HuberBatch(self, pairs, cand_pair, T0)

    def frames(self, p, ns, f, t, dih, sub):
        return Frames(self, p, ns, f, t, dih, sub)

    def hypotheses(self, a, b, ia, ib, tol=25.0):
        """hypotheses::build_with + poses on the device: (pa, pb, poses (k, 12))."""
        ia = np.ascontiguousarray(ia, dtype=np.uint32).reshape(-1)
        ib = np.ascontiguousarray(ib, dtype=np.uint32).reshape(-1)
        cap = max(len(ia) * len(ib), 1)
        pa, pb = np.empty(cap, np.uint32), np.empty(cap, np.uint32)
        poses = np.empty((cap, 12))
        n = C.c_size_t(0)
        _check(lib().sherd_hypotheses(self.ptr, a.ptr, b.ptr, _p(ia, C.c_uint32), len(ia), _p(ib, C.c_uint32), len(ib),
                                      float(tol), cap, _p(pa, C.c_uint32), _p(pb, C.c_uint32), _p(poses, C.c_double),
                                      C.byref(n)))
        k = n.value
        return pa[:k].copy(), pb[:k].copy(), poses[:k].copy()

    def nms(self, order, poses, score, tol, topk, floor):
        """nms::nms over a given walk: the kept items in the order kept."""
        order = np.ascontiguousarray(order, dtype=np.uint32).reshape(-1)
        poses = f64(poses, 12)
        score = np.ascontiguousarray(score, dtype=np.float64).reshape(-1)
        kept = np.empty(max(min(len(order), int(topk) + 1), 1), np.uint32)
        n = C.c_size_t(0)
        _check(lib().sherd_nms(self.ptr, _p(order, C.c_uint32), len(order), _p(poses, C.c_double),
                               _p(score, C.c_double), len(score), float(tol), int(topk), float(floor),
                               _p(kept, C.c_uint32), C.byref(n)))
        return kept[:n.value].copy()

    def match_stage1(self, pairs, stage1=250, stage2=10, tol=25.0):
        """R §5.1-5.5 over many pairs. pairs: dicts with a, b (Frames), optional ia / ib, probe_p,
        probe_n, s1_source, s1_target (Clouds) and scales (12). Returns per pair a dict: n_hyp,
        kept1, s1_T (k, 4, 4), s1_score, kept2, best1."""
        n = len(pairs)
        arr = (MatchPair * max(n, 1))()
        keep = []
        for i, m in enumerate(pairs):
            x = arr[i]
            x.a, x.b = m["a"].ptr, m["b"].ptr
            for side in ("ia", "ib"):
                if m.get(side) is not None:
                    v = np.ascontiguousarray(m[side], dtype=np.uint32).reshape(-1)
                    keep.append(v)
                    setattr(x, side, _p(v, C.c_uint32))
                    setattr(x, "n" + side, len(v))
            pp, pn = f64(m["probe_p"]), f64(m["probe_n"])
            keep += [pp, pn]
            x.probe_p, x.probe_n, x.n_probe = _p(pp, C.c_double), _p(pn, C.c_double), len(pp)
            x.s1_source, x.s1_target = m["s1_source"].ptr, m["s1_target"].ptr
            x.scales = (C.c_double * 12)(*[float(v) for v in m["scales"]])
        res = (MatchResult * max(n, 1))()
        k1 = np.zeros((n, stage1), np.uint32)
        T = np.zeros((n, stage1, 16))
        s1 = np.zeros((n, stage1))
        k2 = np.zeros((n, stage2), np.uint32)
        _check(lib().sherd_match_stage1(self.ptr, arr, n, int(stage1), int(stage2), float(tol), res,
                                        _p(k1, C.c_uint32), _p(T, C.c_double), _p(s1, C.c_double),
                                        _p(k2, C.c_uint32)))
        out = []
        for i in range(n):
            r = res[i]
            out.append(dict(n_hyp=int(r.n_hyp), kept1=k1[i, :r.n_kept1].copy(),
                            s1_T=T[i, :r.n_kept1].reshape(-1, 4, 4).copy(), s1_score=s1[i, :r.n_kept1].copy(),
                            kept2=k2[i, :r.n_kept2].copy(), best1=float(r.best1)))
        return out

    def fracture_cloud(self, V, F, centroids, fracture, thick, res, max_points=150000, seed=0):
        """refine.rs:216-227 fracture_cloud of an original mesh against the working mesh's face
        centroids (face_centroids) and fracture labels."""
        V, F = f64(V).reshape(-1, 3), np.ascontiguousarray(F, dtype=np.uint32).reshape(-1, 3)
        Cc = f64(centroids).reshape(-1, 3)
        fr = np.ascontiguousarray(fracture, dtype=np.uint8).reshape(-1)
        assert len(fr) == len(Cc)
        out = C.c_void_p()
        _check(lib().sherd_fracture_cloud(self.ptr, _p(V, C.c_double), len(V), _p(F, C.c_uint32), len(F),
                                          _p(Cc, C.c_double), _p(fr, C.c_uint8), len(Cc), float(thick), float(res),
                                          int(max_points), int(seed), C.byref(out)))
        return Cloud(self, None, _ptr=out)

    def cloud_from_indices(self, V, F, idx):
        """refine.rs:207-214: the vertices idx of a mesh, with its vertex normals."""
        V, F = f64(V).reshape(-1, 3), np.ascontiguousarray(F, dtype=np.uint32).reshape(-1, 3)
        ix = np.ascontiguousarray(idx, dtype=np.uint32).reshape(-1)
        out = C.c_void_p()
        _check(lib().sherd_cloud_from_indices(self.ptr, _p(V, C.c_double), len(V), _p(F, C.c_uint32), len(F),
                                              _p(ix, C.c_uint32), len(ix), C.byref(out)))
        return Cloud(self, None, _ptr=out)

    def refine_joins(self, pieces, poses, groups, used):
        """refine.rs:270-303 refine_joins. pieces: [(thick, res, Cloud or None)] per fragment;
        poses: (n, 4, 4); groups: lists of fragment ids; used: [(a, b)] in assembly order.
        Returns (poses (n, 4, 4), joins: [dict(fixed, moving, dist, rungs (2, 4, 4), fitness,
        rmse_t)] in walk order)."""
        n = len(pieces)
        pc = (RefinePiece * max(n, 1))()
        for i, (t, r, cl) in enumerate(pieces):
            pc[i] = RefinePiece(float(t), float(r), None if cl is None else cl.ptr)
        P = np.ascontiguousarray(poses, dtype=np.float64).reshape(-1, 16)
        assert len(P) == n
        members = np.ascontiguousarray([m for g in groups for m in g], dtype=np.uint32)
        sizes = np.ascontiguousarray([len(g) for g in groups], dtype=np.uintp)
        us = np.ascontiguousarray(used, dtype=np.uint32).reshape(-1, 2)
        out = np.empty((n, 16))
        joins = (RefinedJoin * max(len(us), 1))()
        nj = C.c_size_t(0)
        _check(lib().sherd_refine_joins(self.ptr, pc, n, _p(P, C.c_double), _p(members, C.c_uint32),
                                        _p(sizes, C.c_size_t), len(sizes), _p(us, C.c_uint32), len(us),
                                        _p(out, C.c_double), joins, C.byref(nj)))
        res = []
        for j in joins[:nj.value]:
            res.append(dict(fixed=int(j.fixed), moving=int(j.moving), dist=np.array(j.dist[:]),
                            rungs=np.array(j.rungs[:]).reshape(2, 4, 4), fitness=float(j.fitness),
                            rmse_t=float(j.rmse_t)))
        return out.reshape(n, 4, 4), res

    def icp_block(self, threads):
        """Fix the ICP workgroup size (64/128/256/512), the wide form (1), or leave the choice to
        the autotuner (0)."""
        _check(lib().sherd_icp_block(self.ptr, int(threads)))

    def icp_batch_ladder(self, pairs, rungs, cand_pair, inits):
        """ladder::climb_all over many pairs. pairs: per rung a list of (source, target, distance)
        (all rungs the same length); rungs: [(estimation, max_iteration)]. Returns an
        (n_cand, n_rungs, 21) array in icp_rung's layout."""
        nr, npairs = len(rungs), len(pairs[0])
        if any(len(pr) != npairs for pr in pairs):
            raise HipError(-1, "every rung needs the same number of pairs")
        arr = np.concatenate([_icp_pairs(pr)[:npairs] for pr in pairs]) if npairs else _icp_pairs([])
        specs = (IcpRungSpec * nr)(*[IcpRungSpec(int(_EST.get(e, e)), int(it)) for e, it in rungs])
        cp = np.ascontiguousarray(cand_pair, dtype=np.uint32)
        inits = np.ascontiguousarray(inits, dtype=np.float64).reshape(-1, 16)
        k = len(inits)
        out = (Registration * max(k * nr, 1))()
        _check(lib().sherd_icp_batch_ladder(self.ptr, _p(arr, IcpPair), npairs, specs, nr, _p(cp, C.c_uint32),
                                            _p(inits, C.c_double), k, out))
        return _registrations(out, k * nr).reshape(k, nr, 21)

    def icp_batch_rung(self, pairs, cand_pair, inits, estimation, max_iteration):
        """One rung over many pairs: pairs [(source, target, distance)] -> (n_cand, 21)."""
        arr = _icp_pairs(pairs)
        cp = np.ascontiguousarray(cand_pair, dtype=np.uint32)
        inits = np.ascontiguousarray(inits, dtype=np.float64).reshape(-1, 16)
        k = len(inits)
        out = (Registration * max(k, 1))()
        _check(lib().sherd_icp_batch_rung(self.ptr, _p(arr, IcpPair), len(pairs), _p(cp, C.c_uint32), _p(inits, C.c_double), k,
                                          int(_EST.get(estimation, estimation)), int(max_iteration), out))
        return _registrations(out, k)

    def fragment(self, fracture, mesh, watertight, frac_area, pf, s, brk_p, brk_ns, margin_p, margin_n):
        return Fragment(self, fracture, mesh, watertight, frac_area, pf, s, brk_p, brk_ns, margin_p, margin_n)

    def verify_batch(self, pairs, cand_pair, cand_T, full=True, retest=False):
        """verify::verify + accept over candidates of many pairs. pairs: [(frag_a, frag_b, scales12)].
        Returns (scores (n, 20), accepted (n,) bool)."""
        pv = (pairs if isinstance(pairs, VerifyPairs) else VerifyPairs(pairs)).arr
        cp = np.ascontiguousarray(cand_pair, dtype=np.uint32)
        T = np.ascontiguousarray(cand_T, dtype=np.float64).reshape(-1, 16)
        n = len(T)
        scores = np.empty((n, 20), np.float64)
        acc = np.empty(n, np.uint8)
        _check(lib().sherd_verify_batch(self.ptr, _p(pv, VerifyPair), len(pairs), _p(cp, C.c_uint32), _p(T, C.c_double), n,
                                        int(full), int(retest), _p(scores, C.c_double), _p(acc, C.c_uint8)))
        return scores, acc.astype(bool)

    def vertex_votes(self, scene, V64, F, thick, retest=False):
        """One inward ray per vertex (sherd_vertex_votes): (good per vertex, fracture label per face)."""
        V64 = f64(V64)
        F = np.ascontiguousarray(F, dtype=np.uint32).reshape(-1, 3)
        good = np.empty(len(V64), np.uint8)
        frac = np.empty(len(F), np.uint8)
        _check(lib().sherd_vertex_votes(self.ptr, scene.ptr, _p(V64, C.c_double), len(V64), _p(F, C.c_uint32), len(F),
                                        float(thick), int(retest), _p(good, C.c_uint8), _p(frac, C.c_uint8)))
        return good, frac

    def thickness(self, scene, V64, F):
        """thickness::estimate_thickness (thickness.rs:100-110): (t, thick_mode) as float32, or None
        when fewer than 100 inward rays hit. scene is the mesh's own (V narrowed to f32)."""
        V64 = f64(V64)
        F = np.ascontiguousarray(F, dtype=np.uint32).reshape(-1, 3)
        t, m, ok = C.c_float(), C.c_float(), C.c_int()
        _check(lib().sherd_thickness(self.ptr, scene.ptr, _p(V64, C.c_double), len(V64), _p(F, C.c_uint32), len(F),
                                     C.byref(t), C.byref(m), C.byref(ok)))
        return (np.float32(t.value), np.float32(m.value)) if ok.value else None

    def thickness_from_hits(self, normals, idx, t_hit, prim):
        """thickness::thickness_from_hits (thickness.rs:139-161) on given rays."""
        normals = f64(normals)
        idx = np.ascontiguousarray(idx, dtype=np.uint32)
        t_hit = np.ascontiguousarray(t_hit, dtype=np.float32)
        prim = np.ascontiguousarray(prim, dtype=np.uint32)
        t, m, ok = C.c_float(), C.c_float(), C.c_int()
        _check(lib().sherd_thickness_from_hits(self.ptr, _p(normals, C.c_double), len(normals), _p(idx, C.c_uint32),
                                               _p(t_hit, C.c_float), _p(prim, C.c_uint32), len(t_hit), C.byref(t),
                                               C.byref(m), C.byref(ok)))
        return (np.float32(t.value), np.float32(m.value)) if ok.value else None

    def smoothed_normals(self, V64, F, rep, near, radius):
        """R §3.4.2's smoothed normal per face (segment.rs:540-546): the ball normal of its voxel
        representative."""
        V64 = f64(V64)
        F = np.ascontiguousarray(F, dtype=np.uint32).reshape(-1, 3)
        rep = np.ascontiguousarray(rep, dtype=np.uint32)
        near = np.ascontiguousarray(near, dtype=np.uint32)
        if len(near) != len(F):
            raise HipError("near needs one entry per face")
        ns = np.empty((len(F), 3))
        _check(lib().sherd_smoothed_normals(self.ptr, _p(V64, C.c_double), len(V64), _p(F, C.c_uint32), len(F),
                                            _p(rep, C.c_uint32), len(rep), _p(near, C.c_uint32), float(radius),
                                            _p(ns, C.c_double)))
        return ns

    def segment_cleanup(self, V64, F, thick, frac_raw, rep, near, boundary_angle=25.0, trace=False):
        """segment::run's steps 3.4.4-3.4.8 (segment.rs:544-635) after the cone vote: the final
        fracture mask (bool per face); with trace, also a dict of the intermediate masks, the shell
        reference normals, has_ref and the number of growth passes."""
        V64 = f64(V64)
        F = np.ascontiguousarray(F, dtype=np.uint32).reshape(-1, 3)
        raw = np.ascontiguousarray(frac_raw, dtype=np.uint8)
        rep = np.ascontiguousarray(rep, dtype=np.uint32)
        near = np.ascontiguousarray(near, dtype=np.uint32)
        n = len(F)
        if len(raw) != n or len(near) != n:
            raise HipError("frac_raw and near need one entry per face")
        labels = np.empty(n, np.uint8)
        tr = None
        if trace:
            bufs = dict(majority=np.empty(n, np.uint8), islands=np.empty(n, np.uint8),
                        reference=np.empty((n, 3)), has_ref=np.empty(n, np.uint8))
            tr = SegTrace(_p(bufs["majority"], C.c_uint8), _p(bufs["islands"], C.c_uint8),
                          _p(bufs["reference"], C.c_double), _p(bufs["has_ref"], C.c_uint8), 0)
        _check(lib().sherd_segment_cleanup(self.ptr, _p(V64, C.c_double), len(V64), _p(F, C.c_uint32), n, float(thick),
                                           _p(raw, C.c_uint8), _p(rep, C.c_uint32), len(rep), _p(near, C.c_uint32),
                                           float(boundary_angle), _p(labels, C.c_uint8),
                                           C.byref(tr) if tr is not None else None))
        if not trace:
            return labels.astype(bool)
        out = dict(majority=bufs["majority"].astype(bool), islands=bufs["islands"].astype(bool),
                   reference=bufs["reference"], has_ref=bufs["has_ref"].astype(bool), passes=tr.passes)
        return labels.astype(bool), out

    def cone_votes(self, scene, V64, F, faces, ns, thick, retest=False):
        """segment::classify_faces' seven-ray vote (segment.rs:318-381) for the listed faces."""
        V64 = f64(V64)
        F = np.ascontiguousarray(F, dtype=np.uint32).reshape(-1, 3)
        fl = None if faces is None else np.ascontiguousarray(faces, dtype=np.int64)
        ns = f64(ns)
        n = len(F) if fl is None else len(fl)
        good = np.empty(n, np.uint8)
        _check(lib().sherd_cone_votes(self.ptr, scene.ptr, _p(V64, C.c_double), len(V64), _p(F, C.c_uint32), len(F),
                                      _p(fl, C.c_int64), _p(ns, C.c_double), n, float(thick), int(retest),
                                      _p(good, C.c_uint8)))
        return good

