"""ctypes view of the oracle (oracle/build/liboracle.so) for the tests and the CPU baseline.

The oracle is the checker: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
load it.
"""

import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "build", "liboracle.so")

OS_KEYS = [
    "tightA", "tightB", "tight", "gapA", "gapB", "gap", "contactA", "contactB", "contact",
    "seam", "gap_limit", "tight_delta", "cont", "cont_n", "pen", "pen_depth",
    "pen_unavailable", "partial", "brk", "brk_best",
]
SCALE_KEYS = ["t", "res", "coarse", "stage1", "tight", "facing", "gap", "seam", "near", "pen", "nms", "icp"]

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = _bind(C.CDLL(LIB_PATH))
    return _lib


def _p(a, ct):
    return a.ctypes.data_as(C.POINTER(ct)) if a is not None else None


class Side(C.Structure):
    _fields_ = [
        ("fracture", C.c_void_p), ("mesh", C.c_void_p), ("watertight", C.c_int),
        ("frac_area", C.c_double),
        ("pf", C.POINTER(C.c_double)), ("n_pf", C.c_size_t),
        ("s", C.POINTER(C.c_double)), ("n_s", C.c_size_t),
        ("brk_p", C.POINTER(C.c_double)), ("brk_ns", C.POINTER(C.c_double)), ("n_brk", C.c_size_t),
        ("margin_p", C.POINTER(C.c_double)), ("margin_n", C.POINTER(C.c_double)),
        ("n_margin", C.c_size_t),
    ]


def _bind(L):
    d, f, u32, i64, u8, sz = C.c_double, C.c_float, C.c_uint32, C.c_int64, C.c_uint8, C.c_size_t
    P = C.POINTER
    L.oracle_nearest_unbounded.argtypes = [P(d), sz, P(d), sz, P(i64), P(d)]
    L.oracle_place.argtypes = [P(d), sz, P(d), C.c_int, P(d)]
    L.oracle_scene_new.restype = C.c_void_p
    L.oracle_scene_new.argtypes = [P(f), sz, P(u32), sz, P(u8)]
    L.oracle_scene_free.argtypes = [C.c_void_p]
    L.oracle_scene_faces.restype = sz
    L.oracle_scene_faces.argtypes = [C.c_void_p]
    L.oracle_scene_aabb.argtypes = [C.c_void_p, P(f), P(f)]
    L.oracle_first_hit.argtypes = [C.c_void_p, P(f), P(f), sz, C.c_int, C.c_int, P(f), P(u32)]
    L.oracle_crossings.restype = u32
    L.oracle_crossings.argtypes = [C.c_void_p, P(f), P(f), C.c_int, C.c_int]
    L.oracle_distance.restype = f
    L.oracle_distance.argtypes = [C.c_void_p, P(f), C.c_int]
    L.oracle_exec_bounded_distance.argtypes = [C.c_void_p, P(d), sz, P(d), d, C.c_int, P(d)]
    L.oracle_exec_inside.argtypes = [C.c_void_p, P(d), sz, P(d), C.c_int, P(f), P(u32)]
    L.oracle_nearest_below.argtypes = [P(d), sz, P(d), sz, d, P(i64), P(d)]
    L.oracle_nearest_f32.argtypes = [P(f), sz, P(f), sz, f, C.c_int, P(i64), P(f)]
    L.oracle_coarse_scores.argtypes = [P(d), P(d), sz, P(d), P(d), sz, P(d), sz, d, d, P(u32), P(d)]
    L.oracle_icp_rung.argtypes = [P(d), sz, P(d), P(d), sz, P(d), sz, C.c_int, d, C.c_int, P(d)]
    L.oracle_icp_correspondences.argtypes = [P(d), sz, P(d), sz, d, P(i64)]
    L.oracle_normal_equations.argtypes = [P(d), sz, P(d), P(d), P(i64), P(d), P(d)]
    L.oracle_scales_for_pair.argtypes = [d, d, P(d)]
    L.oracle_reduce_bits.argtypes = [C.c_size_t, C.c_size_t]
    L.oracle_reduce_bits.restype = C.c_uint32
    L.oracle_verify.argtypes = [P(Side), P(Side), P(d), P(d), C.c_int, C.c_int, P(d)]
    L.oracle_accept.restype = C.c_int
    L.oracle_accept.argtypes = [P(d), P(d)]
    L.oracle_pose_inverse.argtypes = [P(d), P(d)]
    L.oracle_median.restype = d
    L.oracle_median.argtypes = [P(d), sz]
    L.oracle_cone_votes.argtypes = [C.c_void_p, P(d), sz, P(u32), sz, P(i64), P(d), sz, d, C.c_int, P(u8)]
    L.oracle_huber_lm.argtypes = [P(d), sz, P(d), P(d), sz, d, P(d), sz, d, C.c_int, C.c_int, P(d), P(d),
                                  P(u32), P(C.c_int32)]
    L.oracle_set_threads.restype = C.c_int
    L.oracle_set_threads.argtypes = [C.c_int]
    return L


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def set_threads(n):
    return lib().oracle_set_threads(int(n))


class Scene:
    """RayScene::of_mesh / of_subset over f32 vertices (bvh.rs:56-82)."""

    def __init__(self, V, F, mask=None):
        self.V = f32(V).reshape(-1, 3)
        self.F = np.ascontiguousarray(F, dtype=np.uint32).reshape(-1, 3)
        m = None if mask is None else np.ascontiguousarray(mask, dtype=np.uint8)
        self.mask = m
        self.ptr = lib().oracle_scene_new(_p(self.V, C.c_float), len(self.V), _p(self.F, C.c_uint32),
                                          len(self.F), _p(m, C.c_uint8))
        if not self.ptr:
            raise ValueError("a scene needs at least one triangle")

    def __del__(self):
        if getattr(self, "ptr", None):
            lib().oracle_scene_free(self.ptr)
            self.ptr = None

    @property
    def n_faces(self):
        return lib().oracle_scene_faces(self.ptr)

    def aabb(self):
        lo = np.zeros(3, np.float32)
        hi = np.zeros(3, np.float32)
        lib().oracle_scene_aabb(self.ptr, _p(lo, C.c_float), _p(hi, C.c_float))
        return lo, hi

    def first_hit(self, orig, dirs, retest=False, brute=False):
        o = f32(orig).reshape(-1, 3)
        dd = f32(dirs).reshape(-1, 3)
        t = np.empty(len(o), np.float32)
        prim = np.empty(len(o), np.uint32)
        lib().oracle_first_hit(self.ptr, _p(o, C.c_float), _p(dd, C.c_float), len(o), int(retest),
                               int(brute), _p(t, C.c_float), _p(prim, C.c_uint32))
        return t, prim

    def crossings(self, o, d, retest=False, brute=False):
        o = f32(o)
        d = f32(d)
        return lib().oracle_crossings(self.ptr, _p(o, C.c_float), _p(d, C.c_float), int(retest), int(brute))

    def distance(self, p, brute=False):
        p = f32(p)
        return lib().oracle_distance(self.ptr, _p(p, C.c_float), int(brute))

    def bounded_distance(self, pts, T, max_dist, reduce_min=False):
        pts = f64(pts).reshape(-1, 3)
        T = f64(T).reshape(16)
        out = np.empty(1 if reduce_min else len(pts), np.float64)
        lib().oracle_exec_bounded_distance(self.ptr, _p(pts, C.c_double), len(pts), _p(T, C.c_double),
                                           float(max_dist), int(reduce_min), _p(out, C.c_double))
        return out

    def inside(self, pts, T, retest=False):
        pts = f64(pts).reshape(-1, 3)
        T = f64(T).reshape(16)
        depth = np.empty(len(pts), np.float32)
        ins = np.empty(len(pts), np.uint32)
        lib().oracle_exec_inside(self.ptr, _p(pts, C.c_double), len(pts), _p(T, C.c_double), int(retest),
                                 _p(depth, C.c_float), _p(ins, C.c_uint32))
        return depth, ins

    def cone_votes(self, V64, F, faces, ns, thick, retest=False):
        V64 = f64(V64).reshape(-1, 3)
        F = np.ascontiguousarray(F, dtype=np.uint32).reshape(-1, 3)
        faces = np.ascontiguousarray(faces, dtype=np.int64)
        ns = f64(ns).reshape(-1, 3)
        good = np.empty(len(faces), np.uint8)
        lib().oracle_cone_votes(self.ptr, _p(V64, C.c_double), len(V64), _p(F, C.c_uint32), len(F),
                                _p(faces, C.c_int64), _p(ns, C.c_double), len(faces), float(thick),
                                int(retest), _p(good, C.c_uint8))
        return good


def nearest_below(cloud, q, bound):
    cloud = f64(cloud).reshape(-1, 3)
    q = f64(q).reshape(-1, 3)
    idx = np.empty(len(q), np.int64)
    dist = np.empty(len(q), np.float64)
    lib().oracle_nearest_below(_p(cloud, C.c_double), len(cloud), _p(q, C.c_double), len(q), float(bound),
                               _p(idx, C.c_int64), _p(dist, C.c_double))
    return idx, dist


def nearest_unbounded(cloud, q):
    """kdtree.rs:120-123 nearest_distance: (index, distance) of each query's nearest point."""
    cloud = f64(cloud).reshape(-1, 3)
    q = f64(q).reshape(-1, 3)
    idx = np.empty(len(q), np.int64)
    dist = np.empty(len(q), np.float64)
    lib().oracle_nearest_unbounded(_p(cloud, C.c_double), len(cloud), _p(q, C.c_double), len(q),
                                   _p(idx, C.c_int64), _p(dist, C.c_double))
    return idx, dist


def place(pts, T, direction=False):
    """types.rs:202-214: points (or directions) through a 4x4 with fused rows."""
    pts = f64(pts).reshape(-1, 3)
    T = f64(T).reshape(16)
    out = np.empty_like(pts)
    lib().oracle_place(_p(pts, C.c_double), len(pts), _p(T, C.c_double), int(bool(direction)), _p(out, C.c_double))
    return out


def nearest_f32(cloud, q, radius, strict=True):
    cloud = f32(cloud).reshape(-1, 3)
    q = f32(q).reshape(-1, 3)
    idx = np.empty(len(q), np.int64)
    dist = np.empty(len(q), np.float32)
    lib().oracle_nearest_f32(_p(cloud, C.c_float), len(cloud), _p(q, C.c_float), len(q), np.float32(radius),
                             int(strict), _p(idx, C.c_int64), _p(dist, C.c_float))
    return idx, dist


def coarse_scores(target_p, target_n, probe_p, probe_n, poses, radius, normal_agree=0.7):
    """poses: (k, 12) = R row-major then tau, or (k, 4, 4)."""
    poses = np.asarray(poses, dtype=np.float64)
    if poses.ndim == 3:
        poses = np.concatenate([poses[:, :3, :3].reshape(-1, 9), poses[:, :3, 3]], axis=1)
    poses = f64(poses)
    tp, tn, pp, pn = (f64(x).reshape(-1, 3) for x in (target_p, target_n, probe_p, probe_n))
    k = len(poses)
    counts = np.empty(k, np.uint32)
    scores = np.empty(k, np.float64)
    lib().oracle_coarse_scores(_p(tp, C.c_double), _p(tn, C.c_double), len(tp), _p(pp, C.c_double),
                               _p(pn, C.c_double), len(pp), _p(poses, C.c_double), k, float(radius),
                               float(normal_agree), _p(counts, C.c_uint32), _p(scores, C.c_double))
    return scores, counts


def icp_rung(src, tgt_p, tgt_n, inits, estimation, distance, max_iteration):
    src = f64(src).reshape(-1, 3)
    tp = f64(tgt_p).reshape(-1, 3)
    tn = None if tgt_n is None else f64(tgt_n).reshape(-1, 3)
    inits = f64(inits).reshape(-1, 16)
    out = np.empty((len(inits), 21), np.float64)
    est = {"point_to_point": 0, "point_to_plane": 1}.get(estimation, estimation)
    lib().oracle_icp_rung(_p(src, C.c_double), len(src), _p(tp, C.c_double), _p(tn, C.c_double), len(tp),
                          _p(inits, C.c_double), len(inits), int(est), float(distance), int(max_iteration),
                          _p(out, C.c_double))
    return out


def icp_correspondences(moved, tgt_p, distance):
    """icp.rs:361-392: per moved source point the nearest target strictly inside distance, or -1."""
    m, tp = f64(moved).reshape(-1, 3), f64(tgt_p).reshape(-1, 3)
    out = np.empty(len(m), np.int64)
    lib().oracle_icp_correspondences(_p(m, C.c_double), len(m), _p(tp, C.c_double), len(tp), float(distance),
                                     _p(out, C.c_int64))
    return out


def normal_equations(moved, tgt_p, tgt_n, corr):
    """icp.rs:599-640: (JTJ 6x6, JTr 6) over the pairs of corr, source order."""
    m, tp, tn = f64(moved).reshape(-1, 3), f64(tgt_p).reshape(-1, 3), f64(tgt_n).reshape(-1, 3)
    c = np.ascontiguousarray(corr, dtype=np.int64)
    jtj, jtr = np.empty(36), np.empty(6)
    lib().oracle_normal_equations(_p(m, C.c_double), len(m), _p(tp, C.c_double), _p(tn, C.c_double),
                                  _p(c, C.c_int64), _p(jtj, C.c_double), _p(jtr, C.c_double))
    return jtj.reshape(6, 6), jtr


LM_RAN_OUT, LM_CONVERGED, LM_DAMPED_OUT, LM_NO_STEP = 0, 1, 2, 3


def huber_lm(src, tgt_p, tgt_n, radius, T0, delta, iterations, threads=0):
    """The batched Huber LM's FP64 restatement (oracle/huber_lm.cpp) for poses T0 (k x 4 x 4) of one
    pair: (T k x 4 x 4, energy k, steps k, status k)."""
    s, tp, tn = (f64(x).reshape(-1, 3) for x in (src, tgt_p, tgt_n))
    T0 = f64(T0).reshape(-1, 16)
    k = len(T0)
    T = np.empty((k, 16), np.float64)
    E = np.empty(k, np.float64)
    st = np.empty(k, np.uint32)
    status = np.empty(k, np.int32)
    lib().oracle_huber_lm(_p(s, C.c_double), len(s), _p(tp, C.c_double), _p(tn, C.c_double), len(tp), float(radius),
                          _p(T0, C.c_double), k, float(delta), int(iterations), int(threads), _p(T, C.c_double),
                          _p(E, C.c_double), _p(st, C.c_uint32), _p(status, C.c_int32))
    return T.reshape(k, 4, 4), E, st, status


def reduce_bits(n, groups):
    """The self-test reduction's f32 sum bits (selftest.rs:405-460)."""
    return int(lib().oracle_reduce_bits(int(n), int(groups)))


def scales_for_pair(t, res):
    out = np.empty(12, np.float64)
    lib().oracle_scales_for_pair(float(t), float(res), _p(out, C.c_double))
    return out


def pose_inverse(T):
    T = f64(T).reshape(16)
    out = np.empty(16, np.float64)
    lib().oracle_pose_inverse(_p(T, C.c_double), _p(out, C.c_double))
    return out.reshape(4, 4)


def median(values):
    v = f64(values).copy()
    return lib().oracle_median(_p(v, C.c_double), len(v))


class SideData:
    """The arrays R §6 reads of one fragment (verify.rs:223-302), kept alive with its Side."""

    def __init__(self, fracture, mesh, watertight, frac_area, pf, s, brk_p, brk_ns, margin_p, margin_n):
        self.fracture, self.mesh = fracture, mesh
        self.arrays = [f64(x).reshape(-1, 3) for x in (pf, s, brk_p, brk_ns, margin_p, margin_n)]
        pf, s, bp, bn, mp, mn = self.arrays
        self.side = Side(
            fracture.ptr if fracture is not None else None,
            mesh.ptr if mesh is not None else None,
            int(bool(watertight)), float(frac_area),
            _p(pf, C.c_double), len(pf), _p(s, C.c_double), len(s),
            _p(bp, C.c_double), _p(bn, C.c_double), len(bp),
            _p(mp, C.c_double), _p(mn, C.c_double), len(mp),
        )


def verify(a, b, T, scales, full=True, retest=False):
    T = f64(T).reshape(16)
    sc = f64(scales)
    out = np.empty(20, np.float64)
    lib().oracle_verify(C.byref(a.side), C.byref(b.side), _p(T, C.c_double), _p(sc, C.c_double), int(full),
                        int(retest), _p(out, C.c_double))
    return out


def accept(scores, scales):
    s = f64(scores)
    sc = f64(scales)
    return bool(lib().oracle_accept(_p(s, C.c_double), _p(sc, C.c_double)))
