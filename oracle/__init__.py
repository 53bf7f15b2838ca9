"""Oracle bindings (test infrastructure only).

Loads the CPU restatement of the reference GJK/EPA (``oracle/lib``) and, when it
has been built in a container that holds the reference checkout, the reference
CPU sources themselves (``oracle/_ref/libref32.so``, FP32 as shipped).

Only ``tests/``, ``__graft_entry__.smoke()`` and the benchmark's CPU-baseline leg
use this module, always as the checker.  The shipped library never calls it.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB64 = os.path.join(_HERE, "lib", "liboracle64.so")
LIB32 = os.path.join(_HERE, "lib", "liboracle32.so")
REF32 = os.path.join(_HERE, "_ref", "libref32.so")

GJK, EPA, GJK_EPA = 1, 2, 3


def simplex_dtype(bits: int) -> np.dtype:
    """numpy view of gkSimplex (GJK/common.h:84-89) at the given precision."""
    real = np.float64 if bits == 64 else np.float32
    offs = [0, 8, 104, 136] if bits == 64 else [0, 4, 52, 84]
    return np.dtype({
        "names": ["nvrtx", "vrtx", "vrtx_idx", "witnesses"],
        "formats": [np.int32, (real, (4, 3)), (np.int32, (4, 2)), (real, (2, 3))],
        "offsets": offs,
        "itemsize": 184 if bits == 64 else 108,
    })


STATS_DTYPE = np.dtype([("gjk_iters", np.int32), ("epa_iters", np.int32),
                        ("sweeps_a", np.int32), ("sweeps_b", np.int32),
                        ("face_overflow", np.int32)])

_libs: dict = {}


def _lib(bits: int):
    if bits not in _libs:
        path = LIB64 if bits == 64 else LIB32
        if not os.path.exists(path):
            raise RuntimeError(f"oracle library missing: {path} (run __graft_entry__.build())")
        lib = C.CDLL(path)
        lib.oracle_batch.restype = None
        lib.oracle_batch.argtypes = [C.c_int, C.c_int] + [C.c_void_p] * 9 + [C.c_int]
        lib.oracle_batch_f32.restype = None
        lib.oracle_batch_f32.argtypes = [C.c_int, C.c_int] + [C.c_void_p] * 9 + [C.c_int]
        lib.oracle_real_size.restype = C.c_int
        assert lib.oracle_real_size() == bits // 8
        _libs[bits] = lib
    return _libs[bits]


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


@dataclass
class Result:
    distances: np.ndarray
    simplices: np.ndarray
    normals: np.ndarray | None
    stats: np.ndarray | None


def run(mode: int, coords, off, cnt, pa, pb, bits: int = 64, nthreads: int = 1,
        simplices=None, distances=None, stats: bool = False) -> Result:
    """Run the oracle over pairs (pa[i], pb[i]) of bodies in an AoS pool.

    ``coords`` is an (N, 3) pool.  A float32 pool is promoted exactly to the
    oracle's precision pair by pair inside the workers (a large pool is never
    converted as a whole); a pool already in the oracle's precision is read
    in place.  For mode EPA, ``simplices``/``distances`` are the inputs
    (copied).
    """
    real = np.float64 if bits == 64 else np.float32
    n = len(pa)
    per_pair = np.asarray(coords).dtype == np.float32 and real != np.float32
    c = np.ascontiguousarray(coords, dtype=np.float32 if per_pair else real).reshape(-1, 3)
    off = np.ascontiguousarray(off, dtype=np.int64)
    cnt = np.ascontiguousarray(cnt, dtype=np.int32)
    pa = np.ascontiguousarray(pa, dtype=np.int32)
    pb = np.ascontiguousarray(pb, dtype=np.int32)
    sd = simplex_dtype(bits)
    if simplices is None:
        simp = np.zeros(n, dtype=sd)
    else:
        simp = np.array(simplices, dtype=sd, copy=True)
    dist = np.zeros(n, dtype=real) if distances is None else np.array(distances, dtype=real, copy=True)
    normals = np.zeros((n, 3), dtype=real) if mode & EPA else None
    st = np.zeros(n, dtype=STATS_DTYPE) if stats else None
    if n:
        fn = _lib(bits).oracle_batch_f32 if per_pair else _lib(bits).oracle_batch
        fn(mode, n, _p(c), _p(off), _p(cnt), _p(pa), _p(pb), _p(simp), _p(dist), _p(normals),
           _p(st), int(nthreads))
    return Result(dist, simp, normals, st)


def compact(coords, off, cnt, pa, pb, dtype=np.float64):
    """The bodies the pairs (pa[i], pb[i]) name, gathered into a pool of their
    own in `dtype` (exact promotion): (coords, off, cnt, pa, pb) with the
    pairs renumbered.  Lets a timed oracle run read a pool of only the bodies
    it uses, already in its precision."""
    pa = np.asarray(pa, np.int64)
    pb = np.asarray(pb, np.int64)
    bodies, inv = np.unique(np.concatenate([pa, pb]), return_inverse=True)
    cnts = np.asarray(cnt, np.int64)[bodies]
    offs = np.zeros(len(bodies), np.int64)
    if len(bodies):
        offs[1:] = np.cumsum(cnts)[:-1]
    total = int(cnts.sum())
    src = np.repeat(np.asarray(off, np.int64)[bodies] - offs, cnts) + np.arange(total)
    pool = np.asarray(coords).reshape(-1, 3)[src].astype(dtype)
    n = len(pa)
    return (pool, offs, cnts.astype(np.int32), inv[:n].astype(np.int32),
            inv[n:].astype(np.int32))


# ----------------------------------------------------------------- reference

class _Poly32(C.Structure):
    _fields_ = [("numpoints", C.c_int), ("s", C.c_float * 3), ("s_idx", C.c_int),
                ("coord", C.POINTER(C.c_float))]


class _Simplex32(C.Structure):
    _fields_ = [("nvrtx", C.c_int), ("vrtx", (C.c_float * 3) * 4), ("vrtx_idx", (C.c_int * 2) * 4),
                ("witnesses", (C.c_float * 3) * 2)]


assert C.sizeof(_Poly32) == 32 and C.sizeof(_Simplex32) == 108

_ref = None


def ref_available() -> bool:
    return os.path.exists(REF32)


def _ref_lib():
    global _ref
    if _ref is None:
        lib = C.CDLL(REF32)
        lib.compute_minimum_distance.restype = C.c_float
        lib.compute_minimum_distance.argtypes = [_Poly32, _Poly32, C.POINTER(_Simplex32)]
        lib.computeCollisionInformation.restype = None
        lib.computeCollisionInformation.argtypes = [C.POINTER(_Poly32), C.POINTER(_Poly32),
                                                    C.POINTER(_Simplex32), C.POINTER(C.c_float),
                                                    C.POINTER(C.c_float)]
        _ref = lib
    return _ref


def ref_run(mode: int, coords, off, cnt, pa, pb) -> Result:
    """Reference CPU sources (FP32 as shipped), one call per pair.

    Mirrors GJK::CPU::computeDistances / computeEPA (examples/cpu/example.cpp:13-40).
    """
    lib = _ref_lib()
    c = np.ascontiguousarray(coords, dtype=np.float32).reshape(-1, 3)
    base = c.ctypes.data
    n = len(pa)
    sd = simplex_dtype(32)
    out = np.zeros(n, dtype=sd)
    dist = np.zeros(n, dtype=np.float32)
    normals = np.zeros((n, 3), dtype=np.float32) if mode & EPA else None
    fp = C.POINTER(C.c_float)
    for i in range(n):
        a, b = int(pa[i]), int(pb[i])
        p1 = _Poly32(int(cnt[a]), (C.c_float * 3)(), 0, C.cast(base + 12 * int(off[a]), fp))
        p2 = _Poly32(int(cnt[b]), (C.c_float * 3)(), 0, C.cast(base + 12 * int(off[b]), fp))
        s = _Simplex32()
        d = C.c_float(0.0)
        if mode & GJK:
            d.value = lib.compute_minimum_distance(p1, p2, C.byref(s))
        if mode & EPA:
            nrm = (C.c_float * 3)()
            lib.computeCollisionInformation(C.byref(p1), C.byref(p2), C.byref(s), C.byref(d), nrm)
            normals[i] = np.frombuffer(nrm, dtype=np.float32)
        dist[i] = d.value
        out[i] = np.frombuffer(bytes(s), dtype=sd)[0]
    return Result(dist, out, normals, None)


# ------------------------------------------------------------- posed bodies

def pose_matrix(q):
    """FP64 rotation of a float32 quaternion (w, x, y, z), entry by entry in
    the arithmetic the posed GPU path uses."""
    w, x, y, z = (np.float64(v) for v in q)
    return np.array([[1.0 - 2.0 * (y * y + z * z), 2.0 * (x * y - w * z), 2.0 * (x * z + w * y)],
                     [2.0 * (x * y + w * z), 1.0 - 2.0 * (x * x + z * z), 2.0 * (y * z - w * x)],
                     [2.0 * (x * z - w * y), 2.0 * (y * z + w * x), 1.0 - 2.0 * (x * x + y * y)]])


def world_vertices(local, pose):
    """FP64 world vertices R x + t of an (m, 3) float32 body under a pose of
    7 float32 (quaternion, translation): ((r0 x + r1 y) + r2 z) + t."""
    R = pose_matrix(pose[:4])
    t = pose[4:7].astype(np.float64)
    p = np.asarray(local, dtype=np.float64)
    px, py, pz = p[:, 0], p[:, 1], p[:, 2]
    return np.stack([((R[k, 0] * px + R[k, 1] * py) + R[k, 2] * pz) + t[k] for k in range(3)], 1)


def posed_pool(coords, off, cnt, pairs, poses):
    """World-vertex pool for posed pairs: body 2i / 2i+1 = pair i's bodies
    under its two poses, computed in the arithmetic of world_vertices.
    Returns (coords f64, off, cnt, pa, pb)."""
    n = len(pairs)
    bodies = np.asarray(pairs, dtype=np.int64).reshape(-1)        # 2n sides
    side_pose = np.asarray(poses, dtype=np.float32).reshape(2 * n, 7)
    cnts = np.asarray(cnt, dtype=np.int64)[bodies]
    offs = np.zeros(2 * n, np.int64)
    if n:
        offs[1:] = np.cumsum(cnts)[:-1]
    total = int(cnts.sum())
    side = np.repeat(np.arange(2 * n), cnts)                       # side of each world vertex
    src = np.repeat(np.asarray(off, dtype=np.int64)[bodies] - offs, cnts) + np.arange(total)
    p = np.asarray(coords, dtype=np.float64)[src]
    q = side_pose.astype(np.float64)
    w, x, y, z = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    R = [1.0 - 2.0 * (y * y + z * z), 2.0 * (x * y - w * z), 2.0 * (x * z + w * y),
         2.0 * (x * y + w * z), 1.0 - 2.0 * (x * x + z * z), 2.0 * (y * z - w * x),
         2.0 * (x * z - w * y), 2.0 * (y * z + w * x), 1.0 - 2.0 * (x * x + y * y)]
    R = [r[side] for r in R]
    t = [q[side, 4], q[side, 5], q[side, 6]]
    px, py, pz = p[:, 0], p[:, 1], p[:, 2]
    pool = np.stack([((R[3 * k] * px + R[3 * k + 1] * py) + R[3 * k + 2] * pz) + t[k]
                     for k in range(3)], 1)
    return (pool, offs, cnts.astype(np.int32), np.arange(0, 2 * n, 2, dtype=np.int32),
            np.arange(1, 2 * n, 2, dtype=np.int32))


# -------------------------------------------------------------- broad phase

LIBBP = os.path.join(_HERE, "lib", "liboracle_bp.so")


def _bp_lib():
    if "bp" not in _libs:
        if not os.path.exists(LIBBP):
            raise RuntimeError(f"oracle library missing: {LIBBP} (run __graft_entry__.build())")
        lib = C.CDLL(LIBBP)
        lib.oracle_broad_phase.restype = C.c_int64
        lib.oracle_broad_phase.argtypes = [C.c_void_p, C.c_int32, C.c_float, C.c_float, C.c_int32,
                                           C.c_int32, C.c_void_p, C.c_int64]
        lib.oracle_expand_pairs.restype = C.c_int64
        lib.oracle_expand_pairs.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p,
                                            C.c_int32, C.c_void_p, C.c_int64]
        _libs["bp"] = lib
    return _libs["bp"]


def broad_phase(spheres, cell_size: float, boundary: float, grid_size: int,
                max_per_cell: int = 0, max_pairs: int | None = None):
    """Grid broad phase over (n, 4) float32 spheres (x, y, z, radius).
    Returns (pairs (m, 2) int32 with m = min(found, max_pairs), found)."""
    s = np.ascontiguousarray(spheres, dtype=np.float32).reshape(-1, 4)
    lib = _bp_lib()
    if max_pairs is None:
        found = lib.oracle_broad_phase(_p(s), len(s), cell_size, boundary, grid_size,
                                       max_per_cell, None, 0)
        if found < 0:
            raise MemoryError("oracle broad phase scratch")
        max_pairs = found
    out = np.zeros((max(max_pairs, 1), 2), np.int32)
    found = lib.oracle_broad_phase(_p(s), len(s), cell_size, boundary, grid_size, max_per_cell,
                                   _p(out), max_pairs)
    if found < 0:
        raise MemoryError("oracle broad phase scratch")
    return out[:min(found, max_pairs)].copy(), int(found)


def expand_pairs(body_pairs, sub_off, sub_cnt, max_sub: int | None = None):
    """Body pairs -> sub-mesh pairs.  Returns (sub pairs (m, 2) int32, found)."""
    bp = np.ascontiguousarray(body_pairs, dtype=np.int32).reshape(-1, 2)
    so = np.ascontiguousarray(sub_off, dtype=np.int32)
    sc = np.ascontiguousarray(sub_cnt, dtype=np.int32)
    lib = _bp_lib()
    if max_sub is None:
        max_sub = lib.oracle_expand_pairs(_p(bp), len(bp), _p(so), _p(sc), len(sc), None, 0)
    out = np.zeros((max(max_sub, 1), 2), np.int32)
    found = lib.oracle_expand_pairs(_p(bp), len(bp), _p(so), _p(sc), len(sc), _p(out), max_sub)
    return out[:min(found, max_sub)].copy(), int(found)


# ------------------------------------------------------------ convex hulls

LIBHULL = os.path.join(_HERE, "lib", "liboracle_hull.so")
REFMESH = os.path.join(_HERE, "_ref", "libref_mesh.so")


def _hull_lib():
    if "hull" not in _libs:
        if not os.path.exists(LIBHULL):
            raise RuntimeError(f"oracle library missing: {LIBHULL} (run __graft_entry__.build())")
        lib = C.CDLL(LIBHULL)
        lib.oracle_build_hull.restype = C.c_int32
        lib.oracle_build_hull.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_int32]
        lib.oracle_hull_moments.restype = None
        lib.oracle_hull_moments.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p]
        lib.oracle_ingest_batch.restype = None
        lib.oracle_ingest_batch.argtypes = [C.c_void_p] * 3 + [C.c_int32] + [C.c_void_p] * 5 + \
            [C.c_int32]
        for f in ("oracle_hull_com", "oracle_hull_inertia_k"):
            getattr(lib, f).restype = None
            getattr(lib, f).argtypes = [C.c_void_p, C.c_void_p]
        _libs["hull"] = lib
    return _libs["hull"]


def build_hull(points, cap: int | None = None):
    """Incremental hull of (n, 3) float32 points: (nf, 3) int32 faces in the
    reference's order, or None when cap is exceeded."""
    p = np.ascontiguousarray(points, dtype=np.float32).reshape(-1, 3)
    cap = cap or max(4 * len(p), 8)
    faces = np.zeros((cap, 3), np.int32)
    nf = _hull_lib().oracle_build_hull(_p(p), len(p), _p(faces), cap)
    return None if nf < 0 else faces[:nf].copy()


def hull_moments(points, faces):
    """(volume, 4V*centroid x, y, z, second moments x^2, y^2, z^2) in FP64."""
    p = np.ascontiguousarray(points, dtype=np.float32).reshape(-1, 3)
    f = np.ascontiguousarray(faces, dtype=np.int32).reshape(-1, 3)
    m = np.zeros(7, np.float64)
    _hull_lib().oracle_hull_moments(_p(p), _p(f), len(f), _p(m))
    return m


def hull_com(m):
    m = np.ascontiguousarray(m, np.float64)
    c = np.zeros(3, np.float32)
    _hull_lib().oracle_hull_com(_p(m), _p(c))
    return c


def hull_inertia_k(m):
    m = np.ascontiguousarray(m, np.float64)
    k = np.zeros(3, np.float32)
    _hull_lib().oracle_hull_inertia_k(_p(m), _p(k))
    return k


def ingest_batch(points, off, cnt, nthreads: int = 1):
    """The reference's ingestion of every cloud (hull, centring, hull of the
    centred points, inertia factors) in C on nthreads threads.  Returns
    (centred points, faces, face_off, nfaces, inertia_k); capacity 16 n + 64
    faces per cloud."""
    p = np.ascontiguousarray(points, dtype=np.float32).reshape(-1, 3)
    off = np.ascontiguousarray(off, dtype=np.int64)
    cnt = np.ascontiguousarray(cnt, dtype=np.int32)
    n = len(cnt)
    fo = np.zeros(n + 1, np.int64)
    fo[1:] = np.cumsum(16 * cnt.astype(np.int64) + 64)
    faces = np.zeros((max(int(fo[-1]), 1), 3), np.int32)
    nf = np.zeros(n, np.int32)
    centred = np.zeros_like(p)
    k = np.zeros((n, 3), np.float32)
    _hull_lib().oracle_ingest_batch(_p(p), _p(off), _p(cnt), n, _p(faces), _p(fo), _p(nf),
                                    _p(centred), _p(k), nthreads)
    return centred, faces, fo, nf, k


def ref_mesh_available() -> bool:
    return os.path.exists(REFMESH)


class _AtlasVertex(C.Structure):
    _fields_ = [("pos", C.c_float * 3), ("normal", C.c_float * 3), ("uv", C.c_float * 2)]


class _MeshAtlas(C.Structure):
    _fields_ = [("vertices", C.POINTER(_AtlasVertex)), ("num_vertices", C.c_int),
                ("vertex_cap", C.c_int), ("indices", C.c_void_p), ("num_indices", C.c_int),
                ("index_cap", C.c_int), ("meshes", C.c_void_p), ("num_meshes", C.c_int),
                ("mesh_cap", C.c_int)]


def _ref_mesh():
    if "refmesh" not in _libs:
        lib = C.CDLL(REFMESH)
        lib.atlas_add_convex_hull.restype = C.c_int
        lib.atlas_add_convex_hull.argtypes = [C.POINTER(_MeshAtlas), C.c_void_p, C.c_int]
        lib.atlas_init.argtypes = [C.POINTER(_MeshAtlas)]
        lib.atlas_free.argtypes = [C.POINTER(_MeshAtlas)]
        lib.center_hull_verts.argtypes = [C.c_void_p, C.c_int]
        lib.compute_hull_inertia_k.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p,
                                               C.c_void_p]
        lib.gen_gjk_random_hull.restype = C.POINTER(C.c_float)
        lib.gen_gjk_random_hull.argtypes = [C.c_int, C.c_uint, C.POINTER(C.c_int)]
        for g in ("gen_gjk_icosahedron", "gen_gjk_box", "gen_gjk_tetrahedron",
                  "gen_gjk_octahedron"):
            getattr(lib, g).restype = C.POINTER(C.c_float)
            getattr(lib, g).argtypes = [C.POINTER(C.c_int)]
        _libs["refmesh"] = lib
    return _libs["refmesh"]


def ref_hull_corners(points):
    """The reference's hull as face-corner positions, (nf, 3, 3) float32, in its
    face order (through atlas_add_convex_hull, which copies each face's
    corners)."""
    lib = _ref_mesh()
    p = np.ascontiguousarray(points, dtype=np.float32).reshape(-1, 3)
    at = _MeshAtlas()
    lib.atlas_init(C.byref(at))
    try:
        mid = lib.atlas_add_convex_hull(C.byref(at), _p(p), len(p))
        if mid < 0:
            return np.zeros((0, 3, 3), np.float32)
        out = np.array([list(at.vertices[i].pos) for i in range(at.num_vertices)], np.float32)
        return out.reshape(-1, 3, 3)
    finally:
        lib.atlas_free(C.byref(at))


def ref_center(points):
    lib = _ref_mesh()
    p = np.array(points, dtype=np.float32).reshape(-1, 3)
    lib.center_hull_verts(_p(p), len(p))
    return p


def ref_inertia_k(points):
    lib = _ref_mesh()
    p = np.ascontiguousarray(points, dtype=np.float32).reshape(-1, 3)
    k = np.zeros(3, np.float32)
    lib.compute_hull_inertia_k(_p(p), len(p), _p(k[0:]), _p(k[1:]), _p(k[2:]))
    return k


def ref_gen(name: str, n: int = 0, seed: int = 0):
    """The reference's own GJK vertex clouds (gen_gjk_*), (m, 3) float32."""
    lib = _ref_mesh()
    cnt = C.c_int()
    ptr = (lib.gen_gjk_random_hull(n, seed, C.byref(cnt)) if name == "random_hull"
           else getattr(lib, f"gen_gjk_{name}")(C.byref(cnt)))
    out = np.ctypeslib.as_array(ptr, shape=(cnt.value * 3,)).reshape(-1, 3).copy()
    C.CDLL(None).free(C.cast(ptr, C.c_void_p))
    return out
