"""Synthetic batches from libogjk_gen.so (counter-based, deterministic).

A batch holds 2n bodies in one AoS float32 pool; pair i is (body 2i, body 2i+1).
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_LIB = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "lib",
                    "libogjk_gen.so")

SPHERE, ELLIPSOID, NOISY_SPHERE = 0, 2, 3


class _Spec(C.Structure):
    _fields_ = [("family", C.c_int), ("seed", C.c_uint64), ("n_bodies", C.c_int),
                ("nverts", C.c_int), ("nmin", C.c_int), ("nmax", C.c_int),
                ("box", C.c_double), ("sigma", C.c_double)]


_lib = None


def _gen():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            raise RuntimeError(f"{_LIB} missing: run __graft_entry__.build()")
        lib = C.CDLL(_LIB)
        lib.ogjk_gen_counts.restype = C.c_int64
        lib.ogjk_gen_counts.argtypes = [C.POINTER(_Spec), C.c_void_p]
        lib.ogjk_gen_fill.restype = None
        lib.ogjk_gen_fill.argtypes = [C.POINTER(_Spec), C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]
        _lib = lib
    return _lib


@dataclass
class Batch:
    coords: np.ndarray  # (N, 3) float32 AoS pool
    off: np.ndarray     # int64 [bodies], first vertex of each body
    cnt: np.ndarray     # int32 [bodies]
    pa: np.ndarray      # int32 [pairs]
    pb: np.ndarray      # int32 [pairs]

    @property
    def n_pairs(self) -> int:
        return len(self.pa)

    def body(self, b: int) -> np.ndarray:
        return self.coords[self.off[b]:self.off[b] + self.cnt[b]]


def make(family: int, n_pairs: int, seed: int, nverts: int = 64, box: float = 2.0,
         nmin: int = 4, nmax: int = 4096, sigma: float = 0.01, nthreads: int = 8) -> Batch:
    lib = _gen()
    nb = 2 * n_pairs
    spec = _Spec(family, seed, nb, nverts, nmin, nmax, box, sigma)
    cnt = np.zeros(nb, dtype=np.int32)
    total = lib.ogjk_gen_counts(C.byref(spec), cnt.ctypes.data_as(C.c_void_p))
    off = np.zeros(nb, dtype=np.int64)
    if nb > 1:
        np.cumsum(cnt[:-1], out=off[1:])
    coords = np.empty((int(total), 3), dtype=np.float32)
    lib.ogjk_gen_fill(C.byref(spec), cnt.ctypes.data_as(C.c_void_p), off.ctypes.data_as(C.c_void_p),
                      coords.ctypes.data_as(C.c_void_p), int(nthreads))
    pa = np.arange(0, nb, 2, dtype=np.int32)
    pb = pa + 1
    return Batch(coords, off, cnt, pa, pb)


def from_bodies(bodies_a, bodies_b) -> Batch:
    """Pack explicit vertex arrays (lists of (n,3) arrays) into a Batch."""
    bodies = []
    for a, b in zip(bodies_a, bodies_b):
        bodies += [np.asarray(a, dtype=np.float32).reshape(-1, 3),
                   np.asarray(b, dtype=np.float32).reshape(-1, 3)]
    cnt = np.array([len(x) for x in bodies], dtype=np.int32)
    off = np.zeros(len(bodies), dtype=np.int64)
    if len(bodies) > 1:
        np.cumsum(cnt[:-1], out=off[1:])
    coords = np.concatenate(bodies, axis=0) if bodies else np.zeros((0, 3), np.float32)
    pa = np.arange(0, len(bodies), 2, dtype=np.int32)
    return Batch(np.ascontiguousarray(coords), off, cnt, pa, pa + 1)


def posed_requests(n_bodies: int, n_requests: int, seed: int = 4, kmin: int = 1,
                   kmax: int = 10_000, tbox: float = 10.0):
    """Serving requests over a scene of n_bodies: each has K ~ U{kmin..kmax}
    pairs (i, j), i != j uniform, and a fresh pose per pair side: unit
    quaternion (w, x, y, z) = normalised 4 x N(0,1), translation U[-tbox,
    tbox]^3.  Returns a list of (pairs (K, 2) int32, poses (K, 14) float32).
    Drawn 256 requests at a time, each chunk from its own stream spawned from
    `seed` (the result does not depend on how many threads draw them)."""
    from concurrent.futures import ThreadPoolExecutor
    starts = list(range(0, n_requests, 256))
    streams = np.random.SeedSequence(seed).spawn(len(starts))

    def chunk(c):
        rng = np.random.default_rng(streams[c])
        ks = rng.integers(kmin, kmax + 1, size=min(256, n_requests - starts[c]))
        m = int(ks.sum())
        i = rng.integers(0, n_bodies, size=m)
        j = (i + rng.integers(1, n_bodies, size=m)) % n_bodies
        q = rng.normal(size=(m, 2, 4))
        q /= np.linalg.norm(q, axis=2, keepdims=True)
        t = rng.uniform(-tbox, tbox, size=(m, 2, 3))
        poses = np.concatenate([q, t], axis=2).reshape(m, 14).astype(np.float32)
        pairs = np.stack([i, j], 1).astype(np.int32)
        cut = np.cumsum(ks)[:-1]
        return list(zip(np.split(pairs, cut), np.split(poses, cut)))

    with ThreadPoolExecutor(min(16, max(1, len(os.sched_getaffinity(0))))) as ex:
        return [r for part in ex.map(chunk, range(len(starts))) for r in part]


def spheres(n: int, seed: int = 5, box: float = 10.0, rmin: float = 0.25, rmax: float = 1.0,
            outside: float = 0.0):
    """Broad-phase input: (n, 4) float32 spheres, centres U[-box, box]^3 and
    radii U[rmin, rmax]; a fraction `outside` of the centres is pushed up to
    one box width past the faces (those clamp to the border cells)."""
    rng = np.random.default_rng(seed)
    c = rng.uniform(-box, box, size=(n, 3))
    if outside > 0:
        m = rng.random(n) < outside
        c[m] *= rng.uniform(1.0, 2.0, size=(int(m.sum()), 1))
    r = rng.uniform(rmin, rmax, size=(n, 1))
    return np.concatenate([c, r], 1).astype(np.float32)


def body_poses(n_bodies: int, seed: int = 6, tbox: float = 10.0):
    """One pose per body: unit quaternion (w, x, y, z) = normalised 4 x N(0,1),
    translation U[-tbox, tbox]^3.  (n_bodies, 7) float32."""
    rng = np.random.default_rng(seed)
    q = rng.normal(size=(n_bodies, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    t = rng.uniform(-tbox, tbox, size=(n_bodies, 3))
    return np.concatenate([q, t], 1).astype(np.float32)


def body_radii(coords, off, cnt):
    """Bounding radius about the local origin of each body, max |x|_2 over its
    vertices, rounded up to float32 and widened by 2^-20 relative so that it
    bounds the FP64 world vertices of any pose (|R x| = |x| up to the
    rounding of R from a float32 quaternion)."""
    c = np.asarray(coords, np.float64).reshape(-1, 3)
    r2 = (c * c).sum(1)
    off = np.asarray(off, np.int64)
    cnt = np.asarray(cnt, np.int64)
    out = np.zeros(len(cnt), np.float64)
    nz = cnt > 0
    out[nz] = np.sqrt(np.maximum.reduceat(r2, off[nz]))
    r = (out * (1.0 + 2.0 ** -20)).astype(np.float32)
    return np.nextafter(r, np.float32(np.inf))
