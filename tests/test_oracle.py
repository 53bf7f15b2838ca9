"""Pinning the oracle (CPU only).

1. Bit-exact vs the reference sources: the oracle's FP32 instantiation against
   the committed golden vectors produced by the reference CPU code
   (tests/golden/make_golden.py), and against oracle/_ref live where built.
2. FP64 known answers from the reference's README/tests and closed forms.
3. FP64 against mathematics: brute-force closest point of the Minkowski
   difference on tiny hulls, and invariants that a dropped term or wrong sign
   would break (symmetry, witness distance, EPA separation by depth).
"""
import itertools
import os

import numpy as np
import pytest

import cases
import oracle
from ogjk import gen

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref32_golden.npz")
EPS64 = np.finfo(np.float64).eps


def _bits(a):
    a = np.ascontiguousarray(a)
    return a.view(np.uint32 if a.dtype.itemsize == 4 else np.uint64).reshape(len(a), -1)


def test_oracle32_matches_reference_golden():
    g = np.load(GOLDEN, allow_pickle=False)
    coords, off, cnt, pa, pb = g["coords"], g["off"], g["cnt"], g["pa"], g["pb"]
    r = oracle.run(oracle.GJK, coords, off, cnt, pa, pb, bits=32)
    assert np.array_equal(_bits(r.distances.reshape(-1, 1)), _bits(g["gjk_dist"].reshape(-1, 1)))
    assert np.array_equal(r.simplices.view(np.uint8).reshape(-1, 108), g["gjk_simplex"])
    e = oracle.run(oracle.GJK_EPA, coords, off, cnt, pa, pb, bits=32)
    assert np.array_equal(_bits(e.distances.reshape(-1, 1)), _bits(g["epa_dist"].reshape(-1, 1)))
    assert np.array_equal(e.simplices.view(np.uint8).reshape(-1, 108), g["epa_simplex"])
    assert np.array_equal(_bits(e.normals), _bits(g["epa_normals"]))
    assert (g["epa_dist"] < 0).sum() > 100  # EPA paths are exercised


@pytest.mark.skipif(not oracle.ref_available(), reason="reference build (oracle/_ref) not present")
@pytest.mark.parametrize("name,batch", [
    ("sphere64", lambda: gen.make(gen.SPHERE, 1500, seed=21, nverts=64, box=1.29)),
    ("sphere32_literal", lambda: gen.make(gen.SPHERE, 1500, seed=22, nverts=32, box=2.0)),
    ("ellipsoid", lambda: gen.make(gen.ELLIPSOID, 600, seed=23, box=1.5, nmin=4, nmax=512)),
    ("degenerate", lambda: cases.degenerate_batch(2000, seed=24)),
])
def test_oracle32_matches_reference_live(name, batch):
    b = batch()
    r = oracle.ref_run(oracle.GJK_EPA, b.coords, b.off, b.cnt, b.pa, b.pb)
    o = oracle.run(oracle.GJK_EPA, b.coords, b.off, b.cnt, b.pa, b.pb, bits=32)
    assert np.array_equal(r.simplices.view(np.uint8), o.simplices.view(np.uint8)), name
    assert np.array_equal(_bits(r.distances.reshape(-1, 1)), _bits(o.distances.reshape(-1, 1)))
    assert np.array_equal(_bits(r.normals), _bits(o.normals))


def test_oracle64_known_answers():
    b = cases.known_answer_batch()
    r = oracle.run(oracle.GJK_EPA, b.coords, b.off, b.cnt, b.pa, b.pb, bits=64)
    w = r.simplices["witnesses"]
    # README.md:111-115 (GJKUsage): 3.653650, witnesses +-(1.025173, 1.490318, 0.255463)
    assert abs(r.distances[0] - 3.653650) < 5e-7
    assert np.allclose(w[0, 0], [1.025173, 1.490318, 0.255463], atol=5e-7)
    assert np.allclose(w[0, 1], [-1.025173, -1.490318, -0.255463], atol=5e-7)
    assert np.isclose(np.linalg.norm(w[0, 0] - w[0, 1]), r.distances[0], rtol=1e-12)
    # README.md:131-137 (EPAUsage): depth 1.5, witnesses, normal (1, -0, 0), at the
    # README's 6 decimals (the cube is rotated in float, as EPAUsage.cpp does)
    assert abs(-r.distances[1] - 1.5) < 5e-7
    assert np.allclose(w[1, 0], [1.0, 0.5, 0.707107], atol=5e-7)
    assert np.allclose(w[1, 1], [-0.5, 0.5, 0.707107], atol=5e-7)
    assert np.allclose(r.normals[1], [1.0, 0.0, 0.0], atol=1e-7)
    # example.cu:394-580: overlap 1 in x, touching, 3 apart
    assert r.distances[2] == -1.0 and abs(abs(r.normals[2, 0]) - 1.0) < 1e-15
    assert r.distances[3] == 0.0
    assert r.distances[4] == 3.0 and r.simplices["nvrtx"][4] < 4


def _closest_point_to_origin(P):
    """Brute force: min norm over conv(P) by enumerating affine projections on
    every 1-4 point subset and keeping those with non-negative weights."""
    best = np.inf
    for k in range(1, min(4, len(P)) + 1):
        for S in itertools.combinations(range(len(P)), k):
            Q = P[list(S)]
            if k == 1:
                best = min(best, np.linalg.norm(Q[0]))
                continue
            E = Q[1:] - Q[0]
            G = E @ E.T
            if abs(np.linalg.det(G)) < 1e-12:
                continue
            lam = np.linalg.solve(G, -E @ Q[0])
            wts = np.concatenate([[1 - lam.sum()], lam])
            if (wts >= -1e-12).all():
                best = min(best, np.linalg.norm(wts @ Q))
    return best


def test_oracle64_distance_bruteforce_tiny_hulls():
    rng = np.random.default_rng(7)
    A, B = [], []
    for _ in range(60):
        na, nb = rng.integers(1, 5, size=2)
        A.append(rng.normal(size=(na, 3)).astype(np.float32))
        B.append((rng.normal(size=(nb, 3)) + rng.normal(size=3) * 3).astype(np.float32))
    b = gen.from_bodies(A, B)
    r = oracle.run(oracle.GJK, b.coords, b.off, b.cnt, b.pa, b.pb, bits=64)
    for i in range(b.n_pairs):
        P = np.array([a - c for a in A[i].astype(np.float64) for c in B[i].astype(np.float64)])
        d = _closest_point_to_origin(P)
        assert abs(r.distances[i] - d) <= 1e-9 * max(1.0, d), (i, r.distances[i], d)


def test_oracle64_invariants():
    b = gen.make(gen.SPHERE, 400, seed=31, nverts=48, box=1.6)
    r = oracle.run(oracle.GJK_EPA, b.coords, b.off, b.cnt, b.pa, b.pb, bits=64)
    swap = oracle.run(oracle.GJK, b.coords, b.off, b.cnt, b.pb, b.pa, bits=64)
    sep = r.distances > 1e-9
    assert sep.sum() > 50 and (~sep).sum() > 50
    # d(A,B) == d(B,A)
    assert np.allclose(r.distances[sep], swap.distances[sep], rtol=1e-10, atol=1e-12)
    # separated: |w0 - w1| == d, normal = (w1 - w0)/d
    w = r.simplices["witnesses"]
    gap = np.linalg.norm(w[:, 0] - w[:, 1], axis=1)
    assert np.allclose(gap[sep], r.distances[sep], rtol=1e-9, atol=1e-12)
    nrm = (w[sep, 1] - w[sep, 0]) / gap[sep, None]
    assert np.allclose(nrm, r.normals[sep], atol=1e-9)
    # penetrating: moving B by (depth + 0.05) along the normal separates it by ~0.05
    pen = np.where(r.distances < -1e-6)[0][:60]
    A2, B2 = [], []
    for i in pen:
        depth = -r.distances[i]
        n = r.normals[i]
        A2.append(b.body(int(b.pa[i])))
        B2.append((b.body(int(b.pb[i])).astype(np.float64) + (depth + 0.05) * n).astype(np.float32))
    bb = gen.from_bodies(A2, B2)
    rr = oracle.run(oracle.GJK, bb.coords, bb.off, bb.cnt, bb.pa, bb.pb, bits=64)
    assert np.allclose(rr.distances, 0.05, atol=2e-3), rr.distances[:10]


def test_oracle64_epa_witnesses_pinned_by_geometry():
    """EPA witnesses in FP64 against what the geometry fixes.  At convergence
    (EPA.c:635, tolerance 100 * DBL_EPSILON) the closest face lies on the
    boundary of A - B with normal n, so: w0 lies in conv(A) on A's supporting
    plane along n, w1 in conv(B) on B's along -n, (w0 - w1).n = depth, and
    where the origin projects inside that face (most pairs) w0 - w1 = depth n.
    A swapped body index, a wrong barycentric order or sign moves the
    witnesses off these planes or off depth n."""
    b = gen.make(gen.SPHERE, 600, seed=37, nverts=48, box=1.6)
    r = oracle.run(oracle.GJK_EPA, b.coords, b.off, b.cnt, b.pa, b.pb, bits=64, stats=True)
    pen = np.where((r.distances < -1e-6) & (r.stats["epa_iters"] < 64))[0]
    assert len(pen) > 150
    w = r.simplices["witnesses"].astype(np.float64)
    dirs = np.random.default_rng(5).normal(size=(32, 3))
    tol = 1e-12
    aligned = 0
    for i in pen:
        A = b.body(int(b.pa[i])).astype(np.float64)
        B = b.body(int(b.pb[i])).astype(np.float64)
        n = r.normals[i].astype(np.float64)
        depth = -float(r.distances[i])
        assert abs(np.linalg.norm(n) - 1.0) < 1e-14
        hA, lB = (A @ n).max(), (B @ n).min()
        assert -tol <= hA - w[i, 0] @ n <= tol, i
        assert -tol <= w[i, 1] @ n - lB <= tol, i
        assert abs((w[i, 0] - w[i, 1]) @ n - depth) <= tol, i
        assert abs(depth - (hA - lB)) <= tol, i
        # inside the convex hulls: no direction sees a witness beyond the body
        assert np.all(dirs @ w[i, 0] <= (A @ dirs.T).max(axis=0) + tol), i
        assert np.all(dirs @ w[i, 1] <= (B @ dirs.T).max(axis=0) + tol), i
        aligned += np.abs(w[i, 0] - w[i, 1] - depth * n).max() <= 1e-9
    assert aligned >= 0.8 * len(pen), (aligned, len(pen))


def test_oracle64_matches_oracle32_within_fp32_error():
    b = gen.make(gen.SPHERE, 2000, seed=41, nverts=64, box=1.29)
    r64 = oracle.run(oracle.GJK, b.coords, b.off, b.cnt, b.pa, b.pb, bits=64)
    r32 = oracle.run(oracle.GJK, b.coords, b.off, b.cnt, b.pa, b.pb, bits=32)
    sep = (r64.distances > 1e-2) & (r32.distances > 1e-2)
    # FP32 reference tolerances give ~1e-3 relative (openGJK.c:41); FP64 is the checker.
    assert np.allclose(r32.distances[sep], r64.distances[sep], rtol=3e-3)


# ------------------------------------------------------------ posed bodies

def test_world_vertices_closed_forms():
    """The checker's world transform: identity and half-turn quaternions give
    exact closed forms; the vectorised pool equals the per-body transform."""
    rng = np.random.default_rng(3)
    local = rng.normal(size=(50, 3)).astype(np.float32)
    ident = np.array([1, 0, 0, 0, 1.5, -2.0, 0.25], np.float32)
    assert np.array_equal(oracle.world_vertices(local, ident),
                          local.astype(np.float64) + np.array([1.5, -2.0, 0.25]))
    half_z = np.array([0, 0, 0, 1, 0, 0, 0], np.float32)          # 180 degrees about z
    w = oracle.world_vertices(local, half_z)
    assert np.array_equal(w, local.astype(np.float64) * np.array([-1.0, -1.0, 1.0]))
    b = gen.make(gen.ELLIPSOID, 40, seed=5, box=0.0, nmin=8, nmax=200)
    pairs, poses = gen.posed_requests(len(b.cnt), 1, seed=6, kmin=30, kmax=30)[0]
    pool, off, cnt, pa, pb = oracle.posed_pool(b.coords, b.off, b.cnt, pairs, poses)
    for i in range(30):
        for s in range(2):
            k = pairs[i, s]
            ref = oracle.world_vertices(b.coords[b.off[k]:b.off[k] + b.cnt[k]], poses[i, 7 * s:7 * s + 7])
            assert np.array_equal(ref, pool[off[2 * i + s]:off[2 * i + s] + cnt[2 * i + s]])


def test_posed_rigid_invariance():
    """The same rigid motion applied to both bodies leaves distance and depth
    unchanged (up to rounding of the transformed coordinates)."""
    b = gen.make(gen.ELLIPSOID, 200, seed=8, box=0.0, nmin=8, nmax=300)
    pairs, poses = gen.posed_requests(len(b.cnt), 1, seed=9, kmin=200, kmax=200, tbox=1.5)[0]
    base = oracle.run(oracle.GJK_EPA, *oracle.posed_pool(b.coords, b.off, b.cnt, pairs, poses))
    # compose every pose with one extra rotation + translation, in FP64 on the
    # world pool directly
    pool, off, cnt, pa, pb = oracle.posed_pool(b.coords, b.off, b.cnt, pairs, poses)
    R = oracle.pose_matrix(np.array([0.5, 0.5, -0.5, 0.5], np.float32))
    moved = pool @ R.T + np.array([3.0, -1.0, 2.0])
    res = oracle.run(oracle.GJK_EPA, moved, off, cnt, pa, pb)
    assert np.allclose(res.distances, base.distances, rtol=1e-9, atol=1e-9)
