"""Hand-made and tie-heavy test inputs.

Known-answer cases follow the reference's own tests:
  * userP/userQ point sets (examples/usage/userP.dat, userQ.dat) with the
    expected distance 3.653650 (README.md:111-115);
  * the EPAUsage cube pair (examples/usage/EPAUsage.cpp:31-82), expected
    depth 1.5, witnesses (1, .5, .707107) / (-.5, .5, .707107), normal
    (1, -0, 0) (README.md:131-137);
  * the EPATesting cube cases (examples/gpu/example.cu:394-580): cube vs
    cube+1 (depth band 0.8-1.2), touching cube+2 (0 <= d < 0.01), cube+5
    (distance 3, nvrtx < 4).
"""
from __future__ import annotations

import math
import os

import numpy as np

from ogjk import gen

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

CUBE = np.array([[x, y, z] for x in (-1, 1) for y in (-1, 1) for z in (-1, 1)], np.float32)


def read_dat(name: str) -> np.ndarray:
    """Point-set format of examples/main.cpp:31-77: count, then x y z rows."""
    with open(os.path.join(GOLDEN, name)) as f:
        tok = f.read().split()
    n = int(tok[0])
    return np.array(tok[1:1 + 3 * n], dtype=np.float32).reshape(n, 3)


def epa_usage_cubes():
    """Cube and a cube rotated 45 deg about x, y, z then moved +1 in x, built
    in float as examples/usage/EPAUsage.cpp:23-82 does."""
    f = np.float32
    ang = f(f(45.0) * f(math.pi) / f(180.0))
    ca, sa = f(math.cos(float(ang))), f(math.sin(float(ang)))
    c2 = []
    for x in (-1, 1):
        for y in (-1, 1):
            for z in (-1, 1):
                px, py, pz = f(x), f(y), f(z)
                ty = f(py * ca - pz * sa)
                tz = f(py * sa + pz * ca)
                py, pz = ty, tz
                tx = f(px * ca + pz * sa)
                tz = f(-px * sa + pz * ca)
                px, pz = tx, tz
                tx = f(px * ca - py * sa)
                ty = f(px * sa + py * ca)
                px, py = tx, ty
                c2.append([px + f(1.0), py, pz])
    return CUBE.copy(), np.array(c2, np.float32)


def known_answer_batch():
    P, Q = read_dat("userP.dat"), read_dat("userQ.dat")
    c1, c2 = epa_usage_cubes()
    A = [P, c1, CUBE, CUBE, CUBE]
    B = [Q, c2, CUBE + np.float32([1, 0, 0]), CUBE + np.float32([2, 0, 0]), CUBE + np.float32([5, 0, 0])]
    return gen.from_bodies(A, B)


def degenerate_batch(n: int = 4000, seed: int = 5):
    """Tie-heavy and degenerate pairs: lattice cubes, integer clouds, flat
    sets, 1-3 point bodies, duplicated vertices, grid cubes, deep overlap,
    scales from 1e-3 to 1e3."""
    rng = np.random.default_rng(seed)
    A, B = [], []
    g = np.linspace(-1, 1, 5, dtype=np.float32)
    grid = np.array([[x, y, z] for x in g for y in g for z in g if max(abs(x), abs(y), abs(z)) == 1],
                    np.float32)
    for i in range(n):
        k = i % 8
        if k == 0:
            A.append(CUBE); B.append(CUBE + rng.integers(-3, 4, size=3).astype(np.float32))
        elif k == 1:
            n1, n2 = rng.integers(1, 40, size=2)
            A.append(rng.integers(-2, 3, size=(n1, 3)).astype(np.float32))
            B.append(rng.integers(-2, 3, size=(n2, 3)).astype(np.float32))
        elif k == 2:
            n1, n2 = rng.integers(3, 30, size=2)
            a = rng.normal(size=(n1, 3)).astype(np.float32); a[:, 2] = 0
            b = rng.normal(size=(n2, 3)).astype(np.float32); b[:, 2] = rng.integers(-1, 2)
            A.append(a); B.append(b)
        elif k == 3:
            n1, n2 = rng.integers(1, 4, size=2)
            A.append(rng.normal(size=(n1, 3)).astype(np.float32))
            B.append(rng.normal(size=(n2, 3)).astype(np.float32) * np.float32(0.5))
        elif k == 4:
            a = rng.normal(size=(10, 3)).astype(np.float32); a = np.concatenate([a, a[::-1]])
            A.append(a); B.append(a * np.float32(0.7) + rng.normal(size=3).astype(np.float32) * np.float32(0.5))
        elif k == 5:
            A.append(grid); B.append(grid * np.float32(0.5) + rng.integers(-4, 5, size=3).astype(np.float32) * np.float32(0.25))
        elif k == 6:
            A.append(rng.normal(size=(50, 3)).astype(np.float32))
            B.append(rng.normal(size=(50, 3)).astype(np.float32) + np.float32(0.1))
        else:
            s = np.float32(10.0 ** rng.integers(-3, 4))
            A.append(rng.normal(size=(20, 3)).astype(np.float32) * s)
            B.append(rng.normal(size=(20, 3)).astype(np.float32) * s + s * rng.normal(size=3).astype(np.float32))
    return gen.from_bodies(A, B)


def lattice_batch(n: int = 200, seed: int = 7):
    """Large tie-heavy bodies: integer lattice points on cube surfaces (k^3
    grids, k in [6, 20], up to 2,168 vertices, shuffled), so support queries
    along axes and diagonals tie across many vertices and many clusters."""
    rng = np.random.default_rng(seed)
    A, B = [], []
    for _ in range(n):
        bodies = []
        for _ in range(2):
            k = int(rng.integers(6, 21))
            g = np.arange(k, dtype=np.float32) - np.float32((k - 1) / 2)
            pts = np.array([[x, y, z] for x in g for y in g for z in g
                            if max(abs(x), abs(y), abs(z)) == g[-1]], np.float32)
            pts = pts[rng.permutation(len(pts))]
            bodies.append(pts + rng.integers(-k, k + 1, size=3).astype(np.float32))
        A.append(bodies[0])
        B.append(bodies[1])
    return gen.from_bodies(A, B)


# ------------------------------------------------ reference EPA test shapes
#
# The shapes of the reference's EPA tests (examples/gpu/example.cu:122-262,
# cases 5-8 at :593-892).  The reference draws its random points from C
# rand(); here a seeded numpy generator draws the same kind of point set,
# computed in float32 like the reference's gkFloat arithmetic.

def ref_polytope(n: int, offset, seed: int):
    """generatePolytope (example.cu:122-137): n points at radius
    U[1, 1.5] in spherical coordinates theta U[0, 2 pi], phi U[0, pi]."""
    f = np.float32
    rng = np.random.default_rng(seed)
    pi = f(3.14159265358979323846)
    u = rng.random((n, 3)).astype(np.float32)
    theta = u[:, 0] * f(2.0) * pi
    phi = u[:, 1] * pi
    r = f(1.0) + u[:, 2] * f(0.5)
    sp = np.sin(phi.astype(np.float64)).astype(np.float32)
    p = np.stack([r * sp * np.cos(theta.astype(np.float64)).astype(np.float32),
                  r * sp * np.sin(theta.astype(np.float64)).astype(np.float32),
                  r * np.cos(phi.astype(np.float64)).astype(np.float32)], 1)
    return (p + np.asarray(offset, np.float32)).astype(np.float32)


def grid_cube(grid: int = 40, size: float = 1.0, offset=(0.0, 0.0, 0.0)):
    """generateCubeWithGrid (example.cu:145-227): grid x grid points on each
    of the six faces, in the reference's face and loop order, float32."""
    f = np.float32
    sz = f(size)
    t = np.array([-sz + (f(2.0) * sz * f(i)) / f(grid - 1) for i in range(grid)], np.float32)
    a, b = np.meshgrid(t, t, indexing="ij")
    a, b = a.reshape(-1), b.reshape(-1)
    full = np.full_like(a, sz)
    faces = [(full, a, b), (-full, a, b), (a, full, b), (a, -full, b), (a, b, full), (a, b, -full)]
    p = np.concatenate([np.stack(fc, 1) for fc in faces]).astype(np.float32)
    return (p + np.asarray(offset, np.float32)).astype(np.float32)


def rotate45_xyz_shift_x(p, shift: float = 1.0):
    """Rotate by 45 degrees about x, then y, then z, and move +shift in x, in
    float32 as example.cu:663-697 (and EPAUsage.cpp:55-80) does."""
    f = np.float32
    ang = f(f(45.0) * f(math.pi) / f(180.0))
    ca, sa = f(math.cos(float(ang))), f(math.sin(float(ang)))
    px, py, pz = (p[:, k].astype(np.float32) for k in range(3))
    py, pz = py * ca - pz * sa, py * sa + pz * ca
    px, pz = px * ca + pz * sa, -px * sa + pz * ca
    px, py = px * ca - py * sa, px * sa + py * ca
    return np.stack([px + f(shift), py, pz], 1).astype(np.float32)


def sphere_surface(n: int, radius: float, offset, seed: int):
    """generateSphereSurface (example.cu:233-254): theta = 2 pi u,
    phi = acos(2 v - 1), points at the given radius."""
    f = np.float32
    rng = np.random.default_rng(seed)
    u = rng.random(n).astype(np.float32)
    v = rng.random(n).astype(np.float32)
    pi = f(3.14159265358979323846)
    theta = f(2.0) * pi * u
    phi = np.arccos((f(2.0) * v - f(1.0)).astype(np.float64)).astype(np.float32)
    r = f(radius)
    sp = np.sin(phi.astype(np.float64)).astype(np.float32)
    p = np.stack([r * sp * np.cos(theta.astype(np.float64)).astype(np.float32),
                  r * sp * np.sin(theta.astype(np.float64)).astype(np.float32),
                  r * np.cos(phi.astype(np.float64)).astype(np.float32)], 1)
    return (p + np.asarray(offset, np.float32)).astype(np.float32)


def reference_epa_cases():
    """Cases 5-7 of the reference's EPA tests as one batch: 50-point
    polytopes 0.5 apart (case 5), 40x40-grid cubes with the second rotated
    45 degrees about every axis and moved +1 in x (case 6, 9,600 vertices
    each), 1,000-point spheres of radius 2 with centres 1 apart (case 7;
    case 8 is the same pair through separate GJK and EPA calls)."""
    A = [ref_polytope(50, (0, 0, 0), 51), grid_cube(), sphere_surface(1000, 2.0, (0, 0, 0), 71)]
    B = [ref_polytope(50, (0.5, 0, 0), 52), rotate45_xyz_shift_x(grid_cube()),
         sphere_surface(1000, 2.0, (1.0, 0, 0), 72)]
    return gen.from_bodies(A, B)


def epa_cap_pairs(n_search: int = 200_000):
    """Pairs of the uneven workload (4-4096 vertices, seed 2) that the FP64
    oracle's EPA runs to its 64-iteration cap, most of them also running out
    of its 128 face slots (GJK/cpu/EPA.c:766-775, 828-863).  Returns the
    batch of those pairs."""
    import oracle
    b = gen.make(gen.ELLIPSOID, n_search, seed=2, box=1.5, nmin=4, nmax=4096)
    r = oracle.run(oracle.GJK_EPA, b.coords, b.off, b.cnt, b.pa, b.pb, bits=64, nthreads=8,
                   stats=True)
    pick = np.where(r.stats["epa_iters"] >= 64)[0]
    A = [b.body(int(b.pa[i])) for i in pick]
    B = [b.body(int(b.pb[i])) for i in pick]
    return gen.from_bodies(A, B)
