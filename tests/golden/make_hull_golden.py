"""Writes tests/golden/hull_golden.npz from the reference's own mesh builder
(oracle/_ref/libref_mesh.so, built by `make -C oracle ref` from
visualization/rendering/mesh_builder.cpp where it lies).

Per point cloud: the points, the reference hull as face-corner positions in its
face order (atlas_add_convex_hull copies each face's corners), the points after
center_hull_verts, and compute_hull_inertia_k of the raw points.  Clouds: the
reference's own generators (gen_gjk_random_hull at its sizes, icosahedron, box,
tetrahedron, octahedron), random balls, spheres, ellipsoids and noisy spheres,
lattices with coplanar and repeated points, and degenerate sets (fewer than 4,
collinear, coplanar).
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
import oracle  # noqa: E402


def clouds():
    rng = np.random.default_rng(2024)
    out = []
    for n in (4, 8, 16, 24, 32):
        for seed in (1, 7, 99):
            out.append((f"ref_random_hull_{n}_{seed}", oracle.ref_gen("random_hull", n, seed)))
    for g in ("icosahedron", "box", "tetrahedron", "octahedron"):
        out.append((f"ref_{g}", oracle.ref_gen(g)))
    for n in (10, 50, 200, 1000):
        out.append((f"ball_{n}", rng.uniform(-1, 1, (n, 3))))
    for n in (64, 500, 2000):
        u = rng.normal(size=(n, 3))
        out.append((f"sphere_{n}", u / np.linalg.norm(u, axis=1, keepdims=True)))
    u = rng.normal(size=(3000, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    out.append(("noisy_sphere_3000", u + rng.normal(scale=0.01, size=u.shape)))
    u = rng.normal(size=(800, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    out.append(("ellipsoid_800", u * np.array([2.0, 0.7, 0.3]) + np.array([5.0, -3.0, 1.0])))
    g = np.stack(np.meshgrid(*[np.arange(3)] * 3, indexing="ij"), -1).reshape(-1, 3)
    out.append(("lattice_27", g - 1.0))
    g4 = np.stack(np.meshgrid(*[np.arange(4)] * 3, indexing="ij"), -1).reshape(-1, 3)
    out.append(("lattice_64_shuffled", rng.permutation(g4) * 0.5))
    box = oracle.ref_gen("box")
    out.append(("box_repeated", np.concatenate([box, box, box[::-1]])))
    out.append(("three_points", rng.uniform(-1, 1, (3, 3))))
    t = rng.uniform(-1, 1, (10, 1))
    out.append(("collinear_10", t * np.array([[1.0, 2.0, -0.5]])))
    a, b = rng.uniform(-1, 1, (2, 20, 1))
    out.append(("coplanar_20", a * np.array([[1.0, 0.0, 0.3]]) + b * np.array([[0.0, 1.0, -0.2]])))
    return [(name, np.ascontiguousarray(p, dtype=np.float32)) for name, p in out]


def main():
    if not oracle.ref_mesh_available():
        raise SystemExit("oracle/_ref/libref_mesh.so missing: run `make -C oracle ref`")
    names, pts, poff, corners, foff, centred, kk = [], [], [0], [], [0], [], []
    for name, p in clouds():
        c = oracle.ref_hull_corners(p)
        names.append(name)
        pts.append(p)
        poff.append(poff[-1] + len(p))
        corners.append(c.reshape(-1, 9))
        foff.append(foff[-1] + len(c))
        centred.append(oracle.ref_center(p))
        kk.append(oracle.ref_inertia_k(p))
        print(f"{name}: {len(p)} points, {len(c)} faces", flush=True)
    np.savez_compressed(os.path.join(HERE, "hull_golden.npz"), names=np.array(names),
                        points=np.concatenate(pts), point_off=np.array(poff, np.int64),
                        corners=np.concatenate(corners), face_off=np.array(foff, np.int64),
                        centred=np.concatenate(centred), inertia_k=np.stack(kk))


if __name__ == "__main__":
    main()
