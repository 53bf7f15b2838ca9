"""Pinning the convex-hull oracle (CPU only).

The oracle restates the reference's incremental hull and volume moments
(visualization/rendering/mesh_builder.cpp:268-360, 433-481).  It must equal
the reference bit for bit: face lists (as face-corner positions, in order),
centred points and inertia factors, against the golden vectors written by
tests/golden/make_hull_golden.py from the reference's own build, and live
against oracle/_ref/libref_mesh.so where that is built.
"""
import os

import numpy as np
import pytest

import oracle

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "hull_golden.npz")


def golden_cases():
    g = np.load(GOLDEN)
    po, fo = g["point_off"], g["face_off"]
    for i, name in enumerate(g["names"]):
        yield (str(name), g["points"][po[i]:po[i + 1]], g["corners"][fo[i]:fo[i + 1]].reshape(-1, 3, 3),
               g["centred"][po[i]:po[i + 1]], g["inertia_k"][i])


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def oracle_mass(p):
    f = oracle.build_hull(p)
    m = oracle.hull_moments(p, f)
    return f, m, oracle.hull_com(m), oracle.hull_inertia_k(m)


def test_oracle_hull_matches_reference_golden():
    n = 0
    for name, p, corners, centred, k in golden_cases():
        f, m, com, kk = oracle_mass(p)
        assert f is not None, name
        assert f.shape[0] == corners.shape[0], name
        assert np.array_equal(bits(p[f]), bits(corners)), name
        shifted = p - com if m[0] >= 1e-10 else p
        assert np.array_equal(bits(shifted), bits(centred)), name
        assert np.array_equal(bits(kk), bits(k)), name
        n += 1
    assert n >= 30


@pytest.mark.skipif(not oracle.ref_mesh_available(), reason="reference mesh builder not built")
def test_oracle_hull_matches_reference_live():
    rng = np.random.default_rng(77)
    for t in range(40):
        kind = t % 4
        n = int(rng.integers(4, 600))
        if kind == 0:
            p = rng.uniform(-1, 1, (n, 3))
        elif kind == 1:
            u = rng.normal(size=(n, 3))
            p = u / np.linalg.norm(u, axis=1, keepdims=True) * rng.uniform(0.3, 2, 3)
        elif kind == 2:
            p = rng.integers(-2, 3, (n, 3)) * 0.25          # lattice: ties and coplanar faces
        else:
            p = oracle.ref_gen("random_hull", int(rng.integers(4, 33)), int(rng.integers(0, 1000)))
        p = np.ascontiguousarray(p, np.float32)
        f, m, com, kk = oracle_mass(p)
        ref = oracle.ref_hull_corners(p)
        assert np.array_equal(bits(p[f]), bits(ref)), (t, n)
        shifted = p - com if m[0] >= 1e-10 else p
        assert np.array_equal(bits(shifted), bits(oracle.ref_center(p))), (t, n)
        assert np.array_equal(bits(kk), bits(oracle.ref_inertia_k(p))), (t, n)


def test_oracle_hull_closed_forms():
    # unit cube corners: volume 1, centroid at the centre, k = (1/V) integral x^2 dV = 1/12
    # about the centre (so I_xx = m (k_y + k_z) = m / 6)
    c = np.array([[x, y, z] for x in (0, 1) for y in (0, 1) for z in (0, 1)], np.float32)
    f, m, com, k = oracle_mass(c)
    assert len(f) == 12 and abs(m[0] - 1.0) < 1e-12
    assert np.allclose(com, [0.5, 0.5, 0.5]) and np.allclose(k, 1.0 / 12.0, rtol=1e-6)
    # a tetrahedron: 4 faces, volume 1/6
    t = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], np.float32)
    f, m, com, k = oracle_mass(t)
    assert len(f) == 4 and abs(m[0] - 1.0 / 6.0) < 1e-15 and np.allclose(com, 0.25)
    # interior points never appear in faces
    rng = np.random.default_rng(5)
    inner = rng.uniform(0.2, 0.8, (50, 3)).astype(np.float32)
    f2 = oracle.build_hull(np.concatenate([c, inner]))
    assert f2.max() < 8 and len(f2) == 12
    # degenerate sets have no hull
    assert len(oracle.build_hull(c[:3])) == 0
    assert len(oracle.build_hull(np.outer(np.arange(6), [1, 2, 3]).astype(np.float32))) == 0
