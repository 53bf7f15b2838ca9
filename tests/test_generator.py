"""The synthetic workload generator (CPU only)."""
import numpy as np

from ogjk import gen


def test_deterministic_and_independent_of_threads():
    a = gen.make(gen.ELLIPSOID, 300, seed=5, box=1.5, nmin=4, nmax=4096, nthreads=1)
    b = gen.make(gen.ELLIPSOID, 300, seed=5, box=1.5, nmin=4, nmax=4096, nthreads=8)
    assert np.array_equal(a.coords, b.coords) and np.array_equal(a.cnt, b.cnt)
    c = gen.make(gen.ELLIPSOID, 300, seed=6, box=1.5, nmin=4, nmax=4096)
    assert not np.array_equal(a.cnt, c.cnt)


def test_prefix_stability():
    """Body b does not depend on how many bodies are generated."""
    a = gen.make(gen.SPHERE, 50, seed=9, nverts=32, box=2.0)
    b = gen.make(gen.SPHERE, 500, seed=9, nverts=32, box=2.0)
    assert np.array_equal(a.coords, b.coords[:len(a.coords)])


def test_family_shapes():
    s = gen.make(gen.SPHERE, 200, seed=1, nverts=64, box=1.3)
    assert (s.cnt == 64).all()
    for k in range(20):
        pts = s.body(k).astype(np.float64)
        r = np.linalg.norm(pts - pts.mean(0), axis=1)
        assert 0.4 < r.max() < 1.9
    e = gen.make(gen.ELLIPSOID, 4000, seed=2, box=1.5, nmin=4, nmax=4096)
    assert e.cnt.min() >= 4 and e.cnt.max() <= 4096
    assert abs(np.log(e.cnt).mean() - (np.log(4) + np.log(4097)) / 2) < 0.1  # log-uniform
    n = gen.make(gen.NOISY_SPHERE, 10, seed=3, nverts=1000, box=0.0, sigma=0.01)
    rr = np.linalg.norm(n.body(0).astype(np.float64), axis=1)
    assert abs(rr.mean() - 1.0) < 0.01 and 0.005 < rr.std() < 0.02
