"""HIP graph capture of a scene launch (ogjk_graph_capture_scene): every
replay, with eager launches and host synchronisations in between, equals the
eager launch bit for bit."""
import numpy as np
import pytest

import ogjk
from ogjk import gen

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _records(torch, out):
    torch.cuda.synchronize()
    return out["f64"].cpu().numpy().copy(), out["simplices"].cpu().numpy().copy()


@pytest.mark.parametrize("case", ["uniform_single_bucket", "uneven_buckets_clusters"])
def test_graph_replays_match_eager(torch_cuda, case):
    torch = torch_cuda
    if case == "uniform_single_bucket":
        b = gen.make(gen.SPHERE, 3000, seed=21, nverts=64, box=1.3)
        ogjk.set_config(8, 32, False)
        sc = ogjk.DeviceScene(b)
    else:
        b = gen.make(gen.ELLIPSOID, 3000, seed=22, box=1.5, nmin=4, nmax=2048)
        ogjk.set_buckets([64, 512], [8, 16, 64], [32, 32, 64])
        sc = ogjk.DeviceScene(b)
        sc.build_clusters(256)
    try:
        n = b.n_pairs
        eager = sc.collide(3, n, f64_record=True)
        ref_rec, ref_simp = _records(torch, eager)
        g, gout = sc.capture(3, n, f64_record=True)
        for rep in range(4):
            gout["f64"].fill_(-1.0)       # the replay must rewrite every record
            gout["simplices"].fill_(0)
            other = sc.collide(3, n, f64_record=True)   # eager work between replays
            torch.cuda.synchronize()
            g.launch()
            rec, simp = _records(torch, gout)
            assert np.array_equal(rec.view(np.uint64), ref_rec.view(np.uint64)), f"replay {rep}"
            assert np.array_equal(simp, ref_simp), f"replay {rep}"
            orec, _ = _records(torch, other)
            assert np.array_equal(orec.view(np.uint64), ref_rec.view(np.uint64))
        g.close()
    finally:
        ogjk.set_config(8, 32, False)
