"""Write the golden vectors in this folder from the reference CPU sources.

Runs the reference GJK/EPA (GJK/cpu/openGJK.c + GJK/cpu/EPA.c, compiled as
shipped -- FP32 -- into oracle/_ref/libref32.so by oracle/Makefile) on small
seeded inputs and stores inputs and outputs together in ref32_golden.npz.
The oracle's FP32 instantiation must reproduce these bit for bit
(tests/test_oracle.py), which pins the restatement the FP64 checker shares.

Usage (in a checkout with the reference present):  python tests/golden/make_golden.py
"""
from __future__ import annotations

import hashlib
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "opengjk-gpu_amd"), os.path.join(ROOT, "tests")]

import numpy as np  # noqa: E402

import cases  # noqa: E402
import oracle  # noqa: E402
from ogjk import gen  # noqa: E402


def golden_batch():
    parts = [
        gen.make(gen.SPHERE, 96, seed=100, nverts=32, box=1.3),
        gen.make(gen.ELLIPSOID, 64, seed=101, box=1.2, nmin=4, nmax=64),
        gen.make(gen.NOISY_SPHERE, 16, seed=102, nverts=128, box=1.2, sigma=0.01),
        cases.known_answer_batch(),
        cases.degenerate_batch(256, seed=103),
    ]
    A, B = [], []
    for p in parts:
        for i in range(p.n_pairs):
            A.append(p.body(int(p.pa[i])))
            B.append(p.body(int(p.pb[i])))
    return gen.from_bodies(A, B)


def main() -> None:
    if not oracle.ref_available():
        raise SystemExit("oracle/_ref/libref32.so is not built (needs the reference checkout)")
    b = golden_batch()
    g = oracle.ref_run(oracle.GJK, b.coords, b.off, b.cnt, b.pa, b.pb)
    e = oracle.ref_run(oracle.GJK_EPA, b.coords, b.off, b.cnt, b.pa, b.pb)
    sha = hashlib.sha256(b.coords.tobytes()).hexdigest()
    np.savez_compressed(
        os.path.join(HERE, "ref32_golden.npz"), coords=b.coords, off=b.off, cnt=b.cnt, pa=b.pa,
        pb=b.pb, gjk_dist=g.distances, gjk_simplex=g.simplices.view(np.uint8).reshape(-1, 108),
        epa_dist=e.distances, epa_simplex=e.simplices.view(np.uint8).reshape(-1, 108),
        epa_normals=e.normals, coords_sha256=np.array(sha))
    print(f"wrote {b.n_pairs} pairs, {len(b.coords)} vertices, sha256 {sha[:16]}")


if __name__ == "__main__":
    main()
