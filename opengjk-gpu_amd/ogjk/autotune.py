"""Autotuner for the launch shape of a resident scene: size-bucket boundaries,
lanes per pair for GJK and EPA in each bucket, and the vertex count from which
bodies are swept through their clusters.  Every candidate is timed on the
device with HIP events over whole GJK+EPA launches; the best is written to a
JSON table that `apply()` loads.

Search (coordinate descent, about seventy launches):
  1. one bucket: lanes (GJK, EPA) in LANES_1, cluster threshold in CLUSTER_MIN
     or no clusters at all;
  2. two buckets: a small-body bucket split off at each of SPLITS with lanes
     in SMALL_LANES, the large bucket keeping the step-1 choice;
  3. three buckets: the best two-bucket table with a middle bucket split off
     at each larger boundary, lanes in MID_LANES;
  4. the cluster threshold again, on the best table;
  5. pairs per workgroup: for each bucket, the GJK kernel's and then the EPA
     kernel's waves per block in WAVES (pairs per block = 64 * waves / lanes).
"""
from __future__ import annotations

import json
import time

import ogjk

LANES_1 = [(64, 64), (32, 64), (32, 32), (16, 32)]
CLUSTER_MIN = [128, 256, 512, 1024, None]  # None: sweep every body whole
WAVES = [1, 2, 4]
SPLITS = [16, 32, 64, 128, 256]
SMALL_LANES = [(8, 16), (8, 32), (16, 32)]
MID_LANES = [(16, 32), (32, 32), (32, 64)]


def _time(scene, n, reps=3):
    import torch
    out = scene.outputs(n)
    scene.collide(3, n, out=out)
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(reps):
        scene.collide(3, n, out=out)
    b.record()
    torch.cuda.synchronize()
    return a.elapsed_time(b) / reps


def _apply_table(t):
    if len(t["gjk_lanes"]) == 1:
        ogjk.set_config(t["gjk_lanes"][0], t["epa_lanes"][0], False)
    else:
        ogjk.set_buckets(t["max_verts"], t["gjk_lanes"], t["epa_lanes"])
    nb = len(t["gjk_lanes"])
    if "gjk_waves" in t or "epa_waves" in t:
        ogjk.set_block_waves(t.get("gjk_waves", [1] * nb), t.get("epa_waves", [4] * nb))


def _set_clusters(scene, cmin) -> None:
    if cmin is None:
        scene.drop_clusters()
    else:
        scene.use_clusters()
        scene.set_cluster_min(cmin)


def apply(table: dict, scene=None) -> None:
    """Make `table` the launch shape (and the scene's cluster threshold; a
    null threshold sweeps every body whole)."""
    _apply_table(table)
    if scene is not None and scene.clusters is not None:
        _set_clusters(scene, table.get("cluster_min"))


def load(path: str) -> dict:
    with open(path) as f:
        return json.load(f)


def tune(scene, n_pairs: int, log=print) -> dict:
    """Search the launch shape for `scene` (clusters built) over n_pairs pairs."""
    if scene.clusters is None:
        raise ogjk.OgjkError("build the scene's clusters before tuning")
    t0 = time.perf_counter()
    tried = []

    def run(table, cmin):
        _apply_table(table)
        _set_clusters(scene, cmin)
        ms = _time(scene, n_pairs)
        tried.append(dict(table, cluster_min=cmin, ms=ms))
        log(f"  {table['max_verts']} gjk {table['gjk_lanes']} epa {table['epa_lanes']} "
            f"waves {table.get('gjk_waves')} {table.get('epa_waves')} clusters>={cmin}: "
            f"{ms:.3f} ms")
        return ms

    best, best_ms = None, float("inf")
    for lg, le in LANES_1:
        for cmin in CLUSTER_MIN:
            t = {"max_verts": [], "gjk_lanes": [lg], "epa_lanes": [le]}
            ms = run(t, cmin)
            if ms < best_ms:
                best, best_ms, best_c = t, ms, cmin
    large = (best["gjk_lanes"][0], best["epa_lanes"][0])
    for s in SPLITS:
        for lg, le in SMALL_LANES:
            t = {"max_verts": [s], "gjk_lanes": [lg, large[0]], "epa_lanes": [le, large[1]]}
            ms = run(t, best_c)
            if ms < best_ms:
                best, best_ms = t, ms
    if len(best["gjk_lanes"]) == 2:
        s0 = best["max_verts"][0]
        for s in [x for x in (128, 256, 512, 1024) if x > s0]:
            for lg, le in MID_LANES:
                t = {"max_verts": [s0, s], "gjk_lanes": [best["gjk_lanes"][0], lg, large[0]],
                     "epa_lanes": [best["epa_lanes"][0], le, large[1]]}
                ms = run(t, best_c)
                if ms < best_ms:
                    best, best_ms = t, ms
    for cmin in CLUSTER_MIN:
        if cmin != best_c:
            ms = run(best, cmin)
            if ms < best_ms:
                best_ms, best_c = ms, cmin
    nb = len(best["gjk_lanes"])
    best = dict(best, gjk_waves=[1] * nb, epa_waves=[4] * nb)
    for b in range(nb):
        for key in ("gjk_waves", "epa_waves"):
            for w in WAVES:
                if w == best[key][b]:
                    continue
                t = dict(best, **{key: best[key][:b] + [w] + best[key][b + 1:]})
                ms = run(t, best_c)
                if ms < best_ms:
                    best, best_ms = t, ms
    result = dict(best, cluster_min=best_c, ms=best_ms, pairs=n_pairs,
                  candidates=len(tried), search_s=time.perf_counter() - t0)
    apply(result, scene)
    return result


def main() -> int:
    import argparse
    import os
    ap = argparse.ArgumentParser(description="tune the launch shape on a generated workload")
    ap.add_argument("--family", type=int, default=2)
    ap.add_argument("--pairs", type=int, default=1_000_000)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--box", type=float, default=1.5)
    ap.add_argument("--nmin", type=int, default=4)
    ap.add_argument("--nmax", type=int, default=4096)
    ap.add_argument("--out", required=True, help="JSON table to write")
    args = ap.parse_args()
    from ogjk import gen
    b = gen.make(args.family, args.pairs, seed=args.seed, box=args.box, nmin=args.nmin,
                 nmax=args.nmax, nthreads=min(len(os.sched_getaffinity(0)), 16))
    sc = ogjk.DeviceScene(b)
    sc.build_clusters(CLUSTER_MIN[0])
    t = time.perf_counter()
    _time(sc, b.n_pairs, reps=1)  # warm-up
    print(f"# warm-up {time.perf_counter() - t:.2f}s", flush=True)
    res = tune(sc, b.n_pairs, log=lambda m: print(m, flush=True))
    res["workload"] = {"family": args.family, "pairs": args.pairs, "seed": args.seed,
                       "box": args.box, "nmin": args.nmin, "nmax": args.nmax}
    with open(args.out, "w") as f:
        json.dump(res, f, indent=1)
    print(json.dumps(res), flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
