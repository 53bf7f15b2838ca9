#!/usr/bin/env python3
"""Benchmark: batched GJK+EPA pairs/s on one MI355X.

The default run (--workload all) measures every configuration of the
narrow-phase path and prints ONE JSON line:
  headline (value, roofline, cpu_baseline, parity): the uneven workload,
      1,000,000 pairs, vertex counts log-uniform in [4, 4096] per body, points
      on rotated ellipsoids (axes U[0.3,2]), centres U[-1.5,1.5]^3, GJK then
      EPA, against the plain FP64 per-pair kernels (several launch shapes,
      the fastest is the denominator);
  "configs": sub-records of the same form for
      sphere32_gjk: 1,000 pairs x 32 vertices, GJK only (radius U[0.5,1.5],
          centres U[-2,2]^3, 25% overlap measured);
      sphere64: 100,000 pairs x 64 vertices, GJK then EPA (centres
          U[-1.29,1.29]^3, ~60% penetrating);
      noisy_sphere_large: 10,000 pairs x 100,000 vertices (unit-sphere points
          + N(0, 0.01^2), centres U[-1.5,1.5]^3, ~50% penetrating);
      serving: 20,000-body resident scene (8-512 vertices), Poisson requests
          of K ~ U{1..10000} posed pairs, p50/p99 latency at a stated QPS.
Single workloads: --workload ellipsoid_uneven | sphere64 | sphere32_gjk |
noisy_sphere_large | serving | broadphase | scene_step | hulls.

Every pair runs FP32 vertex storage, an FP32 support sweep with an exact FP64
re-rank, FP64 sub-distance and EPA arithmetic; large bodies are swept through
their vertex clusters.  A step is one pass over the whole batch (GJK kernel,
device worklist of penetrating pairs, persistent EPA kernel), inputs resident
in HBM.  K steps are timed between device synchronisations; per-step kernel
times come from HIP events recorded on the launch stream.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [ROOT, os.path.join(ROOT, "opengjk-gpu_amd")]

import numpy as np  # noqa: E402

METRIC = "GJK+EPA pairs/s vs plain FP64 GPU path at 1e-5 rel err; p50/p99 query latency"
HBM_PEAK_GBS = 8000.0
EPS64 = 2.220446049250313e-16
PIPELINE, FUSED, SPLIT = 0, 1, 2  # launch shapes (ogjk_set_config)

# cfg: the configuration's short name; gen: generator arguments; mode: 1 GJK,
# 3 GJK then EPA; lanes: (GJK, EPA) lanes per pair; cluster_min: bodies with at
# least this many vertices are swept through their clusters (None: no
# clusters); profile: the rocprofv3 summary holding the measured HBM traffic
# of one step; graph: also time the step replayed as one HIP graph; min_steps:
# at least this many timed steps, mixed and plain alike (a 30 us step timed
# 10 times is mostly launch jitter).
WORKLOADS = {
    "ellipsoid_uneven": dict(
        cfg="C2", mode=3, gen=dict(family=2, n_pairs=1_000_000, seed=2, box=1.5, nmin=4, nmax=4096),
        desc="1000000 pairs, vertex counts log-uniform in [4,4096] per body, ellipsoids axes "
             "U[0.3,2] rotated, centres U[-1.5,1.5]^3, GJK then EPA",
        lanes=(64, 64), cluster_min=512, profile="gjk_epa_ellipsoid_uneven_1m.md",
        tuning="ellipsoid_uneven_1m.json", graph=True),
    "sphere64": dict(
        cfg="C1", mode=3, gen=dict(family=0, n_pairs=100_000, seed=1, nverts=64, box=1.29),
        desc="100000 pairs x 64-vertex sphere polytopes, radius U[0.5,1.5], centres "
             "U[-1.29,1.29]^3 (~60% penetrating), GJK then EPA",
        lanes=(8, 32), cluster_min=None, profile="gjk_epa_sphere64_100k.md", graph=True),
    "sphere32_gjk": dict(
        cfg="C0", mode=1, gen=dict(family=0, n_pairs=1000, seed=0, nverts=32, box=2.0),
        desc="1000 pairs x 32-vertex sphere polytopes, radius U[0.5,1.5], centres U[-2,2]^3 "
             "(the literal spec: 25% overlap), GJK distance only",
        lanes=(64, 32), cluster_min=None, profile="gjk_sphere32_1k.md", graph=False,
        min_steps=200),
    "noisy_sphere_large": dict(
        cfg="C3", mode=3,
        gen=dict(family=3, n_pairs=10_000, seed=3, nverts=100_000, box=1.5, sigma=0.01),
        desc="10000 pairs x 100000-vertex polytopes, unit-sphere points + N(0,0.01^2), centres "
             "U[-1.5,1.5]^3 (~50% penetrating), GJK then EPA",
        lanes=(64, 64), cluster_min=1024, profile="gjk_epa_noisy_sphere_100k.md", graph=False),
}

# Plain FP64 per-pair kernels over FP64 vertex storage (no FP32 sweep, no
# worklist, no clusters), every one producing the same results as the mixed
# path.  (name, launch shape, GJK lanes, EPA lanes, layout):
#   fused: one kernel, one group per pair runs GJK then EPA;
#   split: the reference's shape, a GJK kernel over every pair then an EPA
#          kernel over every pair (GJK/gpu/openGJK.cu:2968-3004); 8/32 lanes
#          over an AoS pool is the reference's own configuration, 1 GJK lane
#          is one thread per pair.
PLAIN_GJK_EPA = [
    ("fused_16", FUSED, 16, 16, "soa"), ("fused_32", FUSED, 32, 32, "soa"),
    ("fused_64", FUSED, 64, 64, "soa"),
    ("split_8_32_aos", SPLIT, 8, 32, "aos"), ("split_8_32", SPLIT, 8, 32, "soa"),
    ("split_1_32", SPLIT, 1, 32, "soa"),
]
PLAIN_GJK = [
    ("gjk_8_aos", SPLIT, 8, 32, "aos"), ("gjk_8", SPLIT, 8, 32, "soa"),
    ("gjk_16", SPLIT, 16, 32, "soa"), ("gjk_32", SPLIT, 32, 32, "soa"),
    ("gjk_64", SPLIT, 64, 32, "soa"), ("gjk_1", SPLIT, 1, 32, "soa"),
]

# rocprofv3 summaries of the legs that are not in WORKLOADS
LEG_PROFILES = {"broadphase": "broadphase_1m.md", "scene_step": "scene_step_20k.md",
                "hulls": "hulls_100k.md", "serving": "serving_posed_10k.md"}


def log(msg: str) -> None:
    print(f"# {msg}", file=sys.stderr, flush=True)


def cores() -> int:
    return min(len(os.sched_getaffinity(0)), 16)


def profile_value(workload: str, key: str):
    """A figure from the committed rocprofv3 summary of a workload (e.g.
    traffic_bytes_per_step: FETCH_SIZE and WRITE_SIZE scaled by the copy-kernel
    calibration), or None."""
    name = WORKLOADS[workload]["profile"] if workload in WORKLOADS else LEG_PROFILES.get(workload)
    if name is None:
        return None
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.isfile(path):
        return None
    for line in open(path):
        if line.strip().startswith(key + ":"):
            return float(line.split(":", 1)[1].strip())
    return None


def compulsory_bytes(batch, mode: int) -> float:
    """Every vertex read once (12 B FP32 SoA), the two body-table records per
    pair (off 8 + cnt 4 + r1 4 B each), and the results written through the
    gkSimplex ABI (108 B simplex + 4 B distance, + 12 B normal with EPA)."""
    verts = int(batch.cnt[batch.pa].astype(np.int64).sum() + batch.cnt[batch.pb].astype(np.int64).sum())
    n = batch.n_pairs
    return 12.0 * verts + 32.0 * n + (124.0 if mode & 2 else 112.0) * n


def _event_steps(torch, fn, steps: int, warmup: int):
    """Warm up, then run `steps` calls of fn between HIP events on the
    current stream.  Returns (wall seconds, per-step event ms)."""
    stream = torch.cuda.current_stream()
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
           for _ in range(steps)]
    t = time.perf_counter()
    for a, b in evs:
        a.record(stream)
        fn()
        b.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t
    return wall, np.array([a.elapsed_time(b) for a, b in evs])


def _auto_steps(torch, fn, budget_s: float, cap: int) -> int:
    """Steps that take about budget_s (at least 2, at most cap)."""
    fn()
    torch.cuda.synchronize()
    t = time.perf_counter()
    fn()
    torch.cuda.synchronize()
    one = time.perf_counter() - t
    return int(max(2, min(cap, budget_s / max(one, 1e-6))))


# ------------------------------------------------------------ CPU baseline

def cpu_baseline(batch, mode: int, budget_s: float = 15.0):
    """CPU baseline: the FP64 oracle (the reference CPU openGJK + EPA,
    restated) on this host's cores over a bounded sample of the workload,
    the first m pairs with m sized for about budget_s of single-core work.
    The sample's bodies are gathered into an FP64 pool of their own before
    the timer starts, so only GJK/EPA is timed."""
    import oracle
    n = batch.n_pairs
    nth = cores()
    # single-core probe, doubled until it takes 0.5 s, sizes the sample
    probe, el1 = 8, 0.0
    while True:
        probe = min(n, probe * 2)
        sub = oracle.compact(batch.coords, batch.off, batch.cnt, batch.pa[:probe], batch.pb[:probe])
        t = time.perf_counter()
        oracle.run(mode, *sub, bits=64, nthreads=1)
        el1 = time.perf_counter() - t
        if el1 > 0.5 or probe == n:
            break
    per_pair_1c = el1 / probe
    m = int(min(n, max(probe, budget_s / max(per_pair_1c, 1e-9))))
    sub = oracle.compact(batch.coords, batch.off, batch.cnt, batch.pa[:m], batch.pb[:m])
    t = time.perf_counter()
    oracle.run(mode, *sub, bits=64, nthreads=nth)
    el = time.perf_counter() - t
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"value": m / el, "unit": "pairs/s", "cores": nth, "kind": "port",
            "single_core_pairs_per_s": 1.0 / per_pair_1c,
            "single_core_us_per_pair": per_pair_1c * 1e6,
            "sample": f"first {m} pairs of the workload ({m / n:.1%}), FP64 oracle (the reference "
                      f"CPU openGJK+EPA restated, GJK/cpu/openGJK.c + EPA.c), {nth} threads, "
                      f"{el:.2f}s; the sample's bodies gathered into an FP64 pool before the "
                      f"timer; single core {per_pair_1c * 1e6:.2f} us/pair; CPU: {model}"}


def parity_full(batch, mode: int, rec: np.ndarray, simp=None):
    """Every pair of the batch: the GPU's FP64 records against the FP64 oracle
    (run on all host cores over the float32 pool, promoted pair by pair)."""
    import oracle
    t = time.perf_counter()
    ref = oracle.run(mode, batch.coords, batch.off, batch.cnt, batch.pa, batch.pb, bits=64,
                     nthreads=cores(), stats=True)
    el = time.perf_counter() - t
    d, dr = rec[:, 0], ref.distances
    w0r, w1r = ref.simplices["witnesses"][:, 0], ref.simplices["witnesses"][:, 1]
    rel = np.abs(d - dr) / np.maximum(np.abs(dr), 1e-3)
    w = np.maximum(np.linalg.norm(rec[:, 1:4] - w0r, axis=1), np.linalg.norm(rec[:, 4:7] - w1r, axis=1))
    exact = (d == dr) & (rec[:, 1:4] == w0r).all(1) & (rec[:, 4:7] == w1r).all(1)
    out = {"pairs_checked": int(len(d)), "max_rel_err_distance": float(np.nanmax(rel)),
           "max_abs_err_witness": float(np.nanmax(w)), "tolerance": {
               "distance_rel": 1e-5, "floor": 1e-3, "witness_normal_abs": 1e-4}}
    if mode & 2:
        nn = np.linalg.norm(rec[:, 7:10] - ref.normals, axis=1)
        exact &= (rec[:, 7:10] == ref.normals).all(1)
        out["max_abs_err_normal"] = float(np.nanmax(nn))
    if simp is not None:
        exact &= simp["nvrtx"] == ref.simplices["nvrtx"]
        exact &= (simp["vrtx_idx"] == ref.simplices["vrtx_idx"]).all((1, 2))
    st = ref.stats
    out.update(bit_exact_fraction=float(exact.mean()),
               overlap_fraction=float((dr <= EPS64).mean()),
               oracle_s=el, oracle_threads=cores(),
               epa_iterations_p50_p99_max=[int(np.percentile(st["epa_iters"], 50)),
                                           int(np.percentile(st["epa_iters"], 99)),
                                           int(st["epa_iters"].max())] if mode & 2 else None)
    na = batch.cnt[batch.pa].astype(np.float64)
    nb = batch.cnt[batch.pb].astype(np.float64)
    stream_bytes = float((12.0 * (na * st["sweeps_a"] + nb * st["sweeps_b"])).sum())
    return out, stream_bytes


# ---------------------------------------------------------- batch workloads

def batch_leg(name: str, args, headline: bool) -> dict:
    """One batch configuration: the mixed-precision step, its graph replay,
    the plain FP64 per-pair kernels, roofline, CPU baseline and full-batch
    parity.  Returns the record."""
    import torch
    import ogjk
    from ogjk import autotune, gen
    spec = WORKLOADS[name]
    mode = spec["mode"]
    steps = args.steps if headline else max(2, min(args.steps, 10))
    steps = max(steps, spec.get("min_steps", 0))
    warmup = args.warmup if headline else min(args.warmup, 2)
    lg, le = (int(x) for x in args.lanes.split(",")) if (args.lanes and headline) else spec["lanes"]
    cmin = spec["cluster_min"]
    if headline and args.cluster_min is not None:
        cmin = args.cluster_min or None
    table = None
    tpath = os.path.join(ROOT, "opengjk-gpu_amd", "tuning", spec.get("tuning") or "-")
    if not (args.lanes and headline) and os.path.isfile(tpath):
        table = autotune.load(tpath)
        if not (headline and args.cluster_min is not None):
            cmin = table["cluster_min"]  # None: the table sweeps every body whole
    ogjk.set_config(lg, le, PIPELINE)
    wl = dict(spec["gen"])
    if args.pairs and headline:
        wl["n_pairs"] = args.pairs
    t0 = time.perf_counter()
    batch = gen.make(nthreads=cores(), **wl)
    n = batch.n_pairs
    log(f"[{spec['cfg']} {name}] generated {n} pairs, {batch.coords.shape[0]} vertices in "
        f"{time.perf_counter() - t0:.1f}s")
    plain_on = not args.no_plain
    scene = ogjk.DeviceScene(batch, f64="both" if plain_on else False)
    build_s = None
    if cmin:
        # the first cluster build of a process also loads the build kernels'
        # code object: build a 64-pair scene first, so build_s is the cost of
        # this geometry (its 16 B per slot of cluster arrays included)
        warm = ogjk.DeviceScene(gen.make(gen.ELLIPSOID, 64, seed=1, nmin=4, nmax=6000))
        warm.build_clusters(1)
        torch.cuda.synchronize()
        del warm
        t0 = time.perf_counter()
        scene.build_clusters(cmin)
        torch.cuda.synchronize()
        build_s = time.perf_counter() - t0
        log(f"[{spec['cfg']}] clusters built in {build_s:.3f}s")
    if table is not None:
        autotune.apply(table)
    # verification record (untimed)
    out = scene.outputs(n, f64=True)
    scene.collide(mode, n, out=out, f64_record=True)
    torch.cuda.synchronize()
    rec = out["f64"].cpu().numpy()
    simp = ogjk.simplices_from_device(out["simplices"])
    timed_out = scene.outputs(n, f64=False)

    def step():
        scene.collide(mode, n, out=timed_out)

    wall, ms = _event_steps(torch, step, steps, warmup)
    value = n * steps / wall
    k_ms = float(np.mean(ms))
    log(f"[{spec['cfg']}] mixed: {wall / steps * 1e3:.3f} ms/step wall, kernel mean {k_ms:.3f} ms")
    identical = bool(all(torch.equal(timed_out[k], out[k]) for k in ("simplices", "distances",
                                                                    "normals")))
    rec_out = {
        "value": value, "unit": "pairs/s", "steps": steps, "warmup": warmup,
        "ms_per_step": wall / steps * 1e3, "step_ms_p50": float(np.percentile(ms, 50)),
        "step_ms_p99": float(np.percentile(ms, 99)),
        "config": {"workload": spec["desc"], "pairs": n, "vertices": int(batch.coords.shape[0]),
                   "mode": "GJK then EPA" if mode == 3 else "GJK only",
                   "lanes_per_pair": ({"buckets_max_verts": table["max_verts"],
                                       "gjk": table["gjk_lanes"], "epa": table["epa_lanes"],
                                       "gjk_waves_per_block": table.get("gjk_waves"),
                                       "epa_waves_per_block": table.get("epa_waves"),
                                       "table": "opengjk-gpu_amd/tuning/" + spec["tuning"]}
                                      if table else {"gjk": lg, "epa": le}),
                   "clusters": {"min_verts": cmin, "build_s": build_s,
                                "build_timing": "after a warm-up build of a 64-pair scene (the "
                                                "code-object load excluded); includes allocating "
                                                "the cluster arrays"} if cmin else None,
                   "layout": "SoA fp32 scene; GJK kernel -> device worklist -> persistent EPA "
                             "kernel" if mode == 3 else "SoA fp32 scene; GJK kernel"},
        "timed_equals_verification_run": identical,
    }
    if spec["graph"]:
        # the same step captured once as a HIP graph and replayed: no host work
        # per step; its outputs must equal the eager step's bit for bit
        gout = scene.outputs(n, f64=False)
        graph, _ = scene.capture(mode, n, out=gout)
        g_wall, g_ms = _event_steps(torch, graph.launch, steps, warmup)
        rec_out["graph"] = {"ms_per_step": g_wall / steps * 1e3, "pairs_per_s": n * steps / g_wall,
                            "step_ms_p50": float(np.percentile(g_ms, 50)),
                            "identical_to_eager": bool(all(torch.equal(gout[k], timed_out[k])
                                                           for k in ("simplices", "distances",
                                                                     "normals")))}
        graph.close()
        del gout
        log(f"[{spec['cfg']}] graph replay: {rec_out['graph']['ms_per_step']:.3f} ms/step")

    if plain_on:
        plain = {"variants": {}}
        chk = scene.outputs(n, f64=True)
        for vname, shape, gl, el, layout in (PLAIN_GJK_EPA if mode == 3 else PLAIN_GJK):
            ogjk.set_config(gl, el, shape)
            which = "aos" if layout == "aos" else True

            def pstep():
                scene.collide(mode, n, out=timed_out, plain_f64=which)

            psteps = _auto_steps(torch, pstep, 1.5 if headline else 1.0, steps)
            p_wall, p_ms = _event_steps(torch, pstep, psteps, 1)
            # the variant's results equal the mixed path's, pair by pair
            scene.collide(mode, n, out=chk, plain_f64=which, f64_record=True)
            torch.cuda.synchronize()
            same = float((chk["f64"] == out["f64"]).all(dim=1).float().mean())
            plain["variants"][vname] = {
                "pairs_per_s": n * psteps / p_wall, "ms_per_step": p_wall / psteps * 1e3,
                "kernel_ms_mean": float(np.mean(p_ms)), "steps": psteps,
                "launch": ("GJK kernel then EPA kernel, one group per pair each" if shape == SPLIT
                           else "one kernel, one group per pair"),
                "lanes": {"gjk": gl, "epa": el if mode == 3 else None}, "layout": layout,
                "identical_to_mixed_fraction": same}
            log(f"[{spec['cfg']}] plain fp64 {vname}: {p_wall / psteps * 1e3:.3f} ms/step, "
                f"identical {same:.4f}")
        del chk
        ogjk.set_config(lg, le, PIPELINE)
        if table is not None:
            autotune.apply(table)
        best_name, best = max(plain["variants"].items(), key=lambda kv: kv[1]["pairs_per_s"])
        plain.update(best=best_name, pairs_per_s=best["pairs_per_s"],
                     ms_per_step=best["ms_per_step"])
        rec_out["plain_f64"] = plain
        rec_out["speedup_vs_plain_f64"] = value / best["pairs_per_s"]
        # one-shot: the per-geometry cluster build charged to a single step
        one_shot_s = wall / steps + (build_s or 0.0)
        rec_out["speedup_vs_plain_f64_one_shot"] = (n / one_shot_s) / best["pairs_per_s"]
        rec_out["speedup_basis"] = (
            "speedup_vs_plain_f64: steady state, geometry (scene and clusters) resident and "
            "reused across steps; speedup_vs_plain_f64_one_shot: one step plus the cluster "
            "build of its geometry" + ("" if cmin else " (no clusters: the two are equal)"))
    del timed_out
    # roofline of the dominant kernels (the whole step)
    algo = compulsory_bytes(batch, mode)
    rec_out["roofline"] = {
        "bound": "hbm", "achieved": algo / (k_ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
        "unit": "GB/s", "frac": algo / (k_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "traffic": profile_value(name, "traffic_bytes_per_step"),
        "kernel": "k_gjk + k_epa_list (one step)" if mode == 3 else "k_gjk (one step)",
        "kernel_ms_mean": k_ms, "algorithmic_bytes_per_launch": algo,
        "algorithmic_bytes": "compulsory: every vertex once (12 B) + 32 B body tables + results "
                             "(124 B gkSimplex/distance/normal per pair)"}
    del scene
    torch.cuda.empty_cache()
    if not args.no_cpu:
        rec_out["cpu_baseline"] = cpu_baseline(batch, mode, budget_s=15.0 if headline else 8.0)
        log(f"[{spec['cfg']}] cpu baseline: {rec_out['cpu_baseline']['value']:.0f} pairs/s on "
            f"{rec_out['cpu_baseline']['cores']} threads")
        parity, stream_bytes = parity_full(batch, mode, rec, simp)
        rec_out["parity"] = parity
        rec_out["roofline"]["stream_bytes_per_launch"] = stream_bytes
        log(f"[{spec['cfg']}] parity: {parity['pairs_checked']} pairs, bit-exact "
            f"{parity['bit_exact_fraction']:.6f}, oracle {parity['oracle_s']:.1f}s")
    if name == "noisy_sphere_large":
        large_roofline(rec_out, name)
    if name == "sphere32_gjk":
        small_extras(rec_out, batch, args)
    return rec_out


def large_roofline(rec: dict, name: str) -> None:
    """Large hulls are swept through their clusters, so neither every vertex
    once (most are never read) nor the oracle's re-streams (culling skips
    them) is the data the step must move.  The fraction quoted is the
    measured HBM bytes of one step (rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-
    calibrated) over the step's time; the other two figures ride along."""
    r = rec["roofline"]
    traffic = r.get("traffic")
    comp = r["algorithmic_bytes_per_launch"]
    r["compulsory_bytes_per_launch"] = comp
    r["frac_compulsory"] = r["frac"]
    if traffic:
        ach = traffic / (r["kernel_ms_mean"] * 1e-3) / 1e9
        r.update(achieved=ach, frac=ach / HBM_PEAK_GBS, algorithmic_bytes_per_launch=traffic,
                 algorithmic_bytes="measured HBM bytes per step of the culled sweeps (vertices of "
                                   "the clusters opened + cluster bounds + records); every vertex "
                                   "once would be compulsory_bytes_per_launch")


def small_extras(rec: dict, batch, args) -> None:
    """The 1,000-pair GJK configuration is the closest to the reference's
    published figures: also time it end to end through the drop-in (host
    staging, upload, kernel, download) and set it beside
    the reference's published kernel time."""
    import ogjk
    lg, le = WORKLOADS["sphere32_gjk"]["lanes"]
    ogjk.set_config(lg, le, PIPELINE)
    ogjk.collide_host(1, batch)  # warm-up
    reps = 20
    t = time.perf_counter()
    for _ in range(reps):
        ogjk.collide_host(1, batch)
    el = (time.perf_counter() - t) / reps
    rec["end_to_end_dropin"] = {"ms_per_call": el * 1e3, "pairs_per_s": batch.n_pairs / el,
                                "call": "ogjk_collide_host (compute_minimum_distance's path: "
                                        "host staging into kept pinned memory, one upload, "
                                        "kernel, one download)"}
    rec["published_reference"] = {
        "kernel_ms": 0.0917, "pairs": 1000, "vertices": 50, "precision": "FP32",
        "gpu": "RTX 4070 (data file name)", "source": "data/data_32bit_4070:2",
        "note": "kernel-only cudaEvent time of the reference's FP32 warp-parallel GJK on random "
                "sphere points r U[1,1.5], offsets U[-5,5]^3 (mostly separated); not the same "
                "inputs, precision or hardware"}


# ----------------------------------------------------------------- serving

def serving(args, headline: bool = True) -> dict:
    """Serving leg: a resident scene of 20,000 bodies (8-512 vertices,
    log-uniform, ellipsoids in their local frames); requests arrive by a
    Poisson process at --qps, each with K ~ U{1..10000} pairs (i != j
    uniform) and a fresh pose per pair side (uniform unit quaternion,
    translation U[-10,10]^3); GJK+EPA through the serving loop.  Latency is
    submit to results in host memory."""
    import threading
    import torch
    import ogjk
    from ogjk import gen
    lg, le = (int(x) for x in args.lanes.split(",")) if args.lanes else (64, 64)
    # 64 / 64 lanes for hulls of up to 512 vertices: kernel of 10,000 posed pairs 0.188 ms, 0.212
    # at 32 / 64, 0.261 at 16 / 64, 0.253 at 32 / 32
    ogjk.set_config(lg, le, PIPELINE)
    b = gen.make(gen.ELLIPSOID, 10_000, seed=4, box=0.0, nmin=8, nmax=512, nthreads=cores())
    scene = ogjk.DeviceScene(b)
    nb = len(b.cnt)
    n_req = args.requests
    reqs = gen.posed_requests(nb, n_req, seed=4)
    total_pairs = int(sum(len(p) for p, _ in reqs))
    log(f"[C4 serving] scene {nb} bodies, {b.coords.shape[0]} vertices; {n_req} requests, "
        f"{total_pairs} pairs")

    # kernel roofline: one device launch of 10,000 posed pairs, HIP events
    kp, kq = reqs[0]
    while len(kp) < 10_000:
        extra = gen.posed_requests(nb, 1, seed=len(kp) + 99)[0]
        kp, kq = np.concatenate([kp, extra[0]]), np.concatenate([kq, extra[1]])
    kp, kq = torch.from_numpy(kp[:10_000]).cuda(), torch.from_numpy(kq[:10_000]).cuda()
    kout = scene.collide_posed(3, kp, kq)
    torch.cuda.synchronize()
    _, k_ms_arr = _event_steps(torch, lambda: scene.collide_posed(3, kp, kq, out=kout), 20, 2)
    k_ms = float(np.mean(k_ms_arr))
    cnt = b.cnt.astype(np.int64)
    kpn = kp.cpu().numpy()
    k_bytes = float(12 * (cnt[kpn[:, 0]] + cnt[kpn[:, 1]]).sum() + 10_000 * (8 + 56 + 32 + 124))

    if args.kernel_only:
        # profiling: the 10,000-pair posed launch alone, no request loop
        return {"value": 10_000 / (k_ms * 1e-3), "unit": "pairs/s", "steps": 20, "warmup": 2,
                "ms_per_step": k_ms, "config": {"workload": "posed launch of 10000 pairs of the "
                                                            "serving scene"}}

    def run_at(qps, rng_seed, n_run=n_req):
        srv = ogjk.Server(scene, max_batch_pairs=65536, slots=args.slots)
        rng = np.random.default_rng(rng_seed)
        gaps = rng.exponential(1.0 / qps, size=n_run)
        arrive = np.cumsum(gaps)
        tickets = [None] * n_run
        done_ev = threading.Event()

        def submitter():
            t0 = time.perf_counter()
            for k in range(n_run):
                while time.perf_counter() - t0 < arrive[k]:
                    pass
                tickets[k] = srv.submit(*reqs[k])
            done_ev.set()

        th = threading.Thread(target=submitter)
        t_start = time.perf_counter()
        th.start()
        lats = np.zeros(n_run)
        for k in range(n_run):
            while tickets[k] is None:
                time.sleep(0)
            _, lat, _ = srv.wait(tickets[k])
            lats[k] = lat
        wall = time.perf_counter() - t_start
        th.join()
        info = srv.info()
        srv.close()
        return {"qps": qps, "p50_ms": float(np.percentile(lats, 50) * 1e3),
                "p99_ms": float(np.percentile(lats, 99) * 1e3),
                "max_ms": float(lats.max() * 1e3),
                "pairs_per_s": sum(len(reqs[k][0]) for k in range(n_run)) / wall,
                "launches": info["launches"], "slots": info["slots"], "wall_s": wall,
                "requests": n_run}

    rates = [250, 500, 2000, 4000, 6000] if headline else [2000, 4000, 6000]
    # untimed pass first: first launches of every kernel shape the server
    # uses, first use of a server's pinned staging
    run_at(max(rates), 5, min(n_req, 400))
    sweep = []
    # the stated rate runs args.requests (>= 10^4 by default); the others fewer
    n_sweep = min(n_req, args.sweep_requests)
    for qps in sorted(set([args.qps] + rates)):
        r = run_at(qps, 17 + qps, n_req if qps == args.qps else n_sweep)
        sweep.append(r)
        log(f"[C4] qps {qps}: p50 {r['p50_ms']:.3f} ms p99 {r['p99_ms']:.3f} ms, "
            f"{r['pairs_per_s'] / 1e6:.2f} M pairs/s, {r['launches']} launches")
    main_r = next(r for r in sweep if r["qps"] == args.qps)
    achieved = k_bytes / (k_ms * 1e-3) / 1e9
    line = {
        "value": main_r["pairs_per_s"], "unit": "pairs/s", "steps": n_req, "warmup": 0,
        "ms_per_step": main_r["wall_s"] / n_req * 1e3,
        "config": {"workload": "serving: 20000-body scene (8-512 vertices log-uniform), Poisson "
                               f"requests at {args.qps} QPS, K ~ U{{1..10000}} pairs, fresh pose per "
                               "pair side, t ~ U[-10,10]^3", "requests": n_req,
                   "pairs": total_pairs, "lanes_per_pair": {"gjk": lg, "epa": le},
                   "server": {"slots": main_r["slots"], "max_batch_pairs": 65536}},
        "latency_ms": {"p50": main_r["p50_ms"], "p99": main_r["p99_ms"], "max": main_r["max_ms"],
                       "qps": args.qps, "target_p99_ms": 2.0},
        "qps_sweep": sweep,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": profile_value("serving", "traffic_bytes_per_step"),
                     "kernel": "k_gjk + k_epa_list over 10000 posed pairs (one launch)",
                     "kernel_ms_mean": k_ms, "algorithmic_bytes_per_launch": k_bytes,
                     "algorithmic_bytes": "compulsory: vertices once + pair, poses, body table, "
                                          "results"},
    }
    if not args.no_cpu:
        base, parity = serving_cpu_baseline(scene, b, reqs)
        line["cpu_baseline"] = base
        line["parity"] = parity
    return line


def serving_cpu_baseline(scene, b, reqs, max_pairs: int = 200_000):
    """The FP64 oracle on the world vertices of the first requests' pairs (up
    to max_pairs), timed on this host's cores; the server's FP64 records for
    the same requests are compared with it."""
    import oracle
    import ogjk
    nth = cores()
    take, n = [], 0
    for p, q in reqs:
        if n >= max_pairs:
            break
        take.append((p, q))
        n += len(p)
    pairs = np.concatenate([p for p, _ in take])
    poses = np.concatenate([q for _, q in take])
    pool, off, cnt, pa, pb = oracle.posed_pool(b.coords, b.off, b.cnt, pairs, poses)
    t = time.perf_counter()
    ref = oracle.run(3, pool, off, cnt, pa, pb, bits=64, nthreads=nth)
    el = time.perf_counter() - t
    srv = ogjk.Server(scene, max_batch_pairs=65536, slots=2)
    try:
        outs = [srv.wait(srv.submit(p, q, f64=True))[0] for p, q in take]
    finally:
        srv.close()
    rec = np.concatenate([o["f64"] for o in outs])
    d, dr = rec[:, 0], ref.distances
    exact = (d == dr) & (rec[:, 1:4] == ref.simplices["witnesses"][:, 0]).all(1) & \
            (rec[:, 4:7] == ref.simplices["witnesses"][:, 1]).all(1) & (rec[:, 7:10] == ref.normals).all(1)
    rel = np.abs(d - dr) / np.maximum(np.abs(dr), 1e-3)
    parity = {"pairs_checked": int(len(d)), "max_rel_err_distance": float(np.nanmax(rel)),
              "bit_exact_fraction": float(exact.mean()),
              "overlap_fraction": float((dr <= EPS64).mean())}
    base = {"value": len(d) / el, "unit": "pairs/s", "cores": nth, "kind": "port",
            "sample": f"pairs of the first {len(take)} requests ({len(d)} pairs), FP64 oracle on "
                      f"their FP64 world vertices, {nth} threads, {el:.2f}s (transform untimed)"}
    return base, parity


# ------------------------------------------------ legs either side of the path

def broadphase(args) -> dict:
    """Broad-phase leg: 1,000,000 spheres (centres U[-100,100]^3, radii
    U[0.2,1]) on a 100^3 grid of 2.0 cells; a step is one ogjk_broad_phase
    call (cell keys, counting sort, start table, count, scan, emit, total read
    back).  Throughput in objects/s."""
    import torch
    import oracle
    import ogjk
    from ogjk import gen
    n = args.pairs or 1_000_000
    box = 100.0 * (n / 1_000_000) ** (1.0 / 3.0)
    g = max(1, int(box))
    s = gen.spheres(n, seed=12, box=box, rmin=0.2, rmax=1.0)
    grid = ogjk.Grid(2.0 * box / g, box, g)
    d = torch.from_numpy(s).cuda()
    bp = ogjk.BroadPhase(n, grid, max_pairs=8 * n)
    pairs, found = bp(d)
    torch.cuda.synchronize()
    got = pairs.cpu().numpy()
    wall, ms = _event_steps(torch, lambda: bp(d), args.steps, args.warmup)
    ms_step = wall / args.steps * 1e3
    # algorithmic bytes: every sphere read once (16 B), every pair written once (8 B)
    alg = 16.0 * n + 8.0 * found
    achieved = alg / (np.mean(ms) * 1e-3) / 1e9
    line = {
        "metric": "broad-phase objects/s (uniform grid -> gkCollisionPair list)",
        "value": n / (wall / args.steps), "unit": "objects/s", "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True,
        "dtype": "f32",
        "config": {"workload": f"broad phase: {n} spheres, centres U[-{box:.1f},{box:.1f}]^3, radii "
                               f"U[0.2,1], grid {g}^3 cells of {grid.cell_size:.3f}",
                   "objects": n, "pairs_found": found},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": profile_value("broadphase", "traffic_bytes_per_step"),
                     "kernel": "whole broad phase (10 launches + total readback)",
                     "kernel_ms_mean": float(np.mean(ms)), "algorithmic_bytes_per_launch": alg,
                     "algorithmic_bytes": "16 B per sphere read + 8 B per pair written"},
    }
    if not args.no_cpu:
        t = time.perf_counter()
        ref, rfound = oracle.broad_phase(s, grid.cell_size, grid.boundary, grid.grid_size)
        el = time.perf_counter() - t
        line["cpu_baseline"] = {
            "value": n / el, "unit": "objects/s", "cores": 1, "kind": "port",
            "sample": f"the whole workload ({n} spheres), C oracle of the reference grid, one "
                      f"thread, {el:.2f}s"}
        line["parity"] = {"pairs_found": found, "oracle_found": rfound,
                          "identical_list": bool(rfound == found and np.array_equal(got, ref))}
    return line


def scene_step(args) -> dict:
    """Scene-step leg: the serving scene's 20,000 bodies (8-512 vertices,
    log-uniform ellipsoids in local frames), one pose per body (uniform unit
    quaternion, translation U[-40,40]^3), bounding spheres through the device
    broad phase (grid of at most 64^3 cells, cells one largest diameter
    wide), then GJK+EPA on every pair found, written as 40-byte contact
    records.  A step is both phases."""
    import torch
    import oracle
    import ogjk
    from ogjk import gen
    lg, le = (int(x) for x in args.lanes.split(",")) if args.lanes else (64, 64)
    ogjk.set_config(lg, le, PIPELINE)  # 64 / 64 for broad-phase pairs: measured best
    b = gen.make(gen.ELLIPSOID, 10_000, seed=4, box=0.0, nmin=8, nmax=512, nthreads=cores())
    nb = len(b.cnt)
    radii = gen.body_radii(b.coords, b.off, b.cnt)
    tbox = 40.0
    poses = gen.body_poses(nb, seed=7, tbox=tbox)
    grid = ogjk.Grid.for_spheres(float(radii.max()), tbox, 64)
    sc = ogjk.DeviceScene(b)
    step = ogjk.SceneStep(sc, radii, grid, max_pairs=20 * nb, contacts=True)
    dp = torch.from_numpy(poses).cuda()
    pairs, found, out = step(dp)
    torch.cuda.synchronize()
    got = pairs.cpu().numpy()
    n = len(got)
    wall, ms = _event_steps(torch, lambda: step(dp), args.steps, args.warmup)
    # narrow phase alone on the same pairs, for the kernel roofline
    ws = step.workspace
    nwall, nms = _event_steps(torch, lambda: sc.collide_posed_bodies_contacts(
        3, pairs, dp, out=out, workspace=ws), args.steps, args.warmup)
    cnt = b.cnt.astype(np.int64)
    # vertices once per pair side, pair 8 B, two poses 56 B, body table 32 B, contact record 40 B
    alg = float(12 * (cnt[got[:, 0]] + cnt[got[:, 1]]).sum() + n * (8 + 56 + 32 + 40))
    achieved = alg / (np.mean(nms) * 1e-3) / 1e9
    line = {
        "metric": "scene step pairs/s (broad phase + posed GJK+EPA on the pairs found)",
        "value": n / (wall / args.steps), "unit": "pairs/s", "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": wall / args.steps * 1e3, "higher_is_better": True,
        "dtype": "f32+f64",
        "config": {"workload": f"scene step: {nb} bodies (8-512 vertices log-uniform), one pose "
                               f"per body, t ~ U[-{tbox:.0f},{tbox:.0f}]^3, {grid}",
                   "bodies": nb, "pairs": n, "lanes_per_pair": {"gjk": lg, "epa": le}},
        "narrow_phase_ms": float(np.mean(nms)),
        "broad_phase_ms": float(np.mean(ms) - np.mean(nms)),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": profile_value("scene_step", "traffic_bytes_per_step"),
                     "kernel": f"k_gjk + k_epa_list over the {n} pairs of a step",
                     "kernel_ms_mean": float(np.mean(nms)), "algorithmic_bytes_per_launch": alg,
                     "algorithmic_bytes": "compulsory: vertices once + pair, poses, body "
                                          "table, 40-byte contact record"},
    }
    if not args.no_cpu:
        nth = cores()
        spheres = np.concatenate([poses[:, 4:7], radii[:, None]], 1)
        t = time.perf_counter()
        ref_pairs, rfound = oracle.broad_phase(spheres, grid.cell_size, grid.boundary,
                                               grid.grid_size)
        side = np.concatenate([poses[ref_pairs[:, 0]], poses[ref_pairs[:, 1]]], 1)
        t_bp = time.perf_counter() - t
        pool, off, cnt2, pa, pb = oracle.posed_pool(b.coords, b.off, b.cnt, ref_pairs, side)
        t = time.perf_counter()
        ref = oracle.run(3, pool, off, cnt2, pa, pb, bits=64, nthreads=nth)
        t_np = time.perf_counter() - t
        el = t_bp + t_np
        step_out = step.scene.collide_posed_bodies(3, pairs, dp, f64_record=True)
        torch.cuda.synchronize()
        rec = step_out["f64"][:n].cpu().numpy()
        exact = ((rec[:, 0] == ref.distances) &
                 (rec[:, 1:4] == ref.simplices["witnesses"][:, 0]).all(1) &
                 (rec[:, 4:7] == ref.simplices["witnesses"][:, 1]).all(1) &
                 (rec[:, 7:10] == ref.normals).all(1)) if len(ref_pairs) == n else np.zeros(1)
        line["cpu_baseline"] = {
            "value": rfound / el, "unit": "pairs/s", "cores": nth, "kind": "port",
            "sample": f"one whole step: C oracle broad phase (1 thread, {t_bp:.2f}s) + FP64 "
                      f"oracle GJK+EPA on the world vertices of the {rfound} pairs ({nth} "
                      f"threads, {t_np:.2f}s; world transform untimed)"}
        line["parity"] = {"pairs_found": found, "oracle_found": rfound,
                          "identical_pair_list": bool(rfound == found and
                                                      np.array_equal(got, ref_pairs)),
                          "bit_exact_fraction": float(exact.mean()),
                          "overlap_fraction": float((ref.distances <= EPS64).mean())}
    return line


def hulls(args) -> dict:
    """Mesh-ingestion leg: 100,000 point clouds of 8-32 points uniform on the
    unit sphere (the reference's random hull variants, MIN_HULL_VERTS ..
    MAX_HULL_VERTS); a step ingests them all as the reference does each shape
    (convex hull, centre of mass, centring, hull of the centred points and its
    inertia factors).  Throughput in clouds/s."""
    import torch
    import oracle
    import ogjk
    n = args.pairs or 100_000
    rng = np.random.default_rng(5)
    cnt = rng.integers(8, 33, n)
    u = rng.normal(size=(int(cnt.sum()), 3))
    u = (u / np.linalg.norm(u, axis=1, keepdims=True)).astype(np.float32)
    off = np.zeros(n, np.int64)
    off[1:] = np.cumsum(cnt)[:-1]
    pts = torch.from_numpy(u).cuda()
    doff = torch.from_numpy(off).cuda()
    dcnt = torch.from_numpy(cnt.astype(np.int32)).cuda()
    r = ogjk.ingest_hulls(pts, doff, dcnt)
    torch.cuda.synchronize()
    wall, ms = _event_steps(torch, lambda: ogjk.ingest_hulls(pts, doff, dcnt), args.steps,
                            args.warmup)
    # the hull kernel alone, for the roofline: points read once, faces written once
    hw, hms = _event_steps(torch, lambda: ogjk.build_hulls(pts, doff, dcnt), args.steps,
                           args.warmup)
    nf = r["nfaces"].cpu().numpy()
    alg = 12.0 * len(u) + 12.0 * float(nf.clip(min=0).sum()) + 16.0 * n
    achieved = alg / (np.mean(hms) * 1e-3) / 1e9
    line = {
        "metric": "mesh ingestion clouds/s (convex hull, centring, inertia factors)",
        "value": n / (wall / args.steps), "unit": "clouds/s", "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": wall / args.steps * 1e3, "higher_is_better": True,
        "dtype": "f32+f64",
        "config": {"workload": f"hull ingestion: {n} clouds of 8-32 points uniform on the unit "
                               "sphere", "clouds": n, "points": int(len(u)),
                   "faces": int(nf.clip(min=0).sum())},
        "hull_kernel_ms": float(np.mean(hms)),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": profile_value("hulls", "traffic_bytes_per_step"),
                     "kernel": "k_hull (one hull build of every cloud)",
                     "kernel_ms_mean": float(np.mean(hms)), "algorithmic_bytes_per_launch": alg,
                     "algorithmic_bytes": "12 B per point read, 12 B per face written, 16 B "
                                          "per cloud of offsets and counts"},
    }
    if not args.no_cpu:
        nth = cores()
        t = time.perf_counter()
        cen_o, faces_o, fo_o, nf_o, k_o = oracle.ingest_batch(u, off, cnt, nthreads=nth)
        el = time.perf_counter() - t
        faces = r["faces"].cpu().numpy()
        fo = r["face_off"].cpu().numpy()
        k = r["inertia_k"].cpu().numpy()
        cen = r["points"].cpu().numpy()
        same = [nf[i] == nf_o[i] and
                np.array_equal(faces[fo[i]:fo[i] + nf[i]], faces_o[fo_o[i]:fo_o[i] + nf_o[i]])
                for i in range(n)]
        same = np.array(same) & (k.view(np.uint32) == k_o.view(np.uint32)).all(1)
        cen_same = bool(np.array_equal(cen.view(np.uint32), cen_o.view(np.uint32)))
        line["cpu_baseline"] = {
            "value": n / el, "unit": "clouds/s", "cores": nth, "kind": "port",
            "sample": f"all {n} clouds, C oracle of the reference mesh builder (the same "
                      f"sequence), {nth} threads, {el:.2f}s"}
        line["parity"] = {"clouds_checked": n, "bit_exact_fraction": float(np.mean(same)),
                          "centred_points_identical": cen_same}
    return line


# ---------------------------------------------------------------- driver

def _common(rec: dict, args, dtype="f32+f64") -> dict:
    """Fields every printed line carries, in the contract's order."""
    head = {"metric": rec.pop("metric", METRIC), "value": rec.pop("value"),
            "unit": rec.pop("unit"), "n_gpus": 1, "steps": rec.pop("steps"),
            "warmup": rec.pop("warmup"), "ms_per_step": rec.pop("ms_per_step"),
            "higher_is_better": rec.pop("higher_is_better", True), "scaling": "weak",
            "vs_baseline": None, "dtype": rec.pop("dtype", dtype),
            "data": "synthetic (counter-based generator, fixed seeds)",
            "config": rec.pop("config")}
    head.update(rec)
    return head


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="all",
                    choices=["all"] + sorted(WORKLOADS) + ["serving", "broadphase", "scene_step",
                                                           "hulls"])
    ap.add_argument("--qps", type=int, default=1000, help="serving: request rate")
    ap.add_argument("--requests", type=int, default=10000,
                    help="serving: requests at the stated rate (--qps)")
    ap.add_argument("--sweep-requests", type=int, default=2000,
                    help="serving: requests at each other rate of the sweep")
    ap.add_argument("--slots", type=int, default=4,
                    help="serving: launches in flight (at most the process's hardware queues, "
                         "4 by default: more slots share queues and block behind long launches)")
    ap.add_argument("--pairs", type=int, default=0, help="override the pair count")
    ap.add_argument("--kernel-only", action="store_true",
                    help="serving: time the 10,000-pair posed launch only (for profiling)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline and parity")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain-FP64 comparison")
    ap.add_argument("--calibrate", action="store_true",
                    help="also run a 256 MiB elementwise copy (reads and writes 256 MiB) so a "
                         "profiler's FETCH_SIZE/WRITE_SIZE can be scaled against known bytes")
    ap.add_argument("--lanes", default=None,
                    help="GJK,EPA lanes per pair (a single batch workload, serving, scene_step)")
    ap.add_argument("--cluster-min", type=int, default=None,
                    help="cull bodies with at least this many vertices (0: no clusters; "
                         "default per workload)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported: pairs are "
              "independent and would shard by contiguous, cost-balanced ranges of the pair list, "
              "one process per GPU", file=sys.stderr)
        return 2

    import torch
    if not torch.cuda.is_available():
        print("bench.py: no GPU visible", file=sys.stderr)
        return 1
    import ogjk
    ogjk.lib()  # the HIP library must load; there is no other path
    if args.calibrate:
        x = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device="cuda")
        for _ in range(3):
            y = x * 1.0
        torch.cuda.synchronize()
        del x, y
    t_start = time.perf_counter()
    w = args.workload
    if w in ("serving", "broadphase", "scene_step", "hulls"):
        rec = {"serving": serving, "broadphase": broadphase, "scene_step": scene_step,
               "hulls": hulls}[w](args)
        if w == "serving":
            rec["metric"] = METRIC
        print(json.dumps(_common(rec, args)), flush=True)
        return 0
    if w != "all":
        print(json.dumps(_common(batch_leg(w, args, headline=True), args)), flush=True)
        return 0
    # the serving loop runs first, in the state a serving process starts in
    # (its slots' streams get hardware queues of their own)
    serving_rec = serving(args, headline=False)
    # the small batch workloads run before the headline: a latency-bound
    # 1,000-pair step measured ~30% slower right after the 1M-pair leg
    subs = {}
    for name in ("sphere32_gjk", "sphere64"):
        subs[f"{WORKLOADS[name]['cfg']}_{name}"] = batch_leg(name, args, headline=False)
    line = _common(batch_leg("ellipsoid_uneven", args, headline=True), args)
    line["headline"] = ("uneven 1M-pair workload (GJK then EPA); value = steady-state pairs/s "
                        "with the geometry resident; speedup_vs_plain_f64_one_shot charges the "
                        "cluster build to one step")
    name = "noisy_sphere_large"
    subs[f"{WORKLOADS[name]['cfg']}_{name}"] = batch_leg(name, args, headline=False)
    subs["C4_serving"] = serving_rec
    line["configs"] = subs
    line["bench_wall_s"] = time.perf_counter() - t_start
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
