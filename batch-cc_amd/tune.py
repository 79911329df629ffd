"""Autotune the launch table on the GPU and write csrc/bcc_tuning.h.

Three settings, in this order:

1. The culling threshold (one value for all robots): per robot, three
   benchmark-shaped batches (see shaped_batches) are timed with scenes built
   under each candidate threshold; the candidate with the lowest sum of
   time / best time wins.
2. Per robot, the kernel-variant range: the same batches are timed with
   the high- and the low-occupancy variant forced; the range of mean
   obstacle counts per environment in which a launch takes the high variant
   is the candidate with the lowest sum of time / best time when each batch
   runs the variant the range picks for it.
3. Per robot, the ragged launch shape: three seeded ragged batches (long
   pairs, few environments, short pairs: see batch()) are timed as replayed
   hipGraphs for every candidate (target tiles per wave, workgroups per CU);
   the candidate with the lowest mean of time / best time goes into the table.
4. Per robot, the step dimension of small batches' tiles: a batch with fewer
   pairs than the grid has half-waves and long edges (see steps_batch) is
   timed with every candidate step chunk, and without one; the fastest goes
   into the table (0: no step dimension).  `--only-steps` runs this stage
   alone and keeps the rest of the existing table.

Every timing run must reproduce the first run's verdicts.  The library reads
the table when a plan is created with tile_pairs = 0 / groups_per_cu = 0 and
when a scene or launch is set up.

    python batch-cc_amd/tune.py [--out batch-cc_amd/csrc/bcc_tuning.h] [--json timings.json] [--quick]

Rebuild the library (batch-cc_amd/build.py) after writing the table.
"""
from __future__ import annotations

import argparse
import datetime
import importlib.util
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
TILES_PER_WAVE = (4, 8, 16, 32, 64)
GROUPS = (4, 5, 6, 8)
ROBOTS = ("panda", "fetch", "baxter", "xarm7")
PAIRS = {"panda": 400_000, "fetch": 400_000, "baxter": 200_000, "xarm7": 200_000}


def load_pkg():
    spec = importlib.util.spec_from_file_location("batch_cc_amd", os.path.join(HERE, "__init__.py"),
                                                  submodule_search_locations=[HERE])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["batch_cc_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def batch(W, robot: str, kind: str, pairs: int, seed: int):
    """Ragged tuning batches.  "long": U{32..256} small obstacles away from the
    base, lognormal edges (about 40% colliding, free pairs run every step),
    U{16..48} pairs per environment; "few-envs": the same with about 2048 pairs
    per environment; "short": c2's shape (4096 environments of workspace-filling
    mixed obstacles, U{8..256}, uniform edges, Zipf(1.2) pairs per environment:
    nearly every pair collides in its first step)."""
    from batch_cc_amd import PackedScene
    rs = np.random.default_rng([seed, 11])
    if kind == "short":
        E = 4096 if pairs >= 100_000 else max(8, pairs // 200)   # c2's shape at full size
        counts = W.zipf_counts(rs, E)
        envs = [W.mixed_env(rs, int(rs.integers(8, 257)), capsule_frac=0.2) for _ in range(E)]
    else:
        mean = 32 if kind == "long" else 2048
        E = max(8, min(16384, pairs // mean))
        counts = rs.integers(max(1, mean // 2), max(2, 3 * mean // 2 + 1), E).astype(np.int64)
        envs = []
        for _ in range(E):
            n = int(rs.integers(32, 257))
            envs.append(W.scaled_env(rs, n, (32.0 / n) ** (1.0 / 3.0)))
    off = np.zeros(E + 1, np.int64)
    off[1:] = np.cumsum(counts)
    if kind == "short":
        edges = W.uniform_edges(robot, int(off[-1]), rs)
    else:
        edges = W.lognormal_edges(robot, int(off[-1]), rs, median=W.MIXED_EDGE_MEDIAN)
    return PackedScene.from_envs(envs), off, edges


CULL_CANDIDATES = (16, 24, 32, 40, 48, 64)
NEVER = 1 << 30


def shaped_batches(W, robot: str, seed: int, quick: bool):
    """Benchmark-shaped batches of one robot at reduced size, as
    (name, PackedScene, env_off or None (dense), edges, mean obstacles per
    environment weighted by pairs): the robot's dense or ragged benchmark
    workload (panda: 32 obstacles of c0's kinds and uniform edges; fetch:
    U{4..64} mixed obstacles and lognormal edges; baxter: c2's Zipf ragged
    batch), a few-obstacle dense batch (7 per environment, like the
    MotionBenchMaker shelves), and the robot's part of the mixed batch
    (U{32..512} obstacles)."""
    from batch_cc_amd import PackedScene
    rs = np.random.default_rng([seed, 13, ROBOTS.index(robot)])
    f = 0.25 if quick else 1.0
    out = []
    if robot == "baxter":
        _, sc, off, edges = W.make_ragged("c2", seed=seed, scale=0.25 * f)
        out.append(("c2-shaped", sc, off, edges))
    elif robot == "fetch":
        _, sc, edges = W.make_config("c1", seed=seed, scale=0.3 * f)
        out.append(("c1-shaped", sc, None, edges))
    else:
        _, sc, edges = W.make_config("c3", seed=seed, scale=0.1 * f)
        out.append(("c3-shaped", sc, None, edges))
    envs = [W.c0_env(rs, n_cub=4, n_sph=2, n_cyl=1) for _ in range(100)]
    edges = W.uniform_edges(robot, int(4000 * f), rs)
    out.append(("few-obstacles", PackedScene.from_envs(envs), None, edges))
    parts, _, _ = W.make_mixed(seed=seed, scale=0.1 * f)
    for r, sc, off, edges, _ in parts:
        if r == robot:
            out.append(("mixed-part", sc, off, edges))
    res = []
    for name, sc, off, edges in out:
        obs = sc.counts[:, [0, 1, 2, 4, 5]].sum(1).astype(np.float64)   # cylinders are not tested
        w = np.diff(off).astype(np.float64) if off is not None else np.ones(sc.num_envs)
        res.append((name, sc, off, edges, float((obs * w).sum() / max(w.sum(), 1.0))))
    return res


def time_batch(bcc, robot: str, sc, off, edges, iters: int = 5):
    """(ms, verdicts) of one batch on a freshly built scene (culling
    threshold and forced variant as currently set)."""
    with bcc.Scene(sc) as scene:
        de = bcc.DeviceBuffer.from_array(edges)
        if off is None:
            M = edges.shape[0]
            do = bcc.DeviceBuffer(M * sc.num_envs)
            ms = bcc.time_dense(robot, scene, de, M, 32, do, 2, iters)
            return ms, do.to_array(np.uint8, (M, sc.num_envs))
        do = bcc.DeviceBuffer(edges.shape[0])
        plan = bcc.RaggedPlan(robot, scene, off, de, do)
        ms = plan.time(1, iters)
        v = do.to_array(np.uint8, (edges.shape[0],))
        plan.close()
        return ms, v


def tune_cull(bcc, W, seed: int, quick: bool):
    """Culling threshold with the lowest sum of time / best time over every
    robot's shaped batches.  Returns (threshold, timings)."""
    score = {c: 0.0 for c in CULL_CANDIDATES}
    rows = []
    try:
        for robot in ROBOTS:
            for name, sc, off, edges, mean in shaped_batches(W, robot, seed, quick):
                ms_of, ref = {}, None
                for c in CULL_CANDIDATES:
                    bcc.debug_cull_min(c)
                    ms, v = time_batch(bcc, robot, sc, off, edges)
                    if ref is None:
                        ref = v
                    elif not np.array_equal(v, ref):
                        raise SystemExit(f"tune: culling threshold {c} changes verdicts ({robot} {name})")
                    ms_of[c] = ms
                    rows.append({"robot": robot, "batch": name, "cull_min": c, "ms": ms})
                best = min(ms_of.values())
                for c in CULL_CANDIDATES:
                    score[c] += ms_of[c] / best
                print(f"[tune] cull {robot} {name}: " + " ".join(f"{c}:{ms_of[c]:.3f}" for c in CULL_CANDIDATES),
                      flush=True)
    finally:
        bcc.debug_cull_min(-1)
    return min(score, key=score.get), rows


def variant_threshold(bcc, W, robot: str, seed: int, quick: bool, cull_min: int):
    """Range [lo, hi) of mean obstacle counts in which a launch of this robot
    takes the high-occupancy variant: the pair of candidates (0, midpoints
    between the batches' means, never) minimising the sum of time / best time over the shaped
    batches when each runs the variant the range picks (with culling, very
    many obstacles are cheap again and can favour the low-occupancy variant).
    Returns ((lo, hi), rows)."""
    rows, batches = [], []
    bcc.debug_cull_min(cull_min)
    try:
        for name, sc, off, edges, mean in shaped_batches(W, robot, seed, quick):
            t = {}
            ref = None
            for v in (0, 1):
                bcc.debug_force_variant(v)
                ms, got = time_batch(bcc, robot, sc, off, edges)
                if ref is None:
                    ref = got
                elif not np.array_equal(got, ref):
                    raise SystemExit(f"tune: kernel variants disagree for {robot} {name}")
                t[v] = ms
            bcc.debug_force_variant(-1)
            batches.append((mean, t[0], t[1]))
            rows.append({"robot": robot, "batch": name, "mean_obstacles": mean, "ms_low": t[0], "ms_high": t[1]})
            print(f"[tune] {robot} {name} ({mean:.1f} obstacles/env): low {t[0]:.3f} ms, high {t[1]:.3f} ms",
                  flush=True)
    finally:
        bcc.debug_force_variant(-1)
        bcc.debug_cull_min(-1)
    return pick_range(batches), rows


def pick_range(batches) -> tuple:
    """batches: (mean obstacles, ms low variant, ms high variant).  The range
    [lo, hi) of mean obstacle counts that runs the high variant, minimising
    the summed time / best time; bounds lie halfway between consecutive
    batch means (a bound at a batch's own mean would decide that batch by
    rounding), or are 0 / never; ties go to the larger bounds."""
    means = sorted(m for m, _, _ in batches)
    cands = sorted({0, NEVER} | {int(round((a + b) / 2)) for a, b in zip(means, means[1:])})

    def total(lo_thr, hi_thr):
        return sum((hi if lo_thr <= m < hi_thr else lo) / min(lo, hi) for m, lo, hi in batches)

    pairs = [(a, b) for a in cands for b in cands if a < b] + [(NEVER, NEVER)]
    return min(pairs, key=lambda p: (total(*p), -p[0], -p[1]))


def header(table: dict, source: str, thresholds: dict, cull_min: int) -> str:
    rid = {r: i for i, r in enumerate(ROBOTS)}
    lines = [
        "// Launch table: the culling threshold, per robot the kernel-variant",
        "// threshold and the ragged launch shape (persistent workgroups per CU and",
        "// the target number of tiles per wave, from which a plan sizes its tiles).",
        "// Written by batch-cc_amd/tune.py.",
        "#pragma once",
        "",
        "struct bcc_tune_entry {",
        "    int robot;            // BCC_PANDA / BCC_FETCH / BCC_BAXTER / BCC_XARM7",
        "    int tiles_per_wave;   // target tiles per persistent wave",
        "    int groups_per_cu;    // persistent workgroups per CU",
        "    int steps_per_tile;   // step chunk of small batches' tiles (0: no step dimension)",
        "};",
        "",
        "// Exact culling of distant obstacles: a sweep of N robot spheres culls when",
        "// N * (primitives in the environment) is at least this.",
        f"static const int kCullMinWork = {cull_min};",
        "",
        "// Kernel variant per launch: the high-occupancy variant when the mean number",
        "// of tested obstacles per environment (weighted by pairs) is at least the",
        "// first value and below the second, per robot (" + ", ".join(ROBOTS) + ").",
        f"static const int kHiVariantMinObstacles[{len(ROBOTS)}] = {{"
        + ", ".join(str(thresholds[r][0]) for r in ROBOTS) + "};",
        f"static const int kHiVariantMaxObstacles[{len(ROBOTS)}] = {{"
        + ", ".join(str(thresholds[r][1]) for r in ROBOTS) + "};",
        "",
        f"// source: {source}",
        "static const bcc_tune_entry kTuneTable[] = {",
    ]
    for robot in ROBOTS:
        per_wave, groups, note, steps = table[robot]
        note = f"score {note:.3f}" if isinstance(note, float) else note
        lines.append(f"    {{{rid[robot]}, {per_wave}, {groups}, {steps}}},   // {robot}: {note}")
    lines.append("};")
    return "\n".join(lines) + "\n"


STEP_CANDIDATES = (-1, 1, 2, 4, 8, 16)


def steps_batch(W, robot: str, seed: int):
    """Few pairs, long runs: 64 environments of 16-48 mixed obstacles with
    2..16 pairs each (about 600 pairs, fewer than the GPU's half-waves), edges
    uniform in the joint limits, checked at resolution 128 (n ~ 20-40)."""
    from batch_cc_amd import PackedScene
    rs = np.random.default_rng([seed, 17, ROBOTS.index(robot)])
    E = 64
    counts = rs.integers(2, 17, E).astype(np.int64)
    envs = [W.mixed_env(rs, int(rs.integers(16, 49)), capsule_frac=0.2) for _ in range(E)]
    off = np.zeros(E + 1, np.int64)
    off[1:] = np.cumsum(counts)
    return PackedScene.from_envs(envs), off, W.uniform_edges(robot, int(off[-1]), rs), 128


def tune_steps(bcc, W, robot: str, seed: int):
    """Steps per tile for small batches: the fastest candidate (0 = none)."""
    sc, off, edges, res = steps_batch(W, robot, seed)
    n = edges.shape[0]
    ms_of, ref, rows = {}, None, []
    with bcc.Scene(sc) as scene:
        de = bcc.DeviceBuffer.from_array(edges)
        do = bcc.DeviceBuffer(n)
        for c in STEP_CANDIDATES:
            plan = bcc.RaggedPlan(robot, scene, off, de, do, resolution=res, tile_steps=c)
            ms = plan.time(2, 10)
            v = do.to_array(np.uint8, (n,))
            plan.close()
            if ref is None:
                ref = v
            elif not np.array_equal(v, ref):
                raise SystemExit(f"tune: verdicts differ for {robot} steps {c}")
            ms_of[c] = ms
            rows.append({"robot": robot, "batch": "steps", "pairs": n, "tile_steps": c, "ms": ms})
    best = min(ms_of, key=ms_of.get)
    print(f"[tune] {robot} steps: {n} pairs at resolution {res}: " +
          ", ".join(f"{c}: {ms_of[c]:.3f} ms" for c in STEP_CANDIDATES) + f" -> {max(best, 0)}", flush=True)
    return max(best, 0), rows


def read_table(path: str):
    """The current table: culling threshold, variant ranges, launch shapes."""
    import re
    text = open(path).read()
    cull = int(re.search(r"kCullMinWork = (\d+);", text).group(1))
    lo = [int(x) for x in re.search(r"kHiVariantMinObstacles\[\d+\] = \{([^}]*)\}", text).group(1).split(",")]
    hi = [int(x) for x in re.search(r"kHiVariantMaxObstacles\[\d+\] = \{([^}]*)\}", text).group(1).split(",")]
    table = {}
    for m in re.finditer(r"\{(\d+), (\d+), (\d+)(?:, (\d+))?\},\s*// (\w+): (.*)", text):
        table[m.group(5)] = [int(m.group(2)), int(m.group(3)), m.group(6), int(m.group(4) or 0)]
    thresholds = {r: (lo[i], hi[i]) for i, r in enumerate(ROBOTS)}
    return cull, thresholds, table


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=os.path.join(HERE, "csrc", "bcc_tuning.h"))
    ap.add_argument("--json", default=None, help="write every timing here")
    ap.add_argument("--quick", action="store_true", help="fewer candidates and smaller batches")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--only-steps", action="store_true", help="tune the step dimension only; keep the rest")
    a = ap.parse_args()
    bcc = load_pkg()
    from batch_cc_amd import workloads as W
    if a.only_steps:
        cull_min, thresholds, old = read_table(a.out)
        table, timings = {}, []
        for robot in ROBOTS:
            steps, rows = tune_steps(bcc, W, robot, a.seed)
            timings += rows
            per_wave, groups, note, _ = old[robot]
            table[robot] = (per_wave, groups, note, steps)
        src = f"tune.py on gfx950, {datetime.date.today().isoformat()}, seed {a.seed} (step dimension; launch shapes kept)"
        with open(a.out, "w") as f:
            f.write(header(table, src, thresholds, cull_min))
        if a.json:
            with open(a.json, "w") as f:
                json.dump(timings, f)
        print(f"[tune] wrote {a.out}")
        return
    per_waves = (8, 16, 32) if a.quick else TILES_PER_WAVE
    groups = (4, 6) if a.quick else GROUPS
    scale = 0.25 if a.quick else 1.0
    table, timings = {}, []
    cull_min, rows = tune_cull(bcc, W, a.seed, a.quick)
    timings += rows
    print(f"[tune] culling threshold {cull_min}", flush=True)
    thresholds = {}
    for robot in ROBOTS:
        thresholds[robot], rows = variant_threshold(bcc, W, robot, a.seed, a.quick, cull_min)
        timings += rows
        print(f"[tune] {robot}: high-occupancy variant for {thresholds[robot][0]} <= obstacles per environment "
              f"< {thresholds[robot][1]}", flush=True)
    for robot in ROBOTS:
        scores = {}
        for kind in ("long", "few-envs", "short"):
            t0 = time.perf_counter()
            sc, off, edges = batch(W, robot, kind, int(PAIRS[robot] * scale), a.seed)
            n = edges.shape[0]
            ms_of, ref = {}, None
            # the variant the new threshold picks for this batch, and the new culling threshold
            obs = sc.counts[:, [0, 1, 2, 4, 5]].sum(1).astype(np.float64)
            w = np.diff(off).astype(np.float64)
            mean = float((obs * w).sum() / max(w.sum(), 1.0))
            bcc.debug_force_variant(1 if thresholds[robot][0] <= mean < thresholds[robot][1] else 0)
            bcc.debug_cull_min(cull_min)
            with bcc.Scene(sc) as scene:
                de = bcc.DeviceBuffer.from_array(edges)
                do = bcc.DeviceBuffer(n)
                for T in per_waves:
                    for g in groups:
                        plan = bcc.RaggedPlan(robot, scene, off, de, do, tile_pairs=-T, groups_per_cu=g)
                        tile = plan.config[0]
                        ms = plan.time(2, 5)
                        v = do.to_array(np.uint8, (n,))
                        plan.close()
                        if ref is None:
                            ref = v
                        elif not np.array_equal(v, ref):
                            raise SystemExit(f"tune: verdicts differ for {robot} {kind} T {T} groups {g}")
                        ms_of[(T, g)] = ms
                        timings.append({"robot": robot, "batch": kind, "pairs": n, "tiles_per_wave": T,
                                        "tile": tile, "groups_per_cu": g, "ms": ms})
            bcc.debug_force_variant(-1)
            bcc.debug_cull_min(-1)
            best = min(ms_of.values())
            for k, v in ms_of.items():
                scores[k] = scores.get(k, 0.0) + v / best / 3.0
            kb = min(ms_of, key=ms_of.get)
            print(f"[tune] {robot} {kind}: {n} pairs, {sc.num_envs} envs, best tiles/wave {kb[0]} groups {kb[1]} "
                  f"{best:.3f} ms ({time.perf_counter() - t0:.1f} s)", flush=True)
        k = min(scores, key=scores.get)
        steps, rows = tune_steps(bcc, W, robot, a.seed)
        timings += rows
        table[robot] = (k[0], k[1], scores[k], steps)
        print(f"[tune] {robot}: tiles per wave {k[0]}, groups per CU {k[1]} (mean time / best {scores[k]:.3f})",
              flush=True)
    props = bcc.device_name() if hasattr(bcc, "device_name") else "gfx950"
    src = f"tune.py on {props}, {datetime.date.today().isoformat()}, seed {a.seed}" + (" (quick)" if a.quick else "")
    with open(a.out, "w") as f:
        f.write(header(table, src, thresholds, cull_min))
    if a.json:
        with open(a.json, "w") as f:
            json.dump(timings, f)
    print(f"[tune] wrote {a.out}")


if __name__ == "__main__":
    main()
