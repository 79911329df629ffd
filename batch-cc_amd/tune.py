"""Autotune the ragged-batch launch table on the GPU and write csrc/bcc_tuning.h.

Then, per robot, dense batches with 4 .. 128 obstacles per environment (the
benchmark scenes' mixed obstacles spread over the workspace) time
the two compiled kernel variants (high and low occupancy) against each other;
the obstacle count from which the high-occupancy variant wins becomes the
robot's variant threshold.

For every robot, three seeded ragged batches (long pairs, few environments,
short pairs: see batch()) are timed as replayed hipGraphs for every
candidate (target tiles per wave, workgroups per CU); verdicts must be
identical for every candidate.  The candidate with the lowest mean of
time / best time over the three batches goes into the table; the library
reads it when a plan is created with tile_pairs = 0 / groups_per_cu = 0.

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
ROBOTS = ("panda", "fetch", "baxter")
PAIRS = {"panda": 400_000, "fetch": 400_000, "baxter": 200_000}


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


OBSTACLE_COUNTS = (4, 8, 16, 32, 64, 128)


def variant_threshold(bcc, W, robot: str, seed: int, quick: bool):
    """Smallest mean obstacle count from which the high-occupancy kernel
    variant is at least as fast as the low one on dense batches of the
    benchmark scenes' kind: for each count, 256 environments of that many mixed
    obstacles spread over the workspace, once with 4000 lognormal edges (as in
    c1) and once with 2000 uniform edges (as in c0/c3) -- large enough that the
    launch tail does not decide; the two times add up.
    Returns (threshold, timings)."""
    from batch_cc_amd import PackedScene
    rs = np.random.default_rng([seed, 12])
    rows, thr = [], None
    for n_obs in OBSTACLE_COUNTS:
        envs = [W.mixed_env(rs, n_obs) for _ in range(256)]
        sc = PackedScene.from_envs(envs)
        ms = {0: 0.0, 1: 0.0}
        for edges in (W.lognormal_edges(robot, 1000 if quick else 4000, rs),
                      W.uniform_edges(robot, 500 if quick else 2000, rs)):
            with bcc.Scene(sc) as scene:
                de = bcc.DeviceBuffer.from_array(edges)
                do = bcc.DeviceBuffer(edges.shape[0] * sc.num_envs)
                ref = None
                for v in (0, 1):
                    bcc.debug_force_variant(v)
                    ms[v] += bcc.time_dense(robot, scene, de, edges.shape[0], 32, do, 2, 5)
                    got = do.to_array(np.uint8, (edges.shape[0], sc.num_envs))
                    if ref is None:
                        ref = got
                    elif not np.array_equal(got, ref):
                        bcc.debug_force_variant(-1)
                        raise SystemExit(f"tune: kernel variants disagree for {robot} at {n_obs} obstacles")
            bcc.debug_force_variant(-1)
        rows.append({"robot": robot, "obstacles": n_obs, "ms_low": ms[0], "ms_high": ms[1]})
        print(f"[tune] {robot} {n_obs} obstacles/env: low {ms[0]:.3f} ms, high {ms[1]:.3f} ms", flush=True)
    # the first count from which high is faster, or within the 2% timing noise,
    # at every larger count (ties go to the high-occupancy variant, which wins
    # on the full-size benchmark batches)
    for i, n_obs in enumerate(OBSTACLE_COUNTS):
        if all(r["ms_high"] <= 1.02 * r["ms_low"] for r in rows[i:]):
            thr = n_obs
            break
    return (thr if thr is not None else 1 << 30), rows


def header(table: dict, source: str, thresholds: dict) -> str:
    rid = {r: i for i, r in enumerate(ROBOTS)}
    lines = [
        "// Ragged-batch launch table: per robot, the persistent grid (workgroups per",
        "// CU) and the target number of tiles per wave, from which a plan sizes its",
        "// tiles; and the kernel-variant thresholds.  Written by batch-cc_amd/tune.py.",
        "#pragma once",
        "",
        "struct bcc_tune_entry {",
        "    int robot;            // BCC_PANDA / BCC_FETCH / BCC_BAXTER",
        "    int tiles_per_wave;   // target tiles per persistent wave",
        "    int groups_per_cu;    // persistent workgroups per CU",
        "};",
        "",
        "// Kernel variant per launch: the high-occupancy variant when the mean number",
        "// of tested obstacles per environment (weighted by pairs) is at least this,",
        "// per robot (panda, fetch, baxter).",
        "static const int kHiVariantMinObstacles[3] = {"
        + ", ".join(str(thresholds[r]) for r in ROBOTS) + "};",
        "",
        f"// source: {source}",
        "static const bcc_tune_entry kTuneTable[] = {",
    ]
    for robot in ROBOTS:
        per_wave, groups, score = table[robot]
        lines.append(f"    {{{rid[robot]}, {per_wave}, {groups}}},   // {robot}: score {score:.3f}")
    lines.append("};")
    return "\n".join(lines) + "\n"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=os.path.join(HERE, "csrc", "bcc_tuning.h"))
    ap.add_argument("--json", default=None, help="write every timing here")
    ap.add_argument("--quick", action="store_true", help="fewer candidates and smaller batches")
    ap.add_argument("--seed", type=int, default=0)
    a = ap.parse_args()
    bcc = load_pkg()
    from batch_cc_amd import workloads as W
    per_waves = (8, 16, 32) if a.quick else TILES_PER_WAVE
    groups = (4, 6) if a.quick else GROUPS
    scale = 0.25 if a.quick else 1.0
    table, timings = {}, []
    for robot in ROBOTS:
        scores = {}
        for kind in ("long", "few-envs", "short"):
            t0 = time.perf_counter()
            sc, off, edges = batch(W, robot, kind, int(PAIRS[robot] * scale), a.seed)
            n = edges.shape[0]
            ms_of, ref = {}, None
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
            best = min(ms_of.values())
            for k, v in ms_of.items():
                scores[k] = scores.get(k, 0.0) + v / best / 3.0
            kb = min(ms_of, key=ms_of.get)
            print(f"[tune] {robot} {kind}: {n} pairs, {sc.num_envs} envs, best tiles/wave {kb[0]} groups {kb[1]} "
                  f"{best:.3f} ms ({time.perf_counter() - t0:.1f} s)", flush=True)
        k = min(scores, key=scores.get)
        table[robot] = (k[0], k[1], scores[k])
        print(f"[tune] {robot}: tiles per wave {k[0]}, groups per CU {k[1]} (mean time / best {scores[k]:.3f})",
              flush=True)
    thresholds = {}
    for robot in ROBOTS:
        thresholds[robot], rows = variant_threshold(bcc, W, robot, a.seed, a.quick)
        timings += rows
        print(f"[tune] {robot}: high-occupancy variant from {thresholds[robot]} obstacles per environment",
              flush=True)
    props = bcc.device_name() if hasattr(bcc, "device_name") else "gfx950"
    src = f"tune.py on {props}, {datetime.date.today().isoformat()}, seed {a.seed}" + (" (quick)" if a.quick else "")
    with open(a.out, "w") as f:
        f.write(header(table, src, thresholds))
    if a.json:
        with open(a.json, "w") as f:
            json.dump(timings, f)
    print(f"[tune] wrote {a.out}")


if __name__ == "__main__":
    main()
