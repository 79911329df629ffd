"""Benchmark: edge-env checks/sec of the plain batch edge collision checker on
one MI355X, with its roofline and the CPU baseline beside it.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config c1] [--seed 0]

A step is one launch of the dense kernel over the whole synthetic workload
(every edge against every environment), inputs resident in HBM.  Timing is
hipEvents on the launch stream around exactly K launches after W untimed
ones.  The roofline's algorithmic work is the reference's own source-level
FP32 operation count for the same pairs in plain order, counted on the GPU by
the reference-order replay kernel (bcc_work_dense).  The CPU baseline is the
project's C restatement of the reference checker (the reference's CPU path,
vamp, is not available), timed on the host cores on a bounded sample.
Prints one JSON line.
"""
from __future__ import annotations

import argparse
import importlib.util
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
FP32_PEAK_TFLOPS = 157.3     # MI355X FP32 vector peak (datasheet)
BF16_PEAK_TFLOPS = 2500.0    # MI355X dense BF16 MFMA peak (datasheet)
DIM = {"panda": 7, "fetch": 8, "baxter": 14}


def mlp_flops(robot: str) -> int:
    """Algorithmic flops of one inference: 3 layers over the unpadded input
    (64-cell encoding + start and end)."""
    return 2 * ((64 + 2 * DIM[robot]) * 256 + 256 * 256 + 256 * 32)
HBM_PEAK_GBS = 8000.0
METRIC = "edge-env checks/sec at bit-exact verdicts, learned-ordered vs plain, 1 GPU"
# HBM bytes per launch of the dense kernel, measured with rocprofv3 FETCH_SIZE
# (x2.000) + WRITE_SIZE (x1.000), the scales read off a plain copy kernel;
# commands and raw counts in profiles/dense_kernel_fetch_c1.md.  Only the
# default workload is measured; re-measure after kernel changes.
MEASURED_TRAFFIC = {"c1": (2 * 3.156683e6 * 1024 + 8.665456e6 * 1024,
                           "rocprofv3 FETCH_SIZE + WRITE_SIZE, profiles/dense_kernel_fetch_c1.md")}


def load_pkg():
    pkg_dir = os.path.join(REPO, "batch-cc_amd")
    spec = importlib.util.spec_from_file_location("batch_cc_amd", os.path.join(pkg_dir, "__init__.py"),
                                                  submodule_search_locations=[pkg_dir])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["batch_cc_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def cpu_baseline(robot, sc, edges, seconds=10.0, max_threads=16):
    """Time the oracle (CPU restatement) on a seeded random sample of pairs."""
    sys.path.insert(0, os.path.join(REPO, "tests"))
    import oracle_lib
    orc = oracle_lib.Oracle()
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    threads = max(1, min(cores, max_threads))
    E, M = sc.num_envs, edges.shape[0]
    rid = {"panda": 0, "fetch": 1, "baxter": 2}[robot]
    rng = np.random.default_rng(1234)
    n = 2000
    while True:
        pe = rng.integers(0, E, n).astype(np.int32)
        pg = rng.integers(0, M, n).astype(np.int32)
        t0 = time.perf_counter()
        orc.check_pairs(rid, sc, edges, pe, pg, threads=threads)
        dt = time.perf_counter() - t0
        if dt >= seconds or n >= E * M:
            break
        n = int(min(E * M, max(n * 2, n * seconds / max(dt, 1e-3) * 1.1)))
    return {
        "value": n / dt,
        "unit": "edge-env checks/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{n} random (env, edge) pairs of the same workload, {dt:.1f} s, "
                  f"C restatement of the reference checker (vamp unavailable), {threads} threads",
    }


def cpu_baseline_pairs(robot, sc, edges, pair_env, seconds=10.0, max_threads=16):
    """CPU baseline on a ragged batch: a seeded random sample of its pairs."""
    sys.path.insert(0, os.path.join(REPO, "tests"))
    import oracle_lib
    orc = oracle_lib.Oracle()
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    threads = max(1, min(cores, max_threads))
    total = edges.shape[0]
    rid = {"panda": 0, "fetch": 1, "baxter": 2}[robot]
    rng = np.random.default_rng(1234)
    n = 500
    while True:
        idx = rng.integers(0, total, n).astype(np.int32)
        t0 = time.perf_counter()
        orc.check_pairs(rid, sc, edges, pair_env[idx], idx, threads=threads)
        dt = time.perf_counter() - t0
        if dt >= seconds or n >= total:
            break
        n = int(min(total, max(n * 2, n * seconds / max(dt, 1e-3) * 1.1)))
    return {"value": n / dt, "unit": "edge-env checks/s", "cores": threads, "kind": "port",
            "sample": f"{n} random pairs of the ragged batch, {dt:.1f} s, C restatement of the reference "
                      f"checker (vamp unavailable), {threads} threads"}


def learned_leg(bcc, W, a, robot, scene, sc, d_edges, E, M, plain_verdicts, plain_ms):
    """Train the step-ordering model on labels the exact checker writes for
    seeded (env, edge) pairs drawn with a separate edge stream, then time one
    learned step = model inference for every pair + the ordered check."""
    rng = np.random.default_rng([a.seed, 3])
    n_train_edges = max(1, min(M, 100_000))
    train_edges = W.uniform_edges(robot, n_train_edges, rng) if robot != "fetch" else \
        W.lognormal_edges(robot, n_train_edges, rng)
    d_train = bcc.DeviceBuffer.from_array(train_edges)
    n = int(a.train_pairs)
    pe = rng.integers(0, E, n).astype(np.int32)
    pg = rng.integers(0, n_train_edges, n).astype(np.int32)
    t0 = time.perf_counter()
    d_labels = bcc.label_pairs(robot, scene, d_train, pe, pg)
    t_label = time.perf_counter() - t0
    model = bcc.Model(robot, seed=a.seed)
    t0 = time.perf_counter()
    d_x = model.features(scene, d_train, pe, pg)
    losses = model.train(d_x, d_labels, n, epochs=a.train_epochs, batch=8192, lr=2e-3)
    bcc.synchronize()
    t_train = time.perf_counter() - t0
    print(f"[bench] learned: {n} labels in {t_label:.1f} s, {a.train_epochs} epochs in {t_train:.1f} s, "
          f"loss {losses[0]:.3f} -> {losses[-1]:.3f}", file=sys.stderr, flush=True)
    del d_x
    d_out = bcc.DeviceBuffer(E * M)
    # one learned step: order every pair, then the ordered check
    for _ in range(max(1, a.warmup)):
        d_order = model.order_dense(scene, d_edges, M)
        bcc.check_dense_ex(robot, scene, d_edges, M, 32, d_out, d_order)
    bcc.synchronize()
    t_order = []
    for _ in range(a.steps):
        t0 = time.perf_counter()
        d_order = model.order_dense(scene, d_edges, M)
        bcc.synchronize()
        t_order.append((time.perf_counter() - t0) * 1e3)
    order_ms = float(np.median(t_order))
    check_ms = bcc.time_dense_ex(robot, scene, d_edges, M, 32, d_out, d_order, a.warmup, a.steps)
    got = d_out.to_array(np.uint8, (M, E))
    mismatches = int((got != plain_verdicts).sum())
    # rounds per pair, plain vs learned
    d_rounds = bcc.DeviceBuffer(E * M)
    bcc.check_dense_ex(robot, scene, d_edges, M, 32, d_out, None, d_rounds)
    r_plain = d_rounds.to_array(np.uint8, (M, E)).astype(np.float64)
    bcc.check_dense_ex(robot, scene, d_edges, M, 32, d_out, d_order, d_rounds)
    r_learn = d_rounds.to_array(np.uint8, (M, E)).astype(np.float64)
    coll = plain_verdicts == 1
    print(f"[bench] learned step: order {order_ms:.2f} ms + ordered check {check_ms:.3f} ms "
          f"(plain check {plain_ms:.3f} ms); verdict mismatches {mismatches}", file=sys.stderr, flush=True)
    if mismatches:
        print("[bench] learned-ordered verdicts differ from plain order", file=sys.stderr)
        sys.exit(1)
    mfma = mlp_flops(robot) * E * M / (order_ms / 1e3) / 1e12
    return {
        "ms_per_step": order_ms + check_ms,
        "order_ms": order_ms,
        "ordered_check_ms": check_ms,
        "inference_mfma": {"achieved": mfma, "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                           "frac": mfma / BF16_PEAK_TFLOPS, "flops_per_pair": mlp_flops(robot),
                           "note": "wall time of the inference call (fragment packing, encoder, fused kernel)"},
        "speedup_vs_plain": plain_ms / (order_ms + check_ms),
        "check_only_speedup_vs_plain": plain_ms / check_ms,
        "verdict_mismatches_vs_plain": mismatches,
        "mean_rounds_colliding_plain": float(r_plain[coll].mean()) if coll.any() else None,
        "mean_rounds_colliding_learned": float(r_learn[coll].mean()) if coll.any() else None,
        "mean_rounds_all_plain": float(r_plain.mean()),
        "mean_rounds_all_learned": float(r_learn.mean()),
        "train": {"pairs": n, "epochs": a.train_epochs, "label_s": t_label, "train_s": t_train,
                  "loss_first_epoch": float(losses[0]), "loss_last_epoch": float(losses[-1]),
                  "edges": "separate seeded edge stream, same environments"},
    }


def learned_mixed(bcc, W, a, runs, plain_ms) -> dict:
    """Learned leg of the mixed-robot batch: one model per robot, trained on
    labels the exact checker writes for seeded (env, edge) pairs of that
    robot's environments with a separate lognormal edge stream (the batch's
    edge law); one learned step = inference for every pair of every robot
    (bcc_model_order_pairs) + one replay of a group of ordered plans."""
    n = int(a.train_pairs)
    for k, r in enumerate(runs):
        rng = np.random.default_rng([a.seed, 3, k])
        E = r["sc"].num_envs
        n_edges = max(1, min(r["edges"].shape[0], 100_000))
        train = W.lognormal_edges(r["robot"], n_edges, rng, median=W.MIXED_EDGE_MEDIAN)
        d_train = bcc.DeviceBuffer.from_array(train)
        pe = rng.integers(0, E, n).astype(np.int32)
        pg = rng.integers(0, n_edges, n).astype(np.int32)
        t0 = time.perf_counter()
        d_labels = bcc.label_pairs(r["robot"], r["scene"], d_train, pe, pg)
        model = bcc.Model(r["robot"], seed=a.seed)
        d_x = model.features(r["scene"], d_train, pe, pg)
        losses = model.train(d_x, d_labels, n, epochs=a.train_epochs, batch=8192, lr=2e-3)
        bcc.synchronize()
        del d_x, d_labels, d_train
        r["model"], r["losses"] = model, losses
        r["pair_env"] = np.repeat(np.arange(E, dtype=np.int32), np.diff(r["off"]))
        print(f"[bench] learned {r['robot']}: {n} labels, {a.train_epochs} epochs in "
              f"{time.perf_counter() - t0:.1f} s, loss {losses[0]:.3f} -> {losses[-1]:.3f}",
              file=sys.stderr, flush=True)
    # inference time per robot (order every pair of the part)
    order_ms = {}
    for r in runs:
        d_env = bcc.DeviceBuffer.from_array(r["pair_env"])
        r["d_order"] = r["model"].order_pairs(r["scene"], r["d_edges"], d_env)
        ts = []
        for _ in range(max(3, a.steps // 4)):
            t0 = time.perf_counter()
            r["model"].order_pairs(r["scene"], r["d_edges"], d_env, out=r["d_order"])
            ts.append((time.perf_counter() - t0) * 1e3)
        order_ms[r["robot"]] = float(np.median(ts))
    for r in runs:
        r["d_out_l"] = bcc.DeviceBuffer(max(1, r["edges"].shape[0]))
        r["plan_l"] = bcc.RaggedPlan(r["robot"], r["scene"], r["off"], r["d_edges"], r["d_out_l"],
                                     d_order=r["d_order"])
    group = bcc.RaggedGroup([r["plan_l"] for r in runs])
    check_ms = group.time(a.warmup, a.steps)
    mismatches = 0
    for r in runs:
        mismatches += int((r["d_out_l"].to_array(np.uint8, (r["edges"].shape[0],)) != r["v"]).sum())
    group.close()
    for r in runs:
        r["plan_l"].close()
        r["model"].close()
    total_order = sum(order_ms.values())
    print(f"[bench] learned step: order {total_order:.2f} ms ("
          + ", ".join(f"{k} {v:.2f}" for k, v in order_ms.items())
          + f") + ordered group {check_ms:.3f} ms (plain group {plain_ms:.3f} ms); verdict mismatches "
          f"{mismatches}", file=sys.stderr, flush=True)
    if mismatches:
        print("[bench] learned-ordered verdicts differ from plain order", file=sys.stderr)
        sys.exit(1)
    flops = sum(mlp_flops(r["robot"]) * r["edges"].shape[0] for r in runs)
    mfma = flops / (total_order / 1e3) / 1e12
    return {
        "ms_per_step": total_order + check_ms,
        "order_ms": total_order,
        "order_ms_per_robot": order_ms,
        "inference_mfma": {"achieved": mfma, "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                           "frac": mfma / BF16_PEAK_TFLOPS,
                           "note": "wall time of the inference calls (fragment packing, encoder, fused kernel)"},
        "ordered_check_ms": check_ms,
        "speedup_vs_plain": plain_ms / (total_order + check_ms),
        "check_only_speedup_vs_plain": plain_ms / check_ms,
        "verdict_mismatches_vs_plain": mismatches,
        "train": {"pairs_per_robot": n, "epochs": a.train_epochs,
                  "loss_first_epoch": {r["robot"]: float(r["losses"][0]) for r in runs},
                  "loss_last_epoch": {r["robot"]: float(r["losses"][-1]) for r in runs},
                  "edges": "separate seeded lognormal edge stream per robot, same environments"},
    }


def bench_mixed(bcc, W, a) -> dict:
    """c4: panda, fetch and baxter environments interleaved; one ragged plan
    per robot, all replayed as one hipGraph (bcc_ragged_group)."""
    t0 = time.perf_counter()
    parts, env_robot, global_off = W.make_mixed(seed=a.seed, scale=a.scale)
    total = int(global_off[-1])
    print(f"[bench] workload c4: {len(env_robot)} interleaved envs, {total} pairs "
          f"(generated in {time.perf_counter() - t0:.1f} s)", file=sys.stderr, flush=True)
    runs = []
    for robot, sc, off, edges, gidx in parts:
        scene = bcc.Scene(sc)
        d_edges = bcc.DeviceBuffer.from_array(edges)
        d_out = bcc.DeviceBuffer(max(1, edges.shape[0]))
        d_out.zero()
        plan = bcc.RaggedPlan(robot, scene, off, d_edges, d_out)
        runs.append(dict(robot=robot, sc=sc, off=off, edges=edges, gidx=gidx, scene=scene, d_edges=d_edges,
                         d_out=d_out, plan=plan))
    group = bcc.RaggedGroup([r["plan"] for r in runs])
    group.launch()
    bcc.synchronize()
    verdicts = np.zeros(total, np.uint8)
    for r in runs:
        r["v"] = r["d_out"].to_array(np.uint8, (r["edges"].shape[0],))
        verdicts[r["gidx"]] = r["v"]
    coll = float(verdicts.mean())
    ms = group.time(a.warmup, a.steps)
    alone = {r["robot"]: r["plan"].time(1, max(1, a.steps // 2)) for r in runs}
    for r in runs:
        if not np.array_equal(r["d_out"].to_array(np.uint8, (r["edges"].shape[0],)), r["v"]):
            print("[bench] verdicts changed between launches", file=sys.stderr)
            sys.exit(1)
    print(f"[bench] {a.steps} timed group launches: {ms:.3f} ms each; plans alone: "
          + ", ".join(f"{k} {v:.3f} ms" for k, v in alone.items()), file=sys.stderr, flush=True)
    learned = learned_mixed(bcc, W, a, runs, ms) if a.learned else None
    roofline = None
    if not a.no_work:
        flops = trigs = 0.0
        for r in runs:
            pair_env = np.repeat(np.arange(r["sc"].num_envs, dtype=np.int32), np.diff(r["off"]))
            work, trig, wv = bcc.work_pairs(r["robot"], r["scene"], r["d_edges"], pair_env, 32)
            if not np.array_equal(wv, r["v"]):
                print("[bench] work-counter verdicts differ from the ragged kernel", file=sys.stderr)
                sys.exit(1)
            flops += float(work.sum(dtype=np.float64))
            trigs += float(trig.sum(dtype=np.float64))
        achieved = flops / (ms / 1e3) / 1e12
        roofline = {"bound": "valu", "achieved": achieved, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                    "frac": achieved / FP32_PEAK_TFLOPS, "traffic": None,
                    "kernel": "ragged_kernel x3 robots (one hipGraph)",
                    "algorithmic_flops_per_launch": flops, "sincos_per_launch": trigs,
                    "note": "reference source-level FP32 ops in plain order, counted on the GPU by the "
                            "reference-order replay; sin/cos counted separately"}
    cpu = None
    if not a.no_cpu_baseline:
        n_done, t_done, threads = 0, 0.0, 1
        for r in runs:
            pair_env = np.repeat(np.arange(r["sc"].num_envs, dtype=np.int32), np.diff(r["off"]))
            c = cpu_baseline_pairs(r["robot"], r["sc"], r["edges"], pair_env, seconds=4.0)
            n_s = int(c["sample"].split()[0])
            # weight each robot by its share of the batch
            n_done += r["edges"].shape[0]
            t_done += r["edges"].shape[0] / c["value"]
            threads = c["cores"]
        cpu = {"value": n_done / t_done, "unit": "edge-env checks/s", "cores": threads, "kind": "port",
               "sample": f"per robot a random sample of its pairs (about 4 s each, e.g. last {n_s} pairs), "
                         f"combined by each robot's share of the batch; C restatement of the reference "
                         f"checker (vamp unavailable), {threads} threads"}
    desc = (f"{len(env_robot)} envs interleaved panda/fetch/baxter, {total} pairs "
            f"(U{{32..512}} obstacles, lognormal edges)")
    rec = {
        "value": total / (ms / 1e3), "ms_per_step": ms,
        "config": {"workload": f"c4: {desc}, seed {a.seed}", "robot": "panda+fetch+baxter",
                   "envs": int(len(env_robot)), "edges": total,
                   "layout": "three ragged CSR batches (one per robot), one hipGraph with a branch per robot",
                   "pairs": total, "resolution": 32, "order": "plain", "collision_fraction": coll,
                   "per_robot": {r["robot"]: {"pairs": int(r["edges"].shape[0]),
                                              "collision_fraction": float(r["v"].mean()),
                                              "ms_alone": alone[r["robot"]],
                                              "tile_pairs": r["plan"].config[0],
                                              "groups_per_cu": r["plan"].config[1]} for r in runs}},
        "roofline": roofline, "cpu_baseline": cpu,
    }
    if learned is not None:
        learned["value"] = total / (learned["ms_per_step"] / 1e3)
        rec["learned"] = learned
    group.close()
    for r in runs:
        r["plan"].close()
        r["scene"].close()
    return rec


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="c1", choices=["c0", "c1", "c2", "c3", "c4", "mbm"],
                    help="mbm = the README run: 100 bookshelf_small scenes x 1000 Halton edges (panda)")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--scale", type=float, default=1.0, help="env/edge count multiplier (1.0 = full size)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-work", action="store_true", help="skip the reference-order work count")
    ap.add_argument("--no-learned", dest="learned", action="store_false",
                    help="skip training the step-ordering model and timing the learned-ordered check")
    ap.add_argument("--train-pairs", type=int, default=2_000_000)
    ap.add_argument("--train-epochs", type=int, default=8)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run the plain copy kernel on BYTES (counter calibration) and exit")
    a = ap.parse_args()

    if a.copy_calibration:
        bcc = load_pkg()
        src = bcc.DeviceBuffer(a.copy_calibration)
        dst = bcc.DeviceBuffer(a.copy_calibration)
        src.zero()
        for _ in range(3):
            bcc.debug_copy(src, dst, a.copy_calibration)
        print(json.dumps({"copy_calibration_bytes": a.copy_calibration, "launches": 3}))
        return

    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {a.gpus} is not supported "
              "(pairs are independent; multi-GPU sharding is described in DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    bcc = load_pkg()
    from batch_cc_amd import workloads as W

    if a.config == "c4":
        part = bench_mixed(bcc, W, a)
        rec = {"metric": METRIC, "value": part["value"], "unit": "edge-env checks/s", "n_gpus": 1,
               "steps": a.steps, "warmup": a.warmup, "ms_per_step": part["ms_per_step"], "higher_is_better": True,
               "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
               "config": part["config"], "roofline": part["roofline"], "cpu_baseline": part["cpu_baseline"]}
        if "learned" in part:
            rec["learned"] = part["learned"]
        print(json.dumps(rec))
        return

    t0 = time.perf_counter()
    ragged = a.config == "c2"
    if ragged:
        robot, sc, env_off, edges = W.make_ragged(a.config, seed=a.seed, scale=a.scale)
        E, M = sc.num_envs, edges.shape[0]
        pairs = M
        shape = (M,)
        desc = f"{robot}, {E} envs, {M} ragged pairs (Zipf(1.2) edges per env)"
    elif a.config == "mbm":
        from batch_cc_amd import mbm
        robot = "panda"
        problems = os.path.join(REPO, "tests", "golden", "mbm_panda.json")
        sc, edges = mbm.readme_run(problems, robot, max(1, int(100 * a.scale)), max(1, int(1000 * a.scale)))
        E, M = sc.num_envs, edges.shape[0]
        pairs = E * M
        shape = (M, E)
        desc = f"{robot}, {E} MotionBenchMaker bookshelf_small envs x {M} Halton edges"
    else:
        robot, sc, edges = W.make_config(a.config, seed=a.seed, scale=a.scale)
        E, M = sc.num_envs, edges.shape[0]
        pairs = E * M
        shape = (M, E)
        desc = f"{robot}, {E} envs x {M} edges"
    print(f"[bench] workload {a.config}: {desc} = {pairs} pairs "
          f"(generated in {time.perf_counter() - t0:.1f} s)", file=sys.stderr, flush=True)

    scene = bcc.Scene(sc)
    d_edges = bcc.DeviceBuffer.from_array(edges)
    d_out = bcc.DeviceBuffer(pairs)
    d_out.zero()
    plan = None
    if ragged:
        plan = bcc.RaggedPlan(robot, scene, env_off, d_edges, d_out)
        plan.launch()
    else:
        bcc.check_dense_device(robot, scene, d_edges, M, 32, d_out)
    bcc.synchronize()
    verdicts = d_out.to_array(np.uint8, shape)
    coll = float(verdicts.mean())

    wall0 = time.perf_counter()
    if ragged:
        ms = plan.time(a.warmup, a.steps)
    else:
        ms = bcc.time_dense(robot, scene, d_edges, M, 32, d_out, a.warmup, a.steps)
    bcc.synchronize()
    wall = time.perf_counter() - wall0
    print(f"[bench] {a.steps} timed launches: {ms:.3f} ms each (wall incl. warmup {wall:.2f} s)",
          file=sys.stderr, flush=True)
    again = d_out.to_array(np.uint8, shape)
    if not np.array_equal(again, verdicts):
        print("[bench] verdicts changed between launches", file=sys.stderr)
        sys.exit(1)
    value = pairs / (ms / 1e3)

    learned = None
    if a.learned and not ragged:
        learned = learned_leg(bcc, W, a, robot, scene, sc, d_edges, E, M, verdicts, ms)

    roofline = None
    if not a.no_work:
        t1 = time.perf_counter()
        if ragged:
            pair_env = np.repeat(np.arange(E, dtype=np.int32), np.diff(env_off))
            work, trig, wv = bcc.work_pairs(robot, scene, d_edges, pair_env, 32)
        else:
            work, trig, wv = bcc.work_dense(robot, scene, d_edges, M, 32)
        if not np.array_equal(wv, verdicts):
            print("[bench] work-counter verdicts differ from the dense kernel", file=sys.stderr)
            sys.exit(1)
        flops = float(work.sum(dtype=np.float64))
        trigs = float(trig.sum(dtype=np.float64))
        achieved = flops / (ms / 1e3) / 1e12
        roofline = {
            "bound": "valu",
            "achieved": achieved,
            "peak": FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / FP32_PEAK_TFLOPS,
            "traffic": MEASURED_TRAFFIC[a.config][0] if a.config in MEASURED_TRAFFIC and a.scale == 1.0 else None,
            "traffic_source": MEASURED_TRAFFIC[a.config][1] if a.config in MEASURED_TRAFFIC else None,
            "kernel": "ragged_kernel (hipGraph replay)" if ragged else "dense_kernel",
            "algorithmic_flops_per_launch": flops,
            "sincos_per_launch": trigs,
            "note": "reference source-level FP32 ops in plain order, counted on the GPU by the "
                    "reference-order replay (bcc_work_dense); sin/cos counted separately",
        }
        print(f"[bench] work count: {flops:.4g} flops/launch ({time.perf_counter() - t1:.1f} s)",
              file=sys.stderr, flush=True)

    cpu = None
    if not a.no_cpu_baseline and not ragged:
        cpu = cpu_baseline(robot, sc, edges)
    elif not a.no_cpu_baseline:
        cpu = cpu_baseline_pairs(robot, sc, edges, np.repeat(np.arange(E, dtype=np.int32), np.diff(env_off)))
    if cpu is not None:
        print(f"[bench] cpu baseline {cpu['value']:.4g} pairs/s on {cpu['cores']} threads", file=sys.stderr,
              flush=True)

    rec = {
        "metric": METRIC,
        "value": value,
        "unit": "edge-env checks/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "MotionBenchMaker problem file subset + Halton edges" if a.config == "mbm" else "synthetic",
        "config": {
            "workload": f"{a.config}: {desc}, seed {a.seed}",
            "robot": robot,
            "envs": E,
            "edges": M,
            "layout": "ragged CSR, one persistent launch replayed as a hipGraph" if ragged else "dense E x M",
            "pairs": pairs,
            "resolution": 32,
            "order": "plain",
            "collision_fraction": coll,
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    if learned is not None:
        # `value` is the plain-order path; the learned-ordered step (model
        # inference for every pair + ordered check) is reported beside it
        learned["value"] = pairs / (learned["ms_per_step"] / 1e3)
        rec["learned"] = learned
    print(json.dumps(rec))
    scene.close()


if __name__ == "__main__":
    main()
