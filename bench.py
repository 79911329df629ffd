"""Benchmark: edge-env checks/sec on one MI355X, learned-ordered vs plain,
with the roofline and the CPU baseline beside it.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config c3] [--seed 0]

Default workload: c3 (panda, 1000 environments x 100000 edges, 16 cuboids,
8 spheres and 8 cylinders per environment).  A step is one pass of the path
over the whole synthetic batch with inputs resident in HBM:

  learned step (`value`): one fused kernel per batch -- every pair's
      step-0 samples at bins 8 and 24 first, then, for the pairs still
      undecided only, the step-ordering model (BF16 MFMA inside the same
      kernel) picks the two bins to probe next, and the plain step loop
      finishes the rest;
  plain step (`plain_value`): the reference's order, 32 samples per step.

Both are timed with hipEvents on the launch stream around exactly K steps
after W untimed ones.  The model is trained first, on the GPU, on labels the
exact checker writes for a separate seeded edge stream (the pairs the first
probe misses).  Verdicts of every variant are compared with the plain
kernel's (0 mismatches, else exit 1).  Beside the headline: the same fused
kernel without a model (fixed second-probe bins: the method without the
network), perfect probe bins taken from exact labels of the benchmark pairs
themselves (the bound on any order), and how many pairs each stage decides.

Roofline of the step's dominant kernel: before it touches the GPU, the run
re-runs its own workload (one step of each kernel) three times in child
processes under `rocprofv3 --pmc` (SQ counters, FETCH_SIZE, WRITE_SIZE);
`roofline.achieved` is the executed FP32 rate (SQ_INSTS_VALU_FLOPS_FP32 in
units measured on a known-FMA kernel, over the kernel's hipEvent time here)
and `traffic` the HBM bytes per launch (scaled on a known-size copy).  The
reference's own source-level FP32 operation count for the same pairs in plain
order (bcc_work_dense) over the same time is `effective_achieved`.  Without
rocprofv3 (or with --no-counters) the effective figure stands in.  The CPU
baseline is the project's C restatement of the reference checker (the
reference's CPU path, vamp, is not available), timed on the host cores on a
bounded sample.  Prints one JSON line.
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
HBM_PEAK_GBS = 8000.0
DIM = {"panda": 7, "fetch": 8, "baxter": 14}
METRIC = "edge-env checks/sec at bit-exact verdicts, learned-ordered vs plain, 1 GPU"
RID = {"panda": 0, "fetch": 1, "baxter": 2}


def mlp_flops(robot: str) -> int:
    """Algorithmic flops of one inference: 3 layers over the unpadded input
    (64-cell encoding + start and end), hidden width H of the built model."""
    from batch_cc_amd import Model
    h = Model.hidden()
    return 2 * ((64 + 2 * DIM[robot]) * h + h * h + h * 32)


def load_pkg():
    pkg_dir = os.path.join(REPO, "batch-cc_amd")
    spec = importlib.util.spec_from_file_location("batch_cc_amd", os.path.join(pkg_dir, "__init__.py"),
                                                  submodule_search_locations=[pkg_dir])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["batch_cc_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def log(msg: str):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


# ------------------------------------------------------------ CPU baseline

def host_cpu() -> dict:
    """The host's CPU model, core count and the share this process may use
    (its affinity mask, capped by OMP_NUM_THREADS where the host sets it)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        share = len(os.sched_getaffinity(0))
    except AttributeError:
        share = os.cpu_count() or 1
    cap = os.environ.get("OMP_NUM_THREADS", "")
    if cap.isdigit() and int(cap) > 0:
        share = min(share, int(cap))
    return {"model": model, "nproc": os.cpu_count() or 1, "threads": max(1, share)}


def _time_oracle(run, n_start: int, n_max: int, seconds: float):
    """Grow a sample until `run(n)` takes `seconds` (or covers n_max)."""
    n = n_start
    while True:
        t0 = time.perf_counter()
        run(n)
        dt = time.perf_counter() - t0
        if dt >= seconds or n >= n_max:
            return n, dt
        n = int(min(n_max, max(n * 2, n * seconds / max(dt, 1e-3) * 1.1)))


def cpu_baseline(robot, sc, edges, pair_env=None, seconds=10.0, one_thread_seconds=4.0):
    """The oracle (the project's C restatement of the reference checker) on
    seeded random pairs of the same workload: all the host cores this process
    may use, then one thread.  Dense when pair_env is None (pairs = (edge,
    env)); ragged otherwise (pair p = (edge p, env pair_env[p]))."""
    sys.path.insert(0, os.path.join(REPO, "tests"))
    import oracle_lib
    orc = oracle_lib.Oracle()
    cpu = host_cpu()
    rid = RID[robot]
    rng = np.random.default_rng(1234)
    E, M = sc.num_envs, edges.shape[0]
    total = M if pair_env is not None else E * M

    def run(threads):
        def go(n):
            if pair_env is None:
                pe = rng.integers(0, E, n).astype(np.int32)
                pg = rng.integers(0, M, n).astype(np.int32)
            else:
                pg = rng.integers(0, M, n).astype(np.int32)
                pe = pair_env[pg]
            orc.check_pairs(rid, sc, edges, pe, pg, threads=threads)
        return go

    n_all, dt_all = _time_oracle(run(cpu["threads"]), 2000, total, seconds)
    n_one, dt_one = _time_oracle(run(1), 200, total, one_thread_seconds)
    return {
        "value": n_all / dt_all,
        "unit": "edge-env checks/s",
        "cores": cpu["threads"],
        "kind": "port",
        "sample": f"{n_all} random (env, edge) pairs of the same workload in {dt_all:.1f} s on {cpu['threads']} "
                  f"threads; C restatement of the reference checker (vamp unavailable)",
        "one_thread": {"value": n_one / dt_one, "sample": f"{n_one} random pairs in {dt_one:.1f} s"},
        "host": {"cpu_model": cpu["model"], "nproc": cpu["nproc"], "threads_used": cpu["threads"]},
    }


# ------------------------------------------------------------ learned step

# first-probe bins of the fused learned kernel, (32 k + 16) / 2
FIRST_PROBE_MASK = (1 << 8) | (1 << 24)


def train_model(bcc, W, a, robot, scene, E, train_edges, rng):
    """The step-ordering model trained on the GPU on the pairs it will be
    asked about: step-0 labels (written by the exact checker) of seeded (env,
    edge) pairs of a separate edge stream, kept when the fused kernel's first
    probe would miss them (bins 8 and 24 free at step 0), a.train_pairs of
    them from at most 16 times as many labelled pairs."""
    d_train = bcc.DeviceBuffer.from_array(train_edges)
    n = int(a.train_pairs)
    n_big = 16 * n
    pe = rng.integers(0, E, n_big).astype(np.int32)
    pg = rng.integers(0, train_edges.shape[0], n_big).astype(np.int32)
    t0 = time.perf_counter()
    lab = bcc.label_pairs(robot, scene, d_train, pe, pg, steps=1).to_array(np.uint32, (n_big,))
    t_label = time.perf_counter() - t0
    keep = np.nonzero((lab & np.uint32(FIRST_PROBE_MASK)) == 0)[0][:n]
    model = bcc.Model(robot, seed=a.seed)
    t0 = time.perf_counter()
    d_x = model.features(scene, d_train, pe[keep], pg[keep])
    d_labels = bcc.DeviceBuffer.from_array(lab[keep])
    # 64 epochs at lr 3e-3 (MI355X, c3: the second probe decides 5.04M of the
    # 6.95M inferred pairs, against 4.90M after 8 epochs at 2e-3)
    losses = model.train(d_x, d_labels, int(keep.size), epochs=a.train_epochs, batch=8192, lr=3e-3)
    bcc.synchronize()
    t_train = time.perf_counter() - t0
    log(f"learned {robot}: {n_big} step-0 labels in {t_label:.1f} s, {keep.size} of them miss the first probe; "
        f"{a.train_epochs} epochs in {t_train:.1f} s, loss {losses[0]:.3f} -> {losses[-1]:.3f}")
    info = {"labelled_pairs": n_big, "pairs": int(keep.size), "epochs": a.train_epochs,
            "labels": "step-0 samples of each bin (steps=1) of pairs whose first-probe bins (8, 24) are free",
            "label_s": t_label, "train_s": t_train, "loss_first_epoch": float(losses[0]),
            "loss_last_epoch": float(losses[-1]), "edges": "separate seeded edge stream, same environments"}
    return model, info


def perfect_orders(labels: np.ndarray) -> np.ndarray:
    """Order words whose first two bins hold colliding step-0 samples where
    the pair has any (lowest bins first), padded with other distinct bins."""
    L = labels.astype(np.uint64)

    def ctz(x):
        low = x & (~x + np.uint64(1))
        return np.where(x != 0, np.log2(np.maximum(low, 1).astype(np.float64)).astype(np.int64), -1)

    b0 = ctz(L)
    b1 = ctz(L & (L - np.uint64(1)))
    b0 = np.where(b0 < 0, 16, b0)
    b1 = np.where(b1 < 0, (b0 + 16) % 32, b1)
    b2 = np.zeros_like(b0)
    for c in (2, 1, 0):
        b2 = np.where((b0 != c) & (b1 != c), c, b2)
    b3 = np.zeros_like(b0)
    for c in (3, 2, 1, 0):
        b3 = np.where((b0 != c) & (b1 != c) & (b2 != c), c, b3)
    return (b0 | (b1 << 8) | (b2 << 16) | (b3 << 24)).astype(np.uint32)


def probe_decided(bcc, robot, scene, d_edges, M, E, d_order, plain_verdicts):
    """Fraction of colliding pairs the ordered kernel's probe decides."""
    d_out = bcc.DeviceBuffer(E * M)
    d_rounds = bcc.DeviceBuffer(E * M)
    bcc.check_dense_ex(robot, scene, d_edges, M, 32, d_out, d_order, d_rounds)
    r = d_rounds.to_array(np.uint8, (M, E))
    coll = plain_verdicts == 1
    return float((r[coll] == 0).mean()) if coll.any() else None


def stage_counts(stats) -> dict:
    return {"first_probe": int(stats[0]), "second_probe": int(stats[1]), "inferred": int(stats[3]),
            "step_loop": int(stats[4])}


def learned_dense(bcc, W, a, robot, scene, sc, d_edges, E, M, plain_verdicts, plain_ms):
    """The learned step on a dense batch: one fused kernel (first probe at
    evenly spaced bins, in-kernel inference for the undecided pairs only,
    probe of their predicted bins, plain steps), against the same kernel
    without a model and the bound of perfect probe bins."""
    rng = np.random.default_rng([a.seed, 3])
    n_train_edges = max(1, min(M, 100_000))
    edge_law = W.lognormal_edges if robot == "fetch" else W.uniform_edges
    model, train = train_model(bcc, W, a, robot, scene, E, edge_law(robot, n_train_edges, rng), rng)
    d_out = bcc.DeviceBuffer(E * M)
    step_ms = bcc.time_dense_learned(robot, model, scene, d_edges, M, d_out, a.warmup, a.steps)
    mismatches = int((d_out.to_array(np.uint8, (M, E)) != plain_verdicts).sum())
    st = bcc.check_dense_learned(robot, model, scene, d_edges, M, stats=True)
    mismatches += int((st["verdicts"] != plain_verdicts).sum())
    none_ms = bcc.time_dense_learned(robot, None, scene, d_edges, M, d_out, a.warmup, a.steps)
    none_bad = int((d_out.to_array(np.uint8, (M, E)) != plain_verdicts).sum())
    st0 = bcc.check_dense_learned(robot, None, scene, d_edges, M, stats=True)
    none_bad += int((st0["verdicts"] != plain_verdicts).sum())
    mismatches += none_bad
    small = E * M <= bcc.learned_small_pairs()
    route = "small batch: inference of every pair, then the ordered dense kernel" if small else "fused kernel"
    log(f"learned step {step_ms:.3f} ms ({route}; fused-kernel stages {stage_counts(st['stats'])}); no network "
        f"{none_ms:.3f} ms ({stage_counts(st0['stats'])}); plain {plain_ms:.3f} ms; verdict mismatches {mismatches}")
    res = {
        "ms_per_step": step_ms,
        "speedup_vs_plain": plain_ms / step_ms,
        "verdict_mismatches_vs_plain": mismatches,
        "kernel": ("fused_order_kernel (BF16 MFMA inference of every pair) -> ordered_dense_kernel (probe of each "
                   f"pair's 2 top predicted bins, then plain steps): batches of at most {bcc.learned_small_pairs()} "
                   "pairs" if small else
                   "learned_kernel: first probe (bins 8, 24) -> in-kernel BF16 MFMA inference of the undecided "
                   "pairs -> probe of their 2 top predicted bins -> plain steps, one persistent launch"),
        "pairs_by_stage": stage_counts(st["stats"]),
        "train": train,
        "no_network": {"ms_per_step": none_ms, "speedup_vs_plain": plain_ms / none_ms,
                       "verdict_mismatches_vs_plain": none_bad, "pairs_by_stage": stage_counts(st0["stats"]),
                       "note": ("the ordered dense kernel with the invalid word for every pair: the probe takes "
                                "the fixed bins 8 and 24" if small else
                                "the same fused kernel without a model: the second probe takes the fixed bins 0 "
                                "and 16")},
        "network_gain_vs_no_network": none_ms / step_ms,
    }
    if a.bound:
        # perfect bins from exact step-0 labels of every benchmark pair
        t0 = time.perf_counter()
        pe = np.tile(np.arange(E, dtype=np.int32), M)
        pg = np.repeat(np.arange(M, dtype=np.int32), E)
        lab = bcc.label_pairs(robot, scene, d_edges, pe, pg, steps=1).to_array(np.uint32, (E * M,))
        del pe, pg
        d_best = bcc.DeviceBuffer.from_array(perfect_orders(lab))
        del lab
        best_ms = bcc.time_dense_ex(robot, scene, d_edges, M, 32, d_out, d_best, a.warmup, a.steps)
        best_bad = int((d_out.to_array(np.uint8, (M, E)) != plain_verdicts).sum())
        res["oracle_bound"] = {"ordered_check_ms": best_ms, "speedup_vs_plain": plain_ms / best_ms,
                               "verdict_mismatches_vs_plain": best_bad,
                               "probe_decided_colliding": probe_decided(bcc, robot, scene, d_edges, M, E, d_best,
                                                                        plain_verdicts),
                               "note": "every pair probed first at perfect bins from exact step-0 labels of the "
                                       "benchmark pairs (ordered_dense_kernel, no inference): the bound on any "
                                       "learned order"}
        res["oracle_bound_speedup"] = plain_ms / best_ms
        mismatches += best_bad
        log(f"perfect bins {best_ms:.3f} ms ({time.perf_counter() - t0:.1f} s)")
        del d_best
    res["verdict_mismatches_vs_plain"] = mismatches
    if mismatches:
        log("learned-ordered verdicts differ from plain order")
        sys.exit(1)
    model.close()
    return res


def learned_mixed(bcc, W, a, runs, plain_ms, edge_law=None) -> dict:
    """Learned leg of a ragged or mixed-robot batch: one model per robot,
    trained as for the dense configs on that robot's environments with a
    separate edge stream of the batch's edge law (default: c4's lognormal
    law); one learned step = one replay of a group of learned plans (the
    fused kernel, persistent, per robot), against the same group without
    models."""
    for k, r in enumerate(runs):
        rng = np.random.default_rng([a.seed, 3, k])
        n_edges = max(1, min(r["edges"].shape[0], 100_000))
        if edge_law is None:
            train = W.lognormal_edges(r["robot"], n_edges, rng, median=W.MIXED_EDGE_MEDIAN)
        else:
            train = edge_law(r["robot"], n_edges, rng)
        r["model"], r["train"] = train_model(bcc, W, a, r["robot"], r["scene"], r["sc"].num_envs, train, rng)
    res = {}
    legs = (("learned", True),) if a.counter_child else (("learned", True), ("no_network", False))
    for name, with_model in legs:
        plans = []
        for r in runs:
            r["d_out_l"] = bcc.DeviceBuffer(max(1, r["edges"].shape[0]))
            plans.append(bcc.RaggedPlan(r["robot"], r["scene"], r["off"], r["d_edges"], r["d_out_l"], learned=True,
                                        model=r["model"] if with_model else None))
        group = bcc.RaggedGroup(plans)
        ms = group.time(a.warmup, a.steps)
        alone = {r["robot"]: p.time(1, max(1, a.steps // 2)) for r, p in zip(runs, plans)}
        bad = 0
        for r in runs:
            bad += int((r["d_out_l"].to_array(np.uint8, (r["edges"].shape[0],)) != r["v"]).sum())
        group.close()
        for p in plans:
            p.close()
        res[name] = {"ms_per_step": ms, "speedup_vs_plain": plain_ms / ms, "verdict_mismatches_vs_plain": bad,
                     "ms_alone": alone}
        log(f"{name} group {ms:.3f} ms (plain group {plain_ms:.3f} ms; alone "
            + ", ".join(f"{k} {v:.2f}" for k, v in alone.items()) + f"); verdict mismatches {bad}")
    for r in runs:
        r["model"].close()
    if a.counter_child:
        return dict(res["learned"])
    mismatches = res["learned"]["verdict_mismatches_vs_plain"] + res["no_network"]["verdict_mismatches_vs_plain"]
    if mismatches:
        log("learned-ordered verdicts differ from plain order")
        sys.exit(1)
    out = dict(res["learned"])
    out.update({
        "kernel": "learned_kernel per robot (first probe, in-kernel BF16 MFMA inference of the undecided pairs, "
                  "predicted probe, plain steps), one persistent launch per robot, one hipGraph",
        "no_network": res["no_network"],
        "network_gain_vs_no_network": res["no_network"]["ms_per_step"] / res["learned"]["ms_per_step"],
        "verdict_mismatches_vs_plain": mismatches,
        "train": {r["robot"]: r["train"] for r in runs},
    })
    return out


def make_workload(W, a):
    """(robot, scene, edges, env_off or None, E, M, pairs, verdict shape,
    description) of a dense or ragged config."""
    if a.config == "c2":
        robot, sc, env_off, edges = W.make_ragged(a.config, seed=a.seed, scale=a.scale)
        E, M = sc.num_envs, edges.shape[0]
        return (robot, sc, edges, env_off, E, M, M, (M,),
                f"{robot}, {E} envs, {M} ragged pairs (Zipf(1.2) edges per env)")
    if a.config == "mbm":
        from batch_cc_amd import mbm
        robot = "panda"
        problems = os.path.join(REPO, "tests", "golden", "mbm_panda.json")
        sc, edges = mbm.readme_run(problems, robot, max(1, int(100 * a.scale)), max(1, int(1000 * a.scale)))
        E, M = sc.num_envs, edges.shape[0]
        return (robot, sc, edges, None, E, M, E * M, (M, E),
                f"{robot}, {E} MotionBenchMaker bookshelf_small envs x {M} Halton edges")
    robot, sc, edges = W.make_config(a.config, seed=a.seed, scale=a.scale)
    E, M = sc.num_envs, edges.shape[0]
    return robot, sc, edges, None, E, M, E * M, (M, E), f"{robot}, {E} envs x {M} edges"


# ------------------------------------------------------------ counters

def child_args(a) -> list:
    """Arguments of a counter child: the same workload, one step of each kernel."""
    args = ["--config", a.config, "--seed", str(a.seed), "--scale", str(a.scale), "--train-pairs",
            str(a.train_pairs), "--train-epochs", str(a.train_epochs), "--counter-child"]
    if not a.learned:
        args.append("--no-learned")
    return args


def counter_child(bcc, W, C, a):
    """Runs under rocprofv3 --pmc: the known-work kernels, then one launch of
    the workload's plain kernel and, for learned configs, the model trained as
    in the measured run and one learned step (the fused kernel)."""
    C.run_known_work(bcc)
    if a.config == "c4":
        a.steps, a.warmup, a.no_work, a.no_cpu_baseline = 1, 0, True, True
        bench_mixed(bcc, W, a, None, {})
        return
    robot, sc, edges, env_off, E, M, pairs, shape, desc = make_workload(W, a)
    scene = bcc.Scene(sc)
    d_edges = bcc.DeviceBuffer.from_array(edges)
    d_out = bcc.DeviceBuffer(pairs)
    if env_off is not None:
        plan = bcc.RaggedPlan(robot, scene, env_off, d_edges, d_out)
        plan.launch()
        bcc.synchronize()
        plan.close()
        if a.learned:
            run = dict(robot=robot, sc=sc, scene=scene, off=env_off, edges=edges, d_edges=d_edges,
                       v=d_out.to_array(np.uint8, (pairs,)))
            learned_mixed(bcc, W, a, [run], 1.0, edge_law=W.uniform_edges)
    else:
        bcc.check_dense_device(robot, scene, d_edges, M, 32, d_out)
        bcc.synchronize()
        if a.learned:
            rng = np.random.default_rng([a.seed, 3])
            edge_law = W.lognormal_edges if robot == "fetch" else W.uniform_edges
            model, _ = train_model(bcc, W, a, robot, scene, E, edge_law(robot, max(1, min(M, 100_000)), rng), rng)
            bcc.time_dense_learned(robot, model, scene, d_edges, M, d_out, 0, 1)
            bcc.synchronize()
            model.close()
    scene.close()


def build_roofline(C, hw, cal, match, kernel: str, kms: float, ref_flops=None, ref_trigs=None) -> dict:
    """Roofline of the step's dominant kernel(s): executed FP32 rate from the
    counters (unit from the known-work kernel) when they were collected, HBM
    traffic from FETCH_SIZE + WRITE_SIZE (scaled on the known-size copy), and
    the reference-order work rate beside it as the effective figure."""
    eff = ref_flops / (kms / 1e3) / 1e12 if ref_flops else None
    c = C.kernel_counters(hw, cal, match, kms, FP32_PEAK_TFLOPS, HBM_PEAK_GBS) if hw else {}
    rf = {"bound": "valu", "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s", "kernel": kernel, "kernel_ms": kms}
    if "executed_tflops" in c:
        rf["achieved"] = c["executed_tflops"]
        rf["frac"] = c["executed_frac"]
        rf["source"] = ("executed FP32 operations (SQ_INSTS_VALU_FLOPS_FP32 of the kernel's one-step rocprofv3 "
                        "pass x the unit measured on a known-FMA kernel) / the kernel's hipEvent time")
    else:
        rf["achieved"] = eff
        rf["frac"] = eff / FP32_PEAK_TFLOPS if eff is not None else None
        rf["source"] = "reference-order work / hipEvent time (no counters collected)"
    rf["traffic"] = c.get("traffic_bytes_per_launch")
    rf["effective_achieved"] = eff
    rf["effective_frac"] = eff / FP32_PEAK_TFLOPS if eff is not None else None
    rf["algorithmic_flops_per_launch"] = ref_flops
    rf["sincos_per_launch"] = ref_trigs
    if c:
        rf["counters"] = {k: v for k, v in c.items() if k not in ("executed_tflops", "executed_frac")}
        rf["calibration"] = cal
    rf["note"] = ("effective = the reference's source-level FP32 ops for these pairs in plain order (counted on "
                  "the GPU by the reference-order replay) / the kernel time; the kernel executes less than that "
                  "(early exits, culling, probing), so effective_frac can exceed 1")
    return rf


def fp32_fma_rate(bcc) -> dict:
    """Measured FP32 VALU rate of the known-FMA kernel (16 waves per SIMD,
    independent chains, non-packed FMAs as every kernel here is built): the
    practical ceiling beside the datasheet peak."""
    from batch_cc_amd import counters as C
    blocks, iters = 4096, 1024
    out = bcc.DeviceBuffer(4 * 256 * blocks)
    bcc.debug_valu(blocks, iters, 64, out)                    # warm-up
    ms = min(bcc.debug_valu(blocks, iters, 64, out) for _ in range(3))
    flops = (blocks * 4) * 64 * (2 * C.VALU_CHAINS * iters + C.VALU_CHAINS - 1)
    return {"tflops": flops / (ms / 1e3) / 1e12, "ms": ms, "flops": float(flops),
            "kernel": f"valu_kernel: {blocks} x 256 threads, 8 chains x {iters} v_fma_f32 per lane"}


def mfma_rate(hw, match, step_ms: float) -> dict:
    """BF16 matrix-core work of the learned step (the in-kernel inference):
    SQ_INSTS_VALU_MFMA_MOPS_BF16 x 512 flops per launch over the step time,
    against the dense BF16 peak."""
    out = {"peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s", "flops_per_pair": mlp_flops("panda"),
           "note": "executed MFMA flops (padded K: 96 inputs per layer-1 row) of the inference inside the learned "
                   "kernel, averaged over the whole step; inference runs for the undecided pairs only"}
    if hw and "sq" in hw:
        mops = bcc_counters().per_step(hw["sq"], match, "SQ_INSTS_VALU_MFMA_MOPS_BF16")
        if mops:
            out["executed_flops_per_launch"] = 512.0 * mops
            out["achieved"] = 512.0 * mops / (step_ms / 1e3) / 1e12
            out["frac"] = out["achieved"] / BF16_PEAK_TFLOPS
    return out


def is_kernel(prefix: str):
    return lambda name: name.startswith(prefix)


def bcc_counters():
    from batch_cc_amd import counters
    return counters

# ------------------------------------------------------------ workloads

def bench_mixed(bcc, W, a, hw, cal) -> dict:
    """c4: panda, fetch and baxter environments interleaved; one ragged plan
    per robot, all replayed as one hipGraph (bcc_ragged_group).  hw / cal:
    the counter passes and their calibration (or None / {})."""
    t0 = time.perf_counter()
    parts, env_robot, global_off = W.make_mixed(seed=a.seed, scale=a.scale)
    total = int(global_off[-1])
    log(f"workload c4: {len(env_robot)} interleaved envs, {total} pairs (generated in "
        f"{time.perf_counter() - t0:.1f} s)")
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
            log("verdicts changed between launches")
            sys.exit(1)
    log(f"{a.steps} timed group launches: {ms:.3f} ms each; plans alone: "
        + ", ".join(f"{k} {v:.3f} ms" for k, v in alone.items()))
    learned = learned_mixed(bcc, W, a, runs, ms) if a.learned else None
    flops = trigs = None
    if not a.no_work:
        flops = trigs = 0.0
        for r in runs:
            pair_env = np.repeat(np.arange(r["sc"].num_envs, dtype=np.int32), np.diff(r["off"]))
            work, trig, wv = bcc.work_pairs(r["robot"], r["scene"], r["d_edges"], pair_env, 32)
            if not np.array_equal(wv, r["v"]):
                log("work-counter verdicts differ from the ragged kernel")
                sys.exit(1)
            flops += float(work.sum(dtype=np.float64))
            trigs += float(trig.sum(dtype=np.float64))
    # the learned step (learned_kernel per robot) is the measured step; the
    # plain group's ragged_kernel per robot is reported beside it
    plain_rf = build_roofline(bcc_counters(), hw, cal, is_kernel("bcc::ragged_kernel<"),
                              "ragged_kernel x3 robots (one hipGraph, plain order)", ms, flops, trigs)
    if learned:
        roofline = build_roofline(bcc_counters(), hw, cal, is_kernel("bcc::learned_kernel<"),
                                  "learned_kernel x3 robots (one hipGraph: the whole learned step)",
                                  learned["ms_per_step"], flops, trigs)
        plain_rf.pop("calibration", None)
        roofline["plain_kernel"] = plain_rf
        learned["mfma"] = mfma_rate(hw, is_kernel("bcc::learned_kernel<"), learned["ms_per_step"])
    else:
        roofline = plain_rf
    if roofline["achieved"] is None:
        roofline = None
    cpu = None
    if not a.no_cpu_baseline:
        n_done, t_done, c = 0, 0.0, None
        for r in runs:
            pair_env = np.repeat(np.arange(r["sc"].num_envs, dtype=np.int32), np.diff(r["off"]))
            c = cpu_baseline(r["robot"], r["sc"], r["edges"], pair_env, seconds=4.0, one_thread_seconds=1.5)
            # weight each robot by its share of the batch
            n_done += r["edges"].shape[0]
            t_done += r["edges"].shape[0] / c["value"]
        cpu = {"value": n_done / t_done, "unit": "edge-env checks/s", "cores": c["cores"], "kind": "port",
               "sample": "per robot a random sample of its pairs (about 4 s each), combined by each robot's share "
                         "of the batch; C restatement of the reference checker (vamp unavailable)",
               "host": c["host"]}
    desc = (f"{len(env_robot)} envs interleaved panda/fetch/baxter, {total} pairs "
            f"(U{{32..512}} obstacles, lognormal edges)")
    rec = {
        "value": total / (ms / 1e3), "ms_per_step": ms,
        "config": {"workload": f"c4: {desc}, seed {a.seed}", "robot": "panda+fetch+baxter",
                   "envs": int(len(env_robot)), "edges": total,
                   "layout": "three ragged CSR batches (one per robot), one hipGraph with a branch per robot",
                   "pairs": total, "resolution": 32, "collision_fraction": coll,
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
    ap.add_argument("--config", default="c3", choices=["c0", "c1", "c2", "c3", "c4", "mbm"],
                    help="mbm = the README run: 100 bookshelf_small scenes x 1000 Halton edges (panda)")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--scale", type=float, default=1.0, help="env/edge count multiplier (1.0 = full size)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-work", action="store_true", help="skip the reference-order work count")
    ap.add_argument("--no-learned", dest="learned", action="store_false",
                    help="plain order only: skip training and the learned-ordered step")
    ap.add_argument("--no-bound", dest="bound", action="store_false",
                    help="skip the perfect-bin bound (exact step-0 labels of every benchmark pair)")
    ap.add_argument("--train-pairs", type=int, default=2_000_000)
    ap.add_argument("--train-epochs", type=int, default=64)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run the plain copy kernel on BYTES (counter calibration) and exit")
    ap.add_argument("--no-counters", dest="counters", action="store_false",
                    help="skip the rocprofv3 counter passes (roofline from reference-order work only)")
    ap.add_argument("--counter-child", action="store_true", help=argparse.SUPPRESS)
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
    from batch_cc_amd import counters as C

    if a.counter_child:
        counter_child(bcc, W, C, a)
        return

    # counter passes first, while this process has not touched the GPU
    hw, cal = None, {}
    if a.counters:
        t0 = time.perf_counter()
        hw = C.collect(os.path.abspath(__file__), child_args(a), log)
        cal = C.calibration(hw) if hw else {}
        log(f"counters: {len(hw or {})} passes in {time.perf_counter() - t0:.1f} s; calibration {cal}")
    fma = fp32_fma_rate(bcc)
    log(f"known-FMA kernel: {fma['tflops']:.1f} TFLOP/s FP32 (non-packed FMAs)")

    if a.config == "c4":
        part = bench_mixed(bcc, W, a, hw, cal)
        learned = part.get("learned")
        if part["roofline"]:
            part["roofline"]["measured_fp32_fma"] = fma
        rec = {"metric": METRIC, "value": learned["value"] if learned else part["value"], "unit": "edge-env checks/s",
               "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
               "ms_per_step": learned["ms_per_step"] if learned else part["ms_per_step"], "higher_is_better": True,
               "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
               "config": {**part["config"], "order": "learned" if learned else "plain"},
               "plain_value": part["value"], "plain_ms_per_step": part["ms_per_step"],
               "roofline": part["roofline"], "cpu_baseline": part["cpu_baseline"]}
        if learned:
            rec["speedup_vs_plain"] = learned["speedup_vs_plain"]
            rec["verdict_mismatches_vs_plain"] = learned["verdict_mismatches_vs_plain"]
            rec["learned"] = learned
        print(json.dumps(rec))
        return

    t0 = time.perf_counter()
    robot, sc, edges, env_off, E, M, pairs, shape, desc = make_workload(W, a)
    ragged = env_off is not None
    log(f"workload {a.config}: {desc} = {pairs} pairs (generated in {time.perf_counter() - t0:.1f} s)")

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
    log(f"plain: {a.steps} timed launches {ms:.3f} ms each (wall incl. warmup {time.perf_counter() - wall0:.2f} s)")
    if not np.array_equal(d_out.to_array(np.uint8, shape), verdicts):
        log("verdicts changed between launches")
        sys.exit(1)
    plain_value = pairs / (ms / 1e3)

    learned = None
    if a.learned and not ragged:
        learned = learned_dense(bcc, W, a, robot, scene, sc, d_edges, E, M, verdicts, ms)
        learned["value"] = pairs / (learned["ms_per_step"] / 1e3)
    elif a.learned:
        # one learned plan over the ragged batch (edges drawn uniformly, as c2's)
        run = dict(robot=robot, sc=sc, scene=scene, off=env_off, edges=edges, d_edges=d_edges, v=verdicts)
        learned = learned_mixed(bcc, W, a, [run], ms, edge_law=W.uniform_edges)
        learned["value"] = pairs / (learned["ms_per_step"] / 1e3)

    flops = trigs = None
    if not a.no_work:
        t1 = time.perf_counter()
        if ragged:
            pair_env = np.repeat(np.arange(E, dtype=np.int32), np.diff(env_off))
            work, trig, wv = bcc.work_pairs(robot, scene, d_edges, pair_env, 32)
        else:
            work, trig, wv = bcc.work_dense(robot, scene, d_edges, M, 32)
        if not np.array_equal(wv, verdicts):
            log("work-counter verdicts differ from the dense kernel")
            sys.exit(1)
        flops = float(work.sum(dtype=np.float64))
        trigs = float(trig.sum(dtype=np.float64))
        del work, trig, wv
        log(f"work count: {flops:.4g} flops/launch ({time.perf_counter() - t1:.1f} s)")
    # the measured step's kernel: the fused learned kernel, else the plain one
    if learned and ragged:
        roofline = build_roofline(C, hw, cal, is_kernel("bcc::learned_kernel<"),
                                  "learned_kernel (learned plan, hipGraph replay)", learned["ms_per_step"], flops,
                                  trigs)
        roofline["plain_kernel"] = build_roofline(C, hw, cal, is_kernel("bcc::ragged_kernel<"),
                                                  "ragged_kernel (plain plan, hipGraph replay)", ms, flops, trigs)
        roofline["plain_kernel"].pop("calibration", None)
        learned["mfma"] = mfma_rate(hw, is_kernel("bcc::learned_kernel<"), learned["ms_per_step"])
    elif learned:
        roofline = build_roofline(C, hw, cal, is_kernel("bcc::learned_kernel<"),
                                  "learned_kernel (the whole learned step)", learned["ms_per_step"], flops, trigs)
        roofline["plain_kernel"] = build_roofline(C, hw, cal, is_kernel("bcc::dense_kernel<"),
                                                  "dense_kernel (plain order)", ms, flops, trigs)
        roofline["plain_kernel"].pop("calibration", None)
        learned["mfma"] = mfma_rate(hw, is_kernel("bcc::learned_kernel<"), learned["ms_per_step"])
    else:
        roofline = build_roofline(C, hw, cal, is_kernel("bcc::ragged_kernel<" if ragged else "bcc::dense_kernel<"),
                                  "ragged_kernel (hipGraph replay)" if ragged else "dense_kernel", ms, flops, trigs)
    if roofline["achieved"] is None:
        roofline = None
    else:
        roofline["measured_fp32_fma"] = fma

    cpu = None
    if not a.no_cpu_baseline:
        cpu = cpu_baseline(robot, sc, edges, np.repeat(np.arange(E, dtype=np.int32), np.diff(env_off))
                           if ragged else None)
        log(f"cpu baseline {cpu['value']:.4g} pairs/s on {cpu['cores']} threads "
            f"({cpu['one_thread']['value']:.4g} on one; {cpu['host']['cpu_model']}, nproc {cpu['host']['nproc']})")

    step_ms = learned["ms_per_step"] if learned else ms
    rec = {
        "metric": METRIC,
        "value": pairs / (step_ms / 1e3),
        "unit": "edge-env checks/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": step_ms,
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
            "order": ("learned (one fused kernel: probe of bins 8 and 24, in-kernel inference of the undecided "
                      "pairs, probe of their 2 top predicted bins, then plain steps)") if learned else "plain",
            "collision_fraction": coll,
        },
        "plain_value": plain_value,
        "plain_ms_per_step": ms,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    if learned is not None:
        rec["speedup_vs_plain"] = learned["speedup_vs_plain"]
        rec["verdict_mismatches_vs_plain"] = learned["verdict_mismatches_vs_plain"]
        if "oracle_bound_speedup" in learned:
            rec["oracle_bound_speedup"] = learned["oracle_bound_speedup"]
        rec["learned"] = learned
    print(json.dumps(rec))
    scene.close()


if __name__ == "__main__":
    main()
