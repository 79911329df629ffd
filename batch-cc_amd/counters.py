"""Hardware counters of one benchmark workload, for its roofline.

The benchmark process re-runs its own workload in child processes under
`rocprofv3 --pmc`, one counter group per child (rocprofv3 collects a group in
one pass and refuses more counters per block than the hardware has), and
reads back, per kernel, the mean over its dispatches of each counter summed
over the chip.  The children run before the parent touches the GPU.

Every child also runs two known-work kernels, so the counters' units are
measured rather than assumed:
  - valu_kernel (bcc_debug_valu): 8 x iters FMAs in every lane below
    `active`, once with 64 and once with 32 active lanes per wave: fixes the
    unit of SQ_INSTS_VALU_FLOPS_FP32 (FP32 operations per count) and shows
    whether it counts active lanes only;
  - copy_kernel (bcc_debug_copy): reads and writes a known number of bytes:
    fixes the scale of FETCH_SIZE and WRITE_SIZE (KiB per dispatch).
"""
from __future__ import annotations

import csv
import os
import shutil
import signal
import subprocess
import sys
import tempfile
import time
from collections import defaultdict

# counter groups, one rocprofv3 pass each: at most 8 SQ, 4 TCC (FETCH_SIZE
# takes 3, WRITE_SIZE 2) and 2 GRBM counters per pass
PASSES = (
    ("sq", ("SQ_WAVES", "SQ_INSTS_VALU", "SQ_INSTS_VALU_FLOPS_FP32", "SQ_INSTS_VALU_FMA_F32",
            "SQ_INSTS_VALU_MFMA_MOPS_BF16", "SQ_WAVE_CYCLES", "SQ_BUSY_CYCLES", "SQ_VALU_MFMA_BUSY_CYCLES",
            "GRBM_GUI_ACTIVE")),
    ("fetch", ("FETCH_SIZE",)),
    ("write", ("WRITE_SIZE",)),
)

# the known-work kernels' shapes (bcc_debug_valu / bcc_debug_copy)
VALU_BLOCKS = 2048
VALU_ITERS = 4096
VALU_CHAINS = 8
COPY_BYTES = 256 << 20


def valu_known_flops(active_lanes: int) -> float:
    """FP32 operations of one valu_kernel dispatch: per active lane 8 chains
    of VALU_ITERS FMAs (2 operations each) plus 7 adds."""
    waves = VALU_BLOCKS * 256 // 64
    return float(waves * active_lanes * (2 * VALU_CHAINS * VALU_ITERS + VALU_CHAINS - 1))


def run_known_work(bcc):
    """The two known-work kernels (called in every counter child)."""
    out = bcc.DeviceBuffer(4 * 256 * VALU_BLOCKS)
    bcc.debug_valu(VALU_BLOCKS, VALU_ITERS, 64, out)
    bcc.debug_valu(VALU_BLOCKS, VALU_ITERS, 32, out)
    src = bcc.DeviceBuffer(COPY_BYTES)
    dst = bcc.DeviceBuffer(COPY_BYTES)
    src.zero()
    bcc.debug_copy(src, dst, COPY_BYTES)
    bcc.synchronize()


def rocprof() -> str | None:
    p = shutil.which("rocprofv3")
    if p:
        return p
    p = "/opt/rocm/bin/rocprofv3"
    return p if os.path.exists(p) else None


def short_name(kernel: str) -> str:
    """`void bcc::ordered_dense_kernel<bcc::Robot_panda, 2, 6, 1>(...)` ->
    `bcc::ordered_dense_kernel<bcc::Robot_panda, 2, 6, 1>`; anonymous
    namespaces dropped (`(anonymous namespace)::copy_kernel(...)` ->
    `copy_kernel`)."""
    k = kernel.strip().replace("(anonymous namespace)::", "")
    if k.startswith("void "):
        k = k[5:]
    depth = 0
    for i, ch in enumerate(k):
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        elif ch == "(" and depth == 0:
            return k[:i]
    return k


def parse(csv_path: str) -> dict:
    """{kernel short name: {"dispatches": [{counter: value summed over the
    chip}, ...] in dispatch order}}."""
    per = defaultdict(lambda: defaultdict(float))
    names = {}
    with open(csv_path) as f:
        for row in csv.DictReader(f):
            d = int(row["Dispatch_Id"])
            names[d] = short_name(row["Kernel_Name"])
            per[d][row["Counter_Name"]] += float(row["Counter_Value"])
    out = defaultdict(lambda: {"dispatches": []})
    for d in sorted(per):
        out[names[d]]["dispatches"].append(dict(per[d]))
    return dict(out)


def per_step(kernels: dict, match, counter: str) -> float | None:
    """`counter` per step: over the kernels whose short name satisfies
    `match` (one dispatch each per step, e.g. one kernel per robot), the sum
    of each kernel's mean per dispatch."""
    tot, seen = 0.0, False
    for name, k in kernels.items():
        vals = [d[counter] for d in k["dispatches"] if counter in d]
        if match(name) and vals:
            tot += sum(vals) / len(vals)
            seen = True
    return tot if seen else None


# Counters are collected for these kernels only: the measured kernels and the
# known-work ones.  The training kernels (tens of thousands of dispatches per
# model) are left out; with them the mixed-robot child's passes ended without
# a counter file.
MEASURED_KERNELS = ("(learned_kernel|dense_kernel|ragged_kernel|learn_finish_kernel|fused_order_kernel|valu_kernel"
                    "|copy_kernel)")


def collect(bench_py: str, child_args: list, log, timeout: float = 300.0) -> dict | None:
    """Run the counter children; returns {pass name: parsed kernels} (a
    failed pass is left out) or None without rocprofv3."""
    prof = rocprof()
    if prof is None:
        log("counters: rocprofv3 not found; roofline from reference-order work only")
        return None
    results = {}
    for name, counters in PASSES:
        out_dir = tempfile.mkdtemp(prefix=f"bcc_pmc_{name}_", dir="/tmp")
        cmd = [prof, "--pmc", *counters, "--kernel-include-regex", MEASURED_KERNELS, "-d", out_dir, "-o", "p",
               "--output-format", "csv", "--", sys.executable, bench_py, *child_args]
        t0 = time.perf_counter()
        log_path = os.path.join(out_dir, "child.log")
        with open(log_path, "w") as lf:
            p = subprocess.Popen(cmd, cwd="/tmp", env={**os.environ, "TMPDIR": "/tmp"}, stdout=lf, stderr=lf,
                                 start_new_session=True)
            try:
                rc = p.wait(timeout=timeout)
            except subprocess.TimeoutExpired:
                os.killpg(p.pid, signal.SIGKILL)
                p.wait()
                rc = None
        found = [os.path.join(r, f) for r, _, fs in os.walk(out_dir) for f in fs if f.endswith("counter_collection.csv")]
        if rc is None:
            log(f"counters: pass {name} timed out after {timeout:.0f} s (log {log_path})")
            continue
        if not found:
            # rocprofv3 can end with a fault of its own after a complete run;
            # without a counter file the pass is lost either way
            log(f"counters: pass {name} wrote no counter file (exit {rc}, log {log_path})")
            continue
        try:
            results[name] = parse(found[0])
        except (OSError, KeyError, ValueError) as e:
            log(f"counters: pass {name}: unreadable counter file ({e})")
            continue
        shutil.rmtree(out_dir, ignore_errors=True)   # per-dispatch files of every training kernel: large
        log(f"counters: pass {name} ({' '.join(counters)}) in {time.perf_counter() - t0:.1f} s, exit {rc}")
    return results


def calibration(results: dict) -> dict:
    """Units measured on the known-work kernels."""
    cal = {}
    sq = results.get("sq")
    if sq and "valu_kernel" in sq:
        d = sq["valu_kernel"]["dispatches"]
        if len(d) >= 2 and d[0].get("SQ_INSTS_VALU_FLOPS_FP32"):
            full, half = d[0]["SQ_INSTS_VALU_FLOPS_FP32"], d[1].get("SQ_INSTS_VALU_FLOPS_FP32", 0.0)
            cal["fp32_ops_per_count"] = valu_known_flops(64) / full
            cal["half_lanes_ratio"] = half / full
            waves = VALU_BLOCKS * 4
            cal["fma_insts_per_wave"] = d[0].get("SQ_INSTS_VALU_FMA_F32", 0.0) / waves
            cal["fma_insts_expected_per_wave"] = float(VALU_CHAINS * VALU_ITERS)
    for pname, ctr in (("fetch", "FETCH_SIZE"), ("write", "WRITE_SIZE")):
        k = results.get(pname, {}).get("copy_kernel")
        if k and k["dispatches"] and k["dispatches"][0].get(ctr):
            cal[f"{ctr.lower()}_bytes_per_kib"] = COPY_BYTES / k["dispatches"][0][ctr]
    return cal


def kernel_counters(results: dict, cal: dict, match, kernel_ms: float, fp32_peak: float, hbm_peak_gbs: float) -> dict:
    """Executed FP32 rate, HBM traffic and the raw SQ counters per step of the
    kernels `match` selects, over `kernel_ms` (their hipEvent time per step,
    measured by the caller).  Missing passes leave their fields out."""
    sq = results.get("sq", {})
    out = {}
    cnt = per_step(sq, match, "SQ_INSTS_VALU_FLOPS_FP32")
    if cnt is not None and cal.get("fp32_ops_per_count"):
        flops = cnt * cal["fp32_ops_per_count"]
        out["executed_flops_per_launch"] = flops
        out["executed_tflops"] = flops / (kernel_ms / 1e3) / 1e12
        out["executed_frac"] = out["executed_tflops"] / fp32_peak
    fetch = per_step(results.get("fetch", {}), match, "FETCH_SIZE")
    write = per_step(results.get("write", {}), match, "WRITE_SIZE")
    if fetch is not None and write is not None and cal.get("fetch_size_bytes_per_kib") and \
            cal.get("write_size_bytes_per_kib"):
        fb = fetch * cal["fetch_size_bytes_per_kib"]
        wb = write * cal["write_size_bytes_per_kib"]
        out["traffic_bytes_per_launch"] = fb + wb
        out["fetch_bytes_per_launch"] = fb
        out["write_bytes_per_launch"] = wb
        out["hbm_gbs"] = (fb + wb) / (kernel_ms / 1e3) / 1e9
        out["hbm_frac"] = out["hbm_gbs"] / hbm_peak_gbs
    raw = {}
    for c in ("SQ_WAVES", "SQ_INSTS_VALU", "SQ_INSTS_VALU_FMA_F32", "SQ_INSTS_VALU_MFMA_MOPS_BF16", "SQ_WAVE_CYCLES",
              "SQ_BUSY_CYCLES", "SQ_VALU_MFMA_BUSY_CYCLES", "GRBM_GUI_ACTIVE"):
        v = per_step(sq, match, c)
        if v is not None:
            raw[c] = v
    if raw:
        out["sq_per_launch"] = raw
    out["kernels"] = sorted({n for p in results.values() for n in p if match(n)})
    return out
