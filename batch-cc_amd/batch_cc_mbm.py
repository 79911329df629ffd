"""The README run as a command: `batch_cc <robot_name> <num_environments> <num_edges>`
(scripts/batch_cc_mbm.cpp:181-207).

    python batch-cc_amd/batch_cc_mbm.py panda 100 1000 [--problems scripts/panda_problems.json]

Builds the environments from a MotionBenchMaker problem file and the edges
from the Halton sequence exactly as the reference driver does (see mbm.py),
checks every edge against every environment at resolution 32 on the GPU, and
prints what the reference prints: setup, kernel time, edges checked and
throughput (src/batch_cc.cu:264-311), then one line per edge,
"Edge i: v_0 v_1 ... " with 1 = collision (scripts/batch_cc_mbm.cpp:156-162).
Without --problems the subset shipped in tests/golden/mbm_<robot>.json is
used (enough for 100 environments).
"""
from __future__ import annotations

import argparse
import importlib.util
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def load_pkg():
    if "batch_cc_amd" in sys.modules:
        return sys.modules["batch_cc_amd"]
    spec = importlib.util.spec_from_file_location("batch_cc_amd", os.path.join(HERE, "__init__.py"),
                                                  submodule_search_locations=[HERE])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["batch_cc_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(usage="batch_cc_mbm.py <robot_name> <num_environments> <num_edges> [--problems PATH]")
    ap.add_argument("robot_name")
    ap.add_argument("num_environments", type=int)
    ap.add_argument("num_edges", type=int)
    ap.add_argument("--problems", default=None, help="MotionBenchMaker problem file (default: the shipped subset)")
    ap.add_argument("--no-matrix", action="store_true", help="print only the timing lines")
    try:
        a = ap.parse_args(argv)
    except SystemExit:
        print("Usage: ./batch_cc <robot_name> <num_environments> <num_edges>")
        return 1
    if a.robot_name not in ("panda", "fetch"):
        print(f"Unknown robot name: {a.robot_name}")
        return 255
    bcc = load_pkg()
    from batch_cc_amd import mbm
    problems = a.problems or os.path.join(os.path.dirname(HERE), "tests", "golden", f"mbm_{a.robot_name}.json")
    t0 = time.perf_counter()
    sc, edges = mbm.readme_run(problems, a.robot_name, a.num_environments, a.num_edges)
    if sc.num_envs < a.num_environments:
        print(f"note: the problem file holds {sc.num_envs} valid problems; using them all", file=sys.stderr)
    t1 = time.perf_counter()
    with bcc.Scene(sc) as scene:
        t2 = time.perf_counter()
        d_edges = bcc.DeviceBuffer.from_array(edges)
        d_out = bcc.DeviceBuffer(edges.shape[0] * sc.num_envs)
        t3 = time.perf_counter()
        ms = bcc.time_dense(a.robot_name, scene, d_edges, edges.shape[0], 32, d_out, 0, 1)
        results = d_out.to_array(np.uint8, (edges.shape[0], sc.num_envs))
    print(f"Environments Setup time: {(t2 - t1):.6f} s")
    print(f"Setup time: {(t3 - t2) + (t1 - t0):.6f} s")
    print(f"Kernel time: {int(ms * 1e6)} ns")
    print(f"Edges checked: {edges.shape[0] * sc.num_envs}")
    print(f"Throughput: {edges.shape[0] * sc.num_envs / (ms / 1e3):.6g} edges/s")
    if not a.no_matrix:
        out = sys.stdout
        for i in range(edges.shape[0]):
            out.write(f"Edge {i}: " + "".join(f"{int(v)} " for v in results[i]) + "\n")
    return 0


if __name__ == "__main__":
    sys.exit(main())
