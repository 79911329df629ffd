"""Host-side helpers of bench.py (CPU only): the perfect-bin order words of
the learned step's bound and the CPU description the baseline reports."""
import importlib.util
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def _bench():
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(os.path.dirname(HERE), "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_perfect_orders_probe_colliding_bins_first():
    b = _bench()
    rng = np.random.default_rng(0)
    labels = np.concatenate([
        rng.integers(0, 2**32, 5000, dtype=np.uint64).astype(np.uint32),       # dense masks
        (np.uint32(1) << rng.integers(0, 32, 500).astype(np.uint32)),           # one colliding bin
        np.zeros(100, np.uint32),                                               # no colliding bin
        np.full(10, 0xffffffff, np.uint32), np.array([0x80000000, 3, 0x10001], np.uint32),
    ])
    w = b.perfect_orders(labels)
    for lab, word in zip(labels.tolist(), w.tolist()):
        bins = [(word >> (8 * k)) & 0xff for k in range(4)]
        assert len(set(bins)) == 4 and max(bins) < 32, (hex(lab), bins)
        set_bits = [t for t in range(32) if (lab >> t) & 1]
        # the probe's two bins are the two lowest colliding ones, when there are
        assert bins[:min(2, len(set_bits))] == set_bits[:2], (hex(lab), bins)


def test_host_cpu_reports_a_share_of_the_cores():
    b = _bench()
    c = b.host_cpu()
    assert 1 <= c["threads"] <= c["nproc"]
    assert isinstance(c["model"], str) and c["model"]


def _counters():
    spec = importlib.util.spec_from_file_location(
        "bcc_counters", os.path.join(os.path.dirname(HERE), "batch-cc_amd", "counters.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def _write_pmc_csv(path, rows):
    import csv
    with open(path, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"])
        for r in rows:
            w.writerow(r)


def test_counter_parse_sums_instances_and_keeps_dispatch_order(tmp_path):
    C = _counters()
    p = tmp_path / "p_counter_collection.csv"
    _write_pmc_csv(p, [
        (3, "(anonymous namespace)::valu_kernel(float*, int, int, float, float)", "SQ_INSTS_VALU_FLOPS_FP32", 10),
        (3, "(anonymous namespace)::valu_kernel(float*, int, int, float, float)", "SQ_INSTS_VALU_FLOPS_FP32", 6),
        (4, "(anonymous namespace)::valu_kernel(float*, int, int, float, float)", "SQ_INSTS_VALU_FLOPS_FP32", 8),
        (7, "void bcc::dense_kernel<bcc::Robot_panda, 6, 1>(float const*, int)", "SQ_WAVES", 4),
        (9, "void bcc::dense_kernel<bcc::Robot_panda, 6, 1>(float const*, int)", "SQ_WAVES", 6),
        (8, "void bcc::ragged_kernel<bcc::Robot_fetch, 8, 1>(float const*)", "SQ_WAVES", 5),
        (10, "void bcc::ragged_kernel<bcc::Robot_baxter, 4, 1>(float const*)", "SQ_WAVES", 7),
    ])
    k = C.parse(str(p))
    assert [d["SQ_INSTS_VALU_FLOPS_FP32"] for d in k["valu_kernel"]["dispatches"]] == [16, 8]
    assert C.per_step(k, lambda n: n.startswith("bcc::dense_kernel<"), "SQ_WAVES") == 5      # mean of 4 and 6
    # one dispatch per robot per step: the per-robot means add up
    assert C.per_step(k, lambda n: n.startswith("bcc::ragged_kernel<"), "SQ_WAVES") == 12
    assert C.per_step(k, lambda n: n.startswith("bcc::nothing"), "SQ_WAVES") is None


def test_counter_calibration_and_kernel_rates():
    C = _counters()
    full = C.valu_known_flops(64)
    # a counter that counts every active lane's FP32 op once per 2 flops
    hw = {
        "sq": {"valu_kernel": {"dispatches": [{"SQ_INSTS_VALU_FLOPS_FP32": full / 2,
                                               "SQ_INSTS_VALU_FMA_F32": 8 * C.VALU_ITERS * C.VALU_BLOCKS * 4},
                                              {"SQ_INSTS_VALU_FLOPS_FP32": full / 4}]},
               "bcc::dense_kernel<bcc::Robot_panda, 6, 1>": {"dispatches": [{"SQ_INSTS_VALU_FLOPS_FP32": 5e12}]}},
        "fetch": {"copy_kernel": {"dispatches": [{"FETCH_SIZE": C.COPY_BYTES / 1024}]},
                  "bcc::dense_kernel<bcc::Robot_panda, 6, 1>": {"dispatches": [{"FETCH_SIZE": 1000.0}]}},
        "write": {"copy_kernel": {"dispatches": [{"WRITE_SIZE": C.COPY_BYTES / 2048}]},
                  "bcc::dense_kernel<bcc::Robot_panda, 6, 1>": {"dispatches": [{"WRITE_SIZE": 10.0}]}},
    }
    cal = C.calibration(hw)
    assert cal["fp32_ops_per_count"] == 2.0
    assert cal["half_lanes_ratio"] == 0.5
    assert cal["fma_insts_per_wave"] == cal["fma_insts_expected_per_wave"]
    assert cal["fetch_size_bytes_per_kib"] == 1024.0 and cal["write_size_bytes_per_kib"] == 2048.0
    c = C.kernel_counters(hw, cal, lambda n: n.startswith("bcc::dense_kernel<"), 100.0, 157.3, 8000.0)
    assert c["executed_flops_per_launch"] == 1e13
    assert abs(c["executed_tflops"] - 100.0) < 1e-9          # 1e13 flops in 0.1 s
    assert c["traffic_bytes_per_launch"] == 1000.0 * 1024 + 10.0 * 2048
    assert c["kernels"] == ["bcc::dense_kernel<bcc::Robot_panda, 6, 1>"]


def test_counter_children_rebuild_the_same_workload():
    b = _bench()
    import argparse
    a = argparse.Namespace(config="c3", seed=4, scale=0.5, train_pairs=1000, train_epochs=2, learned=False)
    args = b.child_args(a)
    assert args[args.index("--config") + 1] == "c3" and args[args.index("--seed") + 1] == "4"
    assert "--counter-child" in args and "--no-learned" in args
