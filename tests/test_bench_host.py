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
