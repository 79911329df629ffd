"""Tier C: every GPU verdict that differs from the FP64 oracle, listed with
its contact margin.

The audit kernel (bcc_audit_clearance) gives every pair's smallest gated
clearance over all of its samples in FP32.  Its value is pinned against the
FP64 oracle's margin on small batches, and its sign against the product
kernels' verdicts.  On the benchmark configurations the FP64 oracle then
decides every pair within AUDIT_EPS of contact plus a seeded random sample;
every GPU verdict that differs from FP64 must lie within AUDIT_EPS of
contact, and is printed with its margin.  The configurations run at reduced
size by default; BCC_FULL_AUDIT=1 runs them at full size (c1-c4), and
BCC_AUDIT_OUT=<dir> writes each listing there as JSON.
"""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROBOTS = {"panda": 0, "fetch": 1, "baxter": 2}
AUDIT_EPS = 1e-4          # metres: |FP32 clearance| below this goes to the FP64 oracle
FP32_CLEARANCE_TOL = 2e-5  # metres: FP32 audit value vs the FP64 margin
RANDOM_SAMPLE = 2000


def _dense(bcc, cfg, seed, scale):
    from batch_cc_amd import workloads as W
    robot, sc, edges = W.make_config(cfg, seed=seed, scale=scale)
    return robot, sc, edges


@pytest.mark.parametrize("cfg,scale", [("c0", 0.1), ("c1", 0.005)])
def test_audit_clearance_matches_the_fp64_margin(bcc, oracle, gpu, cfg, scale):
    robot, sc, edges = _dense(bcc, cfg, 7, scale)
    with bcc.Scene(sc) as scene:
        d_edges = bcc.DeviceBuffer.from_array(edges)
        clr = bcc.audit_clearance(robot, scene, d_edges, edges.shape[0])
    _, mg = oracle.check_dense64(ROBOTS[robot], sc, edges)
    diff = np.abs(clr.astype(np.float64) - mg)
    print(f"{cfg}: {clr.size} pairs, max |FP32 audit - FP64 margin| = {diff.max():.3e} m, "
          f"collision fraction {float((mg < 0).mean()):.3f}")
    assert diff.max() < FP32_CLEARANCE_TOL


def test_audit_clearance_matches_the_fp64_margin_on_a_ragged_baxter_batch(bcc, oracle, gpu):
    from batch_cc_amd import workloads as W
    robot, sc, off, edges = W.make_ragged("c2", seed=2, scale=1 / 128)
    pair_env = np.repeat(np.arange(sc.num_envs, dtype=np.int32), np.diff(off))
    keep = np.random.default_rng(3).choice(edges.shape[0], min(edges.shape[0], 3000), replace=False)
    keep.sort()
    edges, pair_env = edges[keep], pair_env[keep]
    with bcc.Scene(sc) as scene:
        d_edges = bcc.DeviceBuffer.from_array(edges)
        clr = bcc.audit_clearance(robot, scene, d_edges, edges.shape[0], pair_env)
    _, mg, _, _ = oracle.check_pairs(ROBOTS[robot], sc, edges, pair_env, np.arange(edges.shape[0], dtype=np.int32),
                                     mode=1)
    diff = np.abs(clr.astype(np.float64) - mg)
    print(f"baxter ragged: {clr.size} pairs, max |FP32 audit - FP64 margin| = {diff.max():.3e} m")
    assert diff.max() < FP32_CLEARANCE_TOL


def test_audit_sign_matches_the_product_verdicts(bcc, gpu):
    robot, sc, edges = _dense(bcc, "c3", 1, 0.01)
    with bcc.Scene(sc) as scene:
        d_edges = bcc.DeviceBuffer.from_array(edges)
        clr = bcc.audit_clearance(robot, scene, d_edges, edges.shape[0])
        got = bcc.check_dense(robot, scene, edges)
    far = np.abs(clr) >= 1e-5
    assert np.array_equal((clr[far] < 0).astype(np.uint8), got[far])
    assert 0.0 < got.mean() < 1.0


def _audit(bcc, oracle, robot, sc, edges, verdicts, clr, pair_env, pair_edge, rng):
    """FP64 verdicts of the near-contact pairs and a random sample; returns
    the listing of GPU verdicts that differ from FP64."""
    near = np.flatnonzero(np.abs(clr) < AUDIT_EPS)
    sample = rng.choice(clr.size, min(clr.size, RANDOM_SAMPLE), replace=False)
    idx = np.union1d(near, sample)
    v64, mg, _, _ = oracle.check_pairs(ROBOTS[robot], sc, edges, pair_env[idx], pair_edge[idx], mode=1, threads=16)
    bad = np.flatnonzero(v64 != verdicts[idx])
    listing = [{"pair": int(idx[k]), "env": int(pair_env[idx[k]]), "edge": int(pair_edge[idx[k]]),
                "gpu": int(verdicts[idx[k]]), "fp64": int(v64[k]), "fp64_margin_m": float(mg[k]),
                "fp32_clearance_m": float(clr[idx[k]])} for k in bad]
    in_sample_far = np.isin(idx, sample) & (np.abs(clr[idx]) >= AUDIT_EPS)
    return {"pairs": int(clr.size), "near_contact": int(near.size), "random_sample": int(sample.size),
            "fp64_checked": int(idx.size), "mismatches": listing,
            "sample_far_mismatches": int((in_sample_far & (v64 != verdicts[idx])).sum()),
            "max_abs_margin_of_mismatch": max([abs(x["fp64_margin_m"]) for x in listing], default=0.0)}


def _report(name, rep):
    print(f"Tier C {name}: {rep['pairs']} pairs, {rep['near_contact']} within {AUDIT_EPS} m of contact, "
          f"{rep['fp64_checked']} checked in FP64, {len(rep['mismatches'])} GPU/FP64 mismatches")
    for m in rep["mismatches"]:
        print(f"  pair {m['pair']} (env {m['env']}, edge {m['edge']}): gpu {m['gpu']} fp64 {m['fp64']} "
              f"margin {m['fp64_margin_m']:.3e} m (fp32 clearance {m['fp32_clearance_m']:.3e} m)")
    out = os.environ.get("BCC_AUDIT_OUT")
    if out:
        os.makedirs(out, exist_ok=True)
        with open(os.path.join(out, f"tier_c_{name}.json"), "w") as f:
            json.dump(rep, f, indent=1)
    assert rep["sample_far_mismatches"] == 0
    assert all(abs(m["fp64_margin_m"]) < AUDIT_EPS for m in rep["mismatches"])


FULL = os.environ.get("BCC_FULL_AUDIT") == "1"
CHUNK_PAIRS = 4_000_000    # pairs per audit launch (a full batch runs as several, each reporting)


def _audit_dense_chunked(bcc, robot, scene, edges, E):
    M = edges.shape[0]
    rows = max(1, CHUNK_PAIRS // max(E, 1))
    out = []
    for m0 in range(0, M, rows):
        d = bcc.DeviceBuffer.from_array(np.ascontiguousarray(edges[m0:m0 + rows]))
        out.append(bcc.audit_clearance(robot, scene, d, min(rows, M - m0)).reshape(-1))
        if M > rows:
            print(f"  audit: edges {m0}..{min(M, m0 + rows)} of {M}", flush=True)
    return np.concatenate(out)


def _audit_pairs_chunked(bcc, robot, scene, edges, pair_env):
    n = edges.shape[0]
    out = []
    for p0 in range(0, n, CHUNK_PAIRS):
        d = bcc.DeviceBuffer.from_array(np.ascontiguousarray(edges[p0:p0 + CHUNK_PAIRS]))
        out.append(bcc.audit_clearance(robot, scene, d, min(CHUNK_PAIRS, n - p0), pair_env[p0:p0 + CHUNK_PAIRS]))
        if n > CHUNK_PAIRS:
            print(f"  audit: pairs {p0}..{min(n, p0 + CHUNK_PAIRS)} of {n}", flush=True)
    return np.concatenate(out)


@pytest.mark.parametrize("cfg,scale", [("c1", 1.0 if FULL else 0.05), ("c3", 1.0 if FULL else 0.01)])
def test_tier_c_dense_configs(bcc, oracle, gpu, cfg, scale):
    robot, sc, edges = _dense(bcc, cfg, 0, scale)
    E, M = sc.num_envs, edges.shape[0]
    with bcc.Scene(sc) as scene:
        d_edges = bcc.DeviceBuffer.from_array(edges)
        d_out = bcc.DeviceBuffer(E * M)
        bcc.check_dense_device(robot, scene, d_edges, M, 32, d_out)
        bcc.synchronize()
        verdicts = d_out.to_array(np.uint8, (M * E,))
        clr = _audit_dense_chunked(bcc, robot, scene, edges, E)
    pair_edge = np.repeat(np.arange(M, dtype=np.int32), E)
    pair_env = np.tile(np.arange(E, dtype=np.int32), M)
    rep = _audit(bcc, oracle, robot, sc, edges, verdicts, clr, pair_env, pair_edge, np.random.default_rng(11))
    _report(f"{cfg}{'' if scale == 1.0 else f'_scale{scale}'}", rep)


def test_tier_c_ragged_c2(bcc, oracle, gpu):
    from batch_cc_amd import workloads as W
    robot, sc, off, edges = W.make_ragged("c2", seed=0, scale=1.0 if FULL else 1 / 16)
    pair_env = np.repeat(np.arange(sc.num_envs, dtype=np.int32), np.diff(off))
    n = edges.shape[0]
    with bcc.Scene(sc) as scene:
        d_edges = bcc.DeviceBuffer.from_array(edges)
        d_out = bcc.DeviceBuffer(n)
        plan = bcc.RaggedPlan(robot, scene, off, d_edges, d_out)
        plan.launch()
        bcc.synchronize()
        verdicts = d_out.to_array(np.uint8, (n,))
        plan.close()
        clr = _audit_pairs_chunked(bcc, robot, scene, edges, pair_env)
    rep = _audit(bcc, oracle, robot, sc, edges, verdicts, clr, pair_env, np.arange(n, dtype=np.int32),
                 np.random.default_rng(12))
    _report("c2" if FULL else "c2_scale0.0625", rep)


def test_tier_c_mixed_c4(bcc, oracle, gpu):
    from batch_cc_amd import workloads as W
    parts, _, _ = W.make_mixed(seed=0, scale=1.0 if FULL else 1 / 64)
    for k, (robot, sc, off, edges, _) in enumerate(parts):
        pair_env = np.repeat(np.arange(sc.num_envs, dtype=np.int32), np.diff(off))
        n = edges.shape[0]
        with bcc.Scene(sc) as scene:
            d_edges = bcc.DeviceBuffer.from_array(edges)
            d_out = bcc.DeviceBuffer(max(1, n))
            plan = bcc.RaggedPlan(robot, scene, off, d_edges, d_out)
            plan.launch()
            bcc.synchronize()
            verdicts = d_out.to_array(np.uint8, (n,))
            plan.close()
            clr = _audit_pairs_chunked(bcc, robot, scene, edges, pair_env)
        rep = _audit(bcc, oracle, robot, sc, edges, verdicts, clr, pair_env, np.arange(n, dtype=np.int32),
                     np.random.default_rng(13 + k))
        _report(f"c4_{robot}" + ("" if FULL else "_scale0.015625"), rep)
