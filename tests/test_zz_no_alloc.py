"""The allocation-free ragged check (bcc_check_ragged_noalloc) against the
ragged plan (bcc_ragged_plan_create + launch) on the same CSR batches: bit-
exact verdicts on two calls in a row, the second with a smaller batch, so a
stale tile list or tile counter left in the shared scratch array by the first
call would show.  The small batch is also checked against the oracle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROBOTS = {"panda": 0, "fetch": 1, "baxter": 2}


def _batch(bcc, robot, E, seed, max_edges):
    from batch_cc_amd import workloads as W
    rng = np.random.default_rng(seed)
    counts = rng.integers(0, max_edges, E)
    counts[0] = 0                          # an environment without pairs
    counts[E // 2] = 0
    counts[-1] = max_edges * 5             # one heavy environment
    envs = [W.mixed_env(rng, int(rng.integers(2, 40)), capsule_frac=0.2) for _ in range(E)]
    off = np.zeros(E + 1, np.int64)
    off[1:] = np.cumsum(counts)
    edges = W.uniform_edges(robot, int(off[-1]), rng)
    return bcc.PackedScene.from_envs(envs), off, edges


def _plan_verdicts(bcc, robot, scene, off, d_edges):
    total = int(off[-1])
    d_out = bcc.DeviceBuffer(total)
    d_out.zero()
    plan = bcc.RaggedPlan(robot, scene, off, d_edges, d_out)
    plan.launch()
    bcc.synchronize()
    got = d_out.to_array(np.uint8, (total,))
    plan.close()
    return got


def _noalloc_verdicts(bcc, robot, scene, off, d_edges):
    total = int(off[-1])
    d_off = bcc.DeviceBuffer.from_array(off)
    d_out = bcc.DeviceBuffer.from_array(np.full(total, 0x5A, np.uint8))   # every slot must be written
    bcc.check_ragged_noalloc(robot, scene, d_off, d_edges, total, d_out)
    bcc.synchronize()
    return d_out.to_array(np.uint8, (total,))


@pytest.mark.parametrize("robot", ["panda", "fetch", "baxter"])
def test_noalloc_matches_the_ragged_plan_on_two_calls(bcc, oracle, gpu, robot):
    for call, (E, max_edges) in enumerate([(48, 300), (9, 40)]):
        sc, off, edges = _batch(bcc, robot, E, seed=10 + call, max_edges=max_edges)
        print(f"noalloc {robot} call {call}: {E} envs, {int(off[-1])} pairs", flush=True)
        with bcc.Scene(sc) as scene:
            d_edges = bcc.DeviceBuffer.from_array(edges)
            want = _plan_verdicts(bcc, robot, scene, off, d_edges)
            got = _noalloc_verdicts(bcc, robot, scene, off, d_edges)
        assert set(np.unique(got).tolist()) <= {0, 1}
        assert np.array_equal(got, want), (call, int((got != want).sum()))
        assert 0.0 < want.mean() < 1.0
        if call == 1:
            pair_env = np.repeat(np.arange(E, dtype=np.int32), np.diff(off))
            exp, _, _, _ = oracle.check_pairs(ROBOTS[robot], sc, edges, pair_env,
                                              np.arange(off[-1], dtype=np.int32))
            assert np.array_equal(got, exp), int((got != exp).sum())


def test_noalloc_on_a_zipf_batch_matches_the_ragged_plan(bcc, gpu):
    from batch_cc_amd import workloads as W
    robot, sc, off, edges = W.make_ragged("c2", seed=3, scale=0.125)   # 512 envs, Zipf(1.2) pairs per env
    print(f"noalloc zipf: {sc.num_envs} envs, {int(off[-1])} pairs", flush=True)
    with bcc.Scene(sc) as scene:
        d_edges = bcc.DeviceBuffer.from_array(edges)
        want = _plan_verdicts(bcc, robot, scene, off, d_edges)
        for _ in range(2):
            got = _noalloc_verdicts(bcc, robot, scene, off, d_edges)
            assert np.array_equal(got, want), int((got != want).sum())


def test_noalloc_rejects_oversized_and_empty_inputs(bcc, gpu):
    sc, off, edges = _batch(bcc, "panda", 4, seed=5, max_edges=10)
    with bcc.Scene(sc) as scene:
        d_edges = bcc.DeviceBuffer.from_array(edges)
        d_off = bcc.DeviceBuffer.from_array(off)
        d_out = bcc.DeviceBuffer(int(off[-1]))
        L = bcc.lib()
        rc = L.bcc_check_ragged_noalloc(0, scene.handle, d_off.ptr, d_edges.ptr, 1 << 31, 32, d_out.ptr, None)
        assert rc != 0 and b"2^31" in L.bcc_last_error()
        rc = L.bcc_check_ragged_noalloc(0, scene.handle, None, d_edges.ptr, int(off[-1]), 32, d_out.ptr, None)
        assert rc != 0
        # no pairs: nothing to do, nothing launched
        assert L.bcc_check_ragged_noalloc(0, scene.handle, None, None, 0, 32, None, None) == 0
