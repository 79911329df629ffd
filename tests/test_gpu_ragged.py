"""Ragged batches on the GPU: CSR per-environment edge lists in one
persistent launch, plain and hipGraph-replayed, against the oracle and
against the dense kernel on the same pairs."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROBOTS = {"panda": 0, "fetch": 1, "baxter": 2, "xarm7": 3}


def _ragged_case(bcc, robot, E, seed, max_edges=60):
    from batch_cc_amd import workloads as W
    rng = np.random.default_rng(seed)
    counts = rng.integers(0, max_edges, E)
    counts[0] = 0                                   # an environment without edges
    counts[-1] = max_edges * 3                      # one heavy environment
    envs = [W.mixed_env(rng, int(rng.integers(2, 20)), capsule_frac=0.2) for _ in range(E)]
    off = np.zeros(E + 1, np.int64)
    off[1:] = np.cumsum(counts)
    edges = W.uniform_edges(robot, int(off[-1]), rng)
    return bcc.PackedScene.from_envs(envs), off, edges


@pytest.mark.parametrize("robot", ["panda", "fetch", "baxter", "xarm7"])
def test_ragged_matches_oracle(bcc, oracle, gpu, robot):
    sc, off, edges = _ragged_case(bcc, robot, 24, seed=1)
    pair_env = np.repeat(np.arange(sc.num_envs, dtype=np.int32), np.diff(off).astype(np.int64))
    with bcc.Scene(sc) as scene:
        got = bcc.check_ragged(robot, scene, off, edges)
    exp, _, _, _ = oracle.check_pairs(ROBOTS[robot], sc, edges, pair_env, np.arange(off[-1], dtype=np.int32))
    assert np.array_equal(got, exp), int((got != exp).sum())


def test_ragged_equals_dense_on_the_same_pairs(bcc, gpu):
    from batch_cc_amd import workloads as W
    robot, sc, edges = W.make_config("c0", seed=4, scale=0.1)
    E, M = sc.num_envs, edges.shape[0]
    off = np.arange(E + 1, dtype=np.int64) * M
    redges = np.ascontiguousarray(np.broadcast_to(edges[None], (E, M, 2, 7)).reshape(E * M, 2, 7))
    with bcc.Scene(sc) as scene:
        dense = bcc.check_dense(robot, scene, edges)
        rag = bcc.check_ragged(robot, scene, off, redges)
    assert np.array_equal(rag.reshape(E, M).T, dense)


@pytest.mark.parametrize("tile", [1, 7, 64, 100000])
def test_graph_replays_match_eager(bcc, oracle, gpu, tile):
    sc, off, edges = _ragged_case(bcc, "baxter", 16, seed=2)
    total = int(off[-1])
    with bcc.Scene(sc) as scene:
        de = bcc.DeviceBuffer.from_array(edges)
        do = bcc.DeviceBuffer(total)
        plan = bcc.RaggedPlan("baxter", scene, off, de, do, 32, tile)
        do.zero()
        plan.launch_eager()
        bcc.synchronize()
        eager = do.to_array(np.uint8, (total,))
        for rep in range(4):
            do.zero()                                   # eager work between replays
            bcc.synchronize()
            plan.launch()
            bcc.synchronize()
            assert np.array_equal(do.to_array(np.uint8, (total,)), eager), rep
        plan.close()
    pair_env = np.repeat(np.arange(sc.num_envs, dtype=np.int32), np.diff(off))
    exp, _, _, _ = oracle.check_pairs(2, sc, edges, pair_env, np.arange(total, dtype=np.int32))
    assert np.array_equal(eager, exp)


def test_ragged_work_count_matches_oracle(bcc, oracle, gpu):
    sc, off, edges = _ragged_case(bcc, "fetch", 10, seed=3, max_edges=20)
    pair_env = np.repeat(np.arange(sc.num_envs, dtype=np.int32), np.diff(off))
    with bcc.Scene(sc) as scene:
        de = bcc.DeviceBuffer.from_array(edges)
        w, t, v = bcc.work_pairs("fetch", scene, de, pair_env)
    ev, _, ew, et = oracle.check_pairs(1, sc, edges, pair_env, np.arange(off[-1], dtype=np.int32), mode=2)
    assert np.array_equal(v, ev)
    assert np.array_equal(w, ew) and np.array_equal(t, et)


def test_dense_work_count_matches_oracle(bcc, oracle, gpu):
    from batch_cc_amd import workloads as W
    robot, sc, edges = W.make_config("c0", seed=8, scale=0.05)
    with bcc.Scene(sc) as scene:
        de = bcc.DeviceBuffer.from_array(edges)
        w, t, v = bcc.work_dense(robot, scene, de, edges.shape[0])
    ev, ew, et = oracle.work_dense(0, sc, edges)
    assert np.array_equal(v, ev) and np.array_equal(w, ew) and np.array_equal(t, et)


def test_c2_full_size_graph_is_deterministic_and_matches_sampled_pairs(bcc, oracle, gpu):
    from batch_cc_amd import workloads as W
    robot, sc, off, edges = W.make_ragged("c2", seed=0)
    total = int(off[-1])
    with bcc.Scene(sc) as scene:
        de = bcc.DeviceBuffer.from_array(edges)
        do = bcc.DeviceBuffer(total)
        plan = bcc.RaggedPlan(robot, scene, off, de, do)
        plan.launch()
        bcc.synchronize()
        a = do.to_array(np.uint8, (total,))
        plan.launch()
        bcc.synchronize()
        b = do.to_array(np.uint8, (total,))
        plan.close()
    assert np.array_equal(a, b)
    pair_env = np.repeat(np.arange(sc.num_envs, dtype=np.int32), np.diff(off))
    rng = np.random.default_rng(5)
    idx = rng.integers(0, total, 3000).astype(np.int32)
    exp, _, _, _ = oracle.check_pairs(2, sc, edges, pair_env[idx], idx)
    assert np.array_equal(a[idx], exp)
    print(f"c2: {total} pairs, collision fraction {a.mean():.3f}")


@pytest.mark.parametrize("robot", ["panda", "fetch", "baxter", "xarm7"])
def test_step_chunks_match_the_unsplit_plan(bcc, oracle, gpu, robot):
    """The tile shape's step dimension: pairs split over chunks of 1, 2, 3 or
    7 steps (graph replays and eager launches, resolution 32 and 96) give the
    unsplit plan's verdicts, which match the oracle."""
    sc, off, edges = _ragged_case(bcc, robot, 12, seed=21, max_edges=30)
    total = int(off[-1])
    pair_env = np.repeat(np.arange(sc.num_envs, dtype=np.int32), np.diff(off).astype(np.int64))
    with bcc.Scene(sc) as scene:
        de = bcc.DeviceBuffer.from_array(edges)
        do = bcc.DeviceBuffer(total)
        for res in (32, 96):
            exp, _, _, _ = oracle.check_pairs(ROBOTS[robot], sc, edges, pair_env, np.arange(total, dtype=np.int32),
                                              res)
            plain = bcc.RaggedPlan(robot, scene, off, de, do, resolution=res, tile_steps=-1)
            assert plain.tile_steps == 0
            plain.launch()
            bcc.synchronize()
            assert np.array_equal(do.to_array(np.uint8, (total,)), exp), (robot, res)
            plain.close()
            for steps in (1, 2, 3, 7):
                plan = bcc.RaggedPlan(robot, scene, off, de, do, resolution=res, tile_steps=steps, tile_pairs=4)
                assert plan.tile_steps == steps
                ones = np.ones(total, np.uint8)
                for launch in (plan.launch, plan.launch_eager, plan.launch):
                    # stale verdicts of 1: the plan's zero-fill must clear the split pairs
                    bcc._check(bcc.lib().bcc_memcpy_htod(do.ptr, ones.ctypes.data, total))
                    launch()
                    bcc.synchronize()
                    assert np.array_equal(do.to_array(np.uint8, (total,)), exp), (robot, res, steps)
                plan.close()


def test_small_batches_take_the_tuned_step_chunk(bcc, oracle, gpu):
    """tile_steps = 0: a batch with fewer pairs than the grid has half-waves
    takes the tuning table's step chunk; verdicts equal the oracle's."""
    sc, off, edges = _ragged_case(bcc, "panda", 6, seed=22, max_edges=10)
    total = int(off[-1])
    pair_env = np.repeat(np.arange(sc.num_envs, dtype=np.int32), np.diff(off).astype(np.int64))
    exp, _, _, _ = oracle.check_pairs(0, sc, edges, pair_env, np.arange(total, dtype=np.int32), 128)
    with bcc.Scene(sc) as scene:
        de = bcc.DeviceBuffer.from_array(edges)
        do = bcc.DeviceBuffer(total)
        plan = bcc.RaggedPlan("panda", scene, off, de, do, resolution=128, tile_steps=0)
        plan.launch()
        bcc.synchronize()
        print(f"{total} pairs: step chunk {plan.tile_steps}")
        assert np.array_equal(do.to_array(np.uint8, (total,)), exp)
        plan.close()
