"""Mixed-robot batches (c4 layout at reduced size): one ragged plan per robot
replayed as one graph, against the oracle and against each plan alone; tile
size and grid do not change verdicts."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

RID = {"panda": 0, "fetch": 1, "baxter": 2}


@pytest.fixture(scope="module")
def mixed(bcc):
    from batch_cc_amd import workloads as W
    return W.make_mixed(seed=3, scale=0.02)


def _pair_env(sc, off):
    return np.repeat(np.arange(sc.num_envs, dtype=np.int32), np.diff(off))


def test_group_matches_oracle_and_plans_alone(bcc, oracle, gpu, mixed):
    parts, env_robot, global_off = mixed
    print(f"c4 at scale 0.02: {len(env_robot)} envs, {int(global_off[-1])} pairs", flush=True)
    runs = []
    for robot, sc, off, edges, gidx in parts:
        scene = bcc.Scene(sc)
        de = bcc.DeviceBuffer.from_array(edges)
        do = bcc.DeviceBuffer(edges.shape[0])
        do.zero()
        runs.append((robot, sc, off, edges, gidx, scene, de, do, bcc.RaggedPlan(robot, scene, off, de, do)))
    group = bcc.RaggedGroup([r[-1] for r in runs])
    group.launch()
    bcc.synchronize()
    grouped = [r[7].to_array(np.uint8, (r[3].shape[0],)) for r in runs]
    total = np.zeros(int(global_off[-1]), np.uint8)
    for (robot, sc, off, edges, gidx, scene, de, do, plan), v in zip(runs, grouped):
        total[gidx] = v
        # each plan alone gives the same verdicts
        do.zero()
        plan.launch()
        bcc.synchronize()
        assert np.array_equal(do.to_array(np.uint8, (edges.shape[0],)), v), robot
        # a seeded sample against the oracle
        rng = np.random.default_rng(RID[robot])
        idx = rng.integers(0, edges.shape[0], 4000).astype(np.int32)
        exp = oracle.check_pairs(RID[robot], sc, edges, _pair_env(sc, off)[idx], idx, threads=16)
        exp = exp[0] if isinstance(exp, tuple) else exp
        assert np.array_equal(v[idx], exp), robot
        print(f"{robot}: {edges.shape[0]} pairs, collision fraction {v.mean():.3f}, "
              f"tile/grid {plan.config}")
    # the interleaved view: environment g's pairs are contiguous in global order
    for g in range(len(env_robot)):
        assert global_off[g + 1] >= global_off[g]
    frac = total.mean()
    print(f"overall collision fraction {frac:.3f}")
    assert 0.2 < frac < 0.6
    group.close()
    for r in runs:
        r[-1].close()
        r[5].close()


@pytest.mark.parametrize("tile,groups", [(8, 1), (256, 8), (1, 2)])
def test_tile_size_and_grid_do_not_change_verdicts(bcc, gpu, mixed, tile, groups):
    parts, _, _ = mixed
    robot, sc, off, edges, gidx = parts[2]
    with bcc.Scene(sc) as scene:
        de = bcc.DeviceBuffer.from_array(edges)
        base = bcc.check_ragged(robot, scene, off, edges)
        do = bcc.DeviceBuffer(edges.shape[0])
        plan = bcc.RaggedPlan(robot, scene, off, de, do, tile_pairs=tile, groups_per_cu=groups)
        assert plan.config == (tile, groups)
        plan.launch()
        bcc.synchronize()
        assert np.array_equal(do.to_array(np.uint8, (edges.shape[0],)), base)
        plan.close()
