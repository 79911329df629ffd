"""Learned step ordering on ragged and mixed-robot batches: inference over a
pair list against the dense fused path, ordered ragged plans against plain
plans and the oracle for any order, and a mixed-robot group of ordered plans
(one model per robot) against the plain group."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

RID = {"panda": 0, "fetch": 1, "baxter": 2}


def _pair_env(sc, off):
    return np.repeat(np.arange(sc.num_envs, dtype=np.int32), np.diff(off))


def _pack(o):
    return (o[:, 0] | o[:, 1] << 8 | o[:, 2] << 16 | o[:, 3] << 24).astype(np.uint32)


def _ragged_batch(W, robot, rng, num_envs=24):
    envs = [W.mixed_env(rng, int(rng.integers(8, 48)), capsule_frac=0.2) for _ in range(num_envs)]
    counts = rng.integers(0, 40, num_envs)
    counts[3] = 0                                     # an environment without pairs
    off = np.zeros(num_envs + 1, np.int64)
    off[1:] = np.cumsum(counts)
    from batch_cc_amd import PackedScene
    return PackedScene.from_envs(envs), off, W.uniform_edges(robot, int(off[-1]), rng)


@pytest.mark.parametrize("robot", ["panda", "fetch", "baxter"])
def test_pair_inference_matches_dense_inference(bcc, gpu, robot):
    """bcc_model_order_pairs on the dense batch's own pair list gives the
    dense fused kernel's logits and orders bit for bit; on a ragged list
    (edge p for pair p) it gives the layer-by-layer logits of the same rows."""
    from batch_cc_amd import workloads as W
    import mlp_ref
    rng = np.random.default_rng(21)
    envs = [W.c0_env(rng, n_cub=3, n_sph=2, n_cyl=2) for _ in range(5)]
    sc = bcc.PackedScene.from_envs(envs)
    edges = W.uniform_edges(robot, 37, rng)
    M, E = edges.shape[0], sc.num_envs
    model = bcc.Model(robot, seed=7)
    p = model.params()
    o = 0
    for name, shape in mlp_ref.sizes_of(p.size):
        n = int(np.prod(shape))
        if name.startswith("b"):
            p[o:o + n] = rng.standard_normal(n).astype(np.float32) * 0.1
        o += n
    model.set_params(p)
    with bcc.Scene(sc) as scene:
        de = bcc.DeviceBuffer.from_array(edges)
        order, logits = model.logits_dense(scene, de, M)
        pair = np.arange(M * E)
        d2, l2 = model.order_pairs(scene, de, (pair % E).astype(np.int32), (pair // E).astype(np.int32),
                                   logits=True)
        assert np.array_equal(l2.view(np.uint32), logits.reshape(M * E, 32).view(np.uint32))
        assert np.array_equal(d2.to_array(np.uint32, (M * E,)), order.reshape(-1))
        # ragged: pair p = (env pe[p], edge p)
        pe = rng.integers(0, E, M).astype(np.int32)
        d3, l3 = model.order_pairs(scene, de, pe, None, logits=True)
        x = model.features(scene, de, pe, np.arange(M, dtype=np.int32))
        ref = model.logits(x, M)
        assert np.array_equal(l3.view(np.uint32), ref.view(np.uint32))
        exp = np.array([mlp_ref.spread_top4(ref[m]) for m in range(M)], np.uint32)
        assert np.array_equal(d3.to_array(np.uint32, (M,)), exp)


@pytest.mark.parametrize("robot", ["panda", "fetch", "baxter"])
def test_ordered_ragged_plan_is_exact_for_any_order(bcc, oracle, gpu, robot):
    from batch_cc_amd import workloads as W
    rng = np.random.default_rng([22, RID[robot]])
    sc, off, edges = _ragged_batch(W, robot, rng)
    P = edges.shape[0]
    exp = oracle.check_pairs(RID[robot], sc, edges, _pair_env(sc, off), np.arange(P, dtype=np.int32), threads=16)
    exp = exp[0] if isinstance(exp, tuple) else exp
    print(f"{robot}: {P} pairs over {sc.num_envs} environments, collision fraction {exp.mean():.3f}", flush=True)
    with bcc.Scene(sc) as scene:
        de = bcc.DeviceBuffer.from_array(edges)
        do = bcc.DeviceBuffer(P)
        plain = bcc.RaggedPlan(robot, scene, off, de, do)
        plain.launch()
        bcc.synchronize()
        assert np.array_equal(do.to_array(np.uint8, (P,)), exp)
        plain.close()
        orders = [_pack(np.array([rng.permutation(32)[:4] for _ in range(P)], np.uint32)) for _ in range(2)]
        orders.append(rng.integers(0, 2**32, P, dtype=np.uint64).astype(np.uint32))   # mostly invalid
        orders.append(np.full(P, 0x1f1e1d1c, np.uint32))                              # last bins
        model = bcc.Model(robot, seed=3)
        orders.append(model.order_pairs(scene, de, _pair_env(sc, off)).to_array(np.uint32, (P,)))
        for k, o in enumerate(orders):
            d_order = bcc.DeviceBuffer.from_array(o)
            bcc.debug_force_variant(k % 2)   # both compiled variants of the ordered kernel
            for tile in (0, 1):
                plan = bcc.RaggedPlan(robot, scene, off, de, do, tile_pairs=tile, d_order=d_order)
                for _ in range(2):                    # graph replays
                    do.zero()
                    plan.launch()
                    bcc.synchronize()
                    assert np.array_equal(do.to_array(np.uint8, (P,)), exp), (k, tile)
                do.zero()
                plan.launch_eager()
                bcc.synchronize()
                assert np.array_equal(do.to_array(np.uint8, (P,)), exp), (k, tile)
                plan.close()
            bcc.debug_force_variant(-1)


def test_mixed_group_of_ordered_plans_matches_plain_group(bcc, gpu):
    from batch_cc_amd import workloads as W
    parts, env_robot, global_off = W.make_mixed(seed=5, scale=0.01)
    plain_runs, ordered_runs = [], []
    keep = []
    for robot, sc, off, edges, gidx in parts:
        scene = bcc.Scene(sc)
        de = bcc.DeviceBuffer.from_array(edges)
        P = edges.shape[0]
        d_plain, d_ord = bcc.DeviceBuffer(P), bcc.DeviceBuffer(P)
        model = bcc.Model(robot, seed=11)
        d_order = model.order_pairs(scene, de, _pair_env(sc, off))
        keep += [scene, de, model, d_order]
        plain_runs.append((P, d_plain, bcc.RaggedPlan(robot, scene, off, de, d_plain)))
        ordered_runs.append((P, d_ord, bcc.RaggedPlan(robot, scene, off, de, d_ord, d_order=d_order)))
    results = []
    for runs in (plain_runs, ordered_runs):
        group = bcc.RaggedGroup([r[2] for r in runs])
        for _ in range(2):
            for r in runs:
                r[1].zero()
            group.launch()
            bcc.synchronize()
        results.append([r[1].to_array(np.uint8, (r[0],)) for r in runs])
        group.close()
    for (robot, *_), a, b in zip(parts, *results):
        assert np.array_equal(a, b), robot
        print(f"{robot}: {a.shape[0]} pairs, collision fraction {a.mean():.3f}")
    for runs in (plain_runs, ordered_runs):
        for r in runs:
            r[2].close()
    for k in keep:
        if hasattr(k, "close"):
            k.close()
