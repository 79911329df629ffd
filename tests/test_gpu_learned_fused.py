"""The fused learned step on the GPU: one kernel probes every pair at two
evenly spaced step-0 samples, runs the model only for the pairs left
undecided, probes their predicted bins two at a time and finishes the rest
with the plain step loop.  Checked here: verdicts against the oracle for any model
(none, untrained, trained), both kernel variants and several tile sizes;
every pair's stage against the exact step-0 labels and the order word it was
given; the in-kernel inference against the separate fused inference kernel
(bit-identical logits, so the same words); ragged plans (graph replays and
eager launches) and mixed-robot groups against plain plans; and what the
trained model gains over fixed second-probe bins."""
import numpy as np
import pytest

import mlp_ref

pytestmark = pytest.mark.gpu

FIRST = (8, 24)                 # first-probe bins, (32 k + 16) / 2
FIXED = ((0, 16), (4, 20))      # later probes without a model, (32 k + 0) / 2, (32 k + 8) / 2


def _scene(W, rng, robot, n_env, n_edge, n_cub=3, n_sph=2, n_cyl=2):
    envs = [W.c0_env(rng, n_cub=n_cub, n_sph=n_sph, n_cyl=n_cyl) for _ in range(n_env)]
    return envs, W.uniform_edges(robot, n_edge, rng)


def _biased_model(bcc, robot, seed):
    """An untrained model with random biases, so its logits differ per bin."""
    model = bcc.Model(robot, seed=seed)
    p = model.params()
    rng = np.random.default_rng(seed)
    o = 0
    for name, shape in mlp_ref.sizes_of(p.size):
        n = int(np.prod(shape))
        if name.startswith("b"):
            p[o:o + n] = rng.standard_normal(n).astype(np.float32) * 0.3
        o += n
    model.set_params(p)
    return model


def _far_mask(radius):
    far = 0xffffffff
    for b in FIRST:
        for d in range(-(radius - 1), radius):
            if 0 <= b + d < 32:
                far &= ~(1 << (b + d))
    return far


@pytest.mark.parametrize("robot", ["panda", "fetch", "baxter", "xarm7"])
def test_fused_learned_step_matches_the_oracle(bcc, oracle, gpu, robot):
    from batch_cc_amd import workloads as W
    rid = bcc.ROBOTS[robot]
    rng = np.random.default_rng(31 + rid)
    envs, edges = _scene(W, rng, robot, 9, 70)
    sc = bcc.PackedScene.from_envs(envs)
    M, E = edges.shape[0], sc.num_envs
    exp = oracle.check_dense(rid, sc, edges).astype(bool)
    n = np.array([oracle.steps(edges[m, 0], edges[m, 1], 32)[0] for m in range(M)])
    with bcc.Scene(sc) as scene:
        de = bcc.DeviceBuffer.from_array(edges)
        models = {"none": None, "biased": _biased_model(bcc, robot, 3)}
        try:
            for v in (0, 1):
                bcc.debug_force_variant(v)
                for tile, passes, min_steps in ((0, 2, 2), (16, 3, 1), (100, 1, 2), (256, 3, 5)):
                    bcc.debug_learned_tile(tile)
                    bcc.debug_learned_passes(passes)
                    bcc.debug_learned_min_steps(min_steps)
                    long_ = np.broadcast_to(n[:, None] >= min_steps, exp.shape)
                    for name, model in models.items():
                        r = bcc.check_dense_learned(robot, model, scene, de, M, stage=True, stats=True)
                        assert np.array_equal(r["verdicts"], exp), (v, tile, passes, name)
                        st = r["stage"].astype(np.int64)
                        # a free pair runs all n plain steps: stage passes - 1 + n
                        plain_free = np.broadcast_to(n[:, None] + passes - 1, exp.shape)
                        assert np.all(st[~exp] == plain_free[~exp])
                        assert np.all(st <= n[:, None] + passes - 1)
                        # pairs shorter than min_steps skip inference and the later probes
                        assert not np.any((st >= 1) & (st < passes) & ~long_)
                        s = r["stats"]
                        for p in range(3):
                            assert s[p] == ((st == p).sum() if p < passes else 0), (p, passes)
                        assert s[4] == (st >= passes).sum()
                        inferred = 0 if model is None or passes == 1 else ((st != 0) & long_).sum()
                        assert s[3] == inferred, (s[3], inferred)
        finally:
            bcc.debug_force_variant(-1)
            bcc.debug_learned_tile(0)
            bcc.debug_learned_passes(0)
            bcc.debug_learned_min_steps(0)


@pytest.mark.parametrize("robot", ["panda", "fetch"])
def test_step_loop_orders_match_the_oracle_at_high_resolution(bcc, oracle, gpu, robot):
    """The learned-order kernels' step loops (panda: block order after step 0,
    plain order beyond 64 steps; fetch: plain order) at resolutions 32, 256
    and 1024, where pairs run up to ~200 steps: verdicts equal the oracle's
    at that resolution for the fused learned step and the ordered kernel, and
    a free pair still runs every one of its n steps."""
    from batch_cc_amd import workloads as W
    rid = bcc.ROBOTS[robot]
    rng = np.random.default_rng(71 + rid)
    envs, edges = _scene(W, rng, robot, 6, 30, n_cub=2, n_sph=1, n_cyl=1)
    sc = bcc.PackedScene.from_envs(envs)
    M, E = edges.shape[0], sc.num_envs
    model = _biased_model(bcc, robot, 5)
    try:
        with bcc.Scene(sc) as scene:
            de = bcc.DeviceBuffer.from_array(edges)
            do = bcc.DeviceBuffer(E * M)
            dr = bcc.DeviceBuffer(E * M)
            order = rng.permuted(np.tile(np.arange(32), (M * E, 1)), axis=1)[:, :4].astype(np.uint32)
            d_order = bcc.DeviceBuffer.from_array(order[:, 0] | order[:, 1] << 8 | order[:, 2] << 16 | order[:, 3] << 24)
            bcc.debug_learned_min_steps(1)
            for res in (32, 256, 1024):
                exp = oracle.check_dense(rid, sc, edges, res).astype(bool)
                n = np.array([oracle.steps(edges[m, 0], edges[m, 1], res)[0] for m in range(M)])
                for m in (model, None):
                    r = bcc.check_dense_learned(robot, m, scene, de, M, resolution=res, stage=True)
                    assert np.array_equal(r["verdicts"], exp), (res, m is None)
                    st = r["stage"].astype(np.int64)
                    # passes = 2: a free pair's stage is 1 + n (capped at 255)
                    free_stage = np.minimum(np.broadcast_to(n[:, None] + 1, exp.shape), 255)
                    assert np.all(st[~exp] == free_stage[~exp]), res
                bcc.check_dense_ex(robot, scene, de, M, res, do, d_order, dr)
                assert np.array_equal(do.to_array(np.uint8, (M, E)).astype(bool), exp), res
            print(f"{robot}: steps up to {n.max()} at resolution 1024, collision fraction {exp.mean():.3f}")
    finally:
        bcc.debug_learned_min_steps(0)
        model.close()


@pytest.mark.parametrize("robot", ["panda", "fetch"])
def test_locality_order_changes_no_per_pair_output(bcc, oracle, gpu, robot):
    """Dense learned launches check the edges in a locality order sorted on
    the device and scatter the results back (forced here with mode 2; large
    launches by default); every per-pair output (verdict, stage, inferred
    word) is the same as in the given order, and the verdicts are the
    oracle's."""
    from batch_cc_amd import workloads as W
    rid = bcc.ROBOTS[robot]
    rng = np.random.default_rng(91 + rid)
    envs, edges = _scene(W, rng, robot, 7, 300)
    sc = bcc.PackedScene.from_envs(envs)
    M = edges.shape[0]
    exp = oracle.check_dense(rid, sc, edges).astype(bool)
    model = _biased_model(bcc, robot, 9)
    try:
        with bcc.Scene(sc) as scene:
            de = bcc.DeviceBuffer.from_array(edges)
            bcc.debug_learned_min_steps(1)
            out = {}
            for on in (0, 2):
                bcc.debug_learned_locality(on)
                out[on] = bcc.check_dense_learned(robot, model, scene, de, M, stage=True, words=True, stats=True)
                assert np.array_equal(out[on]["verdicts"], exp), on
            for key in ("stage", "words", "stats"):
                assert np.array_equal(out[0][key], out[2][key]), key
    finally:
        bcc.debug_learned_locality(1)
        bcc.debug_learned_min_steps(0)
        model.close()


@pytest.mark.parametrize("radius", [0, 4])
def test_stages_follow_the_step0_labels_and_the_inferred_words(bcc, gpu, radius):
    """Stage 0 exactly when a first-probe bin's step-0 sample collides; the
    word of every other pair is the spread top-4 of its logits (from the
    separate fused inference kernel) among the bins not yet probed; stage 1
    exactly when one of the word's first two bins collides at step 0, stage 2
    when one of its next two does."""
    from batch_cc_amd import workloads as W
    rng = np.random.default_rng(41)
    envs, edges = _scene(W, rng, "panda", 11, 90, n_cub=4)
    sc = bcc.PackedScene.from_envs(envs)
    M, E = edges.shape[0], sc.num_envs
    far0 = _far_mask(radius)
    with bcc.Scene(sc) as scene:
        de = bcc.DeviceBuffer.from_array(edges)
        pair = np.arange(M * E)
        labels = bcc.label_pairs("panda", scene, de, (pair % E).astype(np.int32), (pair // E).astype(np.int32),
                                 steps=1).to_array(np.uint32, (M * E,)).reshape(M, E)
        model = _biased_model(bcc, "panda", 7)
        _, logits = model.logits_dense(scene, de, M)
        try:
            bcc.debug_learned_far(radius)
            bcc.debug_learned_passes(3)
            bcc.debug_learned_min_steps(1)   # every pair the first probe misses is inferred
            r = bcc.check_dense_learned("panda", model, scene, de, M, stage=True, words=True)
            r0 = bcc.check_dense_learned("panda", None, scene, de, M, stage=True)
        finally:
            bcc.debug_learned_far(0)
            bcc.debug_learned_passes(0)
            bcc.debug_learned_min_steps(0)
    bit = lambda b: (labels >> np.uint32(b)) & np.uint32(1)
    first = (bit(FIRST[0]) | bit(FIRST[1])).astype(bool)
    st, words = r["stage"], r["words"]
    assert np.array_equal(st == 0, first)
    assert np.all(words[first] == 0xffffffff)
    checked = 0
    for m, e in zip(*np.nonzero(~first)):
        w = mlp_ref.spread_top4(logits[m, e], taken0=FIRST, far0=far0)
        assert words[m, e] == w, (m, e, hex(int(words[m, e])), hex(w))
        hit1 = bool(bit(w & 31)[m, e] | bit((w >> 8) & 31)[m, e])
        hit2 = bool(bit((w >> 16) & 31)[m, e] | bit((w >> 24) & 31)[m, e])
        assert (st[m, e] == 1) == hit1, (m, e)
        assert (st[m, e] == 2) == (hit2 and not hit1), (m, e)
        checked += 1
    assert checked > 0
    # without a model the later probes take the fixed bins
    st0 = r0["stage"]
    assert np.array_equal(st0 == 0, first)
    fix1 = (bit(FIXED[0][0]) | bit(FIXED[0][1])).astype(bool)
    fix2 = (bit(FIXED[1][0]) | bit(FIXED[1][1])).astype(bool)
    assert np.array_equal(st0 == 1, ~first & fix1)
    assert np.array_equal(st0 == 2, ~first & ~fix1 & fix2)
    print(f"pairs {M * E}: first probe {first.mean():.3f}, learned probes {(st == 1).mean():.3f} + "
          f"{(st == 2).mean():.3f}, fixed probes {(st0 == 1).mean():.3f} + {(st0 == 2).mean():.3f}")


def _csr(rng, E, sizes):
    off = np.zeros(E + 1, np.int64)
    off[1:] = np.cumsum(sizes)
    return off


@pytest.mark.parametrize("robot", ["panda", "fetch", "baxter", "xarm7"])
def test_learned_plan_matches_the_plain_plan(bcc, oracle, gpu, robot):
    """Ragged batches with empty environments and environments larger than a
    tile: graph replays and eager launches of a learned plan (trained-like
    and no model) give the plain plan's verdicts, which match the oracle on a
    sample."""
    from batch_cc_amd import workloads as W
    rid = bcc.ROBOTS[robot]
    rng = np.random.default_rng(51 + rid)
    E = 10
    envs = [W.c0_env(rng, n_cub=3, n_sph=2, n_cyl=2) for _ in range(E)]
    sc = bcc.PackedScene.from_envs(envs)
    sizes = np.array([0, 1, 300, 37, 0, 600, 5, 256, 257, 90])
    off = _csr(rng, E, sizes)
    total = int(off[-1])
    edges = W.uniform_edges(robot, total, rng)
    with bcc.Scene(sc) as scene:
        de = bcc.DeviceBuffer.from_array(edges)
        dp = bcc.DeviceBuffer(total)
        plain = bcc.RaggedPlan(robot, scene, off, de, dp)
        plain.launch()
        bcc.synchronize()
        exp = dp.to_array(np.uint8, (total,))
        plain.close()
        for model in (None, _biased_model(bcc, robot, 5)):
            for tile in (0, 64, 256):
                dl = bcc.DeviceBuffer(total)
                plan = bcc.RaggedPlan(robot, scene, off, de, dl, tile_pairs=tile, learned=True, model=model)
                for launch in (plan.launch, plan.launch, plan.launch_eager):
                    dl.zero()
                    launch()
                    bcc.synchronize()
                    assert np.array_equal(dl.to_array(np.uint8, (total,)), exp), (tile, model is None)
                plan.close()
    env_of = np.repeat(np.arange(E, dtype=np.int32), sizes)
    ref, _, _, _ = oracle.check_pairs(rid, sc, edges, env_of, np.arange(total, dtype=np.int32))
    assert np.array_equal(exp, ref)


def test_mixed_group_of_learned_plans_matches_plain_group(bcc, gpu):
    from batch_cc_amd import workloads as W
    rng = np.random.default_rng(61)
    plans, bufs, exp = [], [], []
    scenes = []
    for robot in ("panda", "fetch", "baxter"):
        E = 6
        envs = [W.c0_env(rng, n_cub=3, n_sph=2, n_cyl=2) for _ in range(E)]
        sc = bcc.PackedScene.from_envs(envs)
        sizes = rng.integers(0, 400, E)
        off = _csr(rng, E, sizes)
        total = int(off[-1])
        edges = W.uniform_edges(robot, total, rng)
        scene = bcc.Scene(sc)
        scenes.append(scene)
        de = bcc.DeviceBuffer.from_array(edges)
        dp, dl = bcc.DeviceBuffer(max(total, 1)), bcc.DeviceBuffer(max(total, 1))
        p = bcc.RaggedPlan(robot, scene, off, de, dp)
        p.launch()
        bcc.synchronize()
        exp.append(dp.to_array(np.uint8, (total,)))
        p.close()
        plans.append(bcc.RaggedPlan(robot, scene, off, de, dl, learned=True, model=_biased_model(bcc, robot, 9)))
        bufs.append((dl, total, de))
    group = bcc.RaggedGroup(plans)
    for _ in range(3):
        for dl, _, _ in bufs:
            dl.zero()
        group.launch()
        bcc.synchronize()
        for (dl, total, _), e in zip(bufs, exp):
            assert np.array_equal(dl.to_array(np.uint8, (total,)), e)
    group.close()
    for p in plans:
        p.close()
    for s in scenes:
        s.close()


def test_trained_model_gain_in_the_fused_step(bcc, gpu):
    """What the trained network gains inside the fused step, on pairs it was
    not trained on: among the pairs the first probe leaves undecided, the
    share its first two predicted bins decide, against an untrained network's
    bins and the fixed bins 0 and 16.  The network is trained, as the
    benchmark's is, on pairs whose first-probe bins are free at step 0 (the
    only pairs it is asked about).  Verdicts are identical in all three."""
    from batch_cc_amd import workloads as W
    rng = np.random.default_rng(71)
    envs = [W.c0_env(rng, n_cub=4, n_sph=2, n_cyl=2) for _ in range(64)]
    sc = bcc.PackedScene.from_envs(envs)
    train_edges = W.uniform_edges("panda", 4000, rng)
    test_edges = W.uniform_edges("panda", 400, rng)
    E, M = sc.num_envs, test_edges.shape[0]
    with bcc.Scene(sc) as scene:
        dtr = bcc.DeviceBuffer.from_array(train_edges)
        big = 1 << 20
        pe = rng.integers(0, E, big).astype(np.int32)
        pg = rng.integers(0, train_edges.shape[0], big).astype(np.int32)
        lab = bcc.label_pairs("panda", scene, dtr, pe, pg, steps=1).to_array(np.uint32, (big,))
        keep = np.nonzero((lab & np.uint32((1 << FIRST[0]) | (1 << FIRST[1]))) == 0)[0][:65536]
        n = int(keep.size)
        labels = bcc.DeviceBuffer.from_array(lab[keep])
        trained = bcc.Model("panda", seed=5)
        trained.train(trained.features(scene, dtr, pe[keep], pg[keep]), labels, n, epochs=20, batch=2048,
                      lr=2e-3)
        untrained = bcc.Model("panda", seed=5)
        dte = bcc.DeviceBuffer.from_array(test_edges)
        res = {}
        for name, model in (("fixed bins", None), ("untrained", untrained), ("trained", trained)):
            r = bcc.check_dense_learned("panda", model, scene, dte, M, stage=True)
            st = r["stage"]
            undecided = st != 0
            res[name] = (r["verdicts"], float((st[undecided] == 1).mean()))
    print(f"training pairs (first-probe bins free): {n}")
    for name, (_, share) in res.items():
        print(f"{name:10s}: later probes decide {share:.4f} of the pairs the first probe left")
    for name in ("untrained", "trained"):
        assert np.array_equal(res["fixed bins"][0], res[name][0]), name
    assert res["trained"][1] > res["untrained"][1]
    assert res["trained"][1] > res["fixed bins"][1]


@pytest.mark.parametrize("robot", ["panda", "fetch", "baxter", "xarm7"])
def test_small_batches_infer_every_pair_then_run_the_ordered_kernel(bcc, oracle, gpu, robot):
    """A dense learned launch of at most learned_small_pairs() pairs and no
    per-pair diagnostics infers every pair (no model: the invalid word) and
    runs the ordered dense kernel with those words; forcing the limit below
    the batch sends the same launch to the fused kernel.  Both routes, with
    and without a model, through the check and the timing entry points, give
    the oracle's verdicts."""
    from batch_cc_amd import workloads as W
    rid = bcc.ROBOTS[robot]
    rng = np.random.default_rng(71 + rid)
    envs, edges = _scene(W, rng, robot, 9, 70)
    sc = bcc.PackedScene.from_envs(envs)
    M, E = edges.shape[0], sc.num_envs
    assert M * E <= bcc.learned_small_pairs()
    exp = oracle.check_dense(rid, sc, edges).astype(bool)
    with bcc.Scene(sc) as scene:
        de = bcc.DeviceBuffer.from_array(edges)
        model = _biased_model(bcc, robot, 5)
        d_out = bcc.DeviceBuffer(M * E)
        try:
            for limit in (-1, M * E, M * E - 1, 0):
                bcc.debug_learned_small(limit)
                for m in (model, None):
                    r = bcc.check_dense_learned(robot, m, scene, de, M)
                    assert np.array_equal(r["verdicts"], exp), (limit, m is None)
                    d_out.zero()
                    bcc.time_dense_learned(robot, m, scene, de, M, d_out, 1, 2)
                    assert np.array_equal(d_out.to_array(np.uint8, (M, E)).astype(bool), exp), (limit, m is None)
        finally:
            bcc.debug_learned_small(-1)
            model.close()
