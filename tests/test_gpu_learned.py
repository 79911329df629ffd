"""Learned step ordering on the GPU: the MFMA GEMM, encoder and features
against numpy, gradients against the CPU MLP reference, training, model
files, exactness of ordered checking, and what the trained model gains."""
import os

import numpy as np
import pytest

import mlp_ref

pytestmark = pytest.mark.gpu


def _bf(rng, shape, scale=1.0):
    return mlp_ref.bf16_bits((rng.standard_normal(shape) * scale).astype(np.float32))


def _f(bits):
    return (bits.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


@pytest.mark.parametrize("ta,tb", [(False, False), (True, False), (False, True), (True, True)])
def test_mfma_gemm_matches_numpy(bcc, gpu, ta, tb):
    rng = np.random.default_rng(0)
    # the last two shapes take the split-K path (K >= 1024, few output tiles)
    for (m, n, k) in [(16, 16, 16), (70, 45, 37), (128, 256, 96), (1, 33, 5), (96, 256, 4100), (256, 32, 1024)]:
        a = _bf(rng, (k, m) if ta else (m, k))
        b = _bf(rng, (n, k) if tb else (k, n))
        c0 = rng.standard_normal((m, n)).astype(np.float32)
        A = _f(a).T if ta else _f(a)
        B = _f(b).T if tb else _f(b)
        got = bcc.debug_gemm(ta, tb, a, b, m, n, k, alpha=0.5, beta=2.0, c=c0)
        exp = 0.5 * (A @ B) + 2.0 * c0
        assert np.allclose(got, exp, rtol=1e-5, atol=1e-4 * np.sqrt(k)), (m, n, k, np.abs(got - exp).max())


def _encode_ref(env):
    out = np.zeros(64)
    for cell in range(64):
        cx, cy, cz = cell & 3, (cell >> 2) & 3, cell >> 4
        lo = np.array([-1 + 0.5 * cx, -1 + 0.5 * cy, 0.375 * cz])
        hi = lo + np.array([0.5, 0.5, 0.375])
        ext = np.array([0.5, 0.5, 0.375])
        boxes = []
        for s in env["spheres"]:
            boxes.append((s[:3] - s[3], s[:3] + s[3]))
        for c in list(env["capsules"]) + list(env["z_capsules"]):
            p, q = c[:3], c[:3] + c[3:6]
            boxes.append((np.minimum(p, q) - c[6], np.maximum(p, q) + c[6]))
        for b in list(env["cuboids"]) + list(env["z_cuboids"]):
            r = np.abs(b[3:6]) * b[12] + np.abs(b[6:9]) * b[13] + np.abs(b[9:12]) * b[14]
            boxes.append((b[:3] - r, b[:3] + r))
        best = 0.0
        for bl, bh in boxes:
            f = np.prod(np.maximum(np.minimum(bh, hi) - np.maximum(bl, lo), 0) / ext)
            best = max(best, f)
        out[cell] = best
    return out


def test_encoder_and_features_match_numpy(bcc, gpu):
    from batch_cc_amd import workloads as W
    robot, sc, edges = W.make_config("c0", seed=2, scale=0.05)
    with bcc.Scene(sc) as scene:
        enc = bcc.encode_envs(scene)
        for e in range(sc.num_envs):
            assert np.allclose(enc[e], _encode_ref(sc.env(e)), atol=1e-5)
        model = bcc.Model("panda", seed=1)
        de = bcc.DeviceBuffer.from_array(edges)
        rng = np.random.default_rng(3)
        pe = rng.integers(0, sc.num_envs, 300).astype(np.int32)
        pg = rng.integers(0, edges.shape[0], 300).astype(np.int32)
        x = model.features(scene, de, pe, pg).to_array(np.uint16, (300, 96))
    lo, rg = W.LIMITS["panda"]
    scale = np.float32(2.0 / rg.astype(np.float64))   # 2 / range, rounded once
    for i in range(300):
        q = edges[pg[i]]
        f = np.zeros(96, np.float32)
        f[:64] = enc[pe[i]]
        f[64:71] = (q[0] - lo) * scale - np.float32(1.0)
        f[71:78] = (q[1] - lo) * scale - np.float32(1.0)
        assert np.array_equal(x[i], mlp_ref.bf16_bits(f)), i


def test_gradients_match_cpu_reference(bcc, gpu):
    rng = np.random.default_rng(4)
    model = bcc.Model("panda", seed=7)
    p = model.params()
    B = 256
    xb = _bf(rng, (B, 96), 0.6)
    labels = (rng.random((B, 32)) < 0.3).astype(np.uint32)
    lab = (labels << np.arange(32, dtype=np.uint32)).sum(1).astype(np.uint32)
    dx = bcc.DeviceBuffer.from_array(xb)
    dl = bcc.DeviceBuffer.from_array(lab)
    g, loss = model.grads(dx, dl, B)
    ref_loss, ref = mlp_ref.forward_backward(p, _f(xb), lab, rounding=True)
    assert abs(loss - ref_loss) < 1e-3 * max(1.0, abs(ref_loss))
    o = 0
    for name, shape in mlp_ref.sizes_of(p.size):
        n = int(np.prod(shape))
        a, b = g[o:o + n].astype(np.float64), ref[o:o + n]
        rel = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-12)
        assert rel < 1e-2, (name, rel)
        o += n


def _training_set(bcc, scene, sc, edges, robot, n, seed):
    rng = np.random.default_rng(seed)
    pe = rng.integers(0, sc.num_envs, n).astype(np.int32)
    pg = rng.integers(0, edges.shape[0], n).astype(np.int32)
    de = bcc.DeviceBuffer.from_array(edges)
    dl = bcc.label_pairs(robot, scene, de, pe, pg)
    return de, pe, pg, dl


@pytest.mark.parametrize("steps", [0, 1, 2])
def test_labels_match_the_oracle_bin_by_bin(bcc, oracle, gpu, steps):
    """Bit t: one of the first `steps` samples of bin t collides (0: any)."""
    from batch_cc_amd import workloads as W
    robot, sc, edges = W.make_config("c0", seed=5, scale=0.05)
    with bcc.Scene(sc) as scene:
        de, pe, pg, _ = _training_set(bcc, scene, sc, edges, robot, 60, 1)
        lab = bcc.label_pairs(robot, scene, de, pe, pg, steps=steps).to_array(np.uint32, (60,))
    multi = 0
    for p in range(60):
        s, e = edges[pg[p], 0], edges[pg[p], 1]
        n, d = oracle.steps(s, e, 32)
        multi += n > 1
        bits = 0
        for t in range(32):
            for i in range(n if steps == 0 else min(n, steps)):
                if oracle.config_collides(0, sc, int(pe[p]), oracle.sample(s, d, n, t, i)):
                    bits |= 1 << t
                    break
        assert lab[p] == bits, p
    assert multi > 30


def test_training_reduces_loss_and_model_files_round_trip(bcc, gpu, tmp_path):
    from batch_cc_amd import workloads as W
    robot, sc, edges = W.make_config("c0", seed=6, scale=0.2)
    with bcc.Scene(sc) as scene:
        model = bcc.Model(robot, seed=3)
        de, pe, pg, dl = _training_set(bcc, scene, sc, edges, robot, 16384, 2)
        dx = model.features(scene, de, pe, pg)
        losses = model.train(dx, dl, 16384, epochs=12, batch=1024, lr=2e-3)
        print("losses", losses)
        assert losses[-1] < 0.8 * losses[0]
        path = str(tmp_path / "m.bin")
        model.save(path)
        m2 = bcc.Model.load(path)
        assert np.array_equal(m2.params(), model.params())
        assert np.array_equal(m2.logits(dx, 512), model.logits(dx, 512))


def test_ordered_checking_is_exact_for_any_order(bcc, oracle, gpu):
    from batch_cc_amd import workloads as W
    rng = np.random.default_rng(8)
    envs = [W.c0_env(rng, n_cub=3, n_sph=2, n_cyl=2) for _ in range(12)]
    sc = bcc.PackedScene.from_envs(envs)
    edges = W.uniform_edges("panda", 150, rng)
    E, M = sc.num_envs, edges.shape[0]
    exp = oracle.check_dense(0, sc, edges)
    n = np.array([oracle.steps(edges[m, 0], edges[m, 1], 32)[0] for m in range(M)])
    with bcc.Scene(sc) as scene:
        de = bcc.DeviceBuffer.from_array(edges)
        do = bcc.DeviceBuffer(E * M)
        dr = bcc.DeviceBuffer(E * M)
        bcc.check_dense_ex("panda", scene, de, M, 32, do, None, dr)
        plain = do.to_array(np.uint8, (M, E))
        plain_rounds = dr.to_array(np.uint8, (M, E))
        assert np.array_equal(plain, exp)
        orders = []
        for _ in range(3):   # random valid orders
            o = np.array([rng.permutation(32)[:4] for _ in range(M * E)], np.uint32)
            orders.append((o[:, 0] | o[:, 1] << 8 | o[:, 2] << 16 | o[:, 3] << 24).astype(np.uint32))
        orders.append(rng.integers(0, 2**32, M * E, dtype=np.uint64).astype(np.uint32))   # mostly invalid
        orders.append(np.full(M * E, 0x03020100, np.uint32))                              # first bins
        try:
            for v in (0, 1):   # both compiled variants of the ordered kernel
                bcc.debug_force_variant(v)
                for o in orders:
                    do.zero()
                    bcc.check_dense_ex("panda", scene, de, M, 32, do, bcc.DeviceBuffer.from_array(o), dr)
                    assert np.array_equal(do.to_array(np.uint8, (M, E)), exp), v
                    # rounds: 0 = decided by the probe, else 1 + plain steps;
                    # a free pair runs every one of its n steps
                    r = dr.to_array(np.uint8, (M, E))
                    assert np.all(r[exp == 0] == np.broadcast_to(n[:, None] + 1, exp.shape)[exp == 0])
                    assert np.all(r <= n[:, None] + 1)
        finally:
            bcc.debug_force_variant(-1)
    assert np.all(plain_rounds <= n[:, None]) and np.all(plain_rounds[plain == 0] == np.broadcast_to(
        n[:, None], plain.shape)[plain == 0])
    print(f"collision fraction {exp.mean():.3f}")


def test_trained_model_gain_over_untrained_and_plain(bcc, gpu):
    """What the trained network gains, on pairs it was not trained on.  The
    ordered kernel first probes the step-0 samples of a pair's first predicted
    bins (two per pair), several pairs per wave, then runs the plain step loop
    for pairs the probe left undecided.  Compared: plain order (32 samples per
    step), an untrained network's order, evenly spaced probe bins (no
    network) and the trained network's order.  Verdicts are identical in all;
    the trained order decides more colliding pairs in its probe than the
    untrained one and visits fewer samples per colliding pair than plain
    order."""
    from batch_cc_amd import workloads as W
    rng = np.random.default_rng(9)
    envs = [W.c0_env(rng, n_cub=4, n_sph=2, n_cyl=2) for _ in range(64)]
    sc = bcc.PackedScene.from_envs(envs)
    train_edges = W.uniform_edges("panda", 4000, rng)
    test_edges = W.uniform_edges("panda", 400, rng)
    E, M = sc.num_envs, test_edges.shape[0]
    probe = 2
    with bcc.Scene(sc) as scene:
        de, pe, pg, dl = _training_set(bcc, scene, sc, train_edges, "panda", 65536, 4)
        trained = bcc.Model("panda", seed=5)
        dx = trained.features(scene, de, pe, pg)
        trained.train(dx, dl, 65536, epochs=20, batch=2048, lr=2e-3)
        untrained = bcc.Model("panda", seed=5)
        dte = bcc.DeviceBuffer.from_array(test_edges)
        do = bcc.DeviceBuffer(E * M)
        dr = bcc.DeviceBuffer(E * M)
        none = bcc.DeviceBuffer.from_array(np.full(E * M, 0xffffffff, np.uint32))
        res = {}
        for name, order in (("plain", None), ("untrained", untrained.order_dense(scene, dte, M)),
                            ("evenly spaced", none), ("trained", trained.order_dense(scene, dte, M))):
            bcc.check_dense_ex("panda", scene, dte, M, 32, do, order, dr)
            v = do.to_array(np.uint8, (M, E))
            r = dr.to_array(np.uint8, (M, E)).astype(np.int64)[v == 1]
            if order is None:   # rounds of 32 samples up to the first colliding step
                samples = 32 * r
                hit = (r == 1).mean()
            else:               # 0: the probe decided; else 1 + plain steps
                samples = np.where(r == 0, probe, probe + 32 * (r - 1))
                hit = (r == 0).mean()
            res[name] = (v, hit, samples.mean())
    for name, (v, hit, smp) in res.items():
        print(f"{name:13s}: colliding pairs decided by the first visit {hit:.4f}, "
              f"samples visited per colliding pair {smp:.2f}")
    for name in ("untrained", "evenly spaced", "trained"):
        assert np.array_equal(res["plain"][0], res[name][0]), name
    assert res["trained"][1] > res["untrained"][1]
    assert res["trained"][2] < res["plain"][2]


@pytest.mark.parametrize("robot", ["panda", "fetch", "baxter"])
def test_fused_inference_matches_layer_by_layer(bcc, gpu, robot):
    """The fused kernel (features, three MFMA layers, order selection in one
    launch) gives the same logits, bit for bit, as the layer-by-layer GEMM path
    on the same rows, and the order words of the spread top-4 rule
    (mlp_ref.spread_top4) on those logits."""
    from batch_cc_amd import workloads as W
    rng = np.random.default_rng(12)
    envs = [W.c0_env(rng, n_cub=3, n_sph=2, n_cyl=2) for _ in range(7)]
    sc = bcc.PackedScene.from_envs(envs)
    edges = W.uniform_edges(robot, 53, rng)          # 371 pairs: the last wave tile is partial
    M, E = edges.shape[0], sc.num_envs
    model = bcc.Model(robot, seed=4)
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
        x = model.features(scene, de, (pair % E).astype(np.int32), (pair // E).astype(np.int32))
        ref = model.logits(x, M * E).reshape(M, E, 32)
        order2 = model.order_dense(scene, de, M).to_array(np.uint32, (M, E))
    assert np.array_equal(logits.view(np.uint32), ref.view(np.uint32))
    exp = np.array([[mlp_ref.spread_top4(ref[m, e]) for e in range(E)] for m in range(M)], np.uint32)
    assert np.array_equal(order, exp)
    assert np.array_equal(order2, order)
