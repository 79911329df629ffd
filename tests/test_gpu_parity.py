"""GPU parity: every layer of the gfx950 path against the CPU oracle.

Bit-exact (==) for the arithmetic hooks (sin/cos, primitive values, FK
sphere centres, sample configurations) and for the verdicts.  Verdicts that
differ from the FP64 oracle are listed with their contact margin and must lie
within 1e-4 m of contact.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROBOTS = {"panda": 0, "fetch": 1, "baxter": 2, "xarm7": 3}


def test_sincos_bit_exact(bcc, oracle, gpu):
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.uniform(-10, 10, 20000), np.linspace(-7, 7, 4001),
                        [0.0, np.pi / 2, -np.pi / 2, np.pi, -np.pi]]).astype(np.float32)
    got = bcc.debug_sincos(x)
    for i in range(x.size):
        s, c = oracle.sincos(x[i])
        assert got[i, 0] == s and got[i, 1] == c, (i, x[i])


def _prim_inputs(rng, n):
    inp = np.zeros((n, 40), np.float32)
    inp[:, 0:3] = rng.uniform(-1, 1, (n, 3))
    inp[:, 3] = rng.uniform(0.01, 0.25, n)
    inp[:, 4:7] = rng.uniform(-1, 1, (n, 3))
    inp[:, 7] = rng.uniform(0.01, 0.3, n)
    inp[:, 8:11] = rng.uniform(-1, 1, (n, 3))
    inp[:, 11:14] = rng.uniform(-0.6, 0.6, (n, 3))
    zaxis = rng.random(n) < 0.3
    inp[zaxis, 11:13] = 0.0
    inp[:, 14] = rng.uniform(0.02, 0.2, n)
    v = inp[:, 11:14]
    inp[:, 15] = np.float32(1.0 / (v * v).sum(1).astype(np.float32).astype(np.float64))
    for i in range(n):
        A = rng.standard_normal((3, 3))
        Qm, _ = np.linalg.qr(A)
        if rng.random() < 0.3:
            Qm = np.eye(3)
        inp[i, 16:19] = rng.uniform(-1, 1, 3)
        inp[i, 19:28] = np.float32(np.concatenate([Qm[:, 0], Qm[:, 1], Qm[:, 2]]))
        inp[i, 28:31] = rng.uniform(0.05, 0.3, 3)
    inp[:, 32:35] = inp[:, 16:19]
    inp[:, 35:38] = inp[:, 28:31]
    return inp


def test_primitive_values_bit_exact(bcc, oracle, gpu):
    rng = np.random.default_rng(1)
    inp = _prim_inputs(rng, 6000)
    got = bcc.debug_prims(inp)
    L = oracle.L
    for i in range(inp.shape[0]):
        p = inp[i]
        x, y, z, rp = map(float, p[0:4])
        rsq = float(np.float32(p[3] * p[3]))
        sph = np.ascontiguousarray(p[4:8])
        cap = np.ascontiguousarray(p[8:16])
        box = np.ascontiguousarray(p[16:31])
        exp = [L.orc_sphere_val(sph.ctypes.data, x, y, z, rp), L.orc_capsule_val(cap.ctypes.data, x, y, z, rp),
               L.orc_zcapsule_val(cap.ctypes.data, x, y, z, rp), L.orc_cuboid_val(box.ctypes.data, x, y, z, rsq),
               L.orc_zcuboid_val(box.ctypes.data, x, y, z, rsq)]
        assert list(got[i, :5]) == [np.float32(e) for e in exp], i
        ident = np.array_equal(box[3:12], np.float32([1, 0, 0, 0, 1, 0, 0, 0, 1]))
        if ident:   # the aabb special case is the cuboid formula, bit for bit
            assert got[i, 5] == got[i, 3]
        if cap[3] == 0 and cap[4] == 0:
            assert got[i, 1] == got[i, 2]


@pytest.mark.parametrize("robot", ["panda", "fetch", "baxter", "xarm7"])
def test_fk_sphere_centres_bit_exact(bcc, oracle, gpu, robot):
    from batch_cc_amd import workloads as W
    rng = np.random.default_rng(2)
    dim = bcc.DIMS[robot]
    s_a, s_m = W.LIMITS[robot]
    q = (rng.random((1500, dim)) * 1.2 - 0.1).astype(np.float32) * s_m + s_a
    q = q.astype(np.float32)
    got = bcc.debug_fk(robot, q)
    for i in range(q.shape[0]):
        exp = oracle.fk(ROBOTS[robot], q[i])
        assert np.array_equal(got[i], exp), (robot, i, np.abs(got[i] - exp).max())


@pytest.mark.parametrize("robot", ["panda", "fetch", "baxter", "xarm7"])
def test_sample_set_bit_exact(bcc, oracle, gpu, robot):
    from batch_cc_amd import workloads as W
    rng = np.random.default_rng(3)
    e = W.uniform_edges(robot, 300, rng)
    e[:20, 1] = e[:20, 0]                     # zero-length edges: n = 1
    e[20:60] = W.lognormal_edges(robot, 40, rng)
    for res in (32, 1, 7):
        n, cfg0, last = bcc.debug_samples(robot, e, res)
        for i in range(e.shape[0]):
            nn, d = oracle.steps(e[i, 0], e[i, 1], res)
            assert n[i] == nn
            for t in (0, 1, 13, 31):
                assert np.array_equal(cfg0[i, t], oracle.sample(e[i, 0], d, nn, t, 0))
                assert np.array_equal(last[i, t], oracle.sample(e[i, 0], d, nn, t, nn - 1))


def _compare_dense(bcc, oracle, robot, sc, edges, res=32, label=""):
    with bcc.Scene(sc) as scene:
        got = bcc.check_dense(robot, scene, edges, res)
    exp = oracle.check_dense(ROBOTS[robot], sc, edges, res)
    bad = np.argwhere(got != exp)
    assert bad.size == 0, f"{label}: {len(bad)} mismatches, first {bad[:5].tolist()}"
    return got


@pytest.mark.parametrize("cfg,scale", [("c0", 0.3), ("c1", 0.04)])
def test_dense_verdicts_match_oracle_on_benchmark_scenes(bcc, oracle, gpu, cfg, scale):
    from batch_cc_amd import workloads as W
    robot, sc, edges = W.make_config(cfg, seed=11, scale=scale)
    got = _compare_dense(bcc, oracle, robot, sc, edges, label=cfg)
    print(f"{cfg}: {got.size} pairs, collision fraction {got.mean():.3f}")


def _sparse_scene(rng, E, n_obs, size=0.08):
    from batch_cc_amd import workloads as W
    envs = []
    for _ in range(E):
        env = W.mixed_env(rng, n_obs)
        env["spheres"][:, 3] *= size / 0.2 if len(env["spheres"]) else 1
        env["cuboids"][:, 12:15] *= size / 0.3 if len(env["cuboids"]) else 1
        envs.append(env)
    import batch_cc_amd as bcc
    return bcc.PackedScene.from_envs(envs)


@pytest.mark.parametrize("robot", ["panda", "fetch", "baxter", "xarm7"])
def test_dense_verdicts_match_oracle_on_sparse_scenes(bcc, oracle, gpu, robot):
    """Few small obstacles: many free pairs run every step and every gated
    fine sweep, which the dense benchmark scenes rarely reach."""
    from batch_cc_amd import workloads as W
    rng = np.random.default_rng(4)
    sc = _sparse_scene(rng, 12, 3)
    edges = W.uniform_edges(robot, 160, rng)
    got = _compare_dense(bcc, oracle, robot, sc, edges, label=robot)
    print(f"{robot}: collision fraction {got.mean():.3f}")
    assert 0.02 < got.mean() < 0.98


def test_every_primitive_list_and_edge_cases(bcc, oracle, gpu):
    from batch_cc_amd import workloads as W
    rng = np.random.default_rng(5)
    envs = []
    for k in range(10):
        env = W.mixed_env(rng, 6)
        th = rng.uniform(-np.pi, np.pi)
        c, s = np.float32(np.cos(th)), np.float32(np.sin(th))
        A = np.linalg.qr(rng.standard_normal((3, 3)))[0].astype(np.float32)
        env["cuboids"] = np.concatenate([env["cuboids"], np.float32(
            [[*rng.uniform(-1, 1, 2), 0.6, *A[:, 0], *A[:, 1], *A[:, 2], 0.1, 0.05, 0.2]])])
        env["z_cuboids"] = np.float32([[0.5, -0.4, 0.5, c, s, 0, -s, c, 0, 0, 0, 1, 0.1, 0.2, 0.15],
                                       [-0.5, 0.4, 0.3, 1, 0, 0, 0, 1, 0, 0, 0, 1, 0.1, 0.1, 0.1]])
        v = np.float32(rng.uniform(-0.4, 0.4, 3))
        env["capsules"] = np.concatenate([env["capsules"], np.float32(
            [[*rng.uniform(-1, 1, 3), *v, 0.05, np.float32(1.0 / float(np.float32(v @ v)))]])])
        zc = W.cylinder_as_capsule(rng.uniform(-1, 1, 3), 0.07, 0.4)
        env["z_capsules"] = zc[None, :]
        env["cylinders"] = np.float32([[0, 0, 2, 0, 0, -2, 1.5, 0.25]])    # never tested
        envs.append(env)
    envs.append({})                                                      # empty environment
    envs.append({"cylinders": np.float32([[0, 0, 2, 0, 0, -2, 1.5, 0.25]])})
    sc = bcc.PackedScene.from_envs(envs)
    edges = W.uniform_edges("panda", 203, rng)                          # not a multiple of the chunk
    edges[:5, 1] = edges[:5, 0]
    got = _compare_dense(bcc, oracle, "panda", sc, edges, label="mixed")
    assert got[:, -2].sum() == oracle.check_dense(0, bcc.PackedScene.from_envs([{}]), edges).sum()
    for res in (1, 5, 64):
        _compare_dense(bcc, oracle, "panda", sc, edges[:40], res, label=f"res{res}")
    # one env, one edge
    _compare_dense(bcc, oracle, "panda", bcc.PackedScene.from_envs(envs[:1]), edges[:1], label="1x1")


def test_cpp_shim_matches_oracle(bcc, oracle, gpu, tmp_path):
    import os
    import subprocess
    from batch_cc_amd import workloads as W
    repo = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = tmp_path / "shim"
    subprocess.run(["g++", "-O1", "-std=c++17", "-I" + os.path.join(repo, "include"),
                    os.path.join(repo, "tests", "cpp", "shim_driver.cpp"), "-o", str(exe),
                    "-L" + os.path.dirname(bcc.LIB_PATH), "-lbcc", "-Wl,-rpath," + os.path.dirname(bcc.LIB_PATH)],
                   check=True)
    robot, sc, edges = W.make_config("c0", seed=21, scale=0.08)
    with open(tmp_path / "scene.bin", "wb") as f:
        f.write(np.int32(sc.num_envs).tobytes())
        for e in range(sc.num_envs):
            f.write(sc.counts[e].astype(np.int32).tobytes())
            env = sc.env(e)
            for k in bcc.KINDS:
                f.write(np.ascontiguousarray(env[k], np.float32).tobytes())
    with open(tmp_path / "edges.bin", "wb") as f:
        f.write(np.int32(edges.shape[0]).tobytes() + np.int32(7).tobytes() + edges.tobytes())
    r = subprocess.run([str(exe), "panda", str(tmp_path / "scene.bin"), str(tmp_path / "edges.bin")],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.strip().splitlines()
    got = np.array([[int(v) for v in ln.split()] for ln in lines[1:]], np.uint8)
    assert np.array_equal(got, oracle.check_dense(0, sc, edges))


def test_fp64_audit_mismatches_are_near_contact(bcc, oracle, gpu):
    from batch_cc_amd import workloads as W
    robot, sc, edges = W.make_config("c0", seed=5, scale=0.1)
    with bcc.Scene(sc) as scene:
        got = bcc.check_dense(robot, scene, edges)
    v64, mg = oracle.check_dense64(0, sc, edges)
    bad = np.argwhere(got != v64)
    for (m, e) in bad:
        print(f"FP32 vs FP64 mismatch edge {m} env {e}: gpu={got[m, e]} fp64={v64[m, e]} margin={mg[m, e]:.3e} m")
    assert np.all(np.abs(mg[got != v64]) < 1e-4)


def test_full_size_fetch_batch_is_deterministic_and_matches_sampled_pairs(bcc, oracle, gpu):
    """c1 at full size (1000 envs x 10000 edges): two runs agree bit for bit,
    and a seeded sample of pairs matches the oracle."""
    from batch_cc_amd import workloads as W
    robot, sc, edges = W.make_config("c1", seed=0, scale=1.0)
    with bcc.Scene(sc) as scene:
        a = bcc.check_dense(robot, scene, edges)
        b = bcc.check_dense(robot, scene, edges)
    assert np.array_equal(a, b)
    rng = np.random.default_rng(6)
    pe = rng.integers(0, sc.num_envs, 4000).astype(np.int32)
    pg = rng.integers(0, edges.shape[0], 4000).astype(np.int32)
    exp, _, _, _ = oracle.check_pairs(1, sc, edges, pe, pg)
    assert np.array_equal(a[pg, pe], exp)


@pytest.mark.parametrize("robot", ["panda", "fetch", "baxter", "xarm7"])
def test_both_kernel_variants_match_oracle(bcc, oracle, gpu, robot):
    """The high- and low-occupancy kernel variants (dense and ragged), forced
    in turn, give the oracle's verdicts; the automatic choice gives the same."""
    from batch_cc_amd import workloads as W
    rng = np.random.default_rng(31 + ROBOTS[robot])
    envs = [W.mixed_env(rng, int(n)) for n in rng.integers(1, 48, 12)]
    sc = bcc.PackedScene.from_envs(envs)
    edges = W.lognormal_edges(robot, 120, rng)
    exp = oracle.check_dense(ROBOTS[robot], sc, edges, threads=16)
    counts = rng.integers(0, 20, sc.num_envs)
    off = np.zeros(sc.num_envs + 1, np.int64)
    off[1:] = np.cumsum(counts)
    pick = rng.integers(0, edges.shape[0], int(off[-1]))
    env_of = np.repeat(np.arange(sc.num_envs), counts)
    try:
        with bcc.Scene(sc) as scene:
            for v in (0, 1, -1):
                bcc.debug_force_variant(v)
                print(f"{robot}: variant {v}", flush=True)
                assert np.array_equal(bcc.check_dense(robot, scene, edges), exp), v
                assert np.array_equal(bcc.check_ragged(robot, scene, off, edges[pick]), exp[pick, env_of]), v
    finally:
        bcc.debug_force_variant(-1)


@pytest.mark.parametrize("robot", ["panda", "fetch", "baxter", "xarm7"])
def test_high_resolution_matches_oracle(bcc, oracle, gpu, robot):
    """Resolution 1024: uniform edges take ~100-250 steps per pair, and in
    these sparse scenes many pairs are free and run every step.  Each step adds
    delta once to the carried configuration (linear in the step count), in the
    plain dense kernel, the learned-order dense kernel (random orders), and the
    plain and learned-order ragged kernels."""
    from batch_cc_amd import workloads as W
    rng = np.random.default_rng(21)
    sc = _sparse_scene(rng, 6, 3)
    edges = W.uniform_edges(robot, 40, rng)
    E, M = sc.num_envs, edges.shape[0]
    res = 1024
    exp = oracle.check_dense(ROBOTS[robot], sc, edges, res, threads=16)
    n = np.array([oracle.steps(edges[m, 0], edges[m, 1], res)[0] for m in range(M)])
    print(f"{robot}: steps per pair {n.min()}..{n.max()}, collision fraction {exp.mean():.3f}")
    assert n.max() > 64 and 0.02 < exp.mean() < 0.98
    o = np.array([rng.permutation(32)[:4] for _ in range(M * E)], np.uint32)
    order = (o[:, 0] | o[:, 1] << 8 | o[:, 2] << 16 | o[:, 3] << 24).astype(np.uint32)
    off = np.arange(E + 1, dtype=np.int64) * M
    ragged_edges = np.ascontiguousarray(np.tile(edges, (E, 1, 1)))
    with bcc.Scene(sc) as scene:
        got = bcc.check_dense(robot, scene, edges, res)
        de = bcc.DeviceBuffer.from_array(edges)
        do = bcc.DeviceBuffer(E * M)
        bcc.check_dense_ex(robot, scene, de, M, res, do, bcc.DeviceBuffer.from_array(order))
        got_o = do.to_array(np.uint8, (M, E))
        got_r = bcc.check_ragged(robot, scene, off, ragged_edges, res)
        dre = bcc.DeviceBuffer.from_array(ragged_edges)
        dro = bcc.DeviceBuffer(E * M)
        rorder = bcc.DeviceBuffer.from_array(np.ascontiguousarray(order.reshape(M, E).T.reshape(-1)))
        plan = bcc.RaggedPlan(robot, scene, off, dre, dro, resolution=res, d_order=rorder)
        plan.launch()
        got_ro = dro.to_array(np.uint8, (E * M,))
        plan.close()
    assert np.array_equal(got, exp)
    assert np.array_equal(got_o, exp)
    assert np.array_equal(got_r, exp.T.reshape(-1))
    assert np.array_equal(got_ro, exp.T.reshape(-1))


def test_full_size_c0_matches_the_oracle(bcc, oracle, gpu):
    """c0 at full size (panda, 100 envs x 1000 edges): every pair against the oracle."""
    from batch_cc_amd import workloads as W
    robot, sc, edges = W.make_config("c0", seed=0, scale=1.0)
    got = _compare_dense(bcc, oracle, robot, sc, edges, label="c0 full")
    print(f"c0: {got.size} pairs, collision fraction {got.mean():.4f}")


def test_full_size_c3_is_deterministic_and_matches_sampled_pairs(bcc, oracle, gpu):
    """c3 at full size (panda, 1000 envs x 100000 edges = 1e8 pairs): two
    launches agree bit for bit and a seeded sample of 4000 pairs matches the
    oracle."""
    from batch_cc_amd import workloads as W
    robot, sc, edges = W.make_config("c3", seed=0, scale=1.0)
    E, M = sc.num_envs, edges.shape[0]
    with bcc.Scene(sc) as scene:
        de = bcc.DeviceBuffer.from_array(edges)
        da, db = bcc.DeviceBuffer(E * M), bcc.DeviceBuffer(E * M)
        bcc.check_dense_device(robot, scene, de, M, 32, da)
        bcc.check_dense_device(robot, scene, de, M, 32, db)
        a = da.to_array(np.uint8, (M, E))
        b = db.to_array(np.uint8, (M, E))
    assert np.array_equal(a, b)
    rng = np.random.default_rng(7)
    pe = rng.integers(0, E, 4000).astype(np.int32)
    pg = rng.integers(0, M, 4000).astype(np.int32)
    exp, _, _, _ = oracle.check_pairs(0, sc, edges, pe, pg, threads=16)
    assert np.array_equal(a[pg, pe], exp)
    print(f"c3: {a.size} pairs, collision fraction {a.mean():.4f}")


def test_c3_learned_leg_matches_plain_and_the_oracle(bcc, oracle, gpu):
    """c3's learned leg on c3's own scenes (16 cuboids, 8 spheres, 8 cylinders
    per environment) at a tenth of its size (100 envs x 10000 edges): a model
    trained on step-0 labels of a separate edge stream orders the pairs; the
    ordered verdicts equal the plain ones, and a sample equals the oracle."""
    from batch_cc_amd import workloads as W
    robot, sc, edges = W.make_config("c3", seed=0, scale=0.1)
    E, M = sc.num_envs, edges.shape[0]
    rng = np.random.default_rng([0, 3])
    train = W.uniform_edges(robot, 20000, rng)
    n = 200_000
    pe = rng.integers(0, E, n).astype(np.int32)
    pg = rng.integers(0, train.shape[0], n).astype(np.int32)
    with bcc.Scene(sc) as scene:
        dt = bcc.DeviceBuffer.from_array(train)
        dl = bcc.label_pairs(robot, scene, dt, pe, pg, steps=1)
        model = bcc.Model(robot, seed=0)
        losses = model.train(model.features(scene, dt, pe, pg), dl, n, epochs=4, batch=8192, lr=2e-3)
        de = bcc.DeviceBuffer.from_array(edges)
        dp, do = bcc.DeviceBuffer(E * M), bcc.DeviceBuffer(E * M)
        bcc.check_dense_device(robot, scene, de, M, 32, dp)
        order = model.order_dense(scene, de, M)
        bcc.check_dense_ex(robot, scene, de, M, 32, do, order)
        plain = dp.to_array(np.uint8, (M, E))
        learned = do.to_array(np.uint8, (M, E))
    print(f"c3/10: loss {losses[0]:.3f} -> {losses[-1]:.3f}, collision fraction {plain.mean():.4f}")
    assert losses[-1] < losses[0]
    assert np.array_equal(learned, plain)
    s = np.random.default_rng(8)
    spe = s.integers(0, E, 2000).astype(np.int32)
    spg = s.integers(0, M, 2000).astype(np.int32)
    exp, _, _, _ = oracle.check_pairs(0, sc, edges, spe, spg, threads=16)
    assert np.array_equal(learned[spg, spe], exp)
