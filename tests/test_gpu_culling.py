"""Exact culling of distant obstacles in the checking kernels: obstacles of
every list kind planted at signed distances around the cull margin (and
beyond it) from a robot sphere, obstacles at 64-obstacle window boundaries
among many distant ones, huge and non-orthonormal primitives.  The kernels
must give the oracle's verdicts, which never cull.

The kernels cull a sweep only when (spheres in the sweep) x (primitives in the
environment) reaches the threshold stamped into the scene at creation, so the
one-primitive environments below would never cull at the tuned threshold:
every test runs with the threshold forced to 0 (every sweep culls, asserted
through the scene) and with culling off (a threshold no sweep reaches)."""
import contextlib
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

RID = {"panda": 0, "fetch": 1, "baxter": 2}
HERE = os.path.dirname(os.path.abspath(__file__))
ROBOTS_DIR = os.path.join(os.path.dirname(HERE), "batch-cc_amd", "robots")
# signed gap between the primitive's surface and the padded robot sphere, m;
# the kernels cull only beyond 0.01 m (plus a relative term)
OFFSETS = (-2e-3, -1e-4, -1e-6, 0.0, 1e-6, 1e-4, 2e-3, 9.9e-3, 1.01e-2, 2e-2, 5e-2)
KIND_NAMES = ("sphere", "capsule", "z_capsule", "cuboid", "z_cuboid", "aabb", "big_aabb")


CULL = {"always": 0, "never": 1 << 30}


@contextlib.contextmanager
def _scene(bcc, packed, cull):
    """Scene created under a forced culling threshold (restored afterwards)."""
    bcc.debug_cull_min(CULL[cull])
    try:
        scene = bcc.Scene(packed)
    finally:
        bcc.debug_cull_min(-1)
    try:
        assert scene.cull_min == CULL[cull]
        yield scene
    finally:
        scene.close()


def _pad(r):
    return np.float32(np.float64(np.float32(r)) + 0.005)


def _unit(rng):
    v = rng.standard_normal(3)
    return v / np.linalg.norm(v)


def _plant(kind, c, rp, off, rng):
    """(list name, record) of one primitive whose surface lies rp + off from c."""
    c = np.asarray(c, np.float64)
    if kind == "sphere":
        R = 0.05
        u = _unit(rng)
        return "spheres", np.float32([*(c + u * (rp + R + off)), R])
    if kind in ("capsule", "z_capsule"):
        cr, L = 0.04, 0.3
        if kind == "capsule":
            u = _unit(rng)
            w = np.cross(u, _unit(rng))
            w /= np.linalg.norm(w)
        else:
            a = rng.uniform(-np.pi, np.pi)
            u = np.array([np.cos(a), np.sin(a), 0.0])
            w = np.array([0.0, 0.0, 1.0])
        m = c + u * (rp + cr + off)
        v = np.float32(w * L)
        p1 = np.float32(m - w * L / 2)
        rec = np.float32([*p1, *v, cr, np.float32(1.0 / float(np.float32(v @ v)))])
        if kind == "z_capsule":
            rec[3] = rec[4] = 0.0
        return "capsules", rec
    if kind == "cuboid":
        A = np.linalg.qr(rng.standard_normal((3, 3)))[0].T      # rows = axes
        h = rng.uniform(0.05, 0.3, 3)
        k = int(rng.integers(0, 3))
        s = 1.0 if rng.random() < 0.5 else -1.0
        ctr = c - s * A[k] * (h[k] + rp + off)
        return "cuboids", np.float32([*ctr, *A[0], *A[1], *A[2], *h])
    if kind == "z_cuboid":
        a = rng.uniform(-np.pi, np.pi)
        a1 = np.array([np.cos(a), np.sin(a)])
        a2 = np.array([-np.sin(a), np.cos(a)])
        h = rng.uniform(0.05, 0.3, 3)
        if rng.random() < 0.5:
            ctr = c - np.array([*a1, 0.0]) * (h[0] + rp + off)
        else:
            ctr = c - np.array([0.0, 0.0, 1.0]) * (h[2] + rp + off)
        return "z_cuboids", np.float32([*ctr, *a1, 0, *a2, 0, 0, 0, 1, *h])
    h = np.full(3, 100.0) if kind == "big_aabb" else rng.uniform(0.05, 0.3, 3)
    k = int(rng.integers(0, 3))
    e = np.zeros(3)
    e[k] = 1.0 if rng.random() < 0.5 else -1.0
    ctr = c + e * (h[k] + rp + off)
    return "cuboids", np.float32([*ctr, 1, 0, 0, 0, 1, 0, 0, 0, 1, *h])


def _free_configs(oracle, bcc, W, robot, n, rng):
    """Configurations that collide with nothing in an empty scene (no self collision)."""
    empty = bcc.PackedScene.from_envs([{}])
    qs = []
    while len(qs) < n:
        e = W.uniform_edges(robot, 64, rng)
        e[:, 1] = e[:, 0]
        v = oracle.check_dense(RID[robot], empty, e)[:, 0]
        qs += [e[i, 0] for i in np.nonzero(v == 0)[0]]
    return np.stack(qs[:n])


def _tested_spheres(robot):
    model = json.load(open(os.path.join(ROBOTS_DIR, f"{robot}.json")))
    out = []
    for g in model["env_program"]:
        out += [g["gate"]] if g["lone"] else g["fine"]
    return out


@pytest.mark.parametrize("cull", ["always", "never"])
@pytest.mark.parametrize("robot", ["panda", "fetch", "baxter"])
def test_obstacles_at_the_cull_margin_match_oracle(bcc, oracle, gpu, robot, cull):
    from batch_cc_amd import workloads as W
    rng = np.random.default_rng([41, RID[robot]])
    qs = _free_configs(oracle, bcc, W, robot, 8, rng)
    spheres = _tested_spheres(robot)
    envs, own = [], []
    for m, q in enumerate(qs):
        fk = oracle.fk(RID[robot], q)
        for kind in KIND_NAMES:
            for off in OFFSETS:
                s = spheres[int(rng.integers(0, len(spheres)))]
                name, rec = _plant(kind, fk[s, :3], _pad(fk[s, 3]), off, rng)
                envs.append({name: rec[None, :]})
                own.append(m)
    sc = bcc.PackedScene.from_envs(envs)
    edges = np.stack([qs, qs], axis=1).astype(np.float32)     # one configuration per edge
    exp = oracle.check_dense(RID[robot], sc, edges, threads=16)
    with _scene(bcc, sc, cull) as scene:
        got = bcc.check_dense(robot, scene, edges)
        # the same pairs through a ragged plan (one configuration per pair)
        E = sc.num_envs
        off = np.arange(E + 1, dtype=np.int64) * len(qs)
        got_r = bcc.check_ragged(robot, scene, off, np.tile(edges, (E, 1, 1)))
    assert np.array_equal(got_r, exp.T.reshape(-1))
    own = np.array(own)
    mine = exp[own, np.arange(len(envs))]
    print(f"{robot}: {len(envs)} planted environments, own-configuration collisions {int(mine.sum())}, "
          f"all pairs colliding {exp.mean():.3f}")
    assert np.array_equal(got, exp)
    # every primitive planted inside the padded sphere collides with its
    # configuration (one planted just outside may still touch a neighbouring
    # sphere of the same link)
    offs = np.tile(np.array(OFFSETS), len(qs) * len(KIND_NAMES))
    assert np.all(mine[offs < -1e-5] == 1)


@pytest.mark.parametrize("cull", ["always", "never"])
def test_window_boundaries_with_many_distant_obstacles(bcc, oracle, gpu, cull):
    """Environments of 64, 128 and 130 primitives, all far away except one near
    primitive at list positions around the 64-obstacle window edges (the
    kernels' box order: spheres, capsules, z-capsules, cuboids, z-cuboids,
    axis-aligned cuboids)."""
    from batch_cc_amd import workloads as W
    rng = np.random.default_rng(43)
    robot = "panda"
    qs = _free_configs(oracle, bcc, W, robot, 3, rng)
    spheres = _tested_spheres(robot)
    layouts = {   # counts per device list, near kind per list
        130: (40, 30, 0, 10, 5, 45),
        128: (64, 0, 20, 0, 0, 44),
        64: (10, 10, 10, 10, 10, 14),
    }
    kind_of = ("sphere", "capsule", "z_capsule", "cuboid", "z_cuboid", "aabb")
    envs = []
    for total, counts in layouts.items():
        firsts = np.concatenate([[0], np.cumsum(counts)])[:-1]
        positions = sorted({p for p in (0, 62, 63, 64, 65, 126, 127, 128, 129, total - 1) if p < total}
                           | {int(f) for f, n in zip(firsts, counts) if n})
        for pos in positions:
            lst = int(np.searchsorted(firsts, pos, side="right") - 1)
            while counts[lst] == 0:
                lst += 1
            idx = pos - firsts[lst]
            for m, q in enumerate(qs):
                fk = oracle.fk(0, q)
                for off in (-1e-3, 1e-3):
                    env = {"spheres": [], "capsules": [], "z_capsules": [], "cuboids": [], "z_cuboids": []}
                    for L, n in enumerate(counts):
                        for i in range(n):
                            if L == lst and i == idx:
                                s = spheres[int(rng.integers(0, len(spheres)))]
                                name, rec = _plant(kind_of[L], fk[s, :3], _pad(fk[s, 3]), off, rng)
                            else:   # the same kind, 30-60 m away
                                far = np.array([rng.uniform(30, 60), rng.uniform(-60, 60), rng.uniform(0, 60)])
                                name, rec = _plant(kind_of[L], far, 0.01, 0.0, rng)
                            env[name].append(rec)
                    envs.append({k: np.stack(v) for k, v in env.items() if v})
    sc = bcc.PackedScene.from_envs(envs)
    edges = np.stack([qs, qs], axis=1).astype(np.float32)
    exp = oracle.check_dense(0, sc, edges, threads=16)
    with _scene(bcc, sc, cull) as scene:
        got = bcc.check_dense(robot, scene, edges)
    print(f"{len(envs)} environments, colliding pairs {int(exp.sum())}")
    assert exp.sum() > 0
    assert np.array_equal(got, exp)
    with _scene(bcc, sc, cull) as scene:   # the same pairs through a ragged plan
        E = sc.num_envs
        off = np.arange(E + 1, dtype=np.int64) * len(qs)
        ragged_edges = np.tile(edges, (E, 1, 1))
        got_r = bcc.check_ragged(robot, scene, off, ragged_edges)
    assert np.array_equal(got_r, exp.T.reshape(-1))


@pytest.mark.parametrize("robot", ["panda", "fetch", "baxter"])
def test_non_orthonormal_cuboids_are_never_culled(bcc, oracle, gpu, robot):
    """A cuboid whose axes are scaled (not a Euclidean distance) keeps the
    reference formula's verdicts: its box is unbounded, so it is always tested."""
    from batch_cc_amd import workloads as W
    rng = np.random.default_rng([44, RID[robot]])
    qs = _free_configs(oracle, bcc, W, robot, 4, rng)
    spheres = _tested_spheres(robot)
    envs = []
    for q in qs:
        fk = oracle.fk(RID[robot], q)
        for scale in (0.5, 0.9, 1.5, 3.0):
            for off in (-1e-3, 5e-3, 0.05, 0.2, 0.5):
                s = spheres[int(rng.integers(0, len(spheres)))]
                name, rec = _plant("cuboid", fk[s, :3], _pad(fk[s, 3]), off, rng)
                rec[3:12] *= np.float32(scale)
                envs.append({name: rec[None, :]})
    sc = bcc.PackedScene.from_envs(envs)
    edges = np.stack([qs, qs], axis=1).astype(np.float32)
    exp = oracle.check_dense(RID[robot], sc, edges, threads=16)
    with _scene(bcc, sc, "always") as scene:
        got = bcc.check_dense(robot, scene, edges)
    print(f"colliding pairs {int(exp.sum())} of {exp.size}")
    assert np.array_equal(got, exp)


def test_half_wave_boxes_match_numpy(bcc, gpu):
    """The culling step's cross-lane reduction: per half-wave box of the
    masked lanes' spheres, widened by 0.01 + 1e-5 * largest |coordinate|."""
    rng = np.random.default_rng(45)
    full = (1 << 64) - 1
    masks = [full, (1 << 32) - 1, full ^ ((1 << 32) - 1), 1, 1 << 63, (1 << 17) | (1 << 40), 0,
             int(rng.integers(0, 2**63)) * 2 + 1]
    masks += [int.from_bytes(rng.bytes(8), "little") for _ in range(6)]
    for k, mask in enumerate(masks):
        inp = np.zeros((64, 4), np.float32)
        inp[:, :3] = rng.uniform(-3, 3, (64, 3)) * (10.0 ** rng.integers(-1, 3))
        inp[:, 3] = rng.uniform(0.0, 0.3, 64)
        got = bcc.debug_half_boxes(inp, mask)
        for h in range(2):
            lanes = [i for i in range(32 * h, 32 * h + 32) if (mask >> i) & 1]
            if not lanes:
                assert np.all(got[h, 0] == np.inf) and np.all(got[h, 1] == -np.inf), (k, h)
                continue
            c, r = inp[lanes, :3], inp[lanes, 3:4]
            lo = (c - r).min(0).astype(np.float32)
            hi = (c + r).max(0).astype(np.float32)
            big = np.float32(max(np.abs(lo).max(), np.abs(hi).max()))
            delta = np.float32(np.float64(big) * np.float64(np.float32(1e-5)) + np.float64(np.float32(0.01)))
            assert np.allclose(got[h, 0], lo - delta, rtol=0, atol=1e-6 * (1 + np.abs(lo).max())), (k, h)
            assert np.allclose(got[h, 1], hi + delta, rtol=0, atol=1e-6 * (1 + np.abs(hi).max())), (k, h)
            # conservative: the widened box holds every masked sphere
            assert np.all(got[h, 0] <= c - r) and np.all(got[h, 1] >= c + r), (k, h)
