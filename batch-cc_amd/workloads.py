"""Seeded synthetic scenes and edges for the benchmark configurations.

Obstacles are built as the reference's own field values (src/collision/shapes.hh):
  - axis-aligned cuboid: centre, identity axes, half-extents
    (factory::cuboid with zero Euler angles, src/collision/factory.hh:28-62);
  - sphere: centre, radius;
  - z-axis cylinder of radius r and height h centred at c: stored, as every
    reference driver does, as a *capsule* through factory::cylinder::center
    (src/collision/factory.hh:165-185, scripts/batch_cc_mbm.cpp:46-58):
    p1 = c + (0,0,h/2), p2 = c - (0,0,h/2), v = p2 - p1 (float arithmetic),
    rdv = float(1.0 / double(v.v)) (factory.hh:118-129);
  - z-axis capsule: the same record (radius and length drawn like cylinders).
Joint configurations use the reference's limits (src/Robots.hh:17-41,72-98,
128-166): q = u * s_m + s_a.

Configurations (sizes from the benchmark definition):
  c0  panda  100 envs x 1000 edges; 16 cuboids, 8 spheres, 8 cylinders per env
  c1  fetch  1000 envs x 10000 edges; U{4..64} obstacles of mixed kinds;
             lognormal edge lengths (median 0.5 rad, sigma 0.8), clipped
  c2  baxter 4096 envs, Zipf(1.2) edges per env on 1..4096, U{8..256}
             obstacles with 20% capsules (ragged)
  c3  panda  1000 envs x 100000 edges, c0 distributions
  c4  panda/fetch/baxter environments interleaved, 2^24 pairs, U{32..512}
      obstacles, about 40% colliding pairs (make_mixed)
The ragged batch c2 comes from make_ragged().
"""
from __future__ import annotations

import numpy as np

from . import DIMS, PackedScene

# Joint limits: lower = s_a, range = s_m (src/Robots.hh).
LIMITS = {
    "panda": (np.array([-2.9671, -1.8326, -2.9671, -3.1416, -2.9671, -0.0873, -2.9671], np.float32),
              np.array([5.9342, 3.6652, 5.9342, 3.2289, 5.9342, 3.9095999999999997, 5.9342], np.float32)),
    "fetch": (np.array([0.0, -1.6056, -1.221, -3.14159, -2.251, -3.14159, -2.16, -3.14159], np.float32),
              np.array([0.38615, 3.2112, 2.739, 6.28318, 4.502, 6.28318, 4.32, 6.28318], np.float32)),
    "baxter": (np.array([-1.70167993878, -2.147, -3.05417993878, -0.05, -3.059, -1.57079632679, -3.059] * 2,
                        np.float32),
               np.array([3.40335987756, 3.194, 6.10835987756, 2.6679999999999997, 6.118, 3.66479632679, 6.118] * 2,
                        np.float32)),
    # seven arm joints, then the six gripper joints the generated model reads
    "xarm7": (np.array([-2.8, -1.7, -2.8, 0.0, -2.8, 0.0, -2.8] + [0.0] * 6, np.float32),
              np.array([5.6, 3.4, 5.6, 3.0, 5.6, 3.1, 5.6] + [0.85] * 6, np.float32)),
}

CENTRE_LO = np.array([-1.0, -1.0, 0.0], np.float32)
CENTRE_HI = np.array([1.0, 1.0, 1.5], np.float32)


def f32(x):
    return np.float32(x)


def scale_cfg(robot: str, u: np.ndarray) -> np.ndarray:
    """q = u * s_m + s_a in float32 (src/Robots.hh:51-63)."""
    s_a, s_m = LIMITS[robot]
    return (u.astype(np.float32) * s_m + s_a).astype(np.float32)


def cylinder_as_capsule(c, r, h) -> np.ndarray:
    """factory::cylinder::center with zero Euler angles -> capsule record
    (x1 y1 z1 xv yv zv r rdv)."""
    c = np.asarray(c, np.float32)
    half = f32(f32(h) / f32(2.0))
    p1 = np.array([c[0], c[1], f32(c[2] + half)], np.float32)
    p2 = np.array([c[0], c[1], f32(c[2] - half)], np.float32)
    v = (p2 - p1).astype(np.float32)
    dot = f32(f32(f32(v[0] * v[0]) + f32(v[1] * v[1])) + f32(v[2] * v[2]))
    rdv = f32(1.0 / float(dot)) if dot != 0 else f32(np.inf)
    return np.array([p1[0], p1[1], p1[2], v[0], v[1], v[2], f32(r), rdv], np.float32)


def aa_cuboid(c, half) -> np.ndarray:
    c = np.asarray(c, np.float32)
    h = np.asarray(half, np.float32)
    return np.array([c[0], c[1], c[2], 1, 0, 0, 0, 1, 0, 0, 0, 1, h[0], h[1], h[2]], np.float32)


def sphere(c, r) -> np.ndarray:
    c = np.asarray(c, np.float32)
    return np.array([c[0], c[1], c[2], f32(r)], np.float32)


def _centre(rng):
    return rng.uniform(CENTRE_LO, CENTRE_HI).astype(np.float32)


def c0_env(rng, n_cub=16, n_sph=8, n_cyl=8) -> dict:
    cub = [aa_cuboid(_centre(rng), rng.uniform(0.05, 0.3, 3)) for _ in range(n_cub)]
    sph = [sphere(_centre(rng), rng.uniform(0.05, 0.2)) for _ in range(n_sph)]
    cyl = [cylinder_as_capsule(_centre(rng), rng.uniform(0.03, 0.15), rng.uniform(0.1, 0.6)) for _ in range(n_cyl)]
    return {"cuboids": np.array(cub, np.float32).reshape(-1, 15),
            "spheres": np.array(sph, np.float32).reshape(-1, 4),
            "capsules": np.array(cyl, np.float32).reshape(-1, 8)}


def mixed_env(rng, n: int, capsule_frac: float | None = None) -> dict:
    """n obstacles: kinds drawn uniformly from cuboid, sphere, cylinder,
    capsule (or with the given capsule fraction, the rest split evenly)."""
    if capsule_frac is None:
        kinds = rng.integers(0, 4, n)
    else:
        is_cap = rng.random(n) < capsule_frac
        kinds = np.where(is_cap, 3, rng.integers(0, 3, n))
    cub, sph, cap = [], [], []
    for k in kinds:
        if k == 0:
            cub.append(aa_cuboid(_centre(rng), rng.uniform(0.05, 0.3, 3)))
        elif k == 1:
            sph.append(sphere(_centre(rng), rng.uniform(0.05, 0.2)))
        else:  # cylinder (stored as capsule) or capsule: same record
            cap.append(cylinder_as_capsule(_centre(rng), rng.uniform(0.03, 0.15), rng.uniform(0.1, 0.6)))
    return {"cuboids": np.array(cub, np.float32).reshape(-1, 15),
            "spheres": np.array(sph, np.float32).reshape(-1, 4),
            "capsules": np.array(cap, np.float32).reshape(-1, 8)}


def uniform_edges(robot: str, m: int, rng) -> np.ndarray:
    dim = DIMS[robot]
    u = rng.random((m, 2, dim), dtype=np.float32)
    return scale_cfg(robot, u.reshape(-1, dim)).reshape(m, 2, dim)


def lognormal_edges(robot: str, m: int, rng, median=0.5, sigma=0.8) -> np.ndarray:
    dim = DIMS[robot]
    s_a, s_m = LIMITS[robot]
    lo, hi = s_a, (s_a + s_m).astype(np.float32)
    start = scale_cfg(robot, rng.random((m, dim), dtype=np.float32))
    d = rng.standard_normal((m, dim))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    length = rng.lognormal(np.log(median), sigma, m)
    end = np.clip(start + d * length[:, None], lo, hi).astype(np.float32)
    return np.stack([start, end], 1).astype(np.float32)


def zipf_counts(rng, n: int, a: float = 1.2, kmax: int = 4096) -> np.ndarray:
    """n draws of a Zipf(a) law truncated to 1..kmax."""
    k = np.arange(1, kmax + 1, dtype=np.float64)
    p = k ** -a
    p /= p.sum()
    return rng.choice(np.arange(1, kmax + 1), size=n, p=p).astype(np.int64)


def make_ragged(name: str, seed: int = 0, scale: float = 1.0):
    """Ragged batches: returns (robot, PackedScene, env_off[E+1], edges[total,2,dim]).
    c2: baxter, 4096 envs, Zipf(1.2) edges per env on 1..4096, U{8..256}
    obstacles per env with 20% capsules, uniform endpoints."""
    rs = np.random.default_rng([seed, 1])
    re = np.random.default_rng([seed, 2])
    if name != "c2":
        raise ValueError(f"unknown ragged config {name}")
    E = max(1, int(4096 * scale))
    counts = zipf_counts(re, E)
    envs = [mixed_env(rs, int(rs.integers(8, 257)), capsule_frac=0.2) for _ in range(E)]
    off = np.zeros(E + 1, np.int64)
    off[1:] = np.cumsum(counts)
    edges = uniform_edges("baxter", int(off[-1]), re)
    return "baxter", PackedScene.from_envs(envs), off, edges


MIXED_ROBOTS = ("panda", "fetch", "baxter")
MIXED_SIZE_SCALE = 1.0
MIXED_EDGE_MEDIAN = 0.3
MIXED_LO = np.array([-2.5, -2.5, 0.0], np.float32)
MIXED_HI = np.array([2.5, 2.5, 2.0], np.float32)
MIXED_CLEAR_RADIUS = 0.7


def _centres_clear(rng, n: int) -> np.ndarray:
    """n points uniform in MIXED_LO..MIXED_HI outside the vertical cylinder of
    radius MIXED_CLEAR_RADIUS around the z axis, where the robots' bases stand."""
    out = np.zeros((0, 3), np.float32)
    while out.shape[0] < n:
        c = rng.uniform(MIXED_LO, MIXED_HI, (2 * n, 3)).astype(np.float32)
        c = c[c[:, 0] * c[:, 0] + c[:, 1] * c[:, 1] >= np.float32(MIXED_CLEAR_RADIUS ** 2)]
        out = np.concatenate([out, c])
    return out[:n]


def scaled_env(rng, n: int, k: float) -> dict:
    """n obstacles of the mixed kinds (cuboid, sphere, cylinder, capsule) with
    every size multiplied by k, centred away from the robot bases."""
    kinds = rng.integers(0, 4, n)
    cen = _centres_clear(rng, n)
    nc, ns = int((kinds == 0).sum()), int((kinds == 1).sum())
    cub = np.zeros((nc, 15), np.float32)
    cub[:, :3] = cen[kinds == 0]
    cub[:, 3], cub[:, 7], cub[:, 11] = 1, 1, 1
    cub[:, 12:] = rng.uniform(0.05, 0.3, (nc, 3)) * k
    sph = np.zeros((ns, 4), np.float32)
    sph[:, :3] = cen[kinds == 1]
    sph[:, 3] = rng.uniform(0.05, 0.2, ns) * k
    cc = cen[kinds >= 2]
    r = rng.uniform(0.03, 0.15, len(cc)) * k
    h = rng.uniform(0.1, 0.6, len(cc)) * k
    return {"cuboids": cub, "spheres": sph, "capsules": cylinders_as_capsules(cc, r, h)}


def cylinders_as_capsules(c, r, h) -> np.ndarray:
    """cylinder_as_capsule over arrays (elementwise float32, same values)."""
    c = np.asarray(c, np.float32).reshape(-1, 3)
    half = np.asarray(h, np.float32) / np.float32(2.0)
    out = np.zeros((c.shape[0], 8), np.float32)
    out[:, :2] = c[:, :2]
    out[:, 2] = c[:, 2] + half
    vz = (c[:, 2] - half) - out[:, 2]
    out[:, 5] = vz
    dot = (np.float32(0) * np.float32(0) + np.float32(0) * np.float32(0)) + vz * vz
    with np.errstate(divide="ignore"):
        out[:, 7] = (1.0 / dot.astype(np.float64)).astype(np.float32)
    out[:, 6] = np.asarray(r, np.float32)
    return out


def make_mixed(seed: int = 0, scale: float = 1.0, total_pairs: int = 1 << 24, size_scale: float = None):
    """c4: panda, fetch and baxter environments interleaved (env g belongs to
    robot g % 3), 1024 of each at full size, U{32..512} obstacles of mixed
    kinds centred in [-2.5,2.5]^2 x [0,2] outside a 0.7 m column around the
    robot bases, with sizes
    shrinking as (32/n)^(1/3) (the occupied volume does not grow with the
    count) times a constant chosen for about 40% colliding pairs; 2^24 pairs in
    all, U{m/2..3m/2} per environment (m the mean).  Edges follow c1's law
    (uniform start, lognormal length, median 0.3 rad): with uniform endpoints
    the arms' sweep crosses the whole workspace and nearly every pair collides
    (baxter self-collides on 83% of uniform edges with no obstacles at all).

    Returns (parts, env_robot, global_off): parts is one tuple per robot
    (robot, PackedScene, env_off, edges, global_index) where global_index[i]
    is pair i's position in the interleaved batch; global_off is the CSR over
    all environments in interleaved order."""
    k0 = MIXED_SIZE_SCALE if size_scale is None else size_scale
    rs = np.random.default_rng([seed, 1])
    re = np.random.default_rng([seed, 2])
    G = max(3, 3 * int(round(1024 * scale)))
    total = max(G, int(total_pairs * scale))
    m = total / G
    counts = re.integers(max(1, int(m // 2)), max(2, int(3 * m // 2) + 1), G).astype(np.int64)
    # hit the exact total: spread the difference over the environments
    diff = total - int(counts.sum())
    sgn = 1 if diff >= 0 else -1
    q, r = divmod(abs(diff), G)
    counts += sgn * q
    counts[:r] += sgn
    assert counts.min() >= 0 and int(counts.sum()) == total
    global_off = np.zeros(G + 1, np.int64)
    global_off[1:] = np.cumsum(counts)
    env_robot = np.arange(G) % 3
    envs = []
    for g in range(G):
        n = int(rs.integers(32, 513))
        envs.append(scaled_env(rs, n, k0 * (32.0 / n) ** (1.0 / 3.0)))
    parts = []
    for r, robot in enumerate(MIXED_ROBOTS):
        ids = np.nonzero(env_robot == r)[0]
        c = counts[ids]
        off = np.zeros(len(ids) + 1, np.int64)
        off[1:] = np.cumsum(c)
        # global position of each pair: its env's global start + rank in the env
        gidx = np.repeat(global_off[ids] - off[:-1], c) + np.arange(int(off[-1]), dtype=np.int64)
        edges = lognormal_edges(robot, int(off[-1]), re, median=MIXED_EDGE_MEDIAN)
        parts.append((robot, PackedScene.from_envs([envs[g] for g in ids]), off, edges, gidx))
    return parts, env_robot, global_off


def make_config(name: str, seed: int = 0, scale: float = 1.0):
    """Returns (robot, PackedScene, edges[M,2,dim]) for c0/c1/c3, with env and
    edge counts multiplied by `scale` (tests use small scales)."""
    rs = np.random.default_rng([seed, 1])   # scene stream
    re = np.random.default_rng([seed, 2])   # edge stream
    if name == "c0":
        E, M = 100, 1000
        E, M = max(1, int(E * scale)), max(1, int(M * scale))
        envs = [c0_env(rs) for _ in range(E)]
        return "panda", PackedScene.from_envs(envs), uniform_edges("panda", M, re)
    if name == "c1":
        E, M = 1000, 10000
        E, M = max(1, int(E * scale)), max(1, int(M * scale))
        envs = [mixed_env(rs, int(rs.integers(4, 65))) for _ in range(E)]
        return "fetch", PackedScene.from_envs(envs), lognormal_edges("fetch", M, re)
    if name == "c3":
        E, M = 1000, 100000
        E, M = max(1, int(E * scale)), max(1, int(M * scale))
        envs = [c0_env(rs) for _ in range(E)]
        return "panda", PackedScene.from_envs(envs), uniform_edges("panda", M, re)
    raise ValueError(f"unknown dense config {name}")
