"""Pins of the CPU oracle (runs without a GPU).

The reference ships no tests or golden outputs, so the oracle is pinned by:
  - golden FK vectors evaluated from the reference's own generated FK
    expressions (tests/golden/fk_*.json, see make_golden.py);
  - the published Franka Emika Panda DH parameters (joint axes);
  - closed forms and brute force for the primitive tests;
  - the closed form of the reference's sample set (src/batch_cc.cu:103-128);
  - planted scenes whose verdicts follow from geometry (padding, lane-31
    extrapolation past the edge end, cylinders ignored, self-collision).
Each check is chosen so that a plausible slip (wrong sign, dropped padding,
transposed axes, wrong sample index) fails it.
"""
import json
import math
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
ROBOT_IDS = {"panda": 0, "fetch": 1, "baxter": 2, "xarm7": 3}


def load(name):
    with open(os.path.join(GOLD, name)) as f:
        return json.load(f)


# ---------------------------------------------------------------- sin/cos

def test_sincos_within_two_ulp(oracle):
    xs = np.linspace(-4 * np.pi, 4 * np.pi, 40001).astype(np.float32)
    xs = np.concatenate([xs, np.float32([0.0, np.pi / 2, -np.pi / 2, np.pi, 3.0716, -2.9671])]).astype(np.float32)
    worst = 0.0
    for x in xs:
        s, c = oracle.sincos(x)
        ts, tc = math.sin(float(x)), math.cos(float(x))
        for got, true in ((s, ts), (c, tc)):
            ulp = float(np.spacing(np.float32(max(abs(true), 2.0 ** -24))))
            worst = max(worst, abs(float(got) - true) / ulp)
    assert worst <= 2.0, worst
    assert oracle.sincos(0.0) == (0.0, 1.0)


# ---------------------------------------------------------------- FK

TOL64 = {"panda": 1e-9, "fetch": 1e-9, "baxter": 5e-7, "xarm7": 5e-7}


@pytest.mark.parametrize("robot", ["panda", "fetch", "baxter", "xarm7"])
def test_fk_matches_reference_expressions(oracle, robot):
    g = load(f"fk_{robot}.json")
    rid = ROBOT_IDS[robot]
    radii = np.float32(g["radii"])
    C = np.array(g["centres"])            # [sphere][probe][3]
    err64 = 0.0
    err32 = 0.0
    for b, q in enumerate(g["q"]):
        o64 = oracle.fk64(rid, q)
        o32 = oracle.fk(rid, np.float32(q))
        assert np.array_equal(o32[:, 3], radii)
        err64 = max(err64, np.abs(o64[:, :3] - C[:, b, :]).max())
        err32 = max(err32, np.abs(o32[:, :3].astype(np.float64) - C[:, b, :]).max())
    assert err64 <= TOL64[robot], err64
    assert err32 <= 3e-6, err32


def _dh_axes(q):
    """Franka Emika Panda, modified DH (Craig): a, d, alpha per joint."""
    a = [0, 0, 0, 0.0825, -0.0825, 0, 0.088]
    d = [0.333, 0, 0.316, 0, 0.384, 0, 0]
    al = [0, -np.pi / 2, np.pi / 2, np.pi / 2, -np.pi / 2, np.pi / 2, np.pi / 2]
    T = np.eye(4)
    axes = []
    for i in range(7):
        ca, sa = np.cos(al[i]), np.sin(al[i])
        ct, st = np.cos(q[i]), np.sin(q[i])
        Rx = np.array([[1, 0, 0, 0], [0, ca, -sa, 0], [0, sa, ca, 0], [0, 0, 0, 1]])
        Tx = np.eye(4)
        Tx[0, 3] = a[i]
        Rz = np.array([[ct, -st, 0, 0], [st, ct, 0, 0], [0, 0, 1, 0], [0, 0, 0, 1]])
        Tz = np.eye(4)
        Tz[2, 3] = d[i]
        T = T @ Rx @ Tx @ Rz @ Tz
        axes.append((T[:3, 3].copy(), T[:3, 2].copy()))
    return axes


def test_panda_joint_axes_match_published_dh(oracle):
    rng = np.random.default_rng(7)
    for _ in range(20):
        q = rng.uniform(-2.5, 2.5, 7)
        R, t = oracle.frames64(0, q)
        for j, (p, z) in enumerate(_dh_axes(q)):
            assert np.abs(R[j][:, 2] - z).max() < 1e-9, j
            # model origin lies on the DH joint axis
            v = t[j] - p
            assert np.linalg.norm(v - np.dot(v, z) * z) < 1e-9, j


def _model(robot):
    return json.load(open(os.path.join(os.path.dirname(HERE), "batch-cc_amd", "robots", f"{robot}.json")))


def _structure_mismatch(m, s):
    """First difference between a robot model's check program and the
    membership scanned from the reference text (structure_<robot>.json), or
    None.  Compares every gate, every fine sphere and self pair in order, and
    the lone flags (a gate with no nested calls decides by itself)."""
    if len(m["spheres"]) != s["num_spheres"]:
        return f"{len(m['spheres'])} spheres vs {s['num_spheres']}"
    env = [[g["gate"], g["fine"]] for g in m["env_program"]]
    if len(env) != len(s["env_groups"]):
        return f"{len(env)} env groups vs {len(s['env_groups'])}"
    for i, (got, exp) in enumerate(zip(env, s["env_groups"])):
        if got != exp:
            return f"env group {i}: {got} vs {exp}"
        if m["env_program"][i]["lone"] != (not exp[1]):
            return f"env group {i}: lone flag"
    self_ = [[list(g["gate"]), [list(p) for p in g["fine"]]] for g in m["self_program"]]
    if len(self_) != len(s["self_groups"]):
        return f"{len(self_)} self groups vs {len(s['self_groups'])}"
    for i, (got, exp) in enumerate(zip(self_, s["self_groups"])):
        if got != exp:
            return f"self group {i}: {got} vs {exp}"
        if m["self_program"][i]["lone"] != (not exp[1]):
            return f"self group {i}: lone flag"
    return None


@pytest.mark.parametrize("robot", ["panda", "fetch", "baxter", "xarm7"])
def test_check_program_structure_matches_reference_text(bcc, robot):
    s = load(f"structure_{robot}.json")
    m = _model(robot)
    assert len(m["env_program"]) == s["env_top"]
    assert sum(len(g["fine"]) for g in m["env_program"]) == s["env_fine"]
    assert len(m["self_program"]) == s["self_top"]
    assert sum(len(g["fine"]) for g in m["self_program"]) == s["self_fine"]
    assert len(m["program"]) == s["env_top"] + s["self_top"]
    # membership of every group, in order (scanned independently of the
    # extractor that wrote the model)
    assert _structure_mismatch(m, s) is None, _structure_mismatch(m, s)


@pytest.mark.parametrize("robot", ["panda", "fetch", "baxter", "xarm7"])
def test_structure_check_catches_a_moved_fine_sphere(robot):
    """A plausible extractor slip — a fine sphere or self pair attached to the
    neighbouring gate — must fail the membership comparison."""
    import copy
    s = load(f"structure_{robot}.json")
    m = _model(robot)
    gated = [i for i, g in enumerate(m["env_program"]) if g["fine"]]
    for a, b in zip(gated, gated[1:]):
        bad = copy.deepcopy(m)
        bad["env_program"][b]["fine"].insert(0, bad["env_program"][a]["fine"].pop())
        assert _structure_mismatch(bad, s) is not None
    sgated = [i for i, g in enumerate(m["self_program"]) if g["fine"]]
    for a, b in list(zip(sgated, sgated[1:]))[:50]:
        bad = copy.deepcopy(m)
        bad["self_program"][b]["fine"].append(bad["self_program"][a]["fine"].pop())
        assert _structure_mismatch(bad, s) is not None
    # swapping a gate's two spheres, or two fine spheres within a group, too
    if gated:
        bad = copy.deepcopy(m)
        g = bad["env_program"][gated[0]]
        if len(g["fine"]) > 1:
            g["fine"][0], g["fine"][1] = g["fine"][1], g["fine"][0]
            assert _structure_mismatch(bad, s) is not None
    # exchanging two neighbouring groups (a program without gating: xarm7)
    bad = copy.deepcopy(m)
    bad["env_program"][0], bad["env_program"][1] = bad["env_program"][1], bad["env_program"][0]
    assert _structure_mismatch(bad, s) is not None
    bad = copy.deepcopy(m)
    bad["self_program"][0]["gate"] = bad["self_program"][0]["gate"][::-1]
    assert _structure_mismatch(bad, s) is not None


def test_joint_limits_match_reference_text(bcc):
    from batch_cc_amd import workloads as W
    lim = load("limits.json")
    for robot in ("panda", "fetch", "baxter", "xarm7"):
        s_a, s_m = W.LIMITS[robot]
        assert np.array_equal(s_a, np.float32(lim[robot]["s_a"]))
        assert np.array_equal(s_m, np.float32(lim[robot]["s_m"]))


# ---------------------------------------------------------------- primitives

def _rand_rot(rng):
    A = rng.standard_normal((3, 3))
    Qm, Rm = np.linalg.qr(A)
    Qm *= np.sign(np.diag(Rm))
    if np.linalg.det(Qm) < 0:
        Qm[:, 0] *= -1
    return Qm


def test_sphere_and_self_tests_vs_distance(oracle):
    rng = np.random.default_rng(1)
    for _ in range(4000):
        o = np.float32(np.r_[rng.uniform(-1, 1, 3), rng.uniform(0.01, 0.3)])
        p = np.float32(rng.uniform(-1, 1, 3))
        rp = np.float32(rng.uniform(0.01, 0.2))
        v = oracle.L.orc_sphere_val(o.ctypes.data, *map(float, p), float(rp))
        d = np.linalg.norm(p.astype(np.float64) - o[:3]) - (float(rp) + float(o[3]))
        if abs(d) > 1e-5:
            assert (v < 0) == (d < 0)
        a = np.float32(np.r_[p, rp])
        assert oracle.L.orc_self_val(a.ctypes.data, o.ctypes.data) == np.float32(v)


def _seg_dist_brute(p, a, v):
    ts = np.linspace(0.0, 1.0, 200001)
    pts = a[None, :] + ts[:, None] * v[None, :]
    return np.min(np.linalg.norm(pts - p[None, :], axis=1))


def test_capsule_tests_vs_brute_force(oracle):
    rng = np.random.default_rng(2)
    for _ in range(300):
        p1 = rng.uniform(-1, 1, 3)
        v = rng.uniform(-0.6, 0.6, 3)
        dot = np.float32(np.dot(np.float32(v), np.float32(v)))
        cap = np.float32(np.r_[p1, v, rng.uniform(0.02, 0.2), 0.0])
        cap[7] = np.float32(1.0 / float(dot))
        p = np.float32(rng.uniform(-1, 1, 3))
        rp = np.float32(rng.uniform(0.01, 0.2))
        val = oracle.L.orc_capsule_val(cap.ctypes.data, *map(float, p), float(rp))
        d = _seg_dist_brute(p.astype(np.float64), cap[:3].astype(np.float64), cap[3:6].astype(np.float64))
        d -= float(rp) + float(cap[6])
        if abs(d) > 1e-4:
            assert (val < 0) == (d < 0), (val, d)
        d64 = oracle.L.orc_capsule_dist64(cap.ctypes.data, *map(float, p), float(rp), 0)
        assert abs(d64 - d) < 1e-5


def test_z_capsule_is_the_general_formula_on_z_axes(oracle):
    rng = np.random.default_rng(3)
    for _ in range(2000):
        cap = np.float32([*rng.uniform(-1, 1, 3), 0.0, 0.0, rng.uniform(-0.6, 0.6), rng.uniform(0.02, 0.2), 0])
        cap[7] = np.float32(1.0 / float(np.float32(cap[5] * cap[5])))
        p = np.float32(rng.uniform(-1, 1, 3))
        rp = float(np.float32(rng.uniform(0.01, 0.2)))
        a = oracle.L.orc_capsule_val(cap.ctypes.data, *map(float, p), rp)
        b = oracle.L.orc_zcapsule_val(cap.ctypes.data, *map(float, p), rp)
        assert a == b


def test_degenerate_capsule_is_a_sphere_at_p1(oracle):
    cap = np.float32([0.1, 0.2, 0.3, 0, 0, 0, 0.05, np.inf])
    sph = np.float32([0.1, 0.2, 0.3, 0.05])
    for p in ([0.1, 0.2, 0.4], [0.3, 0.2, 0.3], [0.0, 0.0, 0.0]):
        a = oracle.L.orc_capsule_val(cap.ctypes.data, *p, 0.02)
        b = oracle.L.orc_sphere_val(sph.ctypes.data, *p, 0.02)
        assert np.isfinite(a) and abs(a - b) < 1e-6


def _box_dist_opt(p, c, R, h):
    from scipy.optimize import minimize
    f = lambda u: np.sum((c + R @ u - p) ** 2)
    g = lambda u: 2 * R.T @ (c + R @ u - p)
    r = minimize(f, np.zeros(3), jac=g, bounds=[(-hh, hh) for hh in h], method="L-BFGS-B",
                 options={"ftol": 1e-15, "gtol": 1e-12})
    return math.sqrt(max(r.fun, 0.0))


def test_cuboid_tests_vs_optimisation(oracle):
    rng = np.random.default_rng(4)
    for i in range(300):
        R = _rand_rot(rng)
        c = rng.uniform(-1, 1, 3)
        h = rng.uniform(0.05, 0.3, 3)
        box = np.float32(np.r_[c, R[:, 0], R[:, 1], R[:, 2], h])
        p = np.float32(rng.uniform(-1, 1, 3))
        rp = np.float32(rng.uniform(0.01, 0.2))
        val = oracle.L.orc_cuboid_val(box.ctypes.data, *map(float, p), float(rp * rp))
        d = _box_dist_opt(p.astype(np.float64), box[:3].astype(np.float64),
                          np.stack([box[3:6], box[6:9], box[9:12]], 1).astype(np.float64),
                          box[12:15].astype(np.float64)) - float(rp)
        if abs(d) > 1e-4:
            assert (val < 0) == (d < 0), (i, val, d)
    # inside the box always collides
    box = np.float32([0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0, 1, 0.2, 0.2, 0.2])
    assert oracle.L.orc_cuboid_val(box.ctypes.data, 0.05, -0.1, 0.1, 1e-6) < 0
    # on the x axis outside: value = (|x| - h)^2 - r^2
    v = oracle.L.orc_cuboid_val(box.ctypes.data, 0.5, 0.0, 0.0, 0.01)
    assert abs(v - (0.3 ** 2 - 0.01)) < 1e-6


def test_z_cuboid_is_the_general_formula_on_z_boxes(oracle):
    rng = np.random.default_rng(5)
    for _ in range(2000):
        th = rng.uniform(-np.pi, np.pi)
        ct, st = np.float32(np.cos(th)), np.float32(np.sin(th))
        box = np.float32([*rng.uniform(-1, 1, 3), ct, st, 0, -st, ct, 0, 0, 0, 1, *rng.uniform(0.05, 0.3, 3)])
        p = np.float32(rng.uniform(-1, 1, 3))
        rsq = float(np.float32(rng.uniform(0.01, 0.2)) ** 2)
        a = oracle.L.orc_cuboid_val(box.ctypes.data, *map(float, p), rsq)
        b = oracle.L.orc_zcuboid_val(box.ctypes.data, *map(float, p), rsq)
        d = oracle.L.orc_cuboid_dist64(box.ctypes.data, *map(float, p), math.sqrt(rsq), 0)
        if abs(d) > 1e-5:
            assert (a < 0) == (b < 0) == (d < 0)


def test_padding_is_five_millimetres_in_double(oracle):
    for r in (0.06, 0.0235433969646692, 0.1, 0.3):
        assert oracle.L.orc_pad(np.float32(r)) == np.float32(np.float64(np.float32(r)) + 0.005)


# ---------------------------------------------------------------- sample set

def test_step_count_and_delta_closed_form(oracle):
    rng = np.random.default_rng(6)
    for _ in range(2000):
        s = np.float32(rng.uniform(-3, 3, 7))
        e = np.float32(rng.uniform(-3, 3, 7))
        n, d = oracle.steps(s, e, 32)
        dist = np.linalg.norm(s.astype(np.float64) - e)
        x = dist / 31 * 32
        if abs(x - round(x)) > 1e-4:
            assert n == max(math.ceil(x), 1)
        assert np.allclose(d, (e.astype(np.float64) - s) / (31 * n), rtol=1e-6, atol=1e-9)
    s = np.float32([0.1] * 7)
    n, d = oracle.steps(s, s, 32)
    assert n == 1 and np.all(d == 0)
    n, _ = oracle.steps(np.zeros(7, np.float32), np.float32([31.0 / 32, 0, 0, 0, 0, 0, 0]), 32)
    assert n == 1
    n, _ = oracle.steps(np.zeros(7, np.float32), np.float32([1.0, 0, 0, 0, 0, 0, 0]), 32)
    assert n == 2


def test_sample_positions_including_lane31_past_end(oracle):
    rng = np.random.default_rng(8)
    for _ in range(200):
        s = np.float32(rng.uniform(-3, 3, 7))
        e = np.float32(rng.uniform(-3, 3, 7))
        n, d = oracle.steps(s, e, 32)
        assert np.array_equal(oracle.sample(s, d, n, 0, 0), s)
        for t, i in ((31, 0), (31, n - 1), (17, n // 2), (5, n - 1)):
            k = t * n + i
            exact = s.astype(np.float64) + (e.astype(np.float64) - s) * k / (31 * n)
            assert np.abs(oracle.sample(s, d, n, t, i) - exact).max() < 2e-5
        # the last lane-31 sample lies past `end` by (n-1)/(31 n) of the edge
        last = oracle.sample(s, d, n, 31, n - 1).astype(np.float64)
        beyond = (last - e) @ (e.astype(np.float64) - s) / np.dot(e - s.astype(np.float64), e - s)
        assert abs(beyond - (n - 1) / (31 * n)) < 1e-5


# ---------------------------------------------------------------- planted scenes

READY = np.float32([0.0, -0.785, 0.0, -2.356, 0.0, 1.571, 0.785])


def _scene(bcc, env):
    return bcc.PackedScene.from_envs([env])


def _edge(a, b):
    return np.float32([a, b]).reshape(1, 2, -1)


def test_empty_world_free_edge(oracle, bcc):
    sc = _scene(bcc, {})
    assert oracle.check_dense(0, sc, _edge(READY, READY))[0, 0] == 0
    out, mg = oracle.check_dense64(0, sc, _edge(READY, READY))
    assert out[0, 0] == 0 and mg[0, 0] > 0


def _outward_sphere(oracle, q):
    """Index and centre of the robot sphere farthest from the base axis."""
    c = oracle.fk64(0, q)
    r = np.hypot(c[:, 0], c[:, 1]) + c[:, 2] * 0
    i = int(np.argmax(r))
    return i, c[i]


@pytest.mark.parametrize("gap,expect", [(0.003, 1), (0.0049, 1), (0.0051, 0), (0.007, 0)])
def test_padding_decides_near_contact(oracle, bcc, gap, expect):
    i, c = _outward_sphere(oracle, READY)
    u = np.array([c[0], c[1], 0.0])
    u /= np.linalg.norm(u)
    R = 0.02
    centre = c[:3] + u * (c[3] + R + gap)
    sc = _scene(bcc, {"spheres": np.float32([[*centre, R]])})
    got = oracle.check_dense(0, sc, _edge(READY, READY))[0, 0]
    _, mg = oracle.check_dense64(0, sc, _edge(READY, READY))
    assert abs(mg[0, 0] - (gap - 0.005)) < 1e-6
    assert got == expect


def test_collision_only_on_lane31_extrapolated_sample(oracle, bcc):
    a = READY.copy()
    b = READY.copy()
    a[0] = np.float32(-2.5)
    b[0] = np.float32(2.5)                 # sweep the base joint: n = ceil(5/31*32) = 6
    n, d = oracle.steps(a, b, 32)
    assert n == 6
    last = oracle.sample(a, d, n, 31, n - 1)
    cs = oracle.fk64(0, last.astype(np.float64))
    ring = np.hypot(cs[:, 0], cs[:, 1]) > 0.3
    ang = np.where(ring, np.arctan2(cs[:, 1], cs[:, 0]), -np.inf)
    c = cs[int(np.argmax(ang))]            # the sphere leading the motion (+q0 turns +z)
    R = 0.01
    tangent = np.array([-c[1], c[0], 0.0]) / np.hypot(c[0], c[1])   # direction of motion for +q0
    centre = c[:3] + tangent * (c[3] + R + 0.002)   # 3 mm inside the padded contact, ahead of the arm
    sc = _scene(bcc, {"spheres": np.float32([[*centre, R]])})
    # geometry: only samples k > 31 n (past `end`) reach the obstacle
    dists = []
    for k in range(32 * n):
        q = a.astype(np.float64) + (b.astype(np.float64) - a) * k / (31 * n)
        cs = oracle.fk64(0, q)
        dists.append(np.min(np.linalg.norm(cs[:, :3] - centre, axis=1) - cs[:, 3] - R - 0.005))
    dists = np.array(dists)
    assert dists[:31 * n + 1].min() > 0.001 and dists[-1] < 0
    assert oracle.check_dense(0, sc, _edge(a, b))[0, 0] == 1
    # the FP64 margin agrees, and dropping the samples past `end` would miss it
    _, mg = oracle.check_dense64(0, sc, _edge(a, b))
    assert mg[0, 0] < 0
    assert dists[:31 * n + 1].min() > 0


def test_cylinders_are_never_tested(oracle, bcc):
    cyl = np.float32([[0, 0, 1.5, 0, 0, -1.5, 1.0, 1.0 / 2.25]])   # fat cylinder around the robot
    assert oracle.check_dense(0, _scene(bcc, {"cylinders": cyl}), _edge(READY, READY))[0, 0] == 0
    assert oracle.check_dense(0, _scene(bcc, {"capsules": cyl}), _edge(READY, READY))[0, 0] == 1


def test_self_collision_configs(oracle, bcc):
    rng = np.random.default_rng(9)
    sc = _scene(bcc, {})
    found = {0: 0, 1: 0}
    for _ in range(400):
        q = np.float32(rng.uniform(-2.9, 2.9, 7))
        out, mg = oracle.check_dense64(0, sc, _edge(q, q))
        if abs(mg[0, 0]) < 1e-3:
            continue
        got = oracle.check_dense(0, sc, _edge(q, q))[0, 0]
        assert got == int(mg[0, 0] < 0)
        found[got] += 1
    assert found[0] > 10 and found[1] > 10


def test_fp32_vs_fp64_disagree_only_near_contact(oracle, bcc):
    from batch_cc_amd import workloads as W
    robot, sc, edges = W.make_config("c0", seed=3, scale=0.05)
    v32 = oracle.check_dense(0, sc, edges)
    v64, mg = oracle.check_dense64(0, sc, edges)
    bad = v32 != v64
    assert np.all(np.abs(mg[bad]) < 1e-4), mg[bad]


def test_work_count_closed_form_for_a_free_sample(oracle, bcc):
    # empty world, zero-length edge: 32 samples, each runs the full program:
    # FK + one call per env group (2 ops, no obstacles) + 11 per self gate
    # (+ 11 per fine pair of gates that overlap).
    sc = _scene(bcc, {})
    m = json.load(open(os.path.join(os.path.dirname(HERE), "batch-cc_amd", "robots", "panda.json")))
    c = oracle.fk(0, READY)
    fine = 0
    for g in m["self_program"]:
        a, b = g["gate"]
        if np.sum((c[a, :3] - c[b, :3]) ** 2) - (c[a, 3] + c[b, 3]) ** 2 < 0:
            fine += len(g["fine"])
    out, w, t = oracle.work_dense(0, sc, _edge(READY, READY))
    assert out[0, 0] == 0
    per = m["ref_fk_flops"] + 2 * len(m["env_program"]) + 11 * (len(m["self_program"]) + fine)
    assert int(w[0, 0]) == 32 * per
    assert int(t[0, 0]) == 32 * m["ref_fk_trig"]
