"""Input construction of the README run (MotionBenchMaker scenes, Halton
edges) on the CPU: the Halton recurrence against exact radical inverses, the
Euler-angle factories against the problem generator's own quaternions, the
driver's environment walk, the "box" family's cylinder rule, and the oracle's
FP32 verdicts on these scenes against its FP64 margins."""
import os
from fractions import Fraction

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")


@pytest.fixture(scope="module")
def mbm(bcc):
    from batch_cc_amd import mbm
    return mbm


@pytest.fixture(scope="module")
def panda_problems(mbm):
    return mbm.load_problems(os.path.join(GOLDEN, "mbm_panda.json"))


@pytest.fixture(scope="module")
def fetch_problems(mbm):
    return mbm.load_problems(os.path.join(GOLDEN, "mbm_fetch.json"))


def radical_inverse(k: int, b: int) -> Fraction:
    r, f = Fraction(0), Fraction(1, b)
    while k:
        k, digit = divmod(k, b)
        r += digit * f
        f /= b
    return r


def test_halton_draws_are_the_rounded_radical_inverses(mbm):
    # while denominators stay below 2^24 the float recurrence is exact
    # integer arithmetic, so draw k is float(radical inverse of k) per base
    rng = mbm.Halton(8)
    bases = (3, 5, 7, 11, 13, 17, 19, 23)
    for k in range(1, 3001):
        got = rng.next()
        exp = [np.float32(float(radical_inverse(k, b))) for b in bases]
        assert np.array_equal(got, np.array(exp, np.float32)), (k, got, exp)


def test_halton_restarts_with_rotated_bases(mbm):
    rng = mbm.Halton(3)
    rng.MAX_ITERATIONS = 5
    first = [rng.next() for _ in range(5)]
    again = rng.next()          # 6th draw: state reset, bases (5, 7, 3)
    assert np.array_equal(first[0], np.array([1 / 3, 1 / 5, 1 / 7], np.float32))
    assert np.array_equal(again, np.array([1 / 5, 1 / 7, 1 / 3], np.float32))
    rng.reset()
    assert np.array_equal(rng.next(), first[0])


def test_halton_edges_use_two_draws_per_edge_scaled_to_limits(mbm):
    from batch_cc_amd import workloads as W
    e = mbm.halton_edges("panda", 50)
    rng = mbm.Halton(7)
    lo, rg = W.LIMITS["panda"]
    for m in range(50):
        for j in range(2):
            u = rng.next()
            assert np.array_equal(e[m, j], (u * rg + lo).astype(np.float32))
    assert np.all(e >= lo) and np.all(e <= lo + rg)


def _rot(q):
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def _all_obstacles(problems):
    for fam, ps in problems["problems"].items():
        for p in ps:
            for kind in ("cylinder", "box"):
                for o in p[kind]:
                    yield fam, kind, o


@pytest.mark.parametrize("robot", ["panda", "fetch"])
def test_euler_axes_match_the_generator_quaternions(mbm, robot):
    """Cuboid axes from Euler XYZ against the rotation of the quaternion the
    problem generator stored beside the angles (roll and pitch occur in the
    "box" family), and yaw-only rotations keep exact zeros and unit z."""
    problems = mbm.load_problems(os.path.join(GOLDEN, f"mbm_{robot}.json"))
    worst, n_general = 0.0, 0
    for fam, kind, o in _all_obstacles(problems):
        c = mbm.cuboid(o["position"], o["orientation_euler_xyz"], (1, 1, 1))
        A = c[3:12].astype(np.float64).reshape(3, 3).T          # columns = axes
        R = _rot(o["orientation_quat_xyzw"])
        worst = max(worst, np.abs(A - R).max())
        assert np.abs(A.T @ A - np.eye(3)).max() < 5e-7
        roll, pitch, yaw = o["orientation_euler_xyz"]
        if roll == 0 and pitch == 0:
            assert np.array_equal(np.abs(c[[5, 8, 9, 10]]), np.zeros(4, np.float32))
            assert c[11] == 1.0
            assert abs(float(c[3]) - np.cos(np.float32(yaw))) < 3e-7
            assert abs(float(c[4]) - np.sin(np.float32(yaw))) < 3e-7
        else:
            n_general += 1
    assert worst < 3e-7, worst
    assert n_general > 0


def test_cylinders_become_z_capsules_with_exact_endpoints(mbm, panda_problems):
    from batch_cc_amd import workloads as W
    n = 0
    for fam, kind, o in _all_obstacles(panda_problems):
        if kind != "cylinder" or fam == "box":
            continue
        cap = mbm.cylinder_center(o["position"], o["orientation_euler_xyz"], o["radius"], o["length"])
        # yaw about the cylinder's own axis leaves the capsule unchanged
        assert np.array_equal(cap, W.cylinder_as_capsule(o["position"], o["radius"], o["length"]))
        L = np.float32(o["length"])
        assert cap[3] == 0 and cap[4] == 0 and abs(abs(float(cap[5])) - float(L)) <= 2e-7 * float(L) + 1e-7
        assert abs(float(cap[7]) * float(cap[5]) ** 2 - 1.0) < 1e-6
        n += 1
    assert n > 100


def test_general_cylinder_endpoints_follow_the_rotated_axis(mbm, panda_problems):
    for fam, kind, o in _all_obstacles(panda_problems):
        if kind != "cylinder" or fam != "box":
            continue
        cap = mbm.cylinder_center(o["position"], o["orientation_euler_xyz"], o["radius"], o["length"])
        R = _rot(o["orientation_quat_xyzw"])
        p1 = np.array(o["position"]) + R[:, 2] * o["length"] / 2
        assert np.abs(cap[:3] - p1).max() < 1e-6
        assert np.abs(cap[3:6] + R[:, 2] * o["length"]).max() < 1e-6


def test_driver_walk_panda_takes_bookshelf_small(mbm, panda_problems):
    sel = mbm.select_problems(panda_problems, 100)
    assert [f for f, _ in sel] == ["bookshelf_small"] * 100
    assert [p["index"] for _, p in sel] == [p["index"] for p in panda_problems["problems"]["bookshelf_small"]]
    sc = mbm.mbm_scene(panda_problems, 100)
    assert sc.num_envs == 100
    assert int(sc.counts[:, 0].sum()) == 0                      # no spheres in these scenes
    for e, (fam, p) in enumerate(sel):
        assert sc.counts[e, 1] == len(p["cylinder"]) and sc.counts[e, 4] == len(p["box"])


def test_driver_walk_fetch_skips_invalid_and_continues(mbm, fetch_problems):
    sel = mbm.select_problems(fetch_problems, 100)
    fams = [f for f, _ in sel]
    small = fetch_problems["problems"]["bookshelf_small"]
    n_valid = sum(p["valid"] for p in small)
    assert n_valid == 98
    assert fams == ["bookshelf_small"] * 98 + ["bookshelf_tall"] * 2
    assert all(p["valid"] for _, p in sel)


def test_box_family_cylinders_become_cuboids(mbm, panda_problems):
    sel = mbm.select_problems(panda_problems, 10, families={"box"})
    assert len(sel) == 10
    for fam, p in sel:
        env = mbm.problem_dict_to_env(p, fam)
        assert len(env["capsules"]) == 0
        assert len(env["cuboids"]) == len(p["cylinder"]) + len(p["box"])
        for k, o in enumerate(p["cylinder"]):
            r = np.float32(o["radius"])
            assert np.array_equal(env["cuboids"][k][12:], np.array([r, r, r / np.float32(2)], np.float32))
            assert np.array_equal(env["cuboids"][k][:3], np.array(o["position"], np.float32))


def _start_goal_margins(mbm, oracle, problems, rid, families):
    from batch_cc_amd import PackedScene
    margins = []
    for fam, p in mbm.select_problems(problems, 1000, families=families):
        sc = PackedScene.from_envs([mbm.problem_dict_to_env(p, fam)])
        for q in [p["start"]] + p["goals"]:
            q = np.array(q, np.float32)
            _, mg = oracle.check_dense64(rid, sc, np.stack([q, q])[None])
            margins.append(mg[0, 0])
    return np.array(margins)


@pytest.mark.parametrize("robot,rid", [("panda", 0), ("fetch", 1)])
def test_start_and_goal_configurations_are_free_up_to_the_padding(mbm, oracle, robot, rid):
    """Every start and goal was valid in the problem generator's mesh model.
    The sphere model encloses the meshes, so a sphere-model collision at such
    a configuration can only come from the 5 mm padding band: its FP64 margin
    stays above -5 mm.  A wrong Euler convention or cylinder rule rotates the
    shelves and pushes valid configurations deep into them."""
    problems = mbm.load_problems(os.path.join(GOLDEN, f"mbm_{robot}.json"))
    m = _start_goal_margins(mbm, oracle, problems, rid, {"bookshelf_small", "box"})
    print(f"{robot}: {len(m)} configurations, {(m < 0).sum()} in contact, min margin {m.min():.4f} m")
    assert len(m) > 200
    assert m.min() > -0.005
    assert (m < 0).mean() < 0.05


def test_start_goal_invariant_detects_transposed_axes(mbm, oracle, panda_problems, monkeypatch):
    orig = mbm.quat_rotate
    monkeypatch.setattr(mbm, "quat_rotate", lambda q, v: orig((-q[0], -q[1], -q[2], q[3]), v))
    m = _start_goal_margins(mbm, oracle, panda_problems, 0, {"bookshelf_small"})
    assert m.min() < -0.005


def test_oracle_fp32_agrees_with_fp64_on_the_readme_scenes(mbm, panda_problems, oracle):
    sc = mbm.mbm_scene(panda_problems, 20)
    edges = mbm.halton_edges("panda", 100)
    v32 = oracle.check_dense(0, sc, edges)
    v64, mg = oracle.check_dense64(0, sc, edges)
    bad = v32 != v64
    assert np.all(np.abs(mg[bad]) < 1e-4), mg[bad]
    print(f"collision fraction {v32.mean():.3f}, FP32/FP64 disagreements {int(bad.sum())}")
    assert 0.0 < v32.mean() < 1.0


def test_readme_driver_argument_errors_match_the_reference(capsys):
    import importlib.util
    path = os.path.join(os.path.dirname(HERE), "batch-cc_amd", "batch_cc_mbm.py")
    spec = importlib.util.spec_from_file_location("batch_cc_mbm", path)
    drv = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(drv)
    assert drv.main(["panda", "100"]) == 1
    assert "Usage: ./batch_cc <robot_name> <num_environments> <num_edges>" in capsys.readouterr().out
    assert drv.main(["xarm", "1", "1"]) == 255
    assert "Unknown robot name: xarm" in capsys.readouterr().out
