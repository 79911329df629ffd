"""The robot code generator's check schedule: every environment group and
every self-collision group is emitted exactly once in either order, chunks
respect the gate-chunk size, and the tree order is monotone in frame depth
(what lets frame registers die early).  Verdicts do not depend on the order
(an OR over the same gated tests); the GPU parity tests check the built
code."""
import importlib.util
import json
import os

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ROBOTS = os.path.join(REPO, "batch-cc_amd", "robots")


@pytest.fixture(scope="module")
def gen():
    spec = importlib.util.spec_from_file_location("gen_robot_code", os.path.join(ROBOTS, "gen_robot_code.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


@pytest.mark.parametrize("robot", ["panda", "fetch", "baxter", "xarm7"])
@pytest.mark.parametrize("order", ["tree", "program"])
def test_schedule_covers_every_group_once(gen, robot, order):
    model = json.load(open(os.path.join(ROBOTS, f"{robot}.json")))
    items = gen.schedule(model, order)
    env = [g for kind, it in items if kind == "env" for g in it]
    selfs = [it for kind, it in items if kind == "self"]
    n_env = sum(1 for it in model["program"] if it["kind"] == "env")
    assert sorted(env) == list(range(n_env))
    assert sorted(selfs) == list(range(len(model["self_program"])))
    assert all(1 <= len(it) <= gen.GATE_CHUNK for kind, it in items if kind == "env")


@pytest.mark.parametrize("robot", ["panda", "fetch", "baxter", "xarm7"])
def test_tree_order_is_monotone_in_frame_depth(gen, robot):
    model = json.load(open(os.path.join(ROBOTS, f"{robot}.json")))
    rank = gen.tree_rank(model)
    sph = model["spheres"]
    envp = model["env_program"]

    def key(spheres):
        return max(rank[sph[s]["frame"]] for s in spheres)

    last_env = -2
    for kind, it in gen.schedule(model, "tree"):
        if kind == "env":
            k = max(key([envp[g]["gate"]] + envp[g]["fine"]) for g in it)
            assert k >= last_env
            last_env = k
        else:
            # a self group comes after an environment chunk reaching its depth
            # (or after all chunks)
            assert key(model["self_program"][it]["gate"]) <= last_env or last_env == max(
                key([envp[g]["gate"]] + envp[g]["fine"]) for g in range(len(envp)))
    # depth-first ranks cover every joint once, parents before children
    joints = model["joints"]
    assert sorted(rank[j] for j in range(len(joints))) == list(range(len(joints)))
    assert all(rank[j["parent"]] < rank[i] for i, j in enumerate(joints))


def test_generated_sources_match_the_generator(gen, tmp_path):
    """The committed check code is what the generator writes for the
    committed models (default order and chunk sizes)."""
    for robot in ("panda", "fetch", "baxter", "xarm7"):
        model = json.load(open(os.path.join(ROBOTS, f"{robot}.json")))
        committed = open(os.path.join(REPO, "batch-cc_amd", "csrc", "robots", f"{robot}_check.inc")).read()
        assert gen.gen_check(model, "tree") == committed, robot
