"""Endpoint-inclusive verdicts (bcc_check_dense_endpoints): the reference
validator's vamp semantics, scripts/batch_cc_validate.cpp:627-646 -- an edge
is valid when the motions s -> e, s -> s and e -> e all validate -- restated
with the plain sample set for each motion and compared with the oracle
composed the same way."""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROBOTS = {"panda": 0, "fetch": 1, "baxter": 2, "xarm7": 3}
REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _degenerate(edges, k):
    """(x, x) edges from endpoint k (0 = start, 1 = end) of every edge."""
    return np.ascontiguousarray(np.repeat(edges[:, k:k + 1], 2, axis=1))


@pytest.mark.parametrize("robot", ["panda", "fetch", "baxter", "xarm7"])
def test_endpoint_mode_equals_the_oracle_composition(bcc, oracle, gpu, robot):
    from batch_cc_amd import workloads as W
    rng = np.random.default_rng(21)
    envs = [W.mixed_env(rng, int(rng.integers(3, 25)), capsule_frac=0.2) for _ in range(12)]
    sc = bcc.PackedScene.from_envs(envs)
    edges = W.uniform_edges(robot, 300, rng)
    rid = ROBOTS[robot]
    plain = oracle.check_dense(rid, sc, edges)
    at_start = oracle.check_dense(rid, sc, _degenerate(edges, 0))
    at_end = oracle.check_dense(rid, sc, _degenerate(edges, 1))
    # the motion's first sample is its start: the s -> s check is implied
    assert np.all(at_start <= plain)
    want = plain | at_start | at_end
    with bcc.Scene(sc) as scene:
        got = bcc.check_dense_endpoints(robot, scene, edges)
        got_plain = bcc.check_dense(robot, scene, edges)
    assert np.array_equal(got_plain, plain)
    assert np.array_equal(got, want), int((got != want).sum())
    print(f"{robot}: plain {plain.mean():.3f}, endpoint-inclusive {want.mean():.3f}, "
          f"pairs decided by the end configuration alone {int((want & ~plain).sum())}")


def test_endpoint_mode_with_planted_end_collisions(bcc, oracle, gpu):
    """Environment i holds one small sphere at the centre of a robot sphere
    of edge i's end configuration, so that configuration collides there;
    the endpoint-inclusive verdicts still equal the oracle composition."""
    from batch_cc_amd import workloads as W
    robot, rid = "panda", 0
    # spheres the environment program tests: the fine spheres of the groups
    # and the lone gates (some spheres only take part in self checks)
    with open(os.path.join(REPO, "batch-cc_amd", "robots", "panda.json")) as f:
        prog = json.load(f)["env_program"]
    tested = sorted({k for g in prog for k in (g["fine"] or [g["gate"]])})
    rng = np.random.default_rng(5)
    edges = W.uniform_edges(robot, 64, rng)
    planted = []
    for i in range(edges.shape[0]):
        c = oracle.fk(rid, edges[i, 1])[tested[int(rng.integers(0, len(tested)))]]
        planted.append({"spheres": np.array([[c[0], c[1], c[2], 0.01]], np.float32)})
    sc = bcc.PackedScene.from_envs(planted)
    with bcc.Scene(sc) as scene:
        got = bcc.check_dense_endpoints(robot, scene, edges)
    at_end = oracle.check_dense(rid, sc, _degenerate(edges, 1))
    plain = oracle.check_dense(rid, sc, edges)
    assert np.all(np.diag(at_end) == 1)
    assert np.array_equal(got, plain | at_end)
    assert np.all(np.diag(got) == 1)
