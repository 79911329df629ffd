"""The README run on the GPU: `batch_cc panda 100 1000` (100 bookshelf_small
scenes x 1000 Halton edges, resolution 32) and the fetch run over the same
problem files, bit for bit against the oracle, with the FP32/FP64 audit; the
"box" family's rotated cuboids; and the ragged path on the same inputs."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _problems(robot):
    from batch_cc_amd import mbm
    return mbm.load_problems(os.path.join(GOLDEN, f"mbm_{robot}.json"))


@pytest.mark.parametrize("robot,rid", [("panda", 0), ("fetch", 1)])
def test_readme_run_matches_the_oracle(bcc, oracle, gpu, robot, rid):
    from batch_cc_amd import mbm
    sc, edges = mbm.readme_run(_problems(robot), robot, 100, 1000)
    print(f"README run {robot}: {sc.num_envs} envs x {edges.shape[0]} edges", flush=True)
    with bcc.Scene(sc) as scene:
        got = bcc.check_dense(robot, scene, edges)
    exp = oracle.check_dense(rid, sc, edges, threads=16)
    assert np.array_equal(got, exp), np.argwhere(got != exp)[:10]
    v64, mg = oracle.check_dense64(rid, sc, edges, threads=16)
    bad = got != v64
    for (m, e) in np.argwhere(bad)[:20]:
        print(f"FP32 vs FP64 mismatch edge {m} env {e}: gpu={got[m, e]} margin={mg[m, e]:.3e} m")
    assert np.all(np.abs(mg[bad]) < 1e-4)
    print(f"collision fraction {got.mean():.4f}, FP64 disagreements {int(bad.sum())}")


def test_box_family_rotated_cuboids_match_the_oracle(bcc, oracle, gpu):
    from batch_cc_amd import mbm
    sc = mbm.mbm_scene(_problems("panda"), 10, families={"box"})
    edges = mbm.halton_edges("panda", 400)
    with bcc.Scene(sc) as scene:
        got = bcc.check_dense("panda", scene, edges)
    assert np.array_equal(got, oracle.check_dense(0, sc, edges, threads=16))
    print(f"box family collision fraction {got.mean():.4f}")


def test_readme_run_through_the_ragged_plan(bcc, gpu):
    """The same pairs as a ragged batch (each environment with its own slice
    of edges) give the dense verdicts for those pairs."""
    from batch_cc_amd import mbm
    sc, edges = mbm.readme_run(_problems("panda"), "panda", 100, 1000)
    rng = np.random.default_rng(11)
    counts = rng.integers(0, 40, sc.num_envs)
    off = np.zeros(sc.num_envs + 1, np.int64)
    off[1:] = np.cumsum(counts)
    pick = rng.integers(0, edges.shape[0], int(off[-1]))
    with bcc.Scene(sc) as scene:
        dense = bcc.check_dense("panda", scene, edges)
        ragged = bcc.check_ragged("panda", scene, off, edges[pick])
    env_of = np.repeat(np.arange(sc.num_envs), counts)
    assert np.array_equal(ragged, dense[pick, env_of])


@pytest.mark.parametrize("robot,rid", [("panda", 0), ("fetch", 1)])
def test_start_and_goal_configurations_on_the_gpu(bcc, oracle, gpu, robot, rid):
    """Every problem's start and goals as zero-length edges against all 100
    scenes: bit-exact against the oracle; on its own scene each lies outside
    the obstacles up to the 5 mm padding band (FP64 margin > -5 mm)."""
    from batch_cc_amd import mbm
    problems = _problems(robot)
    sel = mbm.select_problems(problems, 100)
    sc = mbm.mbm_scene(problems, 100)
    qs, own = [], []
    for e, (fam, p) in enumerate(sel):
        for q in [p["start"]] + p["goals"]:
            qs.append(np.array(q, np.float32))
            own.append(e)
    q = np.stack(qs)
    edges = np.stack([q, q], 1)
    with bcc.Scene(sc) as scene:
        got = bcc.check_dense(robot, scene, edges)
    assert np.array_equal(got, oracle.check_dense(rid, sc, edges, threads=16))
    _, mg = oracle.check_dense64(rid, sc, edges, threads=16)
    own_margin = mg[np.arange(len(own)), own]
    print(f"{robot}: {len(own)} configurations, {int(got[np.arange(len(own)), own].sum())} in contact with "
          f"their own scene, min margin {own_margin.min():.4f} m; {got.mean():.3f} of all (config, scene) pairs collide")
    assert own_margin.min() > -0.005


def test_readme_driver_prints_the_oracle_matrix(bcc, oracle, gpu):
    """`batch_cc_mbm.py panda 100 200` prints the reference driver's lines; its
    verdict matrix equals the oracle's on the same inputs."""
    import subprocess
    import sys
    from batch_cc_amd import mbm
    repo = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(repo, "batch-cc_amd", "batch_cc_mbm.py"), "panda", "100", "200"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.strip().splitlines()
    assert lines[3] == "Edges checked: 20000" and lines[4].startswith("Throughput: ")
    rows = [ln for ln in lines if ln.startswith("Edge ") and ":" in ln and not ln.startswith("Edges")]
    assert len(rows) == 200 and rows[0].startswith("Edge 0: ")
    got = np.array([[int(v) for v in ln.split(":")[1].split()] for ln in rows], np.uint8)
    sc, edges = mbm.readme_run(_problems("panda"), "panda", 100, 200)
    assert np.array_equal(got, oracle.check_dense(0, sc, edges, threads=16))
