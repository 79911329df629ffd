"""Host-side checks that need no GPU: the C ABI library loads and exports
every declared entry point, argument errors are reported (not crashed on),
the C++ drop-in shim compiles, and the synthetic workloads follow the
reference's input conventions."""
import os
import re
import subprocess

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(REPO, "include", "bcc.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(bcc_\w+)\s*\(", text, flags=re.M)))


def test_library_exports_every_declared_symbol(bcc):
    L = bcc.lib()
    names = declared_functions()
    assert len(names) >= 20
    for n in names:
        assert hasattr(L, n), n
    nm = subprocess.run(["nm", "-D", "--defined-only", bcc.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (bcc_\w+)", nm))
    assert set(names) <= exported


def test_robot_tables(bcc):
    L = bcc.lib()
    assert [L.bcc_robot_dimension(i) for i in range(4)] == [7, 8, 14, 13]
    assert L.bcc_robot_dimension(4) == -1
    assert [L.bcc_robot_num_spheres(i) for i in range(4)] == [70, 125, 100, 74]


def test_scene_rejects_null_pointer_with_positive_count(bcc):
    import ctypes
    L = bcc.lib()
    counts = np.zeros((1, 6), np.int32)
    counts[0, 0] = 2
    offs = np.zeros(1, np.int64)
    h = ctypes.c_void_p()
    rc = L.bcc_scene_create_packed(None, offs.ctypes.data, counts.ctypes.data, 1, ctypes.byref(h))
    assert rc == 1
    assert b"null pointers" in L.bcc_last_error()
    counts[0, 0] = -1
    data = np.zeros(4, np.float32)
    rc = L.bcc_scene_create_packed(data.ctypes.data, offs.ctypes.data, counts.ctypes.data, 1, ctypes.byref(h))
    assert rc == 1 and b"negative" in L.bcc_last_error()


def test_check_rejects_bad_arguments(bcc):
    L = bcc.lib()
    assert L.bcc_check_dense(9, None, None, 1, 32, None) == 1
    assert b"unknown robot" in L.bcc_last_error()
    assert L.bcc_check_dense(0, None, None, 1, 32, None) == 1
    assert b"scene is NULL" in L.bcc_last_error()


def test_python_binding_validates_edge_shape(bcc):
    with pytest.raises(bcc.BCCError):
        bcc._edges_array("panda", np.zeros((3, 2, 8), np.float32))


def test_cpp_shim_compiles_and_throws_like_the_reference(tmp_path, bcc):
    exe = tmp_path / "shim_driver"
    subprocess.run(["g++", "-O1", "-std=c++17", "-I" + os.path.join(REPO, "include"),
                    os.path.join(REPO, "tests", "cpp", "shim_driver.cpp"), "-o", str(exe),
                    "-L" + os.path.dirname(bcc.LIB_PATH), "-lbcc", "-Wl,-rpath," + os.path.dirname(bcc.LIB_PATH)],
                   check=True)
    r = subprocess.run([str(exe), "nullcheck"], capture_output=True, text=True)
    assert r.returncode == 0 and "runtime_error" in r.stdout


def test_robot_traits_header_matches_the_reference_tables(tmp_path):
    """include/Robots.hh (the drop-in traits) against the joint-limit tables
    read from the reference's src/Robots.hh (tests/golden/limits.json)."""
    import json
    src = tmp_path / "traits.cpp"
    src.write_text("""
#include <cstdio>
#include "Robots.hh"
template <class R> void dump() {
    std::printf("%s %d", R::name, (int)R::dimension);
    for (int i = 0; i < R::dimension; ++i) std::printf(" %.9g", (double)R::get_s_m(i));
    for (int i = 0; i < R::dimension; ++i) std::printf(" %.9g", (double)R::get_s_a(i));
    std::printf("\\n");
}
int main() {
    dump<ppln::robots::Panda>(); dump<ppln::robots::Fetch>(); dump<ppln::robots::Baxter>();
    dump<ppln::robots::Xarm7>();
    return 0;
}
""")
    exe = tmp_path / "traits"
    subprocess.run(["g++", "-std=c++17", "-I" + os.path.join(REPO, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    lim = json.load(open(os.path.join(REPO, "tests", "golden", "limits.json")))
    for line in filter(None, out):
        tok = line.split()
        name, dim = tok[0], int(tok[1])
        vals = np.float32([float(v) for v in tok[2:]])
        assert dim == len(lim[name]["s_m"]), name
        assert np.array_equal(vals[:dim], np.float32(lim[name]["s_m"])), name
        assert np.array_equal(vals[dim:], np.float32(lim[name]["s_a"])), name


def test_workloads_follow_reference_conventions(bcc):
    from batch_cc_amd import workloads as W
    robot, sc, edges = W.make_config("c0", seed=0, scale=0.05)
    assert robot == "panda" and sc.num_envs == 5 and edges.shape == (50, 2, 7)
    assert np.all(sc.counts[:, 0] == 8) and np.all(sc.counts[:, 1] == 8) and np.all(sc.counts[:, 4] == 16)
    assert np.all(sc.counts[:, 3] == 0)          # cylinders stored as capsules
    env = sc.env(0)
    cap = env["capsules"]
    assert np.all(cap[:, 3] == 0) and np.all(cap[:, 4] == 0) and np.all(cap[:, 5] < 0)
    assert np.array_equal(cap[:, 7], np.float32(1.0 / (cap[:, 5].astype(np.float32) ** 2).astype(np.float64)))
    cub = env["cuboids"]
    assert np.array_equal(cub[:, 3:12], np.tile(np.float32([1, 0, 0, 0, 1, 0, 0, 0, 1]), (16, 1)))
    lo, rng_ = W.LIMITS["panda"]
    assert np.all(edges >= lo) and np.all(edges <= lo + rng_)
    # deterministic
    _, sc2, e2 = W.make_config("c0", seed=0, scale=0.05)
    assert np.array_equal(sc.data, sc2.data) and np.array_equal(edges, e2)
    robot, sc, edges = W.make_config("c1", seed=0, scale=0.01)
    assert robot == "fetch" and edges.shape[2] == 8
    n = sc.counts[:, [0, 1, 4]].sum(1)
    assert np.all((n >= 4) & (n <= 64))


def test_ragged_plan_and_group_reject_bad_arguments(bcc):
    import ctypes
    L = bcc.lib()
    h = ctypes.c_void_p()
    off = np.zeros(2, np.int64)
    assert L.bcc_ragged_plan_create_tuned(0, None, off.ctypes.data, None, 32, None, 0, 0, ctypes.byref(h)) == 1
    assert b"scene is NULL" in L.bcc_last_error()
    assert L.bcc_ragged_group_create(None, 0, ctypes.byref(h)) == 1
    assert b"at least one plan" in L.bcc_last_error()
    arr = (ctypes.c_void_p * 2)(None, None)
    assert L.bcc_ragged_group_create(arr, 2, ctypes.byref(h)) == 1
    assert b"is NULL" in L.bcc_last_error()
    assert L.bcc_ragged_plan_create_ordered(0, None, off.ctypes.data, None, 32, None, None, 0, 0,
                                            ctypes.byref(h)) == 1
    assert b"scene is NULL" in L.bcc_last_error()
    assert L.bcc_model_order_pairs(None, None, None, 1, None, None, None, None, None) == 1
    assert b"bad arguments" in L.bcc_last_error()
    assert L.bcc_ragged_plan_config(None, None, None) == 1
    assert L.bcc_ragged_group_launch(None, None) == 1
    assert L.bcc_ragged_group_destroy(None) == 0


def test_tuning_table_covers_every_robot():
    text = open(os.path.join(REPO, "batch-cc_amd", "csrc", "bcc_tuning.h")).read()
    body = text[text.index("kTuneTable[]"):]
    rows = re.findall(r"\{\s*(\d+),\s*(\d+),\s*(\d+),\s*(\d+)\s*\}", body)
    assert sorted(int(r) for r, _, _, _ in rows) == [0, 1, 2, 3]
    assert all(1 <= int(t) <= 1024 and 1 <= int(g) <= 32 and 0 <= int(st) <= 1 << 22 for _, t, g, st in rows)
    thr = re.search(r"kHiVariantMinObstacles\[4\] = \{(\d+), (\d+), (\d+), (\d+)\}", text)
    assert thr and all(int(x) >= 0 for x in thr.groups())
    hi = re.search(r"kHiVariantMaxObstacles\[4\] = \{(\d+), (\d+), (\d+), (\d+)\}", text)
    assert hi and all(int(b) > int(a) or int(a) == int(b) == 1 << 30 for a, b in zip(thr.groups(), hi.groups()))
    cull = re.search(r"kCullMinWork = (\d+);", text)
    assert cull and int(cull.group(1)) >= 0


def test_mixed_workload_layout(bcc):
    from batch_cc_amd import workloads as W
    parts, env_robot, goff = W.make_mixed(seed=1, scale=0.01, total_pairs=1 << 20)
    assert [p[0] for p in parts] == ["panda", "fetch", "baxter"]
    assert int(goff[-1]) == int((1 << 20) * 0.01)
    allg = np.concatenate([p[4] for p in parts])
    assert np.array_equal(np.sort(allg), np.arange(goff[-1]))
    for r, (robot, sc, off, edges, gidx) in enumerate(parts):
        ids = np.nonzero(env_robot == r)[0]
        assert sc.num_envs == len(ids) and edges.shape == (int(off[-1]), 2, bcc.DIMS[robot])
        assert np.array_equal(np.diff(off), np.diff(goff)[ids])
        n = sc.counts.sum(1)
        assert np.all((n >= 32) & (n <= 512))
        c = np.concatenate([sc.env(e)["cuboids"][:, :3] for e in range(sc.num_envs)])
        assert np.all(c[:, 0] ** 2 + c[:, 1] ** 2 >= np.float32(W.MIXED_CLEAR_RADIUS ** 2))


def test_force_variant_rejects_out_of_range(bcc):
    L = bcc.lib()
    assert L.bcc_debug_force_variant(2) == 1
    assert b"variant" in L.bcc_last_error()
    assert L.bcc_debug_force_variant(-1) == 0


def test_tuning_hooks_reject_out_of_range(bcc):
    L = bcc.lib()
    assert L.bcc_debug_dense_chunk(-1) == 1
    assert b"edges per workgroup" in L.bcc_last_error()
    assert L.bcc_debug_dense_chunk(5000) == 1
    assert L.bcc_debug_dense_chunk(0) == 0
    assert L.bcc_debug_cull_min(-1) == 0
    assert L.bcc_debug_learned_tile(257) == 1
    assert b"learned tile" in L.bcc_last_error()
    assert L.bcc_debug_learned_tile(0) == 0
    assert L.bcc_debug_learned_far(33) == 1
    assert L.bcc_debug_learned_far(0) == 0
    assert L.bcc_debug_learned_passes(4) == 1
    assert L.bcc_debug_learned_passes(0) == 0
    assert L.bcc_debug_learned_min_steps(-1) == 1
    assert L.bcc_debug_learned_min_steps(0) == 0
    assert L.bcc_debug_learned_locality(3) == 1
    assert b"locality" in L.bcc_last_error()
    assert L.bcc_debug_learned_locality(1) == 0


def test_tuner_variant_range_selection():
    """tune.pick_range on the MI355X timings behind the committed table: panda
    runs the high variant only between its few-obstacle and its c4-shaped
    batch, fetch always, baxter never; a batch is never decided by a bound
    placed at its own mean."""
    import importlib.util
    spec = importlib.util.spec_from_file_location("bcc_tune", os.path.join(REPO, "batch-cc_amd", "tune.py"))
    tune = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(tune)
    never = tune.NEVER
    panda = [(32.0, 1.967, 1.686), (7.0, 2.992, 3.013), (293.5, 2.406, 2.596)]
    assert tune.pick_range(panda) == (20, 163)
    assert tune.pick_range([(34.4, 1.216, 0.991), (7.0, 2.290, 2.043), (246.8, 3.167, 2.778)]) == (0, never)
    assert tune.pick_range([(145.6, 1.597, 1.647), (7.0, 3.087, 3.510), (279.8, 6.592, 8.432)]) == (never, never)
    lo, hi = tune.pick_range(panda)
    assert all(not (m == lo or m == hi) for m, _, _ in panda)
