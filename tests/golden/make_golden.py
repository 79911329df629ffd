"""Write the golden fixtures under tests/golden/ from the reference sources.

Run in a container that has the reference checkout (default
/root/reference); the fixtures are committed so tests never need it.

fk_<robot>.json
    Sphere centres (float64) of every sphere the reference's generated
    fkcc<Robot> tests, evaluated from the reference's own FK expressions
    (src/utils.cuh:342-741 Panda, :10382-24992 Fetch, :24994-38241 Baxter)
    with exact sin/cos, at 24 seeded random joint configurations that were
    NOT used to fit the robot model.  Sphere order = the committed model's
    sphere order (spheres are matched by position, so a fit that maps a
    sphere to the wrong frame or local offset fails the comparison).

structure_<robot>.json
    Counts of the reference's check program read from the source text:
    environment and self-collision calls at each nesting level.

limits.json
    Joint limits (s_a, s_m) per robot as written in src/Robots.hh:17-166.
"""
from __future__ import annotations

import importlib.util
import json
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))


def load_extractor():
    path = os.path.join(REPO, "batch-cc_amd", "robots", "extract_model.py")
    spec = importlib.util.spec_from_file_location("extract_model", path)
    mod = importlib.util.module_from_spec(spec)
    sys.modules["extract_model"] = mod
    spec.loader.exec_module(mod)
    return mod


def main(ref_root="/root/reference"):
    em = load_extractor()
    src = open(os.path.join(ref_root, "src", "utils.cuh")).read()
    for robot in ("panda", "fetch", "baxter"):
        model = json.load(open(os.path.join(REPO, "batch-cc_amd", "robots", f"{robot}.json")))
        dim = em.ROBOT_DIM[robot]
        qname, stmts = em.parse_robot(src, robot)
        rng = np.random.default_rng(12345)
        Q = rng.uniform(-3.0, 3.0, size=(24, dim))
        if robot == "fetch":
            Q[:, 0] = rng.uniform(0.0, 0.4, size=24)   # prismatic torso (metres)
        checks = em.evaluate(qname, stmts, Q)
        # world positions keyed per sphere via the model's sphere order: match
        # each call's evaluated centre to the model sphere with equal radius
        # whose FP64 chain position (computed independently below) agrees.
        seen = {}
        for c in checks:
            calls = [(c.args, 0)] if c.kind == "env" else [(c.args, 0), (c.args, 4)]
            for ch in c.children:
                calls += [(ch.args, 0)] if ch.kind == "env" else [(ch.args, 0), (ch.args, 4)]
            for args, j0 in calls:
                key = em.sphere_key(args, j0)
                if key not in seen:
                    x, y, z, r = args[j0:j0 + 4]
                    seen[key] = (np.stack([x.v, y.v, z.v], 1), float(np.float32(r.v[0])))
        spheres = list(seen.values())
        if len(spheres) != len(model["spheres"]):
            raise SystemExit(f"{robot}: {len(spheres)} spheres in source vs {len(model['spheres'])} in model")
        out = {
            "robot": robot,
            "q": Q.tolist(),
            "centres": [s[0].tolist() for s in spheres],
            "radii": [s[1] for s in spheres],
            "source": f"reference src/utils.cuh fkcc<{em.ROBOT_TYPE[robot]}>, float64, exact sin/cos",
        }
        with open(os.path.join(HERE, f"fk_{robot}.json"), "w") as f:
            json.dump(out, f)
        # structure counts
        env_top = sum(1 for c in checks if c.kind == "env")
        env_fine = sum(len(c.children) for c in checks if c.kind == "env")
        self_top = sum(1 for c in checks if c.kind == "self")
        self_fine = sum(len(c.children) for c in checks if c.kind == "self")
        raw_env = len(re.findall(r"sphere_environment_in_collision\(",
                                 em.function_body(src, robot)))
        raw_self = len(re.findall(r"sphere_sphere_self_collision\(", em.function_body(src, robot)))
        with open(os.path.join(HERE, f"structure_{robot}.json"), "w") as f:
            json.dump({"env_top": env_top, "env_fine": env_fine, "self_top": self_top, "self_fine": self_fine,
                       "env_calls_in_text": raw_env, "self_calls_in_text": raw_self}, f, indent=1)
        print(robot, len(spheres), "spheres;", env_top, env_fine, self_top, self_fine)

    # joint limits straight from the text of src/Robots.hh
    rtext = open(os.path.join(ref_root, "src", "Robots.hh")).read()
    limits = {}
    for robot, typ in (("panda", "Panda"), ("fetch", "Fetch"), ("baxter", "Baxter")):
        body = rtext[rtext.index(f"struct {typ}"):]
        sm = body[body.index("get_s_m"):]
        sm = sm[sm.index("{", sm.index("values[]")) + 1:sm.index("}", sm.index("values[]"))]
        sa = body[body.index("get_s_a"):]
        sa = sa[sa.index("{", sa.index("values[]")) + 1:sa.index("}", sa.index("values[]"))]
        parse = lambda t: [float(v.strip().rstrip("f")) for v in t.split(",") if v.strip()]
        limits[robot] = {"s_m": parse(sm), "s_a": parse(sa)}
    with open(os.path.join(HERE, "limits.json"), "w") as f:
        json.dump(limits, f, indent=1)


if __name__ == "__main__":
    main(*sys.argv[1:])
