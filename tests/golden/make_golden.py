"""Write the golden fixtures under tests/golden/ from the reference sources.

Run in a container that has the reference checkout (default
/root/reference); the fixtures are committed so tests never need it.

fk_<robot>.json
    Sphere centres (float64) of every sphere the reference's generated
    fkcc<Robot> tests, evaluated from the reference's own FK expressions
    (src/utils.cuh:342-741 Panda, :10382-24992 Fetch, :24994-38241 Baxter,
    :38243-44112 Xarm7)
    with exact sin/cos, at 24 seeded random joint configurations that were
    NOT used to fit the robot model.  Sphere order = the committed model's
    sphere order (spheres are matched by position, so a fit that maps a
    sphere to the wrong frame or local offset fails the comparison).

structure_<robot>.json
    The reference's check program read from the source text: counts of the
    environment and self-collision calls at each nesting level, and every
    group's membership (env_groups: [gate, [fine spheres]]; self_groups:
    [[gate a, gate b], [[fine a, fine b], ...]]) in the reference's order.
    The membership comes from scan_program() below, a comment-stripped
    brace-nesting scan of the call sites that shares nothing with
    extract_model.py's parser: a call is a group's gate when an `if` on it
    opens a block holding further calls, and spheres are numbered by first
    appearance of their argument text (gate before fine, a before b), the
    numbering the committed models use.

limits.json
    Joint limits (s_a, s_m) per robot as written in src/Robots.hh:17-166
    and :196-203 (Xarm7: 13 entries each, the joints its generated model
    reads, although its traits declare dimension 7).
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


# fkcc<Robot> bodies in src/utils.cuh (Panda :342-5442, Fetch :10382-24992,
# Baxter :24994-38241, Xarm7 :38244-44110)
FKCC_LINES = {"panda": (342, 5442), "fetch": (10382, 24992), "baxter": (24994, 38241), "xarm7": (38244, 44110)}
_TOKEN = re.compile(r"\b(sphere_environment_in_collision|sphere_sphere_self_collision)\s*\(|[{}]")


def _strip_comments(text: str) -> str:
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    return re.sub(r"//[^\n]*", "", text)


def _close_paren(text: str, i: int) -> int:
    depth = 0
    for k in range(i, len(text)):
        if text[k] == "(":
            depth += 1
        elif text[k] == ")":
            depth -= 1
            if depth == 0:
                return k
    raise ValueError("unbalanced parentheses")


def scan_program(text: str) -> dict:
    """Group membership of one fkcc body, from brace nesting alone."""
    t = _strip_comments(text)
    stack = []            # per open block: index of the call whose `if` opened it, or None
    calls = []            # (kind, argument texts, parent call index or None)
    opener = None
    pos = 0
    while (m := _TOKEN.search(t, pos)):
        tok = m.group(0)
        pos = m.end()
        if tok == "{":
            stack.append(opener)
            opener = None
            continue
        if tok == "}":
            stack.pop()
            continue
        kind = "env" if m.group(1) == "sphere_environment_in_collision" else "self"
        op = m.end() - 1
        cp = _close_paren(t, op)
        args = [re.sub(r"\s+", "", a) for a in t[op + 1:cp].split(",")]
        parent = next((o for o in reversed(stack) if o is not None), None)
        calls.append((kind, args, parent))
        opener = len(calls) - 1 if re.match(r"\s*\)\s*\{", t[cp + 1:]) else None
        pos = cp + 1
    if stack:
        raise ValueError("unbalanced braces")
    index: dict = {}

    def sid(key):
        return index.setdefault(tuple(key), len(index))

    env, self_ = [], []
    for ci, (kind, args, parent) in enumerate(calls):
        if parent is not None:
            continue
        kids = [k for k, c in enumerate(calls) if c[2] == ci]
        if any(c[2] in kids for c in calls):
            raise ValueError("check program nested deeper than gate -> fine")
        if kind == "env":
            env.append([sid(args[1:5]), [sid(calls[k][1][1:5]) for k in kids]])
        else:
            self_.append([[sid(args[0:4]), sid(args[4:8])],
                          [[sid(calls[k][1][0:4]), sid(calls[k][1][4:8])] for k in kids]])
    return {"env_groups": env, "self_groups": self_, "num_spheres": len(index)}


def main(ref_root="/root/reference"):
    em = load_extractor()
    src = open(os.path.join(ref_root, "src", "utils.cuh")).read()
    for robot in ("panda", "fetch", "baxter", "xarm7"):
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
        a, b = FKCC_LINES[robot]
        groups = scan_program("\n".join(src.split("\n")[a - 1:b]))
        with open(os.path.join(HERE, f"structure_{robot}.json"), "w") as f:
            json.dump({"env_top": env_top, "env_fine": env_fine, "self_top": self_top, "self_fine": self_fine,
                       "env_calls_in_text": raw_env, "self_calls_in_text": raw_self,
                       "source": f"src/utils.cuh:{a}-{b}, brace-nesting scan (scan_program)", **groups}, f)
        print(robot, len(spheres), "spheres;", env_top, env_fine, self_top, self_fine)

    # joint limits straight from the text of src/Robots.hh
    rtext = open(os.path.join(ref_root, "src", "Robots.hh")).read()
    limits = {}
    for robot, typ in (("panda", "Panda"), ("fetch", "Fetch"), ("baxter", "Baxter"), ("xarm7", "Xarm7")):
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
