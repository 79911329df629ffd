"""Write tests/golden/mbm_<robot>.json: a subset of the reference's
MotionBenchMaker problem files (scripts/panda_problems.json,
scripts/fetch_problems.json) in their own structure, holding only the fields
the README driver reads (scripts/batch_cc_mbm.cpp:19-93) plus the generator's
quaternion, which pins the Euler-angle conversion, and each problem's start
and goal configurations (valid in the generator's mesh model).

Kept per robot, in the file's family order:
  bookshelf_small  every problem (the README run, `batch_cc panda 100 1000`,
                   takes its 100 environments from this family)
  bookshelf_tall   the first 5 problems (fetch has 98 valid bookshelf_small
                   problems, so a 100-environment run continues here)
  box              the first 10 problems (cylinders become cuboids there,
                   and rotations have roll and pitch)

Numbers are written with Python's shortest round-trip repr, so each parses to
the same double as the original text.

    python tests/golden/make_mbm_fixture.py [/root/reference]
"""
import json
import os
import sys

KEEP = {"bookshelf_small": None, "bookshelf_tall": 5, "box": 10}


def obstacle(o, kind):
    keep = {"sphere": ("position", "radius"),
            "cylinder": ("position", "orientation_euler_xyz", "orientation_quat_xyzw", "radius", "length"),
            "box": ("position", "orientation_euler_xyz", "orientation_quat_xyzw", "half_extents")}[kind]
    return {k: o[k] for k in keep}


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    here = os.path.dirname(os.path.abspath(__file__))
    for robot in ("panda", "fetch"):
        src = os.path.join(ref, "scripts", f"{robot}_problems.json")
        with open(src) as f:
            data = json.load(f)
        out = {"robot": data["robot"], "joints": data["joints"],
               "source": f"scripts/{robot}_problems.json", "problems": {}}
        for fam in sorted(KEEP):
            probs = data["problems"][fam]
            n = KEEP[fam] or len(probs)
            out["problems"][fam] = [
                {"index": p["index"], "valid": p["valid"], "start": p["start"], "goals": p["goals"],
                 **{k: [obstacle(o, k) for o in p[k]] for k in ("sphere", "cylinder", "box")}}
                for p in probs[:n]]
        dst = os.path.join(here, f"mbm_{robot}.json")
        with open(dst, "w") as f:
            json.dump(out, f, separators=(",", ":"))
        print(dst, os.path.getsize(dst))


if __name__ == "__main__":
    main()
