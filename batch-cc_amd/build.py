"""Build the gfx950 library libbcc.so in-tree, and the CPU oracle.

    python batch-cc_amd/build.py            # regenerate robot code, compile, link
    python batch-cc_amd/build.py --no-gen   # skip code generation

Every HIP translation unit is compiled by hipcc for gfx950 only, with
-ffp-contract=off (the arithmetic spec spells out every fused operation
explicitly) and packed-FP32 instruction selection disabled.
"""
from __future__ import annotations

import argparse
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
ROBOTS = ("panda", "fetch", "baxter", "xarm7")
LIB = os.path.join(HERE, "libbcc.so")
OBJDIR = os.path.join(HERE, "_build")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "-O3", "--offload-arch=gfx950", "-fPIC", "-std=c++17",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]
# (source, extra macros, object name): each robot's kernels compile as eleven
# parts in parallel (see csrc/bcc_robot_parts.h)
ROBOT_PARTS = 11
SOURCES = [("bcc_api.hip", (), "bcc_api"), ("bcc_mlp.hip", (), "bcc_mlp")] + [
    (f"robot_{r}.hip", (f"BCC_PART={k}",), f"robot_{r}_p{k}") for r in ROBOTS for k in range(ROBOT_PARTS)]


def run(cmd, **kw):
    p = subprocess.run(cmd, capture_output=True, text=True, **kw)
    out = "\n".join(line for line in (p.stdout + p.stderr).splitlines()
                    if "packed-fp32-ops" not in line and "is not a recognized feature" not in line)
    if p.returncode != 0:
        raise RuntimeError(f"command failed ({p.returncode}): {' '.join(cmd)}\n{out}")
    return out


def generate(order: str = "tree", gate_chunk: int | None = None, fine_chunk: int | None = None,
             first_chunk: int | None = None):
    """Regenerate the robot check code; chunk sizes default to the
    generator's per-robot table."""
    gen = os.path.join(HERE, "robots", "gen_robot_code.py")
    extra = []
    if gate_chunk:
        extra += ["--gate-chunk", str(gate_chunk)]
    if fine_chunk:
        extra += ["--fine-chunk", str(fine_chunk)]
    if first_chunk:
        extra += ["--first-chunk", str(first_chunk)]
    for r in ROBOTS:
        run([sys.executable, gen, os.path.join(HERE, "robots", f"{r}.json"), os.path.join(CSRC, "robots"),
             "--order", order, *extra])


def compile_one(item, objdir=OBJDIR, defines=()):
    src, macros, name = item
    obj = os.path.join(objdir, name + ".o")
    run([HIPCC, *HIP_FLAGS, *[f"-D{d}" for d in (*macros, *defines)], "-c", os.path.join(CSRC, src), "-o", obj],
        cwd=CSRC)
    return obj


def build_oracle():
    run(["make", "-s", "-C", os.path.join(REPO, "oracle")])


def build(gen: bool = True, jobs: int = 8, defines=(), out: str = LIB, order: str = "tree",
          gate_chunk: int | None = None, fine_chunk: int | None = None, first_chunk: int | None = None) -> str:
    """Compile and link; `defines` adds -D macros (tuning variants), `out`
    writes the library elsewhere (the in-tree libbcc.so is the product),
    `order` picks the generated check order (see gen_robot_code.schedule)."""
    if gen:
        generate(order, gate_chunk, fine_chunk, first_chunk)
    objdir = OBJDIR if out == LIB else out + ".objs"
    os.makedirs(objdir, exist_ok=True)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        # largest translation units first
        order_ = sorted(SOURCES, key=lambda it: ("baxter" not in it[0], "fetch" not in it[0]))
        objs = list(ex.map(lambda it: compile_one(it, objdir, defines), order_))
    tmp = out + ".tmp"
    run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, out)
    if out == LIB:
        build_oracle()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--no-gen", action="store_true")
    ap.add_argument("-j", type=int, default=8)
    ap.add_argument("-D", dest="defines", action="append", default=[], help="extra preprocessor macro")
    ap.add_argument("--out", default=LIB, help="library path (default: the in-tree libbcc.so)")
    ap.add_argument("--order", default="tree", choices=["tree", "program"], help="generated check order")
    ap.add_argument("--gate-chunk", type=int, default=None, help="gate spheres per obstacle sweep (default: per robot)")
    ap.add_argument("--fine-chunk", type=int, default=None, help="fine spheres per gated sweep (default: per robot)")
    ap.add_argument("--first-chunk", type=int, default=None, help="gate spheres in the first sweep (default: gate chunk)")
    a = ap.parse_args()
    print(build(gen=not a.no_gen, jobs=a.j, defines=a.defines, out=os.path.abspath(a.out), order=a.order,
                gate_chunk=a.gate_chunk, fine_chunk=a.fine_chunk, first_chunk=a.first_chunk))


if __name__ == "__main__":
    main()
