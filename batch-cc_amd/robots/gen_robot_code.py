"""Generate per-robot source from a robot model JSON (see extract_model.py).

Two outputs per robot:

``csrc/robots/<robot>_model.h``
    Plain C data tables (kinematic tree, sphere table, check programs) in a
    ``bcc_robot_model`` struct.  These are the robot model as data; the CPU
    oracle interprets them with generic loops.

``csrc/robots/<robot>_check.inc``
    Straight-line HIP device code for gfx950: forward kinematics as a chain
    of per-joint 3x3 rotation + translation frames held in registers, and
    the bounding-sphere-gated environment and self-collision program, with
    wave-level early exit.  Sphere centres are computed on demand, right
    before the test that needs them.

The FP32 arithmetic the generated FK performs is the project's fixed spec
(also restated by the oracle in oracle/bcc_oracle.c):

  joint j with parent frame (Rp, tp), fixed (G, g), angle q:
    A[i][k] = sum over l with G[l][k] != 0, ascending l, first term a
              product, later terms fmaf(Rp[i][l], G[l][k], acc)
    t[i]    = tp[i], then fmaf(Rp[i][l], g[l], t[i]) for g[l] != 0
    revolute:  (s, c) = bcc_sincosf(q)
               R[i][0] = fmaf(s, A[i][1], c * A[i][0])
               R[i][1] = fmaf(-s, A[i][0], c * A[i][1])
               R[i][2] = A[i][2]
    prismatic: R = A;  t[i] = fmaf(q, A[i][2], t[i])
  sphere centre: P[i] = t[i], then fmaf(R[i][l], p[l], P[i]) for p[l] != 0

The generator folds operations on the constants 0, 1 and -1 only where the
folded form is IEEE-identical for finite operands (up to the sign of zero,
which no collision predicate can observe).

Usage: python gen_robot_code.py <model.json> <out_dir> [--order tree|program]
                                 [--gate-chunk N] [--fine-chunk N]
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np


def f32(x) -> float:
    return float(np.float32(x))


def flit(x: float) -> str:
    """C float literal that round-trips the float32 value exactly."""
    v = np.float32(x)
    if v == 0:
        return "0.0f"
    s = np.format_float_scientific(v, unique=True, precision=None)
    return f"{s}f" if "e" in s else f"{s}f"


def dlit(x: float) -> str:
    return repr(float(x))


# ------------------------------------------------------------ symbolic FP32

class E:
    """An FP32 value in the generated code: a constant or a named register."""

    def __init__(self, const: float | None = None, name: str | None = None):
        self.const = None if const is None else f32(const)
        self.name = name

    @property
    def is_const(self):
        return self.const is not None

    def code(self) -> str:
        if self.is_const:
            return flit(self.const)
        return self.name


class Emitter:
    def __init__(self):
        self.lines: list[str] = []
        self.tmp = 0
        self.indent = 1

    def line(self, s: str):
        self.lines.append("    " * self.indent + s)

    def new(self, expr: str, hint: str) -> E:
        name = f"{hint}"
        self.line(f"const float {name} = {expr};")
        return E(name=name)

    # exact-foldable helpers -------------------------------------------
    def mul(self, a: E, b: E, hint: str) -> E:
        for x, y in ((a, b), (b, a)):
            if y.is_const:
                if y.const == 0.0:
                    return E(0.0)
                if y.const == 1.0:
                    return x
                if y.const == -1.0:
                    if x.is_const:
                        return E(-x.const)
                    return self.new(f"-{x.code()}", hint)
        return self.new(f"{a.code()} * {b.code()}", hint)

    def fma(self, a: E, b: E, c: E, hint: str) -> E:
        """fmaf(a, b, c) with exact folds."""
        if (a.is_const and a.const == 0.0) or (b.is_const and b.const == 0.0):
            return c
        if c.is_const and c.const == 0.0:
            return self.mul(a, b, hint)
        for x, y in ((a, b), (b, a)):
            if y.is_const and y.const == 1.0:
                return self.new(f"{c.code()} + {x.code()}", hint)
            if y.is_const and y.const == -1.0:
                return self.new(f"{c.code()} - {x.code()}", hint)
        return self.new(f"fmaf({a.code()}, {b.code()}, {c.code()})", hint)

    def neg(self, a: E, hint: str) -> E:
        if a.is_const:
            return E(-a.const)
        return self.new(f"-{a.code()}", hint)


class LazyFK:
    """Emits each joint frame on first request (parents first), at the
    emitter's current position; callers request frames only at function
    scope so later code can use them."""

    def __init__(self, em: Emitter, model: dict, prefix: str = ""):
        self.em = em
        self.model = model
        self.p = prefix
        self.frames = {-1: ([[E(1.0 if i == k else 0.0) for k in range(3)] for i in range(3)],
                            [E(0.0) for _ in range(3)])}

    def __getitem__(self, j):
        if j not in self.frames:
            self.emit(j)
        return self.frames[j]

    def emit(self, j):
        em = self.em
        jt = self.model["joints"][j]
        Rp, tp = self[jt["parent"]]
        G = [[f32(v) for v in row] for row in jt["R"]]
        g = [f32(v) for v in jt["t"]]
        em.line(f"// joint {j} ({jt['type']}), parent frame {jt['parent']}")
        A = [[None] * 3 for _ in range(3)]
        for i in range(3):
            for k in range(3):
                acc = None
                for l in range(3):
                    if G[l][k] == 0.0:
                        continue
                    if acc is None:
                        acc = em.mul(Rp[i][l], E(G[l][k]), f"{self.p}a{j}_{i}{k}_{l}")
                    else:
                        acc = em.fma(Rp[i][l], E(G[l][k]), acc, f"{self.p}a{j}_{i}{k}_{l}")
                A[i][k] = acc if acc is not None else E(0.0)
        t = []
        for i in range(3):
            acc = tp[i]
            for l in range(3):
                if g[l] == 0.0:
                    continue
                acc = em.fma(Rp[i][l], E(g[l]), acc, f"{self.p}t{j}_{i}_{l}")
            t.append(acc)
        if jt["type"] == "revolute":
            em.line(f"float {self.p}s{j}, {self.p}c{j};")
            em.line(f"bcc_sincosf(q[{j}], {self.p}s{j}, {self.p}c{j});")
            s = E(name=f"{self.p}s{j}")
            c = E(name=f"{self.p}c{j}")
            ns = em.neg(s, f"{self.p}ns{j}")
            R = [[None] * 3 for _ in range(3)]
            for i in range(3):
                ca0 = em.mul(c, A[i][0], f"{self.p}ca{j}_{i}0")
                R[i][0] = em.fma(s, A[i][1], ca0, f"{self.p}R{j}_{i}0")
                ca1 = em.mul(c, A[i][1], f"{self.p}ca{j}_{i}1")
                R[i][1] = em.fma(ns, A[i][0], ca1, f"{self.p}R{j}_{i}1")
                R[i][2] = A[i][2]
            self.frames[j] = (R, t)
        else:
            qj = E(name=f"q[{j}]")
            t2 = [em.fma(qj, A[i][2], t[i], f"{self.p}tq{j}_{i}") for i in range(3)]
            self.frames[j] = (A, t2)


def gen_fk(em: Emitter, model: dict):
    """Emit the whole FK chain; returns the frame table."""
    fk = LazyFK(em, model)
    for j in model["joint_order"]:
        fk[j]
    return fk


def centre(em: Emitter, frames, sph: dict, tag: str) -> list[E]:
    R, t = frames[sph["frame"]]
    p = [f32(v) for v in sph["local"]]
    out = []
    for i in range(3):
        acc = t[i]
        for l in range(3):
            if p[l] == 0.0:
                continue
            acc = em.fma(R[i][l], E(p[l]), acc, f"{tag}_{'xyz'[i]}{l}")
        out.append(acc)
    return out


def padded_radius(r: float) -> float:
    """float(double(r) + 0.005): the reference pads env-test radii in double."""
    return f32(np.float64(np.float32(r)) + 0.005)


# Spheres per obstacle sweep, per robot: (first environment chunk, later
# environment chunks, fine spheres of a gated group).  With exact culling,
# larger sweeps pay: the half-wave box reduction and the near-obstacle
# ballot are paid once per sweep and cover more spheres.  Measured on MI355X
# (c1, c3, c2, c4's three parts, the README run, a dense baxter batch) over
# gate chunks 1-8 and fine chunks 2-32; fine 16 takes all but the two largest
# groups in one sweep.  Panda's (2, 4, 16) trades 5% on c3 for 16% on c4's
# panda part; fetch's gate chunk 5 is best on c1 and c4's fetch part (4 and 6
# are 7-8% slower: a register-allocation optimum rather than a trend).
CHUNKS = {"panda": (2, 4, 16), "fetch": (1, 5, 16), "baxter": (1, 4, 16),
          # xarm7's program has no gated groups: 74 lone environment spheres
          "xarm7": (2, 4, 16)}
# Sweeps below the culling threshold run plain counted loops over all
# obstacles (a second copy of each sweep's code) or the culled loop with every
# obstacle marked near.  The second copy pays for panda and fetch; for
# baxter's much longer check program the extra code costs more in
# instruction-cache misses than the plain loops save (MI355X, c2 and c4).
DUAL_SWEEP = {"panda": True, "fetch": True, "baxter": False, "xarm7": False}
# Dense kernels compute each round's samples from the edge (as the ragged and
# learned-order kernels do) or step them by an increment row in LDS: the
# first is faster for fetch (c1 -1.7%), no faster for panda and baxter
# (within the 2% run-to-run spread).  MI355X measurements.
ROUNDS_FROM_EDGE = {"panda": False, "fetch": True, "baxter": False, "xarm7": False}
# The step loops of the learned-order kernels visit a pair's samples in block
# order after step 0 (the samples of one step lie next to each other along the
# edge, so the culling boxes are tight) or in plain order: block order is 4.5%
# faster on c3 (panda), 5% slower on c1 and c4's part (fetch) and even on
# baxter's batches (c2, c4).  MI355X measurements.
BLOCK_STEPS = {"panda": True, "fetch": False, "baxter": False, "xarm7": False}
GATE_CHUNK = 2      # defaults for robots not listed; set per robot by gen_check
FINE_CHUNK = 4
FIRST_CHUNK = 0     # size of the first environment chunk (0: GATE_CHUNK)
ORDER = "tree"      # "tree": kinematic-tree order; "program": the reference's order
_CHUNK_OVERRIDE = False   # set by --gate-chunk / --fine-chunk


def tree_rank(model: dict) -> dict:
    """Depth-first rank of every joint frame (children in index order); the
    base frame -1 ranks first."""
    children: dict[int, list[int]] = {}
    for j, jt in enumerate(model["joints"]):
        children.setdefault(jt["parent"], []).append(j)
    order: list[int] = []

    def visit(p):
        for c in children.get(p, []):
            order.append(c)
            visit(c)

    visit(-1)
    rank = {-1: -1}
    for i, j in enumerate(order):
        rank[j] = i
    return rank


def schedule(model: dict, order: str):
    """Emission order: a list of ("env", [group indices]) chunks of up to
    GATE_CHUNK environment groups and ("self", group index) items.

    "program": environment groups in the reference's program order, then the
    whole self-collision program.  "tree": every item keyed by the deepest
    frame (in depth-first rank) it reads, environment chunks in key order and
    each self group right after the first chunk whose key reaches its own, so
    a frame's registers die once the items on it are done instead of living
    until the self program at the end."""
    sph = model["spheres"]
    envp = model["env_program"]
    prog = [it["index"] for it in model["program"] if it["kind"] == "env"]
    nself = len(model["self_program"])
    if order == "program":
        return ([("env", prog[c:c + GATE_CHUNK]) for c in range(0, len(prog), GATE_CHUNK)]
                + [("self", i) for i in range(nself)])
    rank = tree_rank(model)

    def key(spheres):
        return max(rank[sph[s]["frame"]] for s in spheres)

    env_sorted = sorted(prog, key=lambda gi: (key([envp[gi]["gate"]] + envp[gi]["fine"]), prog.index(gi)))
    first = FIRST_CHUNK or GATE_CHUNK
    chunks = [env_sorted[:first]] + [env_sorted[c:c + GATE_CHUNK] for c in range(first, len(env_sorted), GATE_CHUNK)]
    chunks = [c for c in chunks if c]
    selfs = sorted(range(nself), key=lambda i: (key(model["self_program"][i]["gate"]), i))
    out, si = [], 0
    for ch in chunks:
        k = max(key([envp[gi]["gate"]] + envp[gi]["fine"]) for gi in ch)
        out.append(("env", ch))
        while si < len(selfs) and key(model["self_program"][selfs[si]]["gate"]) <= k:
            out.append(("self", selfs[si]))
            si += 1
    out += [("self", i) for i in selfs[si:]]
    return out


def gen_check(model: dict, order: str | None = None) -> str:
    """check(): environment groups in chunks of GATE_CHUNK gate (or lone)
    spheres per obstacle sweep, each chunk followed by one sweep per gated
    group over its fine spheres, and the self-collision groups (sphere-pair
    gate, then the gated fine cross product), in the order schedule()
    gives.  The evaluation order differs from the reference's interleaving;
    the verdict (an OR over gated tests) does not depend on it."""
    global GATE_CHUNK, FINE_CHUNK, FIRST_CHUNK
    order = order or ORDER
    robot = model["robot"]
    if not _CHUNK_OVERRIDE:
        FIRST_CHUNK, GATE_CHUNK, FINE_CHUNK = CHUNKS.get(robot, (FIRST_CHUNK, GATE_CHUNK, FINE_CHUNK))
    dim = model["dimension"]
    sph = model["spheres"]
    em = Emitter()
    em.line(f"// generated by gen_robot_code.py from the {robot} model ({order} order); do not edit")
    frames = LazyFK(em, model)
    em.line("uint64_t hit = 0;")

    def need_frames(spheres_used):
        for s in spheres_used:
            frames[sph[s]["frame"]]

    cache: dict[int, list[str]] = {}

    def c3(idx, tag, reuse=True):
        if reuse and idx in cache:
            return cache[idx]
        pts = [p.code() for p in centre(em, frames, sph[idx], tag)]
        if reuse:
            cache[idx] = pts
        return pts

    def emit_env_chunk(chunk):
        envp = model["env_program"]
        ng = len(chunk)
        need_frames([envp[gi]["gate"] for gi in chunk] + [f for gi in chunk for f in envp[gi]["fine"]])
        em.line("{")
        em.indent += 1
        gx, gy, gz, grp = [], [], [], []
        for gi in chunk:
            g = envp[gi]
            c = c3(g["gate"], f"s{g['gate']}", reuse=False)
            gx.append(c[0]); gy.append(c[1]); gz.append(c[2])
            grp.append(flit(padded_radius(sph[g["gate"]]["radius"])))
        em.line(f"const float gx[{ng}] = {{{', '.join(gx)}}};")
        em.line(f"const float gy[{ng}] = {{{', '.join(gy)}}};")
        em.line(f"const float gz[{ng}] = {{{', '.join(gz)}}};")
        em.line(f"const float grp[{ng}] = {{{', '.join(grp)}}};")
        em.line(f"bool gh[{ng}];")
        em.line(f"bcc_env_sweep<{ng}, U, kDualSweep>(env, gx, gy, gz, grp, live, gh);")
        for k, gi in enumerate(chunk):
            g = envp[gi]
            em.line(f"// env group {gi}" + (f" ({g['label']})" if g["label"] else ""))
            if g["lone"]:
                em.line(f"{{ const uint64_t d = bcc_group_spread<G>(__ballot(gh[{k}]) & live); hit |= d; live &= ~d; }}")
                continue
            em.line(f"if (uint64_t need = __ballot(gh[{k}]) & live) {{")
            em.indent += 1
            fines = g["fine"]
            for f0 in range(0, len(fines), FINE_CHUNK):
                sub = fines[f0:f0 + FINE_CHUNK]
                nf = len(sub)
                em.line("if (need) {")
                em.indent += 1
                fx, fy, fz, frp = [], [], [], []
                for fi, f in enumerate(sub):
                    c = c3(f, f"f{gi}_{f0 + fi}", reuse=False)
                    fx.append(c[0]); fy.append(c[1]); fz.append(c[2])
                    frp.append(flit(padded_radius(sph[f]["radius"])))
                em.line(f"const float fx[{nf}] = {{{', '.join(fx)}}};")
                em.line(f"const float fy[{nf}] = {{{', '.join(fy)}}};")
                em.line(f"const float fz[{nf}] = {{{', '.join(fz)}}};")
                em.line(f"const float frp[{nf}] = {{{', '.join(frp)}}};")
                em.line(f"const bool h = bcc_env_any<{nf}, U, kDualSweep, G>(env, fx, fy, fz, frp, need);")
                em.line("const uint64_t d = bcc_group_spread<G>(__ballot(h) & need);")
                em.line("hit |= d; live &= ~d; need &= ~d;")
                em.indent -= 1
                em.line("}")
            em.indent -= 1
            em.line("}")
        em.indent -= 1
        em.line("}")
        em.line("if (!live) return hit;")

    def emit_self_group(gi):
        g = model["self_program"][gi]
        a, b = g["gate"]
        need_frames([a, b])
        ca = c3(a, f"sg{a}")
        cb = c3(b, f"sg{b}")
        ra = flit(sph[a]["radius"])
        rb = flit(sph[b]["radius"])
        em.line(f"// self group {gi}" + (f" ({g['label']})" if g["label"] else ""))
        gate = (f"bcc_self_hit({ca[0]}, {ca[1]}, {ca[2]}, {ra}, "
                f"{cb[0]}, {cb[1]}, {cb[2]}, {rb})")
        if g["lone"]:
            em.line(f"{{ const uint64_t d = bcc_group_spread<G>(__ballot({gate}) & live); hit |= d; live &= ~d; }}")
            return
        em.line(f"if (uint64_t need = __ballot({gate}) & live) {{")
        em.indent += 1
        # frames for the fine spheres are recomputed in this (rarely taken)
        # block instead of being kept live through the whole section
        local = LazyFK(em, model, prefix=f"L{gi}_")
        A = []
        B = []
        for x, y in g["fine"]:
            if x not in A:
                A.append(x)
            if y not in B:
                B.append(y)
        assert set(map(tuple, g["fine"])) == {(x, y) for x in A for y in B}, "self group is not a cross product"
        for s_ in A + B:
            local[sph[s_]["frame"]]
        acs = [([p.code() for p in centre(em, local, sph[x], f"fa{gi}_{ai}")], flit(sph[x]["radius"]))
               for ai, x in enumerate(A)]
        for bi, y in enumerate(B):
            em.line("if (need) {")
            em.indent += 1
            yc = [p.code() for p in centre(em, local, sph[y], f"fb{gi}_{bi}")]
            ry = flit(sph[y]["radius"])
            terms = [f"bcc_self_hit({xc[0]}, {xc[1]}, {xc[2]}, {rx}, {yc[0]}, {yc[1]}, {yc[2]}, {ry})"
                     for xc, rx in acs]
            em.line("const bool h = " + "\n                | ".join(terms) + ";")
            em.line("const uint64_t d = bcc_group_spread<G>(__ballot(h) & need);")
            em.line("hit |= d; live &= ~d; need &= ~d;")
            em.indent -= 1
            em.line("}")
        em.indent -= 1
        em.line("}")

    for kind, item in schedule(model, order):
        if kind == "env":
            emit_env_chunk(item)
        else:
            emit_self_group(item)
    em.line("return hit;")

    # all sphere centres (test hook): same FK and centre emission as check()
    em2 = Emitter()
    frames2 = gen_fk(em2, model)
    for si, s in enumerate(sph):
        pts = centre(em2, frames2, s, f"p{si}")
        em2.line(f"out[{4 * si}] = {pts[0].code()}; out[{4 * si + 1}] = {pts[1].code()}; "
                 f"out[{4 * si + 2}] = {pts[2].code()}; out[{4 * si + 3}] = {flit(s['radius'])};")
    fk_fn = [
        "    // Sphere centres and radii, [num_spheres][4] (test hook).",
        f"    static __device__ __forceinline__ void spheres(const float (&q)[{dim}], float* out)",
        "    {",
    ] + ["    " + ln for ln in em2.lines] + ["    }"]

    head = [
        f"// {robot}: generated by batch-cc_amd/robots/gen_robot_code.py from {robot}.json.",
        "// Do not edit by hand.  Included inside namespace bcc by bcc_kernels.hip.",
        f"struct Robot_{robot} {{",
        f"    static constexpr int dim = {dim};",
        f"    static constexpr int num_spheres = {len(sph)};",
        "    // sweeps below the culling threshold get their own plain loops",
        f"    static constexpr bool kDualSweep = {'true' if DUAL_SWEEP.get(robot, True) else 'false'};",
        "    // dense kernels compute each round's samples from the edge",
        f"    static constexpr bool kRoundsFromEdge = {'true' if ROUNDS_FROM_EDGE.get(robot, True) else 'false'};",
        "    // learned-order step loops visit steps after the first in block order",
        f"    static constexpr bool kBlockSteps = {'true' if BLOCK_STEPS.get(robot, False) else 'false'};",
        "    // Collision predicate for one sample per lane.  Lanes form aligned groups of",
        "    // G (32: a half-wave, the reference's 32 samples of one pair per step; fewer:",
        "    // several pairs per half-wave, G samples each).  `live` holds the lanes whose",
        "    // group's pair is still undecided; returns the group-spread mask of lanes",
        "    // whose group found a colliding sample (each collision decides its group).",
        "    // q[j] is joint j of this lane's sample (an array or an LDS row view);",
        "    // U is the obstacle-loop unroll factor of the calling kernel variant.",
        "    template <int U, int G = 32, class Q>",
        "    static __device__ __forceinline__ uint64_t check(const Q& q, const EnvView& env, uint64_t live)",
        "    {",
    ]
    # the check body uses `E` as the env view name in helpers; keep the name `env`
    body = ["    " + ln for ln in em.lines]
    tail = ["    }"] + fk_fn + ["};", ""]
    return "\n".join(head + body + tail)


def gen_model_header(model: dict) -> str:
    r = model["robot"]
    R = r.upper()
    dim = model["dimension"]
    sph = model["spheres"]
    L = []
    L.append(f"/* {r} robot model tables: generated by batch-cc_amd/robots/gen_robot_code.py")
    L.append(f" * from {r}.json ({model['derivation']}).  Do not edit by hand. */")
    L.append(f"#ifndef BCC_{R}_MODEL_H")
    L.append(f"#define BCC_{R}_MODEL_H")
    L.append('#include "bcc_model.h"')
    L.append("")
    p = f"bcc_{r}_"
    L.append(f"static const int {p}joint_order[{dim}] = {{{', '.join(str(j) for j in model['joint_order'])}}};")
    L.append(f"static const int {p}joint_parent[{dim}] = {{{', '.join(str(j['parent']) for j in model['joints'])}}};")
    L.append(f"static const int {p}joint_type[{dim}] = {{"
             + ", ".join("0" if j["type"] == "revolute" else "1" for j in model["joints"]) + "};")
    L.append(f"static const float {p}joint_R[{dim}][9] = {{")
    for j in model["joints"]:
        L.append("    {" + ", ".join(flit(v) for row in j["R"] for v in row) + "},")
    L.append("};")
    L.append(f"static const float {p}joint_t[{dim}][3] = {{")
    for j in model["joints"]:
        L.append("    {" + ", ".join(flit(v) for v in j["t"]) + "},")
    L.append("};")
    L.append(f"static const double {p}joint_R64[{dim}][9] = {{")
    for j in model["joints"]:
        L.append("    {" + ", ".join(dlit(v) for row in j["R"] for v in row) + "},")
    L.append("};")
    L.append(f"static const double {p}joint_t64[{dim}][3] = {{")
    for j in model["joints"]:
        L.append("    {" + ", ".join(dlit(v) for v in j["t"]) + "},")
    L.append("};")
    ns = len(sph)
    L.append(f"static const int {p}sphere_frame[{ns}] = {{{', '.join(str(s['frame']) for s in sph)}}};")
    L.append(f"static const float {p}sphere_p[{ns}][3] = {{")
    for s in sph:
        L.append("    {" + ", ".join(flit(v) for v in s["local"]) + "},")
    L.append("};")
    L.append(f"static const double {p}sphere_p64[{ns}][3] = {{")
    for s in sph:
        L.append("    {" + ", ".join(dlit(v) for v in s["local"]) + "},")
    L.append("};")
    L.append(f"static const float {p}sphere_r[{ns}] = {{{', '.join(flit(s['radius']) for s in sph)}}};")

    env = model["env_program"]
    fine_off = [0]
    fines = []
    for g in env:
        fines.extend(g["fine"])
        fine_off.append(len(fines))
    L.append(f"static const int {p}env_gate[{len(env)}] = {{{', '.join(str(g['gate']) for g in env)}}};")
    L.append(f"static const int {p}env_lone[{len(env)}] = {{{', '.join('1' if g['lone'] else '0' for g in env)}}};")
    L.append(f"static const int {p}env_fine_off[{len(env) + 1}] = {{{', '.join(map(str, fine_off))}}};")
    L.append(f"static const int {p}env_fine[{max(1, len(fines))}] = {{{', '.join(map(str, fines)) or '0'}}};")

    sp = model["self_program"]
    a_off, b_off, As, Bs = [0], [0], [], []
    for g in sp:
        A = []
        B = []
        for x, y in g["fine"]:
            if x not in A:
                A.append(x)
            if y not in B:
                B.append(y)
        As.extend(A)
        Bs.extend(B)
        a_off.append(len(As))
        b_off.append(len(Bs))
    L.append(f"static const int {p}self_gate_a[{len(sp)}] = {{{', '.join(str(g['gate'][0]) for g in sp)}}};")
    L.append(f"static const int {p}self_gate_b[{len(sp)}] = {{{', '.join(str(g['gate'][1]) for g in sp)}}};")
    L.append(f"static const int {p}self_lone[{len(sp)}] = {{{', '.join('1' if g['lone'] else '0' for g in sp)}}};")
    L.append(f"static const int {p}self_a_off[{len(sp) + 1}] = {{{', '.join(map(str, a_off))}}};")
    L.append(f"static const int {p}self_a[{max(1, len(As))}] = {{{', '.join(map(str, As)) or '0'}}};")
    L.append(f"static const int {p}self_b_off[{len(sp) + 1}] = {{{', '.join(map(str, b_off))}}};")
    L.append(f"static const int {p}self_b[{max(1, len(Bs))}] = {{{', '.join(map(str, Bs)) or '0'}}};")

    prog = model["program"]
    L.append(f"static const int {p}prog_kind[{len(prog)}] = {{"
             + ", ".join("0" if it["kind"] == "env" else "1" for it in prog) + "};")
    L.append(f"static const int {p}prog_index[{len(prog)}] = {{{', '.join(str(it['index']) for it in prog)}}};")
    L.append(f"static const int {p}prog_fk_ops[{len(prog)}] = {{{', '.join(str(it['fk_ops_before']) for it in prog)}}};")
    L.append(f"static const int {p}prog_fk_trig[{len(prog)}] = {{{', '.join(str(it['fk_trig_before']) for it in prog)}}};")
    L.append("")
    L.append(f"static const bcc_robot_model bcc_{r}_model = {{")
    L.append(f'    "{r}", {dim}, {ns}, {len(env)}, {len(sp)}, {len(prog)},')
    L.append(f"    {p}joint_order, {p}joint_parent, {p}joint_type, {p}joint_R, {p}joint_t, {p}joint_R64, {p}joint_t64,")
    L.append(f"    {p}sphere_frame, {p}sphere_p, {p}sphere_p64, {p}sphere_r,")
    L.append(f"    {p}env_gate, {p}env_lone, {p}env_fine_off, {p}env_fine,")
    L.append(f"    {p}self_gate_a, {p}self_gate_b, {p}self_lone, {p}self_a_off, {p}self_a, {p}self_b_off, {p}self_b,")
    L.append(f"    {p}prog_kind, {p}prog_index, {p}prog_fk_ops, {p}prog_fk_trig,")
    L.append(f"    {model['ref_fk_flops']}, {model['ref_fk_trig']}")
    L.append("};")
    L.append("")
    L.append(f"#endif /* BCC_{R}_MODEL_H */")
    L.append("")
    return "\n".join(L)


def main(argv):
    global GATE_CHUNK, FINE_CHUNK, FIRST_CHUNK, _CHUNK_OVERRIDE
    order = ORDER
    args = list(argv[1:])
    pos = []
    while args:
        a = args.pop(0)
        if a == "--order":
            order = args.pop(0)
        elif a == "--gate-chunk":
            GATE_CHUNK = int(args.pop(0))
            _CHUNK_OVERRIDE = True
        elif a == "--fine-chunk":
            FINE_CHUNK = int(args.pop(0))
            _CHUNK_OVERRIDE = True
        elif a == "--first-chunk":
            FIRST_CHUNK = int(args.pop(0))
        else:
            pos.append(a)
    argv = [argv[0]] + pos
    if len(argv) != 3 or order not in ("tree", "program") or GATE_CHUNK < 1 or FINE_CHUNK < 1:
        print(__doc__)
        return 2
    model = json.load(open(argv[1]))
    out = argv[2]
    os.makedirs(out, exist_ok=True)
    r = model["robot"]
    with open(os.path.join(out, f"{r}_model.h"), "w") as f:
        f.write(gen_model_header(model))
    with open(os.path.join(out, f"{r}_check.inc"), "w") as f:
        f.write(gen_check(model, order))
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv))
