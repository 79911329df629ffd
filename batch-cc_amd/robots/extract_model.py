"""Derive a kinematic-chain sphere model of a robot from the reference's
generated forward-kinematics + collision-check code.

The reference ships its robot models only as generated straight-line CUDA
(``src/utils.cuh``): ``fkcc<Panda>`` at src/utils.cuh:342-5442,
``fkcc<Fetch>`` at src/utils.cuh:10382-24992, ``fkcc<Baxter>`` at
src/utils.cuh:24994-38241 and ``fkcc<Xarm7>`` at src/utils.cuh:38243-44112.
The Xarm7 code reads 13 joint values (q[0..12]: seven arm joints and the
six joints of the gripper) although its traits declare dimension 7
(src/Robots.hh:190-193); its scale tables hold 13 entries
(src/Robots.hh:196-203), so the model is extracted with 13 joints.  That code is data in disguise: a list of sphere
centres written as closed-form expressions of the joint angles, plus a fixed
program of bounding-sphere-gated environment and self-collision tests.

This tool reads that text (it never compiles or links it), evaluates the
sphere-centre expressions in float64 with exact sin/cos at chosen joint
configurations, and fits this project's own representation:

* a kinematic tree: per joint a parent joint, a fixed transform G (3x3
  rotation + translation) from the parent frame to the joint frame at q=0,
  and a motion (revolute about local z, or prismatic along local z);
* per sphere: the joint frame it is rigidly attached to (-1 = base), its
  local centre and its float32 radius;
* the environment-check program: ordered groups of (gate sphere, fine
  spheres) or lone spheres, exactly as nested in the reference;
* the self-collision program: ordered groups of (gate pair, fine pairs) or
  lone pairs.

The fit is validated on random configurations against the float64
evaluation of the reference expressions (max centre error is written into
the model).  The output JSON is committed; this tool is only needed to
re-derive it and needs the reference sources on disk.

Usage:  python extract_model.py <path/to/utils.cuh> <panda|fetch|baxter|xarm7> <out.json>
"""
from __future__ import annotations

import json
import re
import sys
from dataclasses import dataclass, field

import numpy as np

ROBOT_DIM = {"panda": 7, "fetch": 8, "baxter": 14, "xarm7": 13}
ROBOT_TYPE = {"panda": "Panda", "fetch": "Fetch", "baxter": "Baxter", "xarm7": "Xarm7"}

# ---------------------------------------------------------------- tokenizer

_TOKEN_RE = re.compile(
    r"""
    (?P<bcomment>/\*.*?\*/)
  | (?P<lcomment>//[^\n]*)
  | (?P<num>(?:\d+\.\d*|\.\d+|\d+)(?:[eE][+-]?\d+)?[fF]?)
  | (?P<ident>[A-Za-z_]\w*)
  | (?P<punct>[-+*/=(){}\[\];,<>:!&|])
  | (?P<ws>\s+)
    """,
    re.VERBOSE | re.DOTALL,
)


@dataclass
class Tok:
    kind: str
    text: str


def tokenize(src: str) -> list[Tok]:
    out: list[Tok] = []
    pos = 0
    while pos < len(src):
        m = _TOKEN_RE.match(src, pos)
        if not m:
            raise ValueError(f"cannot tokenize near: {src[pos:pos + 40]!r}")
        kind = m.lastgroup
        pos = m.end()
        if kind in ("ws", "lcomment"):
            continue
        out.append(Tok(kind, m.group(kind)))
    return out


def function_body(src: str, robot: str) -> str:
    """Text between the braces of the live fkcc<Robot> specialisation."""
    typ = ROBOT_TYPE[robot]
    pat = re.compile(r"^[^\n/]*bool\s+fkcc<ppln::robots::" + typ + r">\(volatile float \*\w+", re.M)
    m = pat.search(src)
    if not m:
        raise ValueError(f"fkcc<{typ}> not found")
    start = src.index("{", m.end())
    depth = 0
    i = start
    n = len(src)
    while i < n:
        c = src[i]
        if src.startswith("//", i):
            i = src.index("\n", i)
            continue
        if src.startswith("/*", i):
            i = src.index("*/", i) + 2
            continue
        if c == "{":
            depth += 1
        elif c == "}":
            depth -= 1
            if depth == 0:
                return src[start + 1:i]
        i += 1
    raise ValueError("unbalanced braces")


def function_signature_input(src: str, robot: str) -> str:
    typ = ROBOT_TYPE[robot]
    m = re.search(r"^[^\n/]*bool\s+fkcc<ppln::robots::" + typ + r">\(volatile float \*(\w+)", src, re.M)
    return m.group(1)


# ------------------------------------------------------------------- parser

@dataclass
class Num:
    v: float


@dataclass
class Var:
    name: str


@dataclass
class Index:
    name: str
    idx: int


@dataclass
class Call:
    fn: str
    args: list


@dataclass
class Un:
    op: str
    a: object


@dataclass
class Bin:
    op: str
    a: object
    b: object


@dataclass
class Assign:
    name: str
    expr: object


@dataclass
class If:
    cond: object
    body: list
    label: str | None


@dataclass
class Return:
    value: str


class Parser:
    def __init__(self, toks: list[Tok]):
        self.t = toks
        self.i = 0
        self.pending_label: str | None = None

    def peek(self, k=0):
        j = self.i + k
        while j < len(self.t) and self.t[j].kind == "bcomment":
            j += 1 if k == 0 else 1
        return self.t[j] if j < len(self.t) else None

    def _skip_comments(self):
        while self.i < len(self.t) and self.t[self.i].kind == "bcomment":
            self.pending_label = self.t[self.i].text[2:-2].strip()
            self.i += 1

    def next(self) -> Tok:
        self._skip_comments()
        tok = self.t[self.i]
        self.i += 1
        return tok

    def cur(self) -> Tok | None:
        self._skip_comments()
        return self.t[self.i] if self.i < len(self.t) else None

    def expect(self, text):
        tok = self.next()
        if tok.text != text:
            raise ValueError(f"expected {text!r}, got {tok.text!r} at token {self.i}")

    def parse_block_items(self, until_brace: bool) -> list:
        items = []
        while True:
            tok = self.cur()
            if tok is None:
                if until_brace:
                    raise ValueError("unexpected end")
                return items
            if tok.text == "}":
                if not until_brace:
                    raise ValueError("stray }")
                self.next()
                return items
            items.extend(self.parse_statement())

    def parse_statement(self) -> list:
        tok = self.cur()
        if tok.text == "{":
            self.next()
            return self.parse_block_items(True)
        if tok.text == "if":
            self.next()
            self.pending_label = None
            self.expect("(")
            self._skip_comments()
            label = self.pending_label
            cond = self.parse_expr()
            self.expect(")")
            if self.cur().text == "{":
                self.next()
                body = self.parse_block_items(True)
            else:
                body = self.parse_statement()
            return [If(cond, body, label)]
        if tok.text == "return":
            self.next()
            val = self.next().text
            self.expect(";")
            return [Return(val)]
        if tok.text in ("float", "auto", "double", "const", "volatile"):
            while self.cur().text in ("float", "auto", "double", "const", "volatile"):
                self.next()
        name = self.next()
        if name.kind != "ident":
            raise ValueError(f"unexpected token {name.text!r}")
        self.expect("=")
        expr = self.parse_expr()
        self.expect(";")
        return [Assign(name.text, expr)]

    def parse_expr(self):
        a = self.parse_term()
        while self.cur().text in ("+", "-"):
            op = self.next().text
            b = self.parse_term()
            a = Bin(op, a, b)
        return a

    def parse_term(self):
        a = self.parse_unary()
        while self.cur().text in ("*", "/"):
            op = self.next().text
            b = self.parse_unary()
            a = Bin(op, a, b)
        return a

    def parse_unary(self):
        if self.cur().text == "-":
            self.next()
            return Un("-", self.parse_unary())
        if self.cur().text == "+":
            self.next()
            return self.parse_unary()
        return self.parse_primary()

    def parse_primary(self):
        tok = self.next()
        if tok.kind == "num":
            return Num(float(tok.text.rstrip("fF")))
        if tok.text == "(":
            e = self.parse_expr()
            self.expect(")")
            return e
        if tok.kind == "ident":
            nxt = self.cur()
            if nxt is not None and nxt.text == "(":
                self.next()
                args = []
                if self.cur().text != ")":
                    while True:
                        args.append(self.parse_expr())
                        if self.cur().text == ",":
                            self.next()
                            continue
                        break
                self.expect(")")
                return Call(tok.text, args)
            if nxt is not None and nxt.text == "[":
                self.next()
                idx = self.next()
                self.expect("]")
                return Index(tok.text, int(idx.text))
            return Var(tok.text)
        raise ValueError(f"unexpected token {tok.text!r}")


# ---------------------------------------------------------------- evaluator

_FUNCS = {
    "sin": np.sin, "cos": np.cos, "__sinf": np.sin, "__cosf": np.cos,
    "sinf": np.sin, "cosf": np.cos,
}


@dataclass
class Val:
    v: np.ndarray      # float64 values, shape [B]
    deps: frozenset    # joint indices the value depends on


@dataclass
class Check:
    kind: str                # "env" or "self"
    args: list               # list[Val]; env: x,y,z,r ; self: x,y,z,r,x,y,z,r
    label: str | None
    children: list = field(default_factory=list)
    lone: bool = False       # body is just `return false`
    fk_ops_before: int = 0   # FK arithmetic executed before this check
    fk_trig_before: int = 0


class Evaluator:
    def __init__(self, qname: str, Q: np.ndarray):
        self.qname = qname
        self.Q = Q
        self.env: dict[str, Val] = {}
        self.ops = 0
        self.trig = 0

    def ev(self, e) -> Val:
        B = self.Q.shape[0]
        if isinstance(e, Num):
            return Val(np.full(B, e.v), frozenset())
        if isinstance(e, Var):
            return self.env[e.name]
        if isinstance(e, Index):
            if e.name != self.qname:
                raise ValueError(f"unknown array {e.name}")
            return Val(self.Q[:, e.idx].astype(np.float64), frozenset([e.idx]))
        if isinstance(e, Un):
            a = self.ev(e.a)
            return Val(-a.v, a.deps)
        if isinstance(e, Bin):
            a = self.ev(e.a)
            b = self.ev(e.b)
            if e.op == "+":
                v = a.v + b.v
            elif e.op == "-":
                v = a.v - b.v
            elif e.op == "*":
                v = a.v * b.v
            else:
                v = a.v / b.v
            return Val(v, a.deps | b.deps)
        if isinstance(e, Call):
            if e.fn in _FUNCS:
                a = self.ev(e.args[0])
                return Val(_FUNCS[e.fn](a.v), a.deps)
            raise ValueError(f"unexpected call {e.fn}")
        raise TypeError(e)

    def run(self, stmts: list) -> list[Check]:
        checks: list[Check] = []
        self._run(stmts, checks, depth=0)
        return checks

    def _check_of(self, cond) -> Check:
        if not isinstance(cond, Call):
            raise ValueError("if-condition is not a call")
        if cond.fn == "sphere_environment_in_collision":
            if not (isinstance(cond.args[0], Var) and cond.args[0].name == "environment"):
                raise ValueError("env check without environment argument")
            return Check("env", [self.ev(a) for a in cond.args[1:]], None)
        if cond.fn == "sphere_sphere_self_collision":
            return Check("self", [self.ev(a) for a in cond.args], None)
        raise ValueError(f"unexpected condition call {cond.fn}")

    def _run(self, stmts, checks, depth):
        for s in stmts:
            if isinstance(s, Assign):
                self.env[s.name] = self.ev(s.expr)
                o, t = count_fk_ops([s])
                self.ops += o
                self.trig += t
            elif isinstance(s, If):
                c = self._check_of(s.cond)
                c.label = s.label
                c.fk_ops_before = self.ops
                c.fk_trig_before = self.trig
                body = s.body
                if len(body) == 1 and isinstance(body[0], Return) and body[0].value == "false":
                    c.lone = True
                else:
                    if depth > 0:
                        raise ValueError("nesting deeper than two levels")
                    self._run(body, c.children, depth + 1)
                    for ch in c.children:
                        if not ch.lone:
                            raise ValueError("nested check with a non-trivial body")
                        if ch.kind != c.kind:
                            raise ValueError("gate and fine checks of different kinds")
                checks.append(c)
            elif isinstance(s, Return):
                if s.value != "true" or depth != 0:
                    raise ValueError("unexpected return")
            else:
                raise TypeError(s)


def count_fk_ops(stmts) -> tuple[int, int]:
    """Source-level arithmetic in the FK statements: (binary + - * / ops, sin/cos calls)."""
    ops = 0
    trig = 0

    def walk(e):
        nonlocal ops, trig
        if isinstance(e, Bin):
            ops += 1
            walk(e.a)
            walk(e.b)
        elif isinstance(e, Un):
            walk(e.a)
        elif isinstance(e, Call):
            if e.fn in _FUNCS:
                trig += 1
            for a in e.args:
                walk(a)

    def visit(ss):
        for s in ss:
            if isinstance(s, Assign):
                walk(s.expr)
            elif isinstance(s, If):
                visit(s.body)

    visit(stmts)
    return ops, trig


def parse_robot(src: str, robot: str):
    body = function_body(src, robot)
    qname = function_signature_input(src, robot)
    stmts = Parser(tokenize(body)).parse_block_items(False)
    return qname, stmts


def evaluate(qname, stmts, Q) -> list[Check]:
    return Evaluator(qname, Q).run(stmts)


# ------------------------------------------------------------- model fitting

def _f32(x):
    return float(np.float32(x))


def _kabsch(P, Qp):
    """Rigid transform (R, t) with Qp ~= R P + t (rows are points)."""
    cp = P.mean(0)
    cq = Qp.mean(0)
    H = (P - cp).T @ (Qp - cq)
    U, _, Vt = np.linalg.svd(H)
    d = np.sign(np.linalg.det(Vt.T @ U.T))
    D = np.diag([1.0, 1.0, d])
    R = Vt.T @ D @ U.T
    t = cq - R @ cp
    return R, t


def _snap(x, tol):
    x = np.array(x, dtype=np.float64)
    x[np.abs(x) < tol] = 0.0
    x[np.abs(x - 1.0) < tol] = 1.0
    x[np.abs(x + 1.0) < tol] = -1.0
    return x


def sphere_key(check_args, j0):
    """Key identifying a sphere by its evaluated (x,y,z,r) over the probe set."""
    return tuple(np.round(np.concatenate([a.v for a in check_args[j0:j0 + 4]]), 9))


def build_model(src: str, robot: str, seed: int = 0) -> dict:
    dim = ROBOT_DIM[robot]
    qname, stmts = parse_robot(src, robot)
    rng = np.random.default_rng(seed)

    # Probe set: zero pose, single-joint sweeps, and random poses.
    thetas = [0.35, -0.6, 0.9, 1.3]
    probes = [np.zeros(dim)]
    for j in range(dim):
        for th in thetas:
            q = np.zeros(dim)
            q[j] = th
            probes.append(q)
    n_rand = 64
    probes.extend(rng.uniform(-2.0, 2.0, size=(n_rand, dim)))
    Q = np.array(probes)
    checks = evaluate(qname, stmts, Q)

    # ---- unique spheres
    spheres: list[dict] = []
    index: dict = {}

    def sphere_id(args, j0):
        key = sphere_key(args, j0)
        if key not in index:
            x, y, z, r = args[j0:j0 + 4]
            if r.deps or np.ptp(r.v) != 0.0:
                raise ValueError("radius depends on q")
            index[key] = len(spheres)
            spheres.append({
                "world": np.stack([x.v, y.v, z.v], 1),
                "deps": x.deps | y.deps | z.deps,
                "radius": _f32(r.v[0]),
            })
        return index[key]

    env_prog = []
    self_prog = []
    program = []
    for c in checks:
        program.append({"kind": c.kind, "index": len(env_prog) if c.kind == "env" else len(self_prog),
                        "fk_ops_before": c.fk_ops_before, "fk_trig_before": c.fk_trig_before})
        if c.kind == "env":
            g = sphere_id(c.args, 0)
            fines = [sphere_id(ch.args, 0) for ch in c.children]
            env_prog.append({"label": c.label, "gate": g, "fine": fines, "lone": c.lone})
        else:
            ga = sphere_id(c.args, 0)
            gb = sphere_id(c.args, 4)
            fines = [[sphere_id(ch.args, 0), sphere_id(ch.args, 4)] for ch in c.children]
            self_prog.append({"label": c.label, "gate": [ga, gb], "fine": fines, "lone": c.lone})

    # ---- kinematic tree from dependency sets
    all_deps = [s["deps"] for s in spheres]
    moving = sorted(set().union(*all_deps))
    if moving != list(range(dim)):
        raise ValueError(f"joints without spheres: {set(range(dim)) - set(moving)}")
    anc = {}
    for j in range(dim):
        sets = [d for d in all_deps if j in d]
        inter = frozenset.intersection(*sets)
        anc[j] = inter - {j}
    parent = {}
    for j in range(dim):
        cands = sorted(anc[j], key=lambda k: len(anc[k]))
        parent[j] = cands[-1] if cands else -1
        if cands and anc[parent[j]] | {parent[j]} != anc[j]:
            raise ValueError(f"joint {j}: ancestors do not form a chain")
    depth_of = {j: len(anc[j]) for j in range(dim)}
    order = sorted(range(dim), key=lambda j: (depth_of[j], j))

    # frame (attachment joint) of each sphere = deepest joint it depends on
    for s in spheres:
        d = s["deps"]
        if not d:
            s["frame"] = -1
        else:
            deepest = max(d, key=lambda k: depth_of[k])
            if anc[deepest] | {deepest} != d:
                raise ValueError("sphere dependencies do not form an ancestor chain")
            s["frame"] = deepest

    # probe indices
    def probe_idx(j, k):
        return 1 + j * len(thetas) + k

    # ---- fit each joint's axis at the zero pose (world frame)
    W_R = {-1: np.eye(3)}
    W_t = {-1: np.zeros(3)}
    joints = {}
    for j in order:
        pts = [s for s in spheres if j in s["deps"]]
        P0 = np.array([s["world"][0] for s in pts])
        kind = None
        axis = None
        point = None
        for k, th in enumerate(thetas):
            P1 = np.array([s["world"][probe_idx(j, k)] for s in pts])
            R, t = _kabsch(P0, P1)
            resid = np.abs((P0 @ R.T + t) - P1).max()
            if resid > 1e-6:
                raise ValueError(f"joint {j}: non-rigid motion (resid {resid})")
            ang = np.arccos(np.clip((np.trace(R) - 1) / 2, -1, 1))
            if ang < 1e-9:
                k_kind = "prismatic"
                ax = t / th
                if abs(np.linalg.norm(ax) - 1) > 1e-6:
                    raise ValueError(f"joint {j}: prismatic scale {np.linalg.norm(ax)}")
                pt = None
            else:
                k_kind = "revolute"
                w = np.array([R[2, 1] - R[1, 2], R[0, 2] - R[2, 0], R[1, 0] - R[0, 1]])
                ax = w / np.linalg.norm(w)
                # rotation angle must equal th about ax (sign fixes ax)
                if abs(ang - abs(th)) > 1e-6:
                    raise ValueError(f"joint {j}: angle {ang} vs {th}")
                if th < 0:
                    ax = -ax
                # point on axis: (I - R) p = t, least squares
                A = np.eye(3) - R
                pt, *_ = np.linalg.lstsq(A, t, rcond=None)
            if kind is None:
                kind, axis, point = k_kind, ax, pt
            else:
                if kind != k_kind or np.abs(axis - ax).max() > 1e-6:
                    raise ValueError(f"joint {j}: inconsistent axis fits")
        axis = axis / np.linalg.norm(axis)
        p = parent[j]
        Rp, tp = W_R[p], W_t[p]
        if kind == "revolute":
            # choose the point on the axis closest to the parent frame origin
            point = point + axis * np.dot(tp - point, axis)
            origin = point
        else:
            origin = tp.copy()
        # local x: the parent axis most perpendicular to the joint axis
        best = None
        for c in range(3):
            for sgn in (1.0, -1.0):
                v = sgn * Rp[:, c]
                perp = v - axis * np.dot(v, axis)
                nv = np.linalg.norm(perp)
                score = (nv, sgn > 0, -c)
                if best is None or score > best[0]:
                    best = (score, perp / nv)
        x = best[1]
        y = np.cross(axis, x)
        Rj = np.stack([x, y, axis], 1)
        # fixed transform from the parent frame
        RG = _snap(Rp.T @ Rj, 1e-9)
        tG = _snap(Rp.T @ (origin - tp), 1e-9)
        # re-orthonormalise after snapping
        U, _, Vt = np.linalg.svd(RG)
        RG = U @ Vt
        RG = _snap(RG, 1e-12)
        W_R[j] = Rp @ RG
        W_t[j] = tp + Rp @ tG
        joints[j] = {"parent": p, "type": kind, "R": RG, "t": tG}

    # ---- sphere local centres
    for s in spheres:
        f = s["frame"]
        local = W_R[f].T @ (s["world"][0] - W_t[f])
        s["local"] = _snap(local, 1e-9)

    # ---- validate the chain against every probe
    def fk_world(q):
        Rw = {-1: np.eye(3)}
        tw = {-1: np.zeros(3)}
        for j in order:
            jt = joints[j]
            p = jt["parent"]
            A = Rw[p] @ jt["R"]
            t = tw[p] + Rw[p] @ jt["t"]
            if jt["type"] == "revolute":
                c, sn = np.cos(q[j]), np.sin(q[j])
                Rz = np.array([[c, -sn, 0], [sn, c, 0], [0, 0, 1.0]])
                Rw[j] = A @ Rz
                tw[j] = t
            else:
                Rw[j] = A
                tw[j] = t + A[:, 2] * q[j]
        return Rw, tw

    max_err = 0.0
    for b in range(Q.shape[0]):
        Rw, tw = fk_world(Q[b])
        for s in spheres:
            f = s["frame"]
            pw = tw[f] + Rw[f] @ s["local"]
            max_err = max(max_err, float(np.abs(pw - s["world"][b]).max()))

    # order joints as a list indexed by joint id
    joint_list = []
    for j in range(dim):
        jt = joints[j]
        joint_list.append({
            "parent": int(jt["parent"]),
            "type": jt["type"],
            "R": [[float(v) for v in row] for row in jt["R"]],
            "t": [float(v) for v in jt["t"]],
        })
    model = {
        "robot": robot,
        "dimension": dim,
        "joint_order": [int(j) for j in order],
        "joints": joint_list,
        "spheres": [
            {"frame": int(s["frame"]), "local": [float(v) for v in s["local"]], "radius": s["radius"]}
            for s in spheres
        ],
        "env_program": env_prog,
        "self_program": self_prog,
        "program": program,
        "fit_max_error_m": max_err,
        "ref_fk_flops": count_fk_ops(stmts)[0],
        "ref_fk_trig": count_fk_ops(stmts)[1],
        "derivation": (
            "extract_model.py: float64 evaluation of the reference's generated "
            f"fkcc<{ROBOT_TYPE[robot]}> sphere expressions (src/utils.cuh), fitted to a kinematic tree"
        ),
    }
    return model


def main(argv):
    if len(argv) != 4:
        print(__doc__)
        return 2
    src = open(argv[1]).read()
    model = build_model(src, argv[2])
    with open(argv[3], "w") as f:
        json.dump(model, f, indent=1)
    n_env = sum(len(g["fine"]) + 1 for g in model["env_program"])
    n_self = sum(len(g["fine"]) + 1 for g in model["self_program"])
    print(f"{argv[2]}: {len(model['spheres'])} spheres, {len(model['env_program'])} env groups "
          f"({n_env} calls), {len(model['self_program'])} self groups ({n_self} calls), "
          f"fit max error {model['fit_max_error_m']:.3e} m")
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv))
