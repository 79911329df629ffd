"""MotionBenchMaker problem sets and Halton edges: the inputs of the README run
(`batch_cc panda 100 1000`, scripts/batch_cc_mbm.cpp:19-139), value for value.

Obstacle construction follows the reference factories
(src/collision/factory.hh:28-62 cuboid, :118-129 and :165-185 cylinder) on a
host without FMA instructions, so every product and sum rounds on its own:

  rotation   R = AngleAxis(phi, Z) * AngleAxis(theta, Y) * AngleAxis(rho, X)
             held as a quaternion.  AngleAxis(a, u) -> w = cosf(a/2),
             vec = sinf(a/2) * u; products use the four-lane formula
               x = (ax bw - az by) + (ay bz + aw bx)
               y = (ay bw - ax bz) + (az bx + aw by)
               z = (az bw - ay bx) + (ax by + aw bz)
               w = (aw bw - ax bx) - (az bz + ay by)
  axis_k     q * e_k by  uv = vec x v;  uv = uv + uv;
             out = (v + w uv) + vec x uv
  cylinder   p1 = c + R (0,0,L/2), p2 = c + R (0,0,-L/2) with the rotation
             matrix of q (column 2: 2xz + 2yw, 2yz - 2xw, 1 - (2xx + 2yy));
             v = p2 - p1, rdv = float(1.0 / double(v.v)); stored as a capsule
  sinf/cosf  the host C library's (the reference calls std::sin/std::cos on
             float)

For the yaw-only rotations of every non-"box" family every product that
involves a zero is exact, so the axes are (1 - 2s^2, 2cs, 0), (-2cs,
1 - 2s^2, 0), (0, 0, 1) with s, c the sine and cosine of phi/2, whatever the
evaluation order, and the cylinder's endpoints are c +- (0, 0, L/2).

Environment selection (scripts/batch_cc_mbm.cpp:97-123): families in sorted
name order, valid problems only, until num_envs are taken.  Cylinders of the
"box" family become cuboids with half-extents (r, r, r/2) and the cylinder's
rotation (:33-45); other families' cylinders become capsules (:46-58); boxes
become cuboids (:61-70).  Edges: two draws of rng::Halton<dim> per edge,
each scaled by Robot::scale_cfg (:125-139, src/random/halton.hh:22-121,
src/Robots.hh:51-63).
"""
from __future__ import annotations

import ctypes
import ctypes.util
import json
import os

import numpy as np

from . import DIMS, PackedScene
from .workloads import scale_cfg

F = np.float32
_libm = ctypes.CDLL(ctypes.util.find_library("m"))
for _fn in ("sinf", "cosf"):
    getattr(_libm, _fn).restype = ctypes.c_float
    getattr(_libm, _fn).argtypes = [ctypes.c_float]


def sinf(x) -> np.float32:
    return F(_libm.sinf(float(F(x))))


def cosf(x) -> np.float32:
    return F(_libm.cosf(float(F(x))))


# --- quaternions (x, y, z, w), float32 ------------------------------------

def angle_axis(angle, axis) -> tuple:
    ha = F(0.5) * F(angle)
    s, c = sinf(ha), cosf(ha)
    return (s * F(axis[0]), s * F(axis[1]), s * F(axis[2]), c)


def quat_mul(a, b) -> tuple:
    ax, ay, az, aw = a
    bx, by, bz, bw = b
    return ((ax * bw - az * by) + (ay * bz + aw * bx),
            (ay * bw - ax * bz) + (az * bx + aw * by),
            (az * bw - ay * bx) + (ax * by + aw * bz),
            (aw * bw - ax * bx) - (az * bz + ay * by))


def _cross(a, b):
    return (a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0])


def quat_rotate(q, v) -> tuple:
    vec, w = q[:3], q[3]
    v = tuple(F(t) for t in v)
    uv = _cross(vec, v)
    uv = tuple(t + t for t in uv)
    c2 = _cross(vec, uv)
    return tuple((v[i] + w * uv[i]) + c2[i] for i in range(3))


def euler_xyz_quat(euler) -> tuple:
    rho, theta, phi = (F(t) for t in euler)
    return quat_mul(quat_mul(angle_axis(phi, (0, 0, 1)), angle_axis(theta, (0, 1, 0))), angle_axis(rho, (1, 0, 0)))


# --- factories -------------------------------------------------------------

def cuboid(center, euler_xyz, half_extents) -> np.ndarray:
    """factory::cuboid::array -> (x y z, axis_1, axis_2, axis_3, r1 r2 r3)."""
    q = euler_xyz_quat(euler_xyz)
    axes = [quat_rotate(q, e) for e in ((1, 0, 0), (0, 1, 0), (0, 0, 1))]
    c = [F(t) for t in center]
    h = [F(t) for t in half_extents]
    return np.array(c + [t for a in axes for t in a] + h, np.float32)


def cylinder_center(center, euler_xyz, radius, length) -> np.ndarray:
    """factory::cylinder::center::array -> capsule (x1 y1 z1 xv yv zv r rdv)."""
    x, y, z, w = euler_xyz_quat(euler_xyz)
    tx, ty, tz = F(2) * x, F(2) * y, F(2) * z
    col2 = (tz * x + ty * w, tz * y - tx * w, F(1) - (tx * x + ty * y))
    c = [F(t) for t in center]
    half = F(length) / F(2)
    p1 = [c[i] + col2[i] * half for i in range(3)]
    p2 = [c[i] + col2[i] * (-half) for i in range(3)]
    v = [p2[i] - p1[i] for i in range(3)]
    dot = (v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]
    rdv = F(1.0 / float(dot)) if dot != 0 else F(np.inf)
    return np.array(p1 + v + [F(radius), rdv], np.float32)


def sphere(center, radius) -> np.ndarray:
    return np.array([F(t) for t in center] + [F(radius)], np.float32)


# --- problem files ---------------------------------------------------------

def load_problems(src) -> dict:
    """A problem file (path) or its parsed dictionary."""
    if isinstance(src, (str, os.PathLike)):
        with open(src) as f:
            return json.load(f)
    return src


def problem_dict_to_env(problem: dict, family: str) -> dict:
    """One problem -> {spheres, capsules, cuboids} records
    (scripts/batch_cc_mbm.cpp:19-93)."""
    sph = [sphere(o["position"], o["radius"]) for o in problem.get("sphere", [])]
    cap, cub = [], []
    for o in problem.get("cylinder", []):
        if family == "box":
            r = F(o["radius"])
            cub.append(cuboid(o["position"], o["orientation_euler_xyz"], (r, r, r / F(2))))
        else:
            cap.append(cylinder_center(o["position"], o["orientation_euler_xyz"], o["radius"], o["length"]))
    for o in problem.get("box", []):
        cub.append(cuboid(o["position"], o["orientation_euler_xyz"], o["half_extents"]))
    return {"spheres": np.array(sph, np.float32).reshape(-1, 4),
            "capsules": np.array(cap, np.float32).reshape(-1, 8),
            "cuboids": np.array(cub, np.float32).reshape(-1, 15)}


def select_problems(problems: dict, num_envs: int, families=None) -> list:
    """[(family, problem)] in the driver's order: families sorted by name,
    valid problems only, stopping at num_envs (fewer if the set runs out)."""
    out = []
    for fam in sorted(problems["problems"]):
        if families is not None and fam not in families:
            continue
        for p in problems["problems"][fam]:
            if not p["valid"]:
                continue
            out.append((fam, p))
            if len(out) >= num_envs:
                return out
    return out


def mbm_scene(problems, num_envs: int, families=None) -> PackedScene:
    sel = select_problems(load_problems(problems), num_envs, families)
    return PackedScene.from_envs([problem_dict_to_env(p, fam) for fam, p in sel])


# --- Halton ----------------------------------------------------------------

class Halton:
    """rng::Halton<dim> (src/random/halton.hh:22-121): per-dimension float
    numerator/denominator recurrence with bases 3, 5, 7, ...; after 1e6 draws
    the state restarts with the bases rotated by one."""
    PRIMES = (3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47, 53, 59)
    MAX_ITERATIONS = 1_000_000

    def __init__(self, dim: int, bases=None):
        self.b_init = np.array(bases if bases is not None else self.PRIMES[:dim], np.float32)
        self.reset()

    def reset(self):
        self.iterations = 0
        self.b = self.b_init.copy()
        self.n = np.zeros_like(self.b)
        self.d = np.ones_like(self.b)

    def next(self) -> np.ndarray:
        self.iterations += 1
        if self.iterations > self.MAX_ITERATIONS:
            self.n[:] = 0
            self.d[:] = 1
            self.iterations = 0
            self.b = np.roll(self.b, -1)
        b, n, d = self.b, self.n, self.d
        xf = d - n
        one = xf == F(1)
        y = np.floor(d / b)
        more = ~one & (xf <= y)
        while more.any():
            y = np.where(more, np.floor(y / b), y)
            more &= xf <= y
        self.n = np.where(one, F(1), np.floor((b + F(1)) * y) - xf).astype(np.float32)
        self.d = np.where(one, np.floor(d * b), d).astype(np.float32)
        return (self.n / self.d).astype(np.float32)


def halton_edges(robot: str, num_edges: int) -> np.ndarray:
    """[num_edges, 2, dim]: start and end from consecutive draws, scaled."""
    dim = DIMS[robot]
    rng = Halton(dim)
    u = np.stack([rng.next() for _ in range(2 * num_edges)]).reshape(num_edges, 2, dim)
    return scale_cfg(robot, u.reshape(-1, dim)).reshape(num_edges, 2, dim)


def readme_run(problems, robot: str = "panda", num_envs: int = 100, num_edges: int = 1000):
    """(PackedScene, edges) of `batch_cc <robot> <num_envs> <num_edges>`."""
    return mbm_scene(problems, num_envs), halton_edges(robot, num_edges)
