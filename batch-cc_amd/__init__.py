"""Python binding of the MI355X batch edge collision checker (include/bcc.h).

The checker itself is native: gfx950 HIP kernels behind a C ABI in
``libbcc.so`` next to this file.  This module is a thin ctypes layer for
tests, the benchmark and Python callers.  It never falls back to a CPU
path: if the library is missing, every call raises.

Mirrors the reference's entry point ``batch_cc::batch_cc<Robot>(envs, edges,
resolution, results)`` (src/batch_cc.hh:4-7): `check_dense(robot, scene,
edges, resolution)` returns the edge-major verdict matrix results[edge, env]
(True = the edge collides in that environment, src/batch_cc.cu:131).
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libbcc.so")

ROBOTS = {"panda": 0, "fetch": 1, "baxter": 2, "xarm7": 3}
DIMS = {"panda": 7, "fetch": 8, "baxter": 14, "xarm7": 13}
NUM_SPHERES = {"panda": 70, "fetch": 125, "baxter": 100, "xarm7": 74}
# packed host layout: floats per primitive of kind
# spheres, capsules, z-capsules, cylinders, cuboids, z-cuboids
KIND_STRIDE = (4, 8, 8, 8, 15, 15)
KINDS = ("spheres", "capsules", "z_capsules", "cylinders", "cuboids", "z_cuboids")

BCC_OK = 0

_lib = None


class BCCError(RuntimeError):
    """Raised for any non-zero status of the C ABI (the reference throws
    std::runtime_error for setup errors, src/batch_cc.cu:25-32)."""


def _sig(lib, name, res, args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = args
    return f


def lib():
    """Load libbcc.so (built by batch-cc_amd/build.py)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise BCCError(f"{LIB_PATH} is missing: run `python batch-cc_amd/build.py` (no CPU fallback exists)")
    L = ctypes.CDLL(LIB_PATH)
    P = ctypes.c_void_p
    I = ctypes.c_int
    _sig(L, "bcc_last_error", ctypes.c_char_p, [])
    _sig(L, "bcc_robot_dimension", I, [I])
    _sig(L, "bcc_robot_num_spheres", I, [I])
    _sig(L, "bcc_device_count", I, [ctypes.POINTER(I)])
    _sig(L, "bcc_scene_create", I, [P, I, ctypes.POINTER(P)])
    _sig(L, "bcc_scene_create_packed", I, [P, P, P, I, ctypes.POINTER(P)])
    _sig(L, "bcc_scene_destroy", I, [P])
    _sig(L, "bcc_scene_num_envs", I, [P])
    _sig(L, "bcc_check_dense", I, [I, P, P, I, I, P])
    _sig(L, "bcc_check_dense_async", I, [I, P, P, I, I, P, P])
    _sig(L, "bcc_check_dense_endpoints", I, [I, P, P, I, I, P, P])
    _sig(L, "bcc_time_dense", I, [I, P, P, I, I, P, I, I, P, ctypes.POINTER(ctypes.c_double)])
    _sig(L, "bcc_work_dense", I, [I, P, P, I, I, P, P, P, P])
    _sig(L, "bcc_work_pairs", I, [I, P, P, I, P, I, P, P, P, P])
    _sig(L, "bcc_check_ragged", I, [I, P, P, P, I, P])
    _sig(L, "bcc_audit_clearance", I, [I, P, P, I, P, I, P, P])
    _sig(L, "bcc_check_ragged_noalloc", I, [I, P, P, P, ctypes.c_int64, I, P, P])
    _sig(L, "bcc_ragged_plan_create", I, [I, P, P, P, I, P, I, ctypes.POINTER(P)])
    _sig(L, "bcc_ragged_plan_launch", I, [P, P])
    _sig(L, "bcc_ragged_plan_launch_eager", I, [P, P])
    _sig(L, "bcc_ragged_plan_num_tiles", I, [P])
    _sig(L, "bcc_ragged_plan_time", I, [P, I, I, P, ctypes.POINTER(ctypes.c_double)])
    _sig(L, "bcc_ragged_plan_destroy", I, [P])
    _sig(L, "bcc_ragged_plan_create_tuned", I, [I, P, P, P, I, P, I, I, ctypes.POINTER(P)])
    _sig(L, "bcc_ragged_plan_config", I, [P, ctypes.POINTER(I), ctypes.POINTER(I)])
    _sig(L, "bcc_ragged_plan_create_ordered", I, [I, P, P, P, I, P, P, I, I, ctypes.POINTER(P)])
    _sig(L, "bcc_ragged_group_create", I, [P, I, ctypes.POINTER(P)])
    _sig(L, "bcc_ragged_group_launch", I, [P, P])
    _sig(L, "bcc_ragged_group_time", I, [P, I, I, P, ctypes.POINTER(ctypes.c_double)])
    _sig(L, "bcc_ragged_group_destroy", I, [P])
    _sig(L, "bcc_check_dense_ex", I, [I, P, P, I, I, P, P, P, P])
    _sig(L, "bcc_time_dense_ex", I, [I, P, P, I, I, P, P, I, I, P, ctypes.POINTER(ctypes.c_double)])
    _sig(L, "bcc_model_num_params", I, [])
    _sig(L, "bcc_model_create", I, [I, ctypes.c_uint64, ctypes.POINTER(P)])
    _sig(L, "bcc_model_destroy", I, [P])
    _sig(L, "bcc_model_get_params", I, [P, P])
    _sig(L, "bcc_model_set_params", I, [P, P])
    _sig(L, "bcc_model_save", I, [P, ctypes.c_char_p])
    _sig(L, "bcc_model_load", I, [ctypes.c_char_p, ctypes.POINTER(P)])
    _sig(L, "bcc_encode_envs", I, [P, P])
    _sig(L, "bcc_model_features", I, [P, P, P, ctypes.c_int64, P, P, P, P])
    _sig(L, "bcc_model_order_pairs", I, [P, P, P, ctypes.c_int64, P, P, P, P, P])
    _sig(L, "bcc_label_pairs", I, [I, P, P, I, P, P, I, P, P])
    _sig(L, "bcc_label_pairs_steps", I, [I, P, P, I, P, P, I, I, P, P])
    _sig(L, "bcc_model_grads", I, [P, P, P, I, P, ctypes.POINTER(ctypes.c_float)])
    _sig(L, "bcc_model_train", I, [P, P, P, I, I, I, ctypes.c_float, P])
    _sig(L, "bcc_model_logits", I, [P, P, ctypes.c_int64, P])
    _sig(L, "bcc_model_order_dense", I, [P, P, P, I, P, P])
    _sig(L, "bcc_time_learned_dense", I, [P, P, P, I, I, P, P, I, I, P, ctypes.POINTER(ctypes.c_double),
                                           ctypes.POINTER(ctypes.c_double)])
    _sig(L, "bcc_model_logits_dense", I, [P, P, P, I, P, P, P])
    _sig(L, "bcc_debug_gemm", I, [I, I, P, I, P, I, P, I, I, I, I, ctypes.c_float, ctypes.c_float])
    _sig(L, "bcc_malloc", I, [ctypes.POINTER(P), ctypes.c_size_t])
    _sig(L, "bcc_free", I, [P])
    _sig(L, "bcc_memcpy_htod", I, [P, P, ctypes.c_size_t])
    _sig(L, "bcc_memcpy_dtoh", I, [P, P, ctypes.c_size_t])
    _sig(L, "bcc_memset", I, [P, I, ctypes.c_size_t])
    _sig(L, "bcc_synchronize", I, [])
    _sig(L, "bcc_debug_fk", I, [I, P, I, P])
    _sig(L, "bcc_debug_sincos", I, [P, I, P])
    _sig(L, "bcc_debug_samples", I, [I, P, I, I, P, P, P])
    _sig(L, "bcc_debug_prims", I, [P, I, P])
    _sig(L, "bcc_debug_copy", I, [P, P, ctypes.c_size_t])
    _sig(L, "bcc_debug_valu", I, [I, I, I, P, ctypes.POINTER(ctypes.c_double)])
    _sig(L, "bcc_debug_force_variant", I, [I])
    _sig(L, "bcc_debug_cull_min", I, [I])
    _sig(L, "bcc_debug_scene_cull_min", I, [P, ctypes.POINTER(I)])
    _sig(L, "bcc_debug_dense_chunk", I, [I])
    _sig(L, "bcc_debug_half_boxes", I, [P, ctypes.c_uint64, P])
    _sig(L, "bcc_check_dense_learned", I, [I, P, P, P, I, I, P, P, P, P, P])
    _sig(L, "bcc_time_dense_learned", I, [I, P, P, P, I, I, P, I, I, P, ctypes.POINTER(ctypes.c_double)])
    _sig(L, "bcc_ragged_plan_create_learned", I, [I, P, P, P, P, I, P, I, I, ctypes.POINTER(P)])
    _sig(L, "bcc_ragged_plan_create_shaped", I, [I, P, P, P, I, P, I, I, I, ctypes.POINTER(P)])
    _sig(L, "bcc_ragged_plan_tile_steps", I, [P])
    _sig(L, "bcc_debug_learned_tile", I, [I])
    _sig(L, "bcc_debug_learned_far", I, [I])
    _sig(L, "bcc_debug_learned_passes", I, [I])
    _sig(L, "bcc_debug_learned_min_steps", I, [I])
    _sig(L, "bcc_debug_learned_locality", I, [I])
    _sig(L, "bcc_debug_learned_small", I, [ctypes.c_int64])
    _sig(L, "bcc_learned_small_pairs", ctypes.c_int64, [])
    _lib = L
    return L


def _check(rc):
    if rc != BCC_OK:
        raise BCCError(f"bcc error {rc}: {lib().bcc_last_error().decode()}")


def robot_id(robot) -> int:
    if isinstance(robot, str):
        return ROBOTS[robot]
    return int(robot)


def _ptr(a: np.ndarray):
    return ctypes.c_void_p(a.ctypes.data)


# ------------------------------------------------------------ host scenes

@dataclass
class PackedScene:
    """Environments in the packed host layout of bcc_scene_create_packed:
    env e's floats start at data[env_off[e]] and hold counts[e, k] records
    of kind k (spheres, capsules, z-capsules, cylinders, cuboids, z-cuboids)
    in that order, with 4, 8, 8, 8, 15, 15 floats each."""
    data: np.ndarray      # float32 [total]
    env_off: np.ndarray   # int64 [E]
    counts: np.ndarray    # int32 [E, 6]

    @property
    def num_envs(self) -> int:
        return int(self.counts.shape[0])

    def env(self, e: int) -> dict:
        out = {}
        off = int(self.env_off[e])
        for k, name in enumerate(KINDS):
            n = int(self.counts[e, k])
            out[name] = self.data[off:off + n * KIND_STRIDE[k]].reshape(n, KIND_STRIDE[k])
            off += n * KIND_STRIDE[k]
        return out

    @staticmethod
    def from_envs(envs: list[dict]) -> "PackedScene":
        """envs: list of {kind: float32 array [n, stride]} (missing kinds = 0)."""
        chunks = []
        counts = np.zeros((len(envs), 6), np.int32)
        offs = np.zeros(len(envs), np.int64)
        pos = 0
        for e, env in enumerate(envs):
            offs[e] = pos
            for k, name in enumerate(KINDS):
                a = env.get(name)
                if a is None or len(a) == 0:
                    continue
                a = np.ascontiguousarray(a, dtype=np.float32).reshape(-1, KIND_STRIDE[k])
                counts[e, k] = a.shape[0]
                chunks.append(a.ravel())
                pos += a.size
        data = np.concatenate(chunks) if chunks else np.zeros(1, np.float32)
        return PackedScene(np.ascontiguousarray(data, np.float32), offs, counts)


class Scene:
    """Device-resident environments (bcc_scene handle)."""

    def __init__(self, packed: PackedScene):
        self.packed = packed
        self.handle = ctypes.c_void_p()
        d = np.ascontiguousarray(packed.data, np.float32)
        o = np.ascontiguousarray(packed.env_off, np.int64)
        c = np.ascontiguousarray(packed.counts, np.int32)
        _check(lib().bcc_scene_create_packed(_ptr(d), _ptr(o), _ptr(c), packed.num_envs, ctypes.byref(self.handle)))

    @property
    def num_envs(self) -> int:
        return lib().bcc_scene_num_envs(self.handle)

    @property
    def cull_min(self) -> int:
        """Culling threshold stamped at creation (0: every sweep culls)."""
        v = ctypes.c_int()
        _check(lib().bcc_debug_scene_cull_min(self.handle, ctypes.byref(v)))
        return v.value

    def close(self):
        if self.handle:
            _check(lib().bcc_scene_destroy(self.handle))
            self.handle = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------------ device buffers

class DeviceBuffer:
    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        self.ptr = ctypes.c_void_p()
        _check(lib().bcc_malloc(ctypes.byref(self.ptr), max(self.nbytes, 1)))

    @staticmethod
    def from_array(a: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        b = DeviceBuffer(a.nbytes)
        if a.nbytes:
            _check(lib().bcc_memcpy_htod(b.ptr, _ptr(a), a.nbytes))
        return b

    def to_array(self, dtype, shape) -> np.ndarray:
        out = np.empty(shape, dtype=dtype)
        if out.nbytes:
            _check(lib().bcc_memcpy_dtoh(_ptr(out), self.ptr, out.nbytes))
        return out

    def zero(self):
        _check(lib().bcc_memset(self.ptr, 0, max(self.nbytes, 1)))

    def free(self):
        if self.ptr:
            _check(lib().bcc_free(self.ptr))
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def device_count() -> int:
    n = ctypes.c_int(0)
    rc = lib().bcc_device_count(ctypes.byref(n))
    return n.value if rc == BCC_OK else 0


def synchronize():
    _check(lib().bcc_synchronize())


# ------------------------------------------------------------ checks

def _edges_array(robot, edges) -> np.ndarray:
    dim = DIMS[robot] if isinstance(robot, str) else list(DIMS.values())[robot]
    e = np.ascontiguousarray(edges, dtype=np.float32)
    if e.ndim != 3 or e.shape[1:] != (2, dim):
        raise BCCError(f"edges must be float32 [M, 2, {dim}], got {e.shape}")
    return e


def check_dense(robot, scene: Scene, edges, resolution: int = 32) -> np.ndarray:
    """Verdicts [M, E] (uint8, 1 = collision) for every edge in every env."""
    e = _edges_array(robot, edges)
    M, E = e.shape[0], scene.num_envs
    out = np.zeros((M, E), np.uint8)
    _check(lib().bcc_check_dense(robot_id(robot), scene.handle, _ptr(e), M, int(resolution), _ptr(out)))
    return out


def check_dense_device(robot, scene: Scene, d_edges: DeviceBuffer, num_edges: int, resolution: int,
                       d_out: DeviceBuffer, stream=None):
    _check(lib().bcc_check_dense_async(robot_id(robot), scene.handle, d_edges.ptr, int(num_edges), int(resolution),
                                       d_out.ptr, stream))


def check_dense_endpoints(robot, scene: Scene, edges, resolution: int = 32) -> np.ndarray:
    """Verdicts [M, E] with the reference validator's vamp semantics: the
    edge collides when the motion s -> e or the configuration e collides
    (s is the motion's own first sample)."""
    e = _edges_array(robot, edges)
    M, E = e.shape[0], scene.num_envs
    d_e = DeviceBuffer.from_array(e)
    d_o = DeviceBuffer(max(1, M * E))
    _check(lib().bcc_check_dense_endpoints(robot_id(robot), scene.handle, d_e.ptr, M, int(resolution), d_o.ptr,
                                           None))
    synchronize()
    return d_o.to_array(np.uint8, (M, E))


def time_dense(robot, scene: Scene, d_edges: DeviceBuffer, num_edges: int, resolution: int, d_out: DeviceBuffer,
               warmup: int, iters: int, stream=None) -> float:
    """Average kernel milliseconds over `iters` launches (hipEvents)."""
    ms = ctypes.c_double(0.0)
    _check(lib().bcc_time_dense(robot_id(robot), scene.handle, d_edges.ptr, int(num_edges), int(resolution),
                                d_out.ptr, int(warmup), int(iters), stream, ctypes.byref(ms)))
    return ms.value


def work_dense(robot, scene: Scene, d_edges: DeviceBuffer, num_edges: int, resolution: int = 32):
    """Reference-order operation counts per pair: (work[M,E], trig[M,E], verdicts[M,E])."""
    M, E = int(num_edges), scene.num_envs
    dw = DeviceBuffer(M * E * 8)
    dt = DeviceBuffer(M * E * 8)
    do = DeviceBuffer(M * E)
    _check(lib().bcc_work_dense(robot_id(robot), scene.handle, d_edges.ptr, M, int(resolution), dw.ptr, dt.ptr,
                                do.ptr, None))
    synchronize()
    return (dw.to_array(np.uint64, (M, E)), dt.to_array(np.uint64, (M, E)), do.to_array(np.uint8, (M, E)))


def work_pairs(robot, scene: Scene, d_edges: DeviceBuffer, pair_env: np.ndarray, resolution: int = 32):
    """Reference-order counts for an explicit pair list (pair p = edge p in env pair_env[p])."""
    pe = np.ascontiguousarray(pair_env, np.int32)
    n = pe.size
    dpe = DeviceBuffer.from_array(pe)
    dw = DeviceBuffer(n * 8)
    dt = DeviceBuffer(n * 8)
    do = DeviceBuffer(n)
    _check(lib().bcc_work_pairs(robot_id(robot), scene.handle, d_edges.ptr, n, dpe.ptr, int(resolution), dw.ptr,
                                dt.ptr, do.ptr, None))
    synchronize()
    return dw.to_array(np.uint64, (n,)), dt.to_array(np.uint64, (n,)), do.to_array(np.uint8, (n,))


def audit_clearance(robot, scene: Scene, d_edges: DeviceBuffer, num_edges: int, pair_env=None,
                    resolution: int = 32) -> np.ndarray:
    """Smallest gated clearance (metres, FP32) of every pair over all of its
    samples: [num_edges, num_envs] for a dense batch, [num_pairs] with
    pair_env (pair p = edge p in environment pair_env[p])."""
    n = int(num_edges)
    if pair_env is not None:
        pe = np.ascontiguousarray(pair_env, np.int32)
        if pe.size != n:
            raise BCCError("pair_env must have one entry per pair")
        d_pe = DeviceBuffer.from_array(pe)
        shape = (n,)
    else:
        d_pe = None
        shape = (n, scene.num_envs)
    d_c = DeviceBuffer(4 * max(1, int(np.prod(shape))))
    _check(lib().bcc_audit_clearance(robot_id(robot), scene.handle, d_edges.ptr, n, d_pe.ptr if d_pe else None,
                                     int(resolution), d_c.ptr, None))
    synchronize()
    return d_c.to_array(np.float32, shape)


def check_ragged(robot, scene: Scene, env_off, edges, resolution: int = 32) -> np.ndarray:
    """Verdicts [total] for a CSR batch: env e owns pairs env_off[e]:env_off[e+1]."""
    off = np.ascontiguousarray(env_off, np.int64)
    if off.size != scene.num_envs + 1:
        raise BCCError("env_off must have num_envs + 1 entries")
    e = _edges_array(robot, edges)
    if e.shape[0] != int(off[-1]):
        raise BCCError("edges must have env_off[-1] rows")
    out = np.zeros(e.shape[0], np.uint8)
    _check(lib().bcc_check_ragged(robot_id(robot), scene.handle, _ptr(off), _ptr(e), int(resolution), _ptr(out)))
    return out


def check_ragged_noalloc(robot, scene: Scene, d_env_off: DeviceBuffer, d_edges: DeviceBuffer, num_pairs: int,
                         d_out: DeviceBuffer, resolution: int = 32, stream=None):
    """Allocation-free ragged check on device buffers (asynchronous): the
    tile list is cut on the device into the library's static scratch array.
    d_env_off: int64 [num_envs + 1] CSR offsets on the device."""
    n = int(num_pairs)
    dim = lib().bcc_robot_dimension(robot_id(robot))
    if d_env_off.nbytes < 8 * (scene.num_envs + 1):
        raise BCCError(f"d_env_off holds {d_env_off.nbytes} bytes, {scene.num_envs} envs need {8 * (scene.num_envs + 1)}")
    if d_edges.nbytes < n * 2 * dim * 4 or d_out.nbytes < n:
        raise BCCError(f"d_edges/d_out are too small for {n} pairs")
    _check(lib().bcc_check_ragged_noalloc(robot_id(robot), scene.handle, d_env_off.ptr, d_edges.ptr, n,
                                          int(resolution), d_out.ptr, stream))


class RaggedPlan:
    """Tile list + captured hipGraph for repeated ragged launches on device buffers.
    d_order (optional): a learned visiting order per pair (uint32 per pair,
    e.g. from Model.order_pairs), read at every launch.  learned=True: the
    fused learned kernel (probe, inference of the undecided pairs with
    `model`, probe, plain steps) over a snapshot of `model` taken now
    (model None: fixed second-probe bins)."""

    def __init__(self, robot, scene: Scene, env_off, d_edges: DeviceBuffer, d_out: DeviceBuffer,
                 resolution: int = 32, tile_pairs: int = 0, groups_per_cu: int = 0,
                 d_order: DeviceBuffer | None = None, learned: bool = False, model: "Model | None" = None,
                 tile_steps: int | None = None):
        self.off = np.ascontiguousarray(env_off, np.int64)
        self.robot = robot
        self.scene, self.d_edges, self.d_out, self.d_order = scene, d_edges, d_out, d_order
        self.handle = ctypes.c_void_p()
        # the library reads env_off[num_envs] on the host and the kernels
        # index the device buffers by pair: check both before creating it
        if self.off.ndim != 1 or self.off.size != scene.num_envs + 1:
            raise BCCError("env_off must have num_envs + 1 entries")
        if self.off[0] != 0 or np.any(np.diff(self.off) < 0):
            raise BCCError("env_off must start at 0 and be non-decreasing")
        total = int(self.off[-1])
        dim = lib().bcc_robot_dimension(robot_id(robot))
        if d_edges.nbytes < total * 2 * dim * 4:
            raise BCCError(f"d_edges holds {d_edges.nbytes} bytes, {total} pairs need {total * 2 * dim * 4}")
        if d_out.nbytes < total:
            raise BCCError(f"d_out holds {d_out.nbytes} bytes, {total} pairs need {total}")
        if d_order is not None and d_order.nbytes < 4 * total:
            raise BCCError(f"d_order holds {d_order.nbytes} bytes, {total} pairs need {4 * total}")
        if learned:
            if d_order is not None:
                raise BCCError("a learned plan infers its own order: d_order must be None")
            _check(lib().bcc_ragged_plan_create_learned(robot_id(robot), model.handle if model else None,
                                                        scene.handle, _ptr(self.off), d_edges.ptr, int(resolution),
                                                        d_out.ptr, int(tile_pairs), int(groups_per_cu),
                                                        ctypes.byref(self.handle)))
        elif tile_steps is not None:
            if d_order is not None:
                raise BCCError("a plan with a step dimension takes no order")
            _check(lib().bcc_ragged_plan_create_shaped(robot_id(robot), scene.handle, _ptr(self.off), d_edges.ptr,
                                                       int(resolution), d_out.ptr, int(tile_pairs), int(tile_steps),
                                                       int(groups_per_cu), ctypes.byref(self.handle)))
        elif d_order is None:
            _check(lib().bcc_ragged_plan_create_tuned(robot_id(robot), scene.handle, _ptr(self.off), d_edges.ptr,
                                                      int(resolution), d_out.ptr, int(tile_pairs),
                                                      int(groups_per_cu), ctypes.byref(self.handle)))
        else:
            _check(lib().bcc_ragged_plan_create_ordered(robot_id(robot), scene.handle, _ptr(self.off), d_edges.ptr,
                                                        int(resolution), d_order.ptr, d_out.ptr, int(tile_pairs),
                                                        int(groups_per_cu), ctypes.byref(self.handle)))

    @property
    def num_tiles(self) -> int:
        return lib().bcc_ragged_plan_num_tiles(self.handle)

    @property
    def tile_steps(self) -> int:
        """Steps per tile chunk of a plain plan (0: no step dimension)."""
        return lib().bcc_ragged_plan_tile_steps(self.handle)

    @property
    def config(self) -> tuple:
        """(tile_pairs, groups_per_cu) the plan was built with."""
        t, g = ctypes.c_int(0), ctypes.c_int(0)
        _check(lib().bcc_ragged_plan_config(self.handle, ctypes.byref(t), ctypes.byref(g)))
        return t.value, g.value

    def launch(self, stream=None):
        _check(lib().bcc_ragged_plan_launch(self.handle, stream))

    def launch_eager(self, stream=None):
        _check(lib().bcc_ragged_plan_launch_eager(self.handle, stream))

    def time(self, warmup: int, iters: int, stream=None) -> float:
        ms = ctypes.c_double(0.0)
        _check(lib().bcc_ragged_plan_time(self.handle, int(warmup), int(iters), stream, ctypes.byref(ms)))
        return ms.value

    def close(self):
        if self.handle:
            _check(lib().bcc_ragged_plan_destroy(self.handle))
            self.handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class RaggedGroup:
    """Several ragged plans (e.g. one per robot) replayed as one hipGraph with
    the plans as independent branches.  Keeps the plans alive."""

    def __init__(self, plans: list):
        self.plans = list(plans)
        arr = (ctypes.c_void_p * len(self.plans))(*[p.handle.value for p in self.plans])
        self.handle = ctypes.c_void_p()
        _check(lib().bcc_ragged_group_create(arr, len(self.plans), ctypes.byref(self.handle)))

    def launch(self, stream=None):
        _check(lib().bcc_ragged_group_launch(self.handle, stream))

    def time(self, warmup: int, iters: int, stream=None) -> float:
        ms = ctypes.c_double(0.0)
        _check(lib().bcc_ragged_group_time(self.handle, int(warmup), int(iters), stream, ctypes.byref(ms)))
        return ms.value

    def close(self):
        if self.handle:
            _check(lib().bcc_ragged_group_destroy(self.handle))
            self.handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------------ test hooks

def debug_force_variant(variant: int):
    """-1: automatic kernel variant per launch; 0 / 1: force low / high occupancy."""
    _check(lib().bcc_debug_force_variant(int(variant)))


def debug_dense_chunk(edges: int):
    """Edges per workgroup of dense launches; 0 restores the automatic choice."""
    _check(lib().bcc_debug_dense_chunk(int(edges)))


def debug_cull_min(work: int):
    """Culling threshold (spheres per sweep x primitives per environment) for
    scenes created afterwards; -1 restores the tuned value."""
    _check(lib().bcc_debug_cull_min(int(work)))


def debug_fk(robot, q: np.ndarray) -> np.ndarray:
    q = np.ascontiguousarray(q, np.float32)
    n = q.shape[0]
    ns = NUM_SPHERES[robot] if isinstance(robot, str) else list(NUM_SPHERES.values())[robot]
    dq = DeviceBuffer.from_array(q)
    do = DeviceBuffer(n * ns * 4 * 4)
    _check(lib().bcc_debug_fk(robot_id(robot), dq.ptr, n, do.ptr))
    return do.to_array(np.float32, (n, ns, 4))


def debug_sincos(x: np.ndarray) -> np.ndarray:
    x = np.ascontiguousarray(x, np.float32)
    dx = DeviceBuffer.from_array(x)
    do = DeviceBuffer(x.size * 8)
    _check(lib().bcc_debug_sincos(dx.ptr, x.size, do.ptr))
    return do.to_array(np.float32, (x.size, 2))


def debug_samples(robot, edges: np.ndarray, resolution: int = 32):
    e = _edges_array(robot, edges)
    m, dim = e.shape[0], e.shape[2]
    de = DeviceBuffer.from_array(e)
    dn = DeviceBuffer(m * 4)
    dc = DeviceBuffer(m * 32 * dim * 4)
    dl = DeviceBuffer(m * 32 * dim * 4)
    _check(lib().bcc_debug_samples(robot_id(robot), de.ptr, m, int(resolution), dn.ptr, dc.ptr, dl.ptr))
    return (dn.to_array(np.int32, (m,)), dc.to_array(np.float32, (m, 32, dim)),
            dl.to_array(np.float32, (m, 32, dim)))


def debug_prims(inp: np.ndarray) -> np.ndarray:
    inp = np.ascontiguousarray(inp, np.float32)
    assert inp.ndim == 2 and inp.shape[1] == 40
    di = DeviceBuffer.from_array(inp)
    do = DeviceBuffer(inp.shape[0] * 6 * 4)
    _check(lib().bcc_debug_prims(di.ptr, inp.shape[0], do.ptr))
    return do.to_array(np.float32, (inp.shape[0], 6))


def debug_half_boxes(inp: np.ndarray, mask: int) -> np.ndarray:
    """[64, 4] (x y z rp per lane) -> [2, 2, 3]: (half, lo/hi, xyz)."""
    inp = np.ascontiguousarray(inp, np.float32)
    assert inp.shape == (64, 4)
    di = DeviceBuffer.from_array(inp)
    do = DeviceBuffer(12 * 4)
    _check(lib().bcc_debug_half_boxes(di.ptr, ctypes.c_uint64(mask), do.ptr))
    return do.to_array(np.float32, (2, 2, 3))


def debug_copy(d_src: DeviceBuffer, d_dst: DeviceBuffer, nbytes: int):
    _check(lib().bcc_debug_copy(d_src.ptr, d_dst.ptr, int(nbytes)))


def debug_valu(blocks: int, iters: int, active_lanes: int, d_out: DeviceBuffer) -> float:
    """Known FP32 work: per wave, lanes below active_lanes run 8 x iters FMAs.
    Returns the kernel's hipEvent time in ms."""
    if d_out.nbytes < 4 * 256 * int(blocks):
        raise BCCError("d_out must hold blocks * 256 floats")
    ms = ctypes.c_double(0.0)
    _check(lib().bcc_debug_valu(int(blocks), int(iters), int(active_lanes), d_out.ptr, ctypes.byref(ms)))
    return ms.value


# ------------------------------------------------------------ learned ordering

def check_dense_ex(robot, scene: Scene, d_edges: DeviceBuffer, num_edges: int, resolution: int,
                   d_out: DeviceBuffer, d_order: DeviceBuffer | None = None, d_rounds: DeviceBuffer | None = None,
                   stream=None):
    _check(lib().bcc_check_dense_ex(robot_id(robot), scene.handle, d_edges.ptr, int(num_edges), int(resolution),
                                    d_order.ptr if d_order else None, d_rounds.ptr if d_rounds else None,
                                    d_out.ptr, stream))


def time_dense_ex(robot, scene: Scene, d_edges: DeviceBuffer, num_edges: int, resolution: int, d_out: DeviceBuffer,
                  d_order: DeviceBuffer | None, warmup: int, iters: int, stream=None) -> float:
    ms = ctypes.c_double(0.0)
    _check(lib().bcc_time_dense_ex(robot_id(robot), scene.handle, d_edges.ptr, int(num_edges), int(resolution),
                                   d_order.ptr if d_order else None, d_out.ptr, int(warmup), int(iters), stream,
                                   ctypes.byref(ms)))
    return ms.value


def label_pairs(robot, scene: Scene, d_edges: DeviceBuffer, pair_env, pair_edge, resolution: int = 32,
                steps: int = 0):
    """Bin masks (uint32 per pair) from the exact checker: bit t = one of the
    first `steps` samples of bin t collides (0: any of its samples)."""
    pe = np.ascontiguousarray(pair_env, np.int32)
    pg = np.ascontiguousarray(pair_edge, np.int32)
    n = pe.size
    dpe, dpg = DeviceBuffer.from_array(pe), DeviceBuffer.from_array(pg)
    dl = DeviceBuffer(n * 4)
    _check(lib().bcc_label_pairs_steps(robot_id(robot), scene.handle, d_edges.ptr, n, dpe.ptr, dpg.ptr,
                                       int(resolution), int(steps), dl.ptr, None))
    synchronize()
    return dl


def encode_envs(scene: Scene) -> np.ndarray:
    d = DeviceBuffer(scene.num_envs * 64 * 4)
    _check(lib().bcc_encode_envs(scene.handle, d.ptr))
    return d.to_array(np.float32, (scene.num_envs, 64))


class Model:
    """The learned step-ordering MLP (bcc_model)."""
    IN, OUT = 96, 32

    @staticmethod
    def hidden() -> int:
        """Hidden width H of the compiled model (96 -> H -> H -> 32), from its
        parameter count H^2 + (IN + OUT + 2) H + OUT."""
        n, b = Model.num_params(), Model.IN + Model.OUT + 2
        return int(round((-b + (b * b + 4 * (n - Model.OUT)) ** 0.5) / 2))

    def __init__(self, robot=None, seed: int = 0, handle=None):
        self.handle = ctypes.c_void_p()
        if handle is not None:
            self.handle = handle
        else:
            _check(lib().bcc_model_create(robot_id(robot), int(seed), ctypes.byref(self.handle)))

    @staticmethod
    def num_params() -> int:
        return lib().bcc_model_num_params()

    @staticmethod
    def load(path: str) -> "Model":
        h = ctypes.c_void_p()
        _check(lib().bcc_model_load(path.encode(), ctypes.byref(h)))
        return Model(handle=h)

    def save(self, path: str):
        _check(lib().bcc_model_save(self.handle, path.encode()))

    def params(self) -> np.ndarray:
        out = np.zeros(self.num_params(), np.float32)
        _check(lib().bcc_model_get_params(self.handle, _ptr(out)))
        return out

    def set_params(self, p: np.ndarray):
        p = np.ascontiguousarray(p, np.float32)
        _check(lib().bcc_model_set_params(self.handle, _ptr(p)))

    def features(self, scene: Scene, d_edges: DeviceBuffer, pair_env, pair_edge) -> DeviceBuffer:
        pe = np.ascontiguousarray(pair_env, np.int32)
        pg = np.ascontiguousarray(pair_edge, np.int32)
        dpe, dpg = DeviceBuffer.from_array(pe), DeviceBuffer.from_array(pg)
        dx = DeviceBuffer(pe.size * self.IN * 2)
        _check(lib().bcc_model_features(self.handle, scene.handle, d_edges.ptr, pe.size, dpe.ptr, dpg.ptr, dx.ptr,
                                        None))
        return dx

    def grads(self, d_x: DeviceBuffer, d_labels: DeviceBuffer, rows: int):
        dg = DeviceBuffer(self.num_params() * 4)
        loss = ctypes.c_float(0.0)
        _check(lib().bcc_model_grads(self.handle, d_x.ptr, d_labels.ptr, int(rows), dg.ptr, ctypes.byref(loss)))
        return dg.to_array(np.float32, (self.num_params(),)), loss.value

    def train(self, d_x: DeviceBuffer, d_labels: DeviceBuffer, rows: int, epochs: int, batch: int = 4096,
              lr: float = 1e-3) -> np.ndarray:
        losses = np.zeros(epochs, np.float32)
        _check(lib().bcc_model_train(self.handle, d_x.ptr, d_labels.ptr, int(rows), int(epochs), int(batch),
                                     float(lr), _ptr(losses)))
        return losses

    def logits(self, d_x: DeviceBuffer, rows: int) -> np.ndarray:
        dl = DeviceBuffer(rows * self.OUT * 4)
        _check(lib().bcc_model_logits(self.handle, d_x.ptr, int(rows), dl.ptr))
        return dl.to_array(np.float32, (rows, self.OUT))

    def order_dense(self, scene: Scene, d_edges: DeviceBuffer, num_edges: int) -> DeviceBuffer:
        d = DeviceBuffer(int(num_edges) * scene.num_envs * 4)
        _check(lib().bcc_model_order_dense(self.handle, scene.handle, d_edges.ptr, int(num_edges), d.ptr, None))
        synchronize()
        return d

    def time_learned_dense(self, scene: Scene, d_edges: DeviceBuffer, num_edges: int, d_order: DeviceBuffer,
                           d_out: DeviceBuffer, warmup: int, iters: int, resolution: int = 32) -> tuple:
        """(step ms, inference ms): hipEvent-timed learned steps over a dense
        batch (inference for every pair into d_order, then the ordered check
        into d_out), and the inference alone."""
        step, order = ctypes.c_double(0.0), ctypes.c_double(0.0)
        _check(lib().bcc_time_learned_dense(self.handle, scene.handle, d_edges.ptr, int(num_edges), int(resolution),
                                            d_order.ptr, d_out.ptr, int(warmup), int(iters), None,
                                            ctypes.byref(step), ctypes.byref(order)))
        return step.value, order.value

    def logits_dense(self, scene: Scene, d_edges: DeviceBuffer, num_edges: int):
        """(order uint32 [M, E], logits float32 [M, E, 32]) from the fused kernel."""
        M, E = int(num_edges), scene.num_envs
        d = DeviceBuffer(M * E * 4)
        dl = DeviceBuffer(M * E * self.OUT * 4)
        _check(lib().bcc_model_logits_dense(self.handle, scene.handle, d_edges.ptr, M, d.ptr, dl.ptr, None))
        synchronize()
        return d.to_array(np.uint32, (M, E)), dl.to_array(np.float32, (M, E, self.OUT))

    def order_pairs(self, scene: Scene, d_edges: DeviceBuffer, pair_env, pair_edge: np.ndarray | None = None,
                    logits: bool = False, out: DeviceBuffer | None = None):
        """Top-4 bins of the pairs (pair_env[p], pair_edge[p] or p) from the
        fused kernel: a device buffer of uint32 per pair (`out` if given), and
        with logits=True also the float32 [pairs, 32] logits (host).
        pair_env: int32 host array, or a DeviceBuffer holding one."""
        if isinstance(pair_env, DeviceBuffer):
            d_env, n = pair_env, pair_env.nbytes // 4
        else:
            pe = np.ascontiguousarray(pair_env, np.int32)
            d_env, n = DeviceBuffer.from_array(pe), int(pe.shape[0])
        d_edge = DeviceBuffer.from_array(np.ascontiguousarray(pair_edge, np.int32)) if pair_edge is not None else None
        d = out if out is not None else DeviceBuffer(max(n, 1) * 4)
        if d.nbytes < 4 * n:
            raise ValueError("order buffer too small")
        dl = DeviceBuffer(max(n, 1) * self.OUT * 4) if logits else None
        _check(lib().bcc_model_order_pairs(self.handle, scene.handle, d_edges.ptr, n, d_env.ptr,
                                           d_edge.ptr if d_edge else None, d.ptr, dl.ptr if dl else None, None))
        synchronize()
        if logits:
            return d, dl.to_array(np.float32, (n, self.OUT))
        return d

    def close(self):
        if self.handle:
            _check(lib().bcc_model_destroy(self.handle))
            self.handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def check_dense_learned(robot, model: "Model | None", scene: Scene, d_edges: DeviceBuffer, num_edges: int,
                        resolution: int = 32, d_out: DeviceBuffer | None = None, stage: bool = False,
                        words: bool = False, stats: bool = False) -> dict:
    """One fused learned step over a dense batch (bcc_check_dense_learned).
    Returns {"verdicts": bool [M, E]} plus, on request, "stage" (uint8 [M,
    E]: p = decided by probe p + 1, else passes - 1 + plain steps; passes =
    2 unless set with debug_learned_passes),
    "words" (uint32 [M, E], 0xffffffff where nothing was inferred) and
    "stats" (pairs decided by probes 1, 2 and 3, inferred, sent to the step
    loop)."""
    M, E = int(num_edges), scene.num_envs
    n = M * E
    d = d_out if d_out is not None else DeviceBuffer(max(n, 1))
    if d.nbytes < n:
        raise BCCError("result buffer too small")
    d_stage = DeviceBuffer(max(n, 1)) if stage else None
    d_words = DeviceBuffer.from_array(np.full(max(n, 1), 0xffffffff, np.uint32)) if words else None
    d_stats = DeviceBuffer.from_array(np.zeros(5, np.int32)) if stats else None
    _check(lib().bcc_check_dense_learned(robot_id(robot), model.handle if model else None, scene.handle,
                                         d_edges.ptr, M, int(resolution), d.ptr, d_stage.ptr if d_stage else None,
                                         d_words.ptr if d_words else None, d_stats.ptr if d_stats else None, None))
    synchronize()
    out = {"verdicts": d.to_array(np.uint8, (M, E)).astype(bool)}
    if stage:
        out["stage"] = d_stage.to_array(np.uint8, (M, E))
    if words:
        out["words"] = d_words.to_array(np.uint32, (max(n, 1),))[:n].reshape(M, E)
    if stats:
        out["stats"] = d_stats.to_array(np.int32, (5,))
    return out


def time_dense_learned(robot, model: "Model | None", scene: Scene, d_edges: DeviceBuffer, num_edges: int,
                       d_out: DeviceBuffer, warmup: int, iters: int, resolution: int = 32) -> float:
    """Average ms of a fused learned step (hipEvents on the launch stream)."""
    ms = ctypes.c_double(0.0)
    _check(lib().bcc_time_dense_learned(robot_id(robot), model.handle if model else None, scene.handle, d_edges.ptr,
                                        int(num_edges), int(resolution), d_out.ptr, int(warmup), int(iters), None,
                                        ctypes.byref(ms)))
    return ms.value


def debug_learned_tile(pairs: int):
    _check(lib().bcc_debug_learned_tile(int(pairs)))


def debug_learned_far(radius: int):
    _check(lib().bcc_debug_learned_far(int(radius)))


def debug_learned_passes(passes: int):
    _check(lib().bcc_debug_learned_passes(int(passes)))


def debug_learned_min_steps(steps: int):
    _check(lib().bcc_debug_learned_min_steps(int(steps)))


def debug_learned_small(max_pairs: int):
    """Pair limit of the small-batch learned path (negative: the default)."""
    _check(lib().bcc_debug_learned_small(int(max_pairs)))


def learned_small_pairs() -> int:
    """Dense learned launches of at most this many pairs (no stage, words or
    stats asked for) infer every pair (no model: the invalid word) and run the
    ordered dense kernel."""
    return int(lib().bcc_learned_small_pairs())


def debug_learned_locality(mode: int):
    """0 never, 1 panda launches of 2^22 pairs or more (the default), 2 always."""
    _check(lib().bcc_debug_learned_locality(int(mode)))


def debug_gemm(ta: bool, tb: bool, a: np.ndarray, b: np.ndarray, m: int, n: int, k: int, alpha=1.0, beta=0.0,
               c: np.ndarray | None = None) -> np.ndarray:
    """a, b: uint16 BF16 bit patterns (row-major, shapes as stored)."""
    da, db = DeviceBuffer.from_array(np.ascontiguousarray(a, np.uint16)), DeviceBuffer.from_array(
        np.ascontiguousarray(b, np.uint16))
    c0 = np.zeros((m, n), np.float32) if c is None else np.ascontiguousarray(c, np.float32)
    dc = DeviceBuffer.from_array(c0)
    lda = a.shape[1]
    ldb = b.shape[1]
    _check(lib().bcc_debug_gemm(int(ta), int(tb), da.ptr, lda, db.ptr, ldb, dc.ptr, n, m, n, k, float(alpha),
                                float(beta)))
    return dc.to_array(np.float32, (m, n))
