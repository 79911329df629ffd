"""ctypes wrapper of the CPU oracle (oracle/liboracle.so) -- test-side only."""
import ctypes
import os

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PATH = os.path.join(REPO, "oracle", "liboracle.so")

P = ctypes.c_void_p
I = ctypes.c_int
F = ctypes.c_float
D = ctypes.c_double


def _p(a):
    return ctypes.c_void_p(a.ctypes.data)


class Oracle:
    def __init__(self):
        if not os.path.exists(PATH):
            raise RuntimeError(f"{PATH} missing: run `make -C oracle`")
        L = ctypes.CDLL(PATH)
        sig = lambda n, r, a: (setattr(getattr(L, n), "restype", r), setattr(getattr(L, n), "argtypes", a))
        sig("orc_sincosf", None, [F, ctypes.POINTER(F), ctypes.POINTER(F)])
        sig("orc_steps", I, [P, P, I, I, P])
        sig("orc_sample", None, [P, P, I, I, I, I, P])
        sig("orc_fk_spheres", I, [I, P, P])
        sig("orc_fk_spheres64", I, [I, P, P])
        sig("orc_fk_frames64", I, [I, P, P, P])
        sig("orc_robot_dim", I, [I])
        sig("orc_robot_num_spheres", I, [I])
        sig("orc_pad", F, [F])
        sig("orc_sphere_val", F, [P, F, F, F, F])
        sig("orc_capsule_val", F, [P, F, F, F, F])
        sig("orc_zcapsule_val", F, [P, F, F, F, F])
        sig("orc_cuboid_val", F, [P, F, F, F, F])
        sig("orc_zcuboid_val", F, [P, F, F, F, F])
        sig("orc_self_val", F, [P, P])
        sig("orc_sphere_dist64", D, [P, D, D, D, D])
        sig("orc_capsule_dist64", D, [P, D, D, D, D, I])
        sig("orc_cuboid_dist64", D, [P, D, D, D, D, I])
        sig("orc_config_collides", I, [I, P, P, P, I, P])
        sig("orc_check_dense", I, [I, P, P, P, I, P, I, I, P, I])
        sig("orc_check_dense64", I, [I, P, P, P, I, P, I, I, P, P, I])
        sig("orc_work_dense", I, [I, P, P, P, I, P, I, I, P, P, P, I])
        sig("orc_check_pairs", I, [I, P, P, P, I, P, I, I, P, P, ctypes.c_int64, I, P, P, P, P, I])
        self.L = L

    # -- scalar helpers
    def sincos(self, x):
        s, c = F(), F()
        self.L.orc_sincosf(F(x), ctypes.byref(s), ctypes.byref(c))
        return np.float32(s.value), np.float32(c.value)

    def steps(self, s, e, resolution=32):
        s = np.ascontiguousarray(s, np.float32)
        e = np.ascontiguousarray(e, np.float32)
        d = np.zeros(s.size, np.float32)
        n = self.L.orc_steps(_p(s), _p(e), s.size, resolution, _p(d))
        return n, d

    def sample(self, s, delta, n, t, i):
        s = np.ascontiguousarray(s, np.float32)
        delta = np.ascontiguousarray(delta, np.float32)
        out = np.zeros(s.size, np.float32)
        self.L.orc_sample(_p(s), _p(delta), n, t, i, s.size, _p(out))
        return out

    def fk(self, robot, q):
        q = np.ascontiguousarray(q, np.float32)
        ns = self.L.orc_robot_num_spheres(robot)
        out = np.zeros((ns, 4), np.float32)
        assert self.L.orc_fk_spheres(robot, _p(q), _p(out)) == 0
        return out

    def fk64(self, robot, q):
        q = np.ascontiguousarray(q, np.float64)
        ns = self.L.orc_robot_num_spheres(robot)
        out = np.zeros((ns, 4), np.float64)
        assert self.L.orc_fk_spheres64(robot, _p(q), _p(out)) == 0
        return out

    def frames64(self, robot, q):
        q = np.ascontiguousarray(q, np.float64)
        dim = self.L.orc_robot_dim(robot)
        R = np.zeros((dim, 3, 3))
        t = np.zeros((dim, 3))
        assert self.L.orc_fk_frames64(robot, _p(q), _p(R), _p(t)) == 0
        return R, t

    # -- scenes
    @staticmethod
    def _scene(sc):
        return (np.ascontiguousarray(sc.data, np.float32), np.ascontiguousarray(sc.env_off, np.int64),
                np.ascontiguousarray(sc.counts, np.int32))

    def check_dense(self, robot, sc, edges, resolution=32, threads=8):
        d, o, c = self._scene(sc)
        e = np.ascontiguousarray(edges, np.float32)
        out = np.zeros((e.shape[0], sc.num_envs), np.uint8)
        assert self.L.orc_check_dense(robot, _p(d), _p(o), _p(c), sc.num_envs, _p(e), e.shape[0], resolution,
                                      _p(out), threads) == 0
        return out

    def check_dense64(self, robot, sc, edges, resolution=32, threads=8):
        d, o, c = self._scene(sc)
        e = np.ascontiguousarray(edges, np.float32)
        out = np.zeros((e.shape[0], sc.num_envs), np.uint8)
        mg = np.zeros((e.shape[0], sc.num_envs), np.float64)
        assert self.L.orc_check_dense64(robot, _p(d), _p(o), _p(c), sc.num_envs, _p(e), e.shape[0], resolution,
                                        _p(out), _p(mg), threads) == 0
        return out, mg

    def work_dense(self, robot, sc, edges, resolution=32, threads=8):
        d, o, c = self._scene(sc)
        e = np.ascontiguousarray(edges, np.float32)
        M, E = e.shape[0], sc.num_envs
        out = np.zeros(M * E, np.uint8)
        w = np.zeros(M * E, np.uint64)
        t = np.zeros(M * E, np.uint64)
        assert self.L.orc_work_dense(robot, _p(d), _p(o), _p(c), E, _p(e), M, resolution, _p(out), _p(w), _p(t),
                                     threads) == 0
        return out.reshape(M, E), w.reshape(M, E), t.reshape(M, E)

    def check_pairs(self, robot, sc, edges, pair_env, pair_edge, resolution=32, mode=0, threads=8):
        d, o, c = self._scene(sc)
        e = np.ascontiguousarray(edges, np.float32)
        pe = np.ascontiguousarray(pair_env, np.int32)
        pg = np.ascontiguousarray(pair_edge, np.int32)
        n = pe.size
        out = np.zeros(n, np.uint8)
        mg = np.zeros(n, np.float64)
        w = np.zeros(n, np.uint64)
        t = np.zeros(n, np.uint64)
        assert self.L.orc_check_pairs(robot, _p(d), _p(o), _p(c), sc.num_envs, _p(e), e.shape[0], resolution,
                                      _p(pe), _p(pg), n, mode, _p(out), _p(mg), _p(w), _p(t), threads) == 0
        return out, mg, w, t

    def config_collides(self, robot, sc, env, q):
        d, o, c = self._scene(sc)
        q = np.ascontiguousarray(q, np.float32)
        return self.L.orc_config_collides(robot, _p(d), _p(o), _p(c), env, _p(q))
