"""Shared pieces of the differentiable-flocking tests: the FP64 oracle
(oracle/dflock.c) through ctypes, and seeded inputs."""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DFLOCK_LIB = os.path.join(ROOT, "oracle", "_build", "libdflock.so")


class DflockParams(C.Structure):
    _fields_ = [("radius", C.c_double), ("h2", C.c_double), ("dt", C.c_double), ("damping", C.c_double)]


_D = C.POINTER(C.c_double)


def load_oracle():
    lib = C.CDLL(DFLOCK_LIB)
    lib.oracle_dflock_forward.restype = C.c_int
    lib.oracle_dflock_forward.argtypes = [C.c_uint, _D, _D, _D, C.POINTER(DflockParams), C.c_uint, _D, _D]
    lib.oracle_dflock_gradient.restype = C.c_int
    lib.oracle_dflock_gradient.argtypes = [C.c_uint, _D, _D, _D, C.POINTER(DflockParams), C.c_uint, _D, _D, _D,
                                           _D, _D]
    lib.oracle_dflock_backprop.restype = C.c_int
    lib.oracle_dflock_backprop.argtypes = [C.c_uint, _D, _D, _D, C.POINTER(DflockParams), C.c_uint, _D, _D, _D, _D,
                                           _D]
    lib.oracle_dflock_backprop_along.restype = C.c_int
    lib.oracle_dflock_backprop_along.argtypes = [C.c_uint, _D, _D, _D, C.POINTER(DflockParams), C.c_uint, _D, _D,
                                                 _D, _D, _D]
    return lib


def _p(a):
    return a.ctypes.data_as(_D)


def forward64(lib, p0, v0, weights, prm, steps):
    n = len(p0)
    p0, v0 = np.ascontiguousarray(p0, np.float64), np.ascontiguousarray(v0, np.float64)
    w = np.ascontiguousarray(weights, np.float64)
    p, v = np.zeros((n, 3)), np.zeros((n, 3))
    assert lib.oracle_dflock_forward(n, _p(p0), _p(v0), _p(w), C.byref(prm), steps, _p(p), _p(v))
    return p, v


def gradient64(lib, p0, v0, weights, prm, steps, target):
    n = len(p0)
    p0, v0 = np.ascontiguousarray(p0, np.float64), np.ascontiguousarray(v0, np.float64)
    w = np.ascontiguousarray(weights, np.float64)
    t = np.ascontiguousarray(target, np.float64)
    loss = C.c_double()
    gp, gv, gw = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros(3)
    assert lib.oracle_dflock_gradient(n, _p(p0), _p(v0), _p(w), C.byref(prm), steps, _p(t), C.byref(loss), _p(gp),
                                      _p(gv), _p(gw))
    return loss.value, gp, gv, gw


def backprop64(lib, p0, v0, weights, prm, steps, gp_final, gv_final=None):
    n = len(p0)
    p0, v0 = np.ascontiguousarray(p0, np.float64), np.ascontiguousarray(v0, np.float64)
    w = np.ascontiguousarray(weights, np.float64)
    gpf = np.ascontiguousarray(gp_final, np.float64)
    gvf = None if gv_final is None else np.ascontiguousarray(gv_final, np.float64)
    gp, gv, gw = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros(3)
    assert lib.oracle_dflock_backprop(n, _p(p0), _p(v0), _p(w), C.byref(prm), steps, _p(gpf),
                                      None if gvf is None else _p(gvf), _p(gp), _p(gv), _p(gw))
    return gp, gv, gw


def backprop_along64(lib, states_p, states_v, weights, prm, gp_final, gv_final=None):
    """FP64 adjoint along a given trajectory: states_p/states_v are
    (steps + 1, n, 3)."""
    steps, n = states_p.shape[0] - 1, states_p.shape[1]
    sp = np.ascontiguousarray(states_p, np.float64)
    sv = np.ascontiguousarray(states_v, np.float64)
    w = np.ascontiguousarray(weights, np.float64)
    gpf = np.ascontiguousarray(gp_final, np.float64)
    gvf = None if gv_final is None else np.ascontiguousarray(gv_final, np.float64)
    gp, gv, gw = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros(3)
    assert lib.oracle_dflock_backprop_along(n, _p(sp), _p(sv), _p(w), C.byref(prm), steps, _p(gpf),
                                            None if gvf is None else _p(gvf), _p(gp), _p(gv), _p(gw))
    return gp, gv, gw


def clustered(n, clusters, sigma, box, seed):
    """n agents from `clusters` Gaussian clusters in [0, box)^3 with
    unit-random velocities (float32 values)."""
    g = np.random.Generator(np.random.PCG64(seed))
    centres = g.uniform(3 * sigma, box - 3 * sigma, size=(clusters, 3))
    which = g.integers(0, clusters, size=n)
    p = (centres[which] + g.normal(scale=sigma, size=(n, 3))).astype(np.float32)
    v = g.normal(size=(n, 3))
    v = (v / np.linalg.norm(v, axis=1, keepdims=True)).astype(np.float32)
    return p, v


def mse(p, target):
    return float(np.sum((np.asarray(p, np.float64) - target) ** 2) / p.size)
