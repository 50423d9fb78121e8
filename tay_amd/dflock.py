"""Differentiable flocking on the GPU (include/tay_dflock.h): a thin ctypes
wrapper over the forward/backward pair.  Arrays are (n, 3) float32."""
import ctypes as C

import numpy as np

from . import api


def _f32(a):
    return np.ascontiguousarray(a, np.float32)


class DiffFlock:
    """n agents of the smooth flocking model; the backward pass
    differentiates the most recent forward pass."""

    def __init__(self, lib, n, radius=20.0, h2=1.0, dt=0.1, damping=0.05, cell_size=None, checkpoint_every=10,
                 table_mb=64):
        self.lib = lib
        self.n = n
        self.params = api.TayDflockParams(radius, h2, dt, damping, radius if cell_size is None else cell_size)
        self.flock = lib.tay_dflock_create(n, C.byref(self.params), checkpoint_every, int(table_mb) << 20)
        if not self.flock:
            raise RuntimeError("tay_dflock_create: " + lib.tay_gpu_last_error().decode())

    def _ok(self, r):
        if not r:
            raise RuntimeError(self.lib.tay_gpu_last_error().decode())

    def forward(self, p0, v0, weights, steps):
        p0, v0, w = _f32(p0), _f32(v0), _f32(weights)
        assert p0.shape == (self.n, 3) and v0.shape == (self.n, 3) and w.shape == (3,)
        p, v = np.zeros((self.n, 3), np.float32), np.zeros((self.n, 3), np.float32)
        self._ok(self.lib.tay_dflock_forward(self.flock, p0.ctypes.data, v0.ctypes.data, w.ctypes.data, steps,
                                             p.ctypes.data, v.ctypes.data))
        return p, v

    def backward(self, grad_p, grad_v=None):
        gp = _f32(grad_p)
        gv = _f32(grad_v) if grad_v is not None else None
        gp0, gv0, gw = (np.zeros((self.n, 3), np.float32), np.zeros((self.n, 3), np.float32),
                        np.zeros(3, np.float32))
        self._ok(self.lib.tay_dflock_backward(self.flock, gp.ctypes.data, gv.ctypes.data if gv is not None else None,
                                              gp0.ctypes.data, gv0.ctypes.data, gw.ctypes.data))
        return gp0, gv0, gw

    def times(self):
        f, b = C.c_double(), C.c_double()
        self._ok(self.lib.tay_dflock_times(self.flock, C.byref(f), C.byref(b)))
        return f.value, b.value

    def max_neighbours(self):
        m = C.c_uint()
        self._ok(self.lib.tay_dflock_max_neighbours(self.flock, C.byref(m)))
        return m.value

    def pair_count(self):
        """Neighbour pairs summed over the steps of the last forward call."""
        m = C.c_ulonglong()
        self._ok(self.lib.tay_dflock_pair_count(self.flock, C.byref(m)))
        return m.value

    def close(self):
        if self.flock:
            self.lib.tay_dflock_destroy(self.flock)
            self.flock = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
