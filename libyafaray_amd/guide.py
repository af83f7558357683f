"""Python handle on the path-guiding network (C ABI in include/ymx_guide.h).

Samples and queries are float32 arrays of shape (n, 16): position (3), target
weight c, normal (3), pad, wo (3), pad, sampled direction (3), pad -- the
layout of ymx_guide_sample.  Device arrays are torch tensors.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import native
from .native import check

LEVELS = 16
FEATURES = 2
MLP_WEIGHTS = 64 * 64 * 3 + 48 * 64
MIX_FLOATS = 48
SAMPLE_FLOATS = 16


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise native.YmxError("the guiding network needs a HIP device")
    return torch


def pack_samples(pos, normal, wo, w=None, c=None) -> np.ndarray:
    n = len(pos)
    s = np.zeros((n, SAMPLE_FLOATS), np.float32)
    s[:, 0:3] = pos
    s[:, 3] = 0.0 if c is None else c
    s[:, 4:7] = normal
    s[:, 8:11] = wo
    if w is not None:
        s[:, 12:15] = w
    return s


class GpuGuide:
    def __init__(self, bmin, bmax, log2_table_size=19, n_min=16, n_max=2048, max_batch=1 << 18, seed=0, lr=1e-2,
                 beta1=0.9, beta2=0.99, eps=1e-8):
        self._lib = native.lib()
        self.torch = _torch()
        p = native.YmxGuideParams()
        p.struct_size = C.sizeof(native.YmxGuideParams)
        p.log2_table_size = log2_table_size
        p.n_min, p.n_max = n_min, n_max
        p.bmin = (C.c_float * 3)(*[float(x) for x in bmin])
        p.bmax = (C.c_float * 3)(*[float(x) for x in bmax])
        p.max_batch = max_batch
        p.seed = seed
        p.lr, p.beta1, p.beta2, p.eps = lr, beta1, beta2, eps
        h = C.c_void_p()
        check(self._lib.ymx_guide_create(C.byref(p), C.byref(h)), "ymx_guide_create")
        self._h = h
        self.table = 1 << log2_table_size
        self.max_batch = max_batch

    @property
    def handle(self):
        return self._h

    def close(self):
        if self._h:
            self._lib.ymx_guide_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _stream(self):
        return self.torch.cuda.current_stream().cuda_stream

    def set_weights(self, grid: np.ndarray, mlp: np.ndarray):
        grid = np.ascontiguousarray(grid, np.float32).reshape(-1)
        mlp = np.ascontiguousarray(mlp, np.float32).reshape(-1)
        assert grid.size == LEVELS * self.table * FEATURES and mlp.size == MLP_WEIGHTS
        check(self._lib.ymx_guide_set_weights(self._h, grid.ctypes.data, mlp.ctypes.data), "ymx_guide_set_weights")

    def get_weights(self):
        grid = np.zeros(LEVELS * self.table * FEATURES, np.float32)
        mlp = np.zeros(MLP_WEIGHTS, np.float32)
        check(self._lib.ymx_guide_get_weights(self._h, grid.ctypes.data, mlp.ctypes.data), "ymx_guide_get_weights")
        return grid.reshape(LEVELS, self.table, FEATURES), mlp

    def _dev(self, samples):
        t = self.torch
        if isinstance(samples, np.ndarray):
            samples = t.from_numpy(np.ascontiguousarray(samples, np.float32)).cuda()
        assert samples.dtype == t.float32 and samples.shape[1] == SAMPLE_FLOATS and samples.is_contiguous()
        return samples

    def infer(self, queries, out=None):
        t = self.torch
        q = self._dev(queries)
        n = q.shape[0]
        if out is None:
            out = t.empty((n, MIX_FLOATS), dtype=t.float32, device=q.device)
        check(self._lib.ymx_guide_infer(self._h, q.data_ptr(), out.data_ptr(), n, None, self._stream()),
              "ymx_guide_infer")
        return out

    def backward(self, samples):
        s = self._dev(samples)
        check(self._lib.ymx_guide_backward(self._h, s.data_ptr(), s.shape[0], None, self._stream()),
              "ymx_guide_backward")

    def gradients(self):
        grid = np.zeros(LEVELS * self.table * FEATURES, np.float32)
        mlp = np.zeros(MLP_WEIGHTS, np.float32)
        loss = C.c_double()
        check(self._lib.ymx_guide_get_gradients(self._h, grid.ctypes.data, mlp.ctypes.data, C.byref(loss)),
              "ymx_guide_get_gradients")
        return grid.reshape(LEVELS, self.table, FEATURES), mlp, loss.value

    def adam_step(self):
        check(self._lib.ymx_guide_adam_step(self._h, self._stream()), "ymx_guide_adam_step")

    def train_step(self, samples):
        s = self._dev(samples)
        check(self._lib.ymx_guide_train_step(self._h, s.data_ptr(), s.shape[0], None, self._stream()),
              "ymx_guide_train_step")


def mlp_to_flat(W) -> np.ndarray:
    """[W1, W2, W3 (64 x 64), W4 (48 x 64)] -> the ABI's flat layout."""
    return np.concatenate([np.asarray(w, np.float32).reshape(-1) for w in W])


def flat_to_mlp(flat):
    f = np.asarray(flat, np.float32)
    return [f[0:4096].reshape(64, 64), f[4096:8192].reshape(64, 64), f[8192:12288].reshape(64, 64),
            f[12288:].reshape(48, 64)]
