"""ctypes binding of the profile surrogate (include/mfa_mlp.h, in
libmfa_hip.so): a small MLP trained and evaluated with hand-written BF16 MFMA
kernels on the GPU.  No CPU fallback: the product path fails loudly when the
library is missing."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib

_vp, _i32, _i64, _f = C.c_void_p, C.c_int32, C.c_int64, C.c_float
SIGNATURES = {
    "mfa_mlp_create": [_vp, _i32, C.c_uint64, _vp],
    "mfa_mlp_destroy": [_vp],
    "mfa_mlp_num_params": [_vp, _vp],
    "mfa_mlp_get_params": [_vp, _vp, _i64],
    "mfa_mlp_set_params": [_vp, _vp, _i64],
    "mfa_mlp_forward": [_vp, _vp, _i32, _vp],
    "mfa_mlp_gradient": [_vp, _vp, _vp, _i32, _vp, _vp],
    "mfa_mlp_train": [_vp, _vp, _vp, _i64, _i32, _i32, _f, C.c_uint64, _vp],
    "mfa_mlp_set_graph": [_vp, _i32],
    "mfa_mlp_set_lr_schedule": [_vp, _f],
    "mfa_mlp_gemm_check": [_vp, _vp, _i32, _i32, _i32, _vp],
    "mfa_mlp_last_error": [],
}


class MlpError(RuntimeError):
    pass


def _load():
    lib = _lib.load()
    if not getattr(lib, "_mlp_bound", False):
        for name, args in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = C.c_int
        lib.mfa_mlp_last_error.restype = C.c_char_p
        lib._mlp_bound = True
    return lib


def _check(rc, what):
    if rc != 0:
        raise MlpError("%s failed (%d): %s" % (what, rc, _load().mfa_mlp_last_error().decode(errors="replace")))


def _f32(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float32)
    if shape is not None:
        a = a.reshape(shape)
    return a


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


class Mlp:
    """x (n_in) -> [Linear -> SiLU]^(L-1) -> Linear -> y (n_out)."""

    def __init__(self, widths, seed: int = 1):
        self.lib = _load()
        self.widths = [int(w) for w in widths]
        arr = (C.c_int32 * len(self.widths))(*self.widths)
        h = C.c_void_p()
        _check(self.lib.mfa_mlp_create(arr, len(self.widths), C.c_uint64(seed), C.byref(h)), "mfa_mlp_create")
        self.h = h
        n = C.c_int64()
        _check(self.lib.mfa_mlp_num_params(self.h, C.byref(n)), "num_params")
        self.n_params = n.value

    def close(self):
        if self.h:
            self.lib.mfa_mlp_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self) -> np.ndarray:
        p = np.empty(self.n_params, np.float32)
        _check(self.lib.mfa_mlp_get_params(self.h, _p(p), self.n_params), "get_params")
        return p

    def set_params(self, p):
        p = _f32(p)
        assert p.size == self.n_params
        _check(self.lib.mfa_mlp_set_params(self.h, _p(p), self.n_params), "set_params")

    def forward(self, x) -> np.ndarray:
        x = _f32(x, (-1, self.widths[0]))
        y = np.empty((x.shape[0], self.widths[-1]), np.float32)
        _check(self.lib.mfa_mlp_forward(self.h, _p(x), x.shape[0], _p(y)), "forward")
        return y

    def gradient(self, x, t):
        x = _f32(x, (-1, self.widths[0]))
        t = _f32(t, (-1, self.widths[-1]))
        g = np.empty(self.n_params, np.float32)
        loss = C.c_float()
        _check(self.lib.mfa_mlp_gradient(self.h, _p(x), _p(t), x.shape[0], C.byref(loss), _p(g)), "gradient")
        return loss.value, g

    def set_graph(self, on: bool = True):
        """Train by replaying one captured step (default) or by launching each
        step's kernels; the same bits either way."""
        _check(self.lib.mfa_mlp_set_graph(self.h, 1 if on else 0), "set_graph")

    def train(self, x, t, steps: int, batch: int = 1024, lr: float = 1e-3, seed: int = 7,
              lr_final: float = 1.0) -> np.ndarray:
        """Adam steps on minibatches of x -> t; lr_final < 1: cosine decay of
        the learning rate to lr * lr_final over the steps of this call."""
        x = _f32(x, (-1, self.widths[0]))
        t = _f32(t, (-1, self.widths[-1]))
        _check(self.lib.mfa_mlp_set_lr_schedule(self.h, C.c_float(lr_final)), "set_lr_schedule")
        hist = np.empty(max(steps, 1), np.float32)
        _check(self.lib.mfa_mlp_train(self.h, _p(x), _p(t), x.shape[0], steps, batch, C.c_float(lr),
                                      C.c_uint64(seed), _p(hist)), "train")
        return hist[:steps]


def gemm_check(a, b) -> np.ndarray:
    lib = _load()
    a = _f32(a); b = _f32(b)
    M, K = a.shape
    K2, N = b.shape
    assert K == K2
    c = np.empty((M, N), np.float32)
    _check(lib.mfa_mlp_gemm_check(_p(a), _p(b), M, N, K, _p(c)), "gemm_check")
    return c
