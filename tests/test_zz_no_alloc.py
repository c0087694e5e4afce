"""The allocation-free check_fluid_fluid against the allocating one, through
the C ABI: same V_ff bits on two calls in a row, the second on a smaller
channel, with caller-owned device arrays on a caller-created stream (from
the HIP runtime the library links); a channel beyond the scratch capacity
is refused before any launch."""
import ctypes as C

import numpy as np
import pytest

from mfa_brush_amd import _lib, configs
from mfa_brush_amd.sim import Simulation, to_c

pytestmark = pytest.mark.gpu


def _channel(kind, seed):
    if kind == "c1":
        p, t, r0, v = configs.channel_c1(thermostat=True, rc_dpd=1.0)
    else:
        p, t, r0, v = configs.channel_scaled(6, n_mon=8, lz=14.0, seed=100)
    rng = np.random.default_rng(seed)
    r0 = r0 + rng.normal(0, 0.05, r0.shape)
    r0[:, 0] %= p.boundary[0]
    r0[:, 1] %= p.boundary[1]
    return p, t, r0, v


class _Hip:
    """The HIP runtime the library itself links (libamdhip64.so.7, already
    loaded with libmfa_hip.so): device buffers and a stream of the caller."""

    def __init__(self):
        _lib.load()
        h = C.CDLL("libamdhip64.so.7")
        vp = C.c_void_p
        h.hipMalloc.argtypes = [C.POINTER(vp), C.c_size_t]
        h.hipFree.argtypes = [vp]
        h.hipMemcpy.argtypes = [vp, vp, C.c_size_t, C.c_int]
        h.hipStreamCreate.argtypes = [C.POINTER(vp)]
        h.hipStreamSynchronize.argtypes = [vp]
        h.hipStreamDestroy.argtypes = [vp]
        for f in (h.hipMalloc, h.hipFree, h.hipMemcpy, h.hipStreamCreate, h.hipStreamSynchronize, h.hipStreamDestroy):
            f.restype = C.c_int
        self.h = h

    def ok(self, rc, what):
        assert rc == 0, "%s -> hipError %d" % (what, rc)

    def to_device(self, a):
        a = np.ascontiguousarray(a)
        d = C.c_void_p()
        self.ok(self.h.hipMalloc(C.byref(d), a.nbytes), "hipMalloc")
        self.ok(self.h.hipMemcpy(d, a.ctypes.data_as(C.c_void_p), a.nbytes, 1), "hipMemcpy H2D")
        return d

    def alloc(self, nbytes):
        d = C.c_void_p()
        self.ok(self.h.hipMalloc(C.byref(d), nbytes), "hipMalloc")
        return d

    def to_host(self, d, shape, dtype):
        out = np.empty(shape, dtype)
        self.ok(self.h.hipMemcpy(out.ctypes.data_as(C.c_void_p), d, out.nbytes, 2), "hipMemcpy D2H")
        return out


def _noalloc(hip, lib, p, t, r0, stream):
    d_r0 = hip.to_device(np.ascontiguousarray(r0, np.float64))
    d_t = hip.to_device(np.ascontiguousarray(t, np.int32))
    d_v = hip.to_device(np.array([np.nan]))
    try:
        prm = to_c(p)
        rc = lib.mfa_hip_check_fluid_fluid_noalloc(C.byref(prm), d_r0, d_t, int(p.n_part), 0.0, d_v, stream)
        _lib.check(rc, "check_fluid_fluid_noalloc")
        hip.ok(hip.h.hipStreamSynchronize(stream), "hipStreamSynchronize")
        return float(hip.to_host(d_v, (1,), np.float64)[0])
    finally:
        for d in (d_r0, d_t, d_v):
            hip.h.hipFree(d)


def test_noalloc_check_fluid_fluid_matches_allocating_entry_point():
    lib = _lib.load()
    hip = _Hip()
    stream = C.c_void_p()
    hip.ok(hip.h.hipStreamCreate(C.byref(stream)), "hipStreamCreate")
    got = []
    try:
        for kind, seed in (("c1", 5), ("small", 6)):          # the second call is on a smaller channel
            p, t, r0, v = _channel(kind, seed)
            vn = _noalloc(hip, lib, p, t, r0, stream)
            with Simulation(p) as s:
                s.set_state(r0, v, t)
                vr = s.check_fluid_fluid()
            print("%s: n=%d noalloc %.17g allocating %.17g" % (kind, p.n_part, vn, vr))
            assert vn == vr
            got.append((p.n_part, vn))
    finally:
        hip.h.hipStreamDestroy(stream)
    assert got[1][0] < got[0][0] and got[0][1] != got[1][1]


def test_noalloc_refuses_beyond_capacity():
    lib = _lib.load()
    p, t, r0, v = _channel("small", 1)
    prm = to_c(p)
    dummy = C.c_void_p(1)                      # never dereferenced: the size check comes first
    rc = lib.mfa_hip_check_fluid_fluid_noalloc(C.byref(prm), dummy, dummy, (1 << 20) + 1, 0.0, dummy, None)
    assert rc == -1
    assert b"exceeds the scratch capacity" in lib.mfa_hip_last_error()
