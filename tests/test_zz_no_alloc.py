"""Allocation-free closest hit (pg_trace_noalloc) against pg_trace, through the C ABI.

The caller (this test) owns the device buffers: it allocates them with the HIP runtime the
library itself uses, passes device pointers, and reads the results back. Two calls in a row, the
second smaller, on both BVH widths of the 1.25M-triangle terrain scene (deep enough that the
traversal stack spills out of LDS into the fixed scratch array): every hit equals pg_trace's.
"""
import ctypes

import numpy as np
import pytest

import piper2_amd as pg
from piper2_amd import scenes

pytestmark = pytest.mark.gpu

_H2D, _D2H = 1, 2


def _hip():
    pg.lib()  # loads the runtime the library links; the same instance serves the calls below
    h = ctypes.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipFree.argtypes = [ctypes.c_void_p]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    h.hipStreamCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
    h.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
    h.hipStreamDestroy.argtypes = [ctypes.c_void_p]
    return h


class DeviceBuffer:
    def __init__(self, hip, nbytes):
        self.hip, self.nbytes, self.p = hip, nbytes, ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(self.p), nbytes) == 0

    def put(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        assert self.hip.hipMemcpy(self.p, a.ctypes.data, a.nbytes, _H2D) == 0

    def get(self, dtype, shape):
        out = np.zeros(shape, dtype)
        assert out.nbytes <= self.nbytes
        assert self.hip.hipMemcpy(out.ctypes.data, self.p, out.nbytes, _D2H) == 0
        return out

    def free(self):
        if self.p:
            self.hip.hipFree(self.p)
            self.p = ctypes.c_void_p()


def _rays(rng, n):
    o = np.stack([rng.uniform(-9, 9, n), rng.uniform(0.3, 3.0, n), rng.uniform(-4.5, 4.5, n)], 1)
    d = rng.normal(size=(n, 3))
    d[:, 1] = -np.abs(d[:, 1]) - 0.1
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return np.concatenate([o, d, np.full((n, 1), 1e-3), np.full((n, 1), 1e5)], 1).astype(np.float32)


def test_trace_noalloc_matches_trace():
    hip = _hip()
    dev = pg.Device(0)
    sc = scenes.terrain_spheres()
    s = pg.Scene(dev, sc)
    rng = np.random.default_rng(31)
    n_max = 1 << 18
    bufs = {k: DeviceBuffer(hip, n_max * b) for k, b in (("rays", 32), ("t", 4), ("prim", 4), ("uv", 8))}
    stream = ctypes.c_void_p()
    assert hip.hipStreamCreate(ctypes.byref(stream)) == 0
    try:
        for width in (4, 8):
            s.build_bvh(width=width)
            for call, n in enumerate((n_max, 37_003)):  # the second call is smaller
                rays = _rays(rng, n)
                t_ref, prim_ref, uv_ref = s.trace(rays)
                bufs["rays"].put(rays)
                # first call: the library's stream, waits; second: the caller's stream, enqueues only
                st = None if call == 0 else stream
                pg.check(pg.lib().pg_trace_noalloc(s.h, bufs["rays"].p, n, bufs["t"].p, bufs["prim"].p,
                                                   bufs["uv"].p, st))
                if st is not None:
                    assert hip.hipStreamSynchronize(stream) == 0
                t = bufs["t"].get(np.float32, n)
                prim = bufs["prim"].get(np.uint32, n)
                uv = bufs["uv"].get(np.float32, (n, 2))
                assert np.array_equal(prim, prim_ref), (width, call)
                assert np.array_equal(t, t_ref) and np.array_equal(uv, uv_ref), (width, call)
                assert (prim_ref != 0xFFFFFFFF).mean() > 0.5
    finally:
        hip.hipStreamDestroy(stream)
        for b in bufs.values():
            b.free()
        s.close()
        dev.close()
