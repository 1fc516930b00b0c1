"""The allocation-free entry point against the allocating one, on the GPU.

orion_intersect_closest_noalloc keeps its traversal spill area and overflow
counter in a static per-GPU device array and runs a fixed grid; its hits
must equal orion_intersect_closest's bit for bit.  Two calls in a row on the
1.31M-triangle scene: 2,000,003 rays (every wave of the fixed grid takes
two chunks or more), then 300,001 rays (the array keeps the first call's
contents; the counter is cleared by the call)."""
import ctypes as C

import numpy as np
import pytest

import orion_amd as oa
from test_gpu_parity import random_rays

pytestmark = pytest.mark.gpu


class _Hip:
    """Device buffers through the HIP runtime the library itself links
    (torch bundles another one, which must not be mixed into this process)."""

    def __init__(self):
        h = C.CDLL("libamdhip64.so.7")
        h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        h.hipFree.argtypes = [C.c_void_p]
        h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.h = h
        self.bufs = []

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.h.hipMalloc(C.byref(p), nbytes) == 0
        self.bufs.append(p)
        return p

    def upload(self, a):
        a = np.ascontiguousarray(a)
        p = self.alloc(a.nbytes)
        assert self.h.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), a.nbytes, 1) == 0  # host to device
        return p

    def download(self, p, n, dtype):
        out = np.empty(n, dtype)
        assert self.h.hipMemcpy(out.ctypes.data_as(C.c_void_p), p, out.nbytes, 2) == 0  # device to host
        return out

    def free(self):
        for p in self.bufs:
            self.h.hipFree(p)
        self.bufs = []


def _noalloc(ds, rays, with_uv=True):
    hip = _Hip()
    try:
        n = rays.shape[1]
        dr = hip.upload(np.ascontiguousarray(rays, np.float32))
        t, prim, u, v, tie = (hip.alloc(4 * n) for _ in range(5))
        opt = (lambda x: x) if with_uv else (lambda x: None)
        rc = oa.lib().orion_intersect_closest_noalloc(ds.h, n, dr, t, prim, opt(u), opt(v), opt(tie))
        assert rc == 0, oa.lib().orion_last_error()
        return (hip.download(t, n, np.float32), hip.download(prim, n, np.int32), hip.download(u, n, np.float32),
                hip.download(v, n, np.float32), hip.download(tie, n, np.int32))
    finally:
        hip.free()


def test_noalloc_intersect_equals_allocating(device):
    hs = oa.HostScene.generate("icosphere", seed=1, params=[8, 64, 64])
    ds = oa.DeviceScene(device, hs)
    assert ds.stats()["stack_bound"] <= 48
    for n, seed in ((2_000_003, 5), (300_001, 6)):
        rays = random_rays(hs, n, seed=seed)
        want = ds.intersect_closest(rays)
        got = _noalloc(ds, rays)
        hit = want[1] >= 0
        print(f"noalloc: {n} rays, {hit.mean():.3f} hit")
        assert 0.2 < hit.mean() < 1.0
        for a, b in zip(got, want):
            np.testing.assert_array_equal(np.ascontiguousarray(a).view(np.int32),
                                          np.ascontiguousarray(b).view(np.int32))
    # u, v and tie are optional
    rays = random_rays(hs, 70_000, seed=7)
    t, prim, *_ = _noalloc(ds, rays, with_uv=False)
    want = ds.intersect_closest(rays)
    np.testing.assert_array_equal(t.view(np.int32), want[0].view(np.int32))
    np.testing.assert_array_equal(prim, want[1])
