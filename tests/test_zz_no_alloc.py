"""The allocation-free closest-hit entry point (slrh_intersect_noalloc)
against slrh_intersect on the same device-resident rays: hit distance,
triangle, surface parameters and instance bit-identical, on two calls in a
row, the second with fewer rays (stale scratch contents from the first call
would show there)."""
import ctypes

import numpy as np
import pytest

import slr_amd
from slr_amd import scenes

pytestmark = pytest.mark.gpu


class _DeviceArrays:
    """Device buffers through the HIP runtime (the test allocates; the entry
    point under test does not)."""

    def __init__(self):
        self.hip = ctypes.CDLL("libamdhip64.so")
        self.ptrs = []

    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(p), ctypes.c_size_t(max(4, nbytes))) == 0
        self.ptrs.append(p)
        return p.value

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = self.alloc(a.nbytes)
        assert self.hip.hipMemcpy(ctypes.c_void_p(p), a.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(a.nbytes),
                                  1) == 0   # host to device
        return p

    def get(self, p, n, dtype):
        out = np.zeros(n, dtype)
        assert self.hip.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(p),
                                  ctypes.c_size_t(out.nbytes), 2) == 0   # device to host
        return out

    def sync(self):
        assert self.hip.hipDeviceSynchronize() == 0

    def free(self):
        for p in self.ptrs:
            self.hip.hipFree(p)
        self.ptrs = []


def _trace(sc, rays, noalloc):
    dev = _DeviceArrays()
    try:
        n = len(rays)
        cols = [dev.put(rays[:, k].astype(np.float32)) for k in range(8)]
        outs = [dev.alloc(4 * n) for _ in range(5)]
        # poison the outputs so that an unwritten entry shows
        for o in outs:
            dev.hip.hipMemset(ctypes.c_void_p(o), 0x7F, ctypes.c_size_t(4 * n))
        dev.sync()
        if noalloc:
            sc.intersect_noalloc(cols, outs, n)
            dev.sync()
        else:
            sc.intersect_device(cols, outs, n)
        return (dev.get(outs[0], n, np.float32), dev.get(outs[1], n, np.int32), dev.get(outs[2], n, np.float32),
                dev.get(outs[3], n, np.float32), dev.get(outs[4], n, np.int32))
    finally:
        dev.free()


def _same(a, b):
    for x, y in zip(a, b):
        assert np.array_equal(x.view(np.uint32), y.view(np.uint32))


@pytest.mark.parametrize("case", ["icosphere", "instances"])
def test_noalloc_matches_intersect(gpu_ctx, case):
    if case == "icosphere":
        desc = scenes.cornell_icosphere(64, 64, subdiv=6)
        sizes = (300000, 70001)
    else:
        desc = scenes.instanced_cornell(32, 32, n_instances=10, seed=12, uniform_scale=True, instanced_light=True)
        sizes = (120000, 30011)
    sc = slr_amd.build_scene(gpu_ctx, desc)
    for call, n in enumerate(sizes):
        print("noalloc %s: call %d, %d rays" % (case, call, n), flush=True)
        rays = scenes.random_rays(desc, n, seed=40 + call).astype(np.float32)
        ref = _trace(sc, rays, noalloc=False)
        got = _trace(sc, rays, noalloc=True)
        _same(got, ref)
        assert (ref[1] >= 0).mean() > 0.3


def test_noalloc_refuses_too_many_rays(gpu_ctx):
    sc = slr_amd.build_scene(gpu_ctx, scenes.cornell_box(8, 8))
    with pytest.raises(slr_amd.SlrhError):
        sc.intersect_noalloc([0] * 8, [0] * 4, (1 << 30) + 1)
