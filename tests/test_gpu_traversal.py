"""QBVH ray-batch parity on the GPU against the reference's own QBVH
(libSLR/Accelerator/QBVH.h:355-376 via oracle/_ref): hit primitive IDs
bit-exact, hit t bit-exact (the tolerance the target allows, 1e-5 relative,
is not needed), surface parameters bit-exact; visibility exact."""
import numpy as np
import pytest

from slr_amd import build_scene, scenes
from tests.conftest import need_ref

pytestmark = pytest.mark.gpu


def _compare(desc, ctx, n, seed, mode="mixed"):
    refdrv = need_ref()
    sc = build_scene(ctx, desc)
    ref = refdrv.RefScene(desc)
    ref.build_accel()
    rays = scenes.random_rays(desc, n, seed=seed, mode=mode)
    t, prim, b0, b1 = sc.intersect(rays)
    rt, rprim, rb0, rb1 = ref.intersect(rays)
    miss = rprim < 0
    # ties within 1e-6 in t may legitimately pick either triangle
    same = prim == rprim
    with np.errstate(invalid="ignore"):
        tie = np.abs(t - rt) <= 1e-6 * np.maximum(1.0, np.abs(rt))
    assert np.all(same | tie), \
        "prim mismatch at %s" % np.nonzero(~same)[0][:10]
    hit = ~miss & same
    assert np.array_equal(t[hit].view(np.uint32), rt[hit].view(np.uint32))
    assert np.array_equal(b0[hit].view(np.uint32), rb0[hit].view(np.uint32))
    assert np.array_equal(b1[hit].view(np.uint32), rb1[hit].view(np.uint32))
    assert np.all(prim[miss] == -1)
    # shadow segments to the hit points (QBVH::testVisibility)
    seg = rays.copy()
    seg[:, 6] = 1e-4
    seg[:, 7] = np.where(miss, 10.0, rt * 0.999)
    v = sc.occluded(seg)
    rv = ref.visibility(seg)
    assert np.array_equal(v, rv)
    return float((~miss).mean())


def test_cornell_intersect_parity(gpu_ctx):
    frac = _compare(scenes.cornell_box(64, 64), gpu_ctx, 200000, seed=3)
    assert frac > 0.5   # the box is open towards the camera


def test_cornell_axis_aligned_rays(gpu_ctx):
    _compare(scenes.cornell_box(64, 64), gpu_ctx, 50000, seed=4, mode="axis")


def test_icosphere_intersect_parity(gpu_ctx):
    frac = _compare(scenes.cornell_icosphere(64, 64, subdiv=5), gpu_ctx, 200000, seed=5)
    assert frac > 0.5


def test_empty_batch(gpu_ctx):
    sc = build_scene(gpu_ctx, scenes.cornell_box(8, 8))
    t, prim, b0, b1 = sc.intersect(np.zeros((0, 8), np.float32))
    assert len(t) == 0
