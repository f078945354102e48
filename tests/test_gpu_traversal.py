"""QBVH ray-batch parity on the GPU against the reference's own QBVH
(libSLR/Accelerator/QBVH.h:355-376 via oracle/_ref): hit primitive IDs
bit-exact, hit t bit-exact (the tolerance the target allows, 1e-5 relative,
is not needed), surface parameters bit-exact; visibility exact."""
import numpy as np
import pytest

from slr_amd import build_scene, scenes
from tests.conftest import need_ref

pytestmark = pytest.mark.gpu


def _compare(desc, ctx, n, seed, mode="mixed", standard_bvh=False):
    refdrv = need_ref()
    sc = build_scene(ctx, dict(desc, standard_bvh=standard_bvh))
    ref = refdrv.RefScene(desc)
    ref.build_accel()
    accel = 2 if standard_bvh else 0
    sc.query_accel(standard_bvh)
    rays = scenes.random_rays(desc, n, seed=seed, mode=mode)
    t, prim, b0, b1 = sc.intersect(rays)
    rt, rprim, rb0, rb1 = ref.intersect(rays, accel=accel)
    miss = rprim < 0
    same = prim == rprim
    if standard_bvh:
        # the binary BVH visits triangles in the reference's own order, so
        # even ties at shared edges resolve to the same triangle
        assert np.array_equal(prim, rprim), "prim mismatch at %s" % np.nonzero(~same)[0][:10]
    # ties within 1e-6 in t may legitimately pick either triangle
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
    rv = ref.visibility(seg, accel=accel)
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


@pytest.mark.parametrize("which", ["cornell", "icosphere", "icosphere_axis"])
def test_standard_bvh_ray_batch_parity(gpu_ctx, which):
    """Closest hits and visibility through the binned-SAH binary BVH against
    the reference's StandardBVH(objs, BinnedSAH) traversal
    (libSLR/Accelerator/StandardBVH.h:266-328): primitive IDs, t and surface
    parameters bit-exact, ties included."""
    if which == "cornell":
        _compare(scenes.cornell_box(64, 64), gpu_ctx, 200000, seed=13, standard_bvh=True)
    else:
        _compare(scenes.cornell_icosphere(64, 64, subdiv=5), gpu_ctx, 200000, seed=15,
                 mode="axis" if which.endswith("axis") else "mixed", standard_bvh=True)


def test_standard_bvh_query_switch(gpu_ctx):
    """Switching the query accelerator back returns the QBVH's hits; a scene
    built without the binary BVH refuses the switch."""
    desc = scenes.cornell_box(32, 32)
    sc = build_scene(gpu_ctx, dict(desc, standard_bvh=True))
    rays = scenes.random_rays(desc, 10000, seed=21)
    q = sc.intersect(rays)
    sc.query_accel(True)
    b = sc.intersect(rays)
    sc.query_accel(False)
    q2 = sc.intersect(rays)
    assert all(np.array_equal(x, y) for x, y in zip(q, q2))
    assert np.array_equal(q[1] < 0, b[1] < 0)
    plain = build_scene(gpu_ctx, desc)
    with pytest.raises(Exception):
        plain.query_accel(True)


def test_empty_batch(gpu_ctx):
    sc = build_scene(gpu_ctx, scenes.cornell_box(8, 8))
    t, prim, b0, b1 = sc.intersect(np.zeros((0, 8), np.float32))
    assert len(t) == 0


@pytest.mark.parametrize("name", ["cornell", "icosphere6", "c2"])
def test_traversal_counters_match_an_instrumented_cpu_traversal(gpu_ctx, name):
    """The GPU traversal's node and triangle counters (the algorithmic bytes
    of the roofline are priced on them) equal an independent count over the
    same rays: QBVH::commonProcedure restated on the CPU with counters
    (oracle orc_qbvh_trace) over the same tree, for closest and any hits."""
    from oracle import oracle
    desc = {"cornell": lambda: scenes.cornell_box(16, 16),
            "icosphere6": lambda: scenes.cornell_sphere(16, 16, subdiv=6),
            "c2": lambda: scenes.cornell_icosphere(64, 36, subdiv=8)}[name]()
    sc = build_scene(gpu_ctx, desc)
    rays = scenes.random_rays(desc, 1 << 14, seed=29)
    orc = oracle.OracleScene(desc)
    dump = sc.qbvh_dump()
    (t, prim, _, _), n_cpu, t_cpu = orc.qbvh_trace(dump, rays)
    gt, gprim, _, _ = sc.intersect(rays)
    n_gpu, t_gpu = sc.query_counts()
    assert np.array_equal(gt.view(np.uint32), t.view(np.uint32)) and np.array_equal(gprim, prim)
    assert (n_gpu, t_gpu) == (n_cpu, t_cpu)
    vis, na_cpu, ta_cpu = orc.qbvh_trace(dump, rays, any_hit=True)
    gvis = sc.occluded(rays)
    na_gpu, ta_gpu = sc.query_counts()
    assert np.array_equal(gvis, vis)
    assert (na_gpu, ta_gpu) == (na_cpu, ta_cpu)
    print("counts %s: %.3f nodes, %.3f triangles per closest-hit ray; %.3f, %.3f per any-hit ray" % (
        name, n_cpu / len(rays), t_cpu / len(rays), na_cpu / len(rays), ta_cpu / len(rays)))
