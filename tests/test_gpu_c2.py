"""Parity on the C2 workload's own scene: the Cornell box plus the displaced
icosphere at subdivision 8 (1,310,720 triangles, GGX conductor).

* ray batches: 2^20 seeded rays (half uniformly through the box, half aimed
  at the icosphere) through the GPU QBVH and through the reference's own
  QBVH (libSLR/Accelerator/QBVH.h:355-376 via oracle/_ref, built from the
  same SBVH): hit triangle and t bit-exact except ties within 1e-6 in t,
  surface parameters bit-exact, visibility of the segments to the hits exact;
* the image: the C2 scene at 128x72 (the 1080p view's aspect) against the
  reference PTRenderer's statistics (tests/golden/ref_c2.npz, 32 batches of
  64 spp), within 3 sigma per pixel and XYZ channel;
* serving on the resident C2 scene: a 1920x1080 request rendered alone and
  among 256^2 / 512^2 requests gives the same film bit for bit.
"""
import numpy as np
import pytest

import slr_amd
from slr_amd import scenes
from tests import refcases
from tests.conftest import need_ref

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c2(gpu_ctx):
    desc = scenes.cornell_icosphere(1920, 1080, subdiv=8)
    return desc, slr_amd.build_scene(gpu_ctx, desc)


def c2_rays(desc, n, seed):
    """n rays: half uniform through the scene box, half from random points in
    the box towards jittered icosphere vertices (the dense mesh)."""
    rng = np.random.RandomState(seed)
    half = n // 2
    a = scenes.random_rays(desc, half, seed=seed)
    sphere = np.asarray(desc["meshes"][-1]["pos"], np.float64)
    allp = np.concatenate([m["pos"] for m in desc["meshes"]]).astype(np.float64)
    lo, hi = allp.min(0), allp.max(0)
    org = lo + (hi - lo) * rng.uniform(0.02, 0.98, size=(n - half, 3))
    tgt = sphere[rng.randint(len(sphere), size=n - half)] + rng.normal(scale=1e-3, size=(n - half, 3))
    d = tgt - org
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    b = np.zeros((n - half, 8), np.float32)
    b[:, :3] = org
    b[:, 3:6] = d
    b[:, 7] = np.inf
    return np.concatenate([a, b])


def test_c2_ray_batch_parity(c2):
    refdrv = need_ref()
    desc, sc = c2
    n = 1 << 20
    rays = c2_rays(desc, n, seed=81)
    t, prim, b0, b1 = sc.intersect(rays)
    print("c2 ray batch: GPU done, building the reference QBVH", flush=True)
    ref = refdrv.RefScene(desc)
    ref.build_accel()
    rt, rprim, rb0, rb1 = ref.intersect(rays)
    miss = rprim < 0
    same = prim == rprim
    with np.errstate(invalid="ignore"):
        tie = np.abs(t - rt) <= 1e-6 * np.maximum(1.0, np.abs(rt))
    assert np.all(same | tie), "prim mismatch at %s" % np.nonzero(~(same | tie))[0][:10]
    hit = ~miss & same
    assert np.array_equal(t[hit].view(np.uint32), rt[hit].view(np.uint32))
    assert np.array_equal(b0[hit].view(np.uint32), rb0[hit].view(np.uint32))
    assert np.array_equal(b1[hit].view(np.uint32), rb1[hit].view(np.uint32))
    assert np.all(prim[miss] == -1)
    # most aimed rays end on the icosphere (its triangles come after the box's 36)
    assert (rprim[n // 2:] >= 36).mean() > 0.5
    seg = rays.copy()
    seg[:, 6] = 1e-4
    seg[:, 7] = np.where(miss, 10.0, rt * 0.999)
    assert np.array_equal(sc.occluded(seg), ref.visibility(seg))
    ref.close()
    print("c2 ray batch: %d rays, %.1f%% hits, %d ties" % (n, 100 * (~miss).mean(), int((~same).sum())))


def test_c2_image_matches_reference_3sigma(gpu_ctx):
    make, W, H, B, spp, _, _ = refcases.CASES["c2"]
    sc = slr_amd.build_scene(gpu_ctx, make())
    cmf = slr_amd.cmf_strata()
    gb = []
    for b in range(B):
        film = slr_amd.Film(gpu_ctx, W, H)
        slr_amd.PTRenderer(spp).render(sc, film, seed=300 + b)
        gb.append(refcases.xyz(film.strata(), cmf) / spp)
    refcases.check_3sigma(np.array(gb), refcases.load("c2"), name="c2")


def test_c2_served_1080p_request_same_alone_and_in_mix(gpu_ctx, c2):
    desc, sc = c2
    big = scenes.camera_pose(desc, yaw_deg=4.0, width=1920, height=1080)
    srv = slr_amd.Server(sc, path_state_bytes=8 << 30)
    try:
        alone = slr_amd.Film(gpu_ctx, 1920, 1080)
        assert srv.wait(srv.submit(alone, big, spp=4, seed=123, max_depth=10), 120000) == slr_amd.REQ_DONE
        reqs = scenes.random_requests(desc, 10, seed=9, sizes=((256, 256), (512, 512)), spps=(4, 16))
        films = [slr_amd.Film(gpu_ctx, q["width"], q["height"]) for q in reqs]
        mixed = slr_amd.Film(gpu_ctx, 1920, 1080)
        ts = [srv.submit(f, q["camera"], spp=q["spp"], seed=q["seed"], max_depth=10)
              for f, q in zip(films[:5], reqs[:5])]
        tm = srv.submit(mixed, big, spp=4, seed=123, max_depth=10)
        ts += [srv.submit(f, q["camera"], spp=q["spp"], seed=q["seed"], max_depth=10)
               for f, q in zip(films[5:], reqs[5:])]
        assert [srv.wait(t, 120000) for t in ts + [tm]] == [slr_amd.REQ_DONE] * 11
        a = alone.strata()
        assert np.isfinite(a).all() and a.sum() > 0
        assert np.array_equal(mixed.strata().view(np.uint64), a.view(np.uint64))
        for f in films:
            assert f.strata().sum() > 0
    finally:
        srv.close()
