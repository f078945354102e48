"""Image parity of the GPU wavefront path tracer.

* against the oracle (oracle/slr_oracle.c) with the same seed and slot
  layout: the two share every random number, so films agree to float
  rounding except for the rare path whose branch flips on a last-ulp
  difference (libm cos/sin, FP32 reassociation on the GPU);
* against the compiled reference PTRenderer (oracle/_ref): per pixel and XYZ
  channel within 3 sigma of the combined Monte Carlo error (batch means);
* against the white-furnace closed form.
"""
import numpy as np
import pytest

import slr_amd
from oracle import oracle
from slr_amd import scenes
from tests.conftest import need_ref

pytestmark = pytest.mark.gpu


def _gpu_film(ctx, desc, W, H, spp, seed, K=0, max_depth=0):
    sc = slr_amd.build_scene(ctx, desc)
    film = slr_amd.Film(ctx, W, H)
    st = slr_amd.PTRenderer(spp).render(sc, film, seed=seed, slots_per_pixel=K, max_depth=max_depth)
    return film.strata(), st, film


def test_matches_oracle_same_seed(gpu_ctx):
    W, H, spp, K = 32, 32, 16, 2
    desc = scenes.cornell_box(W, H)
    img, st, _ = _gpu_film(gpu_ctx, desc, W, H, spp, seed=11, K=K)
    ref, rays = oracle.OracleScene(desc).render(W, H, spp, 11, K=K)
    assert st["samples"] == W * H * spp
    rel = np.abs(img - ref) / np.maximum(np.abs(ref), 1e-6)
    # nearly every pixel/stratum equal to FP32 rounding; a handful of paths may diverge
    assert np.mean(rel < 1e-4) > 0.97, np.mean(rel < 1e-4)
    assert abs(img.sum() - ref.sum()) / ref.sum() < 2e-3
    # the two count the same rays up to diverged paths
    assert abs(st["rays"] - rays) / rays < 2e-3


def test_white_furnace(gpu_ctx):
    W = H = 16
    spp = 512
    desc = scenes.furnace(W, H, albedo=0.5, emittance=1.0)
    img, st, _ = _gpu_film(gpu_ctx, desc, W, H, spp, seed=5)
    expect = scenes.pixel_cos4(desc, W, H) * (1.0 / np.pi) / (1 - 0.5) * spp
    r = img.mean(axis=2) / expect
    assert abs(r.mean() - 1) < 0.01, r.mean()


def _xyz(strata):
    from oracle import refdrv
    x, y, z, integ = refdrv.cmf_strata()
    return np.stack([strata @ x, strata @ y, strata @ z], -1) / integ


def test_matches_reference_renderer_3sigma(gpu_ctx):
    """Reference PTRenderer vs GPU on the Cornell box, 32x32, 16 batches of
    64 spp each side; per-pixel XYZ within 3 combined sigma."""
    refdrv = need_ref()
    W = H = 32
    B, spp = 16, 64
    desc = scenes.cornell_box(W, H)
    import os
    threads = max(1, min(16, len(os.sched_getaffinity(0))))
    rb = []
    for b in range(B):
        r = refdrv.RefScene(desc)
        f, _, _ = r.render(W, H, spp, 1000 + b, threads=threads)
        rb.append(_xyz(f.astype(np.float64)) / spp)
        r.close()
    sc = slr_amd.build_scene(gpu_ctx, desc)
    gb = []
    for b in range(B):
        film = slr_amd.Film(gpu_ctx, W, H)
        slr_amd.PTRenderer(spp).render(sc, film, seed=77 + b)
        gb.append(_xyz(film.strata()) / spp)
    rb, gb = np.array(rb), np.array(gb)
    mr, mg = rb.mean(0), gb.mean(0)
    se = np.sqrt(rb.var(0, ddof=1) / B + gb.var(0, ddof=1) / B)
    z = np.abs(mg - mr) / np.maximum(se, 1e-12)
    fail = np.mean(z > 3.0)
    relmse = np.mean((mg - mr) ** 2 / (mr ** 2 + 1e-3))
    # 3 sigma with 15-dof variance estimates: expect ~1% chance failures
    assert fail < 0.03, (fail, relmse)
    tr, tg = rb.sum(axis=(1, 2, 3)), gb.sum(axis=(1, 2, 3))
    se_t = np.sqrt(tr.var(ddof=1) / B + tg.var(ddof=1) / B)
    assert abs(tg.mean() - tr.mean()) < 4 * se_t
    assert relmse < 0.05


def test_slots_per_pixel_do_not_change_the_estimate(gpu_ctx):
    W = H = 32
    desc = scenes.cornell_box(W, H)
    a, _, _ = _gpu_film(gpu_ctx, desc, W, H, 256, seed=1, K=1)
    b, _, _ = _gpu_film(gpu_ctx, desc, W, H, 256, seed=2, K=8)
    assert abs(a.sum() / b.sum() - 1) < 0.01


def test_max_depth_one_is_direct_emission_plus_nee(gpu_ctx):
    W = H = 16
    desc = scenes.cornell_box(W, H)
    full, _, _ = _gpu_film(gpu_ctx, desc, W, H, 64, seed=1)
    short, st, _ = _gpu_film(gpu_ctx, desc, W, H, 64, seed=1, max_depth=1)
    assert short.sum() < full.sum()
    # one bounce: camera ray, no shadow rays beyond the first vertex's ... none
    assert st["shadow_rays"] == 0


def test_film_formats(gpu_ctx):
    W = H = 8
    desc = scenes.cornell_box(W, H)
    strata, _, film = _gpu_film(gpu_ctx, desc, W, H, 16, seed=4)
    xyz = film.xyz(1.0 / 16)
    np.testing.assert_allclose(xyz, _xyz(strata / 16), rtol=1e-5, atol=1e-9)
    rgb = film.srgb8(1.0 / 16)
    assert rgb.dtype == np.uint8 and rgb.shape == (H, W, 3) and rgb.max() > 0
