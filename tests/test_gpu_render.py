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
from tests import refcases

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
    # the film's CMF strata (checked against the reference's in test_host)
    return refcases.xyz(strata, slr_amd.cmf_strata())


def test_matches_reference_renderer_3sigma(gpu_ctx):
    """Reference PTRenderer image of the Cornell box (32 x 32, 32 batches of
    64 spp, tests/golden/ref_cornell32.npz) vs 32 GPU batches of 64 spp."""
    make, W, H, _, spp, _, _ = refcases.CASES["cornell32"]
    sc = slr_amd.build_scene(gpu_ctx, make())
    cmf = slr_amd.cmf_strata()
    gb = []
    for b in range(32):
        film = slr_amd.Film(gpu_ctx, W, H)
        slr_amd.PTRenderer(spp).render(sc, film, seed=100 + b)
        gb.append(refcases.xyz(film.strata(), cmf) / spp)
    refcases.check_3sigma(np.array(gb), refcases.load("cornell32"), name="cornell32")


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
