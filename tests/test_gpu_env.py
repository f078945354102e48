"""Image-based lighting: an equirectangular environment (the reference's
InfiniteSphereNode with an image texture) seen directly, reflected, sampled
by next-event estimation against its importance map and combined with BSDF
sampling by MIS.

* a uniform white environment seen directly has luminance 1 (Meng
  upsampling of sRGB white as a light source, stored as half uvs);
* a Lambert plane of constant albedo filling the view reflects albedo times
  the uniform environment (nothing shadows its hemisphere);
* an open scene under a sky with a sun matches the compiled reference
  PTRenderer per pixel and XYZ channel within 3 combined sigma.
"""
import numpy as np
import pytest

import slr_amd
from slr_amd import scenes
from tests import refcases

pytestmark = pytest.mark.gpu


def _render(ctx, desc, W, H, spp, seed, **kw):
    sc = slr_amd.build_scene(ctx, desc)
    film = slr_amd.Film(ctx, W, H)
    st = slr_amd.PTRenderer(spp).render(sc, film, seed=seed, **kw)
    return film, st


def test_uniform_environment_seen_directly(gpu_ctx):
    W = H = 16
    spp = 64
    desc = scenes.env_furnace(W, H, quad_in_view=False)
    film, st = _render(gpu_ctx, desc, W, H, spp, seed=3)
    y = film.xyz(1.0 / spp)[..., 1] / scenes.pixel_cos4(desc, W, H)
    assert np.isfinite(y).all()
    assert abs(y.mean() - 1) < 0.01, y.mean()
    # every camera ray escapes: no shading of surfaces, no shadow rays
    assert st["shadow_rays"] == 0


def test_uniform_environment_furnace(gpu_ctx):
    W = H = 16
    spp = 256
    albedo = 0.5
    films = []
    for in_view in (False, True):
        desc = scenes.env_furnace(W, H, albedo=albedo, quad_in_view=in_view)
        film, st = _render(gpu_ctx, desc, W, H, spp, seed=8)
        films.append(film.strata().astype(np.float64))
    # a flat reflectance scales every wavelength stratum by the albedo
    r = films[1].sum(axis=(0, 1)) / films[0].sum(axis=(0, 1))
    assert np.all(np.abs(r / albedo - 1) < 0.02), r
    assert abs(films[1].sum() / films[0].sum() / albedo - 1) < 0.005


def test_sky_matches_reference_renderer_3sigma(gpu_ctx):
    """Reference PTRenderer image of an open scene under a sky with a sun
    (32 x 24, 32 batches of 32 spp, tests/golden/ref_sky.npz) vs 32 GPU
    batches of 32 spp."""
    make, W, H, _, spp, _, _ = refcases.CASES["sky"]
    sc = slr_amd.build_scene(gpu_ctx, make())
    cmf = slr_amd.cmf_strata()
    gb = []
    for b in range(32):
        film = slr_amd.Film(gpu_ctx, W, H)
        slr_amd.PTRenderer(spp).render(sc, film, seed=300 + b)
        gb.append(refcases.xyz(film.strata(), cmf) / spp)
    refcases.check_3sigma(np.array(gb), refcases.load("sky"), name="sky")


def test_environment_adds_light(gpu_ctx):
    """The Cornell box is open toward the camera: a sky adds light through
    the opening on top of the ceiling light's contribution."""
    W = H = 16
    desc = scenes.cornell_box(W, H)
    a, _ = _render(gpu_ctx, desc, W, H, 64, seed=2)
    desc["environment"] = {"rgb": scenes.constant_environment(1.0), "scale": 1.0}
    b, st = _render(gpu_ctx, desc, W, H, 64, seed=2)
    assert np.isfinite(b.strata()).all()
    assert b.strata().sum() > a.strata().sum() * 1.01
