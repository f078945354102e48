"""Output stage: the film encoded as ImageSensor::saveImage writes it
(libSLR/Core/ImageSensor.cpp:139-195, libSLR/Helper/bmp_exporter.cpp:15-54),
compared byte for byte with the compiled reference's saveImage run on the
same sums, and progressive 2^k snapshots as PTRenderer::render takes them
(libSLR/Renderer/PTRenderer.cpp:77-94)."""
import numpy as np
import pytest

import slr_amd
from slr_amd import scenes
from tests.conftest import need_ref

pytestmark = pytest.mark.gpu


def _sums(rng, h, w):
    """Stratum sums spread over the tone curve: zeros, tiny, moderate, huge,
    and strata mixes that give negative sRGB before the clamp."""
    a = rng.exponential(1.0, size=(h, w, 16)) * rng.choice([0.0, 1e-4, 0.05, 0.5, 3.0, 40.0], size=(h, w, 1))
    a[rng.uniform(size=(h, w)) < 0.2, :8] = 0.0
    return a.astype(np.float32)


@pytest.mark.parametrize("w,h", [(16, 8), (21, 5)])
def test_bmp_matches_reference_save_image(gpu_ctx, tmp_path, w, h):
    refdrv = need_ref()
    rng = np.random.RandomState(w * 31 + h)
    sums = _sums(rng, h, w)
    for scale, sens in ((1.0 / 16, 1.0), (0.37, 2.5), (1.0, float("inf"))):
        ref = refdrv.save_image(sums, str(tmp_path / "ref.bmp"), scale=scale, sensitivity=sens)
        film = slr_amd.Film(gpu_ctx, w, h)
        film.load(sums.astype(np.float64))
        mine = film.bmp(scale=scale, sensitivity=sens)
        assert len(mine) == len(ref)
        a = np.frombuffer(mine, np.uint8)
        b = np.frombuffer(ref, np.uint8)
        row = 3 * w + w % 4
        keep = np.ones(len(a), bool)
        for y in range(h):   # the reference leaves the row padding uninitialised
            keep[54 + y * row + 3 * w:54 + (y + 1) * row] = False
        assert np.array_equal(a[keep], b[keep]), np.nonzero(a[keep] != b[keep])[0][:10]
        # the same pixels through slrh_film_read(SRGB8), top row first, RGB
        rgb = film.srgb8(scale * (1.0 if np.isinf(sens) else sens))
        px = a[54:].reshape(h, row)[:, :3 * w].reshape(h, w, 3)[::-1, :, ::-1]
        assert np.array_equal(rgb, px)


def test_bmp_of_a_rendered_reference_film(gpu_ctx, tmp_path):
    """A film the reference renderer produced, encoded by both."""
    refdrv = need_ref()
    desc = scenes.cornell_box(24, 16)
    r = refdrv.RefScene(desc)
    f, _, _ = r.render(24, 16, 8, 7, threads=4)
    r.close()
    ref = refdrv.save_image(f, str(tmp_path / "r.bmp"), scale=1.0 / 8)
    film = slr_amd.Film(gpu_ctx, 24, 16)
    film.load(f.astype(np.float64))
    assert film.bmp(scale=1.0 / 8) == ref


def test_progressive_snapshots_at_powers_of_two(gpu_ctx):
    """Snapshots after 1, 2, 4, ... samples per pixel; the accumulated film
    is a render of the whole count: its image total lies within 4 standard
    errors of the reference renderer's (tests/golden/ref_cornell32.npz,
    whose batches have the same 64 spp)."""
    from tests import refcases
    make, W, H, _, spp, _, _ = refcases.CASES["cornell32"]
    sc = slr_amd.build_scene(gpu_ctx, make())
    film = slr_amd.Film(gpu_ctx, W, H)
    seen = []

    def snap(f, done):
        seen.append((done, f.strata().sum(), len(f.bmp(scale=1.0 / done))))
        return False

    st = slr_amd.PTRenderer(spp).render_progressive(sc, film, snap, seed=5)
    assert [d for d, _, _ in seen] == [1, 2, 4, 8, 16, 32, 64]
    assert st["samples"] == W * H * spp
    sums = [x for _, x, _ in seen]
    assert all(b > a for a, b in zip(sums, sums[1:]))
    assert all(n == 54 + 3 * W * H for _, _, n in seen)
    ref = refcases.load("cornell32")
    total = refcases.xyz(film.strata(), slr_amd.cmf_strata()).sum() / spp
    per_batch = float(ref["tot_se"]) * np.sqrt(float(ref["batches"]))
    se = np.sqrt(per_batch ** 2 + float(ref["tot_se"]) ** 2)
    assert abs(total - float(ref["tot_mean"])) < 4 * se, (total, float(ref["tot_mean"]), se)
    # a snapshot that returns true stops the render
    stop = slr_amd.Film(gpu_ctx, W, H)
    calls = []
    slr_amd.PTRenderer(spp).render_progressive(sc, stop, lambda f, d: calls.append(d) or d >= 4, seed=5)
    assert calls == [1, 2, 4]
