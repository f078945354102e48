"""A scene written in SLR's scene language drives the renderer through the C
ABI (slrh_scene_load_text -> slrh_scene_build -> slrh_render): the Cornell
box of the parity tests, written as scene-language text by
slr_amd.scenes.to_scene_text (sRGB spectra, matte and emitter materials,
createMesh under setTransform(root, scale(0.01)) as
TestScenes/Original_Cornell_Box.txt does, the camera under its own node),
renders the reference renderer's image of that scene within 3 sigma, with
the integrator the file's setRenderer names (PT, or BPT)."""
import numpy as np
import pytest

import slr_amd
from slr_amd import scenes
from tests import refcases

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("method,case", [("PT", "cornell32"), ("BPT", "bpt_cornell32")])
def test_scene_file_renders_reference_image(gpu_ctx, method, case):
    make, W, H, B, spp, _, _ = refcases.CASES[case]
    text = scenes.to_scene_text(make(), W, H, spp, method=method)
    sc = slr_amd.Scene(gpu_ctx)
    info = sc.load_text(text)
    sc.build()
    assert (info["width"], info["height"], info["spp"]) == (W, H, spp)
    renderer = {0: slr_amd.PTRenderer, 2: slr_amd.BPTRenderer}[info["integrator"]]
    assert renderer is (slr_amd.BPTRenderer if method == "BPT" else slr_amd.PTRenderer)
    cmf = slr_amd.cmf_strata()
    gb = []
    for b in range(B):
        film = slr_amd.Film(gpu_ctx, W, H)
        renderer(info["spp"]).render(sc, film, seed=info["rng_seed"] + b)
        gb.append(refcases.xyz(film.strata(), cmf) / spp)
        film.close()
    refcases.check_3sigma(np.array(gb), refcases.load(case), name="scene_file_" + method.lower())
    sc.close()
