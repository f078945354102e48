"""The oracle's integrator against the reference renderers' image statistics
(tests/golden/ref_*.npz; 3 sigma per pixel and XYZ channel over batch
means), the white-furnace closed form, and the reference harness's
environment against its closed form."""
import os

import numpy as np

import slr_amd
from oracle import oracle
from slr_amd import scenes
from tests import refcases
from tests.conftest import need_ref


def _xyz(strata, cmf):
    x, y, z, integ = cmf
    return np.stack([strata @ x, strata @ y, strata @ z], -1) / integ


def test_white_furnace_closed_form():
    W = H = 16
    spp = 256
    desc = scenes.furnace(W, H, albedo=0.7, emittance=2.0)
    film, _ = oracle.OracleScene(desc).render(W, H, spp, 9, threads=8)
    expect = scenes.pixel_cos4(desc, W, H) * (2.0 / np.pi) / (1 - 0.7) * spp
    r = film.mean(axis=2) / expect
    assert abs(r.mean() - 1) < 0.01, r.mean()


def test_oracle_matches_reference_renderer():
    """The oracle's Cornell box (32 x 32, 16 batches of 64 spp) against the
    reference PTRenderer's image statistics (tests/golden/ref_cornell32.npz):
    per pixel and XYZ channel within 3 combined standard errors, image total
    within 4 (batch sums are heavy-tailed: light seen through BSDF sampling)."""
    make, W, H, _, spp, _, _ = refcases.CASES["cornell32"]
    cmf = slr_amd.cmf_strata()
    threads = max(1, min(8, len(os.sched_getaffinity(0))))
    orc = oracle.OracleScene(make())
    ob = []
    for b in range(16):
        f, _ = orc.render(W, H, spp, 900 + b, threads=threads)
        ob.append(refcases.xyz(f, cmf) / spp)
    refcases.check_3sigma(np.array(ob), refcases.load("cornell32"))


def test_reference_environment_closed_form():
    """The harness's environment (RGBA16F image, InfiniteSphereNode) renders a
    uniform white sky as luminance 1 per unit cos^4 weight when seen
    directly, and albedo times that off a Lambert plane that fills the view.
    The image is a whole number of the reference sensor's 8-pixel tiles: the
    reference traces whole tiles and files the samples of pixels beyond the
    image into the edge pixels (ImageSensor.cpp:125-130)."""
    refdrv = need_ref()
    W = H = 16
    spp = 32
    cmf = refdrv.cmf_strata()
    out = []
    for in_view in (False, True):
        desc = scenes.env_furnace(W, H, quad_in_view=in_view)
        desc["materials"] = [{"kind": "lambert", "rgb": (0.5, 0.5, 0.5)}]
        r = refdrv.RefScene(desc)
        f, _, _ = r.render(W, H, spp, 3, threads=max(1, min(8, len(os.sched_getaffinity(0)))))
        r.close()
        out.append(_xyz(f.astype(np.float64), cmf)[..., 1] / spp / scenes.pixel_cos4(desc, W, H))
    assert abs(out[0].mean() - 1) < 0.01, out[0].mean()
    assert abs(out[1].mean() - 0.5) < 0.01, out[1].mean()


def test_oracle_volumetric_compat_matches_reference_volumetric_renderer():
    """The reference VolumetricPTRenderer over the Cornell box without media
    (16 x 16, tests/golden/ref_volpt_cornell16.npz) against the oracle in its
    volumetric compat mode (a visible light sample in the shading plane turns
    the sample into NaN, which is not filed), 3 sigma per pixel and channel.
    The same oracle without the compat mode is the path tracer, whose image
    carries about twice the energy: the light is an emitter over a Lambert
    base, so a path that reaches it samples a light point in its own plane,
    and the reference drops every sample that sees the light directly."""
    make, W, H, _, spp, _, _ = refcases.CASES["volpt_cornell16"]
    cmf = slr_amd.cmf_strata()
    threads = max(1, min(8, len(os.sched_getaffinity(0))))
    orc = oracle.OracleScene(make())
    ob, pb = [], []
    for b in range(12):
        f2, _ = orc.render(W, H, spp, 300 + b, threads=threads, volumetric=True)
        ob.append(refcases.xyz(f2, cmf) / spp)
        f3, _ = orc.render(W, H, spp, 300 + b, threads=threads)
        pb.append(refcases.xyz(f3, cmf) / spp)
    ref = refcases.load("volpt_cornell16")
    refcases.check_3sigma(np.array(ob), ref)
    assert float(ref["tot_mean"]) < 0.7 * np.array(pb).sum(axis=(1, 2, 3)).mean()
