"""The oracle against the committed golden fixtures (tests/golden/, written
from the compiled reference by tests/golden/make_golden.py), so that it is
pinned where the compiled reference is absent:

* closest hits of 2048 rays through the Cornell box: t bit-identical to the
  reference QBVH's; triangle and surface parameters identical except where
  two triangles are hit at exactly the same t (a shared edge), which the
  brute-force oracle may resolve to the other triangle;
* the Cornell box image at 16 x 16 against the reference PTRenderer's
  16384-spp image: per pixel and XYZ channel within 3 combined standard
  errors (batch means on both sides)."""
import os

import numpy as np

from oracle import oracle
from slr_amd import scenes

HERE = os.path.dirname(os.path.abspath(__file__))


def golden(name):
    return np.load(os.path.join(HERE, "golden", name))


def test_oracle_hits_match_golden():
    g = golden("cornell_hits.npz")
    t, prim, b0, b1 = oracle.OracleScene(scenes.cornell_box(16, 16)).intersect(g["rays"])
    assert np.array_equal(t.view(np.uint32), g["t"].view(np.uint32))
    diff = prim != g["prim"]
    assert diff.sum() <= 2, diff.sum()
    # a differing triangle is a tie at the same distance: both hit there
    assert np.all(g["prim"][diff] >= 0) and np.all(prim[diff] >= 0)
    same = ~diff
    assert np.array_equal(b0[same], g["b0"][same]) and np.array_equal(b1[same], g["b1"][same])
    assert (g["prim"] >= 0).mean() > 0.8


def test_oracle_image_matches_golden():
    g = golden("cornell_16x16_image.npz")
    W = H = 16
    B, spp = 12, 128
    import slr_amd
    # the CMF strata of the film (checked against the reference's in test_host)
    x, y, z, integ = slr_amd.cmf_strata()
    orc = oracle.OracleScene(scenes.cornell_box(W, H))
    threads = max(1, min(8, len(os.sched_getaffinity(0))))
    out = []
    for b in range(B):
        f, _ = orc.render(W, H, spp, 40 + b, threads=threads)
        out.append(np.stack([f @ x, f @ y, f @ z], -1) / integ / spp)
    out = np.array(out)
    mean, se = out.mean(0), out.std(0, ddof=1) / np.sqrt(B)
    z_ = np.abs(mean - g["mean"]) / np.sqrt(se ** 2 + g["se"] ** 2)
    assert np.mean(z_ > 3) < 0.03, np.mean(z_ > 3)
    relmse = np.mean((mean - g["mean"]) ** 2 / (g["mean"] ** 2 + 1e-3))
    assert relmse < 0.02, relmse


def _fixture_pair(a, b, per_pixel=True):
    """Two committed reference image fixtures of one scene: image totals
    within 4 combined standard errors; with per_pixel also each pixel and XYZ
    channel within 3 combined standard errors for all but 1.5 % of them
    (about 0.4 % fail by chance, the three channels of a pixel together). No
    tile-cluster test here: the two integrators' pixel estimates have tails
    of different weight (light tracing against next-event estimation near
    the light and the walls it grazes), which a normal 3-sigma bound does not
    model; the GPU tests hold each integrator to its own fixture with it."""
    from tests import refcases
    fa, fb = refcases.load(a), refcases.load(b)
    dt = abs(float(fa["tot_mean"]) - float(fb["tot_mean"]))
    assert dt < 4 * np.hypot(float(fa["tot_se"]), float(fb["tot_se"])), (fa["tot_mean"], fb["tot_mean"])
    if not per_pixel:
        return
    se = np.sqrt(fa["se"] ** 2 + fb["se"] ** 2)
    bad = np.abs(fa["mean"] - fb["mean"]) / np.maximum(se, 1e-12) > 3.0
    assert bad.mean() < 0.015, bad.mean()


def test_reference_bpt_and_pt_fixtures_agree():
    """The reference's BPTRenderer and PTRenderer estimate the same images:
    the BPT fixtures (sensor sums plus the per-thread light-to-lens buffers
    the harness adds, as ImageSensor::saveImage does) against the PT fixtures
    of the same scenes. This pins the BPT fixtures' composition. Per pixel on
    the Lambert and Oren-Nayar scenes; on the GGX glass and metal spheres
    only the image totals: there the path tracer's pixel means at 64 spp are
    heavy-tailed (caustics through the sphere that next-event estimation
    cannot sample), so their batch means are not normal and the per-pixel
    3-sigma comparison against BPT does not hold for the reference either."""
    _fixture_pair("bpt_cornell32", "cornell32")
    _fixture_pair("bpt_on", "pt_on")
    for m in ("glass", "metal"):
        _fixture_pair("bpt_" + m, "pt_" + m, per_pixel=False)
