"""Reference images the statistical parity tests compare against.

Each case is a scene, an image size and a number of batches of samples per
pixel that the compiled reference renderer (oracle/_ref) rendered once;
tests/golden/make_golden.py stores per pixel and CIE XYZ channel the mean of
the batch means and its standard error, and the per-batch image totals, in
tests/golden/ref_<case>.npz. The tests render the same scene on the GPU in
batches and compare within 3 combined standard errors (check_3sigma), so
they are deterministic and run without the compiled reference."""
import os

import numpy as np

from slr_amd import scenes

HERE = os.path.dirname(os.path.abspath(__file__))


def field(W, H, per_instance):
    """The small instanced field of the instancing tests."""
    return scenes.instanced_field(W, H, n_instances=16, seed=5, subdiv=3, terrain_n=20, knot=(40, 8),
                                  env_size=(64, 32), sun_radiance=200.0, materials_per_instance=per_instance)


# name -> (reference scene, W, H, batches, spp per batch, first seed, volumetric integrator)
CASES = {
    "cornell32": (lambda: scenes.cornell_box(32, 32), 32, 32, 32, 64, 1000, False),
    "sky": (lambda: scenes.sky_scene(32, 24), 32, 24, 32, 32, 2000, False),
    "instanced": (lambda: dict(scenes.instanced_cornell(32, 32, n_instances=10, seed=12, uniform_scale=True,
                                                        instanced_light=True), compat_instance_shadow=True),
                  32, 32, 32, 32, 4000, False),
    "field": (lambda: dict(field(32, 32, False), compat_instance_shadow=True), 32, 32, 32, 16, 7000, False),
    "medium": (lambda: dict(scenes.cornell_medium(32, 32, n=32, sigma_t=1.0, albedo=0.9, g=0.7),
                            compat_volpt_tangent_nee=True), 32, 32, 32, 16, 6000, True),
    "volpt_cornell16": (lambda: dict(scenes.cornell_box(16, 16), compat_volpt_tangent_nee=True), 16, 16, 24, 32, 700,
                        True),
}


def xyz(strata, cmf):
    x, y, z, integ = cmf
    return np.stack([strata @ x, strata @ y, strata @ z], -1) / integ


def load(name):
    return dict(np.load(os.path.join(HERE, "golden", "ref_%s.npz" % name)))


def check_3sigma(batches, ref, max_fail=0.03, max_relmse=0.05):
    """batches: (B, H, W, 3) XYZ per-batch means of the image under test.
    Per pixel and channel within 3 combined standard errors for all but
    max_fail of them, relMSE against the reference below max_relmse, and the
    image total within 4 combined standard errors of the batch totals.
    Returns (fraction outside 3 sigma, relMSE)."""
    B = len(batches)
    mean = batches.mean(0)
    se = np.sqrt(batches.var(0, ddof=1) / B + ref["se"] ** 2)
    z = np.abs(mean - ref["mean"]) / np.maximum(se, 1e-12)
    fail = float(np.mean(z > 3.0))
    relmse = float(np.mean((mean - ref["mean"]) ** 2 / (ref["mean"] ** 2 + 1e-3)))
    assert fail < max_fail, (fail, relmse)
    tot = batches.sum(axis=(1, 2, 3))
    se_t = np.sqrt(tot.var(ddof=1) / B + float(ref["tot_se"]) ** 2)
    assert abs(tot.mean() - float(ref["tot_mean"])) < 4 * se_t, (tot.mean(), float(ref["tot_mean"]), se_t)
    assert relmse < max_relmse, relmse
    return fail, relmse
