"""Reference images the statistical parity tests compare against.

Each case is a scene, an image size and a number of batches of samples per
pixel that the compiled reference renderer (oracle/_ref) rendered once;
tests/golden/make_golden.py stores per pixel and CIE XYZ channel the mean of
the batch means and its standard error, and the per-batch image totals, in
tests/golden/ref_<case>.npz. The tests render the same scene on the GPU in
batches and compare within 3 combined standard errors (check_3sigma), so
they are deterministic and run without the compiled reference."""
import json
import os

import numpy as np

from slr_amd import scenes

HERE = os.path.dirname(os.path.abspath(__file__))


def field(W, H, per_instance):
    """The small instanced field of the instancing tests."""
    return scenes.instanced_field(W, H, n_instances=16, seed=5, subdiv=3, terrain_n=20, knot=(40, 8),
                                  env_size=(64, 32), sun_radiance=200.0, materials_per_instance=per_instance)


# name -> (reference scene, W, H, batches, spp per batch, first seed, integrator: False PTRenderer,
# True VolumetricPTRenderer, "bpt" BPTRenderer)
CASES = {
    "cornell32": (lambda: scenes.cornell_box(32, 32), 32, 32, 32, 64, 1000, False),
    "sky": (lambda: scenes.sky_scene(32, 24), 32, 24, 32, 32, 2000, False),
    "instanced": (lambda: dict(scenes.instanced_cornell(32, 32, n_instances=10, seed=12, uniform_scale=True,
                                                        instanced_light=True), compat_instance_shadow=True),
                  32, 32, 32, 32, 4000, False),
    "field": (lambda: dict(field(32, 32, False), compat_instance_shadow=True), 32, 32, 32, 16, 7000, False),
    "medium": (lambda: dict(scenes.cornell_medium(32, 32, n=32, sigma_t=1.0, albedo=0.9, g=0.7),
                            compat_volpt_tangent_nee=True), 32, 32, 32, 16, 6000, True),
    # the volumetric configuration's regime: density max(0, fBm5) * sigma_t
    # with sigma_t 20 (optical depth about 9 across the box) on a 64^3 lattice
    "medium_thick": (lambda: dict(scenes.cornell_medium(32, 32, n=64, sigma_t=20.0, albedo=0.9, g=0.7),
                                  compat_volpt_tangent_nee=True), 32, 32, 32, 16, 6500, True),
    # the C2 scene (Cornell box and the 1.31M-triangle displaced icosphere,
    # GGX conductor) at 128x72, the 1920x1080 view's aspect
    "c2": (lambda: scenes.cornell_icosphere(128, 72, subdiv=8), 128, 72, 32, 64, 9000, False),
    "volpt_cornell16": (lambda: dict(scenes.cornell_box(16, 16), compat_volpt_tangent_nee=True), 16, 16, 24, 32, 700,
                        True),
    # the bidirectional integrator (BPTRenderer): the Cornell box, and a
    # displaced sphere with smooth normals of each non-Lambert model, each
    # also rendered by PTRenderer so the two reference integrators can be
    # checked against each other
    "bpt_cornell32": (lambda: scenes.cornell_box(32, 32), 32, 32, 32, 64, 11000, "bpt"),
}
SPHERE_MATERIALS = {
    "glass": {"kind": "ggx_dielectric", "eta_ext": 1.0, "eta_int": 1.5, "alpha": 0.1},
    "metal": {"kind": "ggx_conductor", "eta": 0.2, "k": 3.9, "alpha": 0.08, "alpha_y": 0.4},
    "on": {"kind": "oren_nayar", "rgb": (0.7, 0.5, 0.3), "sigma": 0.3},
}
for _k, (_m, _seed) in {"glass": (SPHERE_MATERIALS["glass"], 12000), "metal": (SPHERE_MATERIALS["metal"], 13000),
                        "on": (SPHERE_MATERIALS["on"], 14000)}.items():
    CASES["bpt_" + _k] = ((lambda m=_m: scenes.cornell_sphere(32, 32, m)), 32, 32, 32, 64, _seed, "bpt")
    CASES["pt_" + _k] = ((lambda m=_m: scenes.cornell_sphere(32, 32, m)), 32, 32, 32, 64, _seed + 500, False)


def xyz(strata, cmf):
    x, y, z, integ = cmf
    return np.stack([strata @ x, strata @ y, strata @ z], -1) / integ


def load(name):
    return dict(np.load(os.path.join(HERE, "golden", "ref_%s.npz" % name)))


def check_3sigma(batches, ref, max_fail=0.01, max_relmse=0.05, tile=8, name=None):
    """batches: (B, H, W, 3) XYZ per-batch means of the image under test.
    Per pixel and channel within 3 combined standard errors for all but
    max_fail of them (about 0.3-0.5 % fail by chance with ~31 degrees of
    freedom per side), no spatial cluster of failures (at most max(6, 4 %)
    of the entries of any tile x tile block, where about 1 is expected by
    chance), relMSE against the reference below max_relmse, and the image
    total within 4 combined standard errors of the batch totals. With a name
    the outcome is appended to $SLRH_PARITY_OUT (JSON lines) when set.
    Returns (fraction outside 3 sigma, relMSE, worst tile's failures)."""
    B = len(batches)
    mean = batches.mean(0)
    se = np.sqrt(batches.var(0, ddof=1) / B + ref["se"] ** 2)
    z = np.abs(mean - ref["mean"]) / np.maximum(se, 1e-12)
    bad = z > 3.0
    fail = float(np.mean(bad))
    relmse = float(np.mean((mean - ref["mean"]) ** 2 / (ref["mean"] ** 2 + 1e-3)))
    H, W = bad.shape[:2]
    worst, limit = 0, 0
    for y in range(0, H, tile):
        for x in range(0, W, tile):
            blk = bad[y:y + tile, x:x + tile]
            worst = max(worst, int(blk.sum()))
            limit = max(limit, max(6, int(0.04 * blk.size)))
    tot = batches.sum(axis=(1, 2, 3))
    se_t = np.sqrt(tot.var(ddof=1) / B + float(ref["tot_se"]) ** 2)
    tot_z = abs(tot.mean() - float(ref["tot_mean"])) / se_t
    if name:
        rec = {"case": name, "pixels": int(H * W), "batches": [int(B), int(ref["batches"])],
               "spp_ref": int(ref["spp"]), "fail_frac": round(fail, 5), "relmse": relmse, "worst_tile": worst,
               "total_z": round(float(tot_z), 3)}
        print("parity %s" % json.dumps(rec))
        out = os.environ.get("SLRH_PARITY_OUT")
        if out:
            with open(out, "a") as f:
                f.write(json.dumps(rec) + "\n")
    assert fail <= max_fail, (fail, relmse)
    assert worst <= limit, ("clustered 3-sigma failures", worst, limit)
    assert tot_z < 4, (tot.mean(), float(ref["tot_mean"]), se_t)
    assert relmse < max_relmse, relmse
    return fail, relmse, worst
