"""Per-material parity of the GPU shading kernels with the oracle
(oracle/slr_oracle.c, whose BSDFs are pinned per query against the
reference's, tests/test_oracle.py): the Cornell box plus a displaced
icosphere with smooth vertex normals in one material, rendered with the same
seed and slot layout on both sides, so they share every random number.

The sphere's shading normals differ from its geometric normals, so the BSDF
front end's shading-normal correction |z / dot(d, ng)|
(libSLR/Core/directional_distribution_functions.h:295-349) is not 1, and
every sample/evaluate/PDF of the material's model runs on the GPU:
Oren-Nayar (BSDF/OrenNayerBRDF.cpp:13-74), Lambert on a curved surface, GGX
conductor isotropic and anisotropic with visible-normal sampling and
FresnelConductor, GGX dielectric with transmission and FresnelDielectric
(BSDF/microfacet_bsdfs.cpp:11-328, Core/directional_distribution_functions.cpp:
118-315).

Films agree to FP32 rounding per pixel stratum except for paths that take a
different branch on a last-ulp difference (libm sin/cos/pow differ between
the GPU and glibc); the fraction of strata within 1e-4 relative is recorded
and must exceed 97 %, the image totals agree within 2e-3."""
import json
import os

import numpy as np
import pytest

import slr_amd
from oracle import oracle
from slr_amd import scenes

pytestmark = pytest.mark.gpu

MATERIALS = {
    "lambert_curved": {"kind": "lambert", "rgb": (0.7, 0.6, 0.5)},
    "oren_nayar": {"kind": "oren_nayar", "rgb": (0.7, 0.5, 0.3), "sigma": 0.3},
    "ggx_conductor": {"kind": "ggx_conductor", "eta": 0.2, "k": 3.9, "alpha": 0.2},
    "ggx_conductor_aniso": {"kind": "ggx_conductor", "eta": 0.2, "k": 3.9, "alpha": 0.08, "alpha_y": 0.4},
    "ggx_dielectric": {"kind": "ggx_dielectric", "eta_ext": 1.0, "eta_int": 1.5, "alpha": 0.1},
    "ggx_dielectric_aniso": {"kind": "ggx_dielectric", "eta_ext": 1.0, "eta_int": 1.5, "alpha": 0.05,
                             "alpha_y": 0.3},
}


@pytest.mark.parametrize("name", sorted(MATERIALS))
def test_material_matches_oracle_same_seed(gpu_ctx, name):
    W, H, spp, K = 32, 32, 16, 2
    desc = scenes.cornell_sphere(W, H, MATERIALS[name])
    sc = slr_amd.build_scene(gpu_ctx, desc)
    film = slr_amd.Film(gpu_ctx, W, H)
    st = slr_amd.PTRenderer(spp).render(sc, film, seed=17, slots_per_pixel=K)
    img = film.strata()
    threads = max(1, min(16, len(os.sched_getaffinity(0))))
    ref, rays = oracle.OracleScene(desc).render(W, H, spp, 17, K=K, threads=threads)
    assert st["samples"] == W * H * spp and np.isfinite(img).all()
    rel = np.abs(img - ref) / np.maximum(np.abs(ref), 1e-6)
    within = float(np.mean(rel < 1e-4))
    tot = abs(img.sum() - ref.sum()) / ref.sum()
    rec = {"case": "same_seed_" + name, "strata_within_1e-4": round(within, 5), "total_rel": float(tot),
           "rays_rel": abs(st["rays"] - rays) / rays}
    print("parity %s" % json.dumps(rec))
    out = os.environ.get("SLRH_PARITY_OUT")
    if out:
        with open(out, "a") as f:
            f.write(json.dumps(rec) + "\n")
    assert within > 0.97, within
    assert tot < 2e-3, tot
    assert abs(st["rays"] - rays) / rays < 2e-3
