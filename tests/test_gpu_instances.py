"""Two-level traversal and shading of instances on the GPU.

* ray batches against the reference's own QBVH over triangles and
  TransformedSurfaceObjects (oracle/_ref): hit triangle and instance ids
  equal except ties within 1e-6 in t, hit t and surface parameters bit-exact
  (the world-to-object rows are the reference's FP32 inverse and the ray is
  transformed in its operation order);
* shadow rays in SLRH_COMPAT_INSTANCE_SHADOW mode equal the reference's
  visibility, which tests the untransformed world ray against the
  object-space prototype (TransformedSurfaceObject::testVisibility,
  libSLR/Core/surface_object.cpp:363-369); by default they are the
  flattened scene's;
* closest hits agree with the same scene flattened into world-space meshes;
* renders agree with the flattened scene (default mode), and with the
  reference PTRenderer within 3 sigma (compat mode), instanced lights and
  per-instance materials included.
"""
import numpy as np
import pytest

import slr_amd
from slr_amd import scenes
from tests import refcases
from tests.conftest import need_ref

pytestmark = pytest.mark.gpu


def _rays(desc, n, seed):
    return scenes.random_rays(desc, n, seed=seed)


@pytest.mark.parametrize("uniform,light", [(True, False), (False, True)])
def test_intersect_matches_reference_qbvh(gpu_ctx, uniform, light):
    refdrv = need_ref()
    desc = scenes.instanced_cornell(16, 16, n_instances=12, seed=4, uniform_scale=uniform, instanced_light=light)
    sc = slr_amd.build_scene(gpu_ctx, desc)
    ref = refdrv.RefScene(desc)
    ref.build_accel()
    rays = _rays(desc, 100000, seed=7)
    t, prim, b0, b1, inst = sc.intersect(rays, with_inst=True)
    rt, rprim, rb0, rb1, rinst = ref.intersect(rays, with_inst=True)
    same = (prim == rprim) & (inst == rinst)
    with np.errstate(invalid="ignore"):
        tie = np.abs(t - rt) <= 1e-6 * np.maximum(1.0, np.abs(rt))
    assert np.all(same | tie), np.nonzero(~(same | tie))[0][:10]
    hit = (rprim >= 0) & same
    assert (inst[hit] >= 0).mean() > 0.2          # a good share of hits are on instances
    assert np.array_equal(t[hit].view(np.uint32), rt[hit].view(np.uint32))
    assert np.array_equal(b0[hit].view(np.uint32), rb0[hit].view(np.uint32))
    assert np.array_equal(b1[hit].view(np.uint32), rb1[hit].view(np.uint32))
    assert np.all(inst[prim < 0] == -1)


def test_compat_shadow_matches_reference_visibility(gpu_ctx):
    refdrv = need_ref()
    desc = scenes.instanced_cornell(16, 16, n_instances=12, seed=4, uniform_scale=False)
    ref = refdrv.RefScene(desc)
    ref.build_accel()
    rays = _rays(desc, 100000, seed=8)
    rt, rprim, _, _ = ref.intersect(rays)
    seg = rays.copy()
    seg[:, 6] = 1e-4
    seg[:, 7] = np.where(rprim < 0, 10.0, rt * 0.999)
    rv = ref.visibility(seg)
    compat = dict(desc, compat_instance_shadow=True)
    v = slr_amd.build_scene(gpu_ctx, compat).occluded(seg)
    assert np.array_equal(v, rv)
    # the default mode is the flattened scene's visibility: the segments end
    # before the nearest hit, so nearly all are unoccluded there, which the
    # reference's instance shadow test does not reproduce
    flat, _, _ = scenes.flatten_instances(desc)
    fv = slr_amd.build_scene(gpu_ctx, flat).occluded(seg)
    dv = slr_amd.build_scene(gpu_ctx, desc).occluded(seg)
    assert np.mean(dv == fv) > 0.999
    assert np.mean(fv) > 0.999 and np.mean(rv) < np.mean(fv)


def test_closest_hits_match_flattened_scene(gpu_ctx):
    desc = scenes.instanced_cornell(16, 16, n_instances=12, seed=6, uniform_scale=False, instanced_light=True)
    flat, maps, first = scenes.flatten_instances(desc)
    rays = _rays(desc, 100000, seed=9)
    t, prim, _, _, inst = slr_amd.build_scene(gpu_ctx, desc).intersect(rays, with_inst=True)
    ft, fprim, _, _ = slr_amd.build_scene(gpu_ctx, flat).intersect(rays)
    mapped = prim.copy()
    for k, inst_desc in enumerate(desc["instances"]):
        sel = inst == k
        mapped[sel] = maps[k][prim[sel] - first[inst_desc["proto"]]]
    agree = (mapped == fprim) & np.isclose(t, ft, rtol=1e-4, atol=1e-6)
    assert np.mean(agree) > 0.999, np.mean(agree)


def _film(ctx, desc, W, H, spp, seed):
    sc = slr_amd.build_scene(ctx, desc)
    f = slr_amd.Film(ctx, W, H)
    slr_amd.PTRenderer(spp).render(sc, f, seed=seed)
    return f.strata().copy()


def test_render_matches_flattened_scene(gpu_ctx):
    """Same seed, same slots: the instanced and the flattened scene draw the
    same random numbers, so the films agree up to rounding in the
    transforms; instanced light, per-instance materials included."""
    W = H = 32
    desc = scenes.instanced_cornell(W, H, n_instances=10, seed=12, uniform_scale=True, instanced_light=True)
    desc["materials"].append({"kind": "oren_nayar", "rgb": (0.2, 0.5, 0.7), "sigma": 0.3})
    desc["instances"][1]["mat"] = len(desc["materials"]) - 1
    desc["instances"][2]["mat"] = 1
    flat, _, _ = scenes.flatten_instances(desc)
    a = _film(gpu_ctx, desc, W, H, 64, seed=21)
    b = _film(gpu_ctx, flat, W, H, 64, seed=21)
    assert np.isfinite(a).all()
    assert abs(a.sum() / b.sum() - 1) < 0.01, a.sum() / b.sum()
    rel = np.abs(a - b) / np.maximum(np.abs(b), 1e-3 * np.abs(b).max())
    assert np.mean(rel < 1e-2) > 0.9, np.mean(rel < 1e-2)


def test_render_matches_reference_renderer_3sigma(gpu_ctx):
    """Reference PTRenderer image (instances through InternalNode +
    ReferenceNode, 32 x 32, 32 batches of 32 spp, tests/golden/ref_instanced.npz)
    vs 32 GPU batches of 32 spp in compat mode."""
    make, W, H, _, spp, _, _ = refcases.CASES["instanced"]
    sc = slr_amd.build_scene(gpu_ctx, make())
    cmf = slr_amd.cmf_strata()
    gb = []
    for b in range(32):
        film = slr_amd.Film(gpu_ctx, W, H)
        slr_amd.PTRenderer(spp).render(sc, film, seed=500 + b)
        gb.append(refcases.xyz(film.strata(), cmf) / spp)
    refcases.check_3sigma(np.array(gb), refcases.load("instanced"), name="instanced")


def _field(W, H, per_instance):
    return refcases.field(W, H, per_instance)


def test_per_instance_materials_equal_per_prototype_materials(gpu_ctx):
    """The instanced field with each instance's material given on the
    instance equals the same scene with one prototype per (mesh, material)
    pair, sample for sample."""
    W = H = 32
    films = []
    for per_instance in (True, False):
        sc = slr_amd.build_scene(gpu_ctx, _field(W, H, per_instance))
        film = slr_amd.Film(gpu_ctx, W, H)
        slr_amd.PTRenderer(8).render(sc, film, seed=3, max_depth=12)
        films.append(film.strata())
    assert np.isfinite(films[0]).all() and films[0].sum() > 0
    assert np.array_equal(films[0], films[1])


def test_instanced_field_matches_reference_renderer_3sigma(gpu_ctx):
    """Instances of four meshes with mixed Lambert, Oren-Nayar, GGX conductor
    and GGX dielectric materials under an equirectangular sky with a sun:
    the reference PTRenderer's image (one prototype per mesh and material,
    32 x 32, 32 batches of 16 spp, tests/golden/ref_field.npz) vs 32 GPU
    batches of 16 spp with the materials on the instances, compat mode."""
    _, W, H, _, spp, _, _ = refcases.CASES["field"]
    sc = slr_amd.build_scene(gpu_ctx, dict(_field(W, H, True), compat_instance_shadow=True))
    cmf = slr_amd.cmf_strata()
    gb = []
    for b in range(32):
        film = slr_amd.Film(gpu_ctx, W, H)
        slr_amd.PTRenderer(spp).render(sc, film, seed=800 + b)
        gb.append(refcases.xyz(film.strata(), cmf) / spp)
    refcases.check_3sigma(np.array(gb), refcases.load("field"), name="field")
