"""Instancing on the host: the world-level SBVH and QBVH over world
triangles and instances, built as the reference builds its aggregate over
triangles and TransformedSurfaceObjects (spatial splits clip an instance's
world box, libSLR/Core/object.h:50-80; an instance costs its prototype
tree's SAH cost, SBVH.h:402,419), are node-for-node the reference's; the
world-to-object rows are the reference's FP32 inverse bit for bit; the
C ABI rejects bad prototypes, materials and singular transforms."""
import numpy as np
import pytest

import slr_amd
from slr_amd import scenes
from tests.conftest import need_ref


@pytest.mark.parametrize("uniform,light", [(True, False), (False, False), (False, True)])
def test_world_tree_matches_reference(host_ctx, uniform, light):
    refdrv = need_ref()
    desc = scenes.instanced_cornell(16, 16, n_instances=12, seed=3, uniform_scale=uniform, instanced_light=light)
    sc = slr_amd.build_scene(host_ctx, desc)
    ref = refdrv.RefScene(desc)
    ref.build_accel()
    mine, theirs = sc.accel_stats(), ref.accel_stats()
    for k in theirs:
        assert mine[k] == theirs[k], (k, mine[k], theirs[k])
    for a, b in zip(sc.sbvh_dump(), ref.sbvh_dump()):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    for a, b in zip(sc.qbvh_dump(), ref.qbvh_dump()):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    # the refs name instances (bit 31) as well as world triangles
    refs = sc.qbvh_dump()[2]
    assert (refs < 0).sum() >= 12 + (1 if light else 0)


def test_instance_inverse_matches_reference(host_ctx):
    refdrv = need_ref()
    desc = scenes.instanced_cornell(16, 16, n_instances=24, seed=9, uniform_scale=False)
    sc = slr_amd.build_scene(host_ctx, desc)
    ref = refdrv.RefScene(desc)
    for k in range(24):
        l2w, w2l = sc.instance_transform(k)
        assert np.array_equal(l2w, np.asarray(desc["instances"][k]["l2w"], np.float32))
        # equal values (the reference's matrix is read back through products
        # with unit vectors, which do not keep the sign of a zero)
        assert np.array_equal(w2l + 0.0, ref.instance_inverse(k, full=True) + 0.0), k
        # and it is an inverse
        m = np.vstack([np.asarray(l2w, np.float64), [0, 0, 0, 1]])
        assert np.allclose(m @ np.asarray(w2l, np.float64), np.eye(4), atol=1e-5)


def test_instance_api_errors(host_ctx):
    sc = slr_amd.Scene(host_ctx)
    mat = sc.lambert(sc.spectrum_constant(0.5))
    v, f = scenes.icosphere(1)
    p = sc.add_prototype(v, f, mat)
    good = np.hstack([np.eye(3), np.zeros((3, 1))])
    with pytest.raises(slr_amd.SlrhError):
        sc.add_instance(p + 1, good)
    with pytest.raises(slr_amd.SlrhError):
        sc.add_instance(p, np.zeros((3, 4)))                 # singular
    bad = good.copy()
    bad[0, 0] = np.nan
    with pytest.raises(slr_amd.SlrhError):
        sc.add_instance(p, bad)
    with pytest.raises(slr_amd.SlrhError):
        sc.add_instance(p, good, material=mat + 5)
    assert sc.add_instance(p, good) == 0
    assert sc.add_instance(p, good, material=mat) == 1
    with pytest.raises(slr_amd.SlrhError):
        sc.instance_transform(2)
    with pytest.raises(slr_amd.SlrhError):
        sc.set_compat(4)
    sc.set_compat(slr_amd.COMPAT_INSTANCE_SHADOW)
    sc.set_camera(np.eye(4), 1.0, 0.5)
    sc.build()
    with pytest.raises(slr_amd.SlrhError):
        sc.add_instance(p, good)                              # after build


def test_flattened_scene_places_the_same_triangles():
    desc = scenes.instanced_cornell(16, 16, n_instances=4, seed=2, uniform_scale=False)
    flat, maps, first = scenes.flatten_instances(desc)
    assert len(flat["meshes"]) == len(desc["meshes"]) + 4
    for k, inst in enumerate(desc["instances"]):
        pr = desc["prototypes"][inst["proto"]]
        m = np.asarray(inst["l2w"], np.float64)
        want = np.asarray(pr["pos"], np.float64) @ m[:, :3].T + m[:, 3]
        np.testing.assert_allclose(flat["meshes"][len(desc["meshes"]) + k]["pos"], want, rtol=1e-6, atol=1e-6)
        assert len(maps[k]) == len(pr["idx"])


def test_instanced_field_generators():
    """The instanced-field meshes have their configured sizes, the scene is
    seed-deterministic, and its per-prototype form carries the same
    instances and materials as the per-instance form."""
    p, f = scenes.terrain()
    assert len(f) == 199712 and f.max() < len(p)
    kp, kn, kf = scenes.torus_knot()
    assert len(kf) == 50000 and kf.max() < len(kp)
    assert np.allclose(np.linalg.norm(kn, axis=1), 1.0)
    kw = dict(n_instances=24, seed=9, subdiv=2, terrain_n=10, knot=(30, 6), env_size=(32, 16))
    a = scenes.instanced_field(32, 32, **kw)
    b = scenes.instanced_field(32, 32, materials_per_instance=False, **kw)
    a2 = scenes.instanced_field(32, 32, **kw)
    assert len(a["instances"]) == len(b["instances"]) == 24 and len(a["prototypes"]) == 4
    kinds = {m["kind"] for m in a["materials"]}
    assert kinds <= {"lambert", "ggx_conductor", "ggx_dielectric", "oren_nayar"} and len(kinds) >= 3
    for ia, ib, ia2 in zip(a["instances"], b["instances"], a2["instances"]):
        assert np.array_equal(ia["l2w"], ib["l2w"]) and np.array_equal(ia["l2w"], ia2["l2w"])
        pb = b["prototypes"][ib["proto"]]
        pa = a["prototypes"][ia["proto"]]
        assert pb["mat"] == ia["mat"]
        assert np.array_equal(pa["pos"], pb["pos"]) and np.array_equal(pa["idx"], pb["idx"])


def test_fbm_density_matches_direct_evaluation():
    n = 24
    d = scenes.fbm_density(n, seed=3, octaves=3, chunk_points=5 * n * n)   # 5 z slices per chunk
    g = np.arange(n) / (n - 1)
    z, y, x = np.meshgrid(g, g, g, indexing="ij")
    p = np.stack([x.ravel(), y.ravel(), z.ravel()], -1)
    # the volumetric configuration's density: max(0, fBm)
    want = np.maximum(0.0, scenes.fbm(p, 3, octaves=3, freq=2.0)).reshape(n, n, n)
    assert np.allclose(d, want.astype(np.float32), rtol=0, atol=1e-6)
