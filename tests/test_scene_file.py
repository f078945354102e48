"""Scene-language reader (slrh_scene_load_text, libSLRSceneGraph's language
and builtins, libSLRSceneGraph/API.cpp:295-1325): the language's rules on
small texts of this suite's own, and the reference's
TestScenes/Original_Cornell_Box.txt (read from the reference checkout when it
is present) against values the file and the reference's assembly rules fix."""
import math
import os
import re

import numpy as np
import pytest

import slr_amd

REF_SCENE = "/root/reference/TestScenes/Original_Cornell_Box.txt"

CAMERA = '''
cam = createNode();
addChild(cam, createPerspectiveCamera("fovY": 0.5));
addChild(root, cam);
'''
MATTE = 'm = createSurfaceMaterial("matte", (SpectrumTexture(Spectrum("Reflectance", 0.5)),));\n'


def load(host_ctx, text):
    sc = slr_amd.Scene(host_ctx)
    info = sc.load_text(text)
    return sc, info


def one_triangle(p0, p1, p2):
    return "addChild(root, createMesh(((%s, (0,0,0), (0,0,0), (0,0)), (%s, (0,0,0), (0,0,0), (0,0)), " \
           "(%s, (0,0,0), (0,0,0), (0,0))), ((m, ((0, 1, 2),)),)));\n" % (p0, p1, p2)


def f32(x):
    return np.float32(x)


def test_arithmetic_and_literal_types(host_ctx):
    """INTEGER and REALNUMBER tokens (SceneLexer.l:29-30), integer arithmetic
    staying integral, mixed arithmetic in double, unary minus binding to its
    term (SceneParser.yy Term rules)."""
    text = MATTE + one_triangle("(1 + 2 * 3, 7 / 2, 10 % 4)", "(7.0 / 2, -2 * 3, 548.8 - 1)", "(.5, 1e1, 2.5e-1)") + CAMERA
    sc, _ = load(host_ctx, text)
    p = sc.geometry()["pos"]
    assert p[0].tolist() == [7.0, 3.0, 2.0]
    assert p[1].tolist() == [3.5, -6.0, float(f32(548.8 - 1))]
    assert p[2].tolist() == [0.5, 10.0, 0.25]
    sc.close()


def test_named_arguments_defaults_and_tuples(host_ctx):
    """Named arguments by key, the rest positional, defaults for what is
    left (mapParamsToArgs); a one-element tuple needs its trailing comma."""
    text = MATTE + one_triangle("(0,0,0)", "(1,0,0)", "(0,1,0)") + '''
cam = createNode();
addChild(cam, createPerspectiveCamera("objDist": 3.0, 2.0, "fovY": 0.25));
addChild(root, cam);
setRenderSettings("height": 48, 64);
setRenderer("method": "PT", ("samples": 12,));
'''
    sc, info = load(host_ctx, text)
    _, cam = sc.camera()
    # the unnamed 2.0 goes to the first parameter not named: sensitivity
    assert cam["sensitivity"] == 2.0 and cam["aspect"] == 1.0
    assert cam["fov_y"] == 0.25 and cam["obj_dist"] == 3.0
    assert cam["img_dist"] == float(f32(0.02)) and cam["lens_radius"] == 0.0
    assert info["sensitivity"] == 2.0
    # numThreads takes the first unnamed integer, width keeps its default
    assert (info["width"], info["height"], info["spp"]) == (1024, 48, 12)
    sc.close()
    sc, info = load(host_ctx, MATTE + one_triangle("(0,0,0)", "(1,0,0)", "(0,1,0)") + CAMERA)
    assert info["sensitivity"] == math.inf   # a pinhole without a sensitivity: 1 / (pi r^2)
    assert (info["width"], info["height"], info["spp"]) == (1024, 1024, 8)
    sc.close()
    # the renderer methods map to the integrators (API.cpp:1222-1254):
    # PT 0, Volumetric PT 1, BPT 2; samples defaults to 8
    for method, integ in (("PT", 0), ("Volumetric PT", 1), ("BPT", 2)):
        sc, info = load(host_ctx, MATTE + one_triangle("(0,0,0)", "(1,0,0)", "(0,1,0)") + CAMERA +
                        'setRenderer("method": "%s");\n' % method)
        assert (info["integrator"], info["spp"]) == (integ, 8)
        sc.close()


def test_spectrum_forms(host_ctx):
    """The five signatures of Spectrum(): (type, value) and (type, space,
    e0, e1, e2) are RGB, (min, max, values) regular, (wls, values)
    irregular, (type, name, idx) the library; overloads are tried in order."""
    text = '''
s0 = Spectrum(400, 700, (0.1, 0.4, 0.7));
s1 = Spectrum((400, 410, 430), (0.1, 0.3, 0.9));
s2 = Spectrum("Illuminant", "E");
s3 = Spectrum("Illuminant", "D65", 0);
m = createSurfaceMaterial("matte", (SpectrumTexture(s0), FloatTexture(0.2)));
m1 = createSurfaceMaterial("matte", (SpectrumTexture(s1),));
e = createSurfaceMaterial("emitter", (m1, createEmitterSurfaceProperty("diffuse", (SpectrumTexture(s3),))));
e2 = createSurfaceMaterial("emitter", (m, createEmitterSurfaceProperty("diffuse", (SpectrumTexture(s2),))));
''' + one_triangle("(0,0,0)", "(1,0,0)", "(0,1,0)") + CAMERA
    sc, info = load(host_ctx, text)
    assert info["num_spectra"] == 4 and info["num_materials"] == 4
    np.testing.assert_allclose(sc.spectrum_evaluate(0, [400, 475, 550, 700]), [0.1, 0.25, 0.4, 0.7], rtol=1e-6)
    # irregular knots 400, 410, 430: tabulated at their common 10 nm spacing,
    # the same piecewise-linear function
    np.testing.assert_allclose(sc.spectrum_evaluate(1, [400, 405, 420, 430, 500]), [0.1, 0.2, 0.6, 0.9, 0.9],
                               rtol=1e-6)
    np.testing.assert_allclose(sc.spectrum_evaluate(2, [360, 830]), [77.0227, 77.0227], rtol=1e-6)
    # CIE D65 (spectrum_library.cpp): 0.0341 at 300 nm, 100 at 560 nm
    np.testing.assert_allclose(sc.spectrum_evaluate(3, [560.0]), [100.0], rtol=1e-6)
    mi = sc.material_info(0)
    assert mi["model"] == 1   # Oren-Nayar: a sigma texture was given
    assert sc.material_info(1)["model"] == 0   # Lambert: no sigma
    assert sc.material_info(2)["emittance"] == 3 and sc.material_info(3)["emittance"] == 2
    sc.close()


def test_node_transforms_compose_parent_first(host_ctx):
    """A node's transform is its parent's times its own (ChainedTransform::
    reduce); vertices go through Matrix4x4 * Point3D, the camera takes its
    node's transform."""
    text = MATTE + '''
n = createNode();
setTransform(root, scale(2));
setTransform(n, translate(1, 0, 0) * scale(1, 3, 1));
addChild(root, n);
addChild(n, createMesh((((0,0,0), (0,0,0), (0,0,0), (0,0)), ((1,1,0), (0,0,0), (0,0,0), (0,0)),
                        ((0,1,1), (0,0,0), (0,0,0), (0,0))), ((m, ((0, 1, 2),)),)));
cam = createNode();
setTransform(cam, rotateY(0.5));
addChild(cam, createPerspectiveCamera());
addChild(root, cam);
'''
    sc, _ = load(host_ctx, text)
    g = sc.geometry()
    assert g["pos"].tolist() == [[2, 0, 0], [4, 6, 0], [2, 6, 2]]
    # face normal of the object-space triangle, through the inverse transpose
    e1, e2 = np.float32([1, 1, 0]), np.float32([0, 1, 1])
    n = np.cross(e1, e2) / np.float32(np.linalg.norm(np.cross(e1, e2)))
    w = n * np.float32([1 / 2, 1 / 6, 1 / 2])
    np.testing.assert_allclose(g["nrm"][0], w / np.linalg.norm(w), atol=1e-6)
    l2w, _ = sc.camera()
    c, s = np.float32(math.cos(0.5)), np.float32(math.sin(0.5))
    np.testing.assert_allclose(l2w[:3, :3], 2 * np.float32([[c, 0, s], [0, 1, 0], [-s, 0, c]]), atol=1e-6)
    sc.close()


@pytest.mark.parametrize("text,message", [
    ("for (i = 0; i < 3; ++i) {}", "line 1: 'for' statements are outside the supported subset"),
    ("\n\nx = load3DModel(\"a.obj\");", "line 3: 'load3DModel' is outside the supported subset"),
    ("m = createSurfaceMaterial(\"metal\", (,));", "material type 'metal' is outside the supported subset"),
    ("setRenderer(\"method\": \"Volumetric BPT\");", "renderer 'Volumetric BPT' is outside the supported subset"),
    ("x = Spectrum(\"Reflectance\");", "the arguments do not match Spectrum()"),
    ("x = (1, 2;", "expected ')'"),
    (MATTE + "addChild(root, createMesh((((1,2,3), (0,0,0), (0,0,0), (0,0)),), ((m, ((0, 1, 2),)),)));",
     "triangle vertex index out of range"),
    (MATTE + "addChild(root, createMesh(((1,2,3),), ((m, ((0, 1, 2),)),)));", "the arguments do not match createVertex()"),
    (MATTE + one_triangle("(0,0,0)", "(1,0,0)", "(0,1,0)"), "the scene has no camera"),
])
def test_errors_name_the_construct_and_line(host_ctx, text, message):
    sc = slr_amd.Scene(host_ctx)
    with pytest.raises(slr_amd.SlrhError) as e:
        sc.load_text(text)
    assert message in str(e.value)
    sc.close()


def _file_vertices(text):
    """The vertex positions of the mesh in the scene file, each coordinate
    evaluated as the language does (reals in double)."""
    body = text[text.index("mesh = createMesh("):]
    out = []
    for m in re.finditer(r"\(\(\s*([^()]*?)\)\s*,\s*\([^()]*\)\s*,\s*\([^()]*\)\s*,\s*\([^()]*\)\)", body):
        out.append([eval(x) for x in m.group(1).split(",")])
    return np.array(out, np.float64)


@pytest.mark.skipif(not os.path.exists(REF_SCENE), reason="reference checkout not present")
def test_original_cornell_box(host_ctx):
    """TestScenes/Original_Cornell_Box.txt: settings, world-space geometry
    (root's scale(0.01) applied in FP32), the normals the reference
    recomputes, materials, spectra and camera."""
    text = open(REF_SCENE).read()
    sc, info = load(host_ctx, text)
    assert (info["width"], info["height"], info["spp"], info["integrator"]) == (1024, 1024, 16384, 0)
    assert info["brightness"] == 10.0 and info["sensitivity"] == math.inf
    assert (info["num_meshes"], info["num_triangles"], info["num_materials"]) == (1, 32, 5)
    g = sc.geometry()
    verts = _file_vertices(text)
    assert verts.shape == (64, 3)
    # Matrix4x4 * Point3D with scale(0.01): 0.01f * x plus zero terms, in FP32
    expect = np.float32(0.01) * verts.astype(np.float32)
    assert np.array_equal(g["pos"], expect)
    # normals (recalculateNormalAndTangents, FP32 as the reference computes
    # them): per vertex the sum of its faces' unit normals (Vector3D
    # normalize: times the float reciprocal of the length), normalised with
    # Normal3D::normalize (a double reciprocal rounded to float); then the
    # transform's inverse transpose (x 100) and normalize() again
    obj = verts.astype(np.float32)
    acc = np.zeros_like(obj)
    for tri in g["idx"]:
        p = obj[tri]
        e1, e2 = p[1] - p[0], p[2] - p[0]
        c = np.float32([e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]])
        c = c * (np.float32(1) / np.sqrt(np.float32(c[0] * c[0] + c[1] * c[1] + c[2] * c[2])))
        for v in tri:
            acc[v] = acc[v] + c
    for v in range(len(acc)):
        n = acc[v]
        n = n * np.float32(1.0 / float(np.sqrt(np.float32(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]))))
        w = np.float32(100) * n
        w = w * (np.float32(1) / np.sqrt(np.float32(w[0] * w[0] + w[1] * w[1] + w[2] * w[2])))
        assert np.array_equal(g["nrm"][v], w), v
    # tangent: makeCoordinateSystem of the normal, orthogonal unit vectors
    assert np.allclose(np.einsum("ij,ij->i", g["nrm"], g["tan"]), 0, atol=1e-6)
    assert np.allclose(np.linalg.norm(g["tan"], axis=1), 1, atol=1e-6)
    # material groups: floor, light, ceiling, back wall, green, red, two blocks
    mats = g["mat"]
    counts = {int(m): int((mats == m).sum()) for m in np.unique(mats)}
    assert counts == {0: 2 + 2 + 2 + 10 + 10, 1: 2, 2: 2, 4: 2}
    for m in range(4):
        assert sc.material_info(m)["model"] == 1   # "matte" with FloatTexture(0.0): Oren-Nayar, sigma 0
        assert sc.material_info(m)["p0"] == 1.0 and sc.material_info(m)["p1"] == 0.0
    light = sc.material_info(4)
    assert light["emittance"] >= 0
    np.testing.assert_allclose(sc.spectrum_evaluate(light["emittance"], [560.0]), [100.0])   # D65
    white = sc.material_info(0)["spec0"]
    np.testing.assert_allclose(sc.spectrum_evaluate(white, [400.0, 404.0, 700.0]), [0.343, 0.445, 0.737], rtol=1e-6)
    l2w, cam = sc.camera()
    assert np.array_equal(l2w[:3, 3], [f32(0.01) * f32(278), f32(0.01) * f32(273), f32(0.01) * f32(-800)])
    assert np.array_equal(np.diag(l2w)[:3], [f32(0.01)] * 3)
    assert cam["fov_y"] == float(f32(0.6860478808)) and cam["obj_dist"] == 4.5 and cam["img_dist"] == 1.0
    sc.close()


def test_scene_text_writer_round_trip(host_ctx):
    """slr_amd.scenes.to_scene_text writes a description the reader turns
    back into the same world: positions within an FP32 ulp of the
    description's (the file's units times scale(0.01) in FP32), the normals
    the reference recomputes equal to the flat faces' normals, the camera at
    the same place."""
    from slr_amd import scenes
    desc = scenes.cornell_box(32, 32)
    sc, info = load(host_ctx, scenes.to_scene_text(desc, 32, 32, 64, method="BPT"))
    assert (info["width"], info["height"], info["spp"], info["integrator"]) == (32, 32, 64, 2)
    g = sc.geometry()
    pos = np.concatenate([m["pos"] for m in desc["meshes"]])
    nrm = np.concatenate([m["nrm"] for m in desc["meshes"]])
    assert g["pos"].shape == pos.shape
    np.testing.assert_allclose(g["pos"], pos, rtol=0, atol=4e-7)
    np.testing.assert_allclose(g["nrm"], nrm, rtol=0, atol=2e-7)
    l2w, cam = sc.camera()
    np.testing.assert_allclose(l2w[:3, 3], np.asarray(desc["camera"]["l2w"])[:3, 3], rtol=1e-6)
    assert cam["fov_y"] == pytest.approx(desc["camera"]["fov_y"], rel=1e-6)
    sc.close()
