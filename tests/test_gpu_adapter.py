"""The reference-side adapter (include/slr_hip_renderer.hpp): a libSLR host
program (oracle/adapter_check.cpp, linked against the compiled reference and
libslr_hip.so by oracle/ref.mk) mirrors a scene into the adapter with the
calls a scene-graph binding makes and renders it through SLR::Renderer::
render(scene, settings). Its film must equal, bit for bit, the same scene
rendered through the Python binding with the same seed (PT), and within the
atomics' rounding (BPT)."""
import os
import struct
import subprocess
import tempfile

import numpy as np
import pytest

import slr_amd
from slr_amd import scenes

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "oracle", "_ref", "adapter_check")


def _dump(desc, path, W, H, spp, seed):
    out = [struct.pack("<5I", W, H, spp, seed, len(desc["materials"]))]
    for m in desc["materials"]:
        kind = {"lambert": 0, "emitter": 1}[m["kind"]]
        out.append(struct.pack("<2I3f", kind, m.get("base", 0), *[float(c) for c in m["rgb"]]))
    out.append(struct.pack("<I", len(desc["meshes"])))
    for mesh in desc["meshes"]:
        pos = np.asarray(mesh["pos"], np.float32)
        v = np.concatenate([pos, np.asarray(mesh["nrm"], np.float32), np.asarray(mesh["tan"], np.float32)], 1)
        idx = np.asarray(mesh["idx"], np.uint32)
        mat = np.full(len(idx), mesh["mat"], np.uint32) if np.isscalar(mesh["mat"]) else np.asarray(mesh["mat"],
                                                                                                    np.uint32)
        out += [struct.pack("<2I", len(pos), len(idx)), v.astype("<f4").tobytes(), idx.astype("<u4").tobytes(),
                mat.astype("<u4").tobytes()]
    cam = desc["camera"]
    l2w = np.asarray(cam["l2w"], np.float32).reshape(4, 4).T.ravel()   # column-major
    out.append(l2w.astype("<f4").tobytes())
    out.append(struct.pack("<6f", cam.get("sensitivity", 1.0), cam["aspect"], cam["fov_y"], cam.get("lens_radius", 0.0),
                           cam.get("img_dist", 1.0), cam.get("obj_dist", 1.0)))
    with open(path, "wb") as f:
        f.write(b"".join(out))


@pytest.mark.skipif(not os.path.exists(EXE), reason="adapter program not built (make -f oracle/ref.mk)")
@pytest.mark.parametrize("integrator", [0, 2])
def test_adapter_renders_like_the_binding(gpu_ctx, integrator):
    W, H, spp, seed = 24, 24, 16, 1234
    desc = scenes.cornell_box(W, H)
    with tempfile.TemporaryDirectory() as d:
        scene_path, film_path = os.path.join(d, "scene.bin"), os.path.join(d, "film.bin")
        _dump(desc, scene_path, W, H, spp, seed)
        print("adapter_check integrator %d" % integrator, flush=True)
        r = subprocess.run([EXE, scene_path, film_path, str(integrator)], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        got = np.fromfile(film_path, np.float64).reshape(H, W, 16)
    sc = slr_amd.build_scene(gpu_ctx, desc)
    film = slr_amd.Film(gpu_ctx, W, H)
    R = slr_amd.BPTRenderer if integrator == 2 else slr_amd.PTRenderer
    R(spp).render(sc, film, seed=seed)
    ref = film.strata()
    assert ref.sum() > 0
    if integrator == 0:
        assert np.array_equal(got, ref)
    else:
        np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())
