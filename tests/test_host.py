"""Host side of the product against the compiled reference: accelerator
topology, spectral upsampling, CMF strata, the C ABI surface and its error
behaviour. No GPU needed (host-only context, device -1)."""
import ctypes
import os
import re

import numpy as np
import pytest

import slr_amd
from slr_amd import scenes
from tests.conftest import need_ref

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    lib = slr_amd.load_library()
    header = open(os.path.join(ROOT, "include", "slr_hip.h")).read()
    declared = re.findall(r"SLRH_API\s+[^;(]*?\b(slrh_\w+)\s*\(", header)
    assert len(declared) >= 25
    for name in declared:
        assert hasattr(lib, name), name
    assert set(declared) == set(slr_amd.EXPORTS)


@pytest.mark.parametrize("which", ["cornell", "ico3", "ico5"])
def test_accelerator_matches_reference_bit_for_bit(host_ctx, which):
    refdrv = need_ref()
    if which == "cornell":
        desc = scenes.cornell_box(8, 8)
    else:
        desc = scenes.cornell_icosphere(8, 8, subdiv=int(which[-1]))
    sc = slr_amd.build_scene(host_ctx, desc)
    ref = refdrv.RefScene(desc)
    ref.build_accel()
    a, b = sc.accel_stats(), ref.accel_stats()
    for k in b:
        assert a[k] == b[k], (k, a[k], b[k])
    for mine, theirs in ((sc.sbvh_dump(), ref.sbvh_dump()), (sc.qbvh_dump(), ref.qbvh_dump())):
        for x, y in zip(mine, theirs):
            assert x.shape == y.shape
            assert np.array_equal(np.asarray(x).view(np.uint32), np.asarray(y).view(np.uint32))


@pytest.mark.parametrize("rgb,kind", [((0.75, 0.75, 0.75), 0), ((0.63, 0.06, 0.04), 0), ((0.15, 0.48, 0.09), 0),
                                      ((17.0, 12.0, 4.0), 1), ((0.0, 0.0, 0.0), 0), ((1.0, 0.0, 1.0), 0),
                                      ((0.01, 0.9, 0.9), 0), ((5.0, 5.0, 5.0), 1)])
def test_rgb_upsampling_matches_reference(host_ctx, rgb, kind):
    refdrv = need_ref()
    sc = slr_amd.Scene(host_ctx)
    sid = sc.spectrum_rgb(rgb, light=(kind == 1))
    lam = np.random.RandomState(0).uniform(355, 835, 2000).astype(np.float32)
    mine = sc.spectrum_evaluate(sid, lam)
    theirs = refdrv.upsample(kind, rgb, lam)
    np.testing.assert_allclose(mine, theirs, rtol=2e-6, atol=1e-7 * max(1.0, float(np.abs(theirs).max())))


def test_cmf_strata_match_reference():
    refdrv = need_ref()
    mine = slr_amd.cmf_strata()
    theirs = refdrv.cmf_strata()
    for a, b in zip(mine[:3], theirs[:3]):
        assert np.array_equal(a, b)
    assert mine[3] == theirs[3]


def test_cmf_strata_integrate_the_cie_table():
    # closed check: the strata partition the 360-830 nm integral of ybar (~106.86 nm)
    x, y, z, integ = slr_amd.cmf_strata()
    assert abs(integ - y.sum()) < 1e-4
    assert 106.0 < integ < 107.5


def test_errors_are_status_codes(host_ctx):
    sc = slr_amd.Scene(host_ctx)
    with pytest.raises(slr_amd.SlrhError):
        sc.lambert(5)   # no such spectrum
    sid = sc.spectrum_constant(0.5)
    m = sc.lambert(sid)
    with pytest.raises(slr_amd.SlrhError):
        sc.add_mesh(np.zeros((3, 3)), np.array([[0, 1, 3]]), m)   # index out of range
    with pytest.raises(slr_amd.SlrhError):
        sc.add_mesh(np.full((3, 3), np.nan), np.array([[0, 1, 2]]), m)
    with pytest.raises(slr_amd.SlrhError):
        sc.intersect(np.zeros((1, 8), np.float32))   # before build
    sc.add_mesh(np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]]), np.array([[0, 1, 2]]), m)
    sc.build()
    with pytest.raises(slr_amd.SlrhError):
        sc.lambert(sid)   # after build
    with pytest.raises(slr_amd.SlrhError):
        sc.intersect(np.zeros((1, 8), np.float32))   # host-only context cannot trace


def test_smooth_normals_when_omitted(host_ctx):
    sc = slr_amd.Scene(host_ctx)
    m = sc.lambert(sc.spectrum_constant(0.5))
    p, f = scenes.icosphere(2)
    sc.add_mesh(p, f, m)
    sc.build()
    st = sc.accel_stats()
    assert st["qbvh_refs"] >= len(f)


def test_stat_key_lists_match_the_header():
    """The ctypes layer sizes its statistics buffers from these lists, and
    the library fills as many doubles as the header documents: a shorter
    list would let the library write past the buffer."""
    hdr = open(os.path.join(ROOT, "include", "slr_hip.h")).read()
    n_render = int(re.search(r"stats, if not\s*\*?\s*NULL, receives (\d+) doubles", hdr).group(1))
    n_server = int(re.search(r"stats\[(\d+)\]: completed requests", hdr).group(1))
    assert len(slr_amd.STAT_KEYS) == n_render == 27
    assert len(slr_amd.SERVER_STAT_KEYS) == n_server == 8
    assert len(set(slr_amd.STAT_KEYS)) == len(slr_amd.STAT_KEYS)


def _triangle_soup(n, seed, duplicates=False):
    """Random triangles in one mesh (duplicated triangles make centroid
    ranges collapse, which takes the builder's median split)."""
    rng = np.random.RandomState(seed)
    c = rng.uniform(-1, 1, size=(n, 3))
    pos = (c[:, None, :] + rng.normal(scale=0.05, size=(n, 3, 3))).reshape(-1, 3)
    if duplicates:
        pos[3 * (n // 2):] = pos[:3 * (n - n // 2)]
    idx = np.arange(3 * n, dtype=np.uint32).reshape(-1, 3)
    nrm = np.tile([0.0, 0.0, 1.0], (3 * n, 1))
    return {"materials": [{"kind": "lambert", "rgb": (0.5, 0.5, 0.5)}],
            "meshes": [{"pos": pos.astype(np.float32), "nrm": nrm.astype(np.float32),
                        "tan": np.tile([1.0, 0.0, 0.0], (3 * n, 1)).astype(np.float32), "idx": idx, "mat": 0}],
            "camera": scenes.cornell_box(8, 8)["camera"]}


@pytest.mark.parametrize("which", ["cornell", "icosphere", "soup", "soup_duplicates"])
def test_standard_bvh_matches_reference_bit_for_bit(host_ctx, which):
    """The binned-SAH binary BVH (StandardBVH(objs, BinnedSAH),
    libSLR/Accelerator/StandardBVH.h:62-193) node for node: boxes, children,
    leaf ranges, split axes, leaf order, depth and SAH cost."""
    refdrv = need_ref()
    desc = {"cornell": lambda: scenes.cornell_box(8, 8), "icosphere": lambda: scenes.cornell_icosphere(8, 8, subdiv=4),
            "soup": lambda: _triangle_soup(3000, 1), "soup_duplicates": lambda: _triangle_soup(2000, 2, True)}[which]()
    desc = dict(desc, standard_bvh=True)
    sc = slr_amd.build_scene(host_ctx, desc)
    boxes, info, refs, depth, cost = sc.standard_bvh_dump()
    ref = refdrv.RefScene(desc)
    ref.build_accel()
    rb, ri, rr, rdepth, rcost = ref.standard_bvh_dump()
    ref.close()
    assert boxes.shape == rb.shape and len(refs) == len(rr)
    assert np.array_equal(boxes.view(np.uint32), rb.view(np.uint32))
    assert np.array_equal(info, ri)
    assert np.array_equal(refs, rr)
    assert depth == rdepth and np.float32(cost) == np.float32(rcost)


def test_context_calls_from_many_threads(host_ctx):
    """Scene assembly from several threads on one context serialises on the
    context's lock, and each thread reads its own last error."""
    import ctypes
    import threading
    lib = host_ctx.lib
    sc = slr_amd.Scene(host_ctx)
    errors = []

    def worker(k):
        try:
            for i in range(200):
                sid = ctypes.c_int()
                assert lib.slrh_spectrum_rgb(sc.h, 0, 0.1 * k, 0.2, 0.3, ctypes.byref(sid)) == 0
                mid = ctypes.c_int()
                assert lib.slrh_material_lambert(sc.h, sid.value, ctypes.byref(mid)) == 0
                # a failure of this thread's own kind, then its own message
                if k % 2:
                    assert lib.slrh_material_lambert(sc.h, -5, ctypes.byref(mid)) < 0
                    assert lib.slrh_last_error(host_ctx.h) == b"lambert: bad spectrum"
                else:
                    assert lib.slrh_spectrum_rgb(sc.h, 9, 0.1, 0.2, 0.3, ctypes.byref(sid)) < 0
                    assert lib.slrh_last_error(host_ctx.h) == b"slrh_spectrum_rgb: bad type"
        except Exception as e:   # reported on the main thread
            errors.append(repr(e))

    threads = [threading.Thread(target=worker, args=(k,)) for k in range(8)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert not errors, errors[:3]
    sc.close()


def test_reference_side_adapter_links():
    """include/slr_hip_renderer.hpp compiles against the reference's headers
    and the adapter program links libSLR and libslr_hip.so (oracle/ref.mk);
    without arguments it prints its usage and exits 2 (no device touched)."""
    import subprocess
    exe = os.path.join(ROOT, "oracle", "_ref", "adapter_check")
    if not os.path.exists(exe):
        pytest.skip("adapter program not built (needs the reference sources: make -f oracle/ref.mk)")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "usage" in r.stderr
