"""Multi-frame batches: several frames rendered in one wavefront (one slot
pool, shared kernel launches) give each frame the film it gets alone.

Slots are seeded from the frame's seed and the slot's index within the
frame, and a slot's samples are filed in the same order either way, so the
FP64 films are bit-identical, not merely close. Also checks that replaying
recorded iteration graphs and launching each kernel give the same films.
"""
import copy

import numpy as np
import pytest

import slr_amd
from slr_amd import scenes

pytestmark = pytest.mark.gpu


def _moved(cam, dx):
    c = copy.deepcopy(cam)
    m = np.array(c["l2w"], dtype=np.float64).reshape(4, 4).copy()
    m[0, 3] += dx
    c["l2w"] = m
    return c


def _frames(ctx, desc):
    cams = [desc["camera"], _moved(desc["camera"], 0.3)]
    sizes = [(40, 24), (24, 32)]
    films = [slr_amd.Film(ctx, w, h) for w, h in sizes]
    frames = [dict(film=f, camera=c, seed=s, spp=p) for f, c, s, p in zip(films, cams, (7, 1234567), (8, 12))]
    return frames


def test_batch_matches_frames_alone(gpu_ctx):
    desc = scenes.cornell_box(40, 24)
    sc = slr_amd.build_scene(gpu_ctx, desc)
    r = slr_amd.PTRenderer(8)
    frames = _frames(gpu_ctx, desc)
    st = r.render_batch(sc, frames, slots_per_pixel=2)
    assert st["samples"] == sum(f["film"].width * f["film"].height * f["spp"] for f in frames)
    together = [f["film"].strata().copy() for f in frames]
    for f, img in zip(frames, together):
        f["film"].clear()
        r.render_batch(sc, [f], slots_per_pixel=2)
        alone = f["film"].strata()
        assert np.isfinite(alone).all() and alone.sum() > 0
        assert np.array_equal(alone.view(np.uint64), img.view(np.uint64))


def test_single_frame_batch_matches_render(gpu_ctx):
    """slrh_render is the one-frame batch: same camera, seed and slots."""
    W, H = 32, 20
    desc = scenes.cornell_box(W, H)
    sc = slr_amd.build_scene(gpu_ctx, desc)
    a = slr_amd.Film(gpu_ctx, W, H)
    b = slr_amd.Film(gpu_ctx, W, H)
    slr_amd.PTRenderer(6).render(sc, a, seed=99, slots_per_pixel=3)
    slr_amd.PTRenderer(6).render_batch(sc, [dict(film=b, camera=desc["camera"], seed=99)], slots_per_pixel=3)
    assert np.array_equal(a.strata().view(np.uint64), b.strata().view(np.uint64))


def test_graph_replay_matches_direct_launches(gpu_ctx):
    W, H = 48, 32
    desc = scenes.cornell_box(W, H)
    sc = slr_amd.build_scene(gpu_ctx, desc)
    a = slr_amd.Film(gpu_ctx, W, H)
    b = slr_amd.Film(gpu_ctx, W, H)
    r = slr_amd.PTRenderer(8)
    # several renders in a row reuse the recorded graph of the same pool
    for seed in (3, 4):
        a.clear()
        b.clear()
        r.render(sc, a, seed=seed, slots_per_pixel=2, graph=True)
        r.render(sc, b, seed=seed, slots_per_pixel=2, graph=False)
        assert np.array_equal(a.strata().view(np.uint64), b.strata().view(np.uint64))


def test_batch_rejects_film_size_mismatch(gpu_ctx):
    desc = scenes.cornell_box(16, 16)
    sc = slr_amd.build_scene(gpu_ctx, desc)
    film = slr_amd.Film(gpu_ctx, 16, 16)
    fr = dict(film=film, camera=desc["camera"], seed=1)
    bad = slr_amd._Frame
    with pytest.raises(slr_amd.SlrhError):
        # a film of another size than the frame asks for
        f2 = slr_amd.Film(gpu_ctx, 8, 8)
        frames = [fr, dict(film=f2, camera=desc["camera"], seed=2)]
        arr = (bad * 2)()
        for i, f in enumerate(frames):
            arr[i].width, arr[i].height, arr[i].spp = 16, 16, 1
            arr[i].aspect, arr[i].fov_y, arr[i].img_dist, arr[i].obj_dist = 1.0, 0.7, 1.0, 1.0
            arr[i].film = f["film"].h
        s = slr_amd._Settings(0, 0, 1, 0, 0, 0, 0, 0, 0, 0)
        sc.ctx.check(sc.lib.slrh_render_batch(sc.h, arr, 2, slr_amd.ctypes.byref(s), None), "slrh_render_batch")


def test_every_traversal_variant_gives_the_same_film(gpu_ctx):
    """The variants differ in stack depth, refill, LDS treelet and step
    fusion only: hits are the reference's in every one, so films are equal."""
    W, H = 64, 48
    desc = scenes.cornell_icosphere(W, H, subdiv=4)
    sc = slr_amd.build_scene(gpu_ctx, desc)
    ref = None
    for v in range(len(slr_amd.trace_variants())):
        f = slr_amd.Film(gpu_ctx, W, H)
        st = slr_amd.PTRenderer(4).render(sc, f, seed=3, slots_per_pixel=2, tune_trace=v + 1)
        assert int(st["trace_variant"]) == v
        bits = f.strata().view(np.uint64).copy()
        if ref is None:
            ref = bits
        assert np.array_equal(bits, ref), v


def test_autotuner_picks_a_variant_and_keeps_the_film(gpu_ctx):
    W, H = 1024, 1024      # 2^20 slots at one slot per pixel: large enough to tune
    desc = scenes.cornell_box(W, H)
    sc = slr_amd.build_scene(gpu_ctx, desc)
    a = slr_amd.Film(gpu_ctx, W, H)
    st = slr_amd.PTRenderer(2).render(sc, a, seed=5, slots_per_pixel=1)
    v = int(st["trace_variant"])
    assert 0 <= v < len(slr_amd.trace_variants())
    # the choice is kept for the scene, and an explicit variant gives the same film
    st2 = slr_amd.PTRenderer(2).render(sc, slr_amd.Film(gpu_ctx, W, H), seed=5, slots_per_pixel=1)
    assert int(st2["trace_variant"]) == v
    b = slr_amd.Film(gpu_ctx, W, H)
    slr_amd.PTRenderer(2).render(sc, b, seed=5, slots_per_pixel=1, tune_trace=(v + 1) % len(slr_amd.trace_variants()) + 1)
    assert np.array_equal(a.strata().view(np.uint64), b.strata().view(np.uint64))


def test_batch_rejects_a_shared_film(gpu_ctx):
    """Slots add into their film cells without atomics, so two frames of one
    wavefront may not accumulate into the same film."""
    desc = scenes.cornell_box(16, 16)
    sc = slr_amd.build_scene(gpu_ctx, desc)
    film = slr_amd.Film(gpu_ctx, 16, 16)
    frames = [dict(film=film, camera=desc["camera"], seed=1), dict(film=film, camera=desc["camera"], seed=2)]
    with pytest.raises(slr_amd.SlrhError, match="share a film"):
        slr_amd.PTRenderer(2).render_batch(sc, frames)


@pytest.mark.parametrize("graph", [True, False])
def test_parallel_shading_gives_the_same_film(gpu_ctx, graph):
    """With SLRH_FLAG_PARALLEL_SHADE the conductor, dielectric and miss
    kernels run on forked streams beside the diffuse one; each reads only
    its own queue and slots, so the film is bit-identical to running them
    in order (all four shading classes under a sky)."""
    W, H = 64, 40
    desc = scenes.instanced_field(W, H, n_instances=24, subdiv=3, terrain_n=24, knot=(40, 8),
                                  env_size=(64, 32), sun_radiance=200.0)
    sc = slr_amd.build_scene(gpu_ctx, desc)
    r = slr_amd.PTRenderer(4)
    a, b = slr_amd.Film(gpu_ctx, W, H), slr_amd.Film(gpu_ctx, W, H)
    for seed in (7, 8):   # a second render replays the recorded graph
        a.clear()
        b.clear()
        r.render(sc, a, seed=seed, slots_per_pixel=2, graph=graph, parallel_shade=True)
        r.render(sc, b, seed=seed, slots_per_pixel=2, graph=graph)
        assert np.array_equal(a.strata().view(np.uint64), b.strata().view(np.uint64))
        assert a.strata().sum() > 0


def test_shading_workgroup_size_keeps_the_film(gpu_ctx):
    """The shading kernels' workgroup size (256 or 128 threads, one of the
    autotuner's dimensions) changes which lanes shade which slots, never a
    slot's arithmetic: films are bit-identical (all shading classes)."""
    W, H = 64, 40
    desc = scenes.instanced_field(W, H, n_instances=24, subdiv=3, terrain_n=24, knot=(40, 8),
                                  env_size=(64, 32), sun_radiance=200.0)
    sc = slr_amd.build_scene(gpu_ctx, desc)
    r = slr_amd.PTRenderer(4)
    a, b = slr_amd.Film(gpu_ctx, W, H), slr_amd.Film(gpu_ctx, W, H)
    sa = r.render(sc, a, seed=9, slots_per_pixel=2, tune_trace=1)
    sb = r.render(sc, b, seed=9, slots_per_pixel=2, tune_trace=1, shade_block_128=True)
    assert sa["shade_block"] == 256 and sb["shade_block"] == 128
    assert np.array_equal(a.strata().view(np.uint64), b.strata().view(np.uint64))


def test_autotuner_reports_its_shading_workgroup(gpu_ctx):
    W, H = 1024, 1024
    desc = scenes.cornell_box(W, H)
    sc = slr_amd.build_scene(gpu_ctx, desc)
    st = slr_amd.PTRenderer(1).render(sc, slr_amd.Film(gpu_ctx, W, H), seed=5, slots_per_pixel=1)
    assert st["shade_block"] in (128, 256)
    st2 = slr_amd.PTRenderer(1).render(sc, slr_amd.Film(gpu_ctx, W, H), seed=6, slots_per_pixel=1)
    assert st2["shade_block"] == st["shade_block"]


def test_pool_width_tuning(gpu_ctx):
    """SLRH_FLAG_TUNE_POOL: the first render pilots 8M/16M/32M live paths
    (8 or 16 slots per pixel at 1024^2) and keeps one for the scene; an
    explicit slots-per-pixel setting gives the same film as the tuned one."""
    W, H = 1024, 1024
    desc = scenes.cornell_box(W, H)
    sc = slr_amd.build_scene(gpu_ctx, desc)
    r = slr_amd.PTRenderer(32)
    a = slr_amd.Film(gpu_ctx, W, H)
    st = r.render(sc, a, seed=4, max_depth=4, tune_pool=True)
    k = int(st["slots_per_pixel"])
    assert k in (8, 16)
    assert st["samples"] == W * H * 32
    st2 = r.render(sc, slr_amd.Film(gpu_ctx, W, H), seed=5, max_depth=4, tune_pool=True)
    assert int(st2["slots_per_pixel"]) == k
    b = slr_amd.Film(gpu_ctx, W, H)
    r.render(sc, b, seed=4, max_depth=4, slots_per_pixel=k)
    assert np.array_equal(a.strata().view(np.uint64), b.strata().view(np.uint64))
    # without the flag: about 32M live paths
    st3 = r.render(sc, slr_amd.Film(gpu_ctx, W, H), seed=6, max_depth=4)
    assert int(st3["slots_per_pixel"]) == 16
