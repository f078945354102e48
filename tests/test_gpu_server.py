"""Continuous-batching server: a request's film is the same alone and in a
mix of concurrent requests (bit-identical: slots are seeded per request and
each slot files its samples in the same order), admission respects the
path-state budget, cancellation frees the slots, and latency statistics are
reported.
"""
import copy
import time

import numpy as np
import pytest

import slr_amd
from slr_amd import scenes

pytestmark = pytest.mark.gpu


def _moved(cam, dx, dy=0.0):
    c = copy.deepcopy(cam)
    m = np.array(c["l2w"], dtype=np.float64).reshape(4, 4).copy()
    m[0, 3] += dx
    m[1, 3] += dy
    c["l2w"] = m
    return c


@pytest.fixture(scope="module")
def cornell(gpu_ctx):
    desc = scenes.cornell_box(64, 64)
    return desc, slr_amd.build_scene(gpu_ctx, desc)


def _bits(film):
    return film.strata().copy().view(np.uint64)


def test_request_same_alone_and_in_mix(gpu_ctx, cornell):
    desc, sc = cornell
    srv = slr_amd.Server(sc, path_state_bytes=256 << 20)
    try:
        cam = _moved(desc["camera"], 0.2)
        alone = slr_amd.Film(gpu_ctx, 48, 40)
        t = srv.submit(alone, cam, spp=8, seed=77, max_depth=10, slots_per_pixel=2)
        assert srv.wait(t, 60000) == slr_amd.REQ_DONE
        ref = _bits(alone)
        assert alone.strata().sum() > 0

        mixed = slr_amd.Film(gpu_ctx, 48, 40)
        others = []
        tickets = []
        for i, (w, h, spp) in enumerate([(64, 64, 16), (32, 24, 4), (56, 16, 8)]):
            f = slr_amd.Film(gpu_ctx, w, h)
            others.append(f)
            tickets.append(srv.submit(f, _moved(desc["camera"], -0.1 * i, 0.05 * i), spp=spp, seed=1000 + i,
                                      max_depth=10))
        tm = srv.submit(mixed, cam, spp=8, seed=77, max_depth=10, slots_per_pixel=2)
        for t in tickets + [tm]:
            assert srv.wait(t, 60000) == slr_amd.REQ_DONE
        assert np.array_equal(_bits(mixed), ref)
        for f in others:
            assert np.isfinite(f.strata()).all() and f.strata().sum() > 0
        st = srv.stats()
        assert st["completed"] == 5 and st["queued"] == 0 and st["running"] == 0
        assert 0 < st["p50_ms"] <= st["p99_ms"]
    finally:
        srv.close()

    # the served film is the batch render of the same frame
    b = slr_amd.Film(gpu_ctx, 48, 40)
    slr_amd.PTRenderer(8).render_batch(sc, [dict(film=b, camera=cam, seed=77)], max_depth=10, slots_per_pixel=2)
    assert np.array_equal(_bits(b), ref)


def test_admission_by_path_state_budget(gpu_ctx, cornell):
    """A budget that holds one 64x64 request at a time still completes all
    requests, each equal to its batch render."""
    desc, sc = cornell
    budget = 64 * 64 * 2 * 420   # one request at 2 slots per pixel
    srv = slr_amd.Server(sc, path_state_bytes=budget)
    try:
        films = [slr_amd.Film(gpu_ctx, 64, 64) for _ in range(3)]
        ts = [srv.submit(f, desc["camera"], spp=4, seed=5 + i, slots_per_pixel=2) for i, f in enumerate(films)]
        assert [srv.wait(t, 60000) for t in ts] == [slr_amd.REQ_DONE] * 3
    finally:
        srv.close()
    for i, f in enumerate(films):
        b = slr_amd.Film(gpu_ctx, 64, 64)
        slr_amd.PTRenderer(4).render_batch(sc, [dict(film=b, camera=desc["camera"], seed=5 + i)], slots_per_pixel=2)
        assert np.array_equal(_bits(f), _bits(b))


def test_request_larger_than_budget_is_refused(gpu_ctx, cornell):
    desc, sc = cornell
    srv = slr_amd.Server(sc, path_state_bytes=8192 * 420)
    try:
        f = slr_amd.Film(gpu_ctx, 128, 128)
        with pytest.raises(slr_amd.SlrhError):
            srv.submit(f, desc["camera"], spp=1)
    finally:
        srv.close()


def test_cancel_frees_slots(gpu_ctx, cornell):
    desc, sc = cornell
    srv = slr_amd.Server(sc, path_state_bytes=64 * 64 * 420)
    try:
        big = slr_amd.Film(gpu_ctx, 64, 64)
        tb = srv.submit(big, desc["camera"], spp=1 << 20, seed=1, slots_per_pixel=1)
        small = slr_amd.Film(gpu_ctx, 32, 32)
        ts = srv.submit(small, desc["camera"], spp=2, seed=2, slots_per_pixel=1)
        time.sleep(0.05)
        srv.cancel(tb)
        assert srv.wait(tb, 60000) == slr_amd.REQ_CANCELLED
        assert srv.wait(ts, 60000) == slr_amd.REQ_DONE
    finally:
        srv.close()
    b = slr_amd.Film(gpu_ctx, 32, 32)
    slr_amd.PTRenderer(2).render_batch(sc, [dict(film=b, camera=desc["camera"], seed=2)], slots_per_pixel=1)
    assert np.array_equal(_bits(small), _bits(b))


def test_requests_sharing_a_film_accumulate_in_turn(gpu_ctx, cornell):
    """Two requests into one film: the second waits for the first (their
    slots would race on the film cells), so the film holds the sum of the
    two renders, bit for bit."""
    desc, sc = cornell
    srv = slr_amd.Server(sc, path_state_bytes=256 << 20)
    try:
        shared = slr_amd.Film(gpu_ctx, 64, 64)
        ts = [srv.submit(shared, desc["camera"], spp=4, seed=s, slots_per_pixel=2) for s in (21, 22)]
        assert [srv.wait(t, 60000) for t in ts] == [slr_amd.REQ_DONE] * 2
    finally:
        srv.close()
    a = slr_amd.Film(gpu_ctx, 64, 64)
    slr_amd.PTRenderer(4).render_batch(sc, [dict(film=a, camera=desc["camera"], seed=21)], slots_per_pixel=2)
    slr_amd.PTRenderer(4).render_batch(sc, [dict(film=a, camera=desc["camera"], seed=22)], slots_per_pixel=2)
    assert np.array_equal(_bits(shared), _bits(a))


def test_automatic_slots_per_pixel_do_not_depend_on_load(gpu_ctx, cornell):
    """Without a fixed slots-per-pixel count a request gets one chosen from
    the request alone, so its film is the same on an idle and a busy server."""
    desc, sc = cornell
    cam = _moved(desc["camera"], 0.15)
    srv = slr_amd.Server(sc, path_state_bytes=256 << 20)
    try:
        idle = slr_amd.Film(gpu_ctx, 40, 32)
        assert srv.wait(srv.submit(idle, cam, spp=12, seed=5), 60000) == slr_amd.REQ_DONE
        busy = slr_amd.Film(gpu_ctx, 40, 32)
        load = [slr_amd.Film(gpu_ctx, 64, 64) for _ in range(4)]
        tl = [srv.submit(f, desc["camera"], spp=16, seed=50 + i) for i, f in enumerate(load)]
        tb = srv.submit(busy, cam, spp=12, seed=5)
        assert [srv.wait(t, 60000) for t in tl + [tb]] == [slr_amd.REQ_DONE] * 5
    finally:
        srv.close()
    assert np.array_equal(_bits(idle), _bits(busy))
