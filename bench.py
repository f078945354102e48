"""Benchmark of the spectral path tracer on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config c2|c1|c3|c4|batch8|serve]
                    [--legs relmse,matched,cpu,serve,c3,c4,bpt | --legs none] [--out FILE]

Workload (default, "c2"): the Cornell box plus a displaced icosphere at
subdivision 8 (1,310,720 triangles, GGX conductor alpha 0.2, eta 0.2,
k 3.9), 1920x1080, 64 samples per pixel, SBVH collapsed to QBVH, 16
wavelengths per path, path length capped at 10 with Russian roulette. One
step = one full frame (132.7M paths) rendered into an FP64 film, scene and
accelerator resident in HBM. The timed region is the K steps.

One JSON line: value = Mrays/s (extension + shadow rays per second of the
whole frame loop), samples/s beside it, the traversal kernel's roofline
(algorithmic bytes from the traversal counters over the mean launch time
from HIP events on its stream, and the DRAM traffic measured by rocprofv3),
and the legs below (untimed; --legs selects them):

  relmse   the C2 scene at 128x72 (the same view) against the reference
           renderer's image statistics (tests/golden/ref_c2.npz, 32 batches
           of 64 spp): relMSE and 3-sigma failures of 32 GPU batches at the
           reference's path-length cap, and the relMSE of the configured
           depth cap of 10 (its truncation bias)
  matched  the GPU frame rate at the reference's own estimator (path cap
           100) on the 1080p C2 frame, the rate the CPU baseline compares to
  cpu      the reference renderer (compiled libSLR, oracle/_ref) timed on
           the host cores on a bounded sample of the same scene
  serve    the resident C2 scene served to seeded requests (pose, resolution
           256^2..1080p, 4/16/64 spp, depth 10) arriving as Poisson
           processes at 1, 4, 16 and 64 requests/s (p50/p99 latency), the
           served-over-plain throughput of all-at-once submission for that
           mix and for small requests (256^2/512^2 x 4/16 spp)
  c3       256 instances of four meshes under a sky, 8 x 1080p x 64 spp in
           one wavefront, one step, with its traversal roofline (instance
           transforms counted)
  c4       the Cornell box around a 256^3 fBm medium, volumetric, 4 x
           1024^2 x 128 spp in one wavefront, one step, with the medium
           tracking kernel's roofline
  bpt      the bidirectional integrator (BPTRenderer): the Cornell box at
           512x512 x 16 spp (C1's scene) beside the reference BPTRenderer on
           the host cores, the C2 scene at 1920x1080 x 4 spp with its
           traversal roofline, and 32 GPU batches of the Cornell box at
           32x32 x 64 spp against the reference BPT image statistics
           (tests/golden/ref_bpt_cornell32.npz)

Other workloads as the headline (--config):
  c1      the Cornell box alone, 512x512, 16 spp, depth 8.
  c3, c4  as the legs, K steps.
  batch8  eight 1080p / 64 spp frames of the c2 scene from different camera
          poses in one wavefront (slrh_render_batch); one step = the batch.
  serve   the serving leg alone (--rate, --requests, --mix select one run).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "Mrays/s + samples/s at matched relMSE; p50/p99 latency of served renders"
ALL_LEGS = ("relmse", "matched", "cpu", "serve", "c3", "c4", "bpt")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def algorithmic_bytes(ext_rays, shadow_rays, nodes, tris, insts=0.0):
    # per extension ray: 4 B queue entry + 32 B ray in + 16 B hit out; per
    # shadow ray: 4 B entry + 32 B ray; 128 B per QBVH node visited; 40 B per
    # triangle tested (36 B positions + 4 B primitive id); 64 B per instance
    # entered (48 B transform + 16 B root and flags)
    return 52.0 * ext_rays + 36.0 * shadow_rays + 128.0 * nodes + 40.0 * tris + 64.0 * insts


def measured_traffic(config):
    """HBM bytes per k_trace launch measured with rocprofv3 (FETCH_SIZE and
    WRITE_SIZE, each scaled by what the counter reads for a 1 GiB copy), as
    committed under profiles/ for this workload; None for other workloads."""
    path = os.path.join(ROOT, "profiles", "k_trace_%s_traffic.json" % config)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return round(json.load(f)["traffic_bytes_per_dispatch"])


def calibration_copy(nbytes):
    """One device-to-device hipMemcpy of nbytes (the runtime's copy kernel),
    straight through the HIP runtime."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    src, dst = ctypes.c_void_p(), ctypes.c_void_p()
    for p in (src, dst):
        if hip.hipMalloc(ctypes.byref(p), ctypes.c_size_t(nbytes)) != 0:
            raise RuntimeError("hipMalloc failed")
    hip.hipMemset(src, 1, ctypes.c_size_t(nbytes))
    hip.hipDeviceSynchronize()
    if hip.hipMemcpy(dst, src, ctypes.c_size_t(nbytes), 3) != 0:   # hipMemcpyDeviceToDevice
        raise RuntimeError("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(src)
    hip.hipFree(dst)


def trace_roofline(tot, config, instanced=False,
                   kernel="k_trace (QBVH closest hit for extension rays + any hit for shadow rays, one launch)"):
    """k_trace's algorithmic bytes per launch over its mean launch time (HIP
    events around each launch on its stream), against the HBM peak; beside
    it the same without the node reads served from the LDS copy of the top
    of the tree, and the DRAM traffic rocprofv3 measured for the workload."""
    insts = tot.get("inst_entered", 0.0) if instanced else 0.0
    nodes = tot["nodes_visited"] + tot["nodes_shadow"]
    tr_bytes = algorithmic_bytes(tot["extend_rays"], tot["shadow_rays"], nodes,
                                 tot["tris_tested"] + tot["tris_shadow"], insts)
    tr_launches = max(1.0, tot["trace_launches"])
    tr_avg = tot["trace_seconds"] / tr_launches
    per_launch = tr_bytes / tr_launches
    achieved = per_launch / tr_avg / 1e9 if tr_avg > 0 else 0.0
    no_lds = (tr_bytes - 128.0 * tot.get("nodes_lds", 0.0)) / tr_launches / tr_avg / 1e9 if tr_avg > 0 else 0.0
    traffic = measured_traffic(config)
    r = {
        "bound": "hbm",
        "kernel": kernel,
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": traffic,
        "avg_launch_ms": round(tr_avg * 1e3, 4),
        "launches": int(tr_launches),
        "bytes_per_launch": round(per_launch),
        "achieved_without_lds_nodes": round(no_lds, 1),
        "frac_without_lds_nodes": round(no_lds / HBM_PEAK_GBS, 4),
        "lds_node_share": round(tot.get("nodes_lds", 0.0) / max(1.0, nodes), 4),
        "nodes_per_ext_ray": round(tot["nodes_visited"] / max(1, tot["extend_rays"]), 3),
        "tris_per_ext_ray": round(tot["tris_tested"] / max(1, tot["extend_rays"]), 3),
        "nodes_per_shadow_ray": round(tot["nodes_shadow"] / max(1, tot["shadow_rays"]), 3),
        "tris_per_shadow_ray": round(tot["tris_shadow"] / max(1, tot["shadow_rays"]), 3),
    }
    if traffic and tr_avg > 0:
        r["dram_achieved"] = round(traffic / tr_avg / 1e9, 1)
        r["frac_dram"] = round(traffic / tr_avg / 1e9 / HBM_PEAK_GBS, 4)
    if instanced:
        r["instances_per_ray"] = round(insts / max(1.0, tot["extend_rays"] + tot["shadow_rays"]), 3)
    return r


def medium_roofline(tot):
    """k_medium's counted bytes per launch over its mean launch time (HIP
    events around each launch): per trace-queue entry 4 B (the entry) + 1 B
    (visibility flag); per tracked extension ray 80 B read (origin, direction,
    hit, stream) and 16 B written (stream); per tracked light sample 80 B
    read (origin, direction, seed words) and 128 B (its contribution, read
    and written) when attenuated; 32 B per density lookup (the 8 lattice
    values of one trilinear lookup), 17 B per scattering event written (hit
    and shading class). The majorant grid of the configurations (32^3 cells)
    is read from each block's LDS copy, so cells entered add no HBM bytes.
    The kernel's counters show it bound by instruction issue, not bytes: the
    fraction is reported against HBM for comparison only. None without a
    medium."""
    if tot.get("medium_lookups", 0) <= 0 or tot.get("medium_seconds", 0) <= 0:
        return None
    launches = max(1.0, tot["trace_launches"])
    entries = tot["extend_rays"] + tot["shadow_rays"]
    shadows = tot["medium_shadows"]
    ext = tot["medium_rays"] - shadows
    b = 5.0 * entries + 96.0 * ext + 208.0 * shadows + 32.0 * tot["medium_lookups"] + 17.0 * tot["medium_events"]
    avg = tot["medium_seconds"] / launches
    achieved = b / launches / avg / 1e9
    return {"bound": "hbm", "kernel": "k_medium (delta tracking of extension rays, ratio tracking of light samples)",
            "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "avg_launch_ms": round(avg * 1e3, 4),
            "bytes_per_launch": round(b / launches),
            "entries_per_launch": round(entries / launches), "tracked_rays_per_launch": round(tot["medium_rays"] / launches),
            "lookups_per_ray": round(tot["medium_lookups"] / max(1.0, tot["medium_rays"]), 3),
            "cells_per_ray": round(tot["medium_cells"] / max(1.0, tot["medium_rays"]), 3),
            "guard_stops": int(tot.get("medium_guard", 0))}


def frame_breakdown(tot, steps):
    return {"trace": round(tot["trace_seconds"] / steps * 1e3, 2),
            "shade": round(tot["shade_seconds"] / steps * 1e3, 2),
            "medium": round(tot.get("medium_seconds", 0.0) / steps * 1e3, 2),
            "regen_compact": round(tot["other_seconds"] / steps * 1e3, 2)}


def host_description():
    """Core counts and CPU model of the host the baseline runs on."""
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    usable = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return {"nproc": usable, "cpu_count": os.cpu_count(), "cpu_model": model}


def cpu_baseline(desc, sample, rays_per_sample, gpu_matched_sps=None):
    """Reference PTRenderer on the host cores over a bounded sample of the
    same scene; excludes its SBVH build (reported separately).
    rays_per_sample: the GPU's count for the same estimator (the reference
    counts no rays)."""
    from oracle import refdrv
    if not refdrv.available():
        return None
    w, h, spp = sample
    host = host_description()
    threads = max(1, min(host["nproc"], 64))
    ref = refdrv.RefScene(desc)
    log("cpu baseline: reference PTRenderer %dx%dx%d spp on %d threads (SBVH build first)" % (w, h, spp, threads))
    film, sec, build_s = ref.render(w, h, spp, 1509761209, threads=threads)
    ref.close()
    sps = w * h * spp / sec
    out = {"value": sps * rays_per_sample / 1e6, "unit": "Mrays/s", "cores": threads, "kind": "reference",
           "samples_per_s": sps, "seconds": sec, "sbvh_build_seconds": build_s, "host": host,
           "sample": "reference libSLR PTRenderer (its fixed path-length cap of 100), same scene at %dx%d, %d spp, "
                     "%d threads; Mrays/s = samples/s x %.3f rays/sample counted on the GPU path for the same "
                     "estimator" % (w, h, spp, threads, rays_per_sample)}
    if gpu_matched_sps:
        out["gpu_over_cpu_same_estimator"] = round(gpu_matched_sps / sps, 1)
    return out


# kernel choices pinned for profiling runs (--tune-trace, --shade-block-128):
# the autotuner's choice can differ from one process to the next
PIN = {}


def run_frames(renderer, sc, film, steps, warmup, seed0=0, max_depth=0):
    """warmup + steps renders of one frame, kernels timed; summed stats and
    wall seconds of the timed steps."""
    for i in range(warmup):
        film.clear()
        renderer.render(sc, film, seed=1000 + seed0 + i, max_depth=max_depth, **PIN)
    tot = {}
    t0 = time.perf_counter()
    for i in range(steps):
        film.clear()
        st = renderer.render(sc, film, seed=seed0 + i, max_depth=max_depth, time_kernels=True, **PIN)
        for k, v in st.items():
            tot[k] = tot.get(k, 0.0) + v
    return tot, time.perf_counter() - t0


def leg_relmse(ctx, sc):
    """The C2 view at 128x72 against the reference's statistics: 32 GPU
    batches of 64 spp at the reference's path cap (parity) and at the
    configured cap of 10 (its truncation bias)."""
    import slr_amd
    from tests import refcases
    ref = refcases.load("c2")
    _, W, H, B, spp, _, _ = refcases.CASES["c2"]
    cmf = slr_amd.cmf_strata()
    out = {"reference": "reference PTRenderer, C2 scene at %dx%d, %d batches x %d spp (tests/golden/ref_c2.npz)"
                        % (W, H, int(ref["batches"]), int(ref["spp"]))}
    for depth, key in ((0, "reference_cap"), (10, "depth_10")):
        gb = []
        for b in range(B):
            film = slr_amd.Film(ctx, W, H)
            slr_amd.PTRenderer(spp).render(sc, film, seed=7000 + b, max_depth=depth)
            gb.append(refcases.xyz(film.strata(), cmf) / spp)
        gb = np.array(gb)
        mean = gb.mean(0)
        se = np.sqrt(gb.var(0, ddof=1) / B + ref["se"] ** 2)
        fail = float(np.mean(np.abs(mean - ref["mean"]) / np.maximum(se, 1e-12) > 3.0))
        relmse = float(np.mean((mean - ref["mean"]) ** 2 / (ref["mean"] ** 2 + 1e-3)))
        out[key] = {"relmse": relmse, "outside_3sigma": round(fail, 5), "gpu_batches": B, "spp_per_batch": spp}
    return out


def _request_films(ctx, reqs):
    import slr_amd
    return [slr_amd.Film(ctx, q["width"], q["height"]) for q in reqs]


def _serve(sc, reqs, films, gaps, budget=8 << 30):
    import slr_amd
    for f in films:
        f.clear()
    srv = slr_amd.Server(sc, path_state_bytes=budget)
    tickets = []
    t0 = time.perf_counter()
    for q, f, gap in zip(reqs, films, gaps):
        if gap > 0:
            time.sleep(gap)
        tickets.append(srv.submit(f, q["camera"], spp=q["spp"], seed=q["seed"], max_depth=10))
    states = [srv.wait(t) for t in tickets]
    dt = time.perf_counter() - t0
    st = srv.stats()
    srv.close()
    if not all(x == slr_amd.REQ_DONE for x in states):
        raise RuntimeError("a served request did not complete")
    return dt, st


def _plain(sc, reqs, films):
    import slr_amd
    t0 = time.perf_counter()
    for q, f in zip(reqs, films):
        f.clear()
        slr_amd.PTRenderer(q["spp"]).render_batch(sc, [dict(film=f, camera=q["camera"], seed=q["seed"])],
                                                  max_depth=10)
    return time.perf_counter() - t0


def leg_serve(ctx, desc, sc, rates=(1.0, 4.0, 16.0, 64.0), counts=(24, 96, 256, 256), n_throughput=160,
              n_small=96):
    """Poisson arrivals at each rate (latency), and all-at-once submission
    against one-at-a-time plain renders (throughput) for the full and the
    small request mix."""
    from slr_amd import scenes
    n_all = max(max(counts), n_throughput)
    reqs = scenes.random_requests(desc, n_all, seed=3)
    films = _request_films(ctx, reqs)
    out = {"mix": "resolution 256^2/512^2/1024^2/1920x1080, spp 4/16/64, depth 10, seeded camera poses"}
    sub, sfilms = reqs[:n_throughput], films[:n_throughput]
    samples = sum(q["width"] * q["height"] * q["spp"] for q in sub)
    t_plain = _plain(sc, sub, sfilms)
    t_sat, st_sat = _serve(sc, sub, sfilms, np.zeros(len(sub)))
    out["throughput"] = {"requests": n_throughput, "plain_samples_per_s": round(samples / t_plain, 1),
                         "served_samples_per_s": round(samples / t_sat, 1),
                         "served_over_plain": round(t_plain / t_sat, 3),
                         "saturated_p50_ms": round(st_sat["p50_ms"], 2), "saturated_p99_ms": round(st_sat["p99_ms"], 2)}
    log("serve: full mix plain %.2fs, served %.2fs" % (t_plain, t_sat))
    small = scenes.random_requests(desc, n_small, seed=5, sizes=((256, 256), (512, 512)), spps=(4, 16))
    small_films = _request_films(ctx, small)
    samples_small = sum(q["width"] * q["height"] * q["spp"] for q in small)
    tp = _plain(sc, small, small_films)
    ts, _ = _serve(sc, small, small_films, np.zeros(len(small)))
    out["small_mix"] = {"requests": n_small, "mix": "256^2/512^2 x 4/16 spp",
                        "plain_samples_per_s": round(samples_small / tp, 1),
                        "served_samples_per_s": round(samples_small / ts, 1), "served_over_plain": round(tp / ts, 3)}
    log("serve: small mix plain %.2fs, served %.2fs" % (tp, ts))
    sweep = []
    rng = np.random.RandomState(11)
    for rate, n in zip(rates, counts):
        gaps = rng.exponential(1.0 / rate, size=n)
        q, f = reqs[:n], films[:n]
        dt, st = _serve(sc, q, f, gaps)
        s = sum(x["width"] * x["height"] * x["spp"] for x in q)
        sweep.append({"rate_per_s": rate, "requests": n, "seconds": round(dt, 2), "p50_ms": round(st["p50_ms"], 2),
                      "p99_ms": round(st["p99_ms"], 2), "mean_ms": round(st["mean_ms"], 2),
                      "samples_per_s": round(s / dt, 1)})
        log("serve: %.0f req/s, %d requests in %.1fs, p50 %.1f ms, p99 %.1f ms" % (rate, n, dt, st["p50_ms"],
                                                                                st["p99_ms"]))
    out["poisson"] = sweep
    for f in films + small_films:
        f.close()
    return out


def _c3_scene(ctx, W=1920, H=1080):
    import slr_amd
    from slr_amd import scenes
    desc = scenes.instanced_field(W, H)
    sc = slr_amd.build_scene(ctx, desc)
    ntris = sum(len(p["idx"]) for p in desc["prototypes"])
    return desc, sc, ntris


def _c4_scene(ctx, W=1024, H=1024):
    import slr_amd
    from slr_amd import scenes
    desc = scenes.cornell_medium(W, H, n=256, sigma_t=20.0, albedo=0.9, g=0.7)
    return desc, slr_amd.build_scene(ctx, desc)


def batch_steps(sc, frames, renderer, max_depth, steps, warmup, slots_per_pixel=0):
    for i in range(warmup):
        for f in frames:
            f["film"].clear()
        renderer.render_batch(sc, frames, max_depth=max_depth, slots_per_pixel=slots_per_pixel, **PIN)
    tot = {}
    t0 = time.perf_counter()
    for i in range(steps):
        for f in frames:
            f["film"].clear()
        st = renderer.render_batch(sc, frames, max_depth=max_depth, time_kernels=True,
                                   slots_per_pixel=slots_per_pixel, **PIN)
        for k, v in st.items():
            tot[k] = tot.get(k, 0.0) + v
        log("batch step %d: %.3fs device, %d rays, %d non-finite" % (i, st["device_seconds"], st["rays"],
                                                                     st["nonfinite"]))
    return tot, time.perf_counter() - t0


def c3_frames(ctx, desc, W=1920, H=1080):
    import slr_amd
    from slr_amd import scenes
    return [dict(film=slr_amd.Film(ctx, W, H), camera=scenes.camera_pose(desc, yaw_deg=-7 + 2 * i), seed=31 + i)
            for i in range(8)]


def c4_frames(ctx, desc, W=1024, H=1024):
    import slr_amd
    from slr_amd import scenes
    return [dict(film=slr_amd.Film(ctx, W, H), camera=scenes.camera_pose(desc, yaw_deg=-6 + 4 * i), seed=41 + i)
            for i in range(4)]


def batch_summary(tot, elapsed, steps, config, instanced=False):
    r = {"mrays_per_s": round(tot["rays"] / elapsed / 1e6, 2), "samples_per_s": round(tot["samples"] / elapsed, 1),
         "ms_per_step": round(elapsed / steps * 1e3, 2), "steps": steps,
         "rays_per_sample": round(tot["rays"] / max(1.0, tot["samples"]), 3),
         "roofline": trace_roofline(tot, config, instanced), "frame_ms_breakdown": frame_breakdown(tot, steps),
         "trace_variant": int(tot["trace_variant"] / max(1, steps)), "nonfinite": int(tot["nonfinite"]),
         "slots_per_pixel": int(tot.get("slots_per_pixel", 0) / max(1, steps)),
         "shade_block": int(tot.get("shade_block", 0) / max(1, steps))}
    if tot.get("medium_lookups", 0) > 0:
        r["medium_events_per_sample"] = round(tot["medium_events"] / max(1.0, tot["samples"]), 3)
        r["medium_roofline"] = medium_roofline(tot)
    return r


def leg_c3(ctx, steps=1, warmup=1):
    import slr_amd
    t0 = time.time()
    desc, sc, ntris = _c3_scene(ctx)
    log("c3 scene built in %.1fs" % (time.time() - t0))
    frames = c3_frames(ctx, desc)
    tot, el = batch_steps(sc, frames, slr_amd.PTRenderer(64), 12, steps, warmup)
    r = batch_summary(tot, el, steps, "c3", instanced=True)
    r["workload"] = ("256 affine instances of {displaced icosphere 1.31M, terrain 200k, torus knot 50k, box 12} "
                     "(%d prototype triangles), 4 BSDFs, 2048x1024 sky + sun, 8 x 1920x1080 x 64 spp, depth 12, "
                     "one wavefront" % ntris)
    for f in frames:
        f["film"].close()
    sc.close()
    return r


def leg_c4(ctx, steps=1, warmup=1):
    import slr_amd
    t0 = time.time()
    desc, sc = _c4_scene(ctx)
    log("c4 scene built in %.1fs" % (time.time() - t0))
    frames = c4_frames(ctx, desc)
    tot, el = batch_steps(sc, frames, slr_amd.VolumetricPTRenderer(128), 32, steps, warmup)
    r = batch_summary(tot, el, steps, "c4")
    r["workload"] = ("cornell box around a 256^3 density grid max(0, fBm5) x sigma_t 20, albedo 0.9, HG g 0.7, "
                     "volumetric, 4 x 1024x1024 x 128 spp, depth 32, one wavefront")
    for f in frames:
        f["film"].close()
    sc.close()
    return r


def cpu_bpt_baseline(desc, sample, rays_per_sample, gpu_sps):
    """Reference BPTRenderer on the host cores over the same scene and size
    as the GPU run (the Cornell box, whose build is instant)."""
    from oracle import refdrv
    if not refdrv.available():
        return None
    w, h, spp = sample
    host = host_description()
    threads = max(1, min(host["nproc"], 64))
    ref = refdrv.RefScene(desc)
    log("cpu baseline: reference BPTRenderer %dx%dx%d spp on %d threads" % (w, h, spp, threads))
    _, sec, _ = ref.render(w, h, spp, 1509761209, threads=threads, bpt=True)
    ref.close()
    sps = w * h * spp / sec
    return {"value": round(sps * rays_per_sample / 1e6, 3), "unit": "Mrays/s", "cores": threads, "kind": "reference",
            "samples_per_s": round(sps, 1), "seconds": round(sec, 3), "host": host,
            "sample": "reference libSLR BPTRenderer, the Cornell box at %dx%d, %d spp, %d threads; Mrays/s = "
                      "samples/s x %.3f rays/sample counted on the GPU for the same estimator" % (
                          w, h, spp, threads, rays_per_sample),
            "gpu_over_cpu": round(gpu_sps / sps, 1)}


def leg_bpt(ctx, sc_c2, c2_desc, with_cpu):
    """The bidirectional integrator: C1's scene and size against the
    reference on the host cores, the C2 scene at 1080p x 4 spp (traversal
    roofline), and its image against the reference BPT statistics."""
    import slr_amd
    from slr_amd import scenes
    from tests import refcases
    out = {}
    W, H, spp = 512, 512, 16
    desc = scenes.cornell_box(W, H)
    sc = slr_amd.build_scene(ctx, desc)
    film = slr_amd.Film(ctx, W, H)
    r = slr_amd.BPTRenderer(spp)
    tot, el = run_frames(r, sc, film, 3, 1)
    c1 = {"workload": "cornell box 512x512 x 16 spp, BPT (no path cap, as the reference)",
          "mrays_per_s": round(tot["rays"] / el / 1e6, 2), "samples_per_s": round(tot["samples"] / el, 1),
          "ms_per_frame": round(el / 3 * 1e3, 2), "rays_per_sample": round(tot["rays"] / tot["samples"], 3),
          "connections_per_sample": round(tot["connections"] / tot["samples"], 3),
          "lens_contributions_per_sample": round(tot["lens_contributions"] / tot["samples"], 3),
          "capped_subpaths": int(tot["capped_subpaths"])}
    out["c1"] = c1
    film.close()
    # relMSE against the reference BPT image statistics
    ref = refcases.load("bpt_cornell32")
    make, w, h, B, s2, _, _ = refcases.CASES["bpt_cornell32"]
    cmf = slr_amd.cmf_strata()
    sc32 = slr_amd.build_scene(ctx, make())
    gb = []
    for b in range(B):
        f = slr_amd.Film(ctx, w, h)
        r2 = slr_amd.BPTRenderer(s2)
        r2.render(sc32, f, seed=8100 + b)
        gb.append(refcases.xyz(f.strata(), cmf) / s2)
        f.close()
    gb = np.array(gb)
    mean = gb.mean(0)
    se = np.sqrt(gb.var(0, ddof=1) / B + ref["se"] ** 2)
    out["relmse"] = {"reference": "reference BPTRenderer, Cornell box 32x32, %d batches x %d spp" % (
        int(ref["batches"]), int(ref["spp"])),
        "relmse": float(np.mean((mean - ref["mean"]) ** 2 / (ref["mean"] ** 2 + 1e-3))),
        "outside_3sigma": round(float(np.mean(np.abs(mean - ref["mean"]) / np.maximum(se, 1e-12) > 3.0)), 5)}
    sc32.close()
    # the C2 scene at 1080p
    W2, H2, spp2 = 1920, 1080, 4
    film = slr_amd.Film(ctx, W2, H2)
    tot, el = run_frames(slr_amd.BPTRenderer(spp2), sc_c2, film, 2, 1)
    out["c2"] = {"workload": "C2 scene (1.31M-triangle GGX icosphere in the Cornell box), 1920x1080 x 4 spp, BPT",
                 "mrays_per_s": round(tot["rays"] / el / 1e6, 2), "samples_per_s": round(tot["samples"] / el, 1),
                 "ms_per_frame": round(el / 2 * 1e3, 2), "rays_per_sample": round(tot["rays"] / tot["samples"], 3),
                 "roofline": trace_roofline(tot, "bpt_c2", kernel="k_bpt_trace (closest hit for light and eye "
                                            "rays, any hit for connection and lens rays, one launch)"),
                 "device_ms_per_frame": round(tot["device_seconds"] / 2 * 1e3, 2),
                 "trace_ms_per_frame": round(tot["trace_seconds"] / 2 * 1e3, 2),
                 "capped_subpaths": int(tot["capped_subpaths"])}
    film.close()
    if with_cpu:
        out["cpu_baseline"] = cpu_bpt_baseline(desc, (W, H, spp), c1["rays_per_sample"], c1["samples_per_s"])
    sc.close()
    return out


def headline_line(args, result):
    line = json.dumps(result)
    print(line, flush=True)
    if args.out:
        with open(args.out, "w") as f:
            f.write(line + "\n")


def batch_config(args, ctx):
    """--config c3 / c4 / batch8: K steps of a multi-frame wavefront."""
    import slr_amd
    from slr_amd import scenes
    inst = False
    if args.config == "c3":
        desc, sc, ntris = _c3_scene(ctx)
        frames, renderer, depth = c3_frames(ctx, desc), slr_amd.PTRenderer(64), 12
        workload = "c3: 256 instances of 4 meshes (%d prototype triangles), 8 x 1080p x 64 spp, depth 12" % ntris
        inst = True
    elif args.config == "c4":
        desc, sc = _c4_scene(ctx)
        frames, renderer, depth = c4_frames(ctx, desc), slr_amd.VolumetricPTRenderer(128), 32
        workload = "c4: 256^3 fBm medium sigma_t 20 in the Cornell box, volumetric, 4 x 1024^2 x 128 spp, depth 32"
    else:
        desc = scenes.cornell_icosphere(1920, 1080, subdiv=8)
        sc = slr_amd.build_scene(ctx, desc)
        frames = [dict(film=slr_amd.Film(ctx, 1920, 1080), camera=scenes.camera_pose(desc, yaw_deg=-14 + 4 * i),
                       seed=17 + i) for i in range(8)]
        renderer, depth = slr_amd.PTRenderer(64), 10
        workload = "batch8: 8 frames x 1920x1080 x 64 spp of the C2 scene from 8 poses, depth 10, one wavefront"
    tot, el = batch_steps(sc, frames, renderer, depth, args.steps, args.warmup, args.slots_per_pixel)
    s = batch_summary(tot, el, args.steps, args.config, inst)
    result = {"metric": METRIC, "value": s["mrays_per_s"], "unit": "Mrays/s", "n_gpus": 1, "steps": args.steps,
              "warmup": args.warmup, "ms_per_step": s["ms_per_step"], "higher_is_better": True, "scaling": "weak",
              "vs_baseline": None, "dtype": "f32 (film f64)", "data": "synthetic: seeded procedural scenes",
              "config": {"workload": workload}}
    result.update({k: v for k, v in s.items() if k not in ("mrays_per_s", "ms_per_step", "steps")})
    headline_line(args, result)


def serve_config(args, ctx):
    import slr_amd
    from slr_amd import scenes
    desc = scenes.cornell_icosphere(1920, 1080, subdiv=8)
    sc = slr_amd.build_scene(ctx, desc)
    if args.mix == "small":
        reqs = scenes.random_requests(desc, args.requests, seed=3, sizes=((256, 256), (512, 512)), spps=(4, 16))
    else:
        reqs = scenes.random_requests(desc, args.requests, seed=3)
    films = _request_films(ctx, reqs)
    samples = sum(q["width"] * q["height"] * q["spp"] for q in reqs)
    t_plain = _plain(sc, reqs, films)
    t_sat, st_sat = _serve(sc, reqs, films, np.zeros(len(reqs)))
    rng = np.random.RandomState(11)
    gaps = rng.exponential(1.0 / args.rate, size=len(reqs)) if args.rate > 0 else np.zeros(len(reqs))
    t_poi, st_poi = _serve(sc, reqs, films, gaps)
    result = {"metric": METRIC, "value": round(samples / t_sat, 1), "unit": "samples/s", "n_gpus": 1, "steps": 1,
              "warmup": 0, "ms_per_step": round(t_sat * 1e3, 2), "higher_is_better": True, "scaling": "weak",
              "vs_baseline": None, "dtype": "f32 (film f64)",
              "data": "synthetic: seeded requests (pose, resolution, spp) on the procedural scene",
              "config": {"workload": "serving %d requests (%s mix, depth 10) on the resident C2 scene"
                                     % (len(reqs), args.mix), "requests": len(reqs), "rate": args.rate},
              "plain_samples_per_s": round(samples / t_plain, 1), "served_over_plain": round(t_plain / t_sat, 3),
              "poisson": {"rate_per_s": args.rate, "seconds": round(t_poi, 3), "p50_ms": round(st_poi["p50_ms"], 2),
                          "p99_ms": round(st_poi["p99_ms"], 2), "mean_ms": round(st_poi["mean_ms"], 2)},
              "saturated": {"p50_ms": round(st_sat["p50_ms"], 2), "p99_ms": round(st_sat["p99_ms"], 2)}}
    headline_line(args, result)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="c2", choices=["c1", "c2", "c3", "c4", "batch8", "serve"])
    ap.add_argument("--legs", default=",".join(ALL_LEGS),
                    help="legs added to the c2 line: comma list of %s, or none" % ",".join(ALL_LEGS))
    ap.add_argument("--no-cpu-baseline", action="store_true", help="same as leaving cpu out of --legs")
    ap.add_argument("--rate", type=float, default=16.0, help="--config serve: request arrivals per second")
    ap.add_argument("--requests", type=int, default=48, help="--config serve: number of requests")
    ap.add_argument("--mix", default="full", choices=["full", "small"])
    ap.add_argument("--tune-trace", type=int, default=0,
                    help="traversal variant + 1 for the timed renders (0: the autotuner's choice)")
    ap.add_argument("--shade-block-128", action="store_true",
                    help="shading kernels in 128-thread workgroups (default: the autotuner's choice, or 256 "
                         "with --tune-trace)")
    ap.add_argument("--slots-per-pixel", type=int, default=0,
                    help="--config c3/c4/batch8: live paths per pixel (0: the library's choice)")
    ap.add_argument("--cpu-sample", default="1920x1080x8")
    ap.add_argument("--out", default="")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="first run one device-to-device copy of this many MiB (a kernel of known traffic that "
                         "scales the FETCH_SIZE / WRITE_SIZE counters of a profiling run)")
    args = ap.parse_args()
    if args.tune_trace:
        PIN["tune_trace"] = args.tune_trace
    if args.shade_block_128:
        PIN["shade_block_128"] = True

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (frames, requests or tiles would be "
              "sharded one per GPU with no exchange); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import slr_amd
    from slr_amd import scenes

    if args.calib_copy_mib > 0:
        calibration_copy(args.calib_copy_mib << 20)
        log("calibration copy: %d MiB device to device" % args.calib_copy_mib)
    ctx = slr_amd.Context(0)
    if args.config == "serve":
        return serve_config(args, ctx)
    if args.config in ("c3", "c4", "batch8"):
        return batch_config(args, ctx)

    legs = set() if args.legs == "none" else set(x for x in args.legs.split(",") if x)
    if args.no_cpu_baseline:
        legs.discard("cpu")
    t0 = time.time()
    if args.config == "c2":
        W, H, spp, depth = 1920, 1080, 64, 10
        desc = scenes.cornell_icosphere(W, H, subdiv=8)
        workload = "cornell+displaced icosphere (1,310,720 tris, GGX conductor), 1920x1080, 64 spp, depth 10, " \
                   "SBVH->QBVH"
    else:
        W, H, spp, depth = 512, 512, 16, 8
        desc = scenes.cornell_box(W, H)
        workload = "cornell box (36 tris, Lambert), 512x512, 16 spp, depth 8, SBVH->QBVH"
        legs = set()
    ntris = sum(len(m["idx"]) for m in desc["meshes"])
    log("scene generated: %d triangles in %.1fs" % (ntris, time.time() - t0))
    t0 = time.time()
    sc = slr_amd.build_scene(ctx, desc)
    build_s = time.time() - t0
    st_acc = sc.accel_stats()
    log("scene built (SBVH+QBVH+upload) in %.2fs" % build_s)
    film = slr_amd.Film(ctx, W, H)
    renderer = slr_amd.PTRenderer(spp)

    # timed region: K full frames, device synchronised on both sides
    # (slrh_render returns after its stream completes)
    tot, elapsed = run_frames(renderer, sc, film, args.steps, args.warmup, max_depth=depth)
    log("timed: %d frames in %.3fs" % (args.steps, elapsed))
    rays, samples = tot["rays"], tot["samples"]
    result = {
        "metric": METRIC,
        "value": round(rays / elapsed / 1e6, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (film f64)",
        "data": "synthetic: seeded procedural scene (icosphere + Perlin fBm displacement, Cornell box)",
        "config": {"workload": workload, "width": W, "height": H, "spp": spp, "triangles": ntris,
                   "wavelengths_per_path": 16, "max_path_length": depth},
        "samples_per_s": round(samples / elapsed, 1),
        "rays_per_sample": round(rays / samples, 3),
        "roofline": trace_roofline(tot, args.config),
        "frame_ms_breakdown": frame_breakdown(tot, args.steps),
        "trace_variant": int(tot["trace_variant"] / max(1, args.steps)),
        "accel": {"build_seconds": round(st_acc["build_seconds"], 3), "qbvh_nodes": int(st_acc["qbvh_nodes"]),
                  "qbvh_refs": int(st_acc["qbvh_refs"]), "qbvh_sah": round(st_acc["qbvh_cost"], 3)},
        "cpu_baseline": None,
    }
    legs_done = {}

    def leg(name, fn):
        if name not in legs:
            return None
        t = time.time()
        try:
            r = fn()
        except Exception as e:   # a leg is reported, never required for the headline
            log("leg %s failed: %r" % (name, e))
            r = {"error": repr(e)}
        legs_done[name] = round(time.time() - t, 1)
        log("leg %s: %.1fs" % (name, legs_done[name]))
        return r

    matched = leg("matched", lambda: (lambda tt: {
        "samples_per_s": round(tt[0]["samples"] / tt[1], 1),
        "mrays_per_s": round(tt[0]["rays"] / tt[1] / 1e6, 2),
        "rays_per_sample": round(tt[0]["rays"] / tt[0]["samples"], 3),
        "ms_per_frame": round(tt[1] / 2 * 1e3, 2),
        "note": "the same 1080p x 64 spp frame with the reference's path-length cap of 100 (its estimator)"})(
        run_frames(renderer, sc, film, 2, 0, seed0=500, max_depth=0)))
    if matched:
        result["same_estimator_as_reference"] = matched
    # the 1080p scene's camera has the 128x72 fixture's aspect: the same view
    r = leg("relmse", lambda: leg_relmse(ctx, sc))
    if r:
        result["relmse"] = r
    if "cpu" in legs:
        w, h, s = [int(x) for x in args.cpu_sample.split("x")]
        rps = matched["rays_per_sample"] if matched and "rays_per_sample" in matched else None
        if rps is None:
            t100, _ = run_frames(renderer, sc, film, 1, 0, seed0=77, max_depth=0)
            rps = t100["rays"] / t100["samples"]
        gps = matched.get("samples_per_s") if matched else None
        result["cpu_baseline"] = leg("cpu", lambda: cpu_baseline(desc, (w, h, s), rps, gps))
    r = leg("serve", lambda: leg_serve(ctx, desc, sc))
    if r:
        result["serve"] = r
    r = leg("bpt", lambda: leg_bpt(ctx, sc, desc, "cpu" in legs))
    if r:
        result["bpt"] = r
    if "c3" in legs or "c4" in legs:
        film.close()
        sc.close()
    r = leg("c3", lambda: leg_c3(ctx))
    if r:
        result["c3"] = r
    r = leg("c4", lambda: leg_c4(ctx))
    if r:
        result["c4"] = r
    result["leg_seconds"] = legs_done
    headline_line(args, result)


if __name__ == "__main__":
    main()
