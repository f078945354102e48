"""Benchmark of the spectral path tracer on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config c2|c1|c3|c4|batch8|serve]
                    [--no-cpu-baseline] [--cpu-sample WxHxSPP] [--out FILE]

Workload (default, "c2"): the Cornell box plus a displaced icosphere at
subdivision 8 (1,310,720 triangles, GGX conductor alpha 0.2, eta 0.2,
k 3.9), 1920x1080, 64 samples per pixel, SBVH collapsed to QBVH, 16
wavelengths per path, path length capped at 10 with Russian roulette. One
step = one full frame (132.7M paths) rendered into an FP64 film, scene and
accelerator resident in HBM.

Prints one JSON line: value = Mrays/s (extension + shadow rays per second of
the whole frame loop), samples/s beside it, the traversal kernel's roofline
from HIP events and the traversal counters, and the reference renderer
(compiled libSLR, oracle/_ref) timed on the host cores on a bounded sample of
the same scene as the CPU baseline.

Other workloads (not the headline line):
  c1      the Cornell box alone, 512x512, 16 spp, depth 8.
  c3      256 instances of four meshes (the icosphere, a 199,712-triangle
          terrain, a 50,000-triangle torus knot, a box) with materials drawn
          from Lambert / GGX conductor / GGX dielectric / Oren-Nayar under a
          2048x1024 sky with a sun; 8 frames of 1080p x 64 spp, depth 12, in
          one wavefront (slrh_render_batch); one step = the batch.
  c4      the Cornell box around a 256^3 fBm density grid (sigma_t 20,
          albedo 0.9, Henyey-Greenstein g 0.7), volumetric integrator,
          4 frames of 1024^2 x 128 spp, depth 32, in one wavefront.
  batch8  eight 1080p / 64 spp frames of the c2 scene from different camera
          poses in one wavefront (slrh_render_batch); one step = the batch.
  serve   the resident scene served to a seeded stream of requests (pose,
          resolution from 256^2..1080p, 4/16/64 spp, depth 10) arriving as a
          Poisson process; reports p50/p99 latency and served samples/s, and
          the same requests rendered one at a time (slrh_render) for the
          plain-GPU comparison. --mix small draws only 256^2/512^2 at 4/16
          spp, where one request alone cannot fill the GPU.
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


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def algorithmic_bytes(ext_rays, shadow_rays, nodes, tris):
    # per extension ray: 4 B queue entry + 32 B ray in + 16 B hit out; per
    # shadow ray: 4 B entry + 32 B ray; 128 B per QBVH node visited; 40 B per
    # triangle tested (36 B positions + 4 B primitive id)
    return 52.0 * ext_rays + 36.0 * shadow_rays + 128.0 * nodes + 40.0 * tris


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


def trace_roofline(tot, config, inst_note=False):
    """k_trace's algorithmic bytes per launch over its mean launch time (HIP
    events around each launch on its stream), against the HBM peak."""
    tr_bytes = algorithmic_bytes(tot["extend_rays"], tot["shadow_rays"], tot["nodes_visited"] + tot["nodes_shadow"],
                                 tot["tris_tested"] + tot["tris_shadow"])
    tr_launches = max(1.0, tot["trace_launches"])
    tr_avg = tot["trace_seconds"] / tr_launches
    achieved = tr_bytes / tr_launches / tr_avg / 1e9 if tr_avg > 0 else 0.0
    r = {
        "bound": "hbm",
        "kernel": "k_trace (QBVH closest hit for extension rays + any hit for shadow rays, one launch)",
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": measured_traffic(config),
        "avg_launch_ms": round(tr_avg * 1e3, 4),
        "launches": int(tr_launches),
        "bytes_per_launch": round(tr_bytes / tr_launches),
        "nodes_per_ext_ray": round(tot["nodes_visited"] / max(1, tot["extend_rays"]), 3),
        "tris_per_ext_ray": round(tot["tris_tested"] / max(1, tot["extend_rays"]), 3),
        "nodes_per_shadow_ray": round(tot["nodes_shadow"] / max(1, tot["shadow_rays"]), 3),
        "tris_per_shadow_ray": round(tot["tris_shadow"] / max(1, tot["shadow_rays"]), 3),
    }
    if inst_note:
        r["note"] = "node and triangle bytes of world and prototype trees; instance transforms (64 B per entry) " \
                    "not counted, so achieved is a lower bound"
    return r


def medium_roofline(tot):
    """k_medium's algorithmic bytes per launch over its mean launch time (HIP
    events around each launch): 32 B per density lookup (the 8 lattice
    values of a trilinear lookup), 4 B per majorant cell entered, 48 B per
    tracked ray (origin, direction, segment, hit). None without a medium."""
    if tot.get("medium_lookups", 0) <= 0 or tot.get("medium_seconds", 0) <= 0:
        return None
    launches = max(1.0, tot["trace_launches"])
    b = 32.0 * tot["medium_lookups"] + 4.0 * tot["medium_cells"] + 48.0 * tot["medium_rays"]
    avg = tot["medium_seconds"] / launches
    achieved = b / launches / avg / 1e9
    return {"bound": "hbm", "kernel": "k_medium (delta tracking of extension rays, ratio tracking of light samples)",
            "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "avg_launch_ms": round(avg * 1e3, 4),
            "bytes_per_launch": round(b / launches),
            "lookups_per_ray": round(tot["medium_lookups"] / max(1.0, tot["medium_rays"]), 3),
            "cells_per_ray": round(tot["medium_cells"] / max(1.0, tot["medium_rays"]), 3)}


def frame_breakdown(tot, steps):
    return {"trace": round(tot["trace_seconds"] / steps * 1e3, 2),
            "shade": round(tot["shade_seconds"] / steps * 1e3, 2),
            "regen_compact": round(tot["other_seconds"] / steps * 1e3, 2)}


def cpu_baseline(desc, sample, rays_per_sample):
    """Reference PTRenderer on the host cores over a bounded sample of the
    same scene; excludes its SBVH build (reported separately).
    rays_per_sample: the GPU's count for the same estimator (the reference
    counts no rays)."""
    from oracle import refdrv
    if not refdrv.available():
        return None
    w, h, spp = sample
    threads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    threads = max(1, min(threads, 64))
    ref = refdrv.RefScene(desc)
    log("cpu baseline: reference PTRenderer %dx%dx%d spp on %d threads (SBVH build first)" % (w, h, spp, threads))
    film, sec, build_s = ref.render(w, h, spp, 1509761209, threads=threads)
    ref.close()
    sps = w * h * spp / sec
    return {"value": sps * rays_per_sample / 1e6, "unit": "Mrays/s", "cores": threads, "kind": "reference",
            "samples_per_s": sps, "seconds": sec, "sbvh_build_seconds": build_s,
            "sample": "reference libSLR PTRenderer (its fixed path-length cap of 100), same scene at %dx%d, %d spp, "
                      "%d threads; Mrays/s = samples/s x "
                      "%.3f rays/sample counted on the GPU path for the same estimator" % (w, h, spp, threads,
                                                                                          rays_per_sample)}


def _scene_c2():
    import slr_amd
    from slr_amd import scenes
    t0 = time.time()
    desc = scenes.cornell_icosphere(1920, 1080, subdiv=8)
    ctx = slr_amd.Context(0)
    sc = slr_amd.build_scene(ctx, desc)
    log("scene built in %.1fs" % (time.time() - t0))
    return desc, ctx, sc


def run_batch(args, sc, frames, renderer, max_depth, config, workload, data, inst_note=False):
    """K steps of one slrh_render_batch call over the frames, kernels timed."""
    for i in range(args.warmup):
        for f in frames:
            f["film"].clear()
        renderer.render_batch(sc, frames, max_depth=max_depth)
        log("warmup %d done" % i)
    tot = {}
    t_start = time.perf_counter()
    for i in range(args.steps):
        for f in frames:
            f["film"].clear()
        st = renderer.render_batch(sc, frames, max_depth=max_depth, time_kernels=True)
        for k, v in st.items():
            tot[k] = tot.get(k, 0.0) + v
        log("step %d: %.3fs device, %d rays, %d non-finite" % (i, st["device_seconds"], st["rays"], st["nonfinite"]))
    elapsed = time.perf_counter() - t_start
    result = {
        "metric": METRIC, "value": round(tot["rays"] / elapsed / 1e6, 2), "unit": "Mrays/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(elapsed / args.steps * 1e3, 2),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32 (film f64)",
        "data": data, "config": dict(workload=workload, **config),
        "samples_per_s": round(tot["samples"] / elapsed, 1),
        "rays_per_sample": round(tot["rays"] / max(1.0, tot["samples"]), 3),
        "roofline": trace_roofline(tot, args.config, inst_note),
        "frame_ms_breakdown": frame_breakdown(tot, args.steps),
        "medium_events_per_sample": round(tot["medium_events"] / max(1.0, tot["samples"]), 3),
        "medium_roofline": medium_roofline(tot),
        "trace_variant": int(tot["trace_variant"] / max(1, args.steps)),
        "cpu_baseline": None,
    }
    print(json.dumps(result), flush=True)
    if args.out:
        with open(args.out, "w") as f:
            f.write(json.dumps(result) + "\n")


def batch8(args):
    """Eight 1080p x 64 spp frames (camera poses yawed -14..14 degrees) in
    one slrh_render_batch call per step."""
    import slr_amd
    from slr_amd import scenes
    desc, ctx, sc = _scene_c2()
    W, H, spp = 1920, 1080, 64
    frames = [dict(film=slr_amd.Film(ctx, W, H), camera=scenes.camera_pose(desc, yaw_deg=-14 + 4 * i), seed=17 + i)
              for i in range(8)]
    run_batch(args, sc, frames, slr_amd.PTRenderer(spp), 10, {"frames": 8, "width": W, "height": H, "spp": spp,
                                                               "max_path_length": 10},
              "8 frames x 1920x1080 x 64 spp of the cornell+displaced icosphere scene in one wavefront "
              "(slrh_render_batch)", "synthetic: seeded procedural scene, 8 camera poses")


def c3(args):
    """256 instances of four meshes under a sky, 8 x 1080p x 64 spp frames
    in one wavefront, depth 12."""
    import slr_amd
    from slr_amd import scenes
    t0 = time.time()
    W, H, spp = 1920, 1080, 64
    desc = scenes.instanced_field(W, H)
    ctx = slr_amd.Context(0)
    sc = slr_amd.build_scene(ctx, desc)
    ntris = sum(len(p["idx"]) for p in desc["prototypes"])
    log("scene built in %.1fs: %d prototype triangles, %d instances" % (time.time() - t0, ntris,
                                                                          len(desc["instances"])))
    frames = [dict(film=slr_amd.Film(ctx, W, H), camera=scenes.camera_pose(desc, yaw_deg=-7 + 2 * i), seed=31 + i)
              for i in range(8)]
    run_batch(args, sc, frames, slr_amd.PTRenderer(spp), 12,
              {"frames": 8, "width": W, "height": H, "spp": spp, "max_path_length": 12, "instances": 256,
               "prototype_triangles": ntris, "environment": "2048x1024 sky gradient + sun 1e4"},
              "256 affine instances of {displaced icosphere 1.31M, terrain 200k, torus knot 50k, box 12} with "
              "Lambert/GGX conductor/GGX dielectric/Oren-Nayar materials under an equirectangular sky, "
              "8 frames x 1920x1080 x 64 spp in one wavefront (slrh_render_batch)",
              "synthetic: seeded instances, materials, meshes and sky", inst_note=True)


def c4(args):
    """The Cornell box around a 256^3 density grid, volumetric integrator,
    4 x 1024^2 x 128 spp frames in one wavefront, depth 32."""
    import slr_amd
    from slr_amd import scenes
    t0 = time.time()
    W = H = 1024
    spp = 128
    desc = scenes.cornell_medium(W, H, n=256, sigma_t=20.0, albedo=0.9, g=0.7)
    log("density grid generated in %.1fs" % (time.time() - t0))
    ctx = slr_amd.Context(0)
    sc = slr_amd.build_scene(ctx, desc)
    frames = [dict(film=slr_amd.Film(ctx, W, H), camera=scenes.camera_pose(desc, yaw_deg=-6 + 4 * i), seed=41 + i)
              for i in range(4)]
    run_batch(args, sc, frames, slr_amd.VolumetricPTRenderer(spp), 32,
              {"frames": 4, "width": W, "height": H, "spp": spp, "max_path_length": 32, "grid": "256^3 f32",
               "sigma_t": 20.0, "albedo": 0.9, "g": 0.7, "majorant_supervoxel": 8},
              "cornell box around a 256^3 fBm density grid (sigma_t 20, albedo 0.9, HG g 0.7), volumetric path "
              "tracing, 4 frames x 1024x1024 x 128 spp in one wavefront (slrh_render_batch)",
              "synthetic: seeded fBm density grid")


def serve(args):
    """Serving: Poisson arrivals of seeded requests on the resident scene."""
    import slr_amd
    from slr_amd import scenes
    desc, ctx, sc = _scene_c2()
    if args.mix == "small":
        reqs = scenes.random_requests(desc, args.requests, seed=3, sizes=((256, 256), (512, 512)), spps=(4, 16))
        mix = "resolutions 256^2/512^2, spp 4/16"
    else:
        reqs = scenes.random_requests(desc, args.requests, seed=3)
        mix = "resolutions 256^2..1920x1080, spp 4/16/64"
    total_samples = sum(q["width"] * q["height"] * q["spp"] for q in reqs)
    rng = np.random.RandomState(11)
    gaps = rng.exponential(1.0 / args.rate, size=len(reqs)) if args.rate > 0 else np.zeros(len(reqs))
    films = [slr_amd.Film(ctx, q["width"], q["height"]) for q in reqs]

    # plain GPU: the same requests one at a time, each a full render of its
    # own (a one-frame slrh_render_batch: slrh_render with the request's camera)
    t0 = time.perf_counter()
    for q, f in zip(reqs, films):
        slr_amd.PTRenderer(q["spp"]).render_batch(sc, [dict(film=f, camera=q["camera"], seed=q["seed"])],
                                                  max_depth=10)
    r_plain = time.perf_counter() - t0
    log("plain: %d requests one at a time in %.2fs" % (len(reqs), r_plain))

    def run_server(gaps):
        for f in films:
            f.clear()
        srv = slr_amd.Server(sc, path_state_bytes=8 << 30)
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
        assert all(x == slr_amd.REQ_DONE for x in states)
        return dt, st

    # saturated: every request submitted at once (served throughput)
    t_sat, st_sat = run_server(np.zeros(len(reqs)))
    log("served, all submitted at once: %.2fs" % t_sat)
    # Poisson arrivals at the given rate (latency under load)
    t_poi, st_poi = run_server(gaps)
    log("served, Poisson %.1f/s: %.2fs, p50 %.1f ms, p99 %.1f ms" % (args.rate, t_poi, st_poi["p50_ms"],
                                                                     st_poi["p99_ms"]))
    samples_s = total_samples / t_sat
    result = {
        "metric": METRIC, "value": round(samples_s, 1), "unit": "samples/s", "n_gpus": 1, "steps": 1,
        "warmup": 0, "ms_per_step": round(t_sat * 1e3, 2), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32 (film f64)",
        "data": "synthetic: seeded requests (pose, resolution, spp) on the procedural scene",
        "config": {"workload": "serving %d requests (%s, depth 10) on the resident cornell+displaced icosphere "
                               "scene" % (len(reqs), mix),
                   "requests": len(reqs), "rate": args.rate, "mix": args.mix},
        "served_samples_per_s": round(samples_s, 1),
        "plain_samples_per_s": round(total_samples / r_plain, 1),
        "served_over_plain": round(r_plain / t_sat, 3),
        "poisson": {"rate_per_s": args.rate, "seconds": round(t_poi, 3), "p50_ms": round(st_poi["p50_ms"], 2),
                    "p99_ms": round(st_poi["p99_ms"], 2), "mean_ms": round(st_poi["mean_ms"], 2)},
        "saturated": {"p50_ms": round(st_sat["p50_ms"], 2), "p99_ms": round(st_sat["p99_ms"], 2)},
    }
    print(json.dumps(result), flush=True)
    if args.out:
        with open(args.out, "w") as f:
            f.write(json.dumps(result) + "\n")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c2", choices=["c1", "c2", "c3", "c4", "batch8", "serve"])
    ap.add_argument("--rate", type=float, default=16.0, help="serve: request arrivals per second")
    ap.add_argument("--requests", type=int, default=48, help="serve: number of requests")
    ap.add_argument("--mix", default="full", choices=["full", "small"],
                    help="serve: request sizes (full: 256^2..1080p x 4/16/64 spp; small: 256^2/512^2 x 4/16 spp)")
    ap.add_argument("--spp", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample", default="1920x1080x8")
    ap.add_argument("--out", default="")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="first run one device-to-device copy of this many MiB (a kernel of known traffic that "
                         "scales the FETCH_SIZE / WRITE_SIZE counters of a profiling run)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (frames, requests or tiles would be "
              "sharded one per GPU with no exchange); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import slr_amd
    from slr_amd import scenes

    if args.calib_copy_mib > 0:
        calibration_copy(args.calib_copy_mib << 20)
        log("calibration copy: %d MiB device to device" % args.calib_copy_mib)
    if args.config == "serve":
        return serve(args)
    if args.config == "batch8":
        return batch8(args)
    if args.config == "c3":
        return c3(args)
    if args.config == "c4":
        return c4(args)

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
    if args.spp:
        spp = args.spp
    ntris = sum(len(m["idx"]) for m in desc["meshes"])
    log("scene generated: %d triangles in %.1fs" % (ntris, time.time() - t0))

    ctx = slr_amd.Context(0)
    t0 = time.time()
    sc = slr_amd.build_scene(ctx, desc)
    build_s = time.time() - t0
    st_acc = sc.accel_stats()
    log("scene built (SBVH+QBVH+upload) in %.2fs: %s" % (build_s, st_acc))
    film = slr_amd.Film(ctx, W, H)
    renderer = slr_amd.PTRenderer(spp)

    for i in range(args.warmup):
        film.clear()
        renderer.render(sc, film, seed=1000 + i, max_depth=depth)
        log("warmup %d done" % i)

    # timed region: K full frames, device synchronised on both sides
    # (slrh_render returns after its stream completes)
    tot = {}
    t_start = time.perf_counter()
    for i in range(args.steps):
        film.clear()
        st = renderer.render(sc, film, seed=i, max_depth=depth, time_kernels=True)
        for k, v in st.items():
            tot[k] = tot.get(k, 0.0) + v
        log("step %d: %.3fs device, %d rays" % (i, st["device_seconds"], st["rays"]))
    elapsed = time.perf_counter() - t_start

    rays = tot["rays"]
    samples = tot["samples"]
    mrays = rays / elapsed / 1e6
    sps = samples / elapsed
    rays_per_sample = rays / samples

    result = {
        "metric": METRIC,
        "value": round(mrays, 2),
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
        "samples_per_s": round(sps, 1),
        "rays_per_sample": round(rays_per_sample, 3),
        "roofline": trace_roofline(tot, args.config),
        "frame_ms_breakdown": frame_breakdown(tot, args.steps),
        "accel": {"build_seconds": round(st_acc["build_seconds"], 3), "qbvh_nodes": int(st_acc["qbvh_nodes"]),
                  "qbvh_refs": int(st_acc["qbvh_refs"]), "qbvh_sah": round(st_acc["qbvh_cost"], 3)},
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        w, h, s = [int(x) for x in args.cpu_sample.split("x")]
        try:
            # the reference caps paths at 100 (PTRenderer.cpp:171-173): its
            # rays per sample are counted on the GPU with that cap, untimed
            f2 = slr_amd.Film(ctx, w, h)
            st100 = renderer.render(sc, f2, seed=77, max_depth=100)
            result["cpu_baseline"] = cpu_baseline(desc, (w, h, s), st100["rays"] / st100["samples"])
        except Exception as e:  # the baseline is reported, never required
            log("cpu baseline failed: %r" % (e,))
    line = json.dumps(result)
    print(line, flush=True)
    if args.out:
        with open(args.out, "w") as f:
            f.write(line + "\n")


if __name__ == "__main__":
    main()
