"""Frames/s of the per-frame splat path on one MI355X.

Sections (all run by default; --sections picks a subset):

  batched  (the headline `value`) 6,000,000 Gaussians of the clustered scene
           (64 anisotropic clusters + 20% on a ground plane, log-scales
           N(-5, 0.7)) stored in the f16 pod (f16 SH / rotation / scale, u8
           opacity), 8 views of 2560x1440 per batch orbiting the scene at
           radius 30, recorded once as a HIP graph over 4 queues and
           replayed; compositing tile autotuned per view, sort tile per batch.
           One step = one batch.
  plain    the same model and resolution through one Viewer, one frame per
           call: frames/s, per-frame latency p50/p99, per-stage GPU times,
           and the per-kernel timing the roofline figures come from (always
           run, whatever --sections names).
  mix      4 resident models of 0.5M/1M/3M/8M Gaussians (clustered scene,
           f16), 32 views per batch at resolutions drawn from
           {640x360, 1280x720, 1920x1080, 3840x2160}, cameras on spheres of
           radius 5-30, graph replay, compositing tile and sort tile
           autotuned per view.
  serve    8 resident models totalling ~40M Gaussians (sizes log-uniform
           1M-12M), a Poisson stream of 400 requests/s for --serve-seconds
           (then 1200 requests/s for --serve-burst-seconds, with several
           requests in flight at once),
           each request (model, camera on a smooth trajectory, resolution of
           the mix set); continuous batching with HBM-budgeted admission:
           p50/p99 latency (completion - arrival) and frames/s.
  sortscale (not run by default) the standalone RadixSorter on 5.2M, 20M
           and 80M random (key, value) pairs with 24, 20, 16, 12 and 8 keys
           per thread: per-sort wall time and algorithmic GB/s (16 B per pair
           per pass); the per-pass kernel times come from a kernel trace of
           this section.

Every frame runs the whole path: preprocess (cull + projection + depth keys)
-> depth radix sort -> tile binning -> tile sort -> compositing. Gaussians are
resident in HBM before any timed region. One JSON line on stdout; progress
on stderr. --gpus N > 1 exits: the viewer runs on one GPU.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# MI355X vector fp32 without packed (two-wide) fp32 instructions, which this
# build disables: half of the 157.3 TFLOP/s packed rate.
VALU_FP32_PEAK_TFLOPS = 78.6

# Algorithmic bytes of the projection kernel per Gaussian read (f16 pod:
# position+colour 16, packed f16 rotation/scale 16, f16 SH 96) and per visible
# Gaussian written (depth key 4 + index 4 + 32-byte splat record + 8-byte tile
# rectangle for the binning).
R_IN = {"single": 220, "half": 128}
R_OUT = 48
# Depth sort: 4 passes x (read key+value 8, write key+value 8) per key, plus
# the one histogram read of the keys (4).
SORT_BYTES_PER_KEY = 68
# Tile sort: 2 passes x (read key+value 8, write key+value 8) per (tile,
# splat) pair; its first histogram comes from the binning.
TILE_SORT_BYTES_PER_PAIR = 32
# Compositing: floating-point operations per blended fragment, a fused
# multiply-add counted as 2. Both blends: the two axis terms
# p.x u'.x + p.y u'.y - k (2 FMAs each, 8), r^2 from them (3), exp2 (1), times
# the opacity (1). The reference's blend (exact, the default): 1 - a (1),
# 255 a (1), per colour channel q (1 - a) + 255 c a (a multiply and an FMA,
# 3) and its rounding (1), the alpha channel's FMA and rounding (3): 30. The
# f32 blend (fast): weight T a (1), colour (3 FMAs, 6), transmittance (1): 21.
FLOPS_PER_FRAGMENT = {"exact": 30, "fast": 21}
# HBM bytes per launch on this workload come from the committed rocprofv3
# summary (FETCH_SIZE and WRITE_SIZE runs, each scaled by what the counter
# reads for a copy of known size); the profile names the kernels it measured.
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "traffic_clustered-6M-1440p-f16_plain.md")


def profile_traffic(kernel):
    """HBM bytes per launch of `kernel` (a substring of its name) from the
    traffic profile's table, and the profile's path; (None, path) when the
    profile has no row for it."""
    rel = os.path.relpath(TRAFFIC_PROFILE, ROOT)
    try:
        lines = open(TRAFFIC_PROFILE).read().splitlines()
    except OSError:
        return None, rel
    for ln in lines:
        cells = [c.strip() for c in ln.strip().strip("|").split("|")]
        if len(cells) >= 5 and kernel in cells[0] and cells[-1].endswith("MB"):
            return int(round(float(cells[-1].split()[0].replace(",", "")) * 1e6)), rel
    return None, rel

MIX_RES = [(640, 360), (1280, 720), (1920, 1080), (3840, 2160)]


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def orbit_pod(gs, w, h, azimuth, radius=30.0, elevation=0.3):
    eye = (radius * math.cos(elevation) * math.sin(azimuth), radius * math.sin(elevation),
           radius * math.cos(elevation) * math.cos(azimuth))
    return gs.CameraPod.look_at(eye, (0.0, 0.0, 0.0), (w, h), math.pi / 3, 0.1, 1e4)


def sphere_pod(gs, rng, w, h, rmin=5.0, rmax=30.0):
    r = rng.uniform(rmin, rmax)
    az = rng.uniform(0.0, 2.0 * math.pi)
    el = math.asin(rng.uniform(-0.9, 0.9))
    return orbit_pod(gs, w, h, az, r, el)


def percentile(xs, q):
    return round(float(np.percentile(np.asarray(xs), q)), 4) if len(xs) else None


# ---------------------------------------------------------------------------
# sections


def run_plain(gs, dev, g, w, h, steps, warmup, tile, blend="exact"):
    viewer = gs.Viewer(dev, gs.TextureFormat.RGBA8_UNORM, g, gs.GaussianPod.HALF)
    viewer.set_tile_size(*tile)
    viewer.set_blend_mode(gs.BlendMode.EXACT if blend == "exact" else gs.BlendMode.FAST)
    target = gs.Texture(dev, w, h, gs.TextureFormat.RGBA8_UNORM)
    pods = [orbit_pod(gs, w, h, 2.0 * math.pi * k / 64) for k in range(64)]
    for k in range(warmup):
        viewer.update_camera_with_pod(pods[k % 64])
        viewer.render(target)
    dev.poll()
    t0 = time.perf_counter()
    for k in range(steps):
        viewer.update_camera_with_pod(pods[k % 64])
        viewer.render(target)
    dev.poll()
    elapsed = time.perf_counter() - t0

    # per-frame latency and per-stage / per-kernel GPU times from HIP events
    # recorded on the viewer's stream
    viewer.set_profiling(True)
    lat, stages, vis, pairs = [], [], [], []
    for k in range(steps):
        viewer.update_camera_with_pod(pods[k % 64])
        t1 = time.perf_counter()
        viewer.render(target)
        dev.poll()
        lat.append((time.perf_counter() - t1) * 1e3)
        stages.append(viewer.stage_times())
        vis.append(int(viewer.indirect_args()[1]))
        pairs.append(viewer.pair_count())
    viewer.set_profiling(False)
    stage_ms = np.array(stages).mean(axis=0)

    # blended fragments of the same frames (the work the compositing roofline
    # is priced on), counted by the exact kernel in a separate pass
    frags = []
    for k in range(min(steps, 8)):
        viewer.update_camera_with_pod(pods[k % 64])
        frags.append(viewer.count_fragments(target))
    viewer.close()
    target.close()
    return {
        "fps": steps / elapsed, "ms_per_frame": 1e3 * elapsed / steps,
        "latency_ms": {"p50": percentile(lat, 50), "p99": percentile(lat, 99)},
        "stage_ms": stage_ms, "visible_mean": float(np.mean(vis)), "pairs_mean": float(np.mean(pairs)),
        "fragments_mean": float(np.mean(frags)),
    }


def run_batched(gs, dev, model, w, h, views, steps, warmup, queues):
    batch = gs.Batch(dev, [model], queues=queues)
    batch.set_layout([(0, w, h)] * views)
    ring = [orbit_pod(gs, w, h, 2.0 * math.pi * k / 64) for k in range(64)]
    pod_sets = [gs.pods_array([ring[(views * s + j) % 64] for j in range(views)]) for s in range(8)]
    batch.autotune([ring[j] for j in range(views)])
    tiles = [batch.tile_size(j) for j in range(views)]
    items = [batch.sort_items(j) for j in range(views)]
    # every camera of the timed loop once, twice over: tile lists that a
    # camera outgrows are enlarged by finish() before the timed region
    for _ in range(2):
        for k in range(8):
            batch.render(pod_sets[k])
        batch.finish()
    for k in range(warmup):
        batch.render(pod_sets[k % 8])
    batch.finish()
    t0 = time.perf_counter()
    for k in range(steps):
        batch.render(pod_sets[k % 8])
    batch.finish()
    elapsed = time.perf_counter() - t0
    batch.close()
    return {"fps": views * steps / elapsed, "ms_per_step": 1e3 * elapsed / steps, "tiles": tiles,
            "sort_items": items}


# The mixed-model batched workload: 4 resident models (clustered scene, f16),
# model i synthesised with seed MIX_SEED + i.
MIX_SIZES = [500_000, 1_000_000, 3_000_000, 8_000_000]
MIX_SEED = 100


def mix_workload(gs, seed=3, sets=4):
    """32 views (model, width, height) at resolutions of MIX_RES and `sets`
    camera sets on spheres of radius 5-30, drawn from one seeded stream."""
    rng = np.random.default_rng(seed)
    views = [(int(rng.integers(0, 4)), *MIX_RES[int(rng.integers(0, 4))]) for _ in range(32)]
    pod_sets = [gs.pods_array([sphere_pod(gs, rng, v[1], v[2]) for v in views]) for _ in range(sets)]
    return views, pod_sets


def run_mix(gs, dev, steps, warmup, queues, seed=3):
    sizes = MIX_SIZES
    models = []
    for i, n in enumerate(sizes):
        models.append(gs.Model(dev, gs.synth_gaussians(2, n, MIX_SEED + i), gs.GaussianPod.HALF))
    views, pod_sets = mix_workload(gs, seed)
    batch = gs.Batch(dev, models, queues=queues)
    batch.set_layout(views)
    batch.autotune([gs.CameraPod.from_bytes(bytes(p)) for p in pod_sets[0]])
    for _ in range(2):
        for k in range(4):
            batch.render(pod_sets[k])
        batch.finish()
    for k in range(warmup):
        batch.render(pod_sets[k % 4])
    batch.finish()
    t0 = time.perf_counter()
    for k in range(steps):
        batch.render(pod_sets[k % 4])
    batch.finish()
    elapsed = time.perf_counter() - t0
    tiles, items = {}, {}
    for j in range(len(views)):
        key = "%dx%d" % batch.tile_size(j)
        tiles[key] = tiles.get(key, 0) + 1
        key = str(batch.sort_items(j))
        items[key] = items.get(key, 0) + 1
    batch.close()
    for m in models:
        m.close()
    pixels = sum(v[1] * v[2] for v in views)
    return {"fps": len(views) * steps / elapsed, "ms_per_batch": 1e3 * elapsed / steps,
            "views": len(views), "megapixels_per_batch": round(pixels / 1e6, 2),
            "gaussians": sum(sizes), "tile_choices": tiles, "sort_items_choices": items}


def serve_requests(gs, rng, sizes, rate, seconds):
    """Poisson arrivals; each request picks a model, a resolution of the mix
    set and the camera of that model's trajectory at its arrival time."""
    n_req = int(rate * seconds * 1.2) + 16
    arrivals = np.cumsum(rng.exponential(1.0 / rate, n_req))
    arrivals = arrivals[arrivals < seconds]
    traj = [(rng.uniform(0, 2 * math.pi), rng.uniform(0.2, 0.6), rng.uniform(8.0, 25.0), rng.uniform(0.1, 0.4))
            for _ in sizes]
    req = np.zeros(len(arrivals), gs.REQUEST_DTYPE)
    for i, t in enumerate(arrivals):
        m = int(rng.integers(0, len(sizes)))
        w, h = MIX_RES[int(rng.integers(0, 4))]
        a0, el0, r0, om = traj[m]
        az = a0 + om * t
        el = el0 + 0.15 * math.sin(0.7 * t + m)
        r = r0 + 3.0 * math.sin(0.3 * t + 2 * m)
        pod = orbit_pod(gs, w, h, az, r, el)
        req[i]["model"], req[i]["width"], req[i]["height"] = m, w, h
        req[i]["arrival_s"] = t
        req[i]["pod"] = np.frombuffer(pod.to_bytes(), np.uint8)
    return req


# The serving workload's models: model i synthesised with seed SERVE_SEED + i.
SERVE_SEED = 200


def serve_model_sizes(rng):
    """8 model sizes log-uniform in 1M-12M, scaled to total 40M Gaussians."""
    raw = np.exp(rng.uniform(math.log(1e6), math.log(12e6), 8))
    return [int(x) for x in np.round(raw * (40e6 / raw.sum()))]


def run_serve(gs, dev, seconds, rate, queues, burst_rate=0.0, burst_seconds=0.0, seed=4):
    rng = np.random.default_rng(seed)
    sizes = serve_model_sizes(rng)
    models = []
    for i, n in enumerate(sizes):
        models.append(gs.Model(dev, gs.synth_gaussians(2, n, SERVE_SEED + i), gs.GaussianPod.HALF))
    log(f"serve: models resident ({sum(sizes) / 1e6:.1f}M Gaussians)")
    req = serve_requests(gs, rng, sizes, rate, seconds)
    server = gs.Server(dev, models, queues=queues)
    # warm every (model, resolution) pair once outside the measured stream
    warm = np.zeros(len(sizes) * len(MIX_RES), gs.REQUEST_DTYPE)
    for i in range(len(warm)):
        m, (w, h) = i // len(MIX_RES), MIX_RES[i % len(MIX_RES)]
        warm[i]["model"], warm[i]["width"], warm[i]["height"] = m, w, h
        warm[i]["pod"] = np.frombuffer(orbit_pod(gs, w, h, 0.3 * i).to_bytes(), np.uint8)
    server.run(warm)
    stats, _ = server.run(req)
    if burst_rate > 0.0 and burst_seconds > 0.0:
        # a denser stream on the same resident models: requests overlap on
        # the device (several in flight), which the 400/s stream rarely needs
        burst = serve_requests(gs, np.random.default_rng(seed + 1), sizes, burst_rate, burst_seconds)
        bstats, _ = server.run(burst)
        bstats["rate_req_per_s"] = burst_rate
        stats["burst"] = bstats
    server.close()
    for m in models:
        m.close()
    stats["gaussians"] = sum(sizes)
    stats["model_sizes"] = sizes
    stats["rate_req_per_s"] = rate
    return stats


SORTSCALE_SIZES = (5_200_000, 20_000_000, 80_000_000)
SORTSCALE_ITEMS = (24, 20, 16, 12, 8)


def run_sortscale(gs, dev, reps=3, seed=0):
    """The standalone sort at several sizes and tile shapes: best of `reps`
    sorts of the same random pairs (uploaded again before each)."""
    rng = np.random.default_rng(seed)
    out = []
    for n in SORTSCALE_SIZES:
        keys = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
        vals = np.arange(n, dtype=np.uint32)
        for items in SORTSCALE_ITEMS:
            s = gs.RadixSorter(dev, n)
            s.set_items(items)
            best = float("inf")
            for _ in range(reps):
                s.write(keys, vals)
                dev.poll()
                t0 = time.perf_counter()
                s.sort()
                dev.poll()
                best = min(best, time.perf_counter() - t0)
            s.close()
            out.append({"pairs": n, "keys_per_thread": items, "ms": round(best * 1e3, 4),
                        "gbs_per_pass": round(16 * n / (best / 4) / 1e9, 1)})
            log(f"sortscale: {n} pairs, {items} keys per thread: {best * 1e3:.3f} ms")
    return out


def cpu_baseline(gs, g, w, h, threads, budget_s):
    """The oracle (a CPU restatement of the reference's path, not the
    reference, which cannot be built here) on frames of the benchmark's own
    workload, until budget_s of wall time has been spent."""
    import oracle as orc

    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from helpers import oracle_inputs_half

    pos, color, sh, cov = oracle_inputs_half(orc, g)
    frames, t0 = 0, time.perf_counter()
    while True:
        pod = orbit_pod(gs, w, h, 2.0 * math.pi * frames / 64)
        frame = orc.frame_from_pod(pod.to_bytes())
        orc.full_frame(frame, pos, color, sh, cov, w, h, orc.TARGET_RGBA8_STEP, threads=threads)
        frames += 1
        el = time.perf_counter() - t0
        if el >= budget_s or frames >= 64:
            break
    try:
        host_cores = len(os.sched_getaffinity(0))
    except AttributeError:
        host_cores = os.cpu_count()
    return {"value": round(frames / el, 4), "unit": "frames/s", "cores": threads, "host_cores": host_cores,
            "kind": "port", "label": "CPU restatement, not the reference",
            "sample": f"{frames} full frame(s) of the benchmark's own 6M-Gaussian 2560x1440 f16 workload (orbit "
                      f"cameras; not the 65,536-Gaussian reference case) through the C oracle (oracle/oracle.c, "
                      f"OpenMP, {threads} threads on a host with {host_cores} available cores) in {el:.1f} s"}


# ---------------------------------------------------------------------------


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40, help="timed batches (batched) / frames (plain)")
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--sections", default="batched,plain,mix,serve")
    ap.add_argument("--views", type=int, default=8, help="views per batch in the batched section")
    ap.add_argument("--queues", type=int, default=4)
    ap.add_argument("--tile", default="16x16", help="tile size of the plain section")
    ap.add_argument("--blend", default="exact", choices=["fast", "exact"],
                    help="blend of the plain section (exact: the reference's per-draw unorm8 rounding, the default "
                         "of every form; fast: f32 front to back, also timed beside it)")
    ap.add_argument("--serve-seconds", type=float, default=60.0)
    ap.add_argument("--serve-rate", type=float, default=400.0)
    ap.add_argument("--serve-queues", type=int, default=8, help="device queues of the serving loop")
    ap.add_argument("--serve-burst-rate", type=float, default=1200.0,
                    help="requests/s of a second, denser stream after the first (0: none)")
    ap.add_argument("--serve-burst-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: the viewer runs on one GPU only (asked for {args.gpus}); run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)
    sections = [s for s in args.sections.split(",") if s]
    for s in sections:
        if s not in ("batched", "plain", "mix", "serve", "sortscale"):
            ap.error(f"unknown section {s}")

    import __graft_entry__

    gs = __graft_entry__.load_package()
    n, w, h = 6_000_000, 2560, 1440
    t_syn = time.perf_counter()
    g = gs.synth_gaussians(2, n, 2)
    log(f"synthesised {n} Gaussians in {time.perf_counter() - t_syn:.1f} s")
    dev = gs.Device(0)
    tile = tuple(int(x) for x in args.tile.split("x"))
    out_extra = {}

    # the plain section always runs: the rooflines and stage times come from it
    plain = run_plain(gs, dev, g, w, h, args.steps, args.warmup, tile, args.blend)
    log(f"plain ({args.blend} blend): {plain['fps']:.1f} frames/s")
    other = "fast" if args.blend == "exact" else "exact"
    plain_other = run_plain(gs, dev, g, w, h, args.steps, args.warmup, tile, other)
    log(f"plain ({other} blend): {plain_other['fps']:.1f} frames/s")
    batched = None
    if "batched" in sections:
        model = gs.Model(dev, g, gs.GaussianPod.HALF)
        batched = run_batched(gs, dev, model, w, h, args.views, args.steps, args.warmup, args.queues)
        model.close()
        log(f"batched: {batched['fps']:.1f} frames/s")
    if "mix" in sections:
        out_extra["mix"] = run_mix(gs, dev, max(args.steps // 4, 4), 2, args.queues)
        log(f"mix: {out_extra['mix']['fps']:.1f} frames/s")
    if "sortscale" in sections:
        out_extra["sortscale"] = run_sortscale(gs, dev)
    if "serve" in sections:
        out_extra["serve"] = run_serve(gs, dev, args.serve_seconds, args.serve_rate, args.serve_queues,
                                       args.serve_burst_rate, args.serve_burst_seconds)
        log(f"serve: p50 {out_extra['serve']['p50_ms']:.2f} ms p99 {out_extra['serve']['p99_ms']:.2f} ms")
        if "burst" in out_extra["serve"]:
            b = out_extra["serve"]["burst"]
            log(f"serve burst {b['rate_req_per_s']:.0f}/s: p50 {b['p50_ms']:.2f} ms p99 {b['p99_ms']:.2f} ms, "
                f"{b['mean_in_flight']:.2f} in flight")

    # rooflines from the plain section's HIP events (same model, resolution
    # and cameras as the batched section)
    stage_ms = plain["stage_ms"]
    v_mean = plain["visible_mean"]
    comp_ms = float(stage_ms[5])
    comp_flops = plain["fragments_mean"] * FLOPS_PER_FRAGMENT[args.blend]
    comp_tflops = comp_flops / (comp_ms * 1e-3) / 1e12
    # the exact blend of a cleared Rgba8Unorm target has its own instantiation
    comp_kernel = ("k_composite_px<true, false, 256, true>" if args.blend == "exact"
                   else "k_composite_px<false, false, 256, false>")
    comp_traffic, traffic_src = profile_traffic(comp_kernel)
    composite = {"bound": "valu", "achieved": round(comp_tflops, 3), "peak": VALU_FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                 "frac": round(comp_tflops / VALU_FP32_PEAK_TFLOPS, 4), "traffic": comp_traffic,
                 "traffic_source": traffic_src,
                 "kernel": comp_kernel,
                 "note": "vector fp32 (exp2 + blend per fragment) against the non-packed VALU peak; no MFMA, "
                         "compositing is not a contraction",
                 "avg_launch_ms": round(comp_ms, 4),
                 "fragments_per_launch": int(plain["fragments_mean"]),
                 "flops_per_fragment": FLOPS_PER_FRAGMENT[args.blend], "blend": args.blend}
    # the projection pass (k_preprocess), priced on the bytes it must move
    # per launch
    pre_ms = float(stage_ms[6])
    pre_bytes = n * R_IN["half"] + v_mean * R_OUT
    pre_gbs = pre_bytes / (pre_ms * 1e-3) / 1e9
    pre_traffic, _ = profile_traffic("k_preprocess")
    projection = {"bound": "hbm", "achieved": round(pre_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                  "frac": round(pre_gbs / HBM_PEAK_GBS, 4), "traffic": pre_traffic,
                  "traffic_source": traffic_src,
                  "kernel": "k_preprocess<1>",
                  "avg_launch_ms": round(pre_ms, 4),
                  "bytes_per_launch": int(pre_bytes),
                  "bytes_per_unit": {"gaussian_read": R_IN["half"], "visible_written": R_OUT},
                  "units_per_launch": {"gaussians": n, "visible": int(v_mean)}}
    # `roofline` is the kernel with the longer time per frame of the two that
    # dominate it (compositing and projection); both are reported
    roofline = max((composite, projection), key=lambda r: r["avg_launch_ms"])
    sort_gbs = SORT_BYTES_PER_KEY * v_mean / (float(stage_ms[1]) * 1e-3) / 1e9
    p_mean = plain["pairs_mean"]
    tile_gbs = TILE_SORT_BYTES_PER_PAIR * p_mean / (float(stage_ms[3]) * 1e-3) / 1e9
    stage_pre_gbs = pre_bytes / (float(stage_ms[0]) * 1e-3) / 1e9
    hbm = {"preprocess_stage": {"achieved_gbs": round(stage_pre_gbs, 1), "frac": round(stage_pre_gbs / HBM_PEAK_GBS, 4),
                                  "bytes_per_launch": int(pre_bytes), "ms": round(float(stage_ms[0]), 4),
                                "note": "k_frame_pre + k_preprocess"},
           "depth_sort": {"achieved_gbs": round(sort_gbs, 1), "frac": round(sort_gbs / HBM_PEAK_GBS, 4),
                          "bytes_per_sort": int(SORT_BYTES_PER_KEY * v_mean), "ms": round(float(stage_ms[1]), 4),
                          "passes": 4, "keys": int(v_mean)},
           "tile_sort": {"achieved_gbs": round(tile_gbs, 1), "frac": round(tile_gbs / HBM_PEAK_GBS, 4),
                         "bytes_per_sort": int(TILE_SORT_BYTES_PER_PAIR * p_mean), "ms": round(float(stage_ms[3]), 4),
                         "passes": 2, "pairs": int(p_mean),
                         "note": "2 passes x 16 B per pair (tile keys of a 160x90 grid: 14 bits), plus the range scan"}}

    names = ["preprocess", "sort", "bin", "tile_sort", "composite", "composite_main_kernel",
             "preprocess_main_kernel"]
    if batched is not None:
        value, ms_step, steps = batched["fps"], batched["ms_per_step"], args.steps
        form = (f"batched: {args.views} views per batch, HIP graph over {args.queues} queues, tile autotuned, "
                f"exact blend")
    else:
        value, ms_step, steps = plain["fps"], plain["ms_per_frame"], args.steps
        form = f"plain single view, {args.blend} blend"
    out = {
        "metric": "frames/s at 6M Gaussians 1440p (batched vs plain); p50/p99 served frame latency",
        "value": round(value, 2),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (clustered scene generator, seeded)",
        "config": {"workload": "clustered-6M-1440p-f16", "gaussians": n, "width": w, "height": h,
                   "storage": "f16 SH/rotation/scale, u8 opacity", "form": form,
                   "blend": "exact: back to front, the Rgba8Unorm target rounded after every splat (the reference's "
                            "fixed-function blend); the default of every form",
                   "visible_mean": int(v_mean)},
        "plain": {"fps": round(plain["fps"], 2), "ms_per_frame": round(plain["ms_per_frame"], 4),
                  "latency_ms": plain["latency_ms"], "tile": args.tile,
                  "stage_ms": {nm: round(float(x), 4) for nm, x in zip(names, stage_ms)}, "blend": args.blend},
        "plain_" + other + "_blend": {
            "fps": round(plain_other["fps"], 2), "ms_per_frame": round(plain_other["ms_per_frame"], 4),
            "latency_ms": plain_other["latency_ms"],
            "stage_ms": {nm: round(float(x), 4) for nm, x in zip(names, plain_other["stage_ms"])},
            "note": "fast: front to back in f32, rounded once; measured against the reference's per-draw rounding "
                    "(tests/test_gpu_parity.py::test_reference_blend_full_size): max 6/255, 51.8 dB on this frame"},
        "roofline": roofline,
        "kernel_rooflines": {"projection": projection, "compositing": composite},
        "hbm_rooflines": hbm,
    }
    if batched is not None:
        out["batched"] = {"fps": round(batched["fps"], 2), "ms_per_batch": round(batched["ms_per_step"], 4),
                          "views": args.views, "tiles": ["%dx%d" % t for t in batched["tiles"]],
                          "sort_items": batched["sort_items"]}
    out.update(out_extra)
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(gs, g, w, h, args.cpu_threads, args.cpu_budget)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
