#!/usr/bin/env python3
"""Benchmark: wavefront path tracing on the 1.31M-triangle displaced icosphere.

Workload (one GPU): level-8 icosphere (1,310,720 triangles + ground + light),
gold GGX per region, quad area light + point light, 1024 x 1024, depth 8, RR
after bounce 3, Sobol.  A step is one wavefront pass of `--spp-per-step`
samples per pixel over the whole image: the default 64 spp (the
configuration's full sample count) puts 67,108,864 camera paths in flight,
about 16 GB of wave state, and the default 2 steps render the configuration
twice.  The scene (BVH, triangles, shading data, Sobol table) is resident in
HBM before timing.

Prints one JSON line: paths/s as `value`, rays/s beside it, the roofline of
the dominant kernel (closest-hit traversal, HBM bound) from HIP events on its
own stream, and the CPU baseline (the oracle -- the C++ restatement of
orion's CPU path with its own BVH in place of Embree) timed on this host.
"""
import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "paths/s + rays/s at 1.3M tris; served-render p50/p99 latency vs plain GPU"
HBM_PEAK_GBS = 8000.0
NODE_BYTES = 64      # 4-wide node: origin, grid steps, 4 child links, 8-bit child boxes
TRI_BYTES = 48       # p0, e1, e2 + primitive id
RAY_IO_BYTES = 48    # 32 B ray in (o, d, tmax, id) + 16 B hit out (t, prim, u, v)
SHADOW_IO_BYTES = 36  # 32 B ray in + 4 B visibility out
SHADE_BYTES = 292     # per shaded path segment
# SPPM photon shading: per photon hit (queue entry 4, hit 12, shading record
# 80, direction 12, beta 12, dimension 4 read; ray 24, beta 12, dimension 4,
# queue 4 written), per in-bounds cell lookup (start, end), per visible point
# scanned (pixel 4, radius 4, position 12), per deposit (material 4, frame 48,
# wo 12 read; Phi 12 + M 4 atomically added)
PHOTON_SEG_BYTES = 168
CELL_RANGE_BYTES = 8
VP_SCAN_BYTES = 20
DEPOSIT_BYTES = 80
HIT_RECORD_BYTES = 72  # a photon hit kept for its iteration's deposits: 36 B written, 36 B read
# a bounce's traversal: the closest-hit and any-hit kernels, launched side by
# side on two streams (one fork / join per bounce)
TRACE_KERNELS = ("k_trace<false, 16, false>", "k_trace<true, 16, false>")
TRAFFIC_FILE = ROOT / "profiles" / "pmc_traffic_icosphere_L8_1024x1024_64spp.json"
SPPM_TRAFFIC_FILE = ROOT / "profiles" / "pmc_traffic_sppm_caustics_512x512_2e20_photons.json"


def load_traffic(a):
    """HBM bytes per dispatch measured with rocprofv3 --pmc (FETCH_SIZE /
    WRITE_SIZE, scaled by a 256 MiB copy) for this exact configuration, or
    None when the configuration differs or the measurement is absent."""
    if not (a.level == 8 and a.res == 1024 and a.depth == 8 and a.spp_per_step == 64):
        return None
    try:
        d = json.loads(TRAFFIC_FILE.read_text())
    except (OSError, ValueError):
        return None
    d["source"] = f"{TRAFFIC_FILE.relative_to(ROOT)} (rocprofv3 --pmc FETCH_SIZE x " \
                  f"{d['copy_calibration']['fetch_scale']}, WRITE_SIZE x " \
                  f"{d['copy_calibration']['write_scale']}, per dispatch)"
    return d


def traversal_traffic(traffic):
    """A bounce's traversal traffic from the rocprofv3 summary: HBM bytes of
    one dispatch of each traversal kernel summed, their fetch-weighted L2 hit
    rate and each kernel's VALU busy; None without both kernels' figures."""
    ts = [traffic["kernels"].get(k) for k in TRACE_KERNELS] if traffic else [None]
    if not all(ts):
        return None
    fetch = [x["fetch_kib"] for x in ts]
    return {"hbm_bytes_per_dispatch": sum(x["hbm_bytes_per_dispatch"] for x in ts),
            "l2_hit_rate": round(sum(x["l2_hit_rate"] * f for x, f in zip(ts, fetch)) / max(sum(fetch), 1), 4),
            "valu_busy": {k: x.get("valu_busy") for k, x in zip(TRACE_KERNELS, ts)}}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--spp-per-step", type=int, default=64)
    ap.add_argument("--level", type=int, default=8)
    ap.add_argument("--res", type=int, default=1024)
    ap.add_argument("--depth", type=int, default=8)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--cpu-seconds", type=float, default=12.0,
                    help="target CPU-baseline work (bounded sample)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--batch-jobs", type=int, default=64, help="batched-launch leg (0 = skip)")
    ap.add_argument("--batch-ceiling", action="store_true",
                    help="batched leg: also time every job in full default waves (the rate batching can reach)")
    ap.add_argument("--serve-requests", type=int, default=24, help="serving leg (0 = skip)")
    ap.add_argument("--serve-policy", choices=["pack", "aged", "arrival", "fair"], default="pack",
                    help="the server's wave order: least remaining work first (optionally credited for "
                         "waiting time), submission order, or fair share")
    ap.add_argument("--serve-load", type=float, default=0.9,
                    help="offered load of the Poisson stream, as a fraction of the plain path's "
                         "one-at-a-time capacity")
    ap.add_argument("--hbm-budget-gib", type=float, default=16.0)
    ap.add_argument("--sppm-iters", type=int, default=64, help="SPPM caustics leg (0 = skip)")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="first run one device-to-device copy of this size (a known-bytes "
                         "kernel to scale the FETCH_SIZE / WRITE_SIZE counters by)")
    return ap.parse_args()


def cpu_baseline(hs, depth, target_s):
    """The oracle on this host's cores over whole 1 spp passes of the image
    until ~target_s seconds of work (at most 64 spp)."""
    sys.path.insert(0, str(ROOT / "tests"))
    from oracle_lib import Oracle
    o = Oracle()
    t0 = time.perf_counter()
    sc = o.scene(hs)
    build_s = time.perf_counter() - t0
    threads = os.cpu_count() or 1
    spp = 0
    t0 = time.perf_counter()
    rays = 0
    while True:
        _, _, ctr = sc.render_path(1, depth, k0=spp, threads=threads)
        spp += 1
        rays += int(ctr[0] + ctr[1])
        el = time.perf_counter() - t0
        if el >= target_s or spp >= 64:
            break
    paths = spp * hs.width * hs.height
    return {
        "value": paths / el, "unit": "paths/s", "cores": threads, "kind": "port",
        "sample": f"oracle (C++ restatement of orion's PathTracing::render, own BVH instead of "
                  f"Embree) on the same scene: {spp} spp x {hs.width}x{hs.height} = {paths} paths, "
                  f"all rows, OpenMP over rows, {el:.1f} s wall; oracle BVH build {build_s:.1f} s "
                  f"not timed",
        "rays_per_s": rays / el,
    }


def c4_job_specs(n_jobs, seed):
    """The batched-launch mix: per job (triangles, width, height, spp, depth,
    scene seed): displaced icosphere of 1k-500k triangles, log-uniform; film
    sides uniform in [128, 1024]; spp uniform in [4, 64]; depth 5-8."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n_jobs):
        tris = float(np.exp(rng.uniform(np.log(1e3), np.log(5e5))))
        W, H = (int(x) for x in rng.integers(128, 1025, 2))
        spp, depth = int(rng.integers(4, 65)), int(rng.integers(5, 9))
        out.append((tris, W, H, spp, depth, seed * 1000 + i))
    return out


def c5_scene_specs(seed):
    """The serving leg's four resident scenes: (triangles, scene seed),
    displaced icospheres of 5.3M, 1.4M, 1.4M and 0.34M target triangles."""
    return [(tris, seed * 100 + i) for i, tris in enumerate([5.3e6, 1.4e6, 1.4e6, 3.4e5])]


def c5_request_specs(n_req, seed):
    """Serving requests: (scene index, width, height, spp, depth), film sides
    uniform in [256, 1024], spp uniform in [16, 256], depth 5-8; the rng
    also draws the arrival gaps afterwards (returned)."""
    rng = np.random.default_rng(seed)
    reqs = []
    for _ in range(n_req):
        k = int(rng.integers(0, 4))
        W, H = (int(x) for x in rng.integers(256, 1025, 2))
        reqs.append((k, W, H, int(rng.integers(16, 257)), int(rng.integers(5, 9))))
    return reqs, rng


def batch_leg(oa, dev, n_jobs, seed, ceiling=False):
    """Batched launch of n_jobs seeded procedural jobs (displaced icosphere of
    1k-500k triangles, log-uniform; film sides uniform in [128, 1024]; spp
    uniform in [4, 64]; depth 5-8; material drawn per mesh) against the plain
    path rendering the same jobs one at a time.  The tuning comes from the
    autotuner on the same mix."""
    jobs, scenes = [], []
    t0 = time.perf_counter()
    for tris, W, H, spp, depth, sseed in c4_job_specs(n_jobs, seed):
        ds = oa.DeviceScene(dev, oa.HostScene.generate("batch_job", seed=sseed, params=[tris, W, H, 3]))
        scenes.append(ds)
        jobs.append((ds, W, H, spp, depth))
    setup_s = time.perf_counter() - t0
    paths = sum(W * H * spp for _, W, H, spp, _ in jobs)
    tris_tot = sum(ds.stats()["n_tris"] for ds in scenes)
    # plain: one job at a time through orion_render_path (render + film
    # readback, as the batched call returns the films), (a) one sample pass
    # over the film per wave -- orion's own loop order, PathTracing::render
    # -- and (b) this library's default multi-sample waves
    for ds, W, H, spp, depth in jobs[:2]:  # warm
        ds.render_path(0, 1, depth)
    t_plain1 = 0.0
    for ds, W, H, spp, depth in jobs:
        ds.film_clear()
        t0 = time.perf_counter()
        ds.render_path(0, spp, depth, spp_per_wave=1)
        ds.film()
        t_plain1 += time.perf_counter() - t0
    plain_films = []
    t_plain = 0.0
    for ds, W, H, spp, depth in jobs:
        ds.film_clear()
        t0 = time.perf_counter()
        ds.render_path(0, spp, depth)
        plain_films.append(ds.film()[0])
        t_plain += time.perf_counter() - t0
    t0 = time.perf_counter()
    tune = oa.autotune(dev, jobs)
    tune_s = time.perf_counter() - t0
    oa.render_batch(dev, jobs[:4], tuning=tune)  # warm
    t0 = time.perf_counter()
    outs, st = oa.render_batch(dev, jobs, tuning=tune)
    t_batch = time.perf_counter() - t0
    equal = sum(int(np.array_equal(a.view(np.int32), b.view(np.int32)))
                for a, (b, _) in zip(plain_films, outs))
    ceil = None
    if ceiling:
        # each job's own scene, film and depth rendered with enough samples to
        # fill a default wave: the job's paths at that full-wave rate are the
        # time a batcher that packs every job into full waves would need
        t_ideal = 0.0
        for ds, W, H, spp, depth in jobs:
            s_full = max(spp, (1 << 26) // (W * H))
            ds.film_clear()
            dev.synchronize()
            t0 = time.perf_counter()
            ds.render_path(0, s_full, depth)
            dev.synchronize()
            dt = time.perf_counter() - t0
            t_ideal += dt * spp / s_full
        ceil = {"method": "per job: the same scene, film and depth with max(spp, 2^26 / pixels) samples "
                          "(at least one full default wave), its time scaled to the job's spp; summed",
                "full_wave_paths_per_s": round(paths / t_ideal, 1),
                "speedup_ceiling_vs_plain": round(t_plain / t_ideal, 3),
                "batched_fraction_of_ceiling": round(t_ideal / t_batch, 3)}
    out = {
        "workload": f"{n_jobs} jobs, displaced icosphere 1k-500k tris log-uniform "
                    f"({tris_tot} tris total), films 128..1024 per side, 4-64 spp, depth 5-8, "
                    f"Lambert/GGX/dielectric per mesh",
        "paths": int(paths),
        "batched_paths_per_s": round(paths / t_batch, 1),
        "plain": "one job at a time through orion_render_path with its default waves (as many "
                 "samples per wave as fit 2^26 paths), film read back per job",
        "plain_paths_per_s": round(paths / t_plain, 1),
        "speedup": round(t_plain / t_batch, 3),
        "batched_s": round(t_batch, 3), "plain_s": round(t_plain, 3),
        "plain_one_pass_per_wave": "one job at a time, one sample pass over the film per wave "
                                   "(orion's loop order, PathTracing::render)",
        "plain_one_pass_per_wave_paths_per_s": round(paths / t_plain1, 1),
        "speedup_vs_plain_one_pass_per_wave": round(t_plain1 / t_batch, 3),
        "plain_one_pass_per_wave_s": round(t_plain1, 3),
        "batched_rays_per_s": round((st.closest_rays + st.shadow_rays) / t_batch, 1),
        "films_bitwise_equal_to_plain": f"{equal}/{n_jobs}",
        "tuning": {"wave_paths": tune.wave_paths, "rays_per_wave": tune.rays_per_wave,
                   "lds_ring": tune.lds_ring, "queue_segment": tune.queue_segment,
                   "use_graph": tune.use_graph},
        "autotune_s": round(tune_s, 2), "setup_s": round(setup_s, 2),
    }
    if ceil:
        out["ceiling"] = ceil
    return out


def sppm_traffic(kernel):
    """HBM bytes per dispatch of an SPPM kernel from the committed rocprofv3
    --pmc summary of this configuration, or None when absent."""
    try:
        return json.loads(SPPM_TRAFFIC_FILE.read_text())["kernels"][kernel]["hbm_bytes_per_dispatch"]
    except (OSError, ValueError, KeyError):
        return None


def sppm_leg(oa, dev, iters):
    """SPPM on the caustics scene: glass icosphere (81,920 triangles, eta 1.5)
    on a matte plane under a 30 degree spot light, 512x512, 2^20 photons per
    iteration, initial radius 0.01, 2^22 hash cells, depth 10."""
    hs = oa.HostScene.generate("caustics", seed=1, params=[6, 512, 512])
    ds = oa.DeviceScene(dev, hs)
    # warm-up: 8 iterations, two photon batches and two camera batches of the
    # timed run's size, so its traversal grids are planned from its first
    # batch on (as the path tracer's warm-up steps plan its grids)
    ds.render_sppm(8, 1 << 20, 0.01, 10, hash_size=1 << 22)
    ds.film_clear()
    dev.synchronize()
    t0 = time.perf_counter()
    st, _ = ds.render_sppm(iters, 1 << 20, 0.01, 10, hash_size=1 << 22, profile=True)
    el = time.perf_counter() - t0
    d = st.as_dict()
    # roofline of the dominant kernel: the photon bounces' closest-hit
    # traversal (k_trace_photons), HIP events around each launch on its stream
    tl = max(d["photon_trace_launches"], 1)
    trace_bytes = (d["photon_rays"] * RAY_IO_BYTES + d["node_visits_photon"] * NODE_BYTES
                   + d["tri_tests_photon"] * TRI_BYTES)
    t_launch = d["photon_trace_ms"] / 1e3 / tl
    achieved = trace_bytes / tl / t_launch / 1e9 if t_launch > 0 else 0.0
    rays = max(d["photon_rays"], 1)
    # photon shading (k_photon_bounce + k_photon_deposit): per photon segment
    # its state and shading record, per hit kept for the deposits its record
    # written and read back, per in-bounds lookup the cell range, per scanned
    # entry the visible point's pixel, radius and position, per deposit its
    # frame, direction and material index read and Phi, M added
    sl = max(d["photon_shade_launches"], 1)
    shade_bytes = (d["photon_rays"] * PHOTON_SEG_BYTES + d["photon_lookups"] * CELL_RANGE_BYTES
                   + d["photon_vp_scanned"] * VP_SCAN_BYTES + d["deposits"] * DEPOSIT_BYTES
                   + d["photon_hits"] * HIT_RECORD_BYTES)
    s_launch = d["photon_shade_ms"] / 1e3 / sl
    return {
        "workload": f"SPPM caustics, glass icosphere ({hs.n_tris} triangles) on a matte plane, "
                    f"spot light, 512x512, {iters} iterations x 2^20 photons, r0 0.01, "
                    f"2^22 hash cells, depth 10",
        "photons_per_s": round(st.photons / el, 1),
        "camera_paths_per_s": round(st.paths / el, 1),
        "ms_per_iteration": round(el * 1e3 / iters, 3),
        "rays_per_s": round((st.closest_rays + st.shadow_rays) / el, 1),
        "deposits": int(st.deposits),
        "seconds": round(el, 3),
        "roofline": {
            "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": sppm_traffic("k_trace_photons<16>"),
            "traffic_source": f"{SPPM_TRAFFIC_FILE.relative_to(ROOT)} (rocprofv3 --pmc FETCH_SIZE x 2.0, "
                              f"WRITE_SIZE x 1.0, per dispatch)",
            "kernel": "k_trace_photons<16>: one photon bounce's closest-hit traversal over a batch of "
                      "iterations' photons (10 bounces per batch)",
            "algorithmic_bytes_per_launch": round(trace_bytes / tl),
            "launch_ms": round(t_launch * 1e3, 4), "launches": int(tl),
            "photon_rays": int(d["photon_rays"]),
            "nodes_per_ray": round(d["node_visits_photon"] / rays, 2),
            "tris_per_ray": round(d["tri_tests_photon"] / rays, 2),
            "photon_shade": {
                "achieved": round(shade_bytes / sl / s_launch / 1e9, 1) if s_launch > 0 else 0.0,
                "unit": "GB/s", "launch_ms": round(s_launch * 1e3, 4),
                "algorithmic_bytes_per_launch": round(shade_bytes / sl),
                "lookups": int(d["photon_lookups"]), "vp_scanned": int(d["photon_vp_scanned"]),
                "bytes": f"{PHOTON_SEG_BYTES} B per photon segment, {HIT_RECORD_BYTES} B per hit kept for "
                         f"deposits, {CELL_RANGE_BYTES} B per cell lookup, {VP_SCAN_BYTES} B per visible point "
                         f"scanned, {DEPOSIT_BYTES} B per deposit",
                "photon_batch": int(d["photon_batch"])},
        },
    }


def serve_leg(oa, dev, n_req, load, budget_gib, seed, policy="pack", also=("aged", "arrival")):
    """Poisson stream of render requests against 4 resident scenes (~8.2M
    triangles: displaced icospheres of 5.2M, 1.3M, 1.3M and 0.33M triangles,
    material drawn per mesh); each request: film sides uniform in [256, 1024],
    spp uniform in [16, 256], depth 5-8.  The plain baselines render each
    request alone (one at a time, first come first served) -- one sample pass
    over the film per wave (orion's loop order), and this library's default
    multi-sample waves; their latencies follow from the measured service
    times on the same arrival schedule.  The arrival rate is `load` x the
    multi-sample plain path's capacity (the one-sample-per-wave plain path
    cannot keep up with it, so its queue grows over the stream)."""
    scenes = [oa.DeviceScene(dev, oa.HostScene.generate("batch_job", seed=sseed, params=[tris, 512, 512, 3]))
              for tris, sseed in c5_scene_specs(seed)]
    tris_tot = sum(ds.stats()["n_tris"] for ds in scenes)
    specs, rng = c5_request_specs(n_req, seed)
    reqs = [(scenes[k], W, H, spp, depth) for k, W, H, spp, depth in specs]
    paths = sum(W * H * spp for _, W, H, spp, _ in reqs)
    oa.render_batch(dev, [(scenes[0], 256, 256, 1, 5)])  # warm
    service, service1 = [], []
    for r in reqs:
        t0 = time.perf_counter()
        oa.render_batch(dev, [r])
        service.append(time.perf_counter() - t0)
        t0 = time.perf_counter()
        oa.render_batch(dev, [r], wave_paths=r[1] * r[2])  # one sample pass per wave
        service1.append(time.perf_counter() - t0)
    capacity = n_req / sum(service)          # requests/s, multi-sample plain
    qps = load * capacity
    gaps = rng.exponential(1.0 / qps, n_req)
    arrive = np.cumsum(gaps) - gaps[0]

    def fifo(svc):  # plain first-come-first-served on the same schedule
        free, lat = 0.0, []
        for a, sv in zip(arrive, svc):
            start = max(a, free)
            free = start + sv
            lat.append(free - a)
        return lat, free - arrive[0]

    lat_plain1, plain1_span = fifo(service1)
    lat_plain, plain_span = fifo(service)
    tune = oa.autotune(dev, [(ds, W, H, min(spp, 2), d) for ds, W, H, spp, d in reqs[:8]])

    def run_server(pol):  # the same requests on the same arrival schedule
        srv = oa.Server(dev, int(budget_gib * (1 << 30)), tuning=tune)
        srv.set_policy({"pack": oa.Server.PACK, "aged": oa.Server.AGED, "arrival": oa.Server.ARRIVAL,
                        "fair": oa.Server.FAIR}[pol])
        t0 = time.perf_counter()
        ids = []
        for a, r in zip(arrive, reqs):
            dt = t0 + a - time.perf_counter()
            if dt > 0:
                time.sleep(dt)
            ids.append(srv.submit(*r))
        srv.drain()
        lat, done = [], []
        for jid in ids:
            _, _, (ts, ta, tf) = srv.poll(jid, want_film=False)
            lat.append(tf - ts)
            done.append(tf)
        sst = srv.stats()
        srv.close()
        return lat, sst, max(done)

    others = {}
    for pol in also:
        if pol != policy:
            l2, s2, _ = run_server(pol)
            others[pol + "_order"] = {
                "p50_s": round(float(np.percentile(l2, 50)), 3), "p99_s": round(float(np.percentile(l2, 99)), 3),
                "speedup_busy": round(sum(service) / (s2["busy_ms"] / 1e3), 3),
                "waves": s2["waves"]}
    lat, sst, span = run_server(policy)
    return {
        "workload": f"{n_req} requests, Poisson at {qps:.2f}/s ({load:g} x multi-sample plain "
                    f"capacity), 4 resident scenes ({tris_tot} tris), films 256..1024 per side, "
                    f"16-256 spp, depth 5-8, HBM budget {budget_gib:g} GiB, wave order {policy}",
        "paths": int(paths),
        "p50_s": round(float(np.percentile(lat, 50)), 3),
        "p99_s": round(float(np.percentile(lat, 99)), 3),
        "plain": "one request at a time, first come first served, through the library's default "
                 "one-job path (multi-sample waves), on the same arrival schedule",
        "plain_p50_s": round(float(np.percentile(lat_plain, 50)), 3),
        "plain_p99_s": round(float(np.percentile(lat_plain, 99)), 3),
        "served_paths_per_s_busy": round(paths / (sst["busy_ms"] / 1e3), 1),
        "plain_paths_per_s": round(paths / sum(service), 1),
        "speedup_busy": round(sum(service) / (sst["busy_ms"] / 1e3), 3),
        "served_span_s": round(span, 3), "plain_span_s": round(plain_span, 3),
        "plain_one_pass_per_wave": "one request at a time, one sample pass over the film per wave "
                                   "(orion's loop order)",
        "plain_one_pass_per_wave_p50_s": round(float(np.percentile(lat_plain1, 50)), 3),
        "plain_one_pass_per_wave_p99_s": round(float(np.percentile(lat_plain1, 99)), 3),
        "plain_one_pass_per_wave_paths_per_s": round(paths / sum(service1), 1),
        "speedup_busy_vs_plain_one_pass_per_wave": round(sum(service1) / (sst["busy_ms"] / 1e3), 3),
        "plain_one_pass_per_wave_span_s": round(plain1_span, 3),
        "waves": sst["waves"], "mixed_waves": sst["mixed_waves"],
        "max_jobs_in_wave": sst["max_jobs_in_wave"],
        # the server caps its waves at 2^25 paths whatever the tuning asks
        "tuning": {"wave_paths_served": min(tune.wave_paths or (1 << 26), 1 << 25),
                   "rays_per_wave": tune.rays_per_wave, "lds_ring": tune.lds_ring,
                   "queue_segment": tune.queue_segment, "use_graph": tune.use_graph},
        **others,
    }


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py measures one GPU: every box in this setup has a single MI355X, so "
              f"--gpus {a.gpus} is not run (the multi-GPU design shards samples / pixels with one "
              f"film reduction per job)", file=sys.stderr)
        sys.exit(2)
    import orion_amd as oa
    if a.copy_calibration_mib > 0:
        import torch
        src = torch.ones(a.copy_calibration_mib << 18, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    t0 = time.perf_counter()
    hs = oa.HostScene.generate("icosphere", seed=a.seed, params=[a.level, a.res, a.res])
    gen_s = time.perf_counter() - t0
    dev = oa.Device(0)
    ds = oa.DeviceScene(dev, hs)
    sst = ds.stats()
    S = a.spp_per_step
    # warmup (different sample indices than the timed region)
    for w in range(a.warmup):
        ds.render_path(1000 + w * S, 1000 + (w + 1) * S, a.depth, spp_per_wave=S)
    ds.film_clear()
    dev.synchronize()
    tot = None
    t0 = time.perf_counter()
    for k in range(a.steps):
        st = ds.render_path(k * S, (k + 1) * S, a.depth, spp_per_wave=S, profile=True)
        d = st.as_dict()
        tot = d if tot is None else {key: tot[key] + d[key] for key in d}
    dev.synchronize()
    el = time.perf_counter() - t0
    print(f"# main leg done: {a.steps} steps in {el:.2f} s", file=sys.stderr, flush=True)
    paths = tot["paths"]
    value = paths / el
    rays = tot["closest_rays"] + tot["shadow_rays"]
    # roofline of a bounce's traversal (the closest-hit and any-hit kernels,
    # concurrent on two streams, timed from the fork to the join on the wave's
    # stream), the dominant work by device time
    kern = {"trace_pair": tot["trace_pair_ms"], "shade": tot["shade_ms"]}
    traffic = load_traffic(a)

    def ray_bytes(kind):
        io = RAY_IO_BYTES if kind == "closest" else SHADOW_IO_BYTES
        return (tot[f"{kind}_rays"] * io + tot[f"node_visits_{kind}"] * NODE_BYTES
                + tot[f"tri_tests_{kind}"] * TRI_BYTES)

    def per_ray(kind):
        rays = max(tot[f"{kind}_rays"], 1)
        return {"rays": tot[f"{kind}_rays"], "bytes_per_ray": round(ray_bytes(kind) / rays, 1),
                "nodes_per_ray": round(tot[f"node_visits_{kind}"] / rays, 2),
                "tris_per_ray": round(tot[f"tri_tests_{kind}"] / rays, 2)}

    launches = max(tot["trace_pair_launches"], 1)
    bytes_alg = ray_bytes("closest") + ray_bytes("shadow")
    per_launch_s = tot["trace_pair_ms"] / 1e3 / launches
    achieved = bytes_alg / launches / per_launch_s / 1e9 if per_launch_s > 0 else 0.0
    t = traversal_traffic(traffic)
    roofline = {
        "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": t["hbm_bytes_per_dispatch"] if t else None,
        "kernel": f"{' + '.join(TRACE_KERNELS)}: a bounce's closest-hit and any-hit traversal (4-wide "
                  f"quantized BVH, LDS ring stack), concurrent on two streams; one launch = one bounce's "
                  f"fork-to-join interval",
        "algorithmic_bytes_per_launch": round(bytes_alg / launches),
        "launch_ms": round(per_launch_s * 1e3, 4),
        "closest": per_ray("closest"), "shadow": per_ray("shadow"),
        "l2_hit_rate": t["l2_hit_rate"] if t else None,
        "valu_busy": t["valu_busy"] if t else None,
        "kernel_ms": {k: round(v, 2) for k, v in kern.items()},
    }
    if traffic:
        roofline["traffic_source"] = traffic["source"]
    # whole-path figure: every ray's traversal bytes plus B_shade per shaded
    # path segment (hit 16 + path state 48 + vertices 72 + material 32 + new
    # and shadow rays 64 + state write 48 + film 12 = 292 B), over the step time
    path_bytes = bytes_alg + tot["shaded_segments"] * SHADE_BYTES
    roofline["whole_path"] = {"achieved": round(path_bytes / el / 1e9, 1), "unit": "GB/s",
                              "frac": round(path_bytes / el / 1e9 / HBM_PEAK_GBS, 4),
                              "bytes_per_path": round(path_bytes / max(tot["paths"], 1), 1),
                              "shaded_segments": tot["shaded_segments"]}
    out = {
        "metric": METRIC, "value": round(value, 1), "unit": "paths/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(el * 1e3 / a.steps, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded displaced icosphere, generated in-process)",
        "config": {
            "workload": f"wavefront path tracing, displaced icosphere level {a.level} "
                        f"({hs.n_tris} triangles) on a ground plane, gold GGX, quad + point light, "
                        f"{a.res}x{a.res}, {S} spp per step, depth {a.depth}, RR after bounce 3",
            "triangles": hs.n_tris, "width": a.res, "height": a.res, "spp": S,
            "spp_rendered_in_timed_region": S * a.steps,
            "max_depth": a.depth, "bvh_nodes": sst["n_nodes"], "bvh_depth": sst["bvh_depth"],
            "stack_bound": sst["stack_bound"],
            "trace_waves_per_simd": sst["trace_waves_per_simd"],
        },
        "rays_per_s": round(rays / el, 1),
        "ref_equiv_rays_per_s": round(tot["ref_equiv_rays"] / el, 1),
        "closest_rays": tot["closest_rays"], "shadow_rays": tot["shadow_rays"],
        "roofline": roofline,
        "scene_gen_s": round(gen_s, 2), "bvh_build_ms": round(sst["build_ms"], 1),
    }
    if a.sppm_iters > 0:
        out["sppm"] = sppm_leg(oa, dev, a.sppm_iters)
        print("# sppm leg done", file=sys.stderr, flush=True)
    if a.batch_jobs > 0:
        out["batched"] = batch_leg(oa, dev, a.batch_jobs, a.seed, a.batch_ceiling)
        print(f"# batched leg done: {out['batched']['speedup']}x", file=sys.stderr, flush=True)
    if a.serve_requests > 0:
        out["serving"] = serve_leg(oa, dev, a.serve_requests, a.serve_load, a.hbm_budget_gib, a.seed,
                                   a.serve_policy)
        print("# serving leg done", file=sys.stderr, flush=True)
    if not a.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(hs, a.depth, a.cpu_seconds)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
