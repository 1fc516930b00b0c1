#!/usr/bin/env python3
"""Benchmark: wavefront path tracing on the 1.31M-triangle displaced icosphere.

Workload (one GPU): level-8 icosphere (1,310,720 triangles + ground + light),
gold GGX per region, quad area light + point light, 1024 x 1024, depth 8, RR
after bounce 3, Sobol.  A step is one wavefront pass of `--spp-per-step`
samples per pixel over the whole image (2 spp = 2,097,152 camera paths); the
default 32 steps render the configuration's full 64 spp.  The scene (BVH,
triangles, shading data, Sobol table) is resident in HBM before timing.

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

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "paths/s + rays/s at 1.3M tris; served-render p50/p99 latency vs plain GPU"
HBM_PEAK_GBS = 8000.0
NODE_BYTES = 64      # BVH2 node: both children's boxes + child links
TRI_BYTES = 48       # p0, e1, e2 + primitive id
RAY_IO_BYTES = 48    # 32 B ray in (o, d, tmax, id) + 16 B hit out (t, prim, u, v)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=32)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--spp-per-step", type=int, default=2)
    ap.add_argument("--level", type=int, default=8)
    ap.add_argument("--res", type=int, default=1024)
    ap.add_argument("--depth", type=int, default=8)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--cpu-seconds", type=float, default=12.0,
                    help="target CPU-baseline work (bounded sample)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
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


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py measures one GPU: every box in this setup has a single MI355X, so "
              f"--gpus {a.gpus} is not run (the multi-GPU design shards samples / pixels with one "
              f"film reduction; see DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    import orion_amd as oa

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
    paths = tot["paths"]
    value = paths / el
    rays = tot["closest_rays"] + tot["shadow_rays"]
    # roofline of the dominant kernel: closest-hit traversal
    kern = {"trace_closest": tot["trace_closest_ms"], "trace_shadow": tot["trace_shadow_ms"],
            "shade": tot["shade_ms"]}
    launches = max(tot["trace_closest_launches"], 1)
    bytes_alg = (tot["closest_rays"] * RAY_IO_BYTES + tot["node_visits_closest"] * NODE_BYTES
                 + tot["tri_tests_closest"] * TRI_BYTES)
    per_launch_bytes = bytes_alg / launches
    per_launch_s = tot["trace_closest_ms"] / 1e3 / launches
    achieved = per_launch_bytes / per_launch_s / 1e9 if per_launch_s > 0 else 0.0
    roofline = {
        "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
        "kernel": "k_trace<closest> (BVH2 traversal, LDS stack)",
        "bytes_per_ray": round(bytes_alg / max(tot["closest_rays"], 1), 1),
        "nodes_per_ray": round(tot["node_visits_closest"] / max(tot["closest_rays"], 1), 2),
        "tris_per_ray": round(tot["tri_tests_closest"] / max(tot["closest_rays"], 1), 2),
        "kernel_ms": {k: round(v, 2) for k, v in kern.items()},
    }
    out = {
        "metric": METRIC, "value": round(value, 1), "unit": "paths/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(el * 1e3 / a.steps, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded displaced icosphere, generated in-process)",
        "config": {
            "workload": f"wavefront path tracing, displaced icosphere level {a.level} "
                        f"({hs.n_tris} triangles) on a ground plane, gold GGX, quad + point light, "
                        f"{a.res}x{a.res}, {S} spp per step, depth {a.depth}, RR after bounce 3",
            "triangles": hs.n_tris, "width": a.res, "height": a.res, "spp": S * a.steps,
            "max_depth": a.depth, "bvh_nodes": sst["n_nodes"], "bvh_depth": sst["bvh_depth"],
        },
        "rays_per_s": round(rays / el, 1),
        "ref_equiv_rays_per_s": round(tot["ref_equiv_rays"] / el, 1),
        "closest_rays": tot["closest_rays"], "shadow_rays": tot["shadow_rays"],
        "roofline": roofline,
        "scene_gen_s": round(gen_s, 2), "bvh_build_ms": round(sst["build_ms"], 1),
    }
    if not a.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(hs, a.depth, a.cpu_seconds)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
