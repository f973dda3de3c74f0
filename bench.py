"""Benchmark: Msamples/s of the GPU path tracer on one MI355X.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]

A step is one full render of the workload's image at its sample count,
inputs resident in HBM, output image in HBM.  Prints one JSON line.

Workloads:
  config1   icosphere(3) on a ground quad, sky cube map 6x64x64 with a
            sun disk, 256x256, 4 spp, 3 bounces, per-pixel PCG streams,
            the reference's estimator (ComputeRadianceForPath) -- the
            configuration whose CPU run is the reference's own workload.
  soup      closest hit only: 2^22 rays vs a 1M-triangle soup (Mrays/s).

The roofline entry is for the dominant kernel of the workload (the
closest-hit traversal, k_extend / k_closest_mesh): algorithmic bytes per
launch = per ray 24 (ray) + 20 (hit record) + 24 per slab test + 36 per
triangle test, the test counts taken from a counting run of the same
launches; divided by the mean launch time from HIP events on the launch
stream.  The cpu_baseline entry times the oracle (the C restatement of the
reference tracer) on a bounded sample of the same workload on this host.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Msamples/s at 1 GPU; time-to-relMSE 1e-3, guided vs plain GPU tracer"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline_config1(threads, spp=16):
    """The oracle's threaded tile renderer (64x64 tiles pulled from an atomic
    counter as src/main.cpp:731-844 does) on config 1 with the reference's
    own BVH and estimator, `spp` samples per pixel."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from _libs import (SPO_HIT_REFERENCE, SPO_RNG_PCG, SPO_TX_LIBM, SpoConfig,
                       oracle)
    from _scenes import load_oracle
    from vk_cinematic_amd import scenes

    desc = scenes.config1(256, 256, env="cube")
    O = oracle()
    s, cam = load_oracle(desc)
    cfg = SpoConfig(rng=SPO_RNG_PCG, tx=SPO_TX_LIBM, hit=SPO_HIT_REFERENCE,
                    estimator=0, samplesPerPixel=spp, bounces=3, radianceClamp=10.0,
                    jitterYFirst=1, seed=0x5EED)
    px = np.zeros(256 * 256 * 4, np.float32)
    O.spo_render_threaded(s, C.byref(cam), C.byref(cfg), 64, 64, threads, px,
                          np.zeros(12, np.uint64))  # warm-up
    secs = []
    for _ in range(3):
        secs.append(O.spo_render_threaded(s, C.byref(cam), C.byref(cfg), 64, 64,
                                          threads, px, np.zeros(12, np.uint64)))
    O.spo_scene_destroy(s)
    sec = float(np.median(secs))
    return {"value": 256 * 256 * spp / sec / 1e6, "unit": "Msamples/s",
            "cores": threads, "kind": "port",
            "sample": f"config1 256x256, {spp} spp, 3 bounces, reference BVH and "
                      f"estimator, median of 3 after 1 warm-up ({sec:.2f} s each)"}


def cpu_baseline_soup(threads, tris, o, d, n_rays=65536):
    """The oracle's restatement of the closest-hit traversal on the same BVH,
    all host threads, on the first n_rays rays."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from _libs import oracle
    import vk_cinematic_amd as V

    nodes, order, _ = V.bvh_build_host(tris)
    rays = np.ascontiguousarray(np.concatenate([o[:n_rays], d[:n_rays]], 1).ravel())
    p = np.zeros(n_rays, np.uint32)
    t = np.zeros(n_rays, np.float32)
    O = oracle()
    t0 = time.perf_counter()
    O.spo_bvh2_closest(np.ascontiguousarray(nodes.ravel()), order,
                       np.ascontiguousarray(tris.ravel()), rays, n_rays, threads, p, t,
                       None, None)
    sec = time.perf_counter() - t0
    return {"value": n_rays / sec / 1e6, "unit": "Mrays/s", "cores": threads,
            "kind": "port",
            "sample": f"{n_rays} of the 2^22 rays, same SAH BVH ({sec:.2f} s)"}


def bench_config1(args, V, scenes):
    dev = V.Device(0)
    desc = scenes.config1(256, 256, env="cube")
    s, cam = scenes.load_gpu(dev, desc)
    tiles = V.ComputeTiles(256, 256, 64, 64)
    spp, bounces = 4, 3
    img = dev.alloc(256 * 256 * 16)
    ctx = V.make_context(cam, s, spp=spp, bounces=bounces, clamp=10.0,
                         estimator=V.EST_PARITY, rng=V.RNG_PCG, seed=0x5EED)
    for _ in range(args.warmup):
        V.path_trace_tiles_device(ctx, tiles, img)
    dev.synchronize()
    ext_ms, ext_n, steps_ms = 0.0, 0, []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ts = time.perf_counter()
        V.path_trace_tiles_device(ctx, tiles, img)
        steps_ms.append((time.perf_counter() - ts) * 1e3)
        tim = dev.last_timings()
        ext_ms += tim["extend"][0]
        ext_n += tim["extend"][1]
    dev.synchronize()
    wall = time.perf_counter() - t0
    # counting run of the same launches: slab and triangle tests, rays
    cctx = V.make_context(cam, s, spp=spp, bounces=bounces, clamp=10.0,
                          estimator=V.EST_PARITY, rng=V.RNG_PCG, seed=0x5EED,
                          collect=True)
    m = V.path_trace_tiles_device(cctx, tiles, img)
    rays = m.values[V.METRIC_RAYS]
    box = m.values[V.METRIC_AABB_TESTS]
    tri = dev.last_counters()["tri_tests"]
    bytes_step = 44 * rays + 24 * box + 36 * tri
    launches = ext_n / args.steps
    per_launch_bytes = bytes_step / launches
    mean_ms = ext_ms / ext_n
    achieved = per_launch_bytes / (mean_ms * 1e-3) / 1e9
    samples = 256 * 256 * spp
    res = {
        "metric": METRIC,
        "value": samples * args.steps / wall / 1e6,
        "unit": "Msamples/s",
        "ms_per_step": wall * 1e3 / args.steps,
        "dtype": "f32",
        "config": {"workload": "config1: icosphere(3) + 20x20 ground quad, sky cube "
                               "6x64x64 (sun 50), 256x256, 4 spp, 3 bounces, "
                               "parity estimator, per-pixel PCG",
                   "image": [256, 256], "spp": spp, "bounces": bounces,
                   "triangles": 1282},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": None,
                     "kernel": "k_extend (closest-hit traversal)",
                     "launch_ms": mean_ms, "algorithmic_bytes_per_launch": per_launch_bytes,
                     "rays_per_step": rays, "slab_tests_per_ray": box / max(rays, 1),
                     "tri_tests_per_ray": tri / max(rays, 1)},
        "extra": {"gpu_ms_per_step": dev.last_timings()["total"][0],
                  "Mrays_per_s": rays * args.steps / wall / 1e6,
                  "step_ms": steps_ms},
    }
    if not args.no_cpu:
        res["cpu_baseline"] = cpu_baseline_config1(min(16, os.cpu_count() or 1))
    img.free()
    s.close()
    dev.close()
    return res


def bench_soup(args, V, scenes):
    dev = V.Device(0)
    tris, o, d = scenes.soup(1_000_000, 1 << 22, seed=7)
    s = V.Scene(dev)
    v, i = scenes.soup_mesh(tris)
    m = s.create_mesh(v, i, smooth=False)
    s.build()
    n = len(o)
    bufs = [dev.upload(np.ascontiguousarray(x)) for x in
            (o[:, 0], o[:, 1], o[:, 2], d[:, 0], d[:, 1], d[:, 2])]
    outp, outt = dev.alloc(4 * n), dev.alloc(4 * n)
    st = V.TraversalStats()
    V.check(V.lib().spg_intersect_closest_device(
        s.handle, m, *[b.ptr for b in bufs], n, outp.ptr, outt.ptr, None, None,
        C.byref(st)))
    ms = C.c_double()
    for _ in range(args.warmup):
        V.check(V.lib().spg_intersect_closest_device(
            s.handle, m, *[b.ptr for b in bufs], n, outp.ptr, outt.ptr, None, None, None))
    t0 = time.perf_counter()
    V.check(V.lib().spg_time_intersect_closest(s.handle, m, *[b.ptr for b in bufs], n,
                                               outp.ptr, outt.ptr, args.steps,
                                               C.byref(ms)))
    wall = time.perf_counter() - t0
    bytes_launch = 44 * n + 24 * st.boxTests + 36 * st.triTests
    achieved = bytes_launch / (ms.value * 1e-3) / 1e9
    res = {
        "metric": "Mrays/s (closest hit)",
        "value": n * args.steps / wall / 1e6,
        "unit": "Mrays/s",
        "ms_per_step": wall * 1e3 / args.steps,
        "dtype": "f32",
        "config": {"workload": "soup: 2^22 rays (origins U[-2,2]^3, directions "
                               "uniform on S^2) vs 1M triangles (centroids U[0,1]^3, "
                               "edge lognormal(-5, 0.5)), SAH BVH2"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": None, "kernel": "k_closest_mesh",
                     "launch_ms": ms.value, "algorithmic_bytes_per_launch": bytes_launch,
                     "slab_tests_per_ray": st.boxTests / n,
                     "tri_tests_per_ray": st.triTests / n},
    }
    if not args.no_cpu:
        res["cpu_baseline"] = cpu_baseline_soup(min(16, os.cpu_count() or 1), tris, o, d)
    for b in bufs + [outp, outt]:
        b.free()
    s.close()
    dev.close()
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="config1", choices=["config1", "soup"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(this build traces on a single MI355X)", file=sys.stderr)
        sys.exit(2)
    import vk_cinematic_amd as V
    from vk_cinematic_amd import scenes

    fn = {"config1": bench_config1, "soup": bench_soup}[args.workload]
    res = fn(args, V, scenes)
    out = {"metric": res.pop("metric"), "value": res.pop("value"),
           "unit": res.pop("unit"), "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": res.pop("ms_per_step"),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": res.pop("dtype"), "data": "synthetic"}
    out.update(res)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
