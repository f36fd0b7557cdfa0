"""Benchmark: wavefront + persistent-extend path tracing of the seeded
incoherent mesh (1,048,576 FP32 triangles, centres uniform in the unit cube,
vertex offsets lognormal(ln 0.004, 0.5), 16 emissive triangles, materials
uniform over diffuse / conductor (fuzz 0.1) / dielectric (IOR 1.5),
procedural 512x256 FP32 equirect sky), 1024x1024, depth 16.

One step = one sample pass over all 1024x1024 pixels; the default 64 steps
are the workload's 64 spp. The scene, its BVH and the film are resident in
HBM before the timed region starts. Prints one JSON line.

  value     Mrays/s = (closest-hit rays + shadow rays) / render seconds
  paths/s   W*H*spp / render seconds (the reference's Mpaths/s definition)
  roofline  extend kernel: algorithmic bytes per ray
            B = 32 + 16 + 80*S + 44*L (28 B ray + 4 B queue slot, 16 B hit
            record, 16 B child refs + 2x32 B child boxes per internal-node
            step, 4 B index + 4 B type + 36 B triangle per leaf test), with
            S and L counted by the kernel itself in a separate counting pass
            over the same samples, divided by the extend time measured with
            HIP events on the library's stream.
  cpu_baseline  the CPU checker (oracle/, a host C++ restatement of the
            reference's device functions) on a band of image rows at 1 spp,
            all host threads.
"""
from __future__ import annotations

import argparse
import importlib
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(hot, cold, mats, sky, desc, depth, seed, budget_s, threads):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle

    t0 = time.time()
    ref = pyoracle.Scene(hot, cold, mats, sky=sky)
    build_s = time.time() - t0
    cam = pyoracle.camera_init(desc)
    H = cam.height
    # calibrate on a thin band, then size the band to ~budget_s
    mid = H // 2
    rows = 8
    _, _, st = ref.render(cam, 1, depth, seed=seed, rows=(mid, mid + rows), threads=threads)
    per_row = max(st["seconds"] / rows, 1e-4)
    rows = int(max(8, min(H, budget_s / per_row)))
    r0 = max(0, mid - rows // 2)
    _, _, st = ref.render(cam, 1, depth, seed=seed, rows=(r0, r0 + rows), threads=threads)
    secs = st["seconds"]
    return {
        "value": (st["extend_rays"] + st["shadow_rays"]) / secs / 1e6,
        "unit": "Mrays/s",
        "paths_per_s": st["paths"] / secs,
        "cores": st["threads"],
        "kind": "port",
        "sample": f"rows {r0}-{r0 + rows - 1} of {desc.width}x{H}, 1 spp, depth {depth}, {secs:.1f} s "
                  f"(BVH build {build_s:.1f} s not timed)",
    }


def copy_calibration(nbytes=1 << 30):
    """One device-to-device copy of `nbytes`, for scaling rocprofv3's
    FETCH_SIZE / WRITE_SIZE counters: a copy kernel of known size reads and
    writes exactly nbytes."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) == 0 and hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(nbytes)) == 0
    for _ in range(3):
        assert hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_calibration_bytes": nbytes, "copies": 3}), flush=True)


def main():
    if "--copy-calibration" in sys.argv:
        copy_calibration()
        return
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=64, help="timed sample passes (spp)")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--width", type=int, default=1024)
    ap.add_argument("--height", type=int, default=1024)
    ap.add_argument("--depth", type=int, default=16)
    ap.add_argument("--tris", type=int, default=1048576)
    ap.add_argument("--mode", default="persistent", choices=["megakernel", "wavefront", "persistent"])
    ap.add_argument("--refill", type=int, default=0)
    ap.add_argument("--persistent-shadow", action="store_true", help="shadow rays on the persistent scheme too")
    ap.add_argument("--batch", type=int, default=0, help="samples in flight per wavefront pass (0 = auto)")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-count", action="store_true", help="skip the traversal-counting pass")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (every GPU box has one GPU); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    mcpt = importlib.import_module("monocuda-pt_amd")
    mode = {"megakernel": mcpt.MODE_MEGAKERNEL, "wavefront": mcpt.MODE_WAVEFRONT,
            "persistent": mcpt.MODE_PERSISTENT}[args.mode]
    W, H, K = args.width, args.height, args.steps
    xf = mcpt.RENDER_PERSISTENT_SHADOW if args.persistent_shadow else 0

    t0 = time.time()
    hot, cold, mats = mcpt.gen_scene(mcpt.GEN_SOUP, seed=args.seed, param=args.tris)
    sky = mcpt.gen_sky(512, 256)
    desc = mcpt.gen_camera(mcpt.GEN_SOUP, W, H)
    cam = mcpt.camera_init(desc)
    log(f"bench: scene {len(hot)} triangles generated in {time.time() - t0:.1f} s")

    with mcpt.Context() as ctx:
        log(f"bench: device {ctx.describe()}")
        scene = ctx.scene(hot, cold, mats, env=mcpt.make_env(sky))
        build = scene.build()
        log(f"bench: LBVH build {build['build_ms']:.2f} ms, depth {build['max_depth']}")
        film = ctx.film(W, H)
        scratch = ctx.film(W, H)
        if args.warmup > 0:
            scene.render(cam, scratch, mode=mode, spp=args.warmup, sample_begin=K, max_depth=args.depth,
                         seed=args.seed, refill_lanes=args.refill, batch=args.batch, flags=xf)
        log(f"bench: warmup {args.warmup} done")
        t1 = time.perf_counter()
        st = scene.render(cam, film, mode=mode, spp=K, sample_begin=0, max_depth=args.depth, seed=args.seed,
                          refill_lanes=args.refill, batch=args.batch, flags=mcpt.RENDER_TIME_KERNELS | xf)
        wall_s = time.perf_counter() - t1
        log(f"bench: timed {K} steps in {wall_s:.3f} s (device {st['total_ms']:.1f} ms)")
        counted = ref_counted = None
        if not args.no_count:
            # the timed kernel's own visit counts (near-first order), then the
            # reference's visit order over the same rays
            scratch.clear()
            counted = scene.render(cam, scratch, mode=mode, spp=K, sample_begin=0, max_depth=args.depth,
                                   seed=args.seed, refill_lanes=args.refill, batch=args.batch, flags=mcpt.RENDER_COUNT_TRAVERSAL | xf)
            assert counted["extend_rays"] == st["extend_rays"], "counting pass traced different rays"
            scratch.clear()
            ref_counted = scene.render(cam, scratch, mode=mode, spp=K, sample_begin=0, max_depth=args.depth,
                                       seed=args.seed, refill_lanes=args.refill,
                                       flags=mcpt.RENDER_COUNT_TRAVERSAL | mcpt.RENDER_REFERENCE_ORDER | xf)
            assert ref_counted["extend_rays"] == st["extend_rays"]
            log("bench: counting passes done")
        img = film.read()
        film.close()
        scratch.close()
        scene.close()

    secs = st["total_ms"] / 1e3
    rays = st["extend_rays"] + st["shadow_rays"]
    result = {
        "metric": "Mrays/s + paths/s at fixed relMSE vs plain; p50/p99 job latency under load",
        "value": rays / secs / 1e6,
        "unit": "Mrays/s",
        "paths_per_s": st["paths"] / secs,
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": wall_s * 1e3 / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded procedural scene)",
        "config": {"workload": f"incoherent mesh {len(hot)} tris + 512x256 sky, {W}x{H}, {K} spp, depth "
                               f"{args.depth}, {args.mode} extend", "mode": args.mode, "width": W, "height": H,
                   "spp": K, "max_depth": args.depth, "triangles": len(hot)},
        "extend_rays": st["extend_rays"],
        "shadow_rays": st["shadow_rays"],
        "kernel_ms": {k: round(st[k], 3) for k in ("camera_ms", "extend_ms", "shade_ms", "shadow_ms",
                                                     "finalize_ms")},
        "lbvh_build_ms": build["build_ms"],
        "image_mean": [float(x) for x in img.reshape(-1, 3).mean(0) / K],
    }
    if counted is not None and st["extend_ms"] > 0:
        n = st["extend_rays"]
        alg_bytes = 48 * n + 80 * counted["extend_steps"] + 44 * counted["extend_leaves"]
        launches = st["extend_launches"]
        avg_ms = st["extend_ms"] / launches
        achieved = alg_bytes / (st["extend_ms"] / 1e3)
        result["roofline"] = {
            "bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
            "frac": achieved / HBM_PEAK, "traffic": None,
            "kernel": "k_extend_persistent" if mode == mcpt.MODE_PERSISTENT else "k_extend",
            "launches": launches, "avg_launch_ms": avg_ms, "alg_bytes_per_launch": alg_bytes / launches,
            "steps_per_ray": counted["extend_steps"] / n, "leaves_per_ray": counted["extend_leaves"] / n,
        }
        # HBM-side traffic per launch, measured by rocprofv3 --pmc FETCH_SIZE /
        # WRITE_SIZE (copy-kernel scaled) for exactly this configuration
        prof = os.path.join(ROOT, "profiles", "extend_soup1M_64spp.json")
        if os.path.exists(prof):
            with open(prof) as f:
                p = json.load(f)
            wl = p["workload"]
            if (wl["width"], wl["height"], wl["spp"], wl["max_depth"], wl["triangles"], wl["mode"], wl["launches"]) == \
                    (W, H, K, args.depth, len(hot), args.mode, launches) and mode == mcpt.MODE_PERSISTENT:
                result["roofline"]["traffic"] = p["k_extend_persistent"]["traffic_bytes_per_launch"]
                result["roofline"]["traffic_source"] = "profiles/extend_soup1M_64spp.json (rocprofv3 --pmc)"
        ref_bytes = 48 * n + 80 * ref_counted["extend_steps"] + 44 * ref_counted["extend_leaves"]
        result["roofline"]["reference_order"] = {
            "steps_per_ray": ref_counted["extend_steps"] / n, "leaves_per_ray": ref_counted["extend_leaves"] / n,
            "equivalent_GBps": ref_bytes / (st["extend_ms"] / 1e3) / 1e9,
            "note": "the reference's left-first visit order does this many bytes of work for the same hits",
        }
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        result["cpu_baseline"] = cpu_baseline(hot, cold, mats, sky, desc, args.depth, args.seed, args.cpu_seconds,
                                              threads)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
