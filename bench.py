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
            B = 32 + 16 + 64*S + 48*L (28 B ray + 4 B queue slot, 16 B hit
            record, one 64 B four-wide node per step, one 48 B leaf record
            per triangle test), with S and L counted by the kernel itself in
            a separate counting pass over the same samples, divided by the
            extend time measured with HIP events on the library's stream.
            The reference's binary walk of the same rays is reported beside
            it as 32 + 16 + 80*S + 44*L (16 B child refs + 2x32 B child
            boxes per step, 4 B index + 4 B type + 36 B triangle per leaf).
  cpu_baseline  the CPU oracle (oracle/, a host C++ restatement of the
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


def host_threads():
    """Threads this process may run on (its CPU affinity set), which on a
    shared box can be fewer than os.cpu_count()."""
    try:
        return len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        return os.cpu_count() or 1


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(hot, cold, mats, sky, desc, depth, seed, budget_s, threads, spheres=None, grid=None, runs=3):
    """The CPU oracle on a band of rows sized to about budget_s / runs per
    run; the median of `runs` timed runs of the same band."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle

    t0 = time.time()
    ref = pyoracle.Scene(hot, cold, mats, sky=sky, spheres=spheres)
    if grid is not None:
        ref.set_density_grid(*grid)
    build_s = time.time() - t0
    cam = pyoracle.camera_init(desc)
    H = cam.height
    # calibrate on a thin band, then size the band to ~budget_s / runs
    mid = H // 2
    rows = 8
    _, _, st = ref.render(cam, 1, depth, seed=seed, rows=(mid, mid + rows), threads=threads)
    per_row = max(st["seconds"] / rows, 1e-4)
    per_run = budget_s / runs
    rows = int(max(8, min(H, per_run / per_row)))
    spp = 1 if rows < H else int(max(1, min(64, per_run / (per_row * H))))
    r0 = max(0, mid - rows // 2)
    times = []
    for _ in range(runs):
        _, _, st = ref.render(cam, spp, depth, seed=seed, rows=(r0, r0 + rows), threads=threads)
        times.append(st["seconds"])
    secs = sorted(times)[len(times) // 2]
    return {
        "value": (st["extend_rays"] + st["shadow_rays"]) / secs / 1e6,
        "unit": "Mrays/s",
        "paths_per_s": st["paths"] / secs,
        "cores": st["threads"],
        "kind": "port",
        "cpu": cpu_model(),
        "runs_s": [round(t, 3) for t in times],
        "sample": f"rows {r0}-{r0 + rows - 1} of {desc.width}x{H}, {spp} spp, depth {depth}, median of {runs} "
                  f"runs of {secs:.1f} s on {st['threads']} threads (BVH build {build_s:.1f} s not timed)",
    }


# ---------------------------------------------------------------- serving
SERVE_SPP = (1, 4, 16, 64)


def serving_workload(mcpt, n_scenes, n_jobs, seed):
    """Resident scenes with triangle counts log-uniform in 10k-2M (the
    incoherent-mesh generator) and a seeded job list: scene uniform,
    square resolution log-uniform in 128^2-1024^2, spp uniform in
    {1, 4, 16, 64}."""
    import numpy as np
    rng = np.random.default_rng(seed)
    tris = np.exp(rng.uniform(np.log(1e4), np.log(2e6), n_scenes)).astype(np.int64)
    jobs = []
    for i in range(n_jobs):
        side = int(round(np.exp(rng.uniform(np.log(128), np.log(1024)))))
        jobs.append({"scene": int(rng.integers(n_scenes)), "side": side, "spp": int(rng.choice(SERVE_SPP)),
                     "seed": 1000 + i})
    gaps = rng.exponential(1.0, n_jobs)  # unit-rate Poisson inter-arrival gaps, scaled per load
    return tris, jobs, gaps


def _pct(xs, q):
    import numpy as np
    return float(np.percentile(np.asarray(xs, np.float64), q)) if len(xs) else None


def _plain_fifo(service_s, arrivals_s):
    """Latencies of one-job-at-a-time FIFO service with known service times
    (Lindley's recursion)."""
    free_at, lat = 0.0, []
    for a, s in zip(arrivals_s, service_s):
        start = max(a, free_at)
        free_at = start + s
        lat.append(free_at - a)
    return lat


def _run_server(mcpt, srv, jobs, cams, depth, arrivals_s=None):
    """Submit `jobs` (all at once, or at `arrivals_s` seconds), release each
    as it finishes; returns per-job latency (s) and the makespan (s)."""
    import threading
    ids = {}
    results = {}
    done = threading.Event()
    lock = threading.Lock()

    def reaper():
        while len(results) < len(jobs):
            with lock:
                pending = [(i, jid) for i, jid in ids.items() if i not in results]
            for i, jid in pending:
                r = srv.poll(jid)
                if r["status"] == mcpt.OK:
                    results[i] = r
                    srv.release(jid)
            time.sleep(0.0005)
        done.set()

    th = threading.Thread(target=reaper, daemon=True)
    th.start()
    t0 = time.perf_counter()
    for i, j in enumerate(jobs):
        if arrivals_s is not None:
            wait = t0 + arrivals_s[i] - time.perf_counter()
            if wait > 0:
                time.sleep(wait)
        jid, _ = srv.submit(j["scene"], cams[j["side"]], j["spp"], max_depth=depth, seed=j["seed"])
        with lock:
            ids[i] = jid
    done.wait()
    th.join()
    lat = [(results[i]["finish_ms"] - results[i]["submit_ms"]) / 1e3 for i in range(len(jobs))]
    first = min(r["submit_ms"] for r in results.values())
    last = max(r["finish_ms"] for r in results.values())
    return lat, (last - first) / 1e3


def _run_plain(mcpt, ctx, scenes, jobs, cams, depth, arrivals_s=None, batch=0):
    """One job at a time with mcpt_render (persistent mode), FIFO; a film
    per job. batch=0 puts all of a job's samples in flight at once; batch=1
    renders one sample per wavefront pass as the reference's
    renderWithWavefront does. Returns per-job latency and service time (s)
    and the makespan."""
    lat, svc = [], []
    totals = {"rays": 0, "device_ms": 0.0}
    t0 = time.perf_counter()
    for i, j in enumerate(jobs):
        if arrivals_s is not None:
            wait = t0 + arrivals_s[i] - time.perf_counter()
            if wait > 0:
                time.sleep(wait)
        ts = time.perf_counter()
        film = ctx.film(j["side"], j["side"])
        st = scenes[j["scene"]].render(cams[j["side"]], film, mode=mcpt.MODE_PERSISTENT, spp=j["spp"],
                                       max_depth=depth, seed=j["seed"], batch=batch)
        film.close()
        te = time.perf_counter()
        svc.append(te - ts)
        lat.append(te - t0 - (arrivals_s[i] if arrivals_s is not None else 0.0))
        totals["rays"] += st["extend_rays"] + st["shadow_rays"]
        totals["device_ms"] += st["total_ms"]
    _run_plain.totals = totals
    return lat, svc, time.perf_counter() - t0


def serving_bench(mcpt, ctx, args):
    import numpy as np
    from concurrent.futures import ThreadPoolExecutor
    tris, jobs, gaps = serving_workload(mcpt, args.serve_scenes, args.serve_jobs, args.serve_seed)
    t0 = time.time()
    with ThreadPoolExecutor(8) as ex:
        gen = list(ex.map(lambda a: mcpt.gen_scene(mcpt.GEN_SOUP, seed=a[0], param=int(a[1])),
                          enumerate(tris, start=100)))
    sky = mcpt.gen_sky(512, 256)
    scenes = []
    for hot, cold, mats in gen:
        sc = ctx.scene(hot, cold, mats, env=mcpt.make_env(sky))
        sc.build()
        scenes.append(sc)
    del gen
    log(f"serving: {len(scenes)} resident scenes, {int(tris.sum())} triangles, ready in {time.time() - t0:.1f} s")
    cams = {j["side"]: mcpt.camera_init(mcpt.gen_camera(mcpt.GEN_SOUP, j["side"], j["side"])) for j in jobs}
    depth = args.serve_depth
    paths = [j["side"] ** 2 * j["spp"] for j in jobs]

    # one job at a time, back to back: service times and plain capacity
    big = max(range(len(jobs)), key=lambda i: paths[i])
    _run_plain(mcpt, ctx, scenes, [jobs[big]] + jobs[:4], cams, depth)  # warm-up (workspace allocation)
    _, svc, plain_span = _run_plain(mcpt, ctx, scenes, jobs, cams, depth)
    plain_cap = len(jobs) / sum(svc)
    pt = _run_plain.totals
    plain_dev = {"Mrays_per_device_s": pt["rays"] / pt["device_ms"] / 1e3, "device_s": pt["device_ms"] / 1e3,
                 "wall_s": sum(svc)}
    log(f"serving: plain back-to-back {plain_cap:.2f} jobs/s {plain_dev}")

    # the reference's structure, one sample per pass, back to back
    _, svc1, _ = _run_plain(mcpt, ctx, scenes, jobs, cams, depth, batch=1)
    plain1_cap = len(jobs) / sum(svc1)
    log(f"serving: plain one sample per pass {plain1_cap:.2f} jobs/s")

    srv = ctx.server(scenes, path_capacity=args.serve_pool, policy=args.serve_policy)
    _run_server(mcpt, srv, jobs[:8], cams, depth)  # warm-up
    s0 = srv.stats()
    _, span = _run_server(mcpt, srv, jobs, cams, depth)
    s1 = srv.stats()
    srv_cap = len(jobs) / span
    srv_dev = {"Mrays_per_device_s": (s1["extend_rays"] + s1["shadow_rays"] - s0["extend_rays"] - s0["shadow_rays"])
               / (s1["busy_ms"] - s0["busy_ms"]) / 1e3, "device_s": (s1["busy_ms"] - s0["busy_ms"]) / 1e3,
               "wall_s": span, "steps": s1["iterations"] - s0["iterations"]}
    log(f"serving: server saturated {srv_cap:.2f} jobs/s {srv_dev}")

    out = {
        "workload": f"{len(scenes)} resident incoherent-mesh scenes ({int(tris.min())}-{int(tris.max())} tris, "
                    f"{int(tris.sum())} total) + 512x256 sky; {len(jobs)} jobs, square resolution log-uniform "
                    f"128^2-1024^2, spp uniform {{1,4,16,64}}, depth {depth}; Poisson arrivals",
        "jobs": len(jobs), "mean_paths_per_job": float(np.mean(paths)),
        "pool_paths": args.serve_pool or 8 << 20,
        "policy": {mcpt.SCHED_SHORTEST_FIRST: "shortest-first", mcpt.SCHED_FIFO: "fifo",
                   mcpt.SCHED_SHORTEST_FIRST_AGED: "shortest-first-aged"}[args.serve_policy],
        "capacity_jobs_per_s": {"server": srv_cap, "plain": plain_cap, "plain_one_sample_per_pass": plain1_cap},
        "capacity_ratio": {"vs_plain": srv_cap / plain_cap, "vs_plain_one_sample_per_pass": srv_cap / plain1_cap},
        "back_to_back": {"server": srv_dev, "plain": plain_dev},
        "loads": [],
    }
    for load in args.serve_loads:
        lam = load * srv_cap
        arr = np.cumsum(gaps) / lam
        arr -= arr[0]
        lat_s, span_s = _run_server(mcpt, srv, jobs, cams, depth, arr)
        lat_p, _, span_p = _run_plain(mcpt, ctx, scenes, jobs, cams, depth, arr)
        p99_s = _pct(lat_s, 99)
        # the plain path's sustainable rate at the server's p99: Lindley
        # replay of the measured back-to-back service times on the same
        # arrival pattern, bisected over the rate
        def rate_at(service, cap):
            # None when even a vanishing arrival rate misses the p99 (the
            # largest jobs alone take longer)
            if _pct(_plain_fifo(service, np.cumsum(gaps) * 1e9), 99) > p99_s:
                return None
            lo, hi = 1e-3, cap
            for _ in range(40):
                mid = 0.5 * (lo + hi)
                a = np.cumsum(gaps) / mid
                if _pct(_plain_fifo(service, a - a[0]), 99) <= p99_s:
                    lo = mid
                else:
                    hi = mid
            return lo
        lo = rate_at(svc, plain_cap)
        lo1 = rate_at(svc1, plain1_cap)
        rec = {
            "load": load, "arrival_jobs_per_s": lam,
            "server": {"jobs_per_s": len(jobs) / span_s, "p50_ms": _pct(lat_s, 50) * 1e3,
                       "p99_ms": p99_s * 1e3},
            "plain": {"jobs_per_s": len(jobs) / span_p, "p50_ms": _pct(lat_p, 50) * 1e3,
                      "p99_ms": _pct(lat_p, 99) * 1e3,
                      "replay_p99_ms": _pct(_plain_fifo(svc, arr), 99) * 1e3},
            "plain_jobs_per_s_at_server_p99": lo,
            "jobs_per_s_ratio_at_equal_p99": lam / lo if lo else None,
            "plain_one_sample_per_pass": {"replay_p99_ms": _pct(_plain_fifo(svc1, arr), 99) * 1e3,
                                          "jobs_per_s_at_server_p99": lo1,
                                          "jobs_per_s_ratio_at_equal_p99": lam / lo1 if lo1 else None},
        }
        out["loads"].append(rec)
        log(f"serving: load {load}: {json.dumps(rec)}")
    st = srv.stats()
    out["server_stats"] = {k: st[k] for k in ("iterations", "paths_started", "extend_rays", "shadow_rays", "busy_ms")}
    srv.close()
    for sc in scenes:
        sc.close()
    return out


# ---------------------------------------------------------- batched scenes
def batch_bench(mcpt, ctx, args):
    """64 seeded scenes (displaced icospheres + random fill, triangle counts
    log-uniform in 1k-500k), each 256x256, 8 spp, depth 8: one batched LBVH
    build + one batched render (autotuned) against 64 sequential builds and
    renders. Films are compared bit for bit."""
    import numpy as np
    from concurrent.futures import ThreadPoolExecutor
    rng = np.random.default_rng(args.batch_seed)
    tris = np.exp(rng.uniform(np.log(1e3), np.log(5e5), args.batch_scenes)).astype(np.int64)
    t0 = time.time()
    with ThreadPoolExecutor(8) as ex:
        data = list(ex.map(lambda a: mcpt.gen_scene(mcpt.GEN_ICOSPHERE, seed=a[0], param=int(a[1])),
                           enumerate(tris, start=200)))
    S, spp, depth = args.batch_size, args.batch_spp, args.batch_depth
    cams = [mcpt.camera_init(mcpt.gen_camera(mcpt.GEN_SOUP, S, S))] * len(data)
    n_tris = sum(len(d[0]) for d in data)
    log(f"batch: {len(data)} scenes, {n_tris} triangles generated in {time.time() - t0:.1f} s")
    # sequential: build + render scene by scene (persistent, all samples in flight)
    seq = [ctx.scene(*d) for d in data]
    films = [ctx.film(S, S) for _ in data]
    seq[0].build()
    seq[0].render(cams[0], films[0], spp=1, sample_begin=1 << 20, max_depth=depth, seed=args.seed)  # warm-up
    films[0].clear()
    seq_build = seq_render = 0.0
    t1 = time.perf_counter()
    for sc, f, cam in zip(seq, films, cams):
        seq_build += sc.build()["build_ms"]
        seq_render += sc.render(cam, f, spp=spp, max_depth=depth, seed=args.seed)["total_ms"]
    seq_wall = time.perf_counter() - t1
    ref = [f.read() for f in films]
    for sc in seq:
        sc.close()
    for f in films:
        f.clear()
    # batched
    bat = [ctx.scene(*d) for d in data]
    t2 = time.perf_counter()
    b_build = mcpt.build_batch(bat)
    tune, search_ms = mcpt.batch_autotune(bat, cams, spp=spp, max_depth=depth, seed=args.seed)
    st = mcpt.render_batch(bat, cams, films, spp=spp, max_depth=depth, seed=args.seed, tune=tune)
    b_wall = time.perf_counter() - t2
    b_render = st["total_ms"]
    equal = all(np.array_equal(f.read(), r) for f, r in zip(films, ref))
    paths = S * S * spp * len(data)
    seq_ms, b_ms = seq_build + seq_render, b_build + b_render
    out = {
        "workload": f"{len(data)} displaced-icosphere scenes ({int(tris.min())}-{int(tris.max())} requested tris, "
                    f"{n_tris} total), {S}x{S}, {spp} spp, depth {depth}",
        "films_identical": equal,
        "sequential": {"build_ms": seq_build, "render_ms": seq_render, "device_ms": seq_ms, "wall_s": seq_wall,
                       "scenes_per_s": len(data) / (seq_ms / 1e3), "paths_per_s": paths / (seq_ms / 1e3)},
        "batched": {"build_ms": b_build, "render_ms": b_render, "device_ms": b_ms, "wall_s": b_wall,
                    "autotune_ms": search_ms, "tune": tune.as_dict(),
                    "scenes_per_s": len(data) / (b_ms / 1e3), "paths_per_s": paths / (b_ms / 1e3),
                    "scenes_per_s_with_autotune": len(data) / ((b_ms + search_ms) / 1e3),
                    "extend_rays": st["extend_rays"], "shadow_rays": st["shadow_rays"]},
        "speedup": seq_ms / b_ms, "speedup_with_autotune": seq_ms / (b_ms + search_ms),
        "build_speedup": seq_build / b_build,
    }
    log(f"batch: sequential {seq_ms:.1f} ms, batched {b_ms:.1f} ms (+ autotune {search_ms:.1f} ms), "
        f"identical films {equal}")
    for sc in bat:
        sc.close()
    for f in films:
        f.close()
    return out


# ---------------------------------------------------------- radiance cache
MLP_FLOPS_FWD = 2 * (4 * 64 * 64 + 3 * 64)  # multiply-adds x 2 per query
MLP_FLOPS_TRAIN = 3 * MLP_FLOPS_FWD         # forward, input gradient, weight gradient
MFMA_BF16_PEAK = 2.5e15


def relmse(img, ref, eps=0.01):
    import numpy as np
    d = img.astype(np.float64) - ref
    return float(np.mean(d * d / (ref * ref + eps)))


def cache_bench(mcpt, ctx, args):
    """Time to relMSE 0.01 against an FP64-accumulated high-spp plain render,
    plain tracer vs the radiance-cached tracer (training included)."""
    import numpy as np
    W, H, depth = args.cache_width, args.cache_height, args.cache_depth
    t0 = time.time()
    hot, cold, mats = mcpt.gen_scene(mcpt.GEN_MEDIUM, seed=args.seed, param=args.cache_tris)
    sky = mcpt.gen_sky(512, 256)
    scene = ctx.scene(hot, cold, mats, env=mcpt.make_env(sky))
    scene.build()
    cam = mcpt.camera_init(mcpt.gen_camera(mcpt.GEN_SOUP, W, H))
    log(f"cache: {len(hot)} triangles + medium box, {W}x{H}, depth {depth}, ready in {time.time() - t0:.1f} s")
    # reference: chunks of FP32 sums added in float64
    film = ctx.film(W, H)
    ref = np.zeros((H, W, 3), np.float64)
    chunk, done, ref_ms = 16, 0, 0.0
    while done < args.cache_ref_spp:
        film.clear()
        st = scene.render(film=film, camera=cam, spp=chunk, sample_begin=done, max_depth=depth, seed=args.seed + 1)
        ref += film.read()
        ref_ms += st["total_ms"]
        done += chunk
        if done % 512 == 0:
            log(f"cache: reference {done} / {args.cache_ref_spp} spp")
    ref /= done
    log(f"cache: reference {done} spp in {ref_ms / 1e3:.1f} s")

    # both tracers render frames of S samples per pixel traced together
    S = args.cache_frame_spp
    grid = [1, 2, 4, 8, 16, 32, 48, 64, 96, 128, 192, 256, 384, 512, 768, 1024]
    checkpoints = [c for c in grid if c % S == 0 and c <= args.cache_max_spp]

    def converge(render_frame, step=S):
        film.clear()
        curve, t_ms, spp = [], 0.0, 0
        for cp in checkpoints:
            if cp % step:
                continue
            while spp < cp:
                t_ms += render_frame(spp)
                spp += step
            e = relmse(film.read() / spp, ref)
            curve.append({"spp": spp, "ms": t_ms, "relmse": e})
            log(f"cache:   spp {spp:5d}  {t_ms:9.1f} ms  relMSE {e:.5f}")
        return curve

    def time_to(curve, target):
        prev = None
        for c in curve:
            if c["relmse"] <= target:
                if prev is None:
                    return c["ms"]
                # relMSE ~ 1/time between checkpoints: interpolate in log-log
                import math
                a, b = prev, c
                f = (math.log(a["relmse"]) - math.log(target)) / (math.log(a["relmse"]) - math.log(b["relmse"]))
                return math.exp(math.log(a["ms"]) + f * (math.log(b["ms"]) - math.log(a["ms"])))
            prev = c
        return None

    log("cache: plain")
    plain = [] if args.cache_only else converge(
        lambda k: scene.render(cam, film, spp=S, sample_begin=k, max_depth=depth, seed=args.seed)["total_ms"])
    # the plain tracer at its own best throughput: 16 samples per pixel in
    # flight per pass
    log("cache: plain, 16 samples per pass")
    plain16 = [] if args.cache_only else converge(
        lambda k: scene.render(cam, film, spp=16, sample_begin=k, max_depth=depth, seed=args.seed)["total_ms"], 16)
    net = ctx.mlp(seed=1)
    agg = {"train_ms": 0.0, "infer_ms": 0.0, "train_samples": 0, "queries": 0, "train_trace_ms": 0.0,
           "render_ms": 0.0}

    def cached_one(k):
        st = scene.render_cached(cam, film, net, spp=S, sample_begin=k, max_depth=depth, seed=args.seed,
                                 train_paths=args.cache_train_paths, terminate_depth=args.cache_terminate,
                                 lr=args.cache_lr, train_steps=args.cache_steps, ema_decay=args.cache_ema,
                                 frame_spp=S)
        for key in agg:
            agg[key] += st[key]
        agg["loss"] = st["loss"]
        return st["total_ms"]

    log("cache: cached")
    warm_ms = 0.0
    if args.cache_warmup:
        # training frames of one sample per pixel through a camera with the
        # same view at 1/div of the resolution: the training paths see the
        # same scene, each frame costs a fraction of a full one
        wdiv = args.cache_warmup_div
        wcam = mcpt.camera_init(mcpt.gen_camera(mcpt.GEN_SOUP, max(1, W // wdiv), max(1, H // wdiv)))
        scratch = ctx.film(wcam.width, wcam.height)
        st = scene.render_cached(wcam, scratch, net, spp=args.cache_warmup, sample_begin=1 << 20, max_depth=depth,
                                 seed=args.seed, train_paths=args.cache_warmup_paths,
                                 terminate_depth=args.cache_terminate, lr=args.cache_lr,
                                 train_steps=args.cache_warmup_steps, ema_decay=args.cache_ema)
        warm_ms = st["total_ms"]
        scratch.close()
        log(f"cache: warm-up {args.cache_warmup} frames at {wcam.width}x{wcam.height} {warm_ms:.1f} ms")
    cached = converge(cached_one)
    for c in cached:
        c["ms"] += warm_ms
    target = 0.01
    tp, tc = time_to(plain, target), time_to(cached, target)
    tp16 = time_to(plain16, target)
    infer_tflops = agg["queries"] * MLP_FLOPS_FWD / (agg["infer_ms"] / 1e3) / 1e12 if agg["infer_ms"] else None
    train_tflops = agg["train_samples"] * MLP_FLOPS_TRAIN / (agg["train_ms"] / 1e3) / 1e12 if agg["train_ms"] else None
    out = {
        "workload": f"{len(hot)} tris incoherent mesh + medium box (sigma_t 0.5, albedo 0.8, g 0.3) + 512x256 sky, "
                    f"{W}x{H}, depth {depth}; reference {done} spp FP64-accumulated",
        "relmse_eps": 0.01,
        "time_to_relmse_ms": {"plain": tp, "cached": tc, "plain_16_per_pass": tp16, "target": target},
        "speedup": (tp / tc) if (tp and tc) else None,
        "speedup_vs_plain_16_per_pass": (tp16 / tc) if (tp16 and tc) else None,
        "frame_spp": S,
        "relmse_at_16spp": {"plain": next((c["relmse"] for c in plain if c["spp"] == 16), None),
                            "cached": next((c["relmse"] for c in cached if c["spp"] == 16), None)},
        "curves": {"plain": plain, "cached": cached, "plain_16_per_pass": plain16},
        "mlp": {"train_samples": agg["train_samples"], "queries": agg["queries"], "train_ms": agg["train_ms"],
                "infer_ms": agg["infer_ms"], "train_trace_ms": agg["train_trace_ms"], "render_ms": agg["render_ms"],
                "last_loss": agg.get("loss"),
                "mfma": {"bound": "mfma", "unit": "TFLOP/s", "peak": MFMA_BF16_PEAK / 1e12,
                         "infer_achieved": infer_tflops, "train_achieved": train_tflops,
                         "infer_frac": infer_tflops / (MFMA_BF16_PEAK / 1e12) if infer_tflops else None,
                         "flops_per_query": MLP_FLOPS_FWD, "flops_per_train_sample": MLP_FLOPS_TRAIN}},
        "cache_opts": {"train_paths": args.cache_train_paths, "terminate_depth": args.cache_terminate,
                       "warmup_frames": args.cache_warmup, "warmup_div": args.cache_warmup_div,
                       "warmup_train_paths": args.cache_warmup_paths, "warmup_steps": args.cache_warmup_steps,
                       "warmup_ms": warm_ms, "frame_spp": S,
                       "lr": args.cache_lr, "train_steps": args.cache_steps, "ema_decay": args.cache_ema},
    }
    film.close()
    net.close()
    scene.close()
    return out


# ----------------------------------------------------------- cloud medium
def cloud_scene(mcpt, seed, dims=(128, 96, 128), voxel=0.016):
    """The Cornell room around a heterogeneous cloud: an isotropic sphere
    enclosing a density grid of soft blobs (the reference's cloud scene
    wraps its NanoVDB volume in a sphere the same way)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    hot, cold, mats = mcpt.gen_scene(mcpt.GEN_CORNELL)
    cloud = np.zeros(1, mcpt.MAT_DTYPE)
    cloud["type"], cloud["albedo"], cloud["density"], cloud["phase_g"] = mcpt.MAT_ISOTROPIC, 0.95, 6.0, 0.6
    mats = np.concatenate([mats, cloud])
    nx, ny, nz = dims
    x = (np.arange(nx) + 0.5) / nx
    y = (np.arange(ny) + 0.5) / ny
    z = (np.arange(nz) + 0.5) / nz
    g = np.zeros((nz, ny, nx), np.float32)
    for _ in range(24):
        c = rng.uniform(0.2, 0.8, 3)
        r = rng.uniform(0.08, 0.25)
        d2 = ((x[None, None, :] - c[0]) ** 2 + (y[None, :, None] - c[1]) ** 2 + (z[:, None, None] - c[2]) ** 2) / r ** 2
        g += (np.maximum(0.0, 1.0 - d2) * rng.uniform(0.5, 1.5)).astype(np.float32)
    extent = np.array(dims, np.float32) * np.float32(voxel)
    center = np.array([0.0, 1.6, -2.2], np.float32)
    origin = center - extent / 2
    sp = np.zeros(1, mcpt.SPHERE_DTYPE)
    sp["center"], sp["radius"], sp["material"] = center, 0.5 * float(np.linalg.norm(extent)), len(mats) - 1
    return hot, cold, mats, sp, (g, origin, voxel)


def cloud_bench(mcpt, args):
    """One frame of the cloud room (delta tracking inside, ratio tracking
    for shadow rays) with the scene resident, timed by HIP events."""
    W = H = args.cloud_size
    K, depth = args.cloud_spp, args.depth
    hot, cold, mats, sp, grid = cloud_scene(mcpt, args.seed)
    desc = mcpt.gen_camera(mcpt.GEN_CORNELL, W, H)
    cam = mcpt.camera_init(desc)
    mode = {"megakernel": mcpt.MODE_MEGAKERNEL, "wavefront": mcpt.MODE_WAVEFRONT,
            "persistent": mcpt.MODE_PERSISTENT}[args.mode]
    with mcpt.Context() as ctx:
        scene = ctx.scene(hot, cold, mats, spheres=sp)
        gmax = scene.set_density_grid(*grid)
        scene.build()
        film = ctx.film(W, H)
        scene.render(cam, film, mode=mode, spp=max(args.warmup, 1), sample_begin=K, max_depth=depth, seed=args.seed)
        film.clear()
        st = scene.render(cam, film, mode=mode, spp=K, max_depth=depth, seed=args.seed)
        img = film.read()
        film.clear()
        kt = scene.render(cam, film, mode=mode, spp=K, max_depth=depth, seed=args.seed,
                          flags=mcpt.RENDER_TIME_KERNELS)
        film.close()
        scene.close()
    secs = st["total_ms"] / 1e3
    rays = st["extend_rays"] + st["shadow_rays"]
    out = {"workload": f"Cornell room + {grid[0].shape[2]}x{grid[0].shape[1]}x{grid[0].shape[0]} density grid "
                       f"cloud (max {gmax:.2f}, sigma 6), {W}x{H}, {K} spp, depth {depth}, {args.mode}",
           "Mrays_per_s": rays / secs / 1e6, "paths_per_s": st["paths"] / secs, "ms_per_spp": st["total_ms"] / K,
           "kernel_ms": {k: round(kt[k], 3) for k in ("camera_ms", "extend_ms", "shade_ms", "shadow_ms",
                                                        "finalize_ms")},
           "serial_ms": kt["total_ms"],
           "extend_rays": st["extend_rays"], "shadow_rays": st["shadow_rays"],
           "grid_MB": grid[0].nbytes / 1e6, "image_mean": float(img.mean() / K)}
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(hot, cold, mats, None, desc, depth, args.seed, args.cpu_seconds,
                                           host_threads(), spheres=sp, grid=grid)
    return out


def mlp_bench(mcpt, ctx, args):
    """The radiance-cache network alone: inference over 2,097,152 queries
    (the paths of a 1920x1080 frame) and training steps of 65,536 samples
    (the workload's training paths per frame), on random features. Wall
    times include the host copies of mcpt_mlp_forward / mcpt_mlp_train;
    kernel durations come from rocprofv3 (profiles/mlp_*.md)."""
    import numpy as np
    rng = np.random.default_rng(args.seed)
    n, nt = 2097152, 65536
    f = rng.uniform(0.0, 1.0, (n, 11)).astype(np.float32)
    t = rng.uniform(0.0, 1.0, (nt, 3)).astype(np.float32)
    net = ctx.mlp(seed=3)
    fwd, trn = [], []
    for _ in range(args.mlp_reps):
        t0 = time.perf_counter()
        net.forward(f)
        fwd.append(time.perf_counter() - t0)
    for _ in range(args.mlp_reps):
        t0 = time.perf_counter()
        loss = net.train(f[:nt], t, batch=nt, lr=1e-3)
        trn.append(time.perf_counter() - t0)
    net.close()
    return {"workload": f"MLP 64-64x4-3 BF16, inference over {n} queries, training steps of {nt} samples",
            "forward_wall_ms_median": 1e3 * sorted(fwd)[len(fwd) // 2],
            "train_step_wall_ms_median": 1e3 * sorted(trn)[len(trn) // 2], "reps": args.mlp_reps, "loss": loss,
            "flop_per_query": 33152, "flop_per_training_sample": 99456}


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
    ap.add_argument("--shadow-per-ray", action="store_true", help="shadow rays one lane per ray (default: persistent)")
    ap.add_argument("--batch", type=int, default=0, help="samples in flight per wavefront pass (0 = auto)")
    ap.add_argument("--group-below", type=int, default=0,
                    help="queues shorter than this walked by 16 lanes per ray (0 = library default, -1 = never)")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-count", action="store_true", help="skip the traversal-counting pass")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--workload", default="all", choices=["all", "extend", "serving", "cache", "batch", "cloud", "mlp"])
    ap.add_argument("--mlp-reps", type=int, default=5)
    ap.add_argument("--cloud-size", type=int, default=512)
    ap.add_argument("--cloud-spp", type=int, default=64)
    ap.add_argument("--batch-scenes", type=int, default=64)
    ap.add_argument("--batch-seed", type=int, default=3)
    ap.add_argument("--batch-size", type=int, default=256)
    ap.add_argument("--batch-spp", type=int, default=8)
    ap.add_argument("--batch-depth", type=int, default=8)
    ap.add_argument("--cache-width", type=int, default=1920)
    ap.add_argument("--cache-height", type=int, default=1080)
    ap.add_argument("--cache-depth", type=int, default=16)
    ap.add_argument("--cache-tris", type=int, default=2097152)
    ap.add_argument("--cache-ref-spp", type=int, default=4096)
    ap.add_argument("--cache-max-spp", type=int, default=256)
    ap.add_argument("--cache-train-paths", type=int, default=65536)
    ap.add_argument("--cache-terminate", type=int, default=2)
    ap.add_argument("--cache-lr", type=float, default=0.005)
    ap.add_argument("--cache-steps", type=int, default=4)
    ap.add_argument("--cache-ema", type=float, default=0.0)
    ap.add_argument("--cache-only", action="store_true", help="skip the plain curve (tuning runs)")
    ap.add_argument("--cache-warmup", type=int, default=4,
                    help="cached frames traced first to train the network; their time counts, their samples "
                         "are not kept")
    ap.add_argument("--cache-warmup-div", type=int, default=2,
                    help="the warm-up frames' camera has the same view at 1/div of the width and height")
    ap.add_argument("--cache-warmup-paths", type=int, default=262144, help="training paths per warm-up frame")
    ap.add_argument("--cache-warmup-steps", type=int, default=32, help="Adam steps per warm-up frame at most")
    ap.add_argument("--cache-frame-spp", type=int, default=1,
                    help="samples per pixel per frame, for the plain and the cached tracer alike")
    ap.add_argument("--serve-jobs", type=int, default=256)
    ap.add_argument("--serve-scenes", type=int, default=32)
    ap.add_argument("--serve-seed", type=int, default=7)
    ap.add_argument("--serve-depth", type=int, default=16)
    ap.add_argument("--serve-pool", type=int, default=0, help="server path pool (0 = library default)")
    ap.add_argument("--serve-policy", type=int, default=1,
                    help="0 shortest-first, 1 FIFO (default: best p99), 2 shortest-first aged")
    ap.add_argument("--serve-loads", type=float, nargs="*", default=[0.5, 0.8])
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (every GPU box has one GPU); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    mcpt = importlib.import_module("monocuda-pt_amd")
    if args.workload == "cloud":
        cl = cloud_bench(mcpt, args)
        print(json.dumps({
            "metric": "Mrays/s through a heterogeneous medium", "value": cl["Mrays_per_s"], "unit": "Mrays/s",
            "n_gpus": 1, "steps": args.cloud_spp, "warmup": args.warmup, "ms_per_step": cl["ms_per_spp"],
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
            "data": "synthetic (seeded procedural density grid)", "config": {"workload": cl["workload"]},
            "cpu_baseline": cl.get("cpu_baseline"), "cloud": cl}), flush=True)
        return
    if args.workload == "batch":
        with mcpt.Context() as ctx:
            bb = batch_bench(mcpt, ctx, args)
        print(json.dumps({
            "metric": "scenes/s and paths/s, batched build + render vs sequential", "value":
            bb["batched"]["scenes_per_s"], "unit": "scenes/s", "n_gpus": 1, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded procedural scenes)",
            "config": {"workload": bb["workload"]}, "batch": bb}), flush=True)
        return
    if args.workload == "mlp":
        with mcpt.Context() as ctx:
            mb = mlp_bench(mcpt, ctx, args)
        print(json.dumps({
            "metric": "MLP forward and training-step wall time", "value": mb["forward_wall_ms_median"], "unit": "ms",
            "n_gpus": 1, "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "bf16+f32",
            "data": "synthetic (random features)", "config": {"workload": mb["workload"]}, "mlp": mb}), flush=True)
        return
    if args.workload == "cache":
        with mcpt.Context() as ctx:
            cb = cache_bench(mcpt, ctx, args)
        print(json.dumps({
            "metric": "time to relMSE 0.01, plain / radiance-cached", "value": cb["speedup"], "unit": "x",
            "n_gpus": 1, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+bf16",
            "data": "synthetic (seeded procedural scene)", "config": {"workload": cb["workload"]}, "cache": cb}),
            flush=True)
        return
    if args.workload == "serving":
        with mcpt.Context() as ctx:
            sv = serving_bench(mcpt, ctx, args)
        print(json.dumps({
            "metric": "jobs/s and p50/p99 job latency under Poisson load", "value": sv["capacity_jobs_per_s"]["server"],
            "unit": "jobs/s", "n_gpus": 1, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f32", "data": "synthetic (seeded procedural scenes and job stream)",
            "config": {"workload": sv["workload"]}, "serving": sv}), flush=True)
        return
    mode = {"megakernel": mcpt.MODE_MEGAKERNEL, "wavefront": mcpt.MODE_WAVEFRONT,
            "persistent": mcpt.MODE_PERSISTENT}[args.mode]
    W, H, K = args.width, args.height, args.steps
    xf = mcpt.RENDER_SHADOW_PER_RAY if args.shadow_per_ray else 0

    t0 = time.time()
    hot, cold, mats = mcpt.gen_scene(mcpt.GEN_SOUP, seed=args.seed, param=args.tris)
    sky = mcpt.gen_sky(512, 256)
    desc = mcpt.gen_camera(mcpt.GEN_SOUP, W, H)
    cam = mcpt.camera_init(desc)
    log(f"bench: scene {len(hot)} triangles generated in {time.time() - t0:.1f} s")

    with mcpt.Context() as ctx:
        log(f"bench: device {ctx.describe()}")
        t_up = time.perf_counter()
        scene = ctx.scene(hot, cold, mats, env=mcpt.make_env(sky))  # host buffers cross PCIe here
        upload_s = time.perf_counter() - t_up
        build = scene.build()
        log(f"bench: LBVH build {build['build_ms']:.2f} ms, depth {build['max_depth']}")
        film = ctx.film(W, H)
        scratch = ctx.film(W, H)
        if args.warmup > 0:
            scene.render(cam, scratch, mode=mode, spp=args.warmup, sample_begin=K, max_depth=args.depth,
                         seed=args.seed, refill_lanes=args.refill, batch=args.batch, flags=xf,
                         group_below=args.group_below)
        log(f"bench: warmup {args.warmup} done")
        t1 = time.perf_counter()
        st = scene.render(cam, film, mode=mode, spp=K, sample_begin=0, max_depth=args.depth, seed=args.seed,
                          refill_lanes=args.refill, batch=args.batch, flags=xf, group_below=args.group_below)
        wall_s = time.perf_counter() - t1
        log(f"bench: timed {K} steps in {wall_s:.3f} s (device {st['total_ms']:.1f} ms)")
        # the same samples again with every launch in order on one stream and
        # timed on its own: the per-kernel split and the extend roofline
        scratch.clear()
        kt = scene.render(cam, scratch, mode=mode, spp=K, sample_begin=0, max_depth=args.depth, seed=args.seed,
                          refill_lanes=args.refill, batch=args.batch, flags=mcpt.RENDER_TIME_KERNELS | xf,
                          group_below=args.group_below)
        assert kt["extend_rays"] == st["extend_rays"]
        log(f"bench: kernel-timed pass {kt['total_ms']:.1f} ms")
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
        t_rd = time.perf_counter()
        img = film.read()
        read_s = time.perf_counter() - t_rd
        film.close()
        scratch.close()
        scene.close()

    secs = st["total_ms"] / 1e3
    rays = st["extend_rays"] + st["shadow_rays"]
    result = {
        "metric": "Mrays/s + paths/s at fixed relMSE vs plain; p50/p99 job latency under load",
        "value": rays / secs / 1e6,
        "unit": "Mrays/s",
        "value_is": ("closest-hit + shadow rays per second of device time for the whole K-sample render of the "
                     "incoherent mesh (the paths/s beside it); the metric's other parts are nested: serving p50/p99 "
                     "and jobs/s in 'serving', the radiance cache's time to relMSE 0.01 with --workload cache"),
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
        # per-kernel device time of the kernel-timed pass (launches in order)
        "kernel_ms": {k: round(kt[k], 3) for k in ("camera_ms", "extend_ms", "shade_ms", "shadow_ms",
                                                     "finalize_ms")},
        "serial_ms": kt["total_ms"],
        "lbvh_build_ms": build["build_ms"],
        # the same job with the scene handed over in host memory and the film
        # read back: scene creation (upload) + build + render + read
        "host_io": {"scene_upload_s": upload_s, "film_read_s": read_s,
                    "Mrays_per_s_incl_pcie": rays / (upload_s + build["build_ms"] / 1e3 + secs + read_s) / 1e6},
        "image_mean": [float(x) for x in img.reshape(-1, 3).mean(0) / K],
    }
    if counted is not None and kt["extend_ms"] > 0:
        n = st["extend_rays"]
        alg_bytes = 48 * n + 64 * counted["extend_steps"] + 48 * counted["extend_leaves"]
        launches = kt["extend_launches"]
        avg_ms = kt["extend_ms"] / launches
        achieved = alg_bytes / (kt["extend_ms"] / 1e3)
        result["roofline"] = {
            "bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
            "frac": achieved / HBM_PEAK, "traffic": None,
            "kernel": "k_extend_persistent" if mode == mcpt.MODE_PERSISTENT else "k_extend",
            "launches": launches, "avg_launch_ms": avg_ms, "alg_bytes_per_launch": alg_bytes / launches,
            "steps_per_ray": counted["extend_steps"] / n, "leaves_per_ray": counted["extend_leaves"] / n,
        }
        # Measured traffic per launch: rocprofv3 --pmc read and write request
        # counts at the L2's fabric interface (TCC_EA0_RDREQ_{32B,64B,128B},
        # TCC_EA0_WRREQ{,_64B}) per closest-hit ray, profiled on this scene,
        # resolution, depth and mode, times this run's rays per launch. They
        # count every L2 miss, whether the Infinity Cache or HBM serves it.
        prof = os.path.join(ROOT, "profiles", "traffic_soup1M_1024x1024_d16.json")
        if os.path.exists(prof) and mode == mcpt.MODE_PERSISTENT:
            with open(prof) as f:
                p = json.load(f)
            wl = p["workload"]
            if (wl["width"], wl["height"], wl["max_depth"], wl["triangles"], wl["mode"]) == \
                    (W, H, args.depth, len(hot), args.mode):
                k = p["kernels"]["k_extend_persistent"]
                tb = (k["read_bytes_per_ray"] + k["write_bytes_per_ray"]) * n / launches
                result["roofline"]["traffic"] = tb
                result["roofline"]["traffic_source"] = (
                    f"profiles/traffic_soup1M_1024x1024_d16.json: {k['read_bytes_per_ray']:.1f} B read + "
                    f"{k['write_bytes_per_ray']:.1f} B written per ray at the L2-fabric interface "
                    f"(rocprofv3 --pmc at {wl['profiled_spp']} spp) x this run's rays per launch")
                result["roofline"]["traffic_GBps"] = tb / (avg_ms / 1e3) / 1e9
                result["roofline"]["traffic_frac"] = tb / (avg_ms / 1e3) / HBM_PEAK
                result["roofline"]["note"] = (
                    "achieved counts every node and leaf a ray reads (the algorithmic unit); BVH reuse is served by "
                    "L2 and the Infinity Cache, so the measured L2-to-fabric traffic is lower")
                if "valu_busy" in k:
                    result["roofline"]["valu_busy"] = k["valu_busy"]
                    result["roofline"]["valu_note"] = (
                        "the kernel's vector ALUs issue on this share of its cycles (rocprofv3 SQ_ACTIVE_INST_VALU x 4 "
                        "over 1024 SIMDs x the dispatch cycles, in the profile): the walk is bound by VALU issue, "
                        "with " + f"{100 * k['valu_active_lane_share']:.0f}% of lanes active")
        ref_bytes = 48 * n + 80 * ref_counted["extend_steps"] + 44 * ref_counted["extend_leaves"]
        result["roofline"]["reference_order"] = {
            "steps_per_ray": ref_counted["extend_steps"] / n, "leaves_per_ray": ref_counted["extend_leaves"] / n,
            "equivalent_GBps": ref_bytes / (kt["extend_ms"] / 1e3) / 1e9,
            "note": "the reference's left-first visit order does this many bytes of work for the same hits",
        }
    if not args.no_cpu:
        threads = host_threads()
        result["cpu_baseline"] = cpu_baseline(hot, cold, mats, sky, desc, args.depth, args.seed, args.cpu_seconds,
                                              threads)
    if args.workload == "all":
        del hot, cold, mats
        with mcpt.Context() as ctx:
            result["serving"] = serving_bench(mcpt, ctx, args)
        with mcpt.Context() as ctx:
            result["batch"] = batch_bench(mcpt, ctx, args)
        # the heterogeneous medium at its default size; its oracle sample is
        # kept short so the default run stays within minutes
        cloud_args = argparse.Namespace(**vars(args))
        cloud_args.cpu_seconds = min(args.cpu_seconds, 8.0)
        result["cloud"] = cloud_bench(mcpt, cloud_args)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
