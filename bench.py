"""Benchmark of the WalkRay hot path on one MI355X.

Headline workload (closest-hit ray queries): a 1024 x 1024-vertex fBm
heightfield (2,093,058 triangles, amplitude 0.2, 6 octaves, seeded), its tree
built by the reference builder, 16,777,216 coherent pinhole rays (4096^2,
pixel centres) plus 16,777,216 incoherent rays (origins uniform in the scene
box, directions uniform on the sphere).  One step = closest hit of both ray
sets, inputs already resident in HBM.  value = Mrays/s over the timed steps,
measured with HIP events on the stream the kernels run on.

The same JSON line also reports, outside `value`, the path-tracing workloads
(Cornell box 256^2 x 16 spp and the 512-object field at 1920x1080 x 64 spp,
megakernel vs wavefront samples/s, with and without next-event estimation),
the light sub-paths and photon hash grid of BDPT / VCM / PPM on the field, the
BidirectionalPathTracer, the VCMTracer and PhotonMap's progressive photon mapping on the Cornell box, and the
served-jobs workload (p50/p99 latency and aggregate samples/s under
Poisson arrivals).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-pt] [--legs ch,pt,lp,bdpt,vcm,ppm,motion,serve,c5]

--legs picks the workloads (closest hit, path tracing, light paths + photon
grid, BidirectionalPathTracer, VCMTracer and PhotonMap on the Cornell box, moving objects (per-pass
rebuilds), serving, and c5: the 43.8M-triangle large scene at 3840x2160 x
256 spp; ptd, not in the default
run: C1 and C3 with the wavefront's kernels launched one by one and no graph
created, for kernel tracers); a run without the closest-hit leg reports value
null (used to profile one leg).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
NODE_BYTES = 32   # node bytes are counted in 32 B units (a 64 B binary or wide node counts 2)
TRI_BYTES = 64   # triangle record: vertices, normal, tie key, triangle id, object, chain record
RAY_IN_BYTES = 24
HIT_OUT_BYTES = 16


TRAFFIC_FILE = "profiles/closest_hit_heightfield_2M.md"


def measured_traffic():
    """HBM bytes per closest-hit call from the committed PMC summary (None when absent)."""
    try:
        with open(os.path.join(ROOT, TRAFFIC_FILE)) as f:
            for line in f:
                if "traffic_bytes_per_call:" in line:
                    return int(line.split("traffic_bytes_per_call:")[1].split()[0])
    except OSError:
        pass
    return None


def measured_copy_gbs(ctx, nbytes=1 << 30, reps=5):
    """HBM bandwidth a plain device-to-device copy reaches: bytes read plus written over
    the best of `reps` copies (HIP events on the context stream)."""
    a, b = ctx.alloc(nbytes), ctx.alloc(nbytes)
    ctx.copy(b, a, nbytes)
    ctx.synchronize()
    best = None
    for _ in range(reps):
        ctx.timer_start()
        ctx.copy(b, a, nbytes)
        ms = ctx.timer_stop()
        best = ms if best is None else min(best, ms)
    del a, b
    return round(2 * nbytes / (best * 1e-3) / 1e9, 1)


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def rng_pick(n, k, seed):
    """k distinct indices of range(n), sorted (a spread-out sample of a ray set)."""
    return np.sort(np.random.default_rng(seed).choice(n, k, replace=False))


def closest_hit_bench(W, S, ctx, steps, warmup, n_side=4096, hf_n=1024):
    t0 = time.time()
    objs = S.heightfield_scene(hf_n, seed=7)
    sc = W.Scene(ctx, objs)
    info = sc.info()
    log(f"[bench] scene: {info['n_triangles']} triangles, {info['n_nodes']} nodes, depth {info['depth']}, "
        f"build {info['build_seconds']:.2f}s ({time.time() - t0:.1f}s with generation)")
    cam = S.heightfield_camera(n_side, n_side)
    coh = S.pinhole_rays(cam)
    lo, hi = S.scene_bounds(objs)
    inc = S.incoherent_rays(n_side * n_side, lo, hi, seed=11)
    n = len(coh)
    d_coh = ctx.to_device(coh)
    d_inc = ctx.to_device(inc)
    d_hits = ctx.alloc(16 * n)
    log(f"[bench] rays resident: 2 x {n} ({time.time() - t0:.1f}s)")

    # autotune the persistent kernel's launch knobs on a sample of both sets (results never depend on them)
    t1 = time.time()
    k = min(n, 1 << 22)
    sample = ctx.to_device(np.concatenate([coh[rng_pick(n, k, 1)], inc[:k]]))
    tuning = sc.autotune_closest_hit(sample, 2 * k)
    del sample
    log(f"[bench] closest-hit autotune: {tuning} ({time.time() - t1:.1f}s)")

    # counter-enabled run on the same rays: box records and triangle records fetched
    counters = {}
    for name, d in (("coherent", d_coh), ("incoherent", d_inc)):
        dc = ctx.to_device(np.zeros(3, np.uint64))
        sc.closest_hit_device(d, n, d_hits, dc)
        ctx.synchronize()
        c = dc.download(np.uint64, 3)
        counters[name] = {"box_records": int(c[0]), "tri_records": int(c[1]), "fallback_rays": int(c[2])}
    hits_inc = d_hits.download(np.uint32, 4 * n).view(W.HIT_DTYPE)
    hit_frac_inc = float((hits_inc["prim"] != W.NONE).mean())
    assert not np.any(hits_inc["prim"] == W.OVERFLOW)

    for _ in range(warmup):
        sc.closest_hit_device(d_coh, n, d_hits)
        sc.closest_hit_device(d_inc, n, d_hits)
    ctx.synchronize()
    # timed region: K steps, each = both ray sets; per-set kernel time from events
    ms_set = {"coherent": 0.0, "incoherent": 0.0}
    ctx.synchronize()
    wall0 = time.perf_counter()
    for _ in range(steps):
        for name, d in (("coherent", d_coh), ("incoherent", d_inc)):
            ctx.timer_start()
            sc.closest_hit_device(d, n, d_hits)
            ms_set[name] += ctx.timer_stop()
    ctx.synchronize()
    wall = time.perf_counter() - wall0
    kern_ms = ms_set["coherent"] + ms_set["incoherent"]
    ms_per_step = kern_ms / steps
    mrays = 2 * n / (ms_per_step * 1e-3) / 1e6

    def alg_bytes(c):
        return n * (RAY_IN_BYTES + HIT_OUT_BYTES) + NODE_BYTES * c["box_records"] + TRI_BYTES * c["tri_records"]

    per_set = {}
    for name in ("coherent", "incoherent"):
        ms = ms_set[name] / steps
        b = alg_bytes(counters[name])
        per_set[name] = {
            "mrays_per_s": n / (ms * 1e-3) / 1e6,
            "ms_per_launch": ms,
            "alg_bytes_per_launch": b,
            "achieved_gbs": b / (ms * 1e-3) / 1e9,
            "frac": b / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "box_records_per_ray": counters[name]["box_records"] / n,
            "tri_records_per_ray": counters[name]["tri_records"] / n,
            "fallback_rays": counters[name]["fallback_rays"],
        }
    # one wr_closest_hit call (persistent traversal kernel + finish kernel) per ray set, two per
    # step: per-call algorithmic bytes and duration are the averages over the two calls
    total_b = alg_bytes(counters["coherent"]) + alg_bytes(counters["incoherent"])
    avg_launch_ms = kern_ms / steps / 2
    roofline = {
        "bound": "hbm",
        "kernel": "k_closest_hit + k_closest_hit_finish (the two kernels of one wr_closest_hit call)",
        "achieved": round(total_b / 2 / (avg_launch_ms * 1e-3) / 1e9, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(total_b / 2 / (avg_launch_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "traffic": measured_traffic(),
        "traffic_source": TRAFFIC_FILE + " (rocprofv3 FETCH_SIZE and WRITE_SIZE per call, calibrated on a 1 GiB copy)",
        "avg_launch_ms": round(avg_launch_ms, 4),
        "alg_bytes_per_launch": int(total_b / 2),
    }
    roofline["launch_tuning"] = {kk: tuning[kk] for kk in ("refill_min", "node_phase_min", "blocks_per_cu", "lds_levels",
                                                           "batch")}
    return {
        "scene": sc, "objs": objs, "coh": coh, "inc": inc, "n": n, "mrays": mrays, "ms_per_step": ms_per_step,
        "wall_s": wall, "per_set": per_set, "roofline": roofline, "info": info, "hit_frac_inc": hit_frac_inc,
        "d_coh": d_coh, "d_inc": d_inc, "d_hits": d_hits,
    }


def cpu_baseline(W, ctx, r, tie_tol=1e-6):
    """The oracle (faithful reference traversal, unpruned) on the whole workload: both
    16,777,216-ray sets, OpenMP over rays on the host cores.  Its answers are then the
    checker of the GPU's on every ray: prim ids, t and barycentrics bit for bit, and the
    rays whose hit has a near-tie (another triangle accepted within |dt| < 1e-6,
    counted by a second oracle pass that is not timed)."""
    from oracle import oracle as O
    threads = min(os.cpu_count() or 1, 16)
    objs, n = r["objs"], r["n"]
    t0 = time.perf_counter()
    osc = O.Scene(objs)
    build_s = time.perf_counter() - t0
    parity = {"rays": 0, "prim_equal": 0, "t_bits_equal": 0, "bary_bits_equal": 0, "near_tie_rays": 0,
              "mismatch_rays": 0, "mismatch_rays_without_near_tie": 0}
    dt = 0.0
    for name, rays, d in (("coherent", r["coh"], r["d_coh"]), ("incoherent", r["inc"], r["d_inc"])):
        t1 = time.perf_counter()
        o = osc.intersect(rays, nthreads=threads)
        dt += time.perf_counter() - t1
        r["scene"].closest_hit_device(d, n, r["d_hits"])
        ctx.synchronize()
        g = r["d_hits"].download(np.uint32, 4 * n).view(W.HIT_DTYPE)
        same_prim = g["prim"] == o["prim"]
        same_t = g["t"].view(np.uint32) == o["t"].view(np.uint32)
        hit = o["prim"] != O.NONE
        same_b = (g["b1"].view(np.uint32) == o["b1"].view(np.uint32)) & \
                 (g["b2"].view(np.uint32) == o["b2"].view(np.uint32))
        same_b |= ~hit  # a miss has no barycentrics
        ties = osc.near_ties(rays, o["prim"], o["t"], tie_tol, nthreads=threads)
        bad = ~(same_prim & same_t & same_b)
        parity["rays"] += n
        parity["prim_equal"] += int(same_prim.sum())
        parity["t_bits_equal"] += int(same_t.sum())
        parity["bary_bits_equal"] += int(same_b.sum())
        parity["near_tie_rays"] += int((ties > 0).sum())
        parity["mismatch_rays"] += int(bad.sum())
        parity["mismatch_rays_without_near_tie"] += int((bad & (ties == 0)).sum())
        log(f"[bench] oracle parity {name}: {int(bad.sum())} of {n} rays differ, {int((ties > 0).sum())} near-ties")
    return {
        "value": round(2 * n / dt / 1e6, 4),
        "unit": "Mrays/s",
        "cores": threads,
        "kind": "port",
        "sample": f"the whole workload: {n} coherent + {n} incoherent rays, oracle KDTree::intersect restatement "
                  f"(unpruned recursive traversal, Scene::intersect with barycentrics), OpenMP over rays; oracle "
                  f"tree build {build_s:.1f}s single-thread not included",
        "build_seconds": round(build_s, 2),
        "parity": parity,
        "near_tie_tol": tie_tol,
    }


def cpu_baseline_pt(objs, cam, spp, depth, seed, passes=None, rebuild=True):
    """The oracle's PathTracer on the host cores: the scene build (the reference
    builder for the global and every object tree), the render including the
    reference's full rebuild before every pass (pathTracer.cpp:18 -> scene.cpp:67-80),
    and the render alone.  passes: the bounded sample of passes timed (all by default)."""
    from oracle import oracle as O
    threads = min(os.cpu_count() or 1, 16)
    passes = passes or spp
    t0 = time.perf_counter()
    osc = O.Scene(objs)
    build_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    osc.render(cam, passes, depth, seed, nthreads=threads)
    render_s = time.perf_counter() - t0
    with_rebuild_s = None
    if rebuild:
        t0 = time.perf_counter()
        for k in range(passes):
            osc.rebuild()
            osc.render(cam, 1, depth, seed, pass0=k, nthreads=threads)
        with_rebuild_s = time.perf_counter() - t0
    samples = cam["width"] * cam["height"] * passes
    out = {"value": round(samples / render_s, 1), "unit": "samples/s", "cores": threads, "kind": "port",
           "sample": f"{passes} of {spp} passes of the {cam['width']}x{cam['height']} image, oracle PathTracer "
                     "restatement, OpenMP over pixels",
           "build_seconds": round(build_s, 3), "render_seconds": round(render_s, 3)}
    if with_rebuild_s is not None:
        out["render_with_per_pass_rebuild_seconds"] = round(with_rebuild_s, 3)
        out["samples_per_s_with_rebuild"] = round(samples / with_rebuild_s, 1)
    return out


def cpu_baseline_vcm(S, passes=2, depth=8, radius=0.01):
    """The oracle's VCMTracer restatement on the host cores, over a bounded sample of the VCM
    leg's workload (2 passes of the same 256x256 Cornell image)."""
    from oracle import oracle as O
    threads = min(os.cpu_count() or 1, 16)
    objs, cam = S.cornell(k=5, width=256, height=256)
    osc = O.Scene(objs)
    t0 = time.perf_counter()
    osc.render_vcm(cam, passes, depth, 1234, radius, nthreads=threads)
    dt = time.perf_counter() - t0
    return {"value": round(cam["width"] * cam["height"] * passes / dt, 1), "unit": "samples/s", "cores": threads,
            "kind": "port", "sample": f"{passes} passes of the same 256x256 image, oracle VCMTracer restatement, "
                                      "OpenMP over pixels"}


def cpu_baseline_ppm(S, passes=4, depth=8, radius=0.01):
    """The oracle's PhotonMap restatement on the host cores, over a bounded sample of the PPM
    leg's workload (4 passes of the same 256x256 Cornell image)."""
    from oracle import oracle as O
    threads = min(os.cpu_count() or 1, 16)
    objs, cam = S.cornell(k=5, width=256, height=256)
    osc = O.Scene(objs)
    t0 = time.perf_counter()
    osc.render_ppm(cam, passes, depth, 1234, radius, nthreads=threads)
    dt = time.perf_counter() - t0
    return {"value": round(cam["width"] * cam["height"] * passes / dt, 1), "unit": "samples/s", "cores": threads,
            "kind": "port", "sample": f"{passes} passes of the same 256x256 image, oracle PhotonMap restatement, "
                                      "OpenMP over pixels"}


def cpu_baseline_bdpt(S, passes=2, depth=8):
    """The oracle's BidirectionalPathTracer restatement on the host cores, over a bounded
    sample of the BDPT leg's workload (2 passes of the same 256x256 Cornell image)."""
    from oracle import oracle as O
    threads = min(os.cpu_count() or 1, 16)
    objs, cam = S.cornell(k=5, width=256, height=256)
    osc = O.Scene(objs)
    t0 = time.perf_counter()
    osc.render_bdpt(cam, passes, depth, 1234, nthreads=threads)
    dt = time.perf_counter() - t0
    return {"value": round(cam["width"] * cam["height"] * passes / dt, 1), "unit": "samples/s", "cores": threads,
            "kind": "port", "sample": f"{passes} passes of the same 256x256 image, oracle BidirectionalPathTracer "
                                      "restatement, OpenMP over pixels"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-pt", action="store_true", help="same as --legs ch")
    ap.add_argument("--legs", default="ch,pt,lp,bdpt,vcm,ppm,motion,serve,c5")
    ap.add_argument("--side", type=int, default=4096)
    ap.add_argument("--copy-bytes", type=int, default=0,
                    help="first run one device-to-device copy of this many bytes (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "multi-GPU sharding is described in DESIGN.md but not built.", file=sys.stderr)
        sys.exit(2)
    from walkray_amd import _lib as W
    from walkray_amd import scenes as S

    ctx = W.Context(0)
    if args.copy_bytes:
        a, b = ctx.alloc(args.copy_bytes), ctx.alloc(args.copy_bytes)
        ctx.copy(b, a, args.copy_bytes)
        ctx.synchronize()
        del a, b
    copy_gbs = measured_copy_gbs(ctx)
    log(f"[bench] 1 GiB device copy: {copy_gbs} GB/s (read + write) against the {HBM_PEAK_GBS:.0f} GB/s nominal peak")
    legs = {"ch"} if args.no_pt else set(args.legs.split(","))
    r = None
    if "ch" in legs:
        r = closest_hit_bench(W, S, ctx, args.steps, args.warmup, n_side=args.side)
        log(f"[bench] closest-hit: {r['mrays']:.1f} Mrays/s ({r['ms_per_step']:.3f} ms/step)")
    extra = {}
    if "pt" in legs:
        from walkray_amd import render as R
        extra["path_tracing"] = R.bench_path_tracing(W, S, ctx, log)
        if not args.no_cpu:
            objs, cam = S.cornell(k=5, width=256, height=256)
            c1 = extra["path_tracing"]["c1"]["cpu_baseline"] = cpu_baseline_pt(objs, cam, 16, 8, 1234)
            log(f"[bench] C1 oracle PathTracer: {c1['value'] / 1e6:.2f} Msamples/s on {c1['cores']} cores, build "
                f"{c1['build_seconds']}s, 16 passes with the per-pass rebuild {c1['render_with_per_pass_rebuild_seconds']}s")
            objs, cam = S.random_field()
            c3 = extra["path_tracing"]["c3"]["cpu_baseline"] = cpu_baseline_pt(objs, cam, 64, 16, 77, passes=1)
            log(f"[bench] C3 oracle PathTracer: {c3['value'] / 1e6:.2f} Msamples/s on {c3['cores']} cores "
                f"(1 of 64 passes), build {c3['build_seconds']}s")
    if "ptd" in legs:  # for kernel tracers: no graph is created or launched
        from walkray_amd import render as R
        extra["path_tracing_direct"] = R.profile_path_tracing(ctx, log)
    if "lp" in legs:
        from walkray_amd import render as R
        extra["light_paths"] = R.bench_light_paths(ctx, log)
    if "bdpt" in legs:
        from walkray_amd import render as R
        extra["bdpt"] = R.bench_bdpt(ctx, log)
        if not args.no_cpu:
            extra["bdpt"]["cpu_baseline"] = cpu_baseline_bdpt(S)
            log(f"[bench] BDPT oracle: {extra['bdpt']['cpu_baseline']['value'] / 1e3:.1f} Ksamples/s on "
                f"{extra['bdpt']['cpu_baseline']['cores']} cores")
    if "vcm" in legs:
        from walkray_amd import render as R
        extra["vcm"] = R.bench_vcm(ctx, log)
        if not args.no_cpu:
            extra["vcm"]["cpu_baseline"] = cpu_baseline_vcm(S)
            log(f"[bench] VCM oracle: {extra['vcm']['cpu_baseline']['value'] / 1e3:.1f} Ksamples/s on "
                f"{extra['vcm']['cpu_baseline']['cores']} cores")
    if "ppm" in legs:
        from walkray_amd import render as R
        extra["ppm"] = R.bench_ppm(ctx, log)
        if not args.no_cpu:
            extra["ppm"]["cpu_baseline"] = cpu_baseline_ppm(S)
            log(f"[bench] PPM oracle: {extra['ppm']['cpu_baseline']['value'] / 1e3:.1f} Ksamples/s on "
                f"{extra['ppm']['cpu_baseline']['cores']} cores")
    if "motion" in legs:
        from walkray_amd import render as R
        extra["motion"] = R.bench_motion(ctx, log)
    if "c5" in legs:  # the large scene: about a minute of host build and rendering
        from walkray_amd import render as R
        extra["large_scene"] = R.bench_large(W, S, ctx, log)
    if "serve" in legs:
        from walkray_amd import serve as V
        extra["serving"] = V.bench_serving(ctx, log)
    out = {
        "metric": "Mrays/s closest-hit + path samples/s at fixed RMSE; p50/p99 job latency",
        "value": round(r["mrays"], 2) if r else None,
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r["ms_per_step"], 4) if r else None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "closest-hit ray queries: 1024x1024 fBm heightfield (2,093,058 triangles, reference "
                        "builder); 16,777,216 coherent 4096^2 pinhole rays + 16,777,216 incoherent rays; "
                        "outputs prim id, t, barycentrics (FP32)",
            "triangles": r["info"]["n_triangles"] if r else None,
            "rays_per_step": 2 * r["n"] if r else None,
        },
    }
    if r:
        out["roofline"] = r["roofline"]
        out["roofline"]["measured_copy_gbs"] = copy_gbs
        out["closest_hit"] = {k: {kk: (round(vv, 4) if isinstance(vv, float) else vv) for kk, vv in v.items()}
                              for k, v in r["per_set"].items()}
    out.update(extra)
    if r and not args.no_cpu:
        log("[bench] cpu baseline (oracle) ...")
        out["cpu_baseline"] = cpu_baseline(W, ctx, r)
        log(f"[bench] cpu baseline: {out['cpu_baseline']['value']} Mrays/s on {out['cpu_baseline']['cores']} cores")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
