"""Path-tracing helpers over the C ABI: image statistics, benchmarks and smoke check."""
from __future__ import annotations

import os
import time

import numpy as np

from . import _lib as W
from . import scenes as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
C1_REFERENCE = os.path.join(ROOT, "tests", "golden", "cornell256_ref.npz")
C3_REFERENCE = os.path.join(ROOT, "tests", "golden", "field1080_ref_sub8.npz")


def film_stats(state):
    """Per pixel/channel mean and standard error of the mean from a film state."""
    n = max(int(state["passes"]), 1)
    s = state["sum"].astype(np.float64)
    q = state["sumsq"].astype(np.float64)
    mean = s / n
    var = np.maximum(q / n - mean * mean, 0.0)
    return mean, np.sqrt(var / n)


def rel_rmse(img, ref):
    """||I - R||_2 / ||R||_2 over all pixels and channels."""
    img = np.asarray(img, np.float64)
    ref = np.asarray(ref, np.float64)
    return float(np.linalg.norm(img - ref) / max(np.linalg.norm(ref), 1e-300))


def expected_rel_rmse(state):
    """Expected ||I - E[I]|| / ||E[I]|| of a film from its own per-pixel sample variance."""
    mean, se = film_stats(state)
    return float(np.sqrt((se ** 2).sum()) / max(np.linalg.norm(mean), 1e-300))


def render_scene(ctx, objs, cam, spp, max_depth, seed, mode, scene=None, wave_slots=0, nee=False, launch=0):
    sc = scene or W.Scene(ctx, objs)
    film = W.Film(ctx, cam["width"], cam["height"])
    st = W.render(sc, cam, film, spp, max_depth, seed, mode, wave_slots=wave_slots, nee=nee, launch=launch)
    out = film.read()
    out["stats"] = st
    return out


def _compare(ctx, sc, cam, spp, depth, seed, log, label, repeats=2, nee=False):
    res = {}
    films = {}
    for name, mode in (("megakernel", W.WR_MEGAKERNEL), ("wavefront", W.WR_WAVEFRONT)):
        best = None
        for _ in range(repeats):
            film = W.Film(ctx, cam["width"], cam["height"])
            st = W.render(sc, cam, film, spp, depth, seed, mode, nee=nee)
            best = st["kernel_ms"] if best is None else min(best, st["kernel_ms"])
            films[name] = film.read()
        samples = cam["width"] * cam["height"] * spp
        res[name] = {"ms": round(best, 3), "samples_per_s": round(samples / (best * 1e-3), 1)}
        log(f"[bench] {label} {name}: {best:.2f} ms, {samples / (best * 1e-3) / 1e6:.1f} Msamples/s")
    res["speedup_wavefront_vs_megakernel"] = round(res["wavefront"]["samples_per_s"] /
                                                   res["megakernel"]["samples_per_s"], 3)
    # same estimator and the same samples: the two films (hence their RMSE) are identical
    res["films_bit_identical"] = bool(np.array_equal(films["megakernel"]["sum"].view(np.uint32),
                                                     films["wavefront"]["sum"].view(np.uint32)))
    res["expected_rel_rmse"] = round(expected_rel_rmse(films["wavefront"]), 5)
    return res, films["wavefront"]


def _at_fixed_rmse(r, spp, cam, target=0.01):
    """Seconds each engine needs for `target` relative RMSE against the reference: the error
    falls as 1/sqrt(spp) (no bias: checked from 16 to 4,096 spp on the Cornell box), so
    spp_needed = spp * (rmse / target)^2 at the measured samples/s."""
    need = spp * (r["rel_rmse_vs_reference"] / target) ** 2
    samples = need * cam["width"] * cam["height"]
    r["at_1pct_rmse"] = {"spp": round(need), **{
        f"{m}_s": round(samples / r[m]["samples_per_s"], 2) for m in ("megakernel", "wavefront")}}


# Path-state bytes per path segment that each wavefront stage reads or writes
# (32-bit SoA words of the path buffers, plus the hit's 64 B triangle record in
# shade); the trace stage adds the fast pass's node (32 B units) and triangle
# (64 B) records counted by its counting instance.
STAGE_BYTES = {
    "trace": 40 + 16,          # o, d, flags, job, inside, contact read; ht, hrec, hobj, flags written
    "finish": 60 + 8,          # o, d, flags, job, ht, hrec, hobj, inside, contact, pixel, pass read; flags, inside written
    "shade_read": 76 + 64,     # 19 words of state and the hit's triangle record
    "shade_write": 52,         # 13 words of a surviving path's next state
    "sample": 12,              # a finished path's sample
    "generate": 52,            # a camera path's 13 words
    "fold": 12 + 72,           # a sample read, the Kahan sum, compensation and square sum read and written
}


def stage_roofline(W_, ctx, sc, cam, spp, depth, seed, log, label, nee=False):
    """One render with WR_RENDER_STAGE_STATS: per iteration kernel, its device time summed
    over the render, the algorithmic bytes it moves (STAGE_BYTES per segment, traversal
    records counted by the trace stage) and the resulting fraction of the HBM peak."""
    film = W.Film(ctx, cam["width"], cam["height"])
    W.render(sc, cam, film, spp, depth, seed, W.WR_WAVEFRONT, nee=nee, launch=W.WR_RENDER_STAGE_STATS)
    st = W.render_stage_stats(ctx)
    T, X, G = st["segments"], st["exits"], st["generated"]
    nodes, tris = st["box_records"], st["tri_records"]
    b = {
        "trace": T * STAGE_BYTES["trace"] + 32 * nodes + 64 * tris,
        "finish": T * STAGE_BYTES["finish"] + G * STAGE_BYTES["sample"],
        "shade": (T - X) * STAGE_BYTES["shade_read"] + (T - G) * STAGE_BYTES["shade_write"] + G * STAGE_BYTES["sample"],
        "shade_exits": X * (STAGE_BYTES["shade_read"] + STAGE_BYTES["shade_write"]),
        "generate": G * STAGE_BYTES["generate"],
        "fold": G * STAGE_BYTES["fold"],
    }
    stages = {}
    for k, ms in st["ms"].items():
        e = {"ms": round(ms, 3)}
        if k in b and ms > 0:
            e["alg_bytes"] = int(b[k])
            e["achieved_gbs"] = round(b[k] / (ms * 1e-3) / 1e9, 1)
            e["frac"] = round(b[k] / (ms * 1e-3) / 1e9 / 8000.0, 4)
        stages[k] = e
    total = sum(st["ms"].values())
    out = {"stages": stages, "total_ms": round(total, 3), "segments": T, "exits": X, "camera_paths": G,
           "box_records_per_segment": round(nodes / max(T, 1), 2), "tri_records_per_segment": round(tris / max(T, 1), 2),
           "launch": "kernels launched one by one with a HIP event around each (the graph replays' sequence)",
           "bytes_per_segment": STAGE_BYTES}
    log(f"[bench] {label} stages (ms): " + ", ".join(f"{k} {v['ms']}" for k, v in stages.items()) +
        f"; trace {stages['trace'].get('achieved_gbs')} GB/s")
    return out


def bench_path_tracing(W_, S_, ctx, log, c3_spp=64):
    out = {}
    objs, cam = S.cornell(k=5, width=256, height=256)
    sc = W.Scene(ctx, objs)
    tune = W.autotune_render(sc, cam, 8, spp=4)
    log(f"[bench] C1 wavefront autotune: {tune}")
    r, film = _compare(ctx, sc, cam, 16, 8, 1234, log, "C1 cornell")
    r["wavefront_tuning"] = tune
    r["workload"] = ("Cornell box + 2 icospheres k=5 (40,974 triangles), 256x256, 16 spp, max depth 8, "
                     "Russian roulette at every diffuse/Phong bounce (reference rule)")
    if os.path.exists(C1_REFERENCE):
        ref = np.load(C1_REFERENCE)
        r["rel_rmse_vs_reference"] = round(rel_rmse(film["mean"], ref["mean"]), 5)
        r["reference"] = f"{int(ref['spp'])}-spp render, FP64 accumulation of per-batch Kahan films"
        _at_fixed_rmse(r, 16, cam)
    r["next_event"], _ = _compare(ctx, sc, cam, 16, 8, 1234, log, "C1 cornell +NEE", nee=True)
    r["stage_roofline"] = stage_roofline(W, ctx, sc, cam, 16, 8, 1234, log, "C1")
    # steady state: enough passes that the pool stays full and the drain tail is a small part;
    # the pool is tuned again on a workload that can fill it (64 spp: 4M paths)
    ss = {"wavefront_tuning": W.autotune_render(sc, cam, 8, spp=64)}
    for name, mode in (("megakernel", W.WR_MEGAKERNEL), ("wavefront", W.WR_WAVEFRONT)):
        film = W.Film(ctx, cam["width"], cam["height"])
        st = W.render(sc, cam, film, 1024, 8, 1234, mode)
        ss[name] = {"ms": round(st["kernel_ms"], 2),
                    "samples_per_s": round(cam["width"] * cam["height"] * 1024 / (st["kernel_ms"] * 1e-3), 1)}
    ss["speedup_wavefront_vs_megakernel"] = round(ss["wavefront"]["samples_per_s"] / ss["megakernel"]["samples_per_s"], 3)
    r["steady_state_1024spp"] = ss
    log(f"[bench] C1 1024 spp: megakernel {ss['megakernel']['samples_per_s'] / 1e6:.1f}, wavefront "
        f"{ss['wavefront']['samples_per_s'] / 1e6:.1f} Msamples/s ({ss['speedup_wavefront_vs_megakernel']}x; "
        f"tuning {ss['wavefront_tuning']})")
    out["c1"] = r
    objs, cam = S.random_field()
    sc = W.Scene(ctx, objs)
    # 8 passes (16.6M paths): enough for the 1M-8M path pools to differ as they do over the full render
    tune = W.autotune_render(sc, cam, 16, spp=8)
    log(f"[bench] C3 wavefront autotune: {tune}")
    r, film = _compare(ctx, sc, cam, c3_spp, 16, 77, log, "C3 field", repeats=1)
    r["wavefront_tuning"] = tune
    r["workload"] = (f"512 random icospheres/boxes on a ground plane ({sc.info()['n_triangles']} triangles), "
                     f"diffuse/Phong/refractive, 2 area lights, 1920x1080, {c3_spp} spp, max depth 16")
    if os.path.exists(C3_REFERENCE):
        ref = np.load(C3_REFERENCE)
        st, off = int(ref["stride"]), int(ref["offset"])
        img = np.asarray(film["mean"]).reshape(cam["height"], cam["width"], 3)[off::st, off::st]
        r["rel_rmse_vs_reference"] = round(rel_rmse(img, ref["mean"]), 5)
        r["reference"] = (f"{int(ref['spp'])}-spp render, FP64 accumulation of per-batch Kahan films, "
                          f"compared on every {st}th pixel in x and y ({img.shape[1]} x {img.shape[0]})")
        _at_fixed_rmse(r, c3_spp, cam)
    r["next_event"], _ = _compare(ctx, sc, cam, c3_spp, 16, 77, log, "C3 field +NEE", repeats=1, nee=True)
    r["stage_roofline"] = stage_roofline(W, ctx, sc, cam, c3_spp, 16, 77, log, "C3")
    out["c3"] = r
    return out


def profile_path_tracing(ctx, log, c3_spp=16):
    """The C1 and C3 renders for a kernel tracer: both engines, the wavefront's kernels
    launched one by one (WR_RENDER_DIRECT) and no graph created, the default engine
    configuration (no autotuning).  Two renders of each, the second the one to read."""
    out = {}
    for name, (objs, cam), spp, depth, seed in (("c1", S.cornell(k=5, width=256, height=256), 16, 8, 1234),
                                                ("c3", S.random_field(), c3_spp, 16, 77)):
        sc = W.Scene(ctx, objs)
        r = {}
        for mode_name, mode, launch in (("megakernel", W.WR_MEGAKERNEL, 0),
                                        ("wavefront", W.WR_WAVEFRONT, W.WR_RENDER_DIRECT)):
            for _ in range(2):
                film = W.Film(ctx, cam["width"], cam["height"])
                st = W.render(sc, cam, film, spp, depth, seed, mode, launch=launch)
            r[mode_name] = {"ms": round(st["kernel_ms"], 3), "iterations": st["iterations"]}
            log(f"[bench] {name} {mode_name} (direct launches): {st['kernel_ms']:.2f} ms")
        out[name] = r
    return out


def bench_light_paths(ctx, log, depth=16, radius=0.01, repeats=3):
    """The front end of BDPT / VCM / PPM on the 512-object field: one light sub-path per
    1080p pixel (maxPathLen 16), the photons VCM keeps (after one leading point per
    pixel, as its vector starts), the hash grid (one cell per pixel, VCM's base radius)
    and one range query per photon.  Device times from HIP events, best of `repeats`."""
    objs, cam = S.random_field()
    sc = W.Scene(ctx, objs)
    npix = cam["width"] * cam["height"]
    maxv = depth + 3
    best = {}

    def timed(name, fn):
        ctx.synchronize()
        ctx.timer_start()
        out = fn()
        ms = ctx.timer_stop()
        best[name] = min(best.get(name, ms), ms)
        return out

    for _ in range(repeats):
        dv, dc = timed("light_paths", lambda: sc.light_paths_device(depth, 11, 0, 0, npix, maxv))
        dp, ds, m = timed("photons", lambda: sc.photons_device(dv, dc, npix, maxv, n_leading=npix, cap=maxv * npix))
        g = timed("grid_build", lambda: W.HashGrid(ctx, dp, m, radius, npix))
        # one query at every stored light vertex (the leading points are not queried: VCM's
        # queries are eye-path vertices)
        nq = m - npix
        dq = W.DeviceView(dp, 12 * npix)
        dq_counts = ctx.alloc(4 * max(nq, 1))
        timed("query", lambda: g.query_device(dq, nq, 0, dq_counts, None))
        counts = dq_counts.download(np.uint32, nq)
        g.close()
        del dv, dc
    r = {
        "workload": (f"512-object field ({sc.info()['n_triangles']} triangles), {npix} light sub-paths "
                     f"(one per 1920x1080 pixel), maxPathLen {depth}; photons after {npix} leading points; "
                     f"hash grid of {npix} cells, radius {radius}; one range query per point"),
        "light_paths_ms": round(best["light_paths"], 3),
        "light_paths_per_s": round(npix / (best["light_paths"] * 1e-3), 1),
        "photons_ms": round(best["photons"], 3),
        "points": int(m),
        "grid_build_ms": round(best["grid_build"], 3),
        "queries": int(nq),
        "query_ms": round(best["query"], 3),
        "queries_per_s": round(nq / (best["query"] * 1e-3), 1),
        "mean_neighbours": round(float(counts.mean()), 3),
    }
    log(f"[bench] light paths: {r['light_paths_ms']} ms ({r['light_paths_per_s'] / 1e6:.1f} M paths/s), photons "
        f"{r['photons_ms']} ms ({m} points), grid build {r['grid_build_ms']} ms, {nq} queries {r['query_ms']} ms")
    return r


def bench_bdpt(ctx, log, spp=16, depth=8, repeats=2):
    """BidirectionalPathTracer on the C1 Cornell box (256^2, 16 spp, maxPathLen 8): GPU
    samples/s, best of `repeats`."""
    objs, cam = S.cornell(k=5, width=256, height=256)
    sc = W.Scene(ctx, objs)
    npix = cam["width"] * cam["height"]
    best, film = None, None
    for _ in range(repeats):
        film = W.Film(ctx, cam["width"], cam["height"])
        st = W.render_bdpt(sc, cam, film, spp, depth, 1234)
        best = st["kernel_ms"] if best is None else min(best, st["kernel_ms"])
    f = film.read()
    r = {"workload": f"BidirectionalPathTracer, Cornell box + 2 icospheres k=5, 256x256, {spp} spp, maxPathLen {depth}",
         "ms": round(best, 3), "samples_per_s": round(npix * spp / (best * 1e-3), 1),
         "mean_radiance": round(float(f["mean"].mean()), 5)}
    log(f"[bench] BDPT C1: {best:.1f} ms, {r['samples_per_s'] / 1e6:.2f} Msamples/s")
    return r


def bench_vcm(ctx, log, spp=16, depth=8, radius=0.01, repeats=2):
    """VCMTracer on the C1 Cornell box (256^2, 16 spp, maxPathLen 8, the reference's default
    base radius 0.01 in the unit box): GPU samples/s, best of `repeats`."""
    objs, cam = S.cornell(k=5, width=256, height=256)
    sc = W.Scene(ctx, objs)
    npix = cam["width"] * cam["height"]
    best, film = None, None
    for _ in range(repeats):
        film = W.Film(ctx, cam["width"], cam["height"])
        st = W.render_vcm(sc, cam, film, spp, depth, 1234, radius)
        best = st["kernel_ms"] if best is None else min(best, st["kernel_ms"])
    f = film.read()
    r = {"workload": f"VCMTracer, Cornell box + 2 icospheres k=5, 256x256, {spp} spp, maxPathLen {depth}, "
                     f"base radius {radius}",
         "ms": round(best, 3), "samples_per_s": round(npix * spp / (best * 1e-3), 1),
         "mean_radiance": round(float(f["mean"].mean()), 5)}
    log(f"[bench] VCM C1: {best:.1f} ms, {r['samples_per_s'] / 1e6:.2f} Msamples/s")
    return r


def bench_ppm(ctx, log, spp=16, depth=8, radius=0.01, repeats=2):
    """PhotonMap's progressive photon mapping on the C1 Cornell box (256^2, 16 spp, light
    paths of maxPathLen 8, base radius 0.01 in the unit box): GPU samples/s, best of `repeats`."""
    objs, cam = S.cornell(k=5, width=256, height=256)
    sc = W.Scene(ctx, objs)
    npix = cam["width"] * cam["height"]
    best, film = None, None
    for _ in range(repeats):
        film = W.Film(ctx, cam["width"], cam["height"])
        st = W.render_ppm(sc, cam, film, spp, depth, 1234, radius)
        best = st["kernel_ms"] if best is None else min(best, st["kernel_ms"])
    f = film.read()
    r = {"workload": f"PhotonMap (progressive photon mapping), Cornell box + 2 icospheres k=5, 256x256, {spp} spp, "
                     f"light paths of maxPathLen {depth}, base radius {radius}",
         "ms": round(best, 3), "samples_per_s": round(npix * spp / (best * 1e-3), 1),
         "mean_radiance": round(float(f["mean"].mean()), 5)}
    log(f"[bench] PPM C1: {best:.1f} ms, {r['samples_per_s'] / 1e6:.2f} Msamples/s")
    return r


def bench_large(W_, S_, ctx, log, spp=256, depth=32):
    """Large scene: 43.8M triangles, 3840x2160, 256 spp, depth 32; build time, both engines."""
    t0 = time.time()
    objs, cam = S.large_scene()
    ntri = sum(len(o["tris"]) for o in objs)
    log(f"[bench] C5 large: {len(objs)} objects, {ntri} triangles generated ({time.time() - t0:.1f}s)")
    t1 = time.time()
    sc = W.Scene(ctx, objs)
    del objs
    info = sc.info()
    log(f"[bench] C5 large: scene built and uploaded in {time.time() - t1:.1f}s (host build {info['build_seconds']:.1f}s, "
        f"{info['device_bytes'] / 2**30:.2f} GiB on the device, depth {info['depth']})")
    tune = W.autotune_render(sc, cam, depth, spp=1)
    log(f"[bench] C5 wavefront autotune: {tune}")
    r, film = _compare(ctx, sc, cam, spp, depth, 2024, log, "C5 large", repeats=1)
    r["wavefront_tuning"] = tune
    r["workload"] = (f"4096^2 fBm heightfield + 2,000 Phong icospheres k=4 + sun disk + sky dome ({ntri} triangles), "
                     f"3840x2160, {spp} spp, max depth {depth}")
    r["build_seconds"] = round(info["build_seconds"], 2)
    r["device_gib"] = round(info["device_bytes"] / 2**30, 3)
    r["mean_radiance"] = round(float(film["mean"].mean()), 5)
    return r


def bench_motion(ctx, log, spp=16):
    """Moving objects (Scene::updateSceneForMotionBlur before every pass, scene.cpp:67-80):
    the Cornell box with its diffuse sphere moving (256^2, max depth 8) and the 512-object
    field with 8 moving objects (480x270, max depth 16), both engines.  Per pass the
    global tree and the moving objects' trees are rebuilt and the static objects' trees
    kept; for comparison, the time of building and uploading every tree of the placed
    scene (what a whole rebuild costs) is measured on the same scene."""
    out = {}
    objs, cam, motions = S.cornell_moving(k=5)
    fobjs, fcam = S.random_field(width=480, height=270)
    fmot = {}
    for i in range(1, 9):  # the first eight field objects slide sideways during the exposure
        fmot[i] = (0.0, 1.0, np.eye(4, dtype=np.float32), S.trs([0.05 * i / 8, 0.0, 0.02], [0, 1, 0], 0.0, [1, 1, 1]))
    for name, (ob, cm, mot, depth) in (("c1_moving_sphere", (objs, cam, motions, 8)),
                                       ("field_8_moving", (fobjs, fcam, fmot, 16))):
        sc = W.Scene(ctx, ob, build=False)
        for o, (t0, t1, a, b) in mot.items():
            sc.set_motion(o, t0, t1, a, b)
        sc.build()
        r = {"workload": f"{len(ob)} objects ({len(mot)} moving), {cm['width']}x{cm['height']}, {spp} spp, "
                         f"max depth {depth}"}
        for mode_name, mode in (("megakernel", W.WR_MEGAKERNEL), ("wavefront", W.WR_WAVEFRONT)):
            film = W.Film(ctx, cm["width"], cm["height"])
            st = W.render(sc, cm, film, spp, depth, 5, mode)
            r[mode_name] = {"ms": round(st["kernel_ms"], 2), "rebuild_ms_per_pass": round(st["build_ms"] / spp, 3),
                            "samples_per_s": round(cm["width"] * cm["height"] * spp / (st["kernel_ms"] * 1e-3), 1)}
        t0 = time.perf_counter()
        for _ in range(3):
            W.Scene(ctx, ob).close()
        r["whole_scene_build_ms"] = round((time.perf_counter() - t0) / 3 * 1e3, 3)
        log(f"[bench] motion {name}: rebuild {r['wavefront']['rebuild_ms_per_pass']} ms per pass (whole scene "
            f"{r['whole_scene_build_ms']} ms); megakernel {r['megakernel']['ms']} ms, wavefront {r['wavefront']['ms']} ms")
        out[name] = r
    return out


def smoke_check(ctx):
    """A 24x24 Cornell render: megakernel and wavefront films must be identical."""
    objs, cam = S.cornell(k=2, width=24, height=24)
    a = render_scene(ctx, objs, cam, 4, 8, 7, W.WR_MEGAKERNEL)
    b = render_scene(ctx, objs, cam, 4, 8, 7, W.WR_WAVEFRONT)
    assert np.array_equal(a["sum"], b["sum"]), "megakernel and wavefront films differ"
    print(f"smoke: render 24x24x4spp megakernel == wavefront, mean radiance {a['mean'].mean():.4f}")
