"""Served render jobs: resident scenes, Poisson arrivals, p50/p99 latency.

Workload (served render jobs): 8 resident procedural scenes of about 10k to 2M
triangles (four fBm heightfields under an emitting sky quad, four fields of
spheres and boxes); jobs arrive as a Poisson process of a stated rate, each
picking a uniform random scene, a resolution from {128^2, 256^2, 512^2,
1024^2}, spp log-uniform in [4, 256] and max depth uniform in {8..16}.
Latency is measured from a job's arrival time to the moment the server
reports its film ready.
"""
from __future__ import annotations

import math
import time

import numpy as np

from . import _lib as W
from . import scenes as S

RESOLUTIONS = (128, 256, 512, 1024)


def heightfield_with_sky(n, seed):
    hf = S.heightfield(n, seed=seed, params=(0.6, 0.55, 0.5))
    sky = S.quad((0.5, 0.5, 1.5), (4.0, 0, 0), (0, 4.0, 0), (0, 0, -1), S.EMISSIVE, (3.0, 3.0, 3.0))
    cam = S.heightfield_camera(256, 256)
    return [hf, sky], cam


def serving_scenes(small=False):
    """The eight resident scenes as (objects, camera at 256^2) pairs."""
    if small:
        hfs, fields = (9, 17, 25, 33), (2, 4, 6, 8)
    else:
        hfs, fields = (72, 256, 724, 1024), (16, 64, 256, 512)
    out = [heightfield_with_sky(n, 11 + i) for i, n in enumerate(hfs)]
    out += [S.random_field(n_objects=k, seed=21 + i, width=256, height=256) for i, k in enumerate(fields)]
    return out


def job_camera(cam, res):
    c = dict(cam)
    c["sight_dist"] = cam["sight_dist"] * res / cam["width"]
    c["width"] = c["height"] = res
    return c


def make_jobs(n_jobs, n_scenes, rate, seed=5, max_res=1024, max_spp=256):
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, n_jobs))
    res_choices = [r for r in RESOLUTIONS if r <= max_res]
    jobs = []
    for i in range(n_jobs):
        spp = int(round(math.exp(rng.uniform(math.log(4), math.log(max_spp)))))
        jobs.append({"arrival": float(arrivals[i]), "scene": int(rng.integers(n_scenes)),
                     "res": int(res_choices[rng.integers(len(res_choices))]), "spp": spp,
                     "depth": int(rng.integers(8, 17)), "seed": 1000 + i})
    return jobs


def run_poisson(server, scene_ids, cams, jobs, release=True):
    """Feeds `jobs` to the server at their arrival times; returns per-job records."""
    start = server.clock()
    pending = list(jobs)
    live = {}
    done = []
    while pending or live:
        now = server.clock() - start
        while pending and pending[0]["arrival"] <= now:
            j = pending.pop(0)
            cam = job_camera(cams[j["scene"]], j["res"])
            jid = server.submit(scene_ids[j["scene"]], cam, j["spp"], j["depth"], j["seed"])
            live[jid] = j
        if not live:
            time.sleep(max(0.0, min(0.002, pending[0]["arrival"] - now)))
            continue
        server.step()
        for jid in list(live):
            st = server.status(jid)
            if st["state"] in (W.JOB_DONE, W.JOB_FAILED, W.JOB_CANCELLED):
                j = live.pop(jid)
                rec = dict(j)
                rec.update(state=st["state"], latency_s=st["done_s"] - (start + j["arrival"]),
                           queue_s=(st["admit_s"] - st["submit_s"]) if st["admit_s"] else None,
                           samples=st["samples"], done_s=st["done_s"] - start)
                done.append(rec)
                if release:
                    server.release(jid)
    return done


def summarize(records):
    ok = [r for r in records if r["state"] == W.JOB_DONE]
    lat = np.array([r["latency_s"] for r in ok]) * 1e3
    samples = sum(r["samples"] for r in ok)
    span = max(r["done_s"] for r in ok) - min(r["arrival"] for r in ok)
    return {"jobs": len(records), "completed": len(ok), "p50_ms": round(float(np.percentile(lat, 50)), 2),
            "p99_ms": round(float(np.percentile(lat, 99)), 2), "mean_ms": round(float(lat.mean()), 2),
            "samples_per_s": round(samples / span, 1), "span_s": round(span, 3)}


def bench_serving_sweep(ctx, log, srv, ids, cams, rates=(8.0, 16.0, 32.0, 64.0), n_jobs=40):
    """Capacity: the same 40-job stream offered at rising Poisson rates, then all at
    once (a burst, every job queued at t = 0).  Latency grows without bound once the
    offered samples/s exceed what the engine sustains; the burst's aggregate
    samples/s is that sustainable rate."""
    out = []
    base = make_jobs(n_jobs, len(ids), 1.0, seed=9)
    for rate in list(rates) + [None]:
        jobs = [dict(j, arrival=(j["arrival"] / rate if rate else 0.0)) for j in base]
        recs = run_poisson(srv, ids, cams, jobs)
        res = summarize(recs)
        offered = sum(r["samples"] for r in recs) / max(jobs[-1]["arrival"], 1e-9) if rate else None
        res.update(rate_jobs_per_s=rate if rate else "burst",
                   offered_samples_per_s=round(offered, 1) if offered else None)
        out.append(res)
        log(f"[bench] serving sweep {res['rate_jobs_per_s']}: p50 {res['p50_ms']} ms, p99 {res['p99_ms']} ms, "
            f"{res['samples_per_s'] / 1e6:.1f} Msamples/s")
    return out


def bench_serving(ctx, log, rate=8.0, n_jobs=80, budget_gib=16):
    t0 = time.time()
    scs = serving_scenes()
    built = [W.Scene(ctx, objs) for objs, _ in scs]
    tris = [b.info()["n_triangles"] for b in built]
    srv = W.Server(ctx, budget_bytes=budget_gib << 30)
    ids = [srv.add_scene(b) for b in built]
    log(f"[bench] serving: 8 scenes ({min(tris)}..{max(tris)} triangles) resident in {time.time() - t0:.1f}s")
    jobs = make_jobs(n_jobs, len(built), rate)
    recs = run_poisson(srv, ids, [c for _, c in scs], jobs)
    res = summarize(recs)
    res.update(rate_jobs_per_s=rate, hbm_budget_gib=budget_gib, scene_triangles=tris,
               workload="served render jobs: 8 resident scenes, Poisson arrivals, res {128..1024}^2, "
                        "spp log-uniform 4-256, depth 8-16")
    log(f"[bench] serving: p50 {res['p50_ms']} ms, p99 {res['p99_ms']} ms, {res['samples_per_s'] / 1e6:.1f} "
        f"Msamples/s over {res['span_s']} s")
    res["sweep"] = bench_serving_sweep(ctx, log, srv, ids, [c for _, c in scs])
    res["sustainable_samples_per_s"] = res["sweep"][-1]["samples_per_s"]
    srv.close()
    return res
