"""Serving: continuous batching of render jobs over one wavefront pool.

Bar: every job's film is bit-identical whether the job runs alone, in a mix of
concurrent jobs on several resident scenes, or through wr_render (megakernel);
admission honours the memory budget and the job-slot limit.
"""
import numpy as np
import pytest

from walkray_amd import _lib as W
from walkray_amd import render as R
from walkray_amd import scenes as S
from walkray_amd import serve as V

pytestmark = pytest.mark.gpu


def _scenes(ctx):
    scs = V.serving_scenes(small=True)
    return scs, [W.Scene(ctx, objs) for objs, _ in scs]


def _drain(srv, jids, limit=100000):
    for _ in range(limit):
        srv.step()
        if all(srv.status(j)["state"] in (W.JOB_DONE, W.JOB_FAILED, W.JOB_CANCELLED) for j in jids):
            return
    raise AssertionError("jobs did not finish")


JOBS = [(0, 48, 6, 9, 101), (4, 40, 10, 12, 102), (2, 32, 3, 8, 103), (5, 56, 5, 16, 104), (7, 24, 7, 10, 105)]


def test_alone_equals_mix_equals_render(ctx):
    scs, built = _scenes(ctx)
    srv = W.Server(ctx)
    ids = [srv.add_scene(b) for b in built]
    cams = [V.job_camera(scs[s][1], res) for s, res, *_ in JOBS]
    mix = [srv.submit(ids[s], cams[k], spp, depth, seed) for k, (s, res, spp, depth, seed) in enumerate(JOBS)]
    _drain(srv, mix)
    mixed = [srv.read(j, cams[k]["width"], cams[k]["height"]) for k, j in enumerate(mix)]
    for k, (s, res, spp, depth, seed) in enumerate(JOBS):
        alone_srv = W.Server(ctx)
        aid = [alone_srv.add_scene(b) for b in built]
        j = alone_srv.submit(aid[s], cams[k], spp, depth, seed)
        _drain(alone_srv, [j])
        alone = alone_srv.read(j, res, res)
        alone_srv.close()
        assert np.array_equal(alone["sum"].view(np.uint32), mixed[k]["sum"].view(np.uint32)), f"job {k}: mix != alone"
        film = W.Film(ctx, res, res)
        W.render(built[s], cams[k], film, spp, depth, seed, W.WR_MEGAKERNEL)
        ref = film.read()
        for key in ("sum", "comp", "sumsq"):
            assert np.array_equal(ref[key].view(np.uint32), mixed[k][key].view(np.uint32)), (k, key)
    st = [srv.status(j) for j in mix]
    assert all(s["state"] == W.JOB_DONE and s["done_s"] >= s["admit_s"] >= s["submit_s"] for s in st)


def test_nee_and_plain_jobs_mixed_equal_render(ctx):
    """Jobs with and without next-event estimation in one pool: the engine shades
    with the NEE instance while any job in the pool asks for it and with the plain
    one otherwise (a plain job admitted after the NEE job finished runs on the
    plain instance, one admitted alongside it on the NEE one); every film equals
    wr_render's with the job's own setting (pathTracer.cpp:39-66)."""
    scs, built = _scenes(ctx)
    srv = W.Server(ctx, max_jobs=2)
    ids = [srv.add_scene(b) for b in built]
    jobs = [(0, 32, 4, 8, 201, True), (2, 48, 12, 10, 202, False), (4, 24, 3, 9, 203, True),
            (5, 40, 6, 12, 204, False), (7, 32, 5, 8, 205, False)]
    cams = [V.job_camera(scs[s][1], res) for s, res, *_ in jobs]
    jids = [srv.submit(ids[s], cams[k], spp, depth, seed, nee=nee)
            for k, (s, res, spp, depth, seed, nee) in enumerate(jobs)]
    _drain(srv, jids)
    for k, (s, res, spp, depth, seed, nee) in enumerate(jobs):
        got = srv.read(jids[k], res, res)
        film = W.Film(ctx, res, res)
        W.render(built[s], cams[k], film, spp, depth, seed, W.WR_MEGAKERNEL, nee=nee)
        ref = film.read()
        for key in ("sum", "comp", "sumsq"):
            assert np.array_equal(ref[key].view(np.uint32), got[key].view(np.uint32)), (k, nee, key)
    srv.close()


def test_small_pool_mix_equals_render(ctx):
    """A path pool smaller than one pass of most jobs: every iteration's slot
    allocation across the running jobs is limited by the free room."""
    scs, built = _scenes(ctx)
    srv = W.Server(ctx, path_slots=2048)
    ids = [srv.add_scene(b) for b in built]
    cams = [V.job_camera(scs[s][1], res) for s, res, *_ in JOBS]
    mix = [srv.submit(ids[s], cams[k], spp, depth, seed) for k, (s, res, spp, depth, seed) in enumerate(JOBS)]
    _drain(srv, mix)
    for k, (s, res, spp, depth, seed) in enumerate(JOBS):
        got = srv.read(mix[k], res, res)
        film = W.Film(ctx, res, res)
        W.render(built[s], cams[k], film, spp, depth, seed, W.WR_MEGAKERNEL)
        ref = film.read()
        for key in ("sum", "comp", "sumsq"):
            assert np.array_equal(ref[key].view(np.uint32), got[key].view(np.uint32)), (k, key)
    srv.close()


def test_budget_and_slots(ctx):
    scs, built = _scenes(ctx)
    cam = V.job_camera(scs[0][1], 64)
    one_job = 3 * 4 * 3 * 64 * 64 + 3 * 4 * 64 * 64 * 8 + 4 * 8
    srv = W.Server(ctx, budget_bytes=int(one_job * 1.5), max_jobs=4)
    sid = srv.add_scene(built[0])
    big = srv.submit(sid, V.job_camera(scs[0][1], 512), 8, 8, 1)  # alone exceeds the budget
    assert srv.status(big)["state"] == W.JOB_FAILED
    a = srv.submit(sid, cam, 8, 8, 2)
    b = srv.submit(sid, cam, 8, 8, 3)
    srv.step()
    assert srv.status(a)["state"] in (W.JOB_RUNNING, W.JOB_DONE)
    assert srv.status(b)["state"] == W.JOB_QUEUED  # waits for memory
    _drain(srv, [a, b])
    sa, sb = srv.status(a), srv.status(b)
    assert sb["admit_s"] >= sa["done_s"]
    with pytest.raises(W.WalkRayError):
        srv.read(big, 512, 512)


def test_poisson_run_small(ctx):
    scs, built = _scenes(ctx)
    srv = W.Server(ctx)
    ids = [srv.add_scene(b) for b in built]
    jobs = V.make_jobs(24, len(built), rate=200.0, max_res=128, max_spp=16)
    recs = V.run_poisson(srv, ids, [c for _, c in scs], jobs)
    summ = V.summarize(recs)
    assert summ["completed"] == 24 and summ["p99_ms"] >= summ["p50_ms"] > 0


def _render_ref(ctx, built, cam, spp, depth, seed):
    film = W.Film(ctx, cam["width"], cam["height"])
    W.render(built, cam, film, spp, depth, seed, W.WR_MEGAKERNEL)
    return film.read()


def _same_film(a, b, what):
    for key in ("sum", "comp", "sumsq"):
        assert np.array_equal(a[key].view(np.uint32), b[key].view(np.uint32)), (what, key)


def test_cancel_running_and_queued(ctx):
    """A job cancelled while running (its paths in the shared pool) and one cancelled
    while queued: both end CANCELLED with no film, every other job of the mix equals
    its wr_render film bit for bit, and a job admitted afterwards into the freed slot
    renders correctly too."""
    scs, built = _scenes(ctx)
    srv = W.Server(ctx, max_jobs=4)
    ids = [srv.add_scene(b) for b in built]
    cams = [V.job_camera(scs[s][1], res) for s, res, *_ in JOBS]
    mix = [srv.submit(ids[s], cams[k], spp, depth, seed) for k, (s, res, spp, depth, seed) in enumerate(JOBS[:3])]
    big_cam = V.job_camera(scs[6][1], 96)
    big = srv.submit(ids[6], big_cam, 400, 12, 77)  # long enough to still run after one step
    queued = srv.submit(ids[1], cams[0], 4, 8, 78)  # a fifth job waits for one of the 4 slots
    srv.step()
    assert srv.status(big)["state"] == W.JOB_RUNNING
    assert srv.status(queued)["state"] == W.JOB_QUEUED
    srv.cancel(big)
    srv.cancel(queued)
    srv.cancel(big)  # cancelling twice does nothing
    for j in (big, queued):
        assert srv.status(j)["state"] == W.JOB_CANCELLED
        with pytest.raises(W.WalkRayError):
            srv.read(j, 96, 96)
    late_cam = V.job_camera(scs[3][1], 40)
    late = srv.submit(ids[3], late_cam, 9, 10, 79)
    _drain(srv, mix + [late])
    for k, (s, res, spp, depth, seed) in enumerate(JOBS[:3]):
        _same_film(srv.read(mix[k], res, res), _render_ref(ctx, built[s], cams[k], spp, depth, seed), k)
    _same_film(srv.read(late, 40, 40), _render_ref(ctx, built[3], late_cam, 9, 10, 79), "late")
    for j in mix:
        with pytest.raises(W.WalkRayError):
            srv.cancel(j)  # finished jobs cannot be cancelled
    srv.release(big)
    srv.release(queued)
    srv.close()


def test_failed_admission_is_isolated(ctx):
    """An admission that fails (injected as a failed device allocation) fails that job
    only: it leaves the queue, the step succeeds, and the jobs behind it run and equal
    their wr_render films bit for bit."""
    scs, built = _scenes(ctx)
    srv = W.Server(ctx)
    ids = [srv.add_scene(b) for b in built]
    cams = [V.job_camera(scs[s][1], res) for s, res, *_ in JOBS]
    srv.inject_admit_failures(1)
    jobs = [srv.submit(ids[s], cams[k], spp, depth, seed) for k, (s, res, spp, depth, seed) in enumerate(JOBS[:3])]
    _drain(srv, jobs)
    assert srv.status(jobs[0])["state"] == W.JOB_FAILED
    for k in (1, 2):
        s, res, spp, depth, seed = JOBS[k]
        assert srv.status(jobs[k])["state"] == W.JOB_DONE
        _same_film(srv.read(jobs[k], res, res), _render_ref(ctx, built[s], cams[k], spp, depth, seed), k)
    srv.release(jobs[0])
    srv.close()


def test_full_size_serving_mix(ctx):
    """The served-jobs workload at full size: the eight resident scenes (10k to 2M
    triangles) and a mix with a 1024^2 job on the 2M-triangle heightfield and jobs on
    the largest fields.  Every film equals its alone-in-a-server film and wr_render's
    (megakernel) bit for bit; two served jobs at reduced resolution (a heightfield and
    a 256-object field) agree with the oracle's PathTracer per pixel."""
    from oracle import oracle as O
    scs = V.serving_scenes()
    built = [W.Scene(ctx, objs) for objs, _ in scs]
    tris = [b.info()["n_triangles"] for b in built]
    assert min(tris) >= 10000 and max(tris) >= 2000000
    jobs = [(3, 1024, 4, 8, 201), (7, 256, 16, 16, 202), (2, 512, 8, 12, 203), (5, 128, 64, 10, 204),
            (1, 64, 16, 8, 205), (6, 64, 16, 12, 206)]
    srv = W.Server(ctx)
    ids = [srv.add_scene(b) for b in built]
    cams = [V.job_camera(scs[s][1], res) for s, res, *_ in jobs]
    mix = [srv.submit(ids[s], cams[k], spp, depth, seed) for k, (s, res, spp, depth, seed) in enumerate(jobs)]
    _drain(srv, mix)
    for k, (s, res, spp, depth, seed) in enumerate(jobs):
        got = srv.read(mix[k], res, res)
        _same_film(got, _render_ref(ctx, built[s], cams[k], spp, depth, seed), k)
        if k in (0, 3):
            alone_srv = W.Server(ctx)
            aid = [alone_srv.add_scene(b) for b in built]
            j = alone_srv.submit(aid[s], cams[k], spp, depth, seed)
            _drain(alone_srv, [j])
            _same_film(alone_srv.read(j, res, res), got, f"alone {k}")
            alone_srv.close()
        if k in (4, 5):
            st = {"sum": got["sum"], "sumsq": got["sumsq"], "passes": spp}
            o_sum, o_sq = O.Scene(scs[s][0]).render(cams[k], spp, depth, seed)
            mg, seg = R.film_stats(st)
            mo = o_sum / spp
            se = np.sqrt(seg ** 2 + np.maximum(o_sq / spp - mo * mo, 0) / spp)
            nz = se > 0
            z = np.abs(mg - mo)[nz] / se[nz]
            assert np.all(np.abs(mg - mo)[~nz] <= 1e-5 * np.maximum(1.0, np.abs(mo[~nz])))
            assert (z > 3).mean() <= 0.01 and z.max() <= 6, (k, (z > 3).mean(), z.max())
    srv.close()
