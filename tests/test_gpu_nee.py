"""GPU next-event estimation (PathTracer's useNextEventEstimation branch,
pathTracer.cpp:39-66) through the C ABI.

Bars, as for the plain estimator:
  * megakernel and wavefront films are bit-identical;
  * GPU vs oracle: per pixel and channel within 3 sigma of the combined standard
    error (at most 1% beyond 3 sigma, none beyond 6), and most pixel sums agree
    closely because the RNG stream (light sample included) is shared;
  * closed forms of the branch (nee_cases.py) within 4 standard errors, exact zeros
    where the branch and the plain estimator both give 0;
  * a served job with the branch equals wr_render's film bit for bit.
"""
import numpy as np
import pytest

from nee_cases import expected_escape, expected_max_depth, nee_scene
from oracle import oracle as O
from walkray_amd import _lib as W
from walkray_amd import render as R
from walkray_amd import scenes as S

pytestmark = pytest.mark.gpu


def _zcheck(gpu, orc_sum, orc_sq, n):
    mg, seg = R.film_stats(gpu)
    mo = orc_sum / n
    vo = np.maximum(orc_sq / n - mo * mo, 0)
    se = np.sqrt(seg ** 2 + vo / n)
    diff = np.abs(mg - mo)
    nz = se > 0
    z = np.zeros_like(diff)
    z[nz] = diff[nz] / se[nz]
    zero_ok = np.all(diff[~nz] <= 1e-5 * np.maximum(1.0, np.abs(mo[~nz])))
    return z, zero_ok


def _same_bits(a, b):
    return all(np.array_equal(a[k].view(np.uint32), b[k].view(np.uint32)) for k in ("sum", "comp", "sumsq"))


@pytest.mark.parametrize("case", ["cornell", "field"])
def test_nee_megakernel_equals_wavefront(ctx, case):
    if case == "cornell":
        objs, cam = S.cornell(k=3, width=40, height=32)
        spp, depth = 12, 8
    else:
        objs, cam = S.random_field(n_objects=48, width=48, height=27, sphere_k=2)
        spp, depth = 8, 5
    print(f"[nee] mega vs wave {case}", flush=True)
    sc = W.Scene(ctx, objs)
    a = R.render_scene(ctx, objs, cam, spp, depth, 99, W.WR_MEGAKERNEL, scene=sc, nee=True)
    b = R.render_scene(ctx, objs, cam, spp, depth, 99, W.WR_WAVEFRONT, scene=sc, nee=True)
    assert _same_bits(a, b)
    c = R.render_scene(ctx, objs, cam, spp, depth, 99, W.WR_WAVEFRONT, scene=sc, wave_slots=1000, nee=True)
    assert _same_bits(a, c)
    plain = R.render_scene(ctx, objs, cam, spp, depth, 99, W.WR_MEGAKERNEL, scene=sc)
    assert not np.array_equal(plain["sum"], a["sum"])  # the branch adds light


@pytest.mark.parametrize("mode", [W.WR_MEGAKERNEL, W.WR_WAVEFRONT])
@pytest.mark.parametrize("case", ["cornell", "field"])
def test_nee_matches_oracle(ctx, mode, case):
    if case == "cornell":
        objs, cam = S.cornell(k=3, width=32, height=32)
        spp, depth = 64, 8
    else:
        objs, cam = S.random_field(n_objects=48, width=48, height=27, sphere_k=2)
        spp, depth = 32, 5
    seed = 2025
    print(f"[nee] oracle parity {case} mode {mode}", flush=True)
    g = R.render_scene(ctx, objs, cam, spp, depth, seed, mode, nee=True)
    s, q = O.Scene(objs).render(cam, spp, depth, seed, nee=True)
    z, zero_ok = _zcheck(g, s, q, spp)
    assert zero_ok
    assert (z > 3).mean() <= 0.01, f"{(z > 3).mean():.4f} of pixel-channels beyond 3 sigma"
    assert z.max() <= 6, f"max z {z.max():.2f}"
    same = np.isclose(g["sum"], s.astype(np.float32), rtol=1e-4, atol=1e-6).mean()
    assert same > 0.5, f"only {same:.3f} of pixel sums agree closely"


@pytest.mark.parametrize("mode", [W.WR_MEGAKERNEL, W.WR_WAVEFRONT])
def test_nee_closed_forms(ctx, mode):
    spp, width = 16, 64
    n = spp * width * width
    for albedo in ((0.5, 0.5, 0.5), (0.3, 0.7, 0.5)):
        objs, cam, h = nee_scene(albedo, width=width)
        print(f"[nee] closed form escape {albedo} mode {mode}", flush=True)
        g = R.render_scene(ctx, objs, cam, spp, 1, 7, mode, nee=True)
        m = g["sum"].reshape(-1, 3).sum(0) / n
        se = np.sqrt(np.maximum(g["sumsq"].reshape(-1, 3).sum(0) / n - m * m, 0) / n)
        want = expected_escape(albedo, h)
        assert np.all(np.abs(m - want) <= 4 * se + 1e-4), (albedo, m, want, se)
        g0 = R.render_scene(ctx, objs, cam, 4, 0, 7, mode, nee=True)  # endIndex 0: nothing connects
        assert np.all(g0["sum"] == 0)
    a, c = 0.6, 0.5
    objs, cam, h = nee_scene((a, a, a), ceiling=c, width=width)
    print(f"[nee] closed form max depth mode {mode}", flush=True)
    g = R.render_scene(ctx, objs, cam, spp, 1, 7, mode, nee=True)
    m = g["sum"].reshape(-1, 3).sum(0) / n
    se = np.sqrt(np.maximum(g["sumsq"].reshape(-1, 3).sum(0) / n - m * m, 0) / n)
    want = expected_max_depth(a, c, h)
    assert np.all(np.abs(m - want) <= 4 * se + 1e-4), (m, want, se)
    for kw in ({"light_up": True}, {"blocker": True}):
        objs, cam, _ = nee_scene((0.5, 0.5, 0.5), width=16, **kw)
        print(f"[nee] zero case {kw} mode {mode}", flush=True)
        g = R.render_scene(ctx, objs, cam, 32, 3, 7, mode, nee=True)
        assert np.all(g["sum"] == 0), kw


def test_nee_served_job_equals_render(ctx):
    objs, cam = S.cornell(k=2, width=48, height=40)
    sc = W.Scene(ctx, objs)
    print("[nee] served job vs render", flush=True)
    ref = R.render_scene(ctx, objs, cam, 10, 6, 31, W.WR_WAVEFRONT, scene=sc, nee=True)
    srv = W.Server(ctx, path_slots=1 << 16)
    sid = srv.add_scene(sc)
    j_nee = srv.submit(sid, cam, 10, 6, 31, nee=True)
    j_plain = srv.submit(sid, cam, 10, 6, 31)
    for _ in range(100000):
        srv.step()
        if all(srv.status(j)["state"] in (W.JOB_DONE, W.JOB_FAILED) for j in (j_nee, j_plain)):
            break
    assert all(srv.status(j)["state"] == W.JOB_DONE for j in (j_nee, j_plain))
    w, h = cam["width"], cam["height"]
    assert _same_bits(ref, srv.read(j_nee, w, h))
    plain = R.render_scene(ctx, objs, cam, 10, 6, 31, W.WR_WAVEFRONT, scene=sc)
    assert _same_bits(plain, srv.read(j_plain, w, h))
    srv.close()
