"""BidirectionalPathTracer on the GPU (wr_render_bdpt) against the oracle's
restatement (orc_render_bdpt) on the same seeds: the same eye and light
sub-paths (counter RNG keyed by pixel, pass and vertex), the same connections
and MIS weights.  The per-pixel sums differ only where a libm last-ulp
difference turns a random decision, and in the summation order of the
light-tracing splats (float atomics on the GPU, pixel order in the oracle):
most pixel sums agree to 1e-4 relative, and the images agree within their
standard errors.  Bar and tolerances as for the path tracer's oracle parity."""
import numpy as np
import pytest

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


@pytest.mark.parametrize("glass,depth", [(False, 3), (True, 6)])
def test_bdpt_matches_oracle(ctx, glass, depth):
    objs, cam = S.cornell(k=2, width=24, height=24)
    if not glass:
        objs = [dict(o) for o in objs]
        objs[5] = dict(objs[5], mat=S.DIFFUSE, params=(0.8, 0.8, 0.8))
    spp, seed = 32, 77
    print(f"[bdpt] oracle parity glass={glass} depth={depth}", flush=True)
    sc = W.Scene(ctx, objs)
    film = W.Film(ctx, cam["width"], cam["height"])
    W.render_bdpt(sc, cam, film, spp, depth, seed)
    g = film.read()
    s, q = O.Scene(objs).render_bdpt(cam, spp, depth, seed)
    z, zero_ok = _zcheck(g, s, q, spp)
    assert zero_ok
    assert (z > 3).mean() <= 0.01, f"{(z > 3).mean():.4f} of pixel-channels beyond 3 sigma"
    assert z.max() <= 6, f"max z {z.max():.2f}"
    same = np.isclose(g["sum"], s.astype(np.float32), rtol=1e-4, atol=1e-6).mean()
    assert same > 0.5, f"only {same:.3f} of pixel sums agree closely"
    assert g["mean"].mean() > 0


def test_bdpt_small_queue_traces_in_place(ctx):
    """A 64-entry visibility queue: nearly every connection is traced in place
    by the lane that forms it; the film equals the queued one up to the float
    atomics' summation order."""
    objs, cam = S.cornell(k=2, width=24, height=24)
    sc = W.Scene(ctx, objs)
    films = []
    for queue in (0, 64):
        film = W.Film(ctx, cam["width"], cam["height"])
        W.render_bdpt(sc, cam, film, 8, 6, 5, queue=queue)
        films.append(film.read()["sum"].astype(np.float64))
    a, b = films
    assert np.allclose(a, b, rtol=1e-4, atol=1e-6), np.abs(a - b).max()
    assert a.sum() > 0


def test_bdpt_arguments(ctx):
    objs, cam = S.cornell(k=1, width=8, height=8)
    sc = W.Scene(ctx, objs)
    film = W.Film(ctx, 8, 8)
    with pytest.raises(W.WalkRayError):
        W.render_bdpt(sc, cam, film, 1, 33, 1)  # maxPathLen above 32
    wrong = W.Film(ctx, 4, 8)
    with pytest.raises(W.WalkRayError):
        W.render_bdpt(sc, cam, wrong, 1, 4, 1)
