"""The autotuners change speed, never results: closest-hit launch knobs at their
extremes and after tuning, and wavefront queue sizes / stage grids after
tuning, all give bit-identical outputs."""
import numpy as np
import pytest

from walkray_amd import _lib as W
from walkray_amd import render as R
from walkray_amd import scenes as S

pytestmark = pytest.mark.gpu


def test_closest_hit_tuning_keeps_results(ctx):
    objs = [S.heightfield(257, seed=9)]
    sc = W.Scene(ctx, objs)
    lo, hi = S.scene_bounds(objs)
    rays = np.concatenate([S.pinhole_rays(S.heightfield_camera(192, 192)),
                           S.incoherent_rays(40000, lo, hi, seed=2)]).astype(np.float32)
    base = sc.closest_hit(rays)
    for refill, node, blocks, levels, batch in ((1, 65, 1, 4, 64), (64, 1, 0, 16, 4096), (7, 13, 3, 8, 100),
                                                (16, 32, 0, 4, 1024)):
        sc.set_closest_hit_tuning(refill, node, blocks, levels, batch)
        h = sc.closest_hit(rays)
        assert np.array_equal(h.view(np.uint32), base.view(np.uint32)), (refill, node, blocks, levels, batch)
    d = ctx.to_device(rays)
    t = sc.autotune_closest_hit(d, len(rays))
    assert 1 <= t["refill_min"] <= 64 and 1 <= t["node_phase_min"] <= 65 and t["ms"] > 0
    assert t["lds_levels"] in (4, 8, 16) and 64 <= t["batch"] <= 4096
    assert np.array_equal(sc.closest_hit(rays).view(np.uint32), base.view(np.uint32))


def test_render_tuning_keeps_films(ctx):
    objs, cam = S.cornell(k=2, width=48, height=40)
    sc = W.Scene(ctx, objs)
    a = R.render_scene(ctx, objs, cam, 6, 8, 3, W.WR_WAVEFRONT, scene=sc)
    t = W.autotune_render(sc, cam, 8, spp=1)
    assert t["wave_slots"] in (1 << 20, 1 << 21, 1 << 22, 1 << 23) and t["wave_grid"] in (1024, 2048, 4096)
    assert t["lds_levels"] in (4, 8, 16)
    b = R.render_scene(ctx, objs, cam, 6, 8, 3, W.WR_WAVEFRONT, scene=sc)
    for k in ("sum", "comp", "sumsq"):
        assert np.array_equal(a[k].view(np.uint32), b[k].view(np.uint32)), k
    # every stack depth of the trace stage, on a small pool and grid
    for levels in (4, 16, 8):
        W.set_render_tuning(ctx, 1 << 12, 1024, levels)
        c = R.render_scene(ctx, objs, cam, 6, 8, 3, W.WR_WAVEFRONT, scene=sc)
        for k in ("sum", "comp", "sumsq"):
            assert np.array_equal(a[k].view(np.uint32), c[k].view(np.uint32)), (levels, k)
    W.set_render_tuning(ctx)
