"""The large-scene workload at its full size: a 4096^2 fBm heightfield plus 2,000
Phong icospheres k=4, a sun disk and a sky dome (43,779,362 triangles),
3840x2160, max depth 32.

- Closest hit (kdtree.cpp:308-353) on sampled camera and incoherent rays:
  prim ids and t bit-identical to the oracle's unpruned traversal of the
  reference tree built over the same triangle soup (Scene::buildKDTree order).
- The PathTracer estimator (pathTracer.cpp:20-70): one full frame at 1 spp,
  megakernel and wavefront films bit-identical.
"""
import numpy as np
import pytest

from oracle import oracle as O
from walkray_amd import _lib as W
from walkray_amd import render as R
from walkray_amd import scenes as S

pytestmark = pytest.mark.gpu


def test_large_scene_closest_hit_and_engines(ctx):
    objs, cam = S.large_scene()
    off = np.cumsum([0] + [len(o["verts"]) for o in objs])
    V = np.concatenate([o["verts"] for o in objs]).astype(np.float32)
    T = np.concatenate([o["tris"].astype(np.uint32) + np.uint32(off[i]) for i, o in enumerate(objs)])
    assert len(T) == 43779362
    sc = W.Scene(ctx, objs)
    del objs
    print(f"large scene built: {sc.info()['build_seconds']:.1f}s host", flush=True)
    rng = np.random.default_rng(3)
    W_, H_ = cam["width"], cam["height"]
    pix = rng.choice(W_ * H_, 16384, replace=False)
    coh = S.pinhole_rays(cam)[pix]
    lo, hi = V.min(axis=0), V.max(axis=0)
    inc = S.incoherent_rays(16384, lo, hi, seed=9)
    rays = np.concatenate([coh, inc])
    hits = sc.closest_hit(rays)
    tree = O.Tree(V, T)
    del V, T
    prim, t = tree.intersect(rays)
    del tree
    assert not np.any(hits["prim"] == W.OVERFLOW)
    bad = np.flatnonzero(hits["prim"] != prim)
    assert bad.size == 0, f"{bad.size} prim mismatches, first {bad[:5]}"
    assert np.array_equal(hits["t"].view(np.uint32), t.view(np.uint32))
    assert (prim != O.NONE).mean() > 0.5
    print("large scene: 32768 rays bit-identical to the oracle", flush=True)
    a = R.render_scene(ctx, None, cam, 1, 32, 17, W.WR_MEGAKERNEL, scene=sc)
    b = R.render_scene(ctx, None, cam, 1, 32, 17, W.WR_WAVEFRONT, scene=sc)
    for k in ("sum", "comp", "sumsq"):
        assert np.array_equal(a[k].view(np.uint32), b[k].view(np.uint32)), k
    assert a["mean"].mean() > 0
