"""A reference-format scene (Config::load, config.cpp:432-575) loaded, built and
traced on the GPU: closest hit bit-identical with the oracle on the loaded
arrays, megakernel and wavefront films identical, and the image consistent with
the procedural scene the files were written from."""
import numpy as np
import pytest

from oracle import oracle as O
from walkray_amd import _lib as W
from walkray_amd import render as R
from walkray_amd import scenes as S

pytestmark = pytest.mark.gpu


def test_loaded_cornell_traces_and_renders(ctx, tmp_path):
    objs, cam = S.cornell(k=2, width=40, height=40)
    cfg = W.SceneConfig(S.write_reference_scene(tmp_path, objs, cam, spp=8, name="Cornell"))
    for k in ("pos", "focus", "up", "sight_dist"):
        assert np.array_equal(np.float32(cfg.camera[k]), np.float32(cam[k])), k
    assert (cfg.camera["width"], cfg.camera["height"]) == (cam["width"], cam["height"])
    assert (cfg.renderer, cfg.spp, cfg.max_len) == ("PathTracer", 8, 20)
    assert len(cfg.objects) == len(objs)
    for a, b in zip(cfg.objects, objs):
        assert a["mat"] == b["mat"]
        assert np.array_equal(a["tris"], b["tris"])
        np.testing.assert_allclose(a["verts"], b["verts"], rtol=0, atol=2e-6)
    sc = W.Scene(ctx)
    cfg.add_to(sc)
    sc.build()
    osc = O.Scene(cfg.objects)
    rays = np.concatenate([S.pinhole_rays(cam), S.incoherent_rays(8192, (0.01, 0.01, 0.01), (0.99, 0.99, 0.99), 4)])
    hits = sc.closest_hit(rays)
    ref = osc.intersect(rays)
    assert np.array_equal(hits["prim"], ref["prim"])
    assert np.array_equal(hits["t"].view(np.uint32), ref["t"].view(np.uint32))
    a = R.render_scene(ctx, None, cam, 8, 8, 5, W.WR_MEGAKERNEL, scene=sc)
    b = R.render_scene(ctx, None, cam, 8, 8, 5, W.WR_WAVEFRONT, scene=sc)
    assert np.array_equal(a["sum"].view(np.uint32), b["sum"].view(np.uint32))
    # the same estimator on geometry equal up to float rounding: same image statistically
    c = R.render_scene(ctx, objs, cam, 8, 8, 5, W.WR_WAVEFRONT)
    ma, mc = a["mean"].mean(), c["mean"].mean()
    assert abs(ma - mc) < 0.05 * mc
    # the film written as the reference's PFM
    W.write_pfm(tmp_path / "out.pfm", a["mean"], 40, 40)
    data = (tmp_path / "out.pfm").read_bytes()
    assert data.startswith(b"PF\n40 40\n-1\n") and len(data) == len(b"PF\n40 40\n-1\n") + 40 * 40 * 12
