"""The C++ host mirror (walkray_amd/host/walkray.hpp) used the way reference code
uses KDTree / Scene / PathTracer, checked against the oracle and against the
Python binding of the same ABI."""
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O
from walkray_amd import _lib as W
from walkray_amd import scenes as S

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(W.LIB_PATH), "mirror_test")


def test_cpp_mirror(ctx, tmp_path):
    if not os.path.exists(BIN):
        pytest.fail("walkray_amd/mirror_test is not built (make -C walkray_amd)")
    hf = S.heightfield(33, seed=3, params=(0.7, 0.6, 0.5))
    lo, hi = S.scene_bounds([hf])
    rays = np.concatenate([S.incoherent_rays(3000, lo, hi, seed=8),
                           S.pinhole_rays(S.heightfield_camera(24, 24))]).astype(np.float32)
    W_, H_, spp, depth = 20, 16, 6, 5
    cam = S.camera((0.5, -0.6, 1.2), (0.5, 0.5, 0.0), (0.0, 0.0, 1.0), W_, H_, sight_dist=22.0)
    hdr = np.array([len(hf["verts"]), len(hf["tris"]), len(rays), W_, H_, spp, depth], np.uint32)
    camv = np.array(list(cam["pos"]) + list(cam["focus"]) + list(cam["up"]) + [cam["sight_dist"]], np.float32)
    inp, out = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(inp, "wb") as f:
        for a in (hdr, hf["verts"], hf["tris"], rays, camv):
            f.write(np.ascontiguousarray(a).tobytes())
    r = subprocess.run([BIN, str(inp), str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    buf = open(out, "rb").read()
    n = len(rays)
    u = np.frombuffer(buf, dtype=np.uint32)
    f = np.frombuffer(buf, dtype=np.float32)
    prim, t = u[0:n], f[n:2 * n]
    prim1, t1 = u[2 * n:3 * n], f[3 * n:4 * n]
    obj, tri, ts = u[4 * n:5 * n], u[5 * n:6 * n], f[6 * n:7 * n]
    inside = u[7 * n:8 * n]
    rgb = f[8 * n:8 * n + 3 * W_ * H_]
    # KDTree: batch and one-ray-at-a-time both equal the oracle bit for bit
    op, ot = O.Tree(hf["verts"], hf["tris"]).intersect(rays)
    assert np.array_equal(prim, op) and np.array_equal(t.view(np.uint32), ot.view(np.uint32))
    assert np.array_equal(prim1, op) and np.array_equal(t1.view(np.uint32), ot.view(np.uint32))
    # Scene: ObjSourceInfo and findInsideObject over the soup + the emitting quad
    light = {"verts": np.array([[-1, -1, 2], [2, -1, 2], [2, 2, 2], [-1, 2, 2]], np.float32),
             "tris": np.array([[0, 2, 1], [0, 3, 2]], np.uint32), "normals": None, "mat": S.EMISSIVE,
             "params": (4.0, 4.0, 4.0)}
    objs = [hf, light]
    osc = O.Scene(objs)
    ref = osc.intersect(rays)
    hitm = ref["prim"] != O.NONE
    assert np.array_equal(obj[hitm], ref["obj"][hitm]) and np.array_equal(tri[hitm], ref["tri"][hitm])
    assert np.array_equal(ts.view(np.uint32), ref["t"].view(np.uint32))
    assert np.array_equal(inside, osc.find_inside(rays))
    # PathTracer::renderPixels equals the Python binding's render of the same job
    sc = W.Scene(ctx, objs)
    film = W.Film(ctx, W_, H_)
    W.render(sc, cam, film, spp, depth, 42, W.WR_WAVEFRONT)
    assert np.array_equal(rgb, film.read()["mean"])
    assert rgb.mean() > 0


def test_cpp_mirror_renders_a_scene_file(ctx, tmp_path):
    """Config::load -> Scene::buildKDTree -> PathTracer::renderPixels -> savePFM through the C++ mirror,
    on the committed Cornell scene files, against the Python binding's render of the same loaded scene."""
    if not os.path.exists(BIN):
        pytest.fail("walkray_amd/mirror_test is not built (make -C walkray_amd)")
    cfg_path = os.path.join(os.path.dirname(__file__), "golden", "scenes", "cornell", "config.xml")
    out = tmp_path / "cornell.pfm"
    r = subprocess.run([BIN, "--config", cfg_path, str(out), "4"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    data = out.read_bytes()
    hdr = b"PF\n64 64\n-1\n"
    assert data[:len(hdr)] == hdr
    img = np.frombuffer(data[len(hdr):], dtype="<f4").reshape(64, 64, 3)[::-1]  # stored bottom row first
    cfg = W.SceneConfig(cfg_path)
    sc = W.Scene(ctx)
    cfg.add_to(sc)
    sc.build()
    film = W.Film(ctx, 64, 64)
    W.render(sc, cfg.camera, film, 4, cfg.max_len, 7)
    mean = film.read()["mean"].reshape(64, 64, 3)
    assert np.array_equal(img.view(np.uint32), mean.view(np.uint32))
