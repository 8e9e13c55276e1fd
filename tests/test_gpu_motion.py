"""Motion blur on the GPU through the C ABI: every pass rebuilds the trees with each
moving object at its time (Scene::updateSceneForMotionBlur, scene.cpp:67-80).

Bars: megakernel and wavefront films bit-identical; the moving-emitter closed form
(each pixel covered a quarter of the time); GPU vs oracle within 3 sigma per pixel
and channel with most pixel sums agreeing closely (the same per-pass placements and
the same RNG stream); the built scene shows the objects at their start time.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from test_motion import _trs, moving_emitter
from walkray_amd import _lib as W
from walkray_amd import render as R
from walkray_amd import scenes as S

pytestmark = pytest.mark.gpu


def _scene(ctx, objs, motions):
    sc = W.Scene(ctx, objs, build=False)
    for obj, (t0, t1, a, b) in motions.items():
        sc.set_motion(obj, t0, t1, a, b)
    sc.build()
    return sc


def test_moving_emitter_closed_form(ctx):
    objs, cam, a, b = moving_emitter(width=32, height=4)
    sc = _scene(ctx, objs, {0: (0.0, 1.0, a, b)})
    spp = 512
    print("[motion] moving emitter", flush=True)
    f = {}
    for mode in (W.WR_MEGAKERNEL, W.WR_WAVEFRONT):
        f[mode] = R.render_scene(ctx, objs, cam, spp, 4, 3, mode, scene=sc)
    assert np.array_equal(f[0]["sum"].view(np.uint32), f[1]["sum"].view(np.uint32))
    m = f[0]["sum"].reshape(-1, 3)[:, 0] / spp
    se = math.sqrt(0.25 * 0.75 / spp)
    assert np.all(np.abs(m - 0.25) < 4.5 * se + 2e-3), m


def _cornell_moving(k=2, width=24, height=24):
    objs, cam = S.cornell(k=k, width=width, height=height)
    # the diffuse sphere (object 4, centred at (0.28, 0.201, 0.35)) as a unit-frame mesh
    # that moves: translated, turned about y and shrunk between the two transforms
    c = np.array([0.28, 0.201, 0.35], np.float32)
    objs[4] = dict(objs[4])
    objs[4]["verts"] = (objs[4]["verts"] - c).astype(np.float32)
    a = _trs(c, [0, 1, 0], 0.0, [1.0, 1.0, 1.0])
    b = _trs(c + np.float32([0.25, 0.1, 0.2]), [0, 1, 0], 1.2, [0.7, 0.9, 0.8])
    return objs, cam, {4: (0.0, 1.0, a, b)}


def test_motion_matches_oracle(ctx):
    objs, cam, motions = _cornell_moving()
    sc = _scene(ctx, objs, motions)
    osc = O.Scene(objs)
    for obj, (t0, t1, a, b) in motions.items():
        osc.set_motion(obj, t0, t1, a, b)
    # the built scene: the objects at their start time, bit-identical closest hits
    rays = S.pinhole_rays(cam)
    g = sc.closest_hit(rays)
    o = osc.intersect(rays)
    assert np.array_equal(g["prim"], o["prim"]) and np.array_equal(g["t"].view(np.uint32), o["t"].view(np.uint32))
    spp, depth, seed = 48, 6, 5
    print("[motion] cornell with a moving sphere vs oracle", flush=True)
    film = R.render_scene(ctx, objs, cam, spp, depth, seed, W.WR_WAVEFRONT, scene=sc)
    mega = R.render_scene(ctx, objs, cam, spp, depth, seed, W.WR_MEGAKERNEL, scene=sc)
    assert np.array_equal(film["sum"].view(np.uint32), mega["sum"].view(np.uint32))
    s, q = osc.render(cam, spp, depth, seed)
    mg, seg = R.film_stats(film)
    mo = s / spp
    so = np.sqrt(np.maximum(q / spp - mo * mo, 0) / spp)
    se = np.sqrt(seg ** 2 + so ** 2)
    z = np.where(se > 0, np.abs(mg - mo) / np.where(se > 0, se, 1), 0)
    assert (z > 3).mean() <= 0.01 and z.max() <= 6, ((z > 3).mean(), z.max())
    same = np.isclose(film["sum"], s.astype(np.float32), rtol=1e-4, atol=1e-6).mean()
    assert same > 0.5, same
    # the moving sphere does change the image
    still = R.render_scene(ctx, objs, cam, spp, depth, seed, W.WR_WAVEFRONT)
    assert not np.array_equal(still["sum"], film["sum"])


def test_motion_from_scene_file(ctx, tmp_path):
    """A <motion> object read by the loader renders as the same object given to
    wr_scene_set_motion by hand (wr_config_add_to_scene wires the motion)."""
    (tmp_path / "strip.obj").write_text("v -0.25 -2 0\nv 0.25 -2 0\nv 0.25 2 0\nv -0.25 2 0\nf 1 2 3\nf 1 3 4\n")
    (tmp_path / "mats.xml").write_text('<Material Name="lamp"><type>Emissive</type><color>(1, 1, 1)</color>'
                                       "</Material>")
    m0 = "[1, 0, 0, -1][0, 1, 0, 0][0, 0, 1, 0][0, 0, 0, 1]"
    m1 = "[0.5, 0, 0, 1][0, 0.5, 0, 0][0, 0, 0.5, 0][0, 0, 0, 1]"
    (tmp_path / "scene.xml").write_text(f"""<Scene Name="M">
<Object><filePath>strip.obj</filePath><Material><path>mats.xml</path><name>lamp</name></Material>
<motion><startTime>0</startTime><endTime>1</endTime><startTransform>{m0}</startTransform>
<endTransform>{m1}</endTransform></motion></Object>
<Camera Name="c"><position>(0, 0, 5)</position><focus>(0, 0, 0)</focus><up>(0, 1, 0)</up>
<width>16</width><height>4</height><sightDist>128</sightDist></Camera></Scene>""")
    (tmp_path / "config.xml").write_text("<Config><Scene><path>scene.xml</path><name>M</name><camera>c</camera>"
                                         "</Scene></Config>")
    cfg = W.SceneConfig(str(tmp_path / "config.xml"))
    a = W.Scene(ctx)
    cfg.add_to(a)
    a.build()
    o = cfg.objects[0]
    t0, t1, ma, mb = o["motion"]
    b = _scene(ctx, [o], {0: (t0, t1, ma, mb)})
    print("[motion] from a scene file", flush=True)
    fa = R.render_scene(ctx, None, cfg.camera, 64, 3, 9, W.WR_WAVEFRONT, scene=a)
    fb = R.render_scene(ctx, None, cfg.camera, 64, 3, 9, W.WR_WAVEFRONT, scene=b)
    assert np.array_equal(fa["sum"].view(np.uint32), fb["sum"].view(np.uint32))
    assert fa["mean"].max() > 0


def _placed(objs, motions, seed, pss):
    """The objects as pass `pss` of a render with `seed` places them: the time of each
    moving object from the counter RNG (object id in the pixel slot, vertex 0xFFFFFFFE),
    Interpolate from the oracle, and Mat4::apply's float operations, in order."""
    out = []
    for oi, o in enumerate(objs):
        v = np.asarray(o["verts"], np.float32)
        if oi in motions:
            t0, t1, a, b = motions[oi]
            tm = float(O.rng4(seed, oi, pss, 0xFFFFFFFE)[0]) if pss is not None else t0
            M = O.motion_interpolate(t0, t1, a, b, tm)
            v = ((v[:, 0:1] * M[:3, 0] + v[:, 1:2] * M[:3, 1]) + v[:, 2:3] * M[:3, 2]) + M[:3, 3]
            v = v.astype(np.float32)
        out.append(dict(o, verts=v))
    return out


def _soup(objs):
    V = np.concatenate([o["verts"] for o in objs]).astype(np.float32)
    off = np.cumsum([0] + [len(o["verts"]) for o in objs])
    T = np.concatenate([o["tris"].astype(np.int64) + off[i] for i, o in enumerate(objs)]).astype(np.uint32)
    return V, T


@pytest.mark.parametrize("pss", [0, 7])
def test_motion_pass_scene(ctx, pss):
    """wr_scene_motion_pass: the scene of one pass.  Its reference trees are the
    oracle's for the placed objects (box bits included), its fast-pass hierarchies,
    refit on the device, are the host refit's node for node, and its closest hits are
    the oracle's on the placed scene, bit for bit."""
    objs, cam, motions = _cornell_moving()
    seed = 11
    sc = _scene(ctx, objs, motions)
    print(f"[motion] pass scene {pss}", flush=True)
    ps = sc.motion_pass(seed, pss)
    built, placed = _placed(objs, motions, seed, None), _placed(objs, motions, seed, pss)
    Vb, T = _soup(built)
    Vp, _ = _soup(placed)
    assert not np.array_equal(Vb, Vp)
    ref = O.Tree(Vp, T).export()
    got = ps.export_tree()
    for k in ("right", "leaf_first", "leaf_count", "tri_list"):
        assert np.array_equal(got[k], ref[k]), k
    assert np.array_equal(got["box"].view(np.uint32), np.asarray(ref["box"], np.float32).view(np.uint32))
    # fast-pass hierarchies: the global tree and the moving object's own tree
    host = W.build_tree_wide(Vb, T, refit_verts=Vp)
    dev = ps.export_wide()
    assert dev["root"] == host["root"]
    assert dev["nodes"].tobytes() == host["nodes"].tobytes()
    for oi in motions:
        host = W.build_tree_wide(built[oi]["verts"], objs[oi]["tris"], refit_verts=placed[oi]["verts"])
        dev = ps.export_wide(oi)
        assert dev["root"] == host["root"] and dev["nodes"].tobytes() == host["nodes"].tobytes()
    # a static object's hierarchy is the built scene's
    assert ps.export_wide(0)["nodes"].tobytes() == sc.export_wide(0)["nodes"].tobytes()
    lo, hi = S.scene_bounds(placed)
    rays = np.concatenate([S.pinhole_rays(cam), S.incoherent_rays(20000, lo, hi, seed=3)])
    g = ps.closest_hit(rays)
    o = O.Scene(placed).intersect(rays)
    assert np.array_equal(g["prim"], o["prim"])
    assert np.array_equal(g["t"].view(np.uint32), o["t"].view(np.uint32))
    hit = g["prim"] != O.NONE
    for k in ("b1", "b2"):
        assert np.array_equal(g[k][hit].view(np.uint32), o[k][hit].view(np.uint32))
    ps.close()


def test_motion_pass_of_static_scene_and_errors(ctx):
    """A scene without moving objects: its pass scene is the scene (trees, hierarchies,
    hits).  An unbuilt scene has no pass; a NaN vertex coordinate is refused at build."""
    objs, cam = S.cornell(k=1, width=16, height=16)
    sc = W.Scene(ctx, objs)
    ps = sc.motion_pass(3, 5)
    for obj in (O.NONE, 4):
        a, b = sc.export_tree(obj), ps.export_tree(obj)
        assert all(np.array_equal(a[k], b[k]) for k in ("right", "leaf_first", "leaf_count", "tri_list"))
        assert np.array_equal(a["box"].view(np.uint32), b["box"].view(np.uint32))
        assert sc.export_wide(obj)["nodes"].tobytes() == ps.export_wide(obj)["nodes"].tobytes()
    rays = S.pinhole_rays(cam)
    g, h = sc.closest_hit(rays), ps.closest_hit(rays)
    assert np.array_equal(g["prim"], h["prim"]) and np.array_equal(g["t"].view(np.uint32), h["t"].view(np.uint32))
    ps.close()
    un = W.Scene(ctx, objs, build=False)
    with pytest.raises(W.WalkRayError):
        un.motion_pass(3, 5)
    bad = [dict(o) for o in objs]
    bad[2] = dict(bad[2], verts=bad[2]["verts"].copy())
    bad[2]["verts"][1, 0] = np.nan
    with pytest.raises(W.WalkRayError, match="NaN"):
        W.Scene(ctx, bad)
