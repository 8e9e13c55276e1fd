"""Scene files on the host: Film::savePFM, Mesh::loadShape (.obj / .shp) and
Config::load, checked on small files written here against closed forms
(film.h:36-42,100-110; mesh.cpp:5-174; config.cpp:432-677; nvMatrix.h)."""
import struct

import numpy as np
import pytest

from walkray_amd import _lib as W


def test_pfm_bytes(tmp_path):
    img = np.arange(2 * 3 * 3, dtype=np.float32).reshape(2, 3, 3)  # 2 rows (top first), 3 columns
    p = tmp_path / "a.pfm"
    W.write_pfm(p, img)
    data = p.read_bytes()
    hdr = b"PF\n3 2\n-1\n"
    assert data[:len(hdr)] == hdr
    body = np.frombuffer(data[len(hdr):], dtype="<f4").reshape(2, 3, 3)
    assert np.array_equal(body[0], img[1]) and np.array_equal(body[1], img[0])  # bottom row first


OBJ = """# comment
v 0 0 0
v 2 0 0
v 0 2 0
v 2 2 2
v 1 1
vt 0.5 0.5
vn 0 0 1
vn 0 1 0
vn 1 0 0
f 1 2 3
f 1//1 2//2 3//3
f 2/1/3 3/1/2 4/1/1
f 1 2 4 3
"""


def test_obj_faces_and_normals(tmp_path):
    p = tmp_path / "m.obj"
    p.write_text(OBJ)
    s = W.load_shape(p, normalize=False)
    # "v 1 1" has three fields and is skipped (mesh.cpp:58-59); a quad keeps its first three vertices
    assert np.array_equal(s["verts"], np.array([[0, 0, 0], [2, 0, 0], [0, 2, 0], [2, 2, 2]], np.float32))
    assert np.array_equal(s["tris"], np.array([[0, 1, 2], [0, 1, 2], [1, 2, 3], [0, 1, 3]], np.uint32))
    assert np.array_equal(s["normals"], np.array([[0, 0, 1], [0, 1, 0], [1, 0, 0]], np.float32))
    # normal indices only for the v//vn and v/vt/vn faces, in face order
    assert np.array_equal(s["normal_tris"], np.array([[0, 1, 2], [2, 1, 0]], np.uint32))


def test_obj_texture_without_normals_is_refused(tmp_path):
    p = tmp_path / "bad.obj"
    p.write_text("v 0 0 0\nv 1 0 0\nv 0 1 0\nvt 0 0\nf 1/1 2/1 3/1\n")
    with pytest.raises(W.WalkRayError, match="texture but no normal"):
        W.load_shape(p, normalize=False)


def test_unitize_closed_form(tmp_path):
    # box [1,5] x [0,2] x [0,1]: centre (3, 1, 0.5), largest extent 4 -> (v - c) / 2, exact in float
    p = tmp_path / "b.obj"
    p.write_text("v 1 0 0\nv 5 2 1\nv 5 0 1\nf 1 2 3\n")
    s = W.load_shape(p, normalize=True)
    assert np.array_equal(s["verts"], np.array([[-1, -0.5, -0.25], [1, 0.5, 0.25], [1, -0.5, 0.25]], np.float32))
    # a single point: zero extent, the division is a no-op (nvVector.h:375-379), only the centring applies
    p.write_text("v 3 3 3\nv 3 3 3\nv 3 3 3\nf 1 2 3\n")
    assert np.array_equal(W.load_shape(p)["verts"], np.zeros((3, 3), np.float32))


def test_shp_reads_n_floats_into_n_vertices(tmp_path):
    # mesh.cpp:20-25: size floats are read into a list resized to size vec3s
    p = tmp_path / "t.shp"
    v = [0, 0, 0, 4, 0, 0, 0, 4, 0]
    f = [0, 1, 2]
    p.write_bytes(struct.pack("<i", 9) + struct.pack("<9f", *v) + struct.pack("<i", 3) + struct.pack("<3I", *f))
    s = W.load_shape(p, normalize=False)
    assert s["verts"].shape == (9, 3) and s["tris"].shape == (3, 3)
    assert np.array_equal(s["verts"][:3], np.array(v, np.float32).reshape(3, 3))
    assert not s["verts"][3:].any()
    assert np.array_equal(s["tris"], np.array([[0, 1, 2], [0, 0, 0], [0, 0, 0]], np.uint32))
    p.write_bytes(struct.pack("<i", 9) + struct.pack("<3f", 1, 2, 3))
    with pytest.raises(W.WalkRayError, match="truncated"):
        W.load_shape(p)


CUBE = "".join(f"v {x} {y} {z}\n" for x in (0, 2) for y in (0, 2) for z in (0, 2)) + "f 1 2 4\nf 1 4 3\nf 5 6 8\n"

GROUPS = """v 0 0 0
v 4 0 0
v 0 4 0
g left
v 0 0 4
v 4 4 4
v 4 0 4
f 4 5 6
g right
v 2 2 2
v 2 0 0
v 0 2 0
f 7 8 9
"""

MATS = """<?xml version="1.0"?>
<!-- materials -->
<Material Name="white"><type>Diffuse</type><color>(0.75, 0.75, 0.75)</color></Material>
<Material Name="shiny"><type>Glossy</type><color>(0.5, 0.25, 1)</color><power>100</power></Material>
<Material Name="glass"><type>Refractive</type><IOR>1.33</IOR></Material>
<Material Name="lamp"><type>Emissive</type><color>(17, 12, 4)</color></Material>
<Material Name="mirror"><type>Reflective</type></Material>
"""


def _scene(tmp_path, objects, extra_cam="", renderer="<renderer>PathTracer</renderer><iter>16</iter>"):
    (tmp_path / "cube.obj").write_text(CUBE)
    (tmp_path / "groups.obj").write_text(GROUPS)
    (tmp_path / "mats.xml").write_text(MATS)
    (tmp_path / "scene.xml").write_text(f"""<Scene Name="Test">
{objects}
<Camera Name="main"><position>(0, 1, 5)</position><focus>(0, 0, 0)</focus><up>(0, 1, 0)</up>
<width>64</width><height>48</height><sightDist>100</sightDist>{extra_cam}</Camera>
</Scene>""")
    (tmp_path / "config.xml").write_text(f"""<Config>
<Scene><path>scene.xml</path><name>Test</name><camera>main</camera></Scene>
{renderer}
</Config>""")
    return str(tmp_path / "config.xml")


def _mat(name):
    return f"<Material><path>mats.xml</path><name>{name}</name></Material>"


def test_config_objects_transforms_camera(tmp_path):
    objs = f"""
<GroupObject><filePath>groups.obj</filePath><transform>[1, 0, 0, 10][0, 1, 0, 0][0, 0, 1, 0][0, 0, 0, 1]</transform>
  <Object Name="right">{_mat("lamp")}<transform>[5, 0, 0, 0][0, 5, 0, 0][0, 0, 5, 0][0, 0, 0, 1]</transform></Object>
  <Object Name="left">{_mat("shiny")}</Object>
</GroupObject>
<Object><filePath>cube.obj</filePath>{_mat("white")}
  <transform>[2, 0, 0, 1][0, 2, 0, 0][0, 0, 2, 0][0, 0, 0, 1]</transform></Object>
<Object><filePath>cube.obj</filePath>{_mat("glass")}</Object>
"""
    cfg = W.SceneConfig(_scene(tmp_path, objs))
    assert cfg.name == "Test" and cfg.renderer == "PathTracer" and cfg.spp == 16 and cfg.max_len == 20
    assert cfg.camera == {"pos": [0, 1, 5], "focus": [0, 0, 0], "up": [0, 1, 0], "width": 64, "height": 48,
                          "sight_dist": 100}
    # group objects first, in <Object> order; then plain objects (config.cpp:452-495)
    mats = [o["mat"] for o in cfg.objects]
    assert mats == [W.WR_EMISSIVE, W.WR_GLOSSY, W.WR_DIFFUSE, W.WR_REFRACTIVE]
    assert cfg.objects[0]["params"] == (17.0, 12.0, 4.0)
    assert cfg.objects[1]["params"] == (0.5, 0.25, 1.0, 100.0)
    assert cfg.objects[3]["params"] == (np.float32(1.33).item(),)
    # groups: the whole file's unitize matrix (scale 2/4 = 0.5, its translation in the bottom row) only scales
    # each group; then the group transform (x + 10) replaces the object's own transform (mesh.cpp:119-125)
    right = cfg.objects[0]
    assert np.array_equal(right["verts"], np.array([[11, 1, 1], [11, 0, 0], [10, 1, 0]], np.float32))
    assert np.array_equal(right["tris"], np.array([[0, 1, 2]], np.uint32))
    left = cfg.objects[1]
    assert np.array_equal(left["verts"], np.array([[10, 0, 2], [12, 2, 2], [12, 0, 2]], np.float32))
    # plain object: unitized to [-1, 1]^3, then x' = 2x + 1, y' = 2y, z' = 2z
    cube = cfg.objects[2]
    xs = sorted(set(cube["verts"][:, 0])), sorted(set(cube["verts"][:, 1]))
    assert xs == ([-1.0, 3.0], [-2.0, 2.0])
    assert cube["corner_normals"] is None  # no vertex normals: flat shading (mesh.h:59-65)


def test_config_refuses_off_path_features(tmp_path):
    obj = f"<Object><filePath>cube.obj</filePath>{_mat('mirror')}</Object>"
    with pytest.raises(W.WalkRayError, match="Reflective"):
        W.SceneConfig(_scene(tmp_path, obj))
    obj = f"<Object><filePath>cube.obj</filePath>{_mat('white')}<motion><startTime>0</startTime></motion></Object>"
    with pytest.raises(W.WalkRayError, match="endTime"):  # an incomplete <motion>
        W.SceneConfig(_scene(tmp_path, obj))
    obj = f"<Object><filePath>cube.obj</filePath>{_mat('white')}</Object>"
    with pytest.raises(W.WalkRayError, match="depth of field"):
        W.SceneConfig(_scene(tmp_path, obj, extra_cam="<lensRadius>1</lensRadius><focalDist>5</focalDist>"))
    cfg = W.SceneConfig(_scene(tmp_path, obj, renderer=""))
    assert cfg.renderer == "" and cfg.spp == 0 and cfg.max_len == 0
    assert cfg.radius == 0.0
    vcm = W.SceneConfig(_scene(tmp_path, obj, renderer="<renderer>VCMTracer</renderer><iter>8</iter><radius>0.004</radius>"))
    assert (vcm.renderer, vcm.spp, vcm.max_len) == ("VCMTracer", 8, 20)
    assert vcm.radius == np.float32(0.004)  # setRadius(atof(...)) (config.cpp:560)


def test_config_normals_follow_inverse_transpose(tmp_path):
    (tmp_path / "n.obj").write_text("v 0 0 0\nv 1 0 0\nv 0 1 0\nvn 0 0 1\nvn 0.6 0 0.8\nvn 0 0.6 0.8\n"
                                    "f 1//1 2//2 3//3\n")
    rot = "[0, -3, 0, 1][2, 0, 0, 0][0, 0, 4, 2][0, 0, 0, 1]"
    obj = f"<Object><filePath>n.obj</filePath>{_mat('white')}<transform>{rot}</transform></Object>"
    cfg = W.SceneConfig(_scene(tmp_path, obj))
    o = cfg.objects[0]
    M = np.array([[0, -3, 0, 1], [2, 0, 0, 0], [0, 0, 4, 2], [0, 0, 0, 1]], np.float64)
    nm = np.linalg.inv(M).T
    n = np.array([[0, 0, 1], [0.6, 0, 0.8], [0, 0.6, 0.8]], np.float64)
    want = (nm[:3, :3] @ n.T).T
    assert o["corner_normals"] is not None
    assert np.allclose(o["corner_normals"][0], want, rtol=1e-6, atol=1e-7)
    # world positions: M (x, y, z, 1) of the unitized triangle (centre (0.5, 0.5, 0), half extent 0.5)
    u = np.array([[-1, -1, 0], [1, -1, 0], [-1, 1, 0]], np.float64)
    wantv = (M[:3, :3] @ u.T).T + M[:3, 3]
    assert np.array_equal(o["verts"], wantv.astype(np.float32))


def test_committed_cornell_files_load():
    """tests/golden/scenes/cornell (written by tests/golden/make_scene_files.py) reads back as the
    procedural Cornell box, vertices equal up to the loader's float rounding."""
    import os
    from walkray_amd import scenes as S
    path = os.path.join(os.path.dirname(__file__), "golden", "scenes", "cornell", "config.xml")
    cfg = W.SceneConfig(path)
    objs, cam = S.cornell(k=2, width=64, height=64)
    assert (cfg.name, cfg.renderer, cfg.spp, cfg.max_len) == ("Cornell", "PathTracer", 16, 20)
    assert len(cfg.objects) == len(objs)
    for a, b in zip(cfg.objects, objs):
        assert a["mat"] == b["mat"] and np.array_equal(a["tris"], b["tris"])
        assert np.allclose(a["verts"], b["verts"], rtol=0, atol=2e-6)
        assert np.allclose(a["params"], np.float32(b["params"]), rtol=0, atol=0)
        assert (a["normals"] is None) == (b["normals"] is None)


def test_config_motion_kept_in_own_space(tmp_path):
    """<motion> (config.cpp:28-45): the object's mesh stays in its own (unitized) space,
    whatever static or group transform it has, and its two transforms and times are kept."""
    m0 = "[1, 0, 0, -1][0, 1, 0, 0][0, 0, 1, 0][0, 0, 0, 1]"
    m1 = "[0, -1, 0, 2][1, 0, 0, 0][0, 0, 2, 0.5][0, 0, 0, 1]"
    mo = (f"<motion><startTime>0.25</startTime><endTime>0.75</endTime><startTransform>{m0}</startTransform>"
          f"<endTransform>{m1}</endTransform></motion>")
    objs = f"""<Object><filePath>cube.obj</filePath>{_mat("white")}{mo}
  <transform>[2, 0, 0, 1][0, 2, 0, 0][0, 0, 2, 0][0, 0, 0, 1]</transform></Object>
<Object><filePath>cube.obj</filePath>{_mat("glass")}</Object>"""
    cfg = W.SceneConfig(_scene(tmp_path, objs))
    moving, still = cfg.objects
    assert still["motion"] is None
    t0, t1, a, b = moving["motion"]
    assert (t0, t1) == (0.25, 0.75)
    assert np.array_equal(a, np.array([[1, 0, 0, -1], [0, 1, 0, 0], [0, 0, 1, 0], [0, 0, 0, 1]], np.float32))
    assert np.array_equal(b, np.array([[0, -1, 0, 2], [1, 0, 0, 0], [0, 0, 2, 0.5], [0, 0, 0, 1]], np.float32))
    # own space: the unitized cube, the static <transform> not applied
    assert sorted(set(moving["verts"][:, 0])) == [-1.0, 1.0]
    assert np.array_equal(moving["verts"], still["verts"])
