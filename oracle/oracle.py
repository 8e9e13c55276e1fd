"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker and the CPU baseline.  The product
(walkray_amd) never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
NONE = 0xFFFFFFFF

_lib = None


def build():
    import subprocess
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run make -C oracle)")
        L = C.CDLL(LIB_PATH)
        vp, u32, u64, f32 = C.c_void_p, C.c_uint32, C.c_uint64, C.c_float
        P = C.POINTER
        L.orc_tree_build.restype = vp
        L.orc_tree_build.argtypes = [vp, u32, vp, u32]
        L.orc_tree_free.argtypes = [vp]
        L.orc_tree_stats.argtypes = [vp, P(u32), P(u32), P(u32), P(u32)]
        L.orc_tree_export.argtypes = [vp, vp, vp, vp, vp, vp]
        L.orc_tree_intersect.argtypes = [vp, vp, u64, vp, vp, C.c_int]
        L.orc_box_intersect.restype = f32
        L.orc_box_intersect.argtypes = [vp, vp]
        L.orc_tri_intersect.restype = f32
        L.orc_tri_intersect.argtypes = [vp, vp, vp, vp]
        L.orc_scene_create.restype = vp
        L.orc_scene_free.argtypes = [vp]
        L.orc_scene_add_object.restype = u32
        L.orc_scene_add_object.argtypes = [vp, vp, u32, vp, u32, vp, C.c_int, vp]
        L.orc_scene_build.argtypes = [vp]
        L.orc_scene_intersect.argtypes = [vp, vp, vp, vp, u64, C.c_int, vp, vp, vp, vp, vp, vp, C.c_int]
        L.orc_scene_find_inside.argtypes = [vp, vp, vp, u64, vp, C.c_int]
        L.orc_scene_near_ties.argtypes = [vp, vp, u64, vp, vp, f32, vp, C.c_int]
        L.orc_world_normal.argtypes = [vp, u32, u32, vp, vp, vp, vp]
        L.orc_render.argtypes = [vp, vp, u32, u64, u32, u32, vp, vp, C.c_int]
        L.orc_render_sample.argtypes = [vp, vp, u32, u64, u32, u32, vp, vp]
        L.orc_render2.argtypes = [vp, vp, u32, u64, u32, u32, u32, vp, vp, C.c_int]
        L.orc_render_sample2.argtypes = [vp, vp, u32, u64, u32, u32, u32, vp, vp]
        L.orc_camera_ray.argtypes = [vp, u32, f32, f32, vp, P(f32)]
        L.orc_scatter.restype = C.c_int
        L.orc_scatter.argtypes = [vp, u32, vp, u32, u32, f32, u32, vp, vp, vp]
        L.orc_light_sample.restype = C.c_int
        L.orc_light_sample.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp]
        L.orc_light_paths.restype = u32
        L.orc_light_paths.argtypes = [vp, u32, u64, u32, u32, u32, u32, vp, vp, C.c_int]
        L.orc_hashgrid_build.restype = vp
        L.orc_hashgrid_build.argtypes = [vp, u64, f32, u32]
        L.orc_hashgrid_export.argtypes = [vp, vp, vp, vp]
        L.orc_hashgrid_query.argtypes = [vp, vp, u64, vp, u64, u32, vp, vp]
        L.orc_hashgrid_free.argtypes = [vp]
        L.orc_scene_set_motion.argtypes = [vp, u32, f32, f32, vp, vp]
        L.orc_motion_interpolate.argtypes = [f32, f32, vp, vp, f32, vp]
        L.orc_render_bdpt.argtypes = [vp, vp, u32, u64, u32, u32, vp, vp, C.c_int]
        L.orc_render_vcm.argtypes = [vp, vp, u32, u64, u32, u32, f32, vp, vp, C.c_int]
        L.orc_vcm_radius.restype = f32
        L.orc_render_ppm.argtypes = [vp, vp, u32, u64, u32, u32, f32, vp, vp, C.c_int]
        L.orc_vcm_radius.argtypes = [f32, u32]
        L.orc_philox4x32.argtypes = [vp, vp, vp]
        L.orc_rng4.argtypes = [u64, u32, u32, u32, vp]
        _lib = L
    return _lib


def vcm_radius(base_radius, pss):
    """VCMTracer's merging radius of pass pss (vcmTracer.cpp:30)."""
    return float(lib().orc_vcm_radius(base_radius, pss))


def _p(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


class Camera(C.Structure):
    _fields_ = [("pos", C.c_float * 3), ("focus", C.c_float * 3), ("up", C.c_float * 3),
                ("width", C.c_uint32), ("height", C.c_uint32), ("sight_dist", C.c_float)]

    @classmethod
    def from_dict(cls, cam):
        c = cls()
        c.pos[:] = cam["pos"]
        c.focus[:] = cam["focus"]
        c.up[:] = cam["up"]
        c.width = cam["width"]
        c.height = cam["height"]
        c.sight_dist = cam["sight_dist"]
        return c


def camera_ray(cam, pixel, u, v):
    """Camera::generateRay's direction for pixel and jitter (u, v), and the sample weight
    |front . d| / |front . d_centre| the camera gives it."""
    c = Camera.from_dict(cam)
    d = np.zeros(3, np.float32)
    w = C.c_float()
    lib().orc_camera_ray(C.byref(c), pixel, u, v, _p(d), C.byref(w))
    return d, w.value


class Tree:
    """KDTree restatement (kdtree.h:9-48): build(BEST) + intersect without Condition."""

    def __init__(self, verts, tris):
        self.verts = np.ascontiguousarray(verts, dtype=np.float32)
        self.tris = np.ascontiguousarray(tris, dtype=np.uint32)
        self.h = lib().orc_tree_build(_p(self.verts), len(self.verts), _p(self.tris), len(self.tris))

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.orc_tree_free(self.h)
            self.h = None

    def stats(self):
        a, b, c, d = (C.c_uint32() for _ in range(4))
        lib().orc_tree_stats(self.h, C.byref(a), C.byref(b), C.byref(c), C.byref(d))
        return {"nodes": a.value, "leaf_tris": b.value, "depth": c.value, "leaves": d.value}

    def export(self):
        st = self.stats()
        n, m = st["nodes"], st["leaf_tris"]
        box = np.zeros((n, 6), np.float32)
        right = np.zeros(n, np.uint32)
        lf = np.zeros(n, np.uint32)
        lc = np.zeros(n, np.uint32)
        tl = np.zeros(max(m, 1), np.uint32)
        lib().orc_tree_export(self.h, _p(box), _p(right), _p(lf), _p(lc), _p(tl))
        return {"box": box, "right": right, "leaf_first": lf, "leaf_count": lc, "tri_list": tl[:m]}

    def intersect(self, rays, nthreads=0):
        rays = np.ascontiguousarray(rays, dtype=np.float32)
        n = len(rays)
        prim = np.empty(n, np.uint32)
        t = np.empty(n, np.float32)
        lib().orc_tree_intersect(self.h, _p(rays), n, _p(prim), _p(t), nthreads)
        return prim, t


def box_intersect(box6, ray6):
    b = np.ascontiguousarray(box6, dtype=np.float32)
    r = np.ascontiguousarray(ray6, dtype=np.float32)
    return float(lib().orc_box_intersect(_p(b), _p(r)))


def tri_intersect(v0, v1, v2, ray6):
    a = [np.ascontiguousarray(x, dtype=np.float32) for x in (v0, v1, v2, ray6)]
    return float(lib().orc_tri_intersect(*[_p(x) for x in a]))


class Scene:
    """Scene restatement (scene.cpp) with the four hot materials."""

    def __init__(self, objs):
        L = lib()
        self.h = L.orc_scene_create()
        self._keep = []
        for o in objs:
            params = np.array(list(o["params"]) + [0.0] * 4, dtype=np.float32)
            normals = o["normals"]
            self._keep += [o["verts"], o["tris"], normals, params]
            L.orc_scene_add_object(self.h, _p(o["verts"]), len(o["verts"]), _p(o["tris"]), len(o["tris"]),
                                   _p(normals), o["mat"], _p(params))
        L.orc_scene_build(self.h)

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.orc_scene_free(self.h)
            self.h = None

    def intersect(self, rays, inside=None, contact=None, noself=False, nthreads=0):
        rays = np.ascontiguousarray(rays, dtype=np.float32)
        n = len(rays)
        out = {k: np.empty(n, np.uint32) for k in ("prim", "obj", "tri")}
        out.update({k: np.empty(n, np.float32) for k in ("t", "b1", "b2")})
        ins = None if inside is None else np.ascontiguousarray(inside, dtype=np.uint32)
        con = None if contact is None else np.ascontiguousarray(contact, dtype=np.uint32)
        lib().orc_scene_intersect(self.h, _p(rays), _p(ins), _p(con), n, int(noself), _p(out["prim"]),
                                  _p(out["obj"]), _p(out["tri"]), _p(out["t"]), _p(out["b1"]),
                                  _p(out["b2"]), nthreads)
        return out

    def find_inside(self, rays, cur=None, nthreads=0):
        rays = np.ascontiguousarray(rays, dtype=np.float32)
        n = len(rays)
        out = np.empty(n, np.uint32)
        cu = None if cur is None else np.ascontiguousarray(cur, dtype=np.uint32)
        lib().orc_scene_find_inside(self.h, _p(rays), _p(cu), n, _p(out), nthreads)
        return out

    def world_normal(self, obj, tri, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        n = np.zeros(3, np.float32)
        u1, u2 = C.c_float(), C.c_float()
        lib().orc_world_normal(self.h, obj, tri, _p(p), _p(n), C.byref(u1), C.byref(u2))
        return n, u1.value, u2.value

    def render(self, cam, spp, max_depth, seed, pass0=0, nthreads=0, sums=None, nee=False, path_vector=False):
        """nee: PathTracer's next-event branch (pathTracer.cpp:39-66); path_vector:
        form the sample from the explicit Path (the restatement the branch uses)."""
        c = Camera.from_dict(cam)
        npx = cam["width"] * cam["height"]
        if sums is None:
            sums = (np.zeros(3 * npx, np.float64), np.zeros(3 * npx, np.float64))
        flags = (NEE if nee else 0) | (PATH_VECTOR if path_vector else 0)
        if flags:
            lib().orc_render2(self.h, C.byref(c), max_depth, seed, pass0, spp, flags, _p(sums[0]), _p(sums[1]),
                              nthreads)
        else:
            lib().orc_render(self.h, C.byref(c), max_depth, seed, pass0, spp, _p(sums[0]), _p(sums[1]), nthreads)
        return sums

    def rebuild(self):
        """Scene::updateSceneForMotionBlur's full rebuild (scene.cpp:67-80): every tree again."""
        lib().orc_scene_build(self.h)

    def near_ties(self, rays, prim, t, tol=1e-6, nthreads=0):
        """Per ray, the other triangles accepted within tol of its hit distance t."""
        rays = np.ascontiguousarray(rays, dtype=np.float32)
        prim = np.ascontiguousarray(prim, dtype=np.uint32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        out = np.empty(len(rays), np.uint32)
        lib().orc_scene_near_ties(self.h, _p(rays), len(rays), _p(prim), _p(t), tol, _p(out), nthreads)
        return out

    def scatter(self, self_obj, ray, obj, tri, dist, inside, rn):
        """AbstractObject::scatter of object self_obj for the ray (o, d) hitting local triangle
        tri of object obj at distance dist (insideObj = inside), from three uniforms."""
        r = np.ascontiguousarray(ray, dtype=np.float32).reshape(6)
        u = np.ascontiguousarray(rn, dtype=np.float32).reshape(3)
        out = np.zeros(14, np.float32)
        ids = np.zeros(2, np.uint32)
        delta = lib().orc_scatter(self.h, self_obj, _p(r), obj, tri, dist, inside, _p(u), _p(out), _p(ids))
        return {"origin": out[0:3], "direction": out[3:6], "radiance": out[6:9], "origin_prob": out[9],
                "direction_prob": out[10], "normal": out[11:14], "inside": int(ids[0]), "contact": int(ids[1]),
                "delta": bool(delta)}

    def render_sample(self, cam, max_depth, seed, pixel, pss, nee=False, path_vector=False):
        c = Camera.from_dict(cam)
        rgb = np.zeros(3, np.float32)
        nv = C.c_uint32()
        flags = (NEE if nee else 0) | (PATH_VECTOR if path_vector else 0)
        if flags:
            lib().orc_render_sample2(self.h, C.byref(c), max_depth, seed, flags, pixel, pss, _p(rgb), C.byref(nv))
        else:
            lib().orc_render_sample(self.h, C.byref(c), max_depth, seed, pixel, pss, _p(rgb), C.byref(nv))
        return rgb, nv.value

    def set_motion(self, obj, t0, t1, start, end):
        """AnimatedTransform(start, t0, end, t1) for object obj (its mesh in its own space)."""
        a = np.ascontiguousarray(start, dtype=np.float32).reshape(16)
        b = np.ascontiguousarray(end, dtype=np.float32).reshape(16)
        lib().orc_scene_set_motion(self.h, obj, t0, t1, _p(a), _p(b))

    def light_paths(self, max_depth, seed, pss, path0, n_paths, max_vertices, nthreads=0):
        """Light sub-paths (genLightSample + samplePath): (records, counts, longest).
        records: (n_paths, max_vertices) of PATH_VERTEX_DTYPE."""
        rec = np.zeros((n_paths, max_vertices), PATH_VERTEX_DTYPE)
        counts = np.zeros(n_paths, np.uint32)
        longest = lib().orc_light_paths(self.h, max_depth, seed, pss, path0, n_paths, max_vertices, _p(rec),
                                        _p(counts), nthreads)
        return rec, counts, longest

    def render_bdpt(self, cam, spp, max_depth, seed, pass0=0, nthreads=0, sums=None):
        """BidirectionalPathTracer passes [pass0, pass0 + spp)."""
        c = Camera.from_dict(cam)
        npx = cam["width"] * cam["height"]
        if sums is None:
            sums = (np.zeros(3 * npx, np.float64), np.zeros(3 * npx, np.float64))
        lib().orc_render_bdpt(self.h, C.byref(c), max_depth, seed, pass0, spp, _p(sums[0]), _p(sums[1]), nthreads)
        return sums

    def render_vcm(self, cam, spp, max_depth, seed, base_radius, pass0=0, nthreads=0, sums=None):
        """VCMTracer passes [pass0, pass0 + spp) with the radius schedule from base_radius."""
        c = Camera.from_dict(cam)
        npx = cam["width"] * cam["height"]
        if sums is None:
            sums = (np.zeros(3 * npx, np.float64), np.zeros(3 * npx, np.float64))
        lib().orc_render_vcm(self.h, C.byref(c), max_depth, seed, pass0, spp, base_radius, _p(sums[0]), _p(sums[1]),
                             nthreads)
        return sums

    def render_ppm(self, cam, spp, max_depth, seed, base_radius, pass0=0, nthreads=0, sums=None):
        """PhotonMap (progressive photon mapping) passes [pass0, pass0 + spp)."""
        c = Camera.from_dict(cam)
        npx = cam["width"] * cam["height"]
        if sums is None:
            sums = (np.zeros(3 * npx, np.float64), np.zeros(3 * npx, np.float64))
        lib().orc_render_ppm(self.h, C.byref(c), max_depth, seed, pass0, spp, base_radius, _p(sums[0]), _p(sums[1]),
                             nthreads)
        return sums

    def light_sample(self, u):
        """Scene::genLightSample from four uniforms; None when nothing can be sampled."""
        u = np.ascontiguousarray(u, dtype=np.float32)
        o, n, rad = np.zeros(3, np.float32), np.zeros(3, np.float32), np.zeros(3, np.float32)
        op, obj, tri = C.c_float(), C.c_uint32(), C.c_uint32()
        ok = lib().orc_light_sample(self.h, _p(u), _p(o), _p(n), _p(rad), C.byref(op), C.byref(obj), C.byref(tri))
        if not ok:
            return None
        return {"origin": o, "normal": n, "radiance": rad, "origin_prob": op.value, "obj": obj.value,
                "tri": tri.value}


PATH_VERTEX_DTYPE = np.dtype([("origin", np.float32, 3), ("direction", np.float32, 3),
                              ("radiance", np.float32, 3), ("origin_prob", np.float32),
                              ("direction_prob", np.float32), ("contact", np.uint32), ("contact_tri", np.uint32),
                              ("inside", np.uint32), ("delta", np.uint32), ("pad", np.uint32)])
EMISSIVE = 3


def photons(scene_mats, rec, counts, n_leading=0):
    """The light points VCM / PPM store (vcmTracer.cpp:42-55, photonMap.cpp:66-86): for
    each path in order, vertices 1.. until one whose contact object is emissive,
    skipping isDeltaDirection ones.  The reference's vector starts with one
    value-initialised point per pixel (vcmTracer.cpp:31), at the origin and
    without a path: n_leading of them come first.  Returns (positions (n, 3),
    source (n, 2) = (path, vertex), NONE for the leading points)."""
    pos = [np.zeros((n_leading, 3), np.float32)]
    src = [np.full((n_leading, 2), NONE, np.uint32)]
    mats = np.asarray(scene_mats)
    for k in range(len(counts)):
        for i in range(1, int(counts[k])):
            v = rec[k, i]
            c = int(v["contact"])
            if c != NONE and mats[c] == EMISSIVE:
                break
            if v["delta"]:
                continue
            pos.append(v["origin"][None, :].astype(np.float32))
            src.append(np.array([[k, i]], np.uint32))
    return np.concatenate(pos), np.concatenate(src)


class HashGrid:
    """HashGrid (hashgrid.hxx): Reserve(n_cells) + Build(points, radius); query = Process."""

    def __init__(self, pos, radius, n_cells):
        self.pos = np.ascontiguousarray(pos, dtype=np.float32)
        self.n, self.n_cells = len(self.pos), int(n_cells)
        self.h = lib().orc_hashgrid_build(_p(self.pos), self.n, radius, self.n_cells)

    def export(self):
        bbox = np.zeros(6, np.float32)
        ends = np.zeros(self.n_cells, np.int32)
        idx = np.zeros(max(self.n, 1), np.int32)
        lib().orc_hashgrid_export(self.h, _p(bbox), _p(ends), _p(idx))
        return bbox, ends, idx[:self.n]

    def query(self, q, cap):
        q = np.ascontiguousarray(q, dtype=np.float32)
        counts = np.zeros(len(q), np.uint32)
        idx = np.full((len(q), cap), -1, np.int32)
        lib().orc_hashgrid_query(self.h, _p(self.pos), self.n, _p(q), len(q), cap, _p(counts), _p(idx))
        return counts, idx

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.orc_hashgrid_free(self.h)
            self.h = None


def motion_interpolate(t0, t1, start, end, time):
    """AnimatedTransform::Interpolate (row-major 4x4 in and out)."""
    a = np.ascontiguousarray(start, dtype=np.float32).reshape(16)
    b = np.ascontiguousarray(end, dtype=np.float32).reshape(16)
    out = np.zeros(16, np.float32)
    lib().orc_motion_interpolate(t0, t1, _p(a), _p(b), time, _p(out))
    return out.reshape(4, 4)


NEE = 1          # orc_render2 flags
PATH_VECTOR = 2
NEE_VERTEX = 0xFFFFFFFF  # vertex counter of the light sample's uniforms


def philox4x32(ctr, key):
    c = np.array(ctr, dtype=np.uint32)
    k = np.array(key, dtype=np.uint32)
    o = np.zeros(4, np.uint32)
    lib().orc_philox4x32(_p(c), _p(k), _p(o))
    return o


def rng4(seed, pixel, pss, vertex):
    o = np.zeros(4, np.float32)
    lib().orc_rng4(seed, pixel, pss, vertex, _p(o))
    return o
