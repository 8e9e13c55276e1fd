"""ctypes binding of libwalkray's C ABI (include/walkray.h).

Loading fails loudly when the HIP library has not been built: there is no
CPU fallback anywhere in this package.
"""
from __future__ import annotations

import ctypes as C
import os
import re

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libwalkray.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "walkray.h")
NONE = 0xFFFFFFFF
OVERFLOW = 0xFFFFFFFE

WR_DIFFUSE, WR_GLOSSY, WR_REFRACTIVE, WR_EMISSIVE = 0, 1, 2, 3
WR_BEST = 1
WR_HASHGRID_NOALLOC_MAX_QUERIES = 1 << 22  # walkray.h

_lib = None


class WalkRayError(RuntimeError):
    pass


class Material(C.Structure):
    _fields_ = [("type", C.c_int32), ("color", C.c_float * 3), ("power", C.c_float), ("ior", C.c_float)]


class SceneInfo(C.Structure):
    _fields_ = [("n_objects", C.c_uint32), ("n_triangles", C.c_uint32), ("n_nodes", C.c_uint32),
                ("n_leaves", C.c_uint32), ("depth", C.c_uint32), ("max_obj_depth", C.c_uint32),
                ("device_bytes", C.c_uint64), ("build_seconds", C.c_double)]


class CameraDesc(C.Structure):
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


class RenderParams(C.Structure):
    _fields_ = [("spp", C.c_uint32), ("max_depth", C.c_uint32), ("seed", C.c_uint64), ("mode", C.c_uint32),
                ("wave_slots", C.c_uint32), ("wave_grid", C.c_uint32), ("flags", C.c_uint32)]


class RenderStats(C.Structure):
    _fields_ = [("paths", C.c_uint64), ("iterations", C.c_uint64), ("kernel_ms", C.c_float), ("build_ms", C.c_float)]


class ServerConfig(C.Structure):
    _fields_ = [("job_memory_budget", C.c_uint64), ("path_slots", C.c_uint32), ("max_jobs", C.c_uint32)]


class JobDescC(C.Structure):
    _fields_ = [("scene", C.c_uint32), ("cam", CameraDesc), ("spp", C.c_uint32), ("max_depth", C.c_uint32),
                ("flags", C.c_uint32), ("seed", C.c_uint64)]


class JobStatus(C.Structure):
    _fields_ = [("state", C.c_int32), ("reserved", C.c_uint32), ("submit_s", C.c_double), ("admit_s", C.c_double),
                ("done_s", C.c_double), ("samples", C.c_uint64), ("device_bytes", C.c_uint64)]


class ChTuning(C.Structure):
    _fields_ = [("refill_min", C.c_uint32), ("node_phase_min", C.c_uint32), ("blocks_per_cu", C.c_uint32),
                ("lds_levels", C.c_uint32), ("batch", C.c_uint32), ("ms", C.c_float)]


class RenderTuning(C.Structure):
    _fields_ = [("wave_slots", C.c_uint32), ("wave_grid", C.c_uint32), ("lds_levels", C.c_uint32), ("ms", C.c_float)]


def autotune_render(scene, cam, max_depth, spp=2):
    """wr_autotune_render: picks the wavefront queue size and stage grid for this context."""
    t = RenderTuning()
    _check(lib().wr_autotune_render(scene.ctx.h, scene.h, C.byref(CameraDesc.from_dict(cam)), max_depth, spp,
                                    C.byref(t)), scene.ctx.h, "wr_autotune_render")
    return {"wave_slots": t.wave_slots, "wave_grid": t.wave_grid, "lds_levels": t.lds_levels, "ms": t.ms}


def set_render_tuning(ctx, wave_slots=0, wave_grid=0, lds_levels=8):
    """wr_set_render_tuning: the context's wavefront configuration (0 = defaults)."""
    t = RenderTuning(wave_slots, wave_grid, lds_levels, 0.0)
    _check(lib().wr_set_render_tuning(ctx.h, C.byref(t)), ctx.h, "wr_set_render_tuning")


class ConfigInfo(C.Structure):
    _fields_ = [("n_objects", C.c_uint32), ("spp", C.c_int32), ("max_len", C.c_int32), ("camera", CameraDesc),
                ("renderer", C.c_char * 32), ("scene_name", C.c_char * 64), ("radius", C.c_float)]


JOB_QUEUED, JOB_RUNNING, JOB_DONE, JOB_FAILED, JOB_CANCELLED = 0, 1, 2, 3, 4


def header_symbols():
    """Function names declared in include/walkray.h."""
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(wr_[a-z0-9_]+)\s*\(", txt)))


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise WalkRayError(f"libwalkray.so is not built ({LIB_PATH}); run __graft_entry__.build() "
                           "or make -C walkray_amd")
    L = C.CDLL(LIB_PATH)
    vp, u32, u64, i32 = C.c_void_p, C.c_uint32, C.c_uint64, C.c_int
    P = C.POINTER
    sig = {
        "wr_ctx_create": (i32, [i32, P(vp)]),
        "wr_ctx_destroy": (None, [vp]),
        "wr_last_error": (C.c_char_p, [vp]),
        "wr_synchronize": (i32, [vp]),
        "wr_device_alloc": (i32, [vp, C.c_size_t, P(vp)]),
        "wr_device_free": (i32, [vp, vp]),
        "wr_memcpy_htod": (i32, [vp, vp, vp, C.c_size_t]),
        "wr_memcpy_dtoh": (i32, [vp, vp, vp, C.c_size_t]),
        "wr_memcpy_dtod": (i32, [vp, vp, vp, C.c_size_t]),
        "wr_timer_start": (i32, [vp]),
        "wr_timer_stop": (i32, [vp, P(C.c_float)]),
        "wr_scene_create": (i32, [vp, P(vp)]),
        "wr_scene_destroy": (None, [vp]),
        "wr_scene_add_mesh": (i32, [vp, vp, u32, vp, u32, vp, P(Material), P(u32)]),
        "wr_scene_build": (i32, [vp, i32]),
        "wr_scene_get_info": (i32, [vp, P(SceneInfo)]),
        "wr_scene_prim_source": (i32, [vp, u32, P(u32), P(u32)]),
        "wr_scene_export_tree": (i32, [vp, u32, P(u32), P(u32), vp, vp, vp, vp, vp]),
        "wr_scene_motion_pass": (i32, [vp, u64, u32, P(vp)]),
        "wr_scene_export_wide": (i32, [vp, u32, P(u32), P(u32), vp]),
        "wr_tree_build": (i32, [vp, u32, vp, u32, i32, P(vp)]),
        "wr_tree_export": (i32, [vp, P(u32), P(u32), P(u32), vp, vp, vp, vp, vp]),
        "wr_tree_free": (None, [vp]),
        "wr_tree_wide": (i32, [vp, P(u32), P(u32), P(u32), vp, vp, vp]),
        "wr_tree_build_refit": (i32, [vp, u32, vp, u32, i32, vp, P(vp)]),
        "wr_closest_hit": (i32, [vp, vp, vp, u64, vp, vp]),
        "wr_closest_hit_noself": (i32, [vp, vp, vp, u64, vp]),
        "wr_inside_object": (i32, [vp, vp, vp, u64, vp]),
        "wr_film_create": (i32, [vp, u32, u32, P(vp)]),
        "wr_film_destroy": (None, [vp]),
        "wr_film_clear": (i32, [vp]),
        "wr_film_read": (i32, [vp, vp, vp, vp, vp, P(u32)]),
        "wr_film_write": (i32, [vp, vp, vp, vp, u32]),
        "wr_render": (i32, [vp, vp, P(CameraDesc), P(RenderParams), vp, P(RenderStats)]),
        "wr_render_bdpt": (i32, [vp, vp, P(CameraDesc), P(RenderParams), vp, P(RenderStats)]),
        "wr_render_vcm": (i32, [vp, vp, P(CameraDesc), P(RenderParams), C.c_float, vp, P(RenderStats)]),
        "wr_vcm_radius": (C.c_float, [C.c_float, u32]),
        "wr_render_ppm": (i32, [vp, vp, P(CameraDesc), P(RenderParams), C.c_float, vp, P(RenderStats)]),
        "wr_server_create": (i32, [vp, P(ServerConfig), P(vp)]),
        "wr_server_destroy": (None, [vp]),
        "wr_server_add_scene": (i32, [vp, vp, P(u32)]),
        "wr_job_submit": (i32, [vp, P(JobDescC), P(u64)]),
        "wr_server_step": (i32, [vp, P(u32)]),
        "wr_job_get_status": (i32, [vp, u64, P(JobStatus)]),
        "wr_job_read": (i32, [vp, u64, vp, vp, vp, vp, P(u32)]),
        "wr_job_release": (i32, [vp, u64]),
        "wr_job_cancel": (i32, [vp, u64]),
        "wr_render_stage_stats": (i32, [vp, vp]),
        "wr_server_inject_admit_failures": (i32, [vp, u32]),
        "wr_hit_sources": (i32, [vp, vp, vp, u64, vp, vp]),
        "wr_server_clock": (C.c_double, [vp]),
        "wr_autotune_closest_hit": (i32, [vp, vp, vp, u64, P(ChTuning)]),
        "wr_scene_set_closest_hit_tuning": (i32, [vp, P(ChTuning)]),
        "wr_autotune_render": (i32, [vp, vp, P(CameraDesc), u32, u32, P(RenderTuning)]),
        "wr_set_render_tuning": (i32, [vp, P(RenderTuning)]),
        "wr_scene_add_mesh_corners": (i32, [vp, vp, u32, vp, u32, vp, P(Material), P(u32)]),
        "wr_config_load": (i32, [C.c_char_p, P(vp)]),
        "wr_config_free": (None, [vp]),
        "wr_config_get_info": (i32, [vp, P(ConfigInfo)]),
        "wr_config_get_object": (i32, [vp, u32, P(u32), P(u32), P(u32), vp, vp, vp, P(Material)]),
        "wr_config_add_to_scene": (i32, [vp, vp]),
        "wr_config_get_motion": (i32, [vp, u32, P(i32), P(C.c_float), P(C.c_float), vp, vp]),
        "wr_shape_load": (i32, [C.c_char_p, i32, P(vp)]),
        "wr_shape_free": (None, [vp]),
        "wr_shape_get": (i32, [vp, P(u32), P(u32), P(u32), P(u32), vp, vp, vp, vp]),
        "wr_pfm_write": (i32, [C.c_char_p, vp, u32, u32]),
        "wr_light_paths": (i32, [vp, vp, u32, u64, u32, u32, u32, u32, vp, vp]),
        "wr_scene_set_motion": (i32, [vp, u32, C.c_float, C.c_float, vp, vp]),
        "wr_motion_interpolate": (i32, [C.c_float, C.c_float, vp, vp, C.c_float, vp]),
        "wr_photons_collect": (i32, [vp, vp, vp, vp, u32, u32, u64, vp, vp, u64, P(u64)]),
        "wr_hashgrid_build": (i32, [vp, vp, u64, C.c_float, u32, P(vp)]),
        "wr_hashgrid_export": (i32, [vp, vp, vp, vp]),
        "wr_hashgrid_query": (i32, [vp, vp, u64, u32, vp, vp]),
        "wr_hashgrid_query_noalloc": (i32, [vp, vp, u64, u32, vp, vp]),
        "wr_hashgrid_destroy": (None, [vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _check(rc, ctx=None, what=""):
    if rc != 0:
        msg = lib().wr_last_error(ctx)
        raise WalkRayError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")


def build_tree_export(verts, tris, nthreads=0):
    """Host-only builder run + canonical export (same format as the oracle's Tree.export)."""
    L = lib()
    v = np.ascontiguousarray(verts, dtype=np.float32)
    t = np.ascontiguousarray(tris, dtype=np.uint32)
    h = C.c_void_p()
    _check(L.wr_tree_build(_p(v), len(v), _p(t), len(t), nthreads, C.byref(h)), None, "wr_tree_build")
    try:
        n, m, d = C.c_uint32(), C.c_uint32(), C.c_uint32()
        _check(L.wr_tree_export(h, C.byref(n), C.byref(m), C.byref(d), None, None, None, None, None), None,
               "wr_tree_export")
        box = np.zeros((n.value, 6), np.float32)
        right, lf, lc = (np.zeros(n.value, np.uint32) for _ in range(3))
        tl = np.zeros(max(m.value, 1), np.uint32)
        _check(L.wr_tree_export(h, C.byref(n), C.byref(m), C.byref(d), _p(box), _p(right), _p(lf), _p(lc),
                                _p(tl)), None, "wr_tree_export")
        return {"box": box, "right": right, "leaf_first": lf, "leaf_count": lc, "tri_list": tl[:m.value],
                "depth": d.value}
    finally:
        L.wr_tree_free(h)


WNODE_DTYPE = np.dtype([("origin", np.float32, 3), ("scale", np.float32, 3), ("q", np.uint8, (6, 4)),
                        ("ref", np.uint32, 4)])


def build_tree_wide(verts, tris, nthreads=0, refit_verts=None):
    """Host-only builder run; returns the fast-pass hierarchy (wide nodes, root, record tables).
    refit_verts: the same triangles at these positions instead, with the first build's
    hierarchy refit to them (the per-pass rebuild of moving objects)."""
    L = lib()
    v = np.ascontiguousarray(verts, dtype=np.float32)
    t = np.ascontiguousarray(tris, dtype=np.uint32)
    h = C.c_void_p()
    _check(L.wr_tree_build(_p(v), len(v), _p(t), len(t), nthreads, C.byref(h)), None, "wr_tree_build")
    if refit_verts is not None:
        v2 = np.ascontiguousarray(refit_verts, dtype=np.float32)
        h2 = C.c_void_p()
        try:
            _check(L.wr_tree_build_refit(_p(v2), len(v2), _p(t), len(t), nthreads, h, C.byref(h2)), None,
                   "wr_tree_build_refit")
        finally:
            L.wr_tree_free(h)
        h = h2
    try:
        nw, nr, root = C.c_uint32(), C.c_uint32(), C.c_uint32()
        _check(L.wr_tree_wide(h, C.byref(nw), C.byref(nr), C.byref(root), None, None, None), None, "wr_tree_wide")
        nodes = np.zeros(max(nw.value, 1), WNODE_DTYPE)
        rf = np.zeros(max(nr.value, 1), np.uint32)
        rt = np.zeros(max(nr.value, 1), np.uint32)
        _check(L.wr_tree_wide(h, C.byref(nw), C.byref(nr), C.byref(root), _p(nodes), _p(rf), _p(rt)), None,
               "wr_tree_wide")
        return {"nodes": nodes[:nw.value], "root": root.value, "record_first": rf[:nr.value],
                "record_tri": rt[:nr.value]}
    finally:
        L.wr_tree_free(h)


def load_shape(path, normalize=True):
    """Mesh::loadShape of one .obj / .shp file (host only): dict of numpy arrays."""
    L = lib()
    h = C.c_void_p()
    _check(L.wr_shape_load(os.fsencode(path), int(bool(normalize)), C.byref(h)), None, "wr_shape_load")
    try:
        nv, nt, nn, nfn = C.c_uint32(), C.c_uint32(), C.c_uint32(), C.c_uint32()
        _check(L.wr_shape_get(h, C.byref(nv), C.byref(nt), C.byref(nn), C.byref(nfn), None, None, None, None), None,
               "wr_shape_get")
        v = np.zeros((nv.value, 3), np.float32)
        f = np.zeros((nt.value, 3), np.uint32)
        vn = np.zeros((nn.value, 3), np.float32)
        fn = np.zeros((nfn.value, 3), np.uint32)
        _check(L.wr_shape_get(h, C.byref(nv), C.byref(nt), C.byref(nn), C.byref(nfn), _p(v), _p(f), _p(vn), _p(fn)),
               None, "wr_shape_get")
        return {"verts": v, "tris": f, "normals": vn, "normal_tris": fn}
    finally:
        L.wr_shape_free(h)


def write_pfm(path, rgb, width=None, height=None):
    """Film::setBuffer + savePFM of an image, top row first: H x W x 3, or flat with width and height."""
    a = np.ascontiguousarray(rgb, dtype=np.float32)
    if width is None or height is None:
        height, width = a.shape[0], a.shape[1]
    if a.size != 3 * width * height:
        raise ValueError(f"write_pfm: {a.size} values for a {width} x {height} RGB image")
    _check(lib().wr_pfm_write(os.fsencode(path), _p(a), width, height), None, "wr_pfm_write")


class SceneConfig:
    """Config::load of a reference scene configuration (host only)."""

    def __init__(self, path):
        self.h = C.c_void_p()
        _check(lib().wr_config_load(os.fsencode(path), C.byref(self.h)), None, "wr_config_load")
        info = ConfigInfo()
        _check(lib().wr_config_get_info(self.h, C.byref(info)), None, "wr_config_get_info")
        c = info.camera
        self.camera = {"pos": list(c.pos), "focus": list(c.focus), "up": list(c.up), "width": c.width,
                       "height": c.height, "sight_dist": c.sight_dist}
        self.spp, self.max_len = info.spp, info.max_len
        self.renderer = info.renderer.decode()
        self.radius = float(info.radius)
        self.name = info.scene_name.decode()
        self.objects = [self._object(i) for i in range(info.n_objects)]

    def _object(self, i):
        L = lib()
        nv, nt, hn = C.c_uint32(), C.c_uint32(), C.c_uint32()
        m = Material()
        _check(L.wr_config_get_object(self.h, i, C.byref(nv), C.byref(nt), C.byref(hn), None, None, None, None),
               None, "wr_config_get_object")
        v = np.zeros((nv.value, 3), np.float32)
        t = np.zeros((nt.value, 3), np.uint32)
        cn = np.zeros((nt.value, 3, 3), np.float32) if hn.value else None
        _check(L.wr_config_get_object(self.h, i, C.byref(nv), C.byref(nt), C.byref(hn), _p(v), _p(t), _p(cn),
                                      C.byref(m)), None, "wr_config_get_object")
        if m.type == WR_REFRACTIVE:
            params = (m.ior,)
        elif m.type == WR_GLOSSY:
            params = tuple(m.color) + (m.power,)
        else:
            params = tuple(m.color)
        # per-vertex normals when every corner of a vertex carries the same normal
        normals = None
        if cn is not None:
            normals = np.zeros_like(v)
            normals[t.reshape(-1)] = cn.reshape(-1, 3)
            if not np.array_equal(normals[t.reshape(-1)], cn.reshape(-1, 3)):
                normals = None
        moving = C.c_int32()
        t0, t1 = C.c_float(), C.c_float()
        a, b = np.zeros(16, np.float32), np.zeros(16, np.float32)
        _check(L.wr_config_get_motion(self.h, i, C.byref(moving), C.byref(t0), C.byref(t1), _p(a), _p(b)), None,
               "wr_config_get_motion")
        motion = (t0.value, t1.value, a.reshape(4, 4), b.reshape(4, 4)) if moving.value else None
        return {"verts": v, "tris": t, "normals": normals, "corner_normals": cn, "mat": m.type,
                "params": tuple(float(x) for x in params), "motion": motion}

    def add_to(self, scene):
        _check(lib().wr_config_add_to_scene(self.h, scene.h), scene.ctx.h, "wr_config_add_to_scene")

    def close(self):
        if self.h:
            lib().wr_config_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceBuffer:
    def __init__(self, ctx, nbytes):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        self.ptr = C.c_void_p()
        _check(lib().wr_device_alloc(ctx.h, max(self.nbytes, 1), C.byref(self.ptr)), ctx.h, "wr_device_alloc")

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= self.nbytes
        _check(lib().wr_memcpy_htod(self.ctx.h, self.ptr, _p(arr), arr.nbytes), self.ctx.h, "wr_memcpy_htod")

    def download(self, dtype, count):
        out = np.empty(count, dtype=dtype)
        assert out.nbytes <= self.nbytes
        _check(lib().wr_memcpy_dtoh(self.ctx.h, _p(out), self.ptr, out.nbytes), self.ctx.h, "wr_memcpy_dtoh")
        return out

    def free(self):
        if self.ptr and self.ptr.value:
            lib().wr_device_free(self.ctx.h, self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class DeviceView:
    """A byte offset into a DeviceBuffer (keeps the buffer alive)."""

    def __init__(self, buf, offset):
        self.buf = buf
        self.ptr = C.c_void_p(buf.ptr.value + int(offset))


class Context:
    """One HIP device + stream (wr_ctx)."""

    def __init__(self, device=0):
        self.h = C.c_void_p()
        _check(lib().wr_ctx_create(device, C.byref(self.h)), None, "wr_ctx_create")

    def alloc(self, nbytes):
        return DeviceBuffer(self, nbytes)

    def to_device(self, arr):
        arr = np.ascontiguousarray(arr)
        b = DeviceBuffer(self, arr.nbytes)
        b.upload(arr)
        return b

    def synchronize(self):
        _check(lib().wr_synchronize(self.h), self.h, "wr_synchronize")

    def copy(self, dst, src, nbytes):
        """Device-to-device copy (asynchronous on the context stream)."""
        _check(lib().wr_memcpy_dtod(self.h, dst.ptr, src.ptr, nbytes), self.h, "wr_memcpy_dtod")

    def timer_start(self):
        _check(lib().wr_timer_start(self.h), self.h, "wr_timer_start")

    def timer_stop(self):
        ms = C.c_float()
        _check(lib().wr_timer_stop(self.h, C.byref(ms)), self.h, "wr_timer_stop")
        return ms.value

    def close(self):
        if self.h:
            lib().wr_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


WR_MEGAKERNEL, WR_WAVEFRONT = 0, 1
WR_RENDER_NEE = 1
WR_RENDER_DIRECT, WR_RENDER_STAGE_STATS = 2, 4
STAGES = ("trace", "finish", "shade", "shade_exits", "plan", "generate", "fold", "advance")


class StageStatsC(C.Structure):
    _fields_ = [("ms", C.c_float * 8), ("segments", C.c_uint64), ("exits", C.c_uint64), ("generated", C.c_uint64),
                ("box_records", C.c_uint64), ("tri_records", C.c_uint64)]


def render_stage_stats(ctx):
    """wr_render_stage_stats: per-stage times and counts of the last WR_RENDER_STAGE_STATS render."""
    s = StageStatsC()
    _check(lib().wr_render_stage_stats(ctx.h, C.byref(s)), ctx.h, "wr_render_stage_stats")
    return {"ms": dict(zip(STAGES, (float(x) for x in s.ms))), "segments": s.segments, "exits": s.exits,
            "generated": s.generated, "box_records": s.box_records, "tri_records": s.tri_records}


class Film:
    """wr_film: Kahan-compensated per-pixel sums of the per-pass samples."""

    def __init__(self, ctx, width, height):
        self.ctx = ctx
        self.width, self.height = int(width), int(height)
        self.h = C.c_void_p()
        _check(lib().wr_film_create(ctx.h, self.width, self.height, C.byref(self.h)), ctx.h, "wr_film_create")

    def clear(self):
        _check(lib().wr_film_clear(self.h), self.ctx.h, "wr_film_clear")

    def read(self):
        n = 3 * self.width * self.height
        mean, s, c, q = (np.zeros(n, np.float32) for _ in range(4))
        passes = C.c_uint32()
        _check(lib().wr_film_read(self.h, _p(mean), _p(s), _p(c), _p(q), C.byref(passes)), self.ctx.h,
               "wr_film_read")
        return {"mean": mean, "sum": s, "comp": c, "sumsq": q, "passes": passes.value}

    def write(self, state):
        _check(lib().wr_film_write(self.h, _p(np.ascontiguousarray(state["sum"], np.float32)),
                                   _p(np.ascontiguousarray(state["comp"], np.float32)),
                                   _p(np.ascontiguousarray(state["sumsq"], np.float32)), int(state["passes"])),
               self.ctx.h, "wr_film_write")

    def close(self):
        if self.h and self.ctx.h:  # a destroyed context already took the device memory with it
            lib().wr_film_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def motion_interpolate(t0, t1, start, end, time):
    """wr_motion_interpolate: AnimatedTransform::Interpolate on the host (row-major 4x4)."""
    a = np.ascontiguousarray(start, dtype=np.float32).reshape(16)
    b = np.ascontiguousarray(end, dtype=np.float32).reshape(16)
    out = np.zeros(16, np.float32)
    _check(lib().wr_motion_interpolate(t0, t1, _p(a), _p(b), time, _p(out)), None, "wr_motion_interpolate")
    return out.reshape(4, 4)


def render(scene, cam, film, spp, max_depth, seed, mode=WR_WAVEFRONT, wave_slots=0, wave_grid=0, nee=False,
           launch=0):
    """PathTracer::renderPixels for `spp` more passes into `film`; returns stats.
    nee: the useNextEventEstimation branch (pathTracer.cpp:39-66).  launch: 0 (graph
    replays), WR_RENDER_DIRECT or WR_RENDER_STAGE_STATS (wavefront only)."""
    cd = CameraDesc.from_dict(cam)
    prm = RenderParams(spp, max_depth, seed, mode, wave_slots, wave_grid, (WR_RENDER_NEE if nee else 0) | launch)
    st = RenderStats()
    _check(lib().wr_render(scene.ctx.h, scene.h, C.byref(cd), C.byref(prm), film.h, C.byref(st)), scene.ctx.h,
           "wr_render")
    return {"paths": st.paths, "iterations": st.iterations, "kernel_ms": st.kernel_ms, "build_ms": st.build_ms}


def render_bdpt(scene, cam, film, spp, max_depth, seed, queue=0):
    """BidirectionalPathTracer::renderPixels for `spp` more passes into `film`; returns stats.
    queue: visibility-queue entries (0 = automatic)."""
    cd = CameraDesc.from_dict(cam)
    prm = RenderParams(spp, max_depth, seed, 0, queue, 0, 0)
    st = RenderStats()
    _check(lib().wr_render_bdpt(scene.ctx.h, scene.h, C.byref(cd), C.byref(prm), film.h, C.byref(st)), scene.ctx.h,
           "wr_render_bdpt")
    return {"paths": st.paths, "kernel_ms": st.kernel_ms}


def render_vcm(scene, cam, film, spp, max_depth, seed, base_radius, queue=0):
    """VCMTracer::renderPixels for `spp` more passes into `film` with the configuration's
    radius `base_radius`; returns stats.  queue: visibility-queue entries (0 = automatic)."""
    cd = CameraDesc.from_dict(cam)
    prm = RenderParams(spp, max_depth, seed, 0, queue, 0, 0)
    st = RenderStats()
    _check(lib().wr_render_vcm(scene.ctx.h, scene.h, C.byref(cd), C.byref(prm), base_radius, film.h, C.byref(st)),
           scene.ctx.h, "wr_render_vcm")
    return {"paths": st.paths, "kernel_ms": st.kernel_ms}


def render_ppm(scene, cam, film, spp, max_depth, seed, base_radius):
    """PhotonMap::renderPixels (progressive photon mapping) for `spp` more passes into `film`
    with the configuration's radius `base_radius`; returns stats."""
    cd = CameraDesc.from_dict(cam)
    prm = RenderParams(spp, max_depth, seed, 0, 0, 0, 0)
    st = RenderStats()
    _check(lib().wr_render_ppm(scene.ctx.h, scene.h, C.byref(cd), C.byref(prm), base_radius, film.h, C.byref(st)),
           scene.ctx.h, "wr_render_ppm")
    return {"paths": st.paths, "kernel_ms": st.kernel_ms}


def vcm_radius(base_radius, pss):
    """wr_vcm_radius: the merging radius of pass pss."""
    return float(lib().wr_vcm_radius(base_radius, pss))


HIT_DTYPE = np.dtype([("prim", np.uint32), ("t", np.float32), ("b1", np.float32), ("b2", np.float32)])
# wr_path_vertex: one Ray of a light path
PATH_VERTEX_DTYPE = np.dtype([("origin", np.float32, 3), ("direction", np.float32, 3),
                              ("radiance", np.float32, 3), ("origin_prob", np.float32),
                              ("direction_prob", np.float32), ("contact", np.uint32), ("contact_tri", np.uint32),
                              ("inside", np.uint32), ("delta", np.uint32), ("pad", np.uint32)])


class HashGrid:
    """wr_hashgrid: HashGrid::Build over device points (kept alive here) and Process as a range query."""

    def __init__(self, ctx, d_pos, n, radius, n_cells):
        self.ctx, self.n, self.n_cells = ctx, int(n), int(n_cells)
        self._pos = d_pos
        self.h = C.c_void_p()
        _check(lib().wr_hashgrid_build(ctx.h, d_pos.ptr, self.n, radius, self.n_cells, C.byref(self.h)), ctx.h,
               "wr_hashgrid_build")

    @classmethod
    def from_points(cls, ctx, pos, radius, n_cells):
        pos = np.ascontiguousarray(pos, dtype=np.float32).reshape(-1, 3)
        return cls(ctx, ctx.to_device(pos), len(pos), radius, n_cells)

    def export(self):
        bbox = np.zeros(6, np.float32)
        ends = np.zeros(self.n_cells, np.int32)
        idx = np.zeros(max(self.n, 1), np.int32)
        _check(lib().wr_hashgrid_export(self.h, _p(bbox), _p(ends), _p(idx)), self.ctx.h, "wr_hashgrid_export")
        return bbox, ends, idx[:self.n]

    def query_device(self, d_q, nq, cap, d_counts, d_idx):
        _check(lib().wr_hashgrid_query(self.h, d_q.ptr, nq, cap, d_counts.ptr, d_idx.ptr if d_idx else None),
               self.ctx.h, "wr_hashgrid_query")

    def query_device_noalloc(self, d_q, nq, cap, d_counts, d_idx):
        """wr_hashgrid_query_noalloc: asynchronous, no device allocation."""
        _check(lib().wr_hashgrid_query_noalloc(self.h, d_q.ptr, nq, cap, d_counts.ptr, d_idx.ptr if d_idx else None),
               self.ctx.h, "wr_hashgrid_query_noalloc")

    def query(self, q, cap):
        q = np.ascontiguousarray(q, dtype=np.float32).reshape(-1, 3)
        nq = len(q)
        dq = self.ctx.to_device(q)
        dc = self.ctx.alloc(4 * max(nq, 1))
        di = self.ctx.alloc(4 * max(nq * cap, 1))
        self.query_device(dq, nq, cap, dc, di)
        counts = dc.download(np.uint32, nq)
        idx = di.download(np.int32, nq * cap).reshape(nq, cap)
        return counts, idx

    def close(self):
        if self.h and self.h.value:
            lib().wr_hashgrid_destroy(self.h)
        self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Scene:
    """wr_scene: objects in world space, built with the reference builder and uploaded."""

    def __init__(self, ctx, objs=None, build=True):
        self.ctx = ctx
        self.h = C.c_void_p()
        _check(lib().wr_scene_create(ctx.h, C.byref(self.h)), ctx.h, "wr_scene_create")
        for o in objs or []:
            self.add(o)
        if objs is not None and build:
            self.build()

    def add(self, o):
        m = Material()
        m.type = o["mat"]
        p = list(o["params"])
        if o["mat"] == WR_REFRACTIVE:
            m.ior = p[0]
        else:
            m.color[:] = p[:3]
            if o["mat"] == WR_GLOSSY:
                m.power = p[3]
        v = np.ascontiguousarray(o["verts"], dtype=np.float32)
        t = np.ascontiguousarray(o["tris"], dtype=np.uint32)
        n = None if o["normals"] is None else np.ascontiguousarray(o["normals"], dtype=np.float32)
        oid = C.c_uint32()
        _check(lib().wr_scene_add_mesh(self.h, _p(v), len(v), _p(t), len(t), _p(n), C.byref(m), C.byref(oid)),
               self.ctx.h, "wr_scene_add_mesh")
        return oid.value

    def build(self):
        _check(lib().wr_scene_build(self.h, WR_BEST), self.ctx.h, "wr_scene_build")

    def info(self):
        i = SceneInfo()
        _check(lib().wr_scene_get_info(self.h, C.byref(i)), self.ctx.h, "wr_scene_get_info")
        return {k: getattr(i, k) for k, _ in SceneInfo._fields_}

    def prim_source(self, prim):
        o, t = C.c_uint32(), C.c_uint32()
        _check(lib().wr_scene_prim_source(self.h, prim, C.byref(o), C.byref(t)), self.ctx.h, "wr_scene_prim_source")
        return o.value, t.value

    def hit_sources(self, d_hits, n):
        """wr_hit_sources over n device hit records: (obj, tri) arrays on the host."""
        do, dt = self.ctx.alloc(4 * max(n, 1)), self.ctx.alloc(4 * max(n, 1))
        _check(lib().wr_hit_sources(self.ctx.h, self.h, d_hits.ptr, n, do.ptr, dt.ptr), self.ctx.h,
               "wr_hit_sources")
        return do.download(np.uint32, n), dt.download(np.uint32, n)

    def motion_pass(self, seed, pass_):
        """wr_scene_motion_pass: the scene as pass `pass_` of a render with `seed` places
        it (a static Scene of its own)."""
        h = C.c_void_p()
        _check(lib().wr_scene_motion_pass(self.h, seed, pass_, C.byref(h)), self.ctx.h, "wr_scene_motion_pass")
        sc = Scene.__new__(Scene)
        sc.ctx = self.ctx
        sc.h = h
        return sc

    def export_wide(self, obj=NONE):
        """wr_scene_export_wide: the device's fast-pass hierarchy of one tree."""
        L = lib()
        nw, root = C.c_uint32(), C.c_uint32()
        _check(L.wr_scene_export_wide(self.h, obj, C.byref(nw), C.byref(root), None), self.ctx.h,
               "wr_scene_export_wide")
        nodes = np.zeros(max(nw.value, 1), WNODE_DTYPE)
        _check(L.wr_scene_export_wide(self.h, obj, C.byref(nw), C.byref(root), _p(nodes)), self.ctx.h,
               "wr_scene_export_wide")
        return {"nodes": nodes[:nw.value], "root": root.value}

    def export_tree(self, obj=NONE):
        L = lib()
        n, m = C.c_uint32(), C.c_uint32()
        _check(L.wr_scene_export_tree(self.h, obj, C.byref(n), C.byref(m), None, None, None, None, None),
               self.ctx.h, "wr_scene_export_tree")
        box = np.zeros((n.value, 6), np.float32)
        right, lf, lc = (np.zeros(n.value, np.uint32) for _ in range(3))
        tl = np.zeros(max(m.value, 1), np.uint32)
        _check(L.wr_scene_export_tree(self.h, obj, C.byref(n), C.byref(m), _p(box), _p(right), _p(lf), _p(lc),
                                      _p(tl)), self.ctx.h, "wr_scene_export_tree")
        return {"box": box, "right": right, "leaf_first": lf, "leaf_count": lc, "tri_list": tl[:m.value]}

    # -- device-pointer entry points (inputs already resident) -------------
    def closest_hit_device(self, d_rays, n, d_hits, d_counters=None):
        _check(lib().wr_closest_hit(self.ctx.h, self.h, d_rays.ptr, n, d_hits.ptr,
                                    d_counters.ptr if d_counters is not None else None),
               self.ctx.h, "wr_closest_hit")

    def autotune_closest_hit(self, d_rays, n):
        """wr_autotune_closest_hit on resident rays: the chosen launch configuration and its time."""
        t = ChTuning()
        _check(lib().wr_autotune_closest_hit(self.ctx.h, self.h, d_rays.ptr, n, C.byref(t)), self.ctx.h,
               "wr_autotune_closest_hit")
        return {"refill_min": t.refill_min, "node_phase_min": t.node_phase_min, "blocks_per_cu": t.blocks_per_cu,
                "lds_levels": t.lds_levels, "batch": t.batch, "ms": t.ms}

    def set_closest_hit_tuning(self, refill_min, node_phase_min, blocks_per_cu=0, lds_levels=8, batch=256):
        t = ChTuning(refill_min, node_phase_min, blocks_per_cu, lds_levels, batch, 0.0)
        _check(lib().wr_scene_set_closest_hit_tuning(self.h, C.byref(t)), self.ctx.h,
               "wr_scene_set_closest_hit_tuning")

    # -- host convenience wrappers ------------------------------------------
    def closest_hit(self, rays, counters=False):
        rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 6)
        n = len(rays)
        dr = self.ctx.to_device(rays)
        dh = self.ctx.alloc(16 * max(n, 1))
        dc = None
        if counters:
            dc = self.ctx.to_device(np.zeros(3, np.uint64))
        self.closest_hit_device(dr, n, dh, dc)
        self.ctx.synchronize()
        hits = dh.download(np.uint32, 4 * n).view(HIT_DTYPE) if n else np.zeros(0, HIT_DTYPE)
        if counters:
            return hits, dc.download(np.uint64, 3)
        return hits

    def set_motion(self, obj, t0, t1, start, end):
        """wr_scene_set_motion (before build): object obj moves from `start` (at t0) to `end` (at t1)."""
        a = np.ascontiguousarray(start, dtype=np.float32).reshape(16)
        b = np.ascontiguousarray(end, dtype=np.float32).reshape(16)
        _check(lib().wr_scene_set_motion(self.h, obj, t0, t1, _p(a), _p(b)), self.ctx.h, "wr_scene_set_motion")

    def light_paths_device(self, max_depth, seed, pss, path0, n_paths, max_vertices):
        """wr_light_paths into new device buffers: (vertices, counts)."""
        dv = self.ctx.alloc(PATH_VERTEX_DTYPE.itemsize * max(n_paths * max_vertices, 1))
        dc = self.ctx.alloc(4 * max(n_paths, 1))
        _check(lib().wr_light_paths(self.ctx.h, self.h, max_depth, seed, pss, path0, n_paths, max_vertices, dv.ptr,
                                    dc.ptr), self.ctx.h, "wr_light_paths")
        return dv, dc

    def light_paths(self, max_depth, seed, pss, path0, n_paths, max_vertices):
        """Light sub-paths on the GPU: (records (n_paths, max_vertices) PATH_VERTEX_DTYPE, counts)."""
        dv, dc = self.light_paths_device(max_depth, seed, pss, path0, n_paths, max_vertices)
        rec = dv.download(PATH_VERTEX_DTYPE, n_paths * max_vertices).reshape(n_paths, max_vertices)
        return rec, dc.download(np.uint32, n_paths)

    def photons_device(self, dv, dc, n_paths, max_vertices, n_leading=0, cap=None):
        """wr_photons_collect: (device positions, device sources, n).  cap None: sized by a first call."""
        n = C.c_uint64()
        if cap is None:
            _check(lib().wr_photons_collect(self.ctx.h, self.h, dv.ptr, dc.ptr, n_paths, max_vertices, n_leading,
                                            None, None, 0, C.byref(n)), self.ctx.h, "wr_photons_collect")
            cap = n.value
        dp = self.ctx.alloc(12 * max(cap, 1))
        ds = self.ctx.alloc(8 * max(cap, 1))
        _check(lib().wr_photons_collect(self.ctx.h, self.h, dv.ptr, dc.ptr, n_paths, max_vertices, n_leading, dp.ptr,
                                        ds.ptr, cap, C.byref(n)), self.ctx.h, "wr_photons_collect")
        return dp, ds, n.value

    def closest_hit_noself(self, rays, inside, contact):
        rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 6)
        n = len(rays)
        pr = np.empty((n, 8), np.uint32)
        pr[:, :6] = rays.view(np.uint32)
        pr[:, 6] = inside
        pr[:, 7] = contact
        dr = self.ctx.to_device(pr)
        dh = self.ctx.alloc(16 * max(n, 1))
        _check(lib().wr_closest_hit_noself(self.ctx.h, self.h, dr.ptr, n, dh.ptr), self.ctx.h,
               "wr_closest_hit_noself")
        self.ctx.synchronize()
        return dh.download(np.uint32, 4 * n).view(HIT_DTYPE) if n else np.zeros(0, HIT_DTYPE)

    def inside_object(self, rays, cur=None):
        rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 6)
        n = len(rays)
        pr = np.empty((n, 8), np.uint32)
        pr[:, :6] = rays.view(np.uint32)
        pr[:, 6] = NONE
        pr[:, 7] = NONE if cur is None else cur
        dr = self.ctx.to_device(pr)
        do = self.ctx.alloc(4 * max(n, 1))
        _check(lib().wr_inside_object(self.ctx.h, self.h, dr.ptr, n, do.ptr), self.ctx.h, "wr_inside_object")
        self.ctx.synchronize()
        return do.download(np.uint32, n) if n else np.zeros(0, np.uint32)

    def close(self):
        if self.h and self.ctx.h:
            lib().wr_scene_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """wr_server: resident scenes, queued jobs, continuous batching in one wavefront pool."""

    def __init__(self, ctx, budget_bytes=0, path_slots=0, max_jobs=0):
        self.ctx = ctx
        self.h = C.c_void_p()
        cfg = ServerConfig(budget_bytes, path_slots, max_jobs)
        _check(lib().wr_server_create(ctx.h, C.byref(cfg), C.byref(self.h)), ctx.h, "wr_server_create")
        self._scenes = []

    def add_scene(self, scene):
        sid = C.c_uint32()
        _check(lib().wr_server_add_scene(self.h, scene.h, C.byref(sid)), self.ctx.h, "wr_server_add_scene")
        self._scenes.append(scene)
        return sid.value

    def submit(self, scene_id, cam, spp, max_depth, seed, nee=False):
        d = JobDescC(scene_id, CameraDesc.from_dict(cam), spp, max_depth, WR_RENDER_NEE if nee else 0, seed)
        jid = C.c_uint64()
        rc = lib().wr_job_submit(self.h, C.byref(d), C.byref(jid))
        if rc != 0 and jid.value == 0:
            _check(rc, self.ctx.h, "wr_job_submit")
        return jid.value

    def step(self):
        n = C.c_uint32()
        _check(lib().wr_server_step(self.h, C.byref(n)), self.ctx.h, "wr_server_step")
        return n.value

    def status(self, jid):
        s = JobStatus()
        _check(lib().wr_job_get_status(self.h, jid, C.byref(s)), self.ctx.h, "wr_job_get_status")
        return {k: getattr(s, k) for k, _ in JobStatus._fields_}

    def read(self, jid, width, height):
        n = 3 * width * height
        mean, s, c, q = (np.zeros(n, np.float32) for _ in range(4))
        passes = C.c_uint32()
        _check(lib().wr_job_read(self.h, jid, _p(mean), _p(s), _p(c), _p(q), C.byref(passes)), self.ctx.h,
               "wr_job_read")
        return {"mean": mean, "sum": s, "comp": c, "sumsq": q, "passes": passes.value}

    def release(self, jid):
        _check(lib().wr_job_release(self.h, jid), self.ctx.h, "wr_job_release")

    def cancel(self, jid):
        _check(lib().wr_job_cancel(self.h, jid), self.ctx.h, "wr_job_cancel")

    def inject_admit_failures(self, n):
        """Test hook: the next n admissions fail as if their device allocation had failed."""
        _check(lib().wr_server_inject_admit_failures(self.h, n), self.ctx.h, "wr_server_inject_admit_failures")

    def clock(self):
        return lib().wr_server_clock(self.h)

    def close(self):
        if self.h and self.ctx.h:
            lib().wr_server_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
