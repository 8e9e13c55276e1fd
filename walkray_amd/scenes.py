"""Procedural, seeded scene and ray generators.

The reference loads its scenes from XML/OBJ files (config.cpp:21-429,
mesh.cpp:5-130); none ship with it, so every workload here is generated
deterministically in world space (identity object transforms, mesh.h:106).

An object is a dict with
    verts   (nv, 3) float32   world-space vertex positions
    tris    (nt, 3) uint32    vertex indices, winding gives the face normal
                              (v1 - v0) x (v2 - v0)  (mesh.h:52-55)
    normals (nv, 3) float32 or None   per-vertex shading normals
    mat     one of DIFFUSE, GLOSSY, REFRACTIVE, EMISSIVE
    params  material parameters: diffuse (r, g, b); glossy (r, g, b, power);
            refractive (ior,); emissive (r, g, b)
"""
from __future__ import annotations

import math

import numpy as np

DIFFUSE, GLOSSY, REFRACTIVE, EMISSIVE = 0, 1, 2, 3
NONE = 0xFFFFFFFF


def _obj(verts, tris, normals, mat, params):
    return {
        "verts": np.ascontiguousarray(verts, dtype=np.float32),
        "tris": np.ascontiguousarray(tris, dtype=np.uint32),
        "normals": None if normals is None else np.ascontiguousarray(normals, dtype=np.float32),
        "mat": int(mat),
        "params": tuple(float(p) for p in params),
    }


# ---------------------------------------------------------------- hashing ---
def _hash_u32(*keys):
    """Integer hash (murmur3 finaliser over a combined key), vectorised, wraps mod 2^32."""
    h = np.uint32(0x9E3779B9)
    with np.errstate(over="ignore"):
        for k in keys:
            k = np.asarray(k).astype(np.uint32)
            h = (h ^ k) * np.uint32(0x85EBCA6B)
            h = h ^ (h >> np.uint32(13))
            h = h * np.uint32(0xC2B2AE35)
            h = h ^ (h >> np.uint32(16))
    return h


def _value_noise(x, y, seed, octave):
    """Lattice value noise in [-1, 1) with smoothstep interpolation."""
    xi = np.floor(x)
    yi = np.floor(y)
    fx = x - xi
    fy = y - yi
    xi = xi.astype(np.int64) & 0xFFFFFFFF
    yi = yi.astype(np.int64) & 0xFFFFFFFF

    def lat(ix, iy):
        return _hash_u32(ix, iy, seed, octave).astype(np.float64) / 2147483648.0 - 1.0

    sx = fx * fx * (3.0 - 2.0 * fx)
    sy = fy * fy * (3.0 - 2.0 * fy)
    v00 = lat(xi, yi)
    v10 = lat(xi + 1, yi)
    v01 = lat(xi, yi + 1)
    v11 = lat(xi + 1, yi + 1)
    a = v00 + sx * (v10 - v00)
    b = v01 + sx * (v11 - v01)
    return a + sy * (b - a)


def fbm(x, y, seed, octaves=6, lacunarity=2.0, gain=0.5, base_freq=4.0):
    total = np.zeros_like(x, dtype=np.float64)
    amp = 1.0
    freq = base_freq
    for o in range(octaves):
        total += amp * _value_noise(x * freq, y * freq, seed, o)
        amp *= gain
        freq *= lacunarity
    return total


# ------------------------------------------------------------ primitives ---
def heightfield(n, seed=1, amplitude=0.2, octaves=6, mat=DIFFUSE, params=(0.7, 0.7, 0.7),
                origin=(0.0, 0.0, 0.0), size=1.0):
    """n x n vertex grid on [0,size]^2 (x, y), z = amplitude * fBm; 2 (n-1)^2 triangles, +z facing."""
    g = np.arange(n, dtype=np.float64) / (n - 1)
    xx, yy = np.meshgrid(g, g)  # yy varies along rows (j), xx along columns (i)
    z = amplitude * fbm(xx, yy, seed, octaves)
    verts = np.stack([origin[0] + size * xx, origin[1] + size * yy, origin[2] + z], axis=-1).reshape(-1, 3)
    i = np.arange(n - 1, dtype=np.uint32)
    ii, jj = np.meshgrid(i, i)
    v00 = (jj * n + ii).ravel()
    v10 = v00 + 1
    v01 = v00 + n
    v11 = v01 + 1
    t1 = np.stack([v00, v10, v11], axis=-1)
    t2 = np.stack([v00, v11, v01], axis=-1)
    tris = np.stack([t1, t2], axis=1).reshape(-1, 3)
    return _obj(verts, tris, None, mat, params)


def icosphere(k, center=(0.0, 0.0, 0.0), radius=1.0, mat=DIFFUSE, params=(0.8, 0.8, 0.8),
              smooth=True, inward=False):
    """Icosahedron subdivided k times: 20*4^k triangles, 10*4^k+2 vertices, outward winding."""
    unit, tris = _unit_icosphere(k, inward)
    pos = np.asarray(center, dtype=np.float64) + radius * unit
    normals = (-unit if inward else unit) if smooth else None
    return _obj(pos, tris.copy(), normals, mat, params)


_ICO_CACHE = {}


def _unit_icosphere(k, inward):
    key = (k, inward)
    if key in _ICO_CACHE:
        return _ICO_CACHE[key]
    t = (1.0 + math.sqrt(5.0)) / 2.0
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t),
         (0, 1, -t), (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    verts = [np.array(p, dtype=np.float64) / np.linalg.norm(p) for p in v]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4),
             (11, 10, 2), (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8),
             (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(k):
        cache = {}
        nf = []

        def mid(a, b):
            key = (a, b) if a < b else (b, a)
            if key not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]

        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    unit = np.array(verts)
    tris = np.array(faces, dtype=np.uint32)
    # make the winding outward: (v1-v0)x(v2-v0) . centroid > 0
    p0, p1, p2 = unit[tris[:, 0]], unit[tris[:, 1]], unit[tris[:, 2]]
    out = np.einsum("ij,ij->i", np.cross(p1 - p0, p2 - p0), p0 + p1 + p2) > 0
    flip = out == inward
    tris[flip] = tris[flip][:, [0, 2, 1]]
    _ICO_CACHE[key] = (unit, tris)
    return unit, tris


def box(lo, hi, mat=DIFFUSE, params=(0.7, 0.7, 0.7), inward=False):
    """Axis-aligned box, 12 triangles, flat shading, outward (or inward) winding."""
    lo = np.asarray(lo, dtype=np.float64)
    hi = np.asarray(hi, dtype=np.float64)
    c = [np.array([hi[0] if i & 1 else lo[0], hi[1] if i & 2 else lo[1], hi[2] if i & 4 else lo[2]])
         for i in range(8)]
    quads = [(0, 2, 6, 4), (1, 5, 7, 3), (0, 4, 5, 1), (2, 3, 7, 6), (0, 1, 3, 2), (4, 6, 7, 5)]
    tris = []
    center = (lo + hi) / 2
    for a, b, cc, d in quads:
        for tri in ((a, b, cc), (a, cc, d)):
            p0, p1, p2 = c[tri[0]], c[tri[1]], c[tri[2]]
            n = np.cross(p1 - p0, p2 - p0)
            outward = np.dot(n, (p0 + p1 + p2) / 3 - center) > 0
            tris.append(tri if outward != inward else (tri[0], tri[2], tri[1]))
    return _obj(np.array(c), np.array(tris, dtype=np.uint32), None, mat, params)


def quad(center, e1, e2, normal, mat, params):
    """Rectangle center +- e1/2 +- e2/2, two triangles whose face normal points along `normal`."""
    c = np.asarray(center, dtype=np.float64)
    e1 = np.asarray(e1, dtype=np.float64) / 2
    e2 = np.asarray(e2, dtype=np.float64) / 2
    p = [c - e1 - e2, c + e1 - e2, c + e1 + e2, c - e1 + e2]
    tris = [(0, 1, 2), (0, 2, 3)]
    if np.dot(np.cross(p[1] - p[0], p[2] - p[0]), normal) < 0:
        tris = [(0, 2, 1), (0, 3, 2)]
    return _obj(np.array(p), np.array(tris, dtype=np.uint32), None, mat, params)


def merge(objs, mat, params):
    """Concatenate flat-shaded meshes into one object."""
    vs, ts, off = [], [], 0
    for o in objs:
        vs.append(o["verts"])
        ts.append(o["tris"].astype(np.int64) + off)
        off += len(o["verts"])
    return _obj(np.concatenate(vs), np.concatenate(ts), None, mat, params)


# ---------------------------------------------------------------- cameras ---
def camera(pos, focus, up, width, height, sight_dist):
    return {"pos": tuple(map(float, pos)), "focus": tuple(map(float, focus)),
            "up": tuple(map(float, up)), "width": int(width), "height": int(height),
            "sight_dist": float(sight_dist)}


# ------------------------------------------------------------------ scenes ---
def cornell(k=5, width=256, height=256):
    """Cornell box [0,1]^3, 12 wall triangles (inward), two k-icospheres, a 0.25^2 light of 17."""
    white, red, green = (0.75, 0.75, 0.75), (0.63, 0.06, 0.04), (0.15, 0.45, 0.09)
    floor = quad((0.5, 0.0, 0.5), (1, 0, 0), (0, 0, 1), (0, 1, 0), DIFFUSE, white)
    ceil = quad((0.5, 1.0, 0.5), (1, 0, 0), (0, 0, 1), (0, -1, 0), DIFFUSE, white)
    back = quad((0.5, 0.5, 0.0), (1, 0, 0), (0, 1, 0), (0, 0, 1), DIFFUSE, white)
    front = quad((0.5, 0.5, 1.0), (1, 0, 0), (0, 1, 0), (0, 0, -1), DIFFUSE, white)
    objs = [
        merge([floor, ceil, back, front], DIFFUSE, white),
        quad((0.0, 0.5, 0.5), (0, 1, 0), (0, 0, 1), (1, 0, 0), DIFFUSE, red),
        quad((1.0, 0.5, 0.5), (0, 1, 0), (0, 0, 1), (-1, 0, 0), DIFFUSE, green),
        quad((0.5, 1.0 - 1e-3, 0.5), (0.25, 0, 0), (0, 0, 0.25), (0, -1, 0), EMISSIVE, (17.0, 17.0, 17.0)),
        icosphere(k, (0.28, 0.201, 0.35), 0.2, DIFFUSE, (0.8, 0.8, 0.8)),
        icosphere(k, (0.72, 0.201, 0.62), 0.2, REFRACTIVE, (1.5,)),
    ]
    cam = camera((0.5, 0.5, 0.97), (0.5, 0.5, 0.0), (0.0, 1.0, 0.0), width, height,
                 sight_dist=250.0 * width / 256.0)
    return objs, cam


def random_field(n_objects=512, seed=3, width=1920, height=1080, sphere_k=3, sphere_frac=0.75):
    """Wavefront workload: seeded icospheres (k=3, 1,280 triangles) and boxes on a ground plane.

    Objects sit on a 32 x 16 jittered grid of unit cells (no overlaps); materials are drawn
    uniformly from diffuse (albedo U[0.1, 0.9] per channel), isotropic Phong (albedo
    U[0.1, 0.9], exponent log-uniform in [10, 1000]) and refractive (IOR U[1.33, 1.7]);
    two 6 x 6 area lights of radiance 10 face down from y = 6.
    """
    rng = np.random.default_rng(seed)
    objs = [quad((0.0, 0.0, 0.0), (40.0, 0, 0), (0, 0, 40.0), (0, 1, 0), DIFFUSE, (0.5, 0.5, 0.5))]
    cols, rows = 32, 16
    cells = rng.permutation(cols * rows)[:n_objects]
    for c in cells:
        cx = -cols / 2 + (c % cols) + 0.5
        cz = -rows / 2 + (c // cols) + 0.5
        kind = rng.integers(3)
        if kind == 0:
            mat, params = DIFFUSE, tuple(rng.uniform(0.1, 0.9, 3))
        elif kind == 1:
            mat, params = GLOSSY, tuple(rng.uniform(0.1, 0.9, 3)) + (float(np.exp(rng.uniform(np.log(10), np.log(1000)))),)
        else:
            mat, params = REFRACTIVE, (float(rng.uniform(1.33, 1.7)),)
        jx, jz = rng.uniform(-0.05, 0.05, 2)
        if rng.random() < sphere_frac:
            r = float(rng.uniform(0.2, 0.42))
            objs.append(icosphere(sphere_k, (cx + jx, r + 1e-3, cz + jz), r, mat, params))
        else:
            h = rng.uniform(0.15, 0.4, 3)
            lo = (cx + jx - h[0], 1e-3, cz + jz - h[2])
            hi = (cx + jx + h[0], 1e-3 + 2 * h[1], cz + jz + h[2])
            objs.append(box(lo, hi, mat, params))
    for x in (-6.0, 6.0):
        objs.append(quad((x, 6.0, 0.0), (6.0, 0, 0), (0, 0, 6.0), (0, -1, 0), EMISSIVE, (10.0, 10.0, 10.0)))
    cam = camera((0.0, 5.0, 18.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), width, height,
                 sight_dist=960.0 / math.tan(math.radians(30.0)) * width / 1920.0)
    return objs, cam


def disk(center, normal, radius, segments, mat, params):
    """A flat disk as a triangle fan, front face along normal."""
    nrm = np.asarray(normal, np.float64)
    nrm = nrm / np.linalg.norm(nrm)
    a = np.cross(nrm, (0.0, 0.0, 1.0) if abs(nrm[2]) < 0.9 else (1.0, 0.0, 0.0))
    a /= np.linalg.norm(a)
    b = np.cross(nrm, a)
    ang = 2 * math.pi * np.arange(segments) / segments
    ring = np.asarray(center) + radius * (np.cos(ang)[:, None] * a + np.sin(ang)[:, None] * b)
    verts = np.concatenate([np.asarray(center, np.float64)[None], ring])
    i = np.arange(segments, dtype=np.uint32)
    tris = np.stack([np.zeros(segments, np.uint32), 1 + i, 1 + (i + 1) % segments], axis=-1)
    # (v1-v0)x(v2-v0) must point along normal
    if np.dot(np.cross(verts[1] - verts[0], verts[2] - verts[0]), nrm) < 0:
        tris = tris[:, [0, 2, 1]]
    return _obj(verts, tris, None, mat, params)


def large_scene(n=4096, n_spheres=2000, width=3840, height=2160, seed=5):
    """Large scene: an n x n fBm heightfield on the unit square (2 (n-1)^2 triangles; 33,538,050
    at n = 4096), n_spheres Phong icospheres k=4 (5,120 triangles each, albedo U[0.1, 0.9],
    exponent log-uniform 10-1000) resting on it, a sun-like emissive disk 20 units away
    (0.5 degree radius) and an emissive sky dome (an inward icosphere k=3 of radius 30)."""
    rng = np.random.default_rng(seed)
    objs = [heightfield(n, seed=seed, amplitude=0.2, params=(0.5, 0.45, 0.4))]
    xy = rng.uniform(0.05, 0.95, (n_spheres, 2))
    r = rng.uniform(0.003, 0.01, n_spheres)
    z = 0.2 * fbm(xy[:, 0], xy[:, 1], seed, 6) + r
    alb = rng.uniform(0.1, 0.9, (n_spheres, 3))
    power = np.exp(rng.uniform(np.log(10.0), np.log(1000.0), n_spheres))
    for k in range(n_spheres):
        objs.append(icosphere(4, (xy[k, 0], xy[k, 1], z[k]), r[k], GLOSSY, tuple(alb[k]) + (power[k],)))
    sun_dir = np.array([math.cos(math.radians(40)) * math.cos(math.radians(30)),
                        math.cos(math.radians(40)) * math.sin(math.radians(30)), math.sin(math.radians(40))])
    centre = np.array([0.5, 0.5, 0.0])
    objs.append(disk(centre + 20.0 * sun_dir, -sun_dir, 20.0 * math.tan(math.radians(0.5)), 32, EMISSIVE,
                     (20000.0, 18000.0, 15000.0)))
    objs.append(icosphere(3, tuple(centre), 30.0, EMISSIVE, (0.4, 0.55, 0.8), smooth=False, inward=True))
    cam = camera((0.5, -0.25, 0.45), (0.5, 0.55, 0.05), (0.0, 0.0, 1.0), width, height,
                 sight_dist=(width / 2) / math.tan(math.radians(30.0)))
    return objs, cam


def trs(t, axis, ang, scale):
    """4x4 row-major transform: translation t, then a rotation by ang about axis, then a
    per-axis scale (applied to a point as T R S p)."""
    a = np.asarray(axis, np.float64)
    x, y, z = a / np.linalg.norm(a)
    c, s = math.cos(ang), math.sin(ang)
    M = np.eye(4)
    M[:3, :3] = np.array([[c + x * x * (1 - c), x * y * (1 - c) - z * s, x * z * (1 - c) + y * s],
                          [y * x * (1 - c) + z * s, c + y * y * (1 - c), y * z * (1 - c) - x * s],
                          [z * x * (1 - c) - y * s, z * y * (1 - c) + x * s, c + z * z * (1 - c)]]) @ np.diag(scale)
    M[:3, 3] = t
    return M.astype(np.float32)


def cornell_moving(k=5, width=256, height=256):
    """The Cornell box with its diffuse sphere (object 4) moving during the exposure:
    translated, turned about y and shrunk between its transforms at times 0 and 1.
    Returns objects (the sphere in its own frame), camera and {object: (t0, t1, start, end)}."""
    objs, cam = cornell(k=k, width=width, height=height)
    c = np.array([0.28, 0.201, 0.35], np.float32)
    objs[4] = dict(objs[4])
    objs[4]["verts"] = (objs[4]["verts"] - c).astype(np.float32)
    a = trs(c, [0, 1, 0], 0.0, [1.0, 1.0, 1.0])
    b = trs(c + np.float32([0.25, 0.1, 0.2]), [0, 1, 0], 1.2, [0.7, 0.9, 0.8])
    return objs, cam, {4: (0.0, 1.0, a, b)}


def heightfield_scene(n=1024, seed=7):
    """Ray-query workload: one fBm heightfield of 2 (n-1)^2 triangles."""
    return [heightfield(n, seed=seed)]


def scene_bounds(objs):
    v = np.concatenate([o["verts"] for o in objs])
    return v.min(axis=0), v.max(axis=0)


# -------------------------------------------------------------------- rays ---
def pinhole_rays(cam, width=None, height=None):
    """Camera rays through pixel centres (camera.cpp:6-23 with u = v = 0.5), float32 (N, 6)."""
    W = width or cam["width"]
    H = height or cam["height"]
    pos = np.array(cam["pos"], dtype=np.float64)
    front = np.array(cam["focus"], dtype=np.float64) - pos
    front /= np.linalg.norm(front)
    right = np.cross(front, np.array(cam["up"], dtype=np.float64))
    right /= np.linalg.norm(right)
    top = np.cross(right, front)
    x = np.arange(W, dtype=np.float64) + 0.5 - W / 2.0
    y = H / 2.0 - np.arange(H, dtype=np.float64) - 0.5
    sd = cam["sight_dist"] * W / cam["width"]
    d = (front[None, None, :] * sd + right[None, None, :] * x[None, :, None]
         + top[None, None, :] * y[:, None, None])
    d /= np.linalg.norm(d, axis=-1, keepdims=True)
    rays = np.empty((H * W, 6), dtype=np.float32)
    rays[:, :3] = pos.astype(np.float32)
    rays[:, 3:] = d.reshape(-1, 3).astype(np.float32)
    return rays


def heightfield_camera(width=4096, height=4096):
    return camera((0.5, -0.35, 0.9), (0.5, 0.45, 0.0), (0.0, 0.0, 1.0), width, height,
                  sight_dist=1.15 * width)


def incoherent_rays(n, lo, hi, seed=11):
    """Origins uniform in the box [lo, hi], directions uniform on the sphere (z = 1-2u, phi = 2 pi v)."""
    idx = np.arange(n, dtype=np.uint64)
    lo = np.asarray(lo, dtype=np.float64)
    hi = np.asarray(hi, dtype=np.float64)

    def u(k):
        return (_hash_u32(idx & np.uint64(0xFFFFFFFF), idx >> np.uint64(32), seed, k) >> np.uint32(8)).astype(
            np.float64) / 16777216.0

    o = lo + (hi - lo) * np.stack([u(0), u(1), u(2)], axis=-1)
    z = 1.0 - 2.0 * u(3)
    phi = 2.0 * math.pi * u(4)
    r = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    d = np.stack([r * np.cos(phi), r * np.sin(phi), z], axis=-1)
    rays = np.empty((n, 6), dtype=np.float32)
    rays[:, :3] = o
    rays[:, 3:] = d
    return rays


# ---- reference scene files ----------------------------------------------------
_MAT_TYPES = {DIFFUSE: "Diffuse", GLOSSY: "Glossy", REFRACTIVE: "Refractive", EMISSIVE: "Emissive"}


def _unitize_frame(verts):
    """Centre and half extent the reference loader computes (mesh.cpp:152-165), in float32."""
    v = np.asarray(verts, np.float32)
    lo, hi = v.min(axis=0), v.max(axis=0)
    c = (lo + hi) / np.float32(2)
    d = hi - lo
    m = max(d[0], d[1])
    m = max(m, d[2])
    return c, np.float32(m) / np.float32(2)


def write_reference_scene(directory, objs, cam, spp=16, name="Scene"):
    """Writes objs and cam as a reference scene configuration that Config::load
    (config.cpp:432-575) reads: config.xml, scene.xml, materials.xml and one
    .obj per object.  The loader unitizes every shape; each object's transform
    scales and translates it back, so the loaded vertices equal the originals
    up to float rounding.  Returns the path of config.xml."""
    import os
    os.makedirs(directory, exist_ok=True)
    f9 = "{:.9g}".format
    mats, objx = [], []
    for i, o in enumerate(objs):
        v = np.asarray(o["verts"], np.float32)
        t = np.asarray(o["tris"], np.int64) + 1
        lines = [f"v {f9(a)} {f9(b)} {f9(c)}" for a, b, c in v]
        if o["normals"] is not None:
            lines += [f"vn {f9(a)} {f9(b)} {f9(c)}" for a, b, c in np.asarray(o["normals"], np.float32)]
            lines += [f"f {a}//{a} {b}//{b} {c}//{c}" for a, b, c in t]
        else:
            lines += [f"f {a} {b} {c}" for a, b, c in t]
        with open(os.path.join(directory, f"object{i}.obj"), "w") as fh:
            fh.write("\n".join(lines) + "\n")
        p = o["params"]
        body = f"<type>{_MAT_TYPES[o['mat']]}</type>"
        if o["mat"] == REFRACTIVE:
            body += f"<IOR>{f9(p[0])}</IOR>"
        else:
            body += f"<color>({f9(p[0])}, {f9(p[1])}, {f9(p[2])})</color>"
            if o["mat"] == GLOSSY:
                body += f"<power>{f9(p[3])}</power>"
        mats.append(f'<Material Name="m{i}">{body}</Material>')
        c, s = _unitize_frame(v)
        tr = (f"[{f9(s)}, 0, 0, {f9(c[0])}][0, {f9(s)}, 0, {f9(c[1])}]"
              f"[0, 0, {f9(s)}, {f9(c[2])}][0, 0, 0, 1]")
        objx.append(f"<Object><filePath>object{i}.obj</filePath>"
                    f"<Material><path>materials.xml</path><name>m{i}</name></Material>"
                    f"<transform>{tr}</transform></Object>")
    vec = lambda x: f"({f9(x[0])}, {f9(x[1])}, {f9(x[2])})"  # noqa: E731
    camx = (f'<Camera Name="main"><position>{vec(cam["pos"])}</position><focus>{vec(cam["focus"])}</focus>'
            f'<up>{vec(cam["up"])}</up><width>{cam["width"]}</width><height>{cam["height"]}</height>'
            f'<sightDist>{f9(cam["sight_dist"])}</sightDist></Camera>')
    with open(os.path.join(directory, "materials.xml"), "w") as fh:
        fh.write("\n".join(mats) + "\n")
    with open(os.path.join(directory, "scene.xml"), "w") as fh:
        fh.write(f'<Scene Name="{name}">\n' + "\n".join(objx) + "\n" + camx + "\n</Scene>\n")
    path = os.path.join(directory, "config.xml")
    with open(path, "w") as fh:
        fh.write(f"<Config>\n<Scene><path>scene.xml</path><name>{name}</name><camera>main</camera></Scene>\n"
                 f"<renderer>PathTracer</renderer><iter>{spp}</iter>\n</Config>\n")
    return path
