"""Closed-form scenes that pin the oracle's refraction, camera and Phong code
where the estimator is not at normal incidence, shared by the oracle tests and
the GPU tests.  Every expected value here is computed in float64 from the
physics or geometry of the case, independently of the oracle's code.

- Oblique glass slab (refractiveObj.h:39-81, util.h:40-54): Snell's law at both
  faces, the unpolarised Fresnel average er = (R_s + R_p) / 2, exit parallel to
  entry with lateral shift h sin(theta_i - theta_t) / cos(theta_t).
- Total internal reflection (refractiveObj.h:52-55): a deterministic mirror
  bounce whose path weight is the cosine term |N . w_r| (ray.cpp:52-58).
- Camera (camera.cpp:6-23,107-122,182-186): ray directions and the sample
  weight cos(theta_jittered) / cos(theta_pixel_centre).
- Isotropic Phong (isotropicPhong.h:20-38): the expected value of the
  reference estimator at oblique incidence as a quadrature of its BRDF.
"""
import math

import numpy as np

from walkray_amd import scenes as S

IOR = 1.5


def snell(theta_i, n1=1.0, n2=IOR):
    return math.asin(n1 / n2 * math.sin(theta_i))


def fresnel_avg(theta_i, n1=1.0, n2=IOR):
    """Unpolarised reflectance (R_s + R_p) / 2 from n1 into n2 at incidence theta_i."""
    st = n1 / n2 * math.sin(theta_i)
    if st >= 1.0:
        return 1.0
    ci, ct = math.cos(theta_i), math.sqrt(1.0 - st * st)
    rs = ((n1 * ci - n2 * ct) / (n1 * ci + n2 * ct)) ** 2
    rp = ((n2 * ci - n1 * ct) / (n2 * ci + n1 * ct)) ** 2
    return (rs + rp) / 2


# ------------------------------------------------------------ oblique slab ---
SLAB_TOP, SLAB_BOTTOM, LIGHT_Z = 0.5, -0.5, -1.5


def slab_scene(theta_deg, small_light=True, le=1.0, width=8):
    """Glass slab z in [-0.5, 0.5] (IOR 1.5); camera in the xz-plane looking at the
    origin of the top face at theta from the normal; emitter facing up at z = -1.5.
    small_light: a 0.3 x 0.3 emitter centred where the once-refracted ray lands
    (ray paths with internal reflections land 2 tan(theta_t) or more further on);
    otherwise an emitter covering the whole plane."""
    th = math.radians(theta_deg)
    tt = snell(th)
    slab = S.box((-50, -50, SLAB_BOTTOM), (50, 50, SLAB_TOP), S.REFRACTIVE, (IOR,))
    land = (SLAB_TOP - SLAB_BOTTOM) * math.tan(tt) + (SLAB_BOTTOM - LIGHT_Z) * math.tan(th)
    if small_light:
        light = S.quad((land, 0, LIGHT_Z), (0.3, 0, 0), (0, 0.3, 0), (0, 0, 1), S.EMISSIVE, (le, le, le))
    else:
        light = S.quad((0, 0, LIGHT_Z), (400, 0, 0), (0, 400, 0), (0, 0, 1), S.EMISSIVE, (le, le, le))
    dist = 3.0
    pos = (-dist * math.sin(th), 0.0, SLAB_TOP + dist * math.cos(th))
    cam = S.camera(pos, (0.0, 0.0, SLAB_TOP), (0.0, 1.0, 0.0), width, width, sight_dist=8000.0 * width)
    return [slab, light], cam


def slab_expected(theta_deg, small_light=True, le=1.0):
    """Small emitter: only the ray transmitted at both faces reaches it, so the pixel is
    Le (1 - er)^2 (each delta scatter has weight 1; the inner face sees theta_t, where
    the reflectance is the same er).  Whole-plane emitter: every path that leaves through
    the bottom counts, Le sum_k (1 - er)^2 er^(2k) = Le (1 - er) / (1 + er)."""
    er = fresnel_avg(math.radians(theta_deg))
    return le * ((1 - er) ** 2 if small_light else (1 - er) / (1 + er))


# ------------------------------------------------ total internal reflection ---
def tir_scene(theta_deg, le=1.0, width=8):
    """Camera inside a glass block (z in [0, 1], IOR 1.5) at height 0.5, looking up at the
    top face at theta from its normal; a 0.4 x 0.4 emitter inside the glass at z = 0.2,
    facing up, centred where the mirror ray from the top face lands."""
    th = math.radians(theta_deg)
    glass = S.box((-50, -50, 0.0), (50, 50, 1.0), S.REFRACTIVE, (IOR,))
    hit_x = 0.5 * math.tan(th)
    land = hit_x + 0.8 * math.tan(th)
    light = S.quad((land, 0, 0.2), (0.4, 0, 0), (0, 0.4, 0), (0, 0, 1), S.EMISSIVE, (le, le, le))
    cam = S.camera((0.0, 0.0, 0.5), (hit_x, 0.0, 1.0), (0.0, 1.0, 0.0), width, width, sight_dist=8000.0 * width)
    return [glass, light], cam


def tir_expected(theta_deg, le=1.0):
    """Beyond the critical angle asin(1/1.5) = 41.8 deg the reference reflects with radiance
    1 and probability 1, so the sample is the cosine term: Le |N . w_r| = Le cos(theta).
    Below it the mirror ray is chosen with probability er (from inside, glass to air)
    and has weight 1: Le er."""
    th = math.radians(theta_deg)
    if IOR * math.sin(th) > 1.0:
        return le * math.cos(th)
    return le * fresnel_avg(th, IOR, 1.0)


# ------------------------------------------------------------------ camera ---
def camera_dir64(cam, pixel, u, v):
    """camera.cpp:7-23 in float64."""
    pos, focus, up = (np.array(cam[k], np.float64) for k in ("pos", "focus", "up"))
    W, H, dist = cam["width"], cam["height"], cam["sight_dist"]
    front = (focus - pos) / np.linalg.norm(focus - pos)
    right = np.cross(front, up)
    right /= np.linalg.norm(right)
    top = np.cross(right, front)
    x, y = pixel % W, pixel // W
    d = front * dist + right * (x + u - W / 2) + top * (H / 2 - y - v)
    return d / np.linalg.norm(d), front


def camera_weight64(cam, pixel, u, v):
    """|front . d(u, v)| / |front . d(0.5, 0.5)| (camera.cpp:107-122,182-186, ray.cpp:52-58)."""
    d, front = camera_dir64(cam, pixel, u, v)
    c, _ = camera_dir64(cam, pixel, 0.5, 0.5)
    return abs(front @ d) / abs(front @ c)


def wide_camera_scene(width=16, height=12, le=1.0):
    """A 90-degree camera (sight_dist = width / 2) facing a plane emitter that fills the view."""
    light = S.quad((0, 0, -1.0), (4000, 0, 0), (0, 4000, 0), (0, 0, 1), S.EMISSIVE, (le, le, le))
    cam = S.camera((0.0, 0.0, 0.0), (0.0, 0.0, -1.0), (0.0, 1.0, 0.0), width, height, sight_dist=width / 2)
    return [light], cam


def wide_camera_expected(cam, le=1.0, m=48):
    """Per pixel, Le E_(u,v)[cos theta_jit] / cos theta_centre: midpoint quadrature over the
    pixel square in float64."""
    W, H, dist = cam["width"], cam["height"], cam["sight_dist"]
    g = (np.arange(m) + 0.5) / m
    u, v = np.meshgrid(g, g, indexing="ij")
    out = np.empty(W * H)
    for p in range(W * H):
        x, y = p % W, p // W
        a, b = x + u - W / 2, H / 2 - y - v
        cj = dist / np.sqrt(dist * dist + a * a + b * b)
        ac, bc = x + 0.5 - W / 2, H / 2 - y - 0.5
        cc = dist / math.sqrt(dist * dist + ac * ac + bc * bc)
        out[p] = le * cj.mean() / cc
    return out


# ------------------------------------------------------------------- Phong ---
def phong_scene(theta_deg, power, albedo=0.8, le=1.0, width=8):
    """Phong floor (y = 0, normal +y) under an emitter plane at y = 1 facing down; camera
    in the xy-plane looking at the origin at theta from the normal."""
    th = math.radians(theta_deg)
    floor = S.quad((0, 0, 0), (4000, 0, 0), (0, 0, 4000), (0, 1, 0), S.GLOSSY, (albedo, albedo, albedo, power))
    light = S.quad((0, 1.0, 0), (4000, 0, 0), (0, 0, 4000), (0, -1, 0), S.EMISSIVE, (le, le, le))
    dist = 0.5
    pos = (-dist * math.sin(th), dist * math.cos(th), 0.0)
    cam = S.camera(pos, (0.0, 0.0, 0.0), (0.0, 0.0, 1.0), width, width, sight_dist=8000.0 * width)
    return [floor, light], cam


def phong_integral(theta_deg, power, n_theta=1500, n_phi=3000):
    """I = int_{upper hemisphere} (n+2)/(2 pi) max(R . w, 0)^n (N . w) dw for the mirror
    direction R of an incident ray at theta from the normal (midpoint rule in (theta, phi)).
    The pixel of phong_scene at depth 1 is Le * albedo * I: Russian roulette divides
    by the survival probability it takes, and the sampling pdf cancels in expectation."""
    th = math.radians(theta_deg)
    R = np.array([math.sin(th), math.cos(th), 0.0])  # mirror of (sin th, -cos th, 0) about +y
    t = (np.arange(n_theta) + 0.5) * (math.pi / 2) / n_theta
    p = (np.arange(n_phi) + 0.5) * (2 * math.pi) / n_phi
    T, P = np.meshgrid(t, p, indexing="ij")
    w = np.stack([np.sin(T) * np.cos(P), np.cos(T), np.sin(T) * np.sin(P)], -1)
    rw = np.maximum(w @ R, 0.0)
    f = (power + 2) / (2 * math.pi) * rw ** power * np.cos(T) * np.sin(T)
    return f.sum() * (math.pi / 2 / n_theta) * (2 * math.pi / n_phi)
