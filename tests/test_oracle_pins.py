"""Pins of the oracle where its estimator leaves normal incidence: oblique
refraction and total internal reflection (refractiveObj.h:27-88, util.h:40-54),
the camera's ray mapping and sample weight (camera.cpp:6-23,107-122,182-186,
ray.cpp:52-58), isotropic Phong off the normal (isotropicPhong.h:20-38,
glossyObj.h:39-62), BidirectionalPathTracer with a glass sphere
(bidirPathTracer.cpp:53-221), VCMTracer (vcmTracer.cpp:6-446) and PhotonMap's
progressive photon mapping (photonMap.cpp:21-248,656-674).  Expected values come from tests/pin_cases.py
(float64 physics and geometry), never from the oracle's own formulas."""
import math

import numpy as np
import pytest

from oracle import oracle as O
from walkray_amd import scenes as S

import pin_cases as P


def _pixel_mean(sc, cam, spp, depth, seed=5):
    s, q = sc.render(cam, spp, depth, seed)
    n = cam["width"] * cam["height"] * spp
    m = s.reshape(-1, 3).sum(0) / n
    var = q.reshape(-1, 3).sum(0) / n - m * m
    return m, np.sqrt(np.maximum(var, 0) / n)


def _first_hit(sc, ray):
    h = sc.intersect(np.asarray([ray], np.float32))
    assert h["prim"][0] != O.NONE
    return int(h["obj"][0]), int(h["tri"][0]), float(h["t"][0])


# ---------------------------------------------------------- oblique glass ---
@pytest.mark.parametrize("theta", [30.0, 60.0])
def test_oblique_slab_scatter_snell_and_fresnel(theta):
    """Both faces of the slab, scattered directly: the reflect probability is the
    unpolarised Fresnel average at theta_i on the way in and at theta_t on the way out
    (equal), the transmitted direction obeys Snell's law in the plane of incidence, the
    exit ray is parallel to the entry ray, its offset is h sin(ti - tt) / cos(tt), and
    the mirror ray is d - 2 (N.d) N.  A flipped rotation axis, a dropped pi - asin
    branch, or swapped indices in R_s / R_p fail one of these."""
    th, tt = math.radians(theta), P.snell(math.radians(theta))
    er = P.fresnel_avg(th)
    objs, cam = P.slab_scene(theta, small_light=False)
    sc = O.Scene(objs)
    o = np.array(cam["pos"], np.float32)
    d = np.array([math.sin(th), 0.0, -math.cos(th)], np.float32)
    obj, tri, t = _first_hit(sc, np.concatenate([o, d]))
    assert obj == 0
    # entering: transmit (rn0 >= er) and reflect (rn0 < er)
    tr = sc.scatter(0, np.concatenate([o, d]), obj, tri, t, O.NONE, (0.999, 0.3, 0.7))
    rf = sc.scatter(0, np.concatenate([o, d]), obj, tri, t, O.NONE, (0.0, 0.3, 0.7))
    np.testing.assert_allclose(tr["direction"], [math.sin(tt), 0, -math.cos(tt)], atol=2e-6)
    np.testing.assert_allclose(rf["direction"], [math.sin(th), 0, math.cos(th)], atol=2e-6)
    assert abs(tr["direction_prob"] - (1 - er)) < 2e-6 and abs(rf["direction_prob"] - er) < 2e-6
    # weight of the chosen branch: radiance * cosine term / directionProb = 1 (delta scatter)
    np.testing.assert_allclose(tr["radiance"] * math.cos(tt) / tr["direction_prob"], 1, rtol=2e-5)
    np.testing.assert_allclose(rf["radiance"] * math.cos(th) / rf["direction_prob"], 1, rtol=2e-5)
    assert tr["delta"] and rf["delta"] and tr["inside"] == 0 and rf["inside"] == O.NONE
    # leaving through the bottom face from inside
    p_in = tr["origin"]
    h = sc.intersect(np.concatenate([p_in, tr["direction"]])[None].astype(np.float32),
                     inside=np.array([0], np.uint32), contact=np.array([0], np.uint32), noself=True)
    assert h["obj"][0] == 0
    out = sc.scatter(0, np.concatenate([p_in, tr["direction"]]), 0, int(h["tri"][0]), float(h["t"][0]), 0,
                     (0.999, 0.3, 0.7))
    back = sc.scatter(0, np.concatenate([p_in, tr["direction"]]), 0, int(h["tri"][0]), float(h["t"][0]), 0,
                      (0.0, 0.3, 0.7))
    np.testing.assert_allclose(out["direction"], d, atol=4e-6)  # parallel to the entry ray
    assert abs(out["direction_prob"] - (1 - er)) < 4e-6 and out["inside"] == O.NONE
    np.testing.assert_allclose(back["direction"], [math.sin(tt), 0, math.cos(tt)], atol=4e-6)
    assert back["inside"] == 0
    q = out["origin"].astype(np.float64)
    np.testing.assert_allclose(q[2], P.SLAB_BOTTOM, atol=1e-6)
    # distance of the exit point from the entry line
    e0, dd = p_in.astype(np.float64), d.astype(np.float64)
    off = np.linalg.norm(np.cross(q - e0, dd))
    want = (P.SLAB_TOP - P.SLAB_BOTTOM) * math.sin(th - tt) / math.cos(tt)
    assert abs(off - want) < 1e-5, (off, want)


@pytest.mark.parametrize("theta", [30.0, 60.0])
@pytest.mark.parametrize("small", [True, False])
def test_oblique_slab_render_closed_form(theta, small):
    """The whole PathTracer estimator through the slab: Le (1 - er)^2 on a small emitter
    under the once-refracted ray, Le (1 - er) / (1 + er) on an emitter plane."""
    objs, cam = P.slab_scene(theta, small_light=small)
    sc = O.Scene(objs)
    m, se = _pixel_mean(sc, cam, 2000, 30)
    want = P.slab_expected(theta, small)
    assert np.all(np.abs(m - want) <= 4 * se + 2e-4), (m, want, se)


@pytest.mark.parametrize("theta", [60.0, 40.0])
def test_total_internal_reflection_weight(theta):
    """From inside the glass beyond the critical angle the reference reflects with
    radiance 1 and probability 1, so the sample is the cosine term, Le cos(theta)
    (an energy loss, not Le); below it the mirror branch has probability er and
    weight 1, Le er."""
    objs, cam = P.tir_scene(theta)
    sc = O.Scene(objs)
    m, se = _pixel_mean(sc, cam, 800, 8)
    want = P.tir_expected(theta)
    if theta > 45:
        assert np.all(se < 1e-4)  # deterministic up to the jitter's change of angle
        np.testing.assert_allclose(m, want, rtol=2e-3)
    else:
        assert np.all(np.abs(m - want) <= 4 * se + 2e-4), (m, want, se)
    # the scatter itself: mirror direction, radiance 1, probabilities 1, the uniforms unused
    th = math.radians(theta)
    if theta > 45:
        o = np.array(cam["pos"], np.float32)
        d = np.array([math.sin(th), 0, math.cos(th)], np.float32)
        obj, tri, t = _first_hit(sc, np.concatenate([o, d]))
        for rn in ((0.0, 0.1, 0.2), (0.999, 0.9, 0.5)):
            r = sc.scatter(0, np.concatenate([o, d]), obj, tri, t, 0, rn)
            np.testing.assert_allclose(r["direction"], [math.sin(th), 0, -math.cos(th)], atol=2e-6)
            assert np.all(r["radiance"] == 1) and r["direction_prob"] == 1 and r["origin_prob"] == 1
            assert r["inside"] == 0 and r["delta"]


# ------------------------------------------------------------------ camera ---
@pytest.mark.parametrize("shape", [(16, 12, 8.0), (7, 5, 3.0), (640, 360, 900.0)])
def test_camera_directions_and_weight(shape):
    """Ray directions of corner, edge and centre pixels at several jitters equal the
    float64 evaluation of camera.cpp:20-23 (y = 0 is the top row), and the sample
    weight equals cos(theta_jit) / cos(theta_centre)."""
    W, H, dist = shape
    cam = S.camera((0.3, -0.2, 2.0), (0.1, 0.4, -1.0), (0.0, 1.0, 0.2), W, H, dist)
    pixels = [0, W - 1, (H - 1) * W, W * H - 1, (H // 2) * W + W // 2, W // 3]
    for p in pixels:
        for u, v in ((0.0, 0.0), (0.5, 0.5), (0.999, 0.25), (0.1, 0.9)):
            d, w = O.camera_ray(cam, p, u, v)
            d64, _ = P.camera_dir64(cam, p, u, v)
            np.testing.assert_allclose(d, d64, atol=3e-6)
            assert abs(w - P.camera_weight64(cam, p, u, v)) < 3e-6
    # orientation: pixel 0 looks up-left, the last pixel down-right
    d0, _ = O.camera_ray(cam, 0, 0.5, 0.5)
    d1, _ = O.camera_ray(cam, W * H - 1, 0.5, 0.5)
    _, front = P.camera_dir64(cam, 0, 0.5, 0.5)
    right = np.cross(front, cam["up"])
    assert d0 @ right < 0 < d1 @ right
    top = np.cross(right, front)
    assert d0 @ top > 0 > d1 @ top


def test_wide_camera_furnace():
    """A 90-degree camera facing a plane emitter: every sample is Le cos(theta_jit) /
    cos(theta_centre), whose per-pixel mean is a float64 quadrature over the pixel.
    Without the vignetting fix the corners would read about 0.6, without the camera's
    cosine term about 1.7."""
    objs, cam = P.wide_camera_scene()
    sc = O.Scene(objs)
    spp = 40000
    s, q = sc.render(cam, spp, 4, 3)
    m = s.reshape(-1, 3)[:, 0] / spp
    se = np.sqrt(np.maximum(q.reshape(-1, 3)[:, 0] / spp - m * m, 0) / spp)
    want = P.wide_camera_expected(cam)
    assert np.all(np.abs(m - want) <= 4 * se + 1e-6), np.abs((m - want) / se).max()
    # near the axis the cosine is concave across a pixel, so the mean there is resolved
    # below 1: taking the vignetting cosine at the jittered ray (a constant 1) fails too
    k = np.argsort(want)[:4]
    dev = 1 - m[k].mean()
    assert dev > 10 * np.sqrt((se[k] ** 2).sum()) / 4, (dev, se[k])
    np.testing.assert_allclose(dev, 1 - want[k].mean(), rtol=0.1)


# ------------------------------------------------------------------- Phong ---
@pytest.mark.parametrize("theta", [0.0, 40.0, 70.0])
def test_phong_scatter_samples_about_normal(theta):
    """Phong scatter at oblique incidence: directions follow cos^n about the normal
    (E[N.w] = (n+1)/(n+2), against about cos(theta) (n+1)/(n+2) if the lobe were sampled
    about the mirror direction), directionProb = p (n+1)/(2 pi) (N.w)^n, radiance =
    albedo (n+2)/(2 pi) max(R.w, 0)^n with R the mirror of the incident ray, every
    sample above the surface."""
    power, alb = 40.0, 0.8
    objs, cam = P.phong_scene(theta, power, alb)
    sc = O.Scene(objs)
    th = math.radians(theta)
    o = np.array([-0.5 * math.sin(th), 0.5 * math.cos(th), 0.0], np.float32)
    d = np.array([math.sin(th), -math.cos(th), 0.0], np.float32)
    obj, tri, t = _first_hit(sc, np.concatenate([o, d]))
    R = np.array([math.sin(th), math.cos(th), 0.0])
    n_side = 48
    cos_n = []
    for i in range(n_side):
        for j in range(n_side):
            r = sc.scatter(0, np.concatenate([o, d]), obj, tri, t, O.NONE,
                           (0.0, (i + 0.5) / n_side, (j + 0.5) / n_side))
            w = r["direction"].astype(np.float64)
            assert abs(np.linalg.norm(w) - 1) < 1e-5 and w[1] > 0
            cos_n.append(w[1])
            pdf = alb * (power + 1) / (2 * math.pi) * w[1] ** power
            val = alb * (power + 2) / (2 * math.pi) * max(R @ w, 0.0) ** power
            assert abs(r["direction_prob"] - pdf) <= 1e-4 * pdf + 1e-30
            assert np.all(np.abs(r["radiance"] - val) <= 1e-3 * val + 1e-30)
    mean = float(np.mean(cos_n))
    assert abs(mean - (power + 1) / (power + 2)) < 2e-3, mean
    # Russian roulette: absorbed with probability 1 - max(albedo)
    a = sc.scatter(0, np.concatenate([o, d]), obj, tri, t, O.NONE, (0.9, 0.5, 0.5))
    assert np.all(a["direction"] == 0) and np.all(a["radiance"] == 0)
    assert abs(a["direction_prob"] - (1 - alb)) < 1e-6


@pytest.mark.parametrize("theta,power", [(60.0, 2.0), (45.0, 1.0), (0.0, 30.0)])
def test_phong_render_oblique_quadrature(theta, power):
    """An oblique camera over a Phong floor under a uniform emitter, one bounce: the pixel
    is Le albedo I with I the float64 quadrature of the Phong BRDF about the mirror
    direction times the cosine (exponents of at most 2 keep the estimator's variance
    finite at oblique incidence, since the lobe is sampled about the normal)."""
    alb = 0.8
    objs, cam = P.phong_scene(theta, power, alb)
    sc = O.Scene(objs)
    m, se = _pixel_mean(sc, cam, 6000, 1)
    want = alb * P.phong_integral(theta, power)
    assert np.all(np.abs(m - want) <= 4 * se + 1e-3), (m, want, se)


# -------------------------------------------------------------------- BDPT ---
def test_bdpt_with_glass_converges_to_path_tracer():
    """BidirectionalPathTracer with the Cornell box's glass sphere kept: paths through the
    delta surfaces come from the unidirectional strategies only, and the MIS weights
    must still sum to one over the strategies that remain.  BDPT at depth 12 against
    PathTracer at depth 40 on the whole image and per pixel."""
    objs, cam = S.cornell(k=1, width=16, height=16)
    sc = O.Scene(objs)
    n1, n2 = 800, 3200
    s1, q1 = sc.render_bdpt(cam, n1, 12, 21)
    s2, q2 = sc.render(cam, n2, 40, 22)
    m1, m2 = s1 / n1, s2 / n2
    v1 = np.maximum(q1 / n1 - m1 * m1, 0) / n1
    v2 = np.maximum(q2 / n2 - m2 * m2, 0) / n2
    diff = m1.sum() - m2.sum()
    se = np.sqrt(v1.sum() + v2.sum())
    assert abs(diff) < 3.5 * se, (m1.sum(), m2.sum(), se)
    z = np.abs(m1 - m2) / np.sqrt(v1 + v2 + 1e-30)
    assert (z < 3.5).mean() > 0.97


# --------------------------------------------------------------------- VCM ---
def test_vcm_radius_schedule():
    """radius = MAX(baseRadius * sqrt(powf(s+1, alpha-1)), EPSILON), alpha = 2/3
    (vcmTracer.cpp:30, vcmTracer.h:39): base * (s+1)^(-1/6), never below 1e-5."""
    for base in (0.01, 0.05, 2.0):
        for s_ in (0, 1, 5, 63, 1000):
            want = base * (s_ + 1) ** (-1.0 / 6.0)
            assert abs(O.vcm_radius(base, s_) - want) <= 1e-6 * want, (base, s_)
    assert O.vcm_radius(1e-7, 0) == np.float32(1e-5)
    assert O.vcm_radius(1e-5, 10) == np.float32(1e-5)


def test_vcm_tiny_radius_equals_bdpt():
    """At the smallest radius the merging strategies weigh nothing (pi r^2 = 3e-10 in a
    unit box) and no light point lies within reach: VCMTracer is then
    BidirectionalPathTracer on the same sub-paths, connections and weights, pixel
    for pixel up to the rounding of the weights' extra (vanishing) terms."""
    objs, cam = S.cornell(k=1, width=16, height=16)
    sc = O.Scene(objs)
    a, _ = sc.render_bdpt(cam, 8, 8, 21)
    b, _ = sc.render_vcm(cam, 8, 8, 21, 1e-7)
    assert a.sum() > 0
    assert np.allclose(a, b, rtol=1e-4, atol=1e-7), np.abs(a - b).max()


def test_vcm_converges_to_bdpt():
    """VCMTracer on the Cornell box with its glass sphere, at a radius where merging
    carries a real share of the image (dropping the merges, leaving out the merge's
    pi r^2 in its probability, or the pixel count in its own weight each darkens the
    image by about 7%), agrees with BidirectionalPathTracer on the whole image: the
    MIS weights over connections and merges sum to one.  The merge's blur is below
    the tolerance at this radius (0.05 in a unit box, shrinking with the passes)."""
    objs, cam = S.cornell(k=1, width=16, height=16)
    sc = O.Scene(objs)
    n = 800
    s1, q1 = sc.render_vcm(cam, n, 12, 23, 0.05)
    s2, q2 = sc.render_bdpt(cam, n, 12, 24)
    m1, m2 = s1 / n, s2 / n
    v1 = np.maximum(q1 / n - m1 * m1, 0) / n
    v2 = np.maximum(q2 / n - m2 * m2, 0) / n
    diff = m1.sum() - m2.sum()
    se = math.sqrt(v1.sum() + v2.sum())
    assert abs(diff) < 4 * se + 0.005 * m2.sum(), (m1.sum(), m2.sum(), se)


# --------------------------------------------------------------------- PPM ---
def _ppm_furnace(alb, le, width=64, height=48):
    """A diffuse floor under an emitter plane (both 4000 wide), a 90-degree camera at half
    height looking straight down."""
    floor = S.quad((0, 0, 0), (4000, 0, 0), (0, 0, 4000), (0, 1, 0), S.DIFFUSE, (alb, alb, alb))
    ceil = S.quad((0, 1, 0), (4000, 0, 0), (0, 0, 4000), (0, -1, 0), S.EMISSIVE, (le, le, le))
    cam = S.camera((0.0, 0.5, 0.0), (0.0, 0.0, 0.0), (0.0, 0.0, 1.0), width, height, sight_dist=width / 2)
    return [floor, ceil], cam


def test_ppm_furnace_density_estimate():
    """Under a uniform emitter plane the floor's irradiance is pi Le everywhere within a
    thousand units of the camera's view, so the kernel density estimate of the photons
    (each carrying Le cos / (pdf_dir pdf_area), the Epanechnikov-type kernel 3 (1-d^2/r^2)^2
    / pi normalised by the photon count and r^2, the diffuse BSDF a/pi) is unbiased at any
    radius: every pixel's expectation is a Le times the camera weight of the pixel (the
    float64 quadrature of cos theta_jit / cos theta_centre).  All pixels of a pass share its
    photons, so the error is measured across passes.  A dropped origin probability, a
    dropped 3/pi or the BSDF's 1/pi missing each moves the ratio by more than 4%."""
    objs, cam = _ppm_furnace(0.5, 2.0)
    sc = O.Scene(objs)
    want = 0.5 * 2.0 * P.wide_camera_expected(cam, m=16)
    ratios = []
    for k in range(400):
        s, _ = sc.render_ppm(cam, 1, 4, 7, 1000.0, pass0=k)
        ratios.append((s.reshape(-1, 3)[:, 0] / want).mean())
    ratios = np.asarray(ratios)
    se = ratios.std() / math.sqrt(len(ratios))
    assert se < 0.008, se
    assert abs(ratios.mean() - 1.0) < 4 * se, (ratios.mean(), se)
