"""GPU closest-hit parity through the C ABI against the CPU oracle.

Bar (exact integer/index work, FP32 arithmetic in the reference order): prim ids
equal, t bit-identical, barycentrics bit-identical.  The oracle follows
kdtree.cpp / scene.cpp / noself.cpp with the same operation order, so any
difference is a bug, not a tolerance question.
"""
import numpy as np
import pytest

from oracle import oracle as O
from walkray_amd import _lib as W
from walkray_amd import scenes as S

pytestmark = pytest.mark.gpu


def _assert_same(hits, ref, label, check_bary=True):
    prim = hits["prim"]
    assert not np.any(prim == W.OVERFLOW), f"{label}: stack overflow reported"
    bad = np.nonzero(prim != ref["prim"])[0]
    assert bad.size == 0, (f"{label}: {bad.size} prim mismatches, first at {bad[:5]}: "
                           f"gpu {prim[bad[:5]]} t {hits['t'][bad[:5]]} vs oracle {ref['prim'][bad[:5]]} "
                           f"t {ref['t'][bad[:5]]}")
    tb = hits["t"].view(np.uint32)
    rb = ref["t"].view(np.uint32)
    badt = np.nonzero(tb != rb)[0]
    assert badt.size == 0, f"{label}: {badt.size} t mismatches, e.g. {hits['t'][badt[:3]]} vs {ref['t'][badt[:3]]}"
    if check_bary:
        hit = prim != O.NONE
        for k in ("b1", "b2"):
            a = hits[k][hit].view(np.uint32)
            b = ref[k][hit].view(np.uint32)
            assert np.array_equal(a, b), f"{label}: barycentric {k} differs on {(a != b).sum()} rays"


def _scene_pair(ctx, objs):
    return W.Scene(ctx, objs), O.Scene(objs)


def test_heightfield_small(ctx):
    objs = [S.heightfield(129, seed=2)]
    sc, osc = _scene_pair(ctx, objs)
    lo, hi = S.scene_bounds(objs)
    rays = np.concatenate([S.pinhole_rays(S.heightfield_camera(64, 64)),
                           S.incoherent_rays(16384, lo, hi, seed=5)])
    hits = sc.closest_hit(rays)
    ref = osc.intersect(rays)
    _assert_same(hits, ref, "heightfield129")
    assert (hits["prim"] != O.NONE).mean() > 0.3


def test_cornell_plain_rays(ctx):
    objs, cam = S.cornell(k=3, width=48, height=48)
    sc, osc = _scene_pair(ctx, objs)
    rays = np.concatenate([S.pinhole_rays(cam), S.incoherent_rays(8192, (0.01, 0.01, 0.01), (0.99, 0.99, 0.99), 9)])
    _assert_same(sc.closest_hit(rays), osc.intersect(rays), "cornell")


def _surface_rays(osc, objs, n, seed):
    """Rays leaving surface points, as the scatter of a path vertex produces them."""
    rng = np.random.default_rng(seed)
    probe = S.incoherent_rays(n * 2, (0.02, 0.02, 0.02), (0.98, 0.98, 0.98), seed)
    h = osc.intersect(probe)
    ok = h["prim"] != O.NONE
    probe = probe[ok][:n]
    t = h["t"][ok][:n]
    obj = h["obj"][ok][:n]
    o = (probe[:, :3] + probe[:, 3:] * t[:, None]).astype(np.float32)
    d = rng.normal(size=(len(o), 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.concatenate([o, d.astype(np.float32)], axis=1).astype(np.float32)
    glass = [i for i, ob in enumerate(objs) if ob["mat"] == S.REFRACTIVE]
    inside = np.full(len(o), O.NONE, np.uint32)
    sel = rng.random(len(o)) < 0.3
    if glass:
        inside[sel] = glass[0]
    return rays, inside, obj.astype(np.uint32)


def test_cornell_noself(ctx):
    objs, _ = S.cornell(k=3)
    sc, osc = _scene_pair(ctx, objs)
    rays, inside, contact = _surface_rays(osc, objs, 8192, 21)
    hits = sc.closest_hit_noself(rays, inside, contact)
    ref = osc.intersect(rays, inside, contact, noself=True)
    _assert_same(hits, ref, "cornell-noself")


def test_inside_object(ctx):
    objs, _ = S.cornell(k=3)
    sc, osc = _scene_pair(ctx, objs)
    rng = np.random.default_rng(4)
    n = 4096
    o = rng.uniform(0.02, 0.98, size=(n, 3))
    # half of the origins inside one of the two spheres
    c = np.where(rng.random(n)[:, None] < 0.5, [0.28, 0.201, 0.35], [0.72, 0.201, 0.62])
    u = rng.normal(size=(n, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    o[: n // 2] = c[: n // 2] + u[: n // 2] * 0.15 * rng.random((n // 2, 1))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.concatenate([o, d], axis=1).astype(np.float32)
    cur = np.where(rng.random(n) < 0.2, 5, O.NONE).astype(np.uint32)
    got = sc.inside_object(rays, cur)
    want = osc.find_inside(rays, cur)
    assert np.array_equal(got, want), f"{(got != want).sum()} inside-object mismatches"
    assert (want == 4).sum() > 100 and (want == 5).sum() > 50


def test_edge_cases(ctx):
    objs = [S.heightfield(17, seed=1)]
    sc, osc = _scene_pair(ctx, objs)
    lo, hi = S.scene_bounds(objs)
    assert len(sc.closest_hit(np.zeros((0, 6), np.float32))) == 0
    for n in (1, 63, 257, 1000):
        rays = S.incoherent_rays(n, lo, hi, seed=n)
        _assert_same(sc.closest_hit(rays), osc.intersect(rays), f"ragged n={n}")
    # axis-aligned directions (zero components), vertical rays, rays starting on the root box faces
    g = np.linspace(0.0, 1.0, 9, dtype=np.float32)
    xx, yy = np.meshgrid(g, g)
    vert = np.stack([xx.ravel(), yy.ravel(), np.full(xx.size, 2.0, np.float32),
                     np.zeros(xx.size, np.float32), np.zeros(xx.size, np.float32),
                     -np.ones(xx.size, np.float32)], axis=1).astype(np.float32)
    horiz = np.stack([np.full(xx.size, -1.0, np.float32), yy.ravel(), xx.ravel() * 0.3 - 0.15,
                      np.ones(xx.size, np.float32), np.zeros(xx.size, np.float32),
                      np.zeros(xx.size, np.float32)], axis=1).astype(np.float32)
    onface = vert.copy()
    onface[:, 2] = hi[2]
    rays = np.concatenate([vert, horiz, onface])
    _assert_same(sc.closest_hit(rays), osc.intersect(rays), "axis-aligned")
    # zero direction: the reference box test skips every face -> miss
    z = np.array([[0.5, 0.5, 1.0, 0, 0, 0]], np.float32)
    h = sc.closest_hit(z)
    assert h["prim"][0] == O.NONE and h["t"][0] == -1.0


def test_single_triangle_and_empty(ctx):
    tri = {"verts": np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32),
           "tris": np.array([[0, 1, 2]], np.uint32), "normals": None, "mat": S.DIFFUSE, "params": (0.5, 0.5, 0.5)}
    sc, osc = _scene_pair(ctx, [tri])
    rays = S.incoherent_rays(2048, (-0.5, -0.5, -1), (1.5, 1.5, 1), seed=2)
    _assert_same(sc.closest_hit(rays), osc.intersect(rays), "single-triangle")
    empty = {"verts": np.zeros((0, 3), np.float32), "tris": np.zeros((0, 3), np.uint32), "normals": None,
             "mat": S.DIFFUSE, "params": (0.5, 0.5, 0.5)}
    sc2 = W.Scene(ctx, [empty])
    h = sc2.closest_hit(rays[:100])
    assert np.all(h["prim"] == O.NONE) and np.all(h["t"] == -1.0)


def test_full_size_heightfield_sampled(ctx):
    """The 2,093,058-triangle C2 heightfield: exact parity on sampled rays of both sets,
    plus size-independent properties over 1M rays of each set."""
    objs = S.heightfield_scene(1024, seed=7)
    sc = W.Scene(ctx, objs)
    info = sc.info()
    assert info["n_triangles"] == 2093058
    tree = O.Tree(objs[0]["verts"], objs[0]["tris"])
    cam = S.heightfield_camera(4096, 4096)
    coh = S.pinhole_rays(cam)
    rng = np.random.default_rng(0)
    sample = rng.choice(len(coh), 32768, replace=False)
    lo, hi = S.scene_bounds(objs)
    inc = S.incoherent_rays(1 << 20, lo, hi, seed=11)
    for name, rays in (("coherent", coh[sample]), ("incoherent", inc[:32768])):
        hits = sc.closest_hit(rays)
        prim, t = tree.intersect(rays)
        assert np.array_equal(hits["prim"], prim), f"{name}: {(hits['prim'] != prim).sum()} prim mismatches"
        assert np.array_equal(hits["t"].view(np.uint32), t.view(np.uint32)), f"{name}: t mismatch"
    # properties on 1M rays: every hit lies on its triangle's plane and inside it
    hits = sc.closest_hit(inc)
    hit = hits["prim"] != O.NONE
    v = objs[0]["verts"].astype(np.float64)
    tr = objs[0]["tris"][hits["prim"][hit]]
    p = inc[hit, :3].astype(np.float64) + inc[hit, 3:] * hits["t"][hit, None]
    rec = v[tr[:, 0]] + hits["b1"][hit, None] * (v[tr[:, 1]] - v[tr[:, 0]]) + \
        hits["b2"][hit, None] * (v[tr[:, 2]] - v[tr[:, 0]])
    assert np.max(np.abs(rec - p)) < 1e-4
    assert np.all(hits["t"][hit] >= 0)
    eps = 1e-3
    assert np.all((hits["b1"][hit] > -eps) & (hits["b2"][hit] > -eps) & (hits["b1"][hit] + hits["b2"][hit] < 1 + eps))


def test_field_plain_noself_inside(ctx):
    """The object-field workload (smaller): plain rays, NoSelf path rays and findInsideObject."""
    objs, cam = S.random_field(n_objects=48, seed=5, width=96, height=54)
    sc, osc = _scene_pair(ctx, objs)
    lo, hi = S.scene_bounds(objs)
    rays = np.concatenate([S.pinhole_rays(cam), S.incoherent_rays(16384, lo, hi, seed=3)])
    _assert_same(sc.closest_hit(rays), osc.intersect(rays), "field")
    rng = np.random.default_rng(8)
    probe = S.incoherent_rays(40000, (-16, 0.0, -8), (16, 1.0, 8), seed=6)
    h = osc.intersect(probe)
    ok = h["prim"] != O.NONE
    o = (probe[ok, :3] + probe[ok, 3:] * h["t"][ok, None]).astype(np.float32)[:16384]
    contact = h["obj"][ok][:16384].astype(np.uint32)
    d = rng.normal(size=(len(o), 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    prays = np.concatenate([o, d.astype(np.float32)], axis=1).astype(np.float32)
    glass = np.array([i for i, ob in enumerate(objs) if ob["mat"] == S.REFRACTIVE], np.uint32)
    inside = np.where(rng.random(len(o)) < 0.3, rng.choice(glass, len(o)), O.NONE).astype(np.uint32)
    _assert_same(sc.closest_hit_noself(prays, inside, contact), osc.intersect(prays, inside, contact, noself=True),
                 "field-noself")
    # origins inside objects (sphere and box centres, jittered) and outside
    cen = np.array([ob["verts"].mean(axis=0) for ob in objs[1:-2]], np.float32)
    oi = cen[rng.integers(0, len(cen), 4096)] + rng.normal(scale=0.05, size=(4096, 3)).astype(np.float32)
    di = rng.normal(size=(4096, 3))
    di /= np.linalg.norm(di, axis=1, keepdims=True)
    irays = np.concatenate([oi, di.astype(np.float32)], axis=1).astype(np.float32)
    cur = np.full(len(irays), O.NONE, np.uint32)
    got = sc.inside_object(irays, cur)
    want = osc.find_inside(irays, cur)
    assert np.array_equal(got, want), f"{(got != want).sum()} inside-object mismatches"
    assert (want != O.NONE).mean() > 0.3


def test_ties_on_coincident_and_coplanar_triangles(ctx):
    """Equal t across leaves and inside a leaf: a soup with repeated and coincident triangles, and
    vertical rays through the vertices and edges of a flat grid (up to six triangles at the same t)."""
    rng = np.random.default_rng(2)
    v = np.round(rng.normal(size=(300, 3)), 1).astype(np.float32)
    t = rng.integers(0, 300, (500, 3)).astype(np.uint32)
    t = np.concatenate([t, t[:100], t[50:150, [1, 2, 0]]])  # repeats, and repeats with rotated corners
    soup = {"verts": v, "tris": t, "normals": None, "mat": S.DIFFUSE, "params": (0.5, 0.5, 0.5)}
    sc, osc = _scene_pair(ctx, [soup])
    lo, hi = S.scene_bounds([soup])
    rays = S.incoherent_rays(20000, lo, hi, seed=12)
    _assert_same(sc.closest_hit(rays), osc.intersect(rays), "soup")
    flat = S.heightfield(33, seed=1, amplitude=0.0)
    sc, osc = _scene_pair(ctx, [flat])
    g = np.arange(0, 65, dtype=np.float32) / 64.0  # vertices (even steps) and edge midpoints (odd steps)
    xx, yy = np.meshgrid(g, g)
    n = xx.size
    vert = np.stack([xx.ravel(), yy.ravel(), np.ones(n, np.float32), np.zeros(n, np.float32),
                     np.zeros(n, np.float32), -np.ones(n, np.float32)], axis=1).astype(np.float32)
    _assert_same(sc.closest_hit(vert), osc.intersect(vert), "flat-grid ties")


def test_hit_sources_batch(ctx):
    """wr_hit_sources maps every hit's global triangle id to (object, triangle of the
    object) as wr_scene_prim_source and the oracle's Scene::intersect do; misses give
    WR_NONE.  An object without triangles sits between two others."""
    objs, cam = S.cornell(k=1, width=24, height=24)
    objs = objs[:3] + [dict(objs[3], verts=objs[3]["verts"][:0], tris=objs[3]["tris"][:0])] + objs[4:]
    sc = W.Scene(ctx, objs)
    rays = np.concatenate([S.pinhole_rays(cam), -S.pinhole_rays(cam)])
    n = len(rays)
    dr = ctx.to_device(rays)
    dh = ctx.alloc(16 * n)
    sc.closest_hit_device(dr, n, dh)
    obj, tri = sc.hit_sources(dh, n)
    hits = dh.download(np.uint32, 4 * n).view(W.HIT_DTYPE)
    o = O.Scene(objs).intersect(rays)
    assert np.array_equal(hits["prim"], o["prim"])
    miss = hits["prim"] == W.NONE
    assert miss.any() and not miss.all()
    assert np.all(obj[miss] == W.NONE) and np.all(tri[miss] == W.NONE)
    assert np.array_equal(obj[~miss], o["obj"][~miss]) and np.array_equal(tri[~miss], o["tri"][~miss])
    for k in np.flatnonzero(~miss)[::97]:
        assert sc.prim_source(int(hits["prim"][k])) == (obj[k], tri[k])


def _grazing_rays(objs, n, seed, eps_max=3e-4):
    """Rays nearly in the plane of a random triangle (|n . d| below eps_max), aimed through
    its centroid from a random distance: the regime where the reference's t is
    ill-conditioned, for the fast pass's pruning margin."""
    rng = np.random.default_rng(seed)
    V = np.concatenate([o["verts"] for o in objs]).astype(np.float64)
    T, off = [], 0
    for o in objs:
        T.append(o["tris"].astype(np.int64) + off)
        off += len(o["verts"])
    T = np.concatenate(T)
    pick = rng.integers(len(T), size=n)
    a, b, c = V[T[pick, 0]], V[T[pick, 1]], V[T[pick, 2]]
    nrm = np.cross(b - a, c - a)
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    tan = np.cross(nrm, rng.normal(size=(n, 3)))
    tan /= np.linalg.norm(tan, axis=1, keepdims=True)
    eps = rng.uniform(-eps_max, eps_max, n)[:, None]
    d = tan * np.sqrt(1 - eps * eps) + nrm * eps
    cen = (a + b + c) / 3
    o = cen - d * rng.uniform(0.01, 0.5, n)[:, None] + nrm * rng.uniform(-1e-4, 1e-4, n)[:, None]
    return np.concatenate([o, d], 1).astype(np.float32)


@pytest.mark.parametrize("which", ["heightfield", "field"])
def test_near_grazing_rays(ctx, which):
    """131,072 rays at |n . d| < 3e-4 to a triangle they pass next to: prim, t and
    barycentrics bit-identical to the oracle's unpruned traversal."""
    if which == "heightfield":
        objs = [S.heightfield(257, seed=4)]
    else:
        objs, _ = S.random_field(n_objects=48, width=16, height=16, sphere_k=2)
    sc, osc = _scene_pair(ctx, objs)
    rays = _grazing_rays(objs, 1 << 17, seed=5)
    hits = sc.closest_hit(rays)
    ref = osc.intersect(rays)
    _assert_same(hits, ref, f"grazing-{which}")
    assert (ref["prim"] != O.NONE).mean() > 0.3
