"""GPU light sub-paths (genLightSample + samplePath), photon collection and the
photon hash grid through the C ABI, against the oracle.

Bars:
  * light paths: the same vertex count and the same integer fields (contact and
    inside objects, contact triangle, delta flag) on at least 99% of paths, the
    float fields within 1e-4 relative there; many paths bit-identical (the shared
    counter RNG; the rest differ by the last-ulp results of libm sin/cos/pow, 0.42
    bit-identical on the Cornell box);
  * photon collection on the GPU's own records equals the oracle's rule, bit for bit;
  * hash grid: bbox, mCellEnds and mIndices equal to the oracle's Build, and every
    query's particle list (order and repeats included) equal to its Process; at full
    size (one light path per 1080p pixel), structure and query sets against a
    brute-force neighbour search.
"""
import numpy as np
import pytest

from oracle import oracle as O
from walkray_amd import _lib as W
from walkray_amd import scenes as S

pytestmark = pytest.mark.gpu

INT_FIELDS = ("contact", "contact_tri", "inside", "delta")
FLOAT_FIELDS = ("origin", "direction", "radiance", "origin_prob", "direction_prob")


def _compare_paths(g, gc, o, oc):
    n = len(gc)
    same_count = gc == oc
    agree = np.zeros(n, bool)
    exact = np.zeros(n, bool)
    for k in np.nonzero(same_count)[0]:
        c = int(gc[k])
        a, b = g[k, :c], o[k, :c]
        if not all(np.array_equal(a[f], b[f]) for f in INT_FIELDS):
            continue
        agree[k] = all(np.allclose(a[f], b[f], rtol=1e-4, atol=1e-5) for f in FLOAT_FIELDS)
        exact[k] = all(np.array_equal(a[f].view(np.uint32), b[f].view(np.uint32)) for f in FLOAT_FIELDS)
    return agree.mean(), exact.mean()


@pytest.mark.parametrize("case", ["cornell", "field"])
def test_light_paths_match_oracle(ctx, case):
    if case == "cornell":
        objs, _ = S.cornell(k=3, width=8, height=8)
        depth = 8
    else:
        objs, _ = S.random_field(n_objects=64, width=16, height=9, sphere_k=2)
        depth = 6
    print(f"[lp] light paths vs oracle: {case}", flush=True)
    sc = W.Scene(ctx, objs)
    n, maxv = 20000, depth + 3
    g, gc = sc.light_paths(depth, 77, 2, 1000, n, maxv)
    o, oc, longest = O.Scene(objs).light_paths(depth, 77, 2, 1000, n, maxv)
    assert longest <= maxv and gc.max() <= maxv
    agree, exact = _compare_paths(g, gc, o, oc)
    print(f"[lp] {case}: {agree:.4f} agree, {exact:.4f} bit-identical", flush=True)
    assert agree >= 0.99 and exact >= 0.25


def test_light_path_continuity_on_gpu(ctx):
    """Each GPU vertex after the first starts at the GPU NoSelf closest hit of the ray before it."""
    objs, _ = S.cornell(k=3, width=8, height=8)
    sc = W.Scene(ctx, objs)
    g, gc = sc.light_paths(8, 3, 0, 0, 4000, 11)
    rays, ins, con, nxt = [], [], [], []
    for k in range(len(gc)):
        for i in range(int(gc[k]) - 1):
            a, b = g[k, i], g[k, i + 1]
            if b["delta"] and np.linalg.norm(b["direction"]) < 0.5:  # terminateRay
                continue
            rays.append(np.concatenate([a["origin"], a["direction"]]))
            ins.append(a["inside"])
            con.append(a["contact"])
            nxt.append(b["origin"])
    rays = np.array(rays, np.float32)
    h = sc.closest_hit_noself(rays, np.array(ins, np.uint32), np.array(con, np.uint32))
    assert np.all(h["prim"] != W.NONE)
    p = rays[:, :3] + h["t"][:, None] * rays[:, 3:]
    assert np.allclose(p, np.array(nxt), atol=1e-5)


def test_photons_collect_equals_rule(ctx):
    objs, _ = S.random_field(n_objects=64, width=16, height=9, sphere_k=2)
    sc = W.Scene(ctx, objs)
    n, maxv, lead = 30000, 9, 1234
    dv, dc = sc.light_paths_device(6, 5, 1, 0, n, maxv)
    dp, ds, m = sc.photons_device(dv, dc, n, maxv, n_leading=lead)
    rec = dv.download(W.PATH_VERTEX_DTYPE, n * maxv).reshape(n, maxv)
    cnt = dc.download(np.uint32, n)
    pos, src = O.photons([o["mat"] for o in objs], rec, cnt, n_leading=lead)
    assert m == len(pos)
    assert np.array_equal(dp.download(np.float32, 3 * m).reshape(-1, 3).view(np.uint32), pos.view(np.uint32))
    assert np.array_equal(ds.download(np.uint32, 2 * m).reshape(-1, 2), src)
    # a cap below the total writes a prefix and still reports the total
    dp2, ds2, m2 = sc.photons_device(dv, dc, n, maxv, n_leading=lead, cap=lead + 100)
    assert m2 == m
    assert np.array_equal(dp2.download(np.float32, 3 * (lead + 100)).reshape(-1, 3), pos[:lead + 100])


@pytest.mark.parametrize("n_cells", [1024, 97, 1])
def test_hashgrid_equals_oracle(ctx, n_cells):
    rng = np.random.default_rng(n_cells)
    pos = rng.uniform(-2, 3, (3000, 3)).astype(np.float32)
    pos[:50] = 0.0  # leading points at the origin, as VCM's vector starts
    q = np.concatenate([rng.uniform(-2.5, 3.5, (500, 3)), pos[100:300] + rng.normal(0, 0.05, (200, 3))])
    q = q.astype(np.float32)
    g = W.HashGrid.from_points(ctx, pos, 0.3, n_cells)
    o = O.HashGrid(pos, 0.3, n_cells)
    for a, b in zip(g.export(), o.export()):
        assert np.array_equal(a, b)
    cap = 4096 if n_cells > 1 else 24000
    gc, gi = g.query(q, cap)
    oc, oi = o.query(q, cap)
    assert np.array_equal(gc, oc) and gc.max() <= cap
    for k in range(len(q)):
        assert np.array_equal(gi[k, :gc[k]], oi[k, :oc[k]]), k


def test_hashgrid_lattice_and_empty(ctx):
    ax = np.arange(8, dtype=np.float32) * 0.5
    pos = np.stack(np.meshgrid(ax, ax, ax, indexing="ij"), -1).reshape(-1, 3)
    g = W.HashGrid.from_points(ctx, pos, 0.5, 97)
    q = np.array([[1.5, 1.5, 1.5], [2.0, 1.0, 1.5], [-0.1, 1.0, 1.0], [3.6, 1.0, 1.0]], np.float32)
    c, i = g.query(q, 64)
    assert len(set(i[0, :c[0]].tolist())) == 7 and len(set(i[1, :c[1]].tolist())) == 7
    assert c[2] == 0 and c[3] == 0
    e = W.HashGrid.from_points(ctx, np.zeros((0, 3), np.float32), 0.5, 16)
    bbox, ends, idx = e.export()
    assert np.all(ends == 0) and len(idx) == 0 and bbox[0] == np.float32(1e36)
    c, _ = e.query(q, 4)
    assert np.all(c == 0)


def _cells(pos, bbox, r, n_cells):
    inv = np.float32(1) / (np.float32(r) * np.float32(2))
    dm = (pos - bbox[:3]).astype(np.float32)
    ci = np.floor(inv * dm).astype(np.int32).astype(np.uint32)
    h = (ci[:, 0] * np.uint32(73856093)) ^ (ci[:, 1] * np.uint32(19349663)) ^ (ci[:, 2] * np.uint32(83492791))
    return h % np.uint32(n_cells)


def test_full_size_photon_grid(ctx):
    """One light path per 1080p pixel of the 512-object field (maxPathLen 16), the photons
    VCM keeps (after one leading point per pixel), a grid with one cell per pixel and
    VCM's base radius: structure, and query sets against a brute-force search."""
    from scipy.spatial import cKDTree
    objs, cam = S.random_field()
    sc = W.Scene(ctx, objs)
    npix = cam["width"] * cam["height"]
    depth, maxv, r = 16, 19, 0.01
    print("[lp] full size: light paths", flush=True)
    dv, dc = sc.light_paths_device(depth, 11, 0, 0, npix, maxv)
    dp, ds, m = sc.photons_device(dv, dc, npix, maxv, n_leading=npix)
    del dv
    print(f"[lp] full size: {m} points, building the grid", flush=True)
    g = W.HashGrid(ctx, dp, m, r, npix)
    bbox, ends, idx = g.export()
    pos = dp.download(np.float32, 3 * m).reshape(-1, 3)
    assert ends[-1] == m and np.all(np.diff(ends) >= 0)
    assert np.array_equal(np.sort(idx), np.arange(m))
    cell = _cells(pos, bbox, r, npix)
    starts = np.concatenate([[0], ends[:-1]])
    owner = np.repeat(np.arange(npix, dtype=np.uint32), ends - starts)
    assert np.array_equal(cell[idx], owner)  # every point sits in its cell's range
    seg_ok = np.all((np.diff(idx) > 0) | (np.diff(owner) != 0))  # input order inside a cell
    assert seg_ok
    rng = np.random.default_rng(1)
    qs = pos[npix:][rng.integers(0, m - npix, 3000)] + rng.normal(0, 0.004, (3000, 3)).astype(np.float32)
    qs = qs[np.linalg.norm(qs, axis=1) > 3 * r].astype(np.float32)  # away from the leading points' cell
    counts, got = g.query(qs, 8192)
    tree = cKDTree(pos.astype(np.float64))
    inside = np.all((qs >= bbox[:3]) & (qs <= bbox[3:]), axis=1)
    for k in range(len(qs)):
        cand = tree.query_ball_point(qs[k].astype(np.float64), r * 1.001)
        d = np.sqrt(((qs[k] - pos[cand]) ** 2).sum(-1, dtype=np.float32), dtype=np.float32)
        want = set(np.asarray(cand)[(d * d) <= np.float32(r) * np.float32(r)].tolist())
        got_k = set(got[k, :counts[k]].tolist())
        assert got_k == (want if inside[k] else set()), k
    assert counts.mean() > 0.5 and counts.max() >= 3


@pytest.mark.parametrize("n_cells,cap", [(1, 5000), (3, 20000)])
def test_hashgrid_long_ranges_in_order(ctx, n_cells, cap):
    """Buckets far longer than one step of the workgroup query (4,096 points per
    iteration): 10,000 points, most of them within the radius of the queries, so every
    query walks several iterations of its ranges; the lists equal the oracle's in order
    (repeats through hash collisions included) up to the cap, and the counts beyond it."""
    rng = np.random.default_rng(31 + n_cells)
    pos = rng.uniform(0.0, 0.2, (10000, 3)).astype(np.float32)
    pos[::7] = 0.05  # coincident points
    q = np.concatenate([rng.uniform(0.0, 0.2, (40, 3)), np.full((4, 3), 0.05)]).astype(np.float32)
    g = W.HashGrid.from_points(ctx, pos, 0.12, n_cells)
    o = O.HashGrid(pos, 0.12, n_cells)
    gc, gi = g.query(q, cap)
    oc, oi = o.query(q, cap)
    assert np.array_equal(gc, oc) and gc.min() > 4096
    assert (gc > cap).any()  # some lists are cut at the cap
    for k in range(len(q)):
        m = min(int(gc[k]), cap)
        assert np.array_equal(gi[k, :m], oi[k, :m]), k
