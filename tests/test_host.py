"""Host-side checks that need no GPU: the C-ABI library loads and exports every
declared symbol, the product builder equals the oracle's builder node for node,
and the procedural scene generators produce the sizes the workloads state."""
import ctypes
import os

import numpy as np
import pytest

from oracle import oracle as O
from walkray_amd import _lib as W
from walkray_amd import scenes as S


def test_library_exports_every_header_symbol():
    L = ctypes.CDLL(W.LIB_PATH)
    names = W.header_symbols()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_binding_covers_header():
    W.lib()
    import inspect
    src = inspect.getsource(W.lib)
    for n in W.header_symbols():
        assert f'"{n}"' in src, f"{n} has no ctypes signature"


def test_no_gpu_errors_are_reported_not_crashed():
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is present")
    with pytest.raises(W.WalkRayError):
        W.Context(0)


def _soup(objs):
    V = np.concatenate([o["verts"] for o in objs])
    T, off = [], 0
    for o in objs:
        T.append(o["tris"].astype(np.int64) + off)
        off += len(o["verts"])
    return V, np.concatenate(T).astype(np.uint32)


def _same(a, b):
    for k in ("right", "leaf_first", "leaf_count", "tri_list"):
        assert np.array_equal(a[k], b[k]), k
    # bit for bit: the sign of a zero bound is the reference fold's (first of equal values)
    assert np.array_equal(np.asarray(a["box"], np.float32).view(np.uint32),
                          np.asarray(b["box"], np.float32).view(np.uint32))


@pytest.mark.parametrize("case", ["heightfield", "cornell", "field", "soup", "tiny"])
def test_product_builder_equals_oracle(case):
    if case == "heightfield":
        objs = [S.heightfield(129, seed=4)]
    elif case == "cornell":
        objs = S.cornell(k=3)[0]
    elif case == "field":
        objs = S.random_field(n_objects=24, seed=1)[0]
    elif case == "soup":
        rng = np.random.default_rng(2)
        v = np.round(rng.normal(size=(300, 3)), 1).astype(np.float32)  # coincident coordinates
        t = rng.integers(0, 300, (500, 3)).astype(np.uint32)  # degenerate and repeated triangles
        objs = [{"verts": v, "tris": t, "normals": None, "mat": 0, "params": (0.5,) * 3}]
    else:
        objs = [{"verts": np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32),
                 "tris": np.array([[0, 1, 2]], np.uint32), "normals": None, "mat": 0, "params": (0.5,) * 3}]
    V, T = _soup(objs)
    _same(O.Tree(V, T).export(), W.build_tree_export(V, T))
    # threads change speed, not the tree
    _same(W.build_tree_export(V, T, nthreads=1), W.build_tree_export(V, T, nthreads=8))


def test_builder_large_heightfield_equals_oracle():
    hf = S.heightfield(400, seed=7)
    _same(O.Tree(hf["verts"], hf["tris"]).export(), W.build_tree_export(hf["verts"], hf["tris"]))


def test_builder_signed_zeros_and_nan():
    """Signed zeros everywhere (the box fold keeps the first of equal values, so the
    bound's sign is the first vertex's): the oracle's tree bit for bit, whatever the
    thread count.  A NaN coordinate leaves the split rule's median undefined
    (std::nth_element needs a strict weak order): the build is refused."""
    rng = np.random.default_rng(5)
    v = rng.choice(np.array([0.0, -0.0, 1.0, -1.0, 0.5], np.float32), size=(4000, 3))
    t = rng.integers(0, 4000, (70000, 3)).astype(np.uint32)
    for nth in (1, 8):
        _same(O.Tree(v, t).export(), W.build_tree_export(v, t, nthreads=nth))
    v2 = v.copy()
    v2[int(t[5, 2]), 1] = np.nan
    for nth in (1, 8):
        with pytest.raises(W.WalkRayError, match="NaN"):
            W.build_tree_export(v2, t, nthreads=nth)


def test_builder_rejects_bad_indices():
    v = np.zeros((3, 3), np.float32)
    with pytest.raises(W.WalkRayError):
        W.build_tree_export(v, np.array([[0, 1, 3]], np.uint32))


def test_workload_sizes():
    assert len(S.heightfield(1024, seed=7)["tris"]) == 2093058
    objs, cam = S.cornell(k=5)
    assert sum(len(o["tris"]) for o in objs) == 40974
    assert (cam["width"], cam["height"]) == (256, 256)
    ico = S.icosphere(5)
    assert len(ico["tris"]) == 20480 and len(ico["verts"]) == 10242
    objs, cam = S.random_field()
    n = sum(len(o["tris"]) for o in objs)
    assert 450_000 < n < 550_000 and len(objs) == 515
    mats = [o["mat"] for o in objs[1:-2]]
    assert min(mats.count(m) for m in (S.DIFFUSE, S.GLOSSY, S.REFRACTIVE)) > 120


def test_icosphere_and_box_winding_outward():
    for o in (S.icosphere(2, (1, 2, 3), 0.5), S.box((0, 0, 0), (1, 2, 3))):
        v = o["verts"].astype(np.float64)
        c = v.mean(axis=0)
        a, b, cc = v[o["tris"][:, 0]], v[o["tris"][:, 1]], v[o["tris"][:, 2]]
        n = np.cross(b - a, cc - a)
        assert np.all(np.einsum("ij,ij->i", n, (a + b + cc) / 3 - c) > 0)


def _wide_walk(wide, V, T):
    """Walks the fast-pass hierarchy; returns {record: number of visits} and checks, for every
    child box, that it contains the boxes of all triangles below it (decode: fma(q, scale, origin))."""
    nodes = wide["nodes"]
    tri_lo = V[T].min(axis=1)
    tri_hi = V[T].max(axis=1)
    rec_tri = wide["record_tri"]
    seen = {}

    def below(ref):  # records under a reference
        if ref & 0x80000000:
            assert (ref >> 27) & 0xF == 1, "wide leaves hold one triangle"
            rec = ref & 0x07FFFFFF
            seen[rec] = seen.get(rec, 0) + 1
            return [rec]
        nd = nodes[ref]
        out = []
        for c in range(4):
            r = int(nd["ref"][c])
            if r == 0xFFFFFFFF:
                continue
            recs = below(r)
            q = nd["q"][:, c].astype(np.float64)
            box = (q * nd["scale"][[0, 1, 2, 0, 1, 2]].astype(np.float64)
                   + nd["origin"][[0, 1, 2, 0, 1, 2]].astype(np.float64)).astype(np.float32)
            tris = rec_tri[recs]
            assert np.all(box[:3] <= tri_lo[tris].min(axis=0)), "child box must contain its triangles"
            assert np.all(box[3:] >= tri_hi[tris].max(axis=0)), "child box must contain its triangles"
            out += recs
        return out

    below(int(wide["root"]))
    return seen


@pytest.mark.parametrize("case", ["heightfield", "cornell", "field", "soup", "tiny"])
def test_wide_hierarchy_covers_every_record_once(case):
    if case == "heightfield":
        objs = [S.heightfield(65, seed=4)]
    elif case == "cornell":
        objs = S.cornell(k=2)[0]
    elif case == "field":
        objs = S.random_field(n_objects=12, seed=1)[0]
    elif case == "soup":
        rng = np.random.default_rng(2)
        v = np.round(rng.normal(size=(300, 3)), 1).astype(np.float32)
        t = rng.integers(0, 300, (500, 3)).astype(np.uint32)
        objs = [{"verts": v, "tris": t, "normals": None, "mat": 0, "params": (0.5,) * 3}]
    else:
        objs = [{"verts": np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32),
                 "tris": np.array([[0, 1, 2]], np.uint32), "normals": None, "mat": 0, "params": (0.5,) * 3}]
    V, T = _soup(objs)
    wide = W.build_tree_wide(V, T)
    ex = W.build_tree_export(V, T)
    # records are the reference leaves' triangles in left-to-right leaf order
    assert np.array_equal(wide["record_tri"], ex["tri_list"])
    seen = _wide_walk(wide, V, T)
    assert sorted(seen) == list(range(len(ex["tri_list"])))
    assert set(seen.values()) == {1}
    # record -> first record of its reference leaf (the tie key)
    leaves = ex["right"] == 0xFFFFFFFF
    for f, c in zip(ex["leaf_first"][leaves], ex["leaf_count"][leaves]):
        assert np.all(wide["record_first"][f:f + c] == f)


def test_wide_hierarchy_threads_change_speed_not_result():
    o = S.heightfield(300, seed=6)  # 178,802 triangles: the parallel build path
    a = W.build_tree_wide(o["verts"], o["tris"], nthreads=1)
    b = W.build_tree_wide(o["verts"], o["tris"], nthreads=8)
    assert a["root"] == b["root"]
    assert np.array_equal(a["nodes"].view(np.uint8), b["nodes"].view(np.uint8))
    assert np.array_equal(a["record_first"], b["record_first"])


@pytest.mark.parametrize("case", ["field", "heightfield"])
def test_refit_hierarchy_after_motion(case):
    """The per-pass rebuild of moving objects: the reference tree is built anew for the
    new positions (records in its leaf order, the same as a fresh build) and the previous
    fast-pass hierarchy is refit to them: every record still appears exactly once and
    every child box contains its triangles at their new positions."""
    rng = np.random.default_rng(3)
    if case == "field":
        objs = S.random_field(n_objects=16, seed=2)[0]
        V, T = _soup(objs)
        V2 = V.copy()
        off = np.cumsum([0] + [len(o["verts"]) for o in objs])
        for i in (1, 4, 9):  # three objects move and turn
            M = S.trs(rng.uniform(-0.3, 0.3, 3), rng.normal(size=3), rng.uniform(-1, 1), [1.0, 1.0, 1.0])
            p = V[off[i]:off[i + 1]].astype(np.float64)
            V2[off[i]:off[i + 1]] = (p @ M[:3, :3].T.astype(np.float64) + M[:3, 3]).astype(np.float32)
    else:
        hf = S.heightfield(65, seed=4)
        V, T = hf["verts"], hf["tris"]
        V2 = V.copy()
        V2[:, 2] += rng.uniform(-0.05, 0.05, len(V)).astype(np.float32)
    refit = W.build_tree_wide(V, T, refit_verts=V2)
    fresh = W.build_tree_wide(V2, T)
    assert np.array_equal(refit["record_tri"], fresh["record_tri"])
    assert np.array_equal(refit["record_first"], fresh["record_first"])
    seen = _wide_walk(refit, V2, T)
    assert sorted(seen) == list(range(len(T))) and set(seen.values()) == {1}
