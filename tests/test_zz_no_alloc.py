"""wr_hashgrid_query_noalloc (the entry point for callers that cannot allocate
device memory) against wr_hashgrid_query: two calls in a row on the shared
fixed work list, the second with fewer queries, each through the C ABI."""
import numpy as np
import pytest

from walkray_amd import _lib as W

pytestmark = pytest.mark.gpu


def _both(ctx, g, q, cap):
    nq = len(q)
    dq = ctx.to_device(q)
    want_c, want_i = ctx.alloc(4 * nq), ctx.alloc(4 * max(nq * cap, 1))
    got_c, got_i = ctx.alloc(4 * nq), ctx.alloc(4 * max(nq * cap, 1))
    g.query_device(dq, nq, cap, want_c, want_i)
    g.query_device_noalloc(dq, nq, cap, got_c, got_i)
    ctx.synchronize()
    wc, gc = want_c.download(np.uint32, nq), got_c.download(np.uint32, nq)
    wi = want_i.download(np.uint32, nq * cap).reshape(nq, cap)
    gi = got_i.download(np.uint32, nq * cap).reshape(nq, cap)
    assert np.array_equal(wc, gc)
    for k in range(nq):
        m = min(int(wc[k]), cap)
        assert np.array_equal(wi[k, :m], gi[k, :m]), k
    return wc


@pytest.mark.parametrize("n_cells", [1, 97, 4096])
def test_noalloc_query_equals_query(ctx, n_cells):
    rng = np.random.default_rng(7 + n_cells)
    pos = rng.uniform(-2, 3, (6000, 3)).astype(np.float32)
    pos[:2500] = 0.0  # a bucket above 2,048 points: its queries take the deferred (workgroup) path
    g = W.HashGrid.from_points(ctx, pos, 0.3, n_cells)
    q1 = np.concatenate([rng.uniform(-2.5, 3.5, (700, 3)), np.zeros((40, 3)),
                         pos[3000:3300] + rng.normal(0, 0.05, (300, 3))]).astype(np.float32)
    cap = 3000
    c1 = _both(ctx, g, q1, cap)
    assert c1.max() > 2048  # the deferred path ran
    # second call, fewer queries: the list's counter must restart at zero
    q2 = np.concatenate([np.zeros((3, 3)), rng.uniform(-2.5, 3.5, (60, 3))]).astype(np.float32)
    c2 = _both(ctx, g, q2, 64)
    assert c2[:3].min() > 64  # counts beyond the cap are reported, lists truncated


def test_noalloc_query_limits(ctx):
    g = W.HashGrid.from_points(ctx, np.zeros((4, 3), np.float32), 0.5, 8)
    dq = ctx.alloc(12)
    dc = ctx.alloc(4)
    with pytest.raises(W.WalkRayError):
        g.query_device_noalloc(dq, W.WR_HASHGRID_NOALLOC_MAX_QUERIES + 1, 4, dc, dc)
    g.query_device_noalloc(dq, 0, 4, dc, dc)  # nothing to do
