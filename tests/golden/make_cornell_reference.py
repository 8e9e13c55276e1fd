"""Writes tests/golden/cornell256_ref.npz: the Cornell-box workload (40,974
triangles, 256x256, max depth 8) rendered to 65,536 spp on the GPU, as the
reference image that relative RMSE is stated against.

The passes are rendered in batches on a Kahan FP32 film each and the batch
sums are accumulated in FP64 on the host.  The seed differs from the one the
benchmark renders with, so the reference shares no samples with the image it
is compared to.

Usage (on a machine with the GPU): python tests/golden/make_cornell_reference.py [spp]
"""
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from walkray_amd import _lib as W  # noqa: E402
from walkray_amd import scenes as S  # noqa: E402

SEED = 987654321
BATCH = 512


def main():
    spp = int(sys.argv[1]) if len(sys.argv) > 1 else 65536
    ctx = W.Context(0)
    objs, cam = S.cornell(k=5, width=256, height=256)
    sc = W.Scene(ctx, objs)
    n = 3 * 256 * 256
    acc = np.zeros(n, np.float64)
    film = W.Film(ctx, 256, 256)
    zero = np.zeros(n, np.float32)
    t0 = time.time()
    for b in range(spp // BATCH):
        film.write({"sum": zero, "comp": zero, "sumsq": zero, "passes": b * BATCH})
        W.render(sc, cam, film, BATCH, 8, SEED, W.WR_WAVEFRONT)
        acc += film.read()["sum"].astype(np.float64)
        if b % 16 == 0:
            print(f"batch {b}: {(b + 1) * BATCH} spp, {time.time() - t0:.1f}s", flush=True)
    mean = (acc / spp).astype(np.float32)
    out = os.path.join(ROOT, "tests", "golden", "cornell256_ref.npz")
    np.savez_compressed(out, mean=mean, spp=np.int64(spp), seed=np.int64(SEED))
    print(f"wrote {out}: mean radiance {mean.mean():.5f}, {time.time() - t0:.1f}s", flush=True)


if __name__ == "__main__":
    main()
