"""Writes tests/golden/field1080_ref_sub8.npz: the 512-object field workload
(scenes.random_field, 1920x1080, max depth 16) rendered to 65,536 spp on the
GPU, kept at every 8th pixel in x and y (240 x 135 pixels, the pixel centres
(8i + 4, 8j + 4)) as the reference that relative RMSE is stated against on the
same pixels.

The passes are rendered in batches on a Kahan FP32 film each and the batch
sums are accumulated in FP64 on the host.  The seed differs from the one the
benchmark renders with, so the reference shares no samples with the image it
is compared to.

Usage (on a machine with the GPU): python tests/golden/make_field_reference.py [spp]
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
STRIDE, OFFSET = 8, 4


def subsample(mean, width, height):
    """The reference's pixel subset of a row-major, top-row-first RGB image."""
    img = np.asarray(mean).reshape(height, width, 3)
    return img[OFFSET::STRIDE, OFFSET::STRIDE]


def main():
    spp = int(sys.argv[1]) if len(sys.argv) > 1 else 65536
    ctx = W.Context(0)
    objs, cam = S.random_field()
    sc = W.Scene(ctx, objs)
    w, h = cam["width"], cam["height"]
    n = 3 * w * h
    acc = np.zeros(n, np.float64)
    film = W.Film(ctx, w, h)
    zero = np.zeros(n, np.float32)
    t0 = time.time()
    for b in range(spp // BATCH):
        film.write({"sum": zero, "comp": zero, "sumsq": zero, "passes": b * BATCH})
        W.render(sc, cam, film, BATCH, 16, SEED, W.WR_WAVEFRONT)
        acc += film.read()["sum"].astype(np.float64)
        if b % 16 == 0:
            print(f"batch {b}: {(b + 1) * BATCH} spp, {time.time() - t0:.1f}s", flush=True)
    mean = subsample(acc / spp, w, h).astype(np.float32)
    out = os.path.join(ROOT, "tests", "golden", "field1080_ref_sub8.npz")
    np.savez_compressed(out, mean=mean, spp=np.int64(spp), seed=np.int64(SEED), stride=np.int64(STRIDE),
                        offset=np.int64(OFFSET))
    print(f"wrote {out}: {mean.shape}, mean radiance {mean.mean():.5f}, {time.time() - t0:.1f}s", flush=True)


if __name__ == "__main__":
    main()
