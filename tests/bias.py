"""Per-seed tile statistics shared by the bias tests."""

import numpy as np

TILES = 4


def tile_means(img, tiles=TILES):
    """Per-channel means of tiles x tiles image tiles (row-major), with the integer tile
    bounds the reference harness's seeds command uses."""
    h, w = img.shape[:2]
    out = np.zeros((tiles * tiles, 3))
    for ty in range(tiles):
        for tx in range(tiles):
            t = img[ty * h // tiles:(ty + 1) * h // tiles, tx * w // tiles:(tx + 1) * w // tiles].astype(np.float64)
            out[ty * tiles + tx] = t.reshape(-1, 3).mean(0)
    return out


def welch_z(a, b, floor=1e-5):
    """z of mean(a) - mean(b) along axis 0 for independent samples. The standard error has
    a floor of `floor` times the mean magnitude, so a region that is (nearly) constant on
    both sides, such as a background of one value, is compared at fp32 resolution rather
    than dividing by a zero variance."""
    va, vb = a.var(0, ddof=1) / len(a), b.var(0, ddof=1) / len(b)
    scale = floor * np.maximum(np.abs(a.mean(0)), np.abs(b.mean(0)))
    den = np.sqrt(va + vb + scale ** 2)
    return np.where(den > 0, (a.mean(0) - b.mean(0)) / np.where(den > 0, den, 1.0), 0.0)
