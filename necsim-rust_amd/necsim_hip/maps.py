"""Seeded synthetic spatially-explicit landscape (habitat + sparse dispersal map).

Habitat: every cell uniform in 0..=16 (SplitMix64 of the seed and the cell
index, modulo 17), then exactly round(zero_fraction * cells) cells, chosen by
a second seeded hash, forced to 0.
Dispersal: for each habitable origin, a truncated Gaussian kernel
w = exp(-d^2 / (2 sigma^2)) over habitable targets with centre distance
d <= cutoff (no wrap).  Weights are quantised to multiples of 2^-30 so that
the map is bit-identical on every platform whatever libm rounds exp to.
Rows list targets in ascending column order, i.e. exactly the positive
entries that the reference's dense Array2D row iteration visits
(packed_separable_alias/mod.rs:94-118).
"""
import math

import numpy as np

M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(x):
    """SplitMix64 finaliser on a uint64 array (wrapping arithmetic)."""
    with np.errstate(over="ignore"):
        z = x + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def kernel_weights(sigma, cutoff):
    """w[k] for squared distances k = 0..floor(cutoff^2), quantised to 2^-30."""
    kmax = int(math.floor(cutoff * cutoff))
    w = np.zeros(kmax + 1)
    for k in range(kmax + 1):
        v = math.exp(-k / (2.0 * sigma * sigma)) * 2.0**30
        frac = v - math.floor(v)
        assert abs(frac - 0.5) > 1e-6, "quantisation too close to a tie"
        w[k] = math.floor(v + 0.5) / 2.0**30
    return w


def generate_map(width, height, seed=42, zero_fraction=0.2, sigma=3.0, cutoff=12.0):
    """Returns (habitat[u32, H*W], row_ptr[u64, H*W+1], col[u32], weight[f64])."""
    cells = width * height
    idx = np.arange(cells, dtype=np.uint64)
    base = splitmix64(np.array([seed], dtype=np.uint64))[0]
    habitat = (splitmix64(idx ^ base) % np.uint64(17)).astype(np.uint32)
    n_zero = int(round(zero_fraction * cells))
    order = np.lexsort((idx, splitmix64(idx + splitmix64(np.array([base ^ np.uint64(1)]))[0])))
    habitat[order[:n_zero]] = 0

    wtab = kernel_weights(sigma, cutoff)
    r = int(math.floor(cutoff))
    offs = [(dy, dx) for dy in range(-r, r + 1) for dx in range(-r, r + 1)
            if dx * dx + dy * dy <= cutoff * cutoff]  # ascending (dy, dx) -> ascending column
    ys, xs = np.divmod(np.arange(cells, dtype=np.int64), width)
    hab_ok = habitat > 0
    valid_cols, valid_w, valid_mask = [], [], []
    for dy, dx in offs:
        ty, tx = ys + dy, xs + dx
        inb = (ty >= 0) & (ty < height) & (tx >= 0) & (tx < width)
        t = np.where(inb, ty * width + tx, 0)
        ok = inb & hab_ok & hab_ok[t]
        valid_mask.append(ok)
        valid_cols.append(t.astype(np.uint32))
        valid_w.append(wtab[dy * dy + dx * dx])
    mask = np.stack(valid_mask, axis=1)            # [cells, n_off], row-major = CSR order
    cols = np.stack(valid_cols, axis=1)[mask]
    weight = np.broadcast_to(np.array(valid_w), mask.shape)[mask].astype(np.float64)
    counts = mask.sum(axis=1)
    row_ptr = np.zeros(cells + 1, dtype=np.uint64)
    row_ptr[1:] = np.cumsum(counts)
    return habitat, row_ptr, cols.astype(np.uint32), weight
