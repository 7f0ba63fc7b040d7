"""Random-projection sketches of long result vectors, for golden fixtures.

A sketch of a vector v is S v, where the 64 rows of S are Rademacher (+-1)
vectors drawn from numpy's default_rng with fixed seeds, divided by sqrt(64).
E ||S v||^2 = ||v||^2, so ||S a - S b|| / ||S b|| estimates the relative L2
distance ||a - b|| / ||b|| within about 20% (1 sigma ~ sqrt(2/64)): enough to
hold a result to a tolerance from a fixture of 64 numbers instead of the
vector itself.  Used by tests/golden/make_sketches.py (oracle side) and the
large-case GPU tests (device side) with the same seeds.
"""
import numpy as np

NPROBES = 64
SEED = 20260531


def sketch(v, seed=SEED, k=NPROBES):
    v = np.ascontiguousarray(v, dtype=np.float64).ravel()
    out = np.empty(k)
    for j in range(k):
        r = np.random.default_rng(seed + j).integers(0, 2, size=v.size, dtype=np.int8)
        out[j] = np.dot(v, r.astype(np.float64) * 2.0 - 1.0)
    return out / np.sqrt(k)


def rel_sketch(sa, sb):
    """Estimated relative L2 distance of the vectors whose sketches are sa, sb."""
    sa, sb = np.asarray(sa), np.asarray(sb)
    return float(np.linalg.norm(sa - sb) / np.linalg.norm(sb))


def smooth_direction(shape, seed, sigma, margin):
    """A smooth random field: Gaussian-filtered white noise, zero within
    `margin` cells of the edges (the C-PML and the receiver rows), largest
    |value| 1."""
    from scipy.ndimage import gaussian_filter

    rng = np.random.default_rng(seed)
    f = gaussian_filter(rng.standard_normal(shape), sigma)
    mz, mx = margin
    w = np.zeros(shape)
    w[mz[0]:shape[0] - mz[1], mx:shape[1] - mx] = 1.0
    w = gaussian_filter(w, sigma / 2.0)
    f = f * w
    return f / np.abs(f).max()
