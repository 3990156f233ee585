"""Seeded synthetic inputs for the benchmark workloads.

All randomness comes from a counter-based generator (SplitMix64 of
``(seed, stream, index)``), so every array is a pure function of its seed and
identical wherever it is generated.  These generators have no counterpart in
the reference (which reads uniform properties per layer, VFCommon.c:506-588);
they produce the heterogeneous inputs the workloads call for:

* ``lognormal_E``: E_e = exp(m + s g_e), m = -s^2/2 (so E[E] = 1); g a
  unit-variance Gaussian random field with covariance exp(-r^2/(2 l^2)):
  white noise convolved with a separable Gaussian of sigma_f = l/sqrt(2)
  cells, truncated at 4 sigma_f, noise padded by the truncation width.
* ``random_pennies``: centres U[0.3,0.7]^3, normals uniform on the sphere
  (theta = 360 u1, phi = acos(1 - 2 u2) in degrees, VFCracks.c:109-111),
  radii U[r0, r1], thickness 3h.
* ``smooth_field``: the same GRF on nodes, affinely mapped to [lo, hi].
"""
import math

import numpy as np
from scipy.ndimage import convolve1d

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(x):
    x = (x + np.uint64(0x9E3779B97F4A7C15)) & _M64
    z = x
    z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & _M64
    z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & _M64
    return z ^ (z >> np.uint64(31))


def uniform(seed, stream, n, offset=0):
    """n doubles in [0,1) for key (seed, stream), indices offset..offset+n-1."""
    with np.errstate(over="ignore"):
        key = splitmix64(np.uint64(seed) ^ splitmix64(np.uint64(stream) + np.uint64(0x632BE59BD9B4E019)))
        idx = np.arange(offset, offset + n, dtype=np.uint64)
        bits = splitmix64(key ^ splitmix64(idx))
    return (bits >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def normal(seed, stream, n):
    """Standard normals by Box-Muller on pairs of counter uniforms."""
    m = (n + 1) // 2
    u1 = uniform(seed, stream, m, 0)
    u2 = uniform(seed, stream, m, m)
    r = np.sqrt(-2.0 * np.log1p(-u1))
    z = np.concatenate([r * np.cos(2 * math.pi * u2), r * np.sin(2 * math.pi * u2)])
    return z[:n]


def grf(shape, corr_len, seed, stream):
    """Unit-variance Gaussian random field on a grid of `shape` (z, y, x)."""
    sf = corr_len / math.sqrt(2.0)
    half = int(math.ceil(4.0 * sf))
    t = np.arange(-half, half + 1, dtype=np.float64)
    k = np.exp(-0.5 * (t / sf) ** 2)
    k /= math.sqrt((k * k).sum())  # unit variance per separable pass
    pad = tuple(s + 2 * half for s in shape)
    w = normal(seed, stream, int(np.prod(pad))).reshape(pad)
    for ax in range(3):
        w = convolve1d(w, k, axis=ax, mode="constant")
    return np.ascontiguousarray(w[half:half + shape[0], half:half + shape[1], half:half + shape[2]])


def lognormal_E(n_elem, sigma, corr_len, seed):
    """Per-element Young's modulus (N_e, DMDA cell order)."""
    nex, ney, nez = n_elem
    g = grf((nez, ney, nex), corr_len, seed, stream=1)
    m = -0.5 * sigma * sigma
    return np.exp(m + sigma * g).reshape(-1)


def smooth_field(n_nodes, corr_len, lo, hi, seed, stream=2):
    nx, ny, nz = n_nodes
    g = grf((nz, ny, nx), corr_len, seed, stream).reshape(-1)
    gmin, gmax = g.min(), g.max()
    return lo + (hi - lo) * (g - gmin) / (gmax - gmin if gmax > gmin else 1.0)


def random_pennies(count, h, seed, r_range=(0.1, 0.25), stream=3):
    u = uniform(seed, stream, 6 * count).reshape(count, 6)
    cracks = []
    for i in range(count):
        c = 0.3 + 0.4 * u[i, :3]
        theta = 360.0 * u[i, 3]
        phi = math.degrees(math.acos(1.0 - 2.0 * u[i, 4]))
        r = r_range[0] + (r_range[1] - r_range[0]) * u[i, 5]
        cracks.append(dict(center=tuple(float(x) for x in c), r=float(r), theta=theta, phi=phi,
                           thickness=3.0 * h))
    return cracks
