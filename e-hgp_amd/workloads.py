"""Seeded synthetic point clouds for the benchmark and test configurations.

All clouds are binary64, generated from splitmix64 streams, and free of
duplicate points (a repeated point is redrawn).
"""

from __future__ import annotations

import numpy as np

_MASK = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed: int):
        self.state = seed & _MASK

    def next(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & _MASK
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _MASK
        return z ^ (z >> 31)

    def uniform(self, count: int) -> np.ndarray:
        out = np.empty(count, dtype=np.float64)
        for i in range(count):
            out[i] = (self.next() >> 11) * (1.0 / (1 << 53))
        return out

    def numpy_rng(self) -> np.random.Generator:
        return np.random.default_rng(self.next())


def _dedupe(points: np.ndarray, redraw) -> np.ndarray:
    while True:
        view = np.ascontiguousarray(points).view(np.dtype((np.void, 24))).ravel()
        _, first = np.unique(view, return_index=True)
        if len(first) == len(points):
            return points
        dup = np.setdiff1d(np.arange(len(points)), first)
        points[dup] = redraw(len(dup))


def uniform_cloud(n: int, seed: int) -> np.ndarray:
    """n points i.i.d. uniform in [0,1)^3."""
    rng = SplitMix64(seed).numpy_rng()
    pts = rng.random((n, 3))
    return _dedupe(pts, lambda m: rng.random((m, 3)))


def jittered_lattice(n: int = 50_000, side: int = 37, sigma: float = 0.1, seed: int = 0,
                     spacing: float = 1.0) -> np.ndarray:
    """Seeded subsample of a side^3 cubic lattice, each site displaced by
    N(0, (sigma * spacing)^2) per axis."""
    rng = SplitMix64(seed).numpy_rng()
    sites = side ** 3
    if n > sites:
        raise ValueError("more points than lattice sites")
    keep = rng.choice(sites, size=n, replace=False)
    idx = np.stack([keep // (side * side), (keep // side) % side, keep % side], axis=1)
    pts = idx.astype(np.float64) * spacing + rng.normal(0.0, sigma * spacing, size=(n, 3))
    return _dedupe(pts, lambda m: rng.normal(0.0, sigma * spacing, size=(m, 3)) + side * spacing / 2)


def _haar_rotation(rng: np.random.Generator) -> np.ndarray:
    q, r = np.linalg.qr(rng.normal(size=(3, 3)))
    q = q * np.sign(np.diag(r))
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    return q


def gaussian_mixture(n: int = 200_000, blobs: int = 20, noise_fraction: float = 0.15,
                     box: float = 10.0, snap_bits: int | None = 12, seed: int = 0,
                     eig_range=(0.01, 1.0)) -> np.ndarray:
    """(1-noise) of the points from `blobs` anisotropic Gaussians (means uniform
    in [0, box]^3, covariance R diag(lambda) R^T with lambda log-uniform in
    eig_range and R Haar-random), the rest uniform in [0, box]^3; coordinates
    snapped to the 2^-snap_bits grid when snap_bits is set."""
    rng = SplitMix64(seed).numpy_rng()
    n_blob = int(round(n * (1.0 - noise_fraction)))
    means = rng.random((blobs, 3)) * box
    lo, hi = np.log(eig_range[0]), np.log(eig_range[1])
    comps = []
    for _ in range(blobs):
        lam = np.exp(rng.uniform(lo, hi, size=3))
        comps.append(_haar_rotation(rng) * np.sqrt(lam))
    which = rng.integers(0, blobs, size=n_blob)

    def draw_blob(idx):
        z = rng.normal(size=(len(idx), 3))
        return means[idx] + np.einsum("nij,nj->ni", np.stack([comps[i] for i in idx]), z)

    pts = np.concatenate([draw_blob(which), rng.random((n - n_blob, 3)) * box])

    def snap(a):
        if snap_bits is None:
            return a
        s = float(1 << snap_bits)
        return np.round(a * s) / s

    pts = snap(pts)
    return _dedupe(pts, lambda m: snap(rng.random((m, 3)) * box))


def batch_sizes(count: int = 256, lo: int = 1_000, hi: int = 50_000, seed: int = 0) -> np.ndarray:
    rng = SplitMix64(seed).numpy_rng()
    return np.exp(rng.uniform(np.log(lo), np.log(hi), size=count)).astype(np.int64)


def batch_cloud(size: int, seed: int) -> np.ndarray:
    """Either a jittered lattice or a 5-20 blob mixture, chosen by seed."""
    rng = SplitMix64(seed).numpy_rng()
    if rng.random() < 0.5:
        side = int(np.ceil((size * 1.3) ** (1 / 3)))
        return jittered_lattice(size, side=side, sigma=0.1, seed=seed)
    return gaussian_mixture(size, blobs=int(rng.integers(5, 21)), noise_fraction=0.0,
                            box=10.0, snap_bits=None, seed=seed)
