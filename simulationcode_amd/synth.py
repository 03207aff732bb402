"""Seeded synthetic inputs of the benchmark configurations.

Positions: simple-cubic (or square) lattice filling the box, uniform jitter,
wrapped into [0, L).  Diameters: the polydisperse recipe of
generate_diameters! (src/InitialConfiguration.jl:1-11): regular quantiles of
P(sigma) ~ sigma^-3 on [sigma_min, sigma_ratio*sigma_min] through the inverse
CDF, then a seeded shuffle.  Velocities: Maxwell-Boltzmann at kBT/m with the
centre-of-mass velocity removed (src/MD.jl:683-686 draws randn!*sqrt(kBT/m)).
"""
import math

import numpy as np


def box_size(N, rho, dims):
    """(N / rho)^(1/dims), as the reference scripts compute it."""
    return (N / rho) ** (1.0 / dims)


def lattice_positions(N, dims, L, jitter, rng):
    n_side = int(math.ceil(N ** (1.0 / dims) - 1e-9))
    a = L / n_side
    idx = np.arange(N)
    coords = []
    for c in range(dims):
        coords.append((idx // (n_side ** c)) % n_side)
    base = (np.stack(coords, axis=1).astype(np.float64) + 0.5) * a
    r = base + rng.uniform(-jitter, jitter, size=(N, dims))
    r -= np.floor(r / L) * L
    return np.ascontiguousarray(r)


def berthier_diameters(N, rng, sigma_ratio=2.219):
    """src/InitialConfiguration.jl:1-11 (LinRange quantiles, then shuffle)."""
    A = 0.5 + 1.0 / (sigma_ratio - 1.0)
    smin = (1.0 + sigma_ratio) / (2.0 * sigma_ratio)
    P = np.linspace(1.0 / (2 * N), 1.0 - 1.0 / (2 * N), N)
    D = np.sqrt(1.0 / (smin ** -2 - 2.0 * P / A))
    return np.ascontiguousarray(rng.permutation(D))


def maxwell_boltzmann(N, dims, kBT, m, rng):
    v = rng.standard_normal((N, dims)) * math.sqrt(kBT / m)
    v -= v.mean(axis=0)
    return np.ascontiguousarray(v)
