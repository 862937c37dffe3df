"""Seeded Gaussian random fields for synthetic grid-defined jet models."""
import numpy as np


def gaussian_random_field(shape, seed, spectral_index=-11.0 / 3.0):
    """Zero-mean, unit-variance random field on a regular grid with power
    spectrum P(k) ~ k^spectral_index (white noise filtered in Fourier space)."""
    rng = np.random.default_rng(seed)
    white = rng.standard_normal(shape)
    k0 = np.fft.fftfreq(shape[0])
    k1 = np.fft.rfftfreq(shape[1])
    K = np.sqrt(k0[:, None] ** 2 + k1[None, :] ** 2)
    K[0, 0] = np.inf
    axes = tuple(range(len(shape)))
    f = np.fft.irfftn(np.fft.rfftn(white, axes=axes) * K ** (spectral_index / 2.0), s=shape, axes=axes)
    f -= f.mean()
    s = f.std()
    return f / s if s > 0 else f


def jet_grids(nz, nr, z1_pc=20.0, r1_pc=0.5, seeds=(1, 2, 3), sigma=0.3, b1=0.1, n1=500.0, gamma=10.0,
              r_floor_pc=0.05):
    """B [G], N [cm^-3], Gamma node grids [nz, nr] over z in [0, z1], r in
    [0, r1] (parsec): log F = log F_conical(r_sph) + sigma G with the conical
    profiles B = b1 (r/pc)^-1, N = n1 (r/pc)^-2, Gamma constant (clamped >= 1),
    r_sph = max(hypot(z, r), r_floor).  Returns (B, N, Gamma), (z0, dz, r0, dr)."""
    z = np.linspace(0.0, z1_pc, nz)
    r = np.linspace(0.0, r1_pc, nr)
    Z, R = np.meshgrid(z, r, indexing="ij")
    rs = np.maximum(np.hypot(Z, R), r_floor_pc)
    B = b1 / rs * np.exp(sigma * gaussian_random_field((nz, nr), seeds[0]))
    N = n1 / rs ** 2 * np.exp(sigma * gaussian_random_field((nz, nr), seeds[1]))
    G = np.maximum(1.0, gamma * np.exp(sigma * gaussian_random_field((nz, nr), seeds[2])))
    return (B, N, G), (0.0, z[1] - z[0], 0.0, r[1] - r[0])
