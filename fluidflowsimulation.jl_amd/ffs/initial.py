"""Initial fields (host-side data generators; the reference has none built in).

random_spectrum: seeded random-phase solenoidal velocity with shell spectrum
E(k) ~ k^4 exp(-2 (k/k0)^2), dealiased, scaled to 1/2<|u|^2> = energy.
random_scalar: Gaussian random scalar with the same spectral shape, given rms.
taylor_green: u1 = cos x sin y sin z, u2 = -sin x cos y sin z, u3 = 0.
All return real arrays on the grid x_i = 2 pi L i / n (index order x, y, z).
"""
import os

import numpy as np
import scipy.fft

__all__ = ["taylor_green", "random_spectrum", "random_scalar"]


def _coords(nx, ny, nz, lx, ly, lz):
    x = 2 * np.pi * lx * np.arange(nx) / nx
    y = 2 * np.pi * ly * np.arange(ny) / ny
    z = 2 * np.pi * lz * np.arange(nz) / nz
    return np.meshgrid(x, y, z, indexing="ij")


def taylor_green(nx, ny=None, nz=None, lx=1.0, ly=1.0, lz=1.0):
    ny = ny or nx
    nz = nz or nx
    x, y, z = _coords(nx, ny, nz, lx, ly, lz)
    return np.array([np.cos(x) * np.sin(y) * np.sin(z), -np.sin(x) * np.cos(y) * np.sin(z), np.zeros_like(x)])


def _kgrid(nx, ny, nz, lx, ly, lz):
    kx = np.fft.rfftfreq(nx, 1.0 / nx) / lx
    ky = np.fft.fftfreq(ny, 1.0 / ny) / ly
    kz = np.fft.fftfreq(nz, 1.0 / nz) / lz
    return kx[:, None, None], ky[None, :, None], kz[None, None, :]


def _dealias_mask(nx, ny, nz, lx, ly, lz, cutoff):
    kx, ky, kz = _kgrid(nx, ny, nz, lx, ly, lz)
    return ((kx**2 > (cutoff * kx.max()) ** 2) | (ky**2 > (cutoff * np.abs(ky).max()) ** 2)
            | (kz**2 > (cutoff * np.abs(kz).max()) ** 2))


_WORKERS = max(1, min(16, os.cpu_count() or 1))


def _rfft(a):
    return scipy.fft.rfftn(a, axes=(2, 1, 0), workers=_WORKERS)


def _irfft(c, shape):
    return scipy.fft.irfftn(c, s=shape[::-1], axes=(2, 1, 0), workers=_WORKERS)


def random_spectrum(nx, ny=None, nz=None, seed=0, k0=4.0, energy=0.5, lx=1.0, ly=1.0, lz=1.0, cutoff=2 / 3):
    """Seeded solenoidal random field: white noise filtered to amplitude
    sqrt(E(k) / (4 pi k^2)), Leray-projected, dealiased, energy-normalised."""
    ny = ny or nx
    nz = nz or nx
    rng = np.random.default_rng(seed)
    noise = rng.standard_normal((3, nx, ny, nz))
    kx, ky, kz = _kgrid(nx, ny, nz, lx, ly, lz)
    k2 = kx * kx + ky * ky + kz * kz
    k = np.sqrt(k2)
    with np.errstate(divide="ignore", invalid="ignore"):
        amp = np.where(k > 0, np.sqrt(k**4 * np.exp(-2 * (k / k0) ** 2) / (4 * np.pi * k2)), 0.0)
    mask = _dealias_mask(nx, ny, nz, lx, ly, lz, cutoff)
    uh = np.array([_rfft(c) * amp for c in noise])
    with np.errstate(divide="ignore", invalid="ignore"):
        p = np.where(k2 > 0, (kx * uh[0] + ky * uh[1] + kz * uh[2]) / k2, 0)
    uh = np.array([uh[0] - p * kx, uh[1] - p * ky, uh[2] - p * kz])
    uh[:, mask] = 0
    uh[:, 0, 0, 0] = 0
    u = np.array([_irfft(c, (nx, ny, nz)) for c in uh])
    e = 0.5 * np.mean(np.sum(u * u, axis=0))
    return u * np.sqrt(energy / e)


def random_scalar(nx, ny=None, nz=None, seed=1, k0=4.0, rms=0.01, lx=1.0, ly=1.0, lz=1.0, cutoff=2 / 3):
    """Seeded Gaussian random scalar with spectrum ~ k^4 exp(-2 (k/k0)^2), dealiased, given rms."""
    ny = ny or nx
    nz = nz or nx
    rng = np.random.default_rng(seed)
    noise = rng.standard_normal((nx, ny, nz))
    kx, ky, kz = _kgrid(nx, ny, nz, lx, ly, lz)
    k2 = kx * kx + ky * ky + kz * kz
    k = np.sqrt(k2)
    with np.errstate(divide="ignore", invalid="ignore"):
        amp = np.where(k > 0, np.sqrt(k**4 * np.exp(-2 * (k / k0) ** 2) / (4 * np.pi * k2)), 0.0)
    ch = _rfft(noise) * amp
    ch[_dealias_mask(nx, ny, nz, lx, ly, lz, cutoff)] = 0
    ch[0, 0, 0] = 0
    r = _irfft(ch, (nx, ny, nz))
    return r * (rms / np.sqrt(np.mean(r * r)))
