"""Host twin of the device's seeded random-phase initial fields (TEST INFRASTRUCTURE ONLY).

The reference has no field generator (a case loads u1.0 ... written
elsewhere), so these fields are this project's own, defined here and in the
HIP kernel `k_seed` alike:

* mode (ix, iy, iz) (signed integer indices, FFT order) of component c draws
  xi = (re, im) from two SplitMix64 hashes of (seed, c, ix, iy, iz): with
  s = mix(seed) and ctr = ((c+1) << 48) | (ix & 0xffff) << 32 | (iy & 0xffff) << 16 | (iz & 0xffff),
  h1 = mix(s ^ mix(2 ctr)), h2 = mix(s ^ mix(2 ctr + 1)), u_j = ((h_j >> 11) + 1/2) 2^-53,
  re + i im = sqrt(-2 ln u1) exp(2 pi i u2)  (Box-Muller);
* on the kx = 0 plane the half with iy < 0, or iy = 0 and iz < 0, takes the
  conjugate of its mirror mode, so the field is real;
* velocity: u_hat = A(|k|) (I - k k / |k|^2) xi, scalar: A(|k|) xi, with
  A(k) = k exp(-(k/k0)^2) / sqrt(4 pi) (shell energy ~ k^4 exp(-2 (k/k0)^2)),
  zero at k = 0 and on the dealiased modes;
* scaled so that 1/2 sum_w |u_hat|^2 = energy (velocity) or sum_w |rho_hat|^2 = rms^2
  (w = 2 for kx > 0): the Parseval means of the physical field.

SplitMix64 is Vigna's published mixer (the finaliser of java.util.SplittableRandom).
"""
import numpy as np

__all__ = ["mix64", "gauss", "seeded_velocity_hat", "seeded_scalar_hat"]

_U = np.uint64


def mix64(z):
    """SplitMix64 output function applied to the state z (uint64 array): z + golden gamma, then the finaliser."""
    z = np.asarray(z, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = z + _U(0x9E3779B97F4A7C15)
        z = (z ^ (z >> _U(30))) * _U(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> _U(27))) * _U(0x94D049BB133111EB)
    return z ^ (z >> _U(31))


def gauss(seed, c, ix, iy, iz):
    """Complex Gaussian per mode (broadcast over the index arrays)."""
    m16 = np.int64(0xFFFF)
    ctr = ((np.int64(c + 1) << 48) | ((np.asarray(ix, dtype=np.int64) & m16) << 32)
           | ((np.asarray(iy, dtype=np.int64) & m16) << 16) | (np.asarray(iz, dtype=np.int64) & m16))
    ctr = ctr.astype(np.uint64)
    s = mix64(_U(int(seed) & (2**64 - 1)))
    with np.errstate(over="ignore"):
        h1 = mix64(s ^ mix64(_U(2) * ctr))
        h2 = mix64(s ^ mix64(_U(2) * ctr + _U(1)))
    u1 = ((h1 >> _U(11)).astype(np.float64) + 0.5) * 2.0**-53
    u2 = ((h2 >> _U(11)).astype(np.float64) + 0.5) * 2.0**-53
    r = np.sqrt(-2.0 * np.log(u1))
    th = 6.283185307179586 * u2
    return r * np.cos(th), r * np.sin(th)


def _indices(grid):
    nx, ny, nz = grid.NRX, grid.NY, grid.NZ
    ix = np.arange(nx // 2 + 1)[:, None, None]
    jy = np.arange(ny)
    jz = np.arange(nz)
    iy = np.where(jy < (ny + 1) // 2, jy, jy - ny)[None, :, None]
    iz = np.where(jz < (nz + 1) // 2, jz, jz - nz)[None, None, :]
    ix, iy, iz = np.broadcast_arrays(ix, iy, iz)
    conj = (ix == 0) & ((iy < 0) | ((iy == 0) & (iz < 0)))
    return ix, np.where(conj, -iy, iy), np.where(conj, -iz, iz), conj


def _amp(grid, k0):
    k2 = np.broadcast_to(grid.K2, (grid.NX, grid.NY, grid.NZ))
    k = np.sqrt(k2)
    with np.errstate(invalid="ignore", divide="ignore"):
        a = np.where(k2 > 0, k * np.exp(-(k / k0) * (k / k0)) / np.sqrt(4.0 * 3.141592653589793), 0.0)
    return np.where(grid.DEALIAS, 0.0, a), k2


def seeded_velocity_hat(grid, seed, k0=4.0, energy=0.5):
    """[u1_hat, u2_hat, u3_hat], each (nx/2+1, ny, nz) complex, as the device generates them."""
    ix, iy, iz, conj = _indices(grid)
    amp, k2 = _amp(grid, k0)
    xr, xi = [], []
    for c in range(3):
        a, b = gauss(seed, c, ix, iy, iz)
        xr.append(a)
        xi.append(np.where(conj, -b, b))
    kv = [np.broadcast_to(k, k2.shape) for k in (grid.kx, grid.ky, grid.kz)]
    with np.errstate(invalid="ignore", divide="ignore"):
        pr = np.where(k2 > 0, (kv[0] * xr[0] + kv[1] * xr[1] + kv[2] * xr[2]) / k2, 0.0)
        pi = np.where(k2 > 0, (kv[0] * xi[0] + kv[1] * xi[1] + kv[2] * xi[2]) / k2, 0.0)
    u = [amp * ((xr[c] - pr * kv[c]) + 1j * (xi[c] - pi * kv[c])) for c in range(3)]
    tot = sum(float(np.sum(grid.W * np.abs(x) ** 2)) for x in u)
    f = np.sqrt(energy / (0.5 * tot))
    return [f * x for x in u]


def seeded_scalar_hat(grid, seed, k0=4.0, rms=0.01):
    """rho_hat (nx/2+1, ny, nz) complex, as the device generates it."""
    ix, iy, iz, conj = _indices(grid)
    amp, _ = _amp(grid, k0)
    a, b = gauss(seed, 3, ix, iy, iz)
    r = amp * (a + 1j * np.where(conj, -b, b))
    tot = float(np.sum(grid.W * np.abs(r) ** 2))
    return r * np.sqrt(rms * rms / tot)
