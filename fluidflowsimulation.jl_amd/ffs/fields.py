"""Field files: raw little-endian binary64 of the padded real buffer.

`write("u1.$it", s.u.rr.x)` (output.jl:6-8) writes the PaddedArray
(2*(nx/2+1), ny, nz) column-major with no header; the FluidFields reader
(un-vendored) is taken to accept the padded and the unpadded (nx, ny, nz)
layouts, told apart by the file size.
"""
import os

import numpy as np

__all__ = ["padded_shape", "pad", "unpad", "read_field", "write_field"]


def padded_shape(nx, ny, nz):
    """Fortran-order shape of the padded real view."""
    return (2 * (nx // 2 + 1), ny, nz)


def pad(a):
    """(nx, ny, nz) real array -> padded (2*(nx//2+1), ny, nz), padding zeroed."""
    nx, ny, nz = a.shape
    out = np.zeros(padded_shape(nx, ny, nz), dtype=np.float64, order="F")
    out[:nx] = a
    return out


def unpad(a, nx):
    return np.asfortranarray(a[:nx])


def read_field(path, nx, ny, nz):
    """Returns the padded (Fortran-ordered) array from a padded or unpadded file."""
    n = os.path.getsize(path) // 8
    data = np.fromfile(path, dtype="<f8")
    if n == 2 * (nx // 2 + 1) * ny * nz:
        return data.reshape(padded_shape(nx, ny, nz), order="F")
    if n == nx * ny * nz:
        return pad(data.reshape((nx, ny, nz), order="F"))
    raise ValueError(f"{path}: {n} doubles match neither the padded nor the unpadded {nx}x{ny}x{nz} layout")


def write_field(path, padded):
    np.asarray(padded, dtype="<f8").ravel(order="F").tofile(path)
