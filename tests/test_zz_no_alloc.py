"""ffs_get_field_noalloc (the real-space read-out that allocates no device memory) against
ffs_get_field(FFS_SPACE_REAL) through the C ABI: bitwise equal, on a 256 x 256 x 64 FP64 run (two
z-plane slabs through the read-out scratch) and then, on the same GPU, on a smaller 32^3 FP32 run
with density (the scratch still holds the first call's values), then the first run again."""
import ctypes

import numpy as np
import pytest

import ffs
from ffs import _lib

from conftest import print_case

pytestmark = pytest.mark.gpu


def check_fields(s, fields):
    for f in fields:
        want = s.get_field(f, _lib.FFS_SPACE_REAL)
        got = s.get_field_noalloc(f)
        assert got.shape == want.shape
        assert np.array_equal(got, want), (f, np.abs(got - want).max())
        # a caller-owned buffer is filled in place
        out = np.full(s.padded_size, np.nan)
        s.get_field_noalloc(f, out=out)
        assert np.array_equal(out.reshape(want.shape, order="F"), want)


def test_noalloc_read_out_matches_get_field():
    print_case("no-alloc read-out: 256x256x64 fp64 (2 slabs), then 32^3 fp32 + density, then the first again")
    pa = ffs.Params(nx=256, ny=256, nz=64, nu=2e-3, dt=1e-3, variable_dt=False, dt_stat=1 << 30, dt_spec=1 << 30,
                    write_time=1 << 30, integrator="Adams_Bashforth3rdO")
    pb = ffs.Params(nx=32, ny=32, nz=32, nu=5e-3, dt=2e-3, variable_dt=False, dt_stat=1 << 30, dt_spec=1 << 30,
                    write_time=1 << 30, integrator="Adams_Bashforth3rdO", precision="fp32", density=True, pr=1.0,
                    density_gradient=1.0, z_acceleration=-1.0, reference_density=1.0)
    with ffs.Simulation(pa) as a, ffs.Simulation(pb) as b:
        a.seed_velocity(5)
        b.seed_velocity(6)
        b.seed_density(7)
        a.initialize()
        b.initialize()
        a.advance(2)
        b.advance(3)
        check_fields(a, [_lib.FFS_FIELD_U1, _lib.FFS_FIELD_U3])
        check_fields(b, [_lib.FFS_FIELD_U2, _lib.FFS_FIELD_RHO])
        check_fields(a, [_lib.FFS_FIELD_U2])
        # the read-out leaves the run's state alone
        u = a.velocity_hat()
        a.get_field_noalloc(_lib.FFS_FIELD_U1)
        assert all(np.array_equal(x, y) for x, y in zip(u, a.velocity_hat()))
        small = np.zeros(8)
        with pytest.raises(_lib.FfsError, match="2\\*\\(nx/2\\+1\\)\\*ny\\*nz"):
            a._check(a._lib.ffs_get_field_noalloc(a._h, _lib.FFS_FIELD_U1,
                                                  small.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), 8))
        with pytest.raises(_lib.FfsError, match="state fields"):
            a.get_field_noalloc(_lib.FFS_FIELD_T11)
        with pytest.raises(_lib.FfsError, match="no density"):
            a.get_field_noalloc(_lib.FFS_FIELD_RHO)
