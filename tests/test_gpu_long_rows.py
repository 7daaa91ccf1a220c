"""Parity at long transform lengths on thin grids: every compiled FFT length of
the x pass (nx up to 1024) and of the column passes (ny, nz up to 1024) against the
oracle, for the plain, LES and Boussinesq fields, FP64 and FP32, over three
steps.  The thin grids keep the oracle cheap while each transform length runs
the same kernel variant as a cubic grid of that size.  The step is fixed and
small: set_dt's CFL uses the y spacing only (time_step_functions.jl:86-129),
so on a thin grid it would pick a dt far beyond the explicit viscous limit of
the long axis, and the runs would amplify roundoff instead of comparing it."""
import numpy as np
import pytest

import ffs
from ffs.initial import random_scalar, random_spectrum
from oracle.ffs_oracle import Grid, OracleSim

from conftest import print_case

pytestmark = pytest.mark.gpu

# (nx, ny, nz, x_domain): the horizontal spectrum bins ky on the x wavenumbers unless kx is the finer
# grid (Globals.jl:148-152), so the long-y case takes a longer x box
# (the 256 and 128 column lengths are the ones the 256^3 and 128^3 runs use; on plain FP64 steps the z
# length also selects the fused forward-z + update kernel of that length)
SHAPES = [(512, 16, 16, 1.0), (256, 16, 16, 1.0), (16, 512, 16, 2.0), (16, 16, 512, 1.0), (192, 16, 16, 1.0),
          (96, 24, 16, 1.0), (16, 256, 16, 2.0), (16, 16, 256, 1.0), (128, 16, 16, 1.0), (16, 128, 16, 2.0),
          (16, 16, 128, 1.0), (16, 192, 16, 2.0), (16, 16, 192, 1.0), (16, 96, 16, 2.0), (16, 16, 96, 1.0),
          (16, 48, 16, 2.0), (16, 16, 48, 1.0)]
# the radix-5 lengths and 384 / 768 (2^k 3), along each axis
NEW_LENGTHS = [40, 80, 160, 320, 384, 640, 768, 1024]
SHAPES += [(n, 16, 16, 1.0) for n in NEW_LENGTHS] + [(16, n, 16, 2.0) for n in NEW_LENGTHS] + \
    [(16, 16, n, 1.0) for n in NEW_LENGTHS]
MODELS = ["plain", "les", "density"]


DT = 1e-3


def stable_nu(nx, ny, nz, lx, nu=2e-3):
    """nu = 2e-3, or less where the longest axis's highest retained mode would pass explicit AB3's viscous
    limit at dt = 1e-3 (nu k^2 dt < ~0.545): beyond it AB3 amplifies roundoff by orders of magnitude in three
    steps, and FP32 runs (LES most, whose stress fills the high modes) then differ from the FP64 oracle by that
    noise.  dt stays 1e-3, so the nonlinear term moves the field as much as on the short axes."""
    kmax = max(2 * np.pi * (n // 3) / L for n, L in ((nx, lx), (ny, 1.0), (nz, 1.0)))
    return min(nu, 0.25 / (DT * kmax**2))


def les_dt(nx, ny, nz):
    """The LES stress is explicit: its eddy viscosity (c Delta)^2 |S| with the Delta of a thin, anisotropic grid
    passes the explicit limit on the high modes of an axis longer than 512, which then amplify roundoff in both
    precisions alike (measured at y length 768, dt = 1e-3: FP64 1.1e-12, FP32 1.3e-4 against the oracle; at 512:
    3.7e-14 and 6.6e-6).  The LES runs there take dt scaled by (512 / n)^2."""
    n = max(nx, ny, nz)
    return DT * min(1.0, (512.0 / n) ** 2)


def x_pass_fits(nx, model, prec):
    """The general x pass holds a row's field pairs in one block of at most 256 threads (nx / 16 per pair
    at 16 points per thread): with LES (6 pairs) a row of 768 does not fit, nor one of 1024 in FP64 (FP32 rows
    of 1024 take the register pass, one row per wavefront), and creating those runs fails."""
    return not (model == "les" and (nx == 768 or (nx == 1024 and prec == "fp64")))


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("model", MODELS)
@pytest.mark.parametrize("shape", SHAPES, ids=["x".join(map(str, s[:3])) for s in SHAPES])
def test_long_transforms_match_oracle(shape, model, prec):
    nx, ny, nz, lx = shape
    print_case(f"long rows {shape} {model} {prec}")
    g = Grid(nx, ny, nz, lx, 1.0, 1.0)
    u0 = random_spectrum(nx, ny, nz, seed=5, lx=lx)
    les_c = 0.16 if model == "les" else None
    dens = {"Pr": 0.7, "gradient": 1.0, "g": -1.0} if model == "density" else None
    nu = stable_nu(nx, ny, nz, lx)
    dt = les_dt(nx, ny, nz) if model == "les" else DT
    ref = OracleSim(g, nu, dt=dt, variable_dt=False, dt_stat=1 << 30, les_c=les_c, density=dens)
    ref.set_velocity(u0)
    p = ffs.Params(nx=nx, ny=ny, nz=nz, x_domain=lx, nu=nu, integrator="Adams_Bashforth3rdO", variable_dt=False, dt=dt,
                   dt_stat=1 << 30, dt_spec=1 << 30, write_time=1 << 30, precision=prec)
    if les_c is not None:
        p.les_model, p.smagorinsky_c = "Smagorinsky", les_c
    rho0 = None
    if dens is not None:
        rho0 = random_scalar(nx, ny, nz, seed=6, lx=lx)
        ref.set_density(rho0)
        p.density, p.pr, p.density_gradient, p.z_acceleration, p.reference_density = True, 0.7, 1.0, -1.0, 1.0
    if not x_pass_fits(nx, model, prec):
        with pytest.raises(ffs._lib.FfsError, match="exceed one block|field pairs per row"):
            ffs.Simulation(p)
        return
    with ffs.Simulation(p) as s:
        s.set_velocity(*u0)
        if rho0 is not None:
            s.set_density(rho0)
        ref.initialize()
        s.initialize()
        ref.advance_in_time(3)
        s.advance(3)
        tol = 1e-11 if prec == "fp64" else 2e-5
        got = s.velocity_hat()
        err = max(rel(got[c], ref.uh[c]) for c in range(3))
        assert err < tol, ("velocity", err)
        if rho0 is not None:
            rerr = rel(s.density_hat(), ref.rhoh)
            assert rerr < tol, ("density", rerr)


CUTS = [(256, 0.9), (256, 1.0), (64, 1.0), (32, 0.8)]


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("nx,cutoff", CUTS, ids=[f"{n}-cut{c}" for n, c in CUTS])
def test_other_cutoffs_match_oracle(nx, cutoff, prec):
    """A wider retained set (cutoff 0.8, 0.9) and no dealiasing at all (cutoff 1: the x Nyquist mode is
    kept, so the x pass takes its register variant that builds and splits rows by position)."""
    print_case(f"cutoff {cutoff} nx={nx} {prec}")
    ny = nz = 16
    g = Grid(nx, ny, nz, 1.0, 1.0, 1.0, cutoff=cutoff)
    u0 = random_spectrum(nx, ny, nz, seed=9, cutoff=cutoff)
    ref = OracleSim(g, 2e-3, dt=1e-3, variable_dt=False, dt_stat=1 << 30)
    ref.set_velocity(u0)
    p = ffs.Params(nx=nx, ny=ny, nz=nz, nu=2e-3, integrator="Adams_Bashforth3rdO", variable_dt=False, dt=1e-3,
                   dt_stat=1 << 30, dt_spec=1 << 30, write_time=1 << 30, precision=prec, cutoff=cutoff)
    with ffs.Simulation(p) as s:
        s.set_velocity(*u0)
        ref.initialize()
        s.initialize()
        ref.advance_in_time(3)
        s.advance(3)
        got = s.velocity_hat()
        err = max(rel(got[c], ref.uh[c]) for c in range(3))
        assert err < (1e-11 if prec == "fp64" else 2e-5), err
