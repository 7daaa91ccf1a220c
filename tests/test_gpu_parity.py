"""GPU parity: the HIP path through the C ABI against the CPU oracle on the
same inputs.  FP64 tolerance: fields 1e-12 relative (max-norm over the
component, relative to the component's max), statistics 1e-10 relative."""
import numpy as np
import pytest

import ffs
from ffs import _lib
from ffs.initial import random_spectrum, taylor_green
from oracle.ffs_oracle import Grid, OracleSim

from conftest import print_case

pytestmark = pytest.mark.gpu

FIELD_TOL = 1e-12
STATS_TOL = 1e-10


def make(n, nu, dt, u0, variable_dt=False, shape=None, dom=(1.0, 1.0, 1.0), dt_stat=1, forcing_kf=None,
         cfl=0.5, write_time=1 << 30):
    nx, ny, nz = shape or (n, n, n)
    g = Grid(nx, ny, nz, *dom)
    ref = OracleSim(g, nu, dt=dt, variable_dt=variable_dt, cfl=cfl, dt_stat=dt_stat, forcing_kf=forcing_kf)
    ref.set_velocity(u0)
    p = ffs.Params(nx=nx, ny=ny, nz=nz, x_domain=dom[0], y_domain=dom[1], z_domain=dom[2], nu=nu, dt=dt,
                   integrator="Adams_Bashforth3rdO", variable_dt=variable_dt, cfl=cfl, dt_stat=dt_stat,
                   dt_spec=1 << 30, write_time=write_time,
                   forcing="shellForcing" if forcing_kf else None, forcing_kf=forcing_kf or 2.5)
    s = ffs.Simulation(p)
    s.set_velocity(*u0)
    return ref, s


def rel_field_err(a, b):
    return max(np.abs(a[c] - b[c]).max() / max(np.abs(b[c]).max(), 1e-300) for c in range(3))


def test_roundtrip_transform():
    for shape in [(16, 16, 16), (32, 24, 48), (64, 64, 64)]:
        print_case(f"transform round trip {shape}")
        nx, ny, nz = shape
        g = Grid(nx, ny, nz)
        rng = np.random.default_rng(7)
        a = rng.standard_normal((3, nx, ny, nz))
        p = ffs.Params(nx=nx, ny=ny, nz=nz, nu=0.01, dt=0.01, integrator="Adams_Bashforth3rdO", variable_dt=False,
                       dt_stat=1 << 20, dt_spec=1 << 20, write_time=1 << 20)
        with ffs.Simulation(p) as s:
            s.set_velocity(*a)
            for c in range(3):
                ref = g.myfourier(a[c])
                got = s.get_field(c, _lib.FFS_SPACE_FOURIER)
                err = np.abs(got - ref).max() / np.abs(ref).max()
                assert err < 1e-13, (shape, c, err)
                back = s.get_field(c, _lib.FFS_SPACE_REAL)
                assert np.all(back[nx:] == 0)
                refr = g.irfft(ref)
                err_r = np.abs(back[:nx] - refr).max() / np.abs(refr).max()
                assert err_r < 1e-13, (shape, c, err_r)


@pytest.mark.parametrize("prec,pair", [("fp64", 2), ("fp64", 1), ("fp64", 0), ("fp32", 1)])
def test_umax_is_the_nan_propagating_maximum(prec, pair):
    """find_max (models.jl: maximum(abs, u), which propagates NaN) as the x pass reduces it for set_dt (variable
    dt): on a random field max|u_i| equals the maximum over the real-space components; one NaN makes it NaN."""
    n = 32
    p = ffs.Params(nx=n, ny=n, nz=n, nu=1e-3, integrator="Adams_Bashforth3rdO", variable_dt=True, cfl=0.5,
                   dt_stat=1 << 30, dt_spec=1 << 30, write_time=1 << 30, precision=prec)
    u0 = random_spectrum(n, n, n, seed=5)
    tol = 1e-12 if prec == "fp64" else 1e-5
    for plant in (False, True):
        print_case(f"{prec} x_pair {pair} {'one NaN' if plant else 'finite'}")
        with ffs.Simulation(p) as s:
            s.set_plan({"x_pair": pair})
            u = [np.array(c, copy=True) for c in u0]
            if plant:
                u[1][3, 2, 1] = np.nan
            s.set_velocity(*u)
            s.initialize()
            umax = s.scalar(_lib.FFS_SCALAR_UMAX)
            if plant:
                assert np.isnan(umax)
            else:
                want = max(np.abs(s.get_field(c)[:n]).max() for c in range(3))
                assert abs(umax - want) <= tol * want, (umax, want)


def test_unpadded_and_padded_inputs_agree():
    n = 16
    u0 = random_spectrum(n, seed=3)
    p = ffs.Params(nx=n, ny=n, nz=n, nu=0.01, dt=0.01, integrator="Adams_Bashforth3rdO", variable_dt=False,
                   dt_stat=1 << 20, dt_spec=1 << 20, write_time=1 << 20)
    with ffs.Simulation(p) as a, ffs.Simulation(p) as b:
        a.set_velocity(*u0)
        b.set_velocity(*[ffs.pad(c) for c in u0])
        for c in range(3):
            assert np.array_equal(a.get_field(c, _lib.FFS_SPACE_FOURIER), b.get_field(c, _lib.FFS_SPACE_FOURIER))


@pytest.mark.parametrize("n", [16, 24, 32, 48, 64])
def test_taylor_green_one_step(n):
    print_case(f"taylor-green one step n={n}")
    nu, dt = 0.0333333, 0.05
    g = Grid(n, n, n)
    ref, s = make(n, nu, dt, taylor_green(n))
    with s:
        ref.initialize()
        s.initialize()
        ref.advance_in_time(1)
        s.advance_in_time(1)
        err = rel_field_err(s.velocity_hat(), ref.uh)
        assert err < FIELD_TOL, err


@pytest.mark.parametrize("shape", [(32, 32, 32), (48, 48, 48), (32, 24, 48), (64, 32, 16)])
def test_random_field_steps(shape):
    print_case(f"random field 5 steps {shape}")
    nx, ny, nz = shape
    u0 = random_spectrum(nx, ny, nz, seed=11)
    ref, s = make(nx, 0.01, 0.01, u0, shape=shape)
    with s:
        ref.initialize()
        s.initialize()
        for step in range(5):
            ref.advance_in_time(1)
            s.advance_in_time(1)
            err = rel_field_err(s.velocity_hat(), ref.uh)
            assert err < FIELD_TOL * (step + 1), (step, err)


def test_noncubic_domain():
    shape = (32, 16, 24)
    dom = (1.0, 0.5, 2.0)
    print_case(f"non-cubic domain {shape} {dom}")
    u0 = random_spectrum(*shape, seed=5, lx=dom[0], ly=dom[1], lz=dom[2])
    ref, s = make(32, 0.02, 0.005, u0, shape=shape, dom=dom)
    with s:
        ref.initialize()
        s.initialize()
        ref.advance_in_time(3)
        s.advance_in_time(3)
        assert rel_field_err(s.velocity_hat(), ref.uh) < 3 * FIELD_TOL


def test_taylor_green_stats_40_steps():
    n, nu, dt = 32, 0.0333333, 0.05
    print_case("taylor-green 32^3 stats every step, 40 steps")
    ref, s = make(n, nu, dt, taylor_green(n), dt_stat=1)
    with s:
        ref.initialize()
        s.initialize()
        ref.advance_in_time(40)
        s.advance_in_time(40)
        rows = s.pop_stats()
    assert len(rows) == 40
    for (st, sc, nm), (ost, osc, onm) in zip(rows, ref.records):
        assert st[0] == ost[0]
        assert abs(st[1] - ost[1]) <= 1e-14 * max(1.0, abs(ost[1]))
        # energies relative to k, dissipations relative to diss (u3 = 0 exactly for TG: roundoff floor)
        for col, scale in ((5, ost[11]), (6, ost[11]), (7, ost[11]), (8, ost[11]), (9, ost[11]), (11, ost[11]),
                           (12, ost[14]), (13, ost[14]), (14, ost[14]), (21, ost[14])):
            assert abs(st[col] - ost[col]) <= STATS_TOL * abs(ost[col]) + 1e-15 * abs(scale), (st[0], col, st[col],
                                                                                                ost[col])
        # transfer terms: absolute tolerance on the scale of the dissipation
        for col in (15, 16, 17, 18):
            assert abs(st[col] - ost[col]) <= 1e-12 * abs(ost[14]) + STATS_TOL * abs(ost[col]), (col, st[col], ost[col])
        # length and time scales that are not roundoff for TG (its integral-length plane sums vanish)
        for i in (2, 3, 4, 5, 6, 11, 12, 22, 23):
            assert abs(sc[i] - osc[i]) <= 1e-9 * abs(osc[i]), (i, sc[i], osc[i])


def test_variable_dt_and_forcing():
    n = 32
    print_case("random field, variable dt (cfl 0.5) + shell forcing, 12 steps")
    u0 = random_spectrum(n, seed=2)
    ref, s = make(n, 0.005, 0.0, u0, variable_dt=True, cfl=0.5, forcing_kf=2.5, dt_stat=3)
    with s:
        ref.initialize()
        s.initialize()
        for step in range(12):
            ref.advance_in_time(1)
            s.advance_in_time(1)
            assert abs(s.dt - ref.dt) <= 1e-13 * ref.dt, (step, s.dt, ref.dt)
            assert abs(s.time - ref.time) <= 1e-13 * ref.time
            err = rel_field_err(s.velocity_hat(), ref.uh)
            assert err < FIELD_TOL * (step + 1), (step, err)
        rows = s.pop_stats()
    assert [int(r[0][0]) for r in rows] == [int(r[0][0]) for r in ref.records]


def test_spectrum_matches_oracle():
    n = 32
    print_case("spectrum_u + calculate_specHV")
    u0 = random_spectrum(n, seed=4)
    ref, s = make(n, 0.01, 0.01, u0)
    with s:
        ref.initialize()
        s.initialize()
        ref.advance_in_time(2)
        s.advance_in_time(2)
        got = s.spectrum()
    exp = ref.spectrum()
    for (gh, gv), (eh, ev) in zip(got, exp):
        np.testing.assert_allclose(gh, eh, rtol=1e-11, atol=1e-14 * np.abs(eh).max())
        np.testing.assert_allclose(gv, ev, rtol=1e-11, atol=1e-14 * np.abs(ev).max())


def test_output_event():
    n = 16
    print_case("output event cadence")
    ref, s = make(n, 0.01, 0.01, taylor_green(n), write_time=4)
    with s:
        s.initialize()
        done, ev = s.advance(10)
        assert (done, ev) == (4, _lib.FFS_EVENT_OUTPUT) and s.iteration == 4
        done, ev = s.advance(6)
        assert (done, ev) == (4, _lib.FFS_EVENT_OUTPUT) and s.iteration == 8
        done, ev = s.advance(2)
        assert (done, ev) == (2, _lib.FFS_EVENT_NONE) and s.iteration == 10


def shell_spectrum(uh, g):
    """3D shell spectrum E(k) binned by round(|k|/dk) with Parseval weights."""
    k = np.sqrt(g.K2)
    dk = g.KX[1]
    bins = np.rint(k / dk).astype(int)
    e = 0.5 * g.W * sum(np.abs(c) ** 2 for c in uh)
    return np.bincount(bins.ravel(), weights=np.where(g.RANGE, e, 0).ravel())


@pytest.mark.parametrize("n", [32, 64])
def test_fp32_bulk_against_fp64_oracle(n):
    """FP32 fields/FFTs with FP64 coefficients and reductions: spectrum error <= 1e-5."""
    print_case(f"fp32 bulk vs fp64 oracle n={n}")
    u0 = random_spectrum(n, seed=8)
    g = Grid(n, n, n)
    ref = OracleSim(g, 0.01, dt=0.002, dt_stat=5)
    ref.set_velocity(u0)
    p = ffs.Params(nx=n, ny=n, nz=n, nu=0.01, dt=0.002, integrator="Adams_Bashforth3rdO", variable_dt=False,
                   dt_stat=5, dt_spec=1 << 30, write_time=1 << 30, precision="fp32")
    with ffs.Simulation(p) as s:
        s.set_velocity(*u0)
        ref.initialize()
        s.initialize()
        ref.advance_in_time(20)
        s.advance_in_time(20)
        uh = s.velocity_hat()
        rows = s.pop_stats()
    err = rel_field_err(uh, ref.uh)
    assert err < 2e-5, err
    e32, e64 = shell_spectrum(uh, g), shell_spectrum(ref.uh, g)
    sel = e64 >= 1e-6 * e64.max()
    spec_err = np.max(np.abs(e32[sel] - e64[sel]) / e64[sel])
    print(f"fp32 field err {err:.2e} spectrum err {spec_err:.2e}")
    assert spec_err <= 1e-5, spec_err
    for (st, _, _), (ost, _, _) in zip(rows, ref.records):
        assert abs(st[11] - ost[11]) <= 1e-6 * ost[11] and abs(st[14] - ost[14]) <= 1e-5 * ost[14]


def make_ext(n, nu, dt, u0, integrator="AB3", les_c=None, density=None, rho0=None, variable_dt=False, cfl=0.5,
             dt_stat=1, precision="fp64"):
    g = Grid(n, n, n)
    ref = OracleSim(g, nu, dt=dt, variable_dt=variable_dt, cfl=cfl, dt_stat=dt_stat, integrator=integrator,
                    les_c=les_c, density=density)
    ref.set_velocity(u0)
    if density is not None:
        ref.set_density(rho0)
    p = ffs.Params(nx=n, ny=n, nz=n, nu=nu, dt=dt, variable_dt=variable_dt, cfl=cfl, dt_stat=dt_stat,
                   dt_spec=1 << 30, write_time=1 << 30, precision=precision,
                   integrator="Adams_Bashforth3rdO" if integrator == "AB3" else "ETD3rdO")
    if les_c is not None:
        p.les_model, p.smagorinsky_c = "Smagorinsky", les_c
    if density is not None:
        p.density, p.pr, p.density_gradient = True, density["Pr"], density["gradient"]
        p.z_acceleration, p.reference_density = density["g"], 1.0
    s = ffs.Simulation(p)
    s.set_velocity(*u0)
    if density is not None:
        s.set_density(rho0)
    return ref, s


def rel_err(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


# ETD3's closed-form coefficients (set_ABCt!) cancel down to O((c dt)^3) near the series switch
# |c dt| = 1e-4.  With a fixed dt both sides evaluate them from identical inputs (no FMA contraction
# on the device) and agree to roundoff.  With a variable dt, dt = cfl dx / max|u| inherits the
# 1-ulp differences of two FFT implementations, and the closed forms amplify a relative change of
# dt by up to ~1/(c dt)^2 ~ 1e8: the variable-dt ETD field tolerance is therefore 1e-8.
ETD_TOL = FIELD_TOL
ETD_TOL_VARIABLE_DT = 1e-8


@pytest.mark.parametrize("variable_dt", [False, True])
def test_etd3_steps(variable_dt):
    n = 32
    print_case(f"ETD3 random field, variable_dt={variable_dt}")
    ref, s = make_ext(n, 0.01, 0.01, random_spectrum(n, seed=21), integrator="ETD3", variable_dt=variable_dt,
                      dt_stat=2)
    with s:
        ref.initialize()
        s.initialize()
        for step in range(6):
            ref.advance_in_time(1)
            s.advance_in_time(1)
            err = rel_field_err(s.velocity_hat(), ref.uh)
            print(f"etd step {step} err {err:.2e}")
            assert err < (ETD_TOL_VARIABLE_DT if variable_dt else ETD_TOL * (step + 1)), (step, err)
            # dt = cfl dx / max|u| follows the field difference
            assert abs(s.dt - ref.dt) <= (1e-9 if variable_dt else 1e-15) * ref.dt
        rows = s.pop_stats()
    for (st, _, _), (ost, _, _) in zip(rows, ref.records):
        assert abs(st[11] - ost[11]) <= STATS_TOL * ost[11] and abs(st[14] - ost[14]) <= STATS_TOL * ost[14]


def test_smagorinsky_les():
    n = 32
    print_case("Smagorinsky LES c=0.16, AB3")
    ref, s = make_ext(n, 1e-3, 0.005, random_spectrum(n, seed=31), les_c=0.16, dt_stat=1)
    with s:
        ref.initialize()
        s.initialize()
        for step in range(5):
            ref.advance_in_time(1)
            s.advance_in_time(1)
            err = rel_field_err(s.velocity_hat(), ref.uh)
            assert err < FIELD_TOL * (step + 1), (step, err)
        rows = s.pop_stats()
        hdr = s.header(_lib.FFS_TABLE_STATS).split(",")
    assert hdr[-4:] == ["prh", "prv", "praniso", "pr"]
    for (st, _, _), (ost, _, _) in zip(rows, ref.records):
        assert len(st) == len(ost) == len(hdr)
        for col in (11, 14, 21, 22, 23, 25):  # k, diss, tdiss, prh, prv, pr
            assert abs(st[col] - ost[col]) <= STATS_TOL * abs(ost[col]) + 1e-14 * abs(ost[14]), (col, st[col], ost[col])
        assert ost[25] > 0  # the Smagorinsky stress drains energy


@pytest.mark.parametrize("integrator", ["AB3", "ETD3"])
def test_boussinesq(integrator):
    n = 32
    print_case(f"Boussinesq N=1 Pr=1, {integrator}")
    from ffs.initial import random_scalar

    dens = {"Pr": 1.0, "gradient": 1.0, "g": 1.0}
    ref, s = make_ext(n, 0.005, 0.01, random_spectrum(n, seed=41), integrator=integrator, density=dens,
                      rho0=random_scalar(n, seed=42, rms=0.1), dt_stat=1)
    tol = FIELD_TOL if integrator == "AB3" else ETD_TOL
    with s:
        ref.initialize()
        s.initialize()
        for step in range(5):
            ref.advance_in_time(1)
            s.advance_in_time(1)
            err = rel_field_err(s.velocity_hat(), ref.uh)
            rerr = rel_err(s.density_hat(), ref.rhoh)
            assert err < tol * (step + 1) and rerr < tol * (step + 1), (step, err, rerr)
        rows = s.pop_stats()
        hdr = s.header(_lib.FFS_TABLE_STATS).split(",")
        shdr = s.header(_lib.FFS_TABLE_SCALES).split(",")
        nhdr = s.header(_lib.FFS_TABLE_NUMBERS).split(",")
    assert hdr[22:] == ["rho", "rhop2", "drhodxp2", "drhodyp2", "drhodzp2", "rhodiss", "trhodiss", "bflux"]
    for (st, sc, nm), (ost, osc, onm) in zip(rows, ref.records):
        assert len(st) == len(hdr) and len(sc) == len(shdr) and len(nm) == len(nhdr)
        for col in (11, 14, 23, 24, 25, 26, 27, 28, 29):
            assert abs(st[col] - ost[col]) <= STATS_TOL * abs(ost[col]) + 1e-14 * abs(ost[23]), (col, st[col], ost[col])
        for i in range(2, len(osc)):
            if np.isfinite(osc[i]) and abs(osc[i]) > 1e-12:
                assert abs(sc[i] - osc[i]) <= 1e-8 * abs(osc[i]), (i, sc[i], osc[i])


def test_boussinesq_variable_dt_fp32():
    """Stratified run in the FP32 bulk path with ETD3 and variable dt (density CFL limit) against FP64."""
    n = 32
    print_case("Boussinesq FP32 ETD3 variable dt")
    from ffs.initial import random_scalar

    dens = {"Pr": 1.0, "gradient": 1.0, "g": 1.0}
    u0, r0 = random_spectrum(n, seed=51), random_scalar(n, seed=52, rms=0.01)
    ref, s = make_ext(n, 1e-3, 0.0, u0, integrator="ETD3", density=dens, rho0=r0, variable_dt=True, dt_stat=4,
                      precision="fp32")
    with s:
        ref.initialize()
        s.initialize()
        ref.advance_in_time(12)
        s.advance_in_time(12)
        err = rel_field_err(s.velocity_hat(), ref.uh)
        rerr = rel_err(s.density_hat(), ref.rhoh)
        assert abs(s.dt - ref.dt) <= 1e-5 * ref.dt
        assert ref.dt < 0.5 * 2 * np.pi / n / ref.umax + 1e-12  # the CFL limit is active
    print(f"fp32 stratified: field err {err:.2e} density err {rerr:.2e}")
    assert err < 2e-5 and rerr < 2e-5, (err, rerr)
