"""Pins the CPU oracle (oracle/ffs_oracle.py) without the reference, which
cannot run here (Julia, un-vendored FluidFields/FFTW): every check below is
fixed by mathematics -- closed forms, brute-force definitions, invariants --
and is chosen so that a plausible transcription slip (a dropped term, a
wrong sign, a transposed index) fails at least one of them."""
import math

import numpy as np
import pytest

from oracle.ffs_oracle import Grid, OracleSim, fftfreq, proj, rfftfreq, taylor_green


# ---------------------------------------------------------------- grid (Globals.jl)
def test_grid_wavenumbers_closed_form():
    assert np.array_equal(rfftfreq(8, 1.0), [0, 1, 2, 3, 4])
    assert np.array_equal(fftfreq(8, 1.0), [0, 1, 2, 3, -4, -3, -2, -1])
    assert np.array_equal(fftfreq(7, 1.0), [0, 1, 2, 3, -3, -2, -1])
    assert np.allclose(rfftfreq(8, 2.0), [0, 0.5, 1, 1.5, 2])
    # same as numpy's definitions
    for n in (6, 7, 16, 48):
        assert np.allclose(fftfreq(n, 1.0), np.fft.fftfreq(n, 1.0 / n))
        assert np.allclose(rfftfreq(n, 1.0)[: n // 2 + 1], np.fft.rfftfreq(n, 1.0 / n))


def test_grid_dealias_mask_and_ranges_256():
    g = Grid(256, 256, 256)
    # 2/3 rule: |k| <= 85 kept, 86.. dealiased in every direction (box, not sphere)
    assert not g.DEALIAS[85, 0, 0] and g.DEALIAS[86, 0, 0]
    assert not g.DEALIAS[0, 85, 0] and g.DEALIAS[0, 86, 0]
    assert not g.DEALIAS[0, 256 - 85, 0] and g.DEALIAS[0, 256 - 86, 0]
    assert not g.DEALIAS[85, 85, 85]  # corner of the box is kept
    # XRANGE = 1:86, YRANGE = 1:86 ++ 169:256 (3 extra zero rows, Globals.jl:165-166)
    assert len(g.XRANGE) == 86 and g.XRANGE[-1] == 85
    assert len(g.YRANGE) == 174 and g.YRANGE[86] == 168
    # every kept mode lies in the loop ranges
    assert not np.any(~g.DEALIAS & ~g.RANGE)


def test_grid_small_counts():
    g = Grid(64, 64, 64)
    kept = ~g.DEALIAS
    assert kept.sum() == 22 * 43 * 43  # kx 0..21, ky/kz -21..21


# ---------------------------------------------------------------- transforms
def brute_rfft3(a):
    nx, ny, nz = a.shape
    out = np.zeros((nx // 2 + 1, ny, nz), dtype=complex)
    n = np.arange
    ex = np.exp(-2j * np.pi * np.outer(n(nx // 2 + 1), n(nx)) / nx)
    ey = np.exp(-2j * np.pi * np.outer(n(ny), n(ny)) / ny)
    ez = np.exp(-2j * np.pi * np.outer(n(nz), n(nz)) / nz)
    for kx in range(nx // 2 + 1):
        for ky in range(ny):
            for kz in range(nz):
                out[kx, ky, kz] = np.sum(a * ex[kx][:, None, None] * ey[ky][None, :, None] * ez[kz][None, None, :])
    return out


def test_rfft_matches_bruteforce_dft():
    rng = np.random.default_rng(0)
    a = rng.standard_normal((6, 4, 6))
    g = Grid(6, 4, 6)
    assert np.allclose(g.rfft(a), brute_rfft3(a), atol=1e-12)
    # unnormalised backward: irfft(rfft(a)) = N a
    assert np.allclose(g.irfft(g.rfft(a)), a * a.size, atol=1e-11)


def test_myfourier_single_mode_and_scaling():
    g = Grid(16, 16, 16)
    x, y, z = g.coords()
    a = np.cos(x) + 2 * np.sin(3 * y) + np.cos(2 * x + z)
    c = g.myfourier(a)
    assert abs(c[1, 0, 0] - 0.5) < 1e-15
    assert abs(c[0, 3, 0] - (-1j)) < 1e-15  # 2 sin(3y) -> -i at ky=3, +i at ky=-3
    assert abs(c[0, 16 - 3, 0] - 1j) < 1e-15
    assert abs(c[2, 0, 1] - 0.5) < 1e-15
    # a mode above the 2/3 cutoff is removed
    b = np.cos(7 * x)
    assert np.abs(g.myfourier(b)).max() < 1e-14  # roundoff only


def test_parseval_weights():
    """<u^2> = sum_{kx>=0} (1+[kx>0]) |u_k|^2 (stats.jl:300-301) for a dealiased field."""
    g = Grid(16, 16, 16)
    rng = np.random.default_rng(1)
    c = g.myfourier(rng.standard_normal((16, 16, 16)))
    r = g.irfft(c)
    lhs = np.mean(r * r)
    rhs = np.sum(g.W * np.abs(c) ** 2)
    assert abs(lhs - rhs) < 1e-13 * lhs


# ---------------------------------------------------------------- curl / projection
def test_curl_and_projection_single_mode():
    g = Grid(16, 16, 16)
    x, y, z = g.coords()
    # u = (0, 0, sin(x)) -> w = curl u = (0, -cos x, 0); projection keeps it (div-free)
    u = np.array([0 * x, 0 * x, np.sin(x)])
    s = OracleSim(g, nu=0.0, dt=0.01)
    s.set_velocity(u)
    kx, ky, kz = g.kx, g.ky, g.kz
    uh = s.uh
    w = np.array([1j * (ky * uh[2] - kz * uh[1]), 1j * (kz * uh[0] - kx * uh[2]), 1j * (kx * uh[1] - ky * uh[0])])
    wr = np.array([g.irfft(c) for c in w])
    assert np.allclose(wr[1], -np.cos(x), atol=1e-13) and np.allclose(wr[0], 0) and np.allclose(wr[2], 0)
    # gradient field is removed entirely by the Leray projection
    grad = np.array([np.cos(x + 2 * y) * 1, np.cos(x + 2 * y) * 2, 0 * x])
    gh = np.array([g.myfourier(c) for c in grad])
    assert np.abs(g.project(gh)).max() < 1e-15
    # projected fields are divergence-free: k . u = 0
    rng = np.random.default_rng(2)
    rh = np.array([g.myfourier(rng.standard_normal((16, 16, 16))) for _ in range(3)])
    ph = g.project(rh)
    assert np.abs(kx * ph[0] + ky * ph[1] + kz * ph[2]).max() < 1e-15


# ---------------------------------------------------------------- Taylor-Green closed forms
@pytest.fixture(scope="module")
def tg32():
    g = Grid(32, 32, 32)
    s = OracleSim(g, 0.0333333, dt=0.05, dt_stat=1)
    s.set_velocity(taylor_green(g))
    s.initialize()
    s.advance_in_time(1)
    return s


def test_taylor_green_energy_dissipation_t0(tg32):
    st = tg32.records[0][0]
    assert st[0] == 0 and st[1] == 0
    assert abs(st[11] - 0.125) < 1e-15  # k
    assert abs(st[14] - 6 * 0.0333333 * 0.125) < 1e-16  # diss = 6 nu k
    assert abs(st[5] - 0.125) < 1e-15 and abs(st[6] - 0.125) < 1e-15 and st[7] < 1e-30  # <u1^2>, <u2^2>, <u3^2>
    # kh = (u1p2+u2p2)/2, kv = 0; dissh = diss
    assert abs(st[8] - 0.125) < 1e-15 and st[9] < 1e-30 and abs(st[12] - st[14]) < 1e-18


def test_taylor_green_nonlinear_term_t0():
    """u x w for this TG variant is a gradient plus a 2k-mode field; its projected
    part and the pressure are known in closed form: p = (cos 2x + cos 2y)(cos 2z - 2)/16."""
    g = Grid(32, 32, 32)
    x, y, z = g.coords()
    s = OracleSim(g, nu=0.0, dt=0.01)
    s.set_velocity(taylor_green(g))
    s.initialize()
    # nonlinear term in the convective form: -(u.grad)u = u x w - grad(|u|^2/2); with
    # -(u.grad)u - grad p + P[...] the pressure solves lap p = -div((u.grad)u)
    ur = np.array([g.irfft(c) for c in s.uh])
    kx, ky, kz = g.kx, g.ky, g.kz
    # rhs before projection = F(u x w); pressure-free part = P(F(u x w))
    w = np.array([1j * (ky * s.uh[2] - kz * s.uh[1]), 1j * (kz * s.uh[0] - kx * s.uh[2]),
                  1j * (kx * s.uh[1] - ky * s.uh[0])])
    wr = np.array([g.irfft(c) for c in w])
    cross = np.array([ur[1] * wr[2] - ur[2] * wr[1], ur[2] * wr[0] - ur[0] * wr[2], ur[0] * wr[1] - ur[1] * wr[0]])
    ch = np.array([g.myfourier(c) for c in cross])
    ke = g.myfourier(0.5 * (ur * ur).sum(axis=0))
    # u x w = grad(|u|^2/2) + (-(u.grad)u); the analytic pressure gradient: -(u.grad)u - grad p is div-free
    p = (np.cos(2 * x) + np.cos(2 * y)) * (np.cos(2 * z) - 2) / 16
    ph = g.myfourier(p)
    conv = np.array([ch[0] - 1j * kx * ke, ch[1] - 1j * ky * ke, ch[2] - 1j * kz * ke])  # -(u.grad)u
    resid = np.array([conv[0] - 1j * kx * ph, conv[1] - 1j * ky * ph, conv[2] - 1j * kz * ph])
    div = kx * resid[0] + ky * resid[1] + kz * resid[2]
    assert np.abs(div).max() < 1e-13
    # and the projection of u x w equals conv - grad p exactly
    assert np.allclose(g.project(ch), resid, atol=1e-14)


def test_energy_budget_unforced():
    """dk/dt = -eps (the nonlinear term conserves energy)."""
    g = Grid(32, 32, 32)
    s = OracleSim(g, 0.01, dt=0.002, dt_stat=1)
    from ffs.initial import random_spectrum

    s.set_velocity(random_spectrum(32, seed=9))
    s.initialize()
    s.advance_in_time(6)
    k = np.array([r[0][11] for r in s.records])
    eps = np.array([r[0][14] for r in s.records])
    dk = (k[4] - k[2]) / (2 * 0.002)
    assert abs(dk + eps[3]) < 2e-3 * eps[3]
    # total nonlinear + pressure transfer vanishes
    for r in s.records:
        st = r[0]
        assert abs(st[15] + st[16]) < 1e-12 * st[14]
        assert abs(st[17] + st[18]) < 1e-12 * st[14]


def test_constant_power_forcing_injects_p():
    g = Grid(32, 32, 32)
    from ffs.initial import random_spectrum

    s = OracleSim(g, 0.0, dt=1e-4, dt_stat=1, forcing_kf=2.5, forcing_power=0.1)
    s.set_velocity(random_spectrum(32, seed=3))
    s.initialize()
    s.advance_in_time(3)
    k = [r[0][11] for r in s.records]
    # inviscid: dk/dt = P  (to O(dt))
    assert abs((k[2] - k[1]) / 1e-4 - 0.1) < 1e-3


# ---------------------------------------------------------------- AB3
def test_ab3_constant_step_limit():
    s = OracleSim(Grid(8, 8, 8), 0.0, dt=0.05)
    At, Bt, Ct = s.ab3_coefficients()
    assert abs(At - 23 / 12 * 0.05) < 1e-16
    assert abs(Bt + 16 / 12 * 0.05) < 1e-16
    assert abs(Ct - 5 / 12 * 0.05) < 1e-16


def test_ab3_variable_step_exact_for_quadratics():
    """Variable-step AB3 integrates f(t) = a + b t + c t^2 exactly."""
    s = OracleSim(Grid(8, 8, 8), 0.0, dt=0.1)
    s.dt, s.dt2, s.dt3 = 0.1, 0.07, 0.13
    At, Bt, Ct = s.ab3_coefficients()
    f = lambda t: 1.0 + 2.0 * t - 3.0 * t * t
    F = lambda t: t + t * t - t**3
    t0 = 0.0
    approx = At * f(t0) + Bt * f(t0 - s.dt2) + Ct * f(t0 - s.dt2 - s.dt3)
    assert abs(approx - (F(t0 + s.dt) - F(t0))) < 1e-15


def test_viscous_decay_single_mode_startup_order():
    """A single Fourier mode decays as exp(-nu k^2 t).  The reference starts AB3
    with fm1 = fm2 = rhs (an Euler step, Adams_Bashforth.jl:157-164) and an AB2
    step (fix_fm2!), so the global error is second order: halving dt divides it by 4."""
    g = Grid(16, 16, 16)
    x, y, z = g.coords()
    nu = 0.1
    errs = []
    for dt in (0.04, 0.02):
        s = OracleSim(g, nu, dt=dt)
        s.set_velocity(np.array([np.sin(y), 0 * x, 0 * x]))  # ky = 1, div-free, u x w = grad
        s.initialize()
        nsteps = int(round(0.8 / dt))
        s.advance_in_time(nsteps)
        amp = abs(s.uh[0][0, 1, 0]) * 2
        errs.append(abs(amp - math.exp(-nu * 0.8)))
    order = math.log(errs[0] / errs[1], 2)
    assert 1.9 < order < 2.1, (errs, order)


def test_spectrum_sum_rules():
    """sum of specH bins weighted as in calculate_specHV reproduces the energy marginals."""
    g = Grid(16, 16, 16)
    from ffs.initial import random_spectrum

    s = OracleSim(g, 0.01, dt=0.01)
    s.set_velocity(random_spectrum(16, seed=1))
    (hH, hV), (vH, vV) = s.spectrum()
    u = s.uh
    eh = np.sum(np.where(g.RANGE, (proj(u[0], u[0]) + proj(u[1], u[1])) / 2, 0))
    e_i1 = np.sum(np.where(g.RANGE, (proj(u[0], u[0]) + proj(u[1], u[1])) / 2, 0)[0])
    # V: sum over bins = e(i=1) + 2 dkz * rest
    assert abs(hV.sum() - (e_i1 + 2 * g.DKZ * (eh - e_i1))) < 1e-14
    # H: h counts every mode twice (its kx bin and its ky bin)
    assert abs(hH.sum() - (e_i1 + 2 * g.MAXDKH * (eh - e_i1))) < 1e-14


# ---------------------------------------------------------------- ETD3
def test_etd_exact_for_linear_decay():
    """u = (sin y, 0, 0): u x w is a gradient, so the projected rhs vanishes and
    ETD integrates exp(-nu k^2 t) exactly for any dt (ETD1, ETD2 and ETD3 steps)."""
    g = Grid(16, 16, 16)
    x, y, z = g.coords()
    nu, dt = 0.3, 0.25
    s = OracleSim(g, nu, dt=dt, integrator="ETD3")
    s.set_velocity(np.array([np.sin(y), 0 * x, 0 * x]))
    s.initialize()
    s.advance_in_time(6)
    amp = abs(s.uh[0][0, 1, 0]) * 2
    assert abs(amp - math.exp(-nu * 6 * dt)) < 1e-14


def test_etd3_coefficients_constant_forcing_and_continuity():
    """For y' = c y + f with constant f, sum(At, Bt, Ct) = (e^{c dt} - 1)/c; the
    series and closed-form branches meet at -c dt = 1e-4 (set_ABCt!, ETD.jl:181)."""
    s = OracleSim(Grid(8, 8, 8), 0.0, dt=0.01, integrator="ETD3")
    s.dt, s.dt2, s.dt3 = 0.01, 0.012, 0.009
    c = np.array([-0.5, -10.0, -200.0, -1e-4 / 0.01 * 0.999, -1e-4 / 0.01 * 1.001])
    At, Bt, Ct = s.etd_coefficients(c)
    exact = np.expm1(c * s.dt) / c
    assert np.allclose(At + Bt + Ct, exact, rtol=1e-9)
    # the two branches agree at the switch to the accuracy the closed form allows there
    assert abs(At[3] - At[4]) / abs(At[3]) < 1e-6 and abs(Ct[3] - Ct[4]) / abs(Ct[3]) < 1e-3
    # constant step, c -> 0: AB3 weights
    s.dt = s.dt2 = s.dt3 = 0.01
    A0, B0, C0 = s.etd_coefficients(np.array([0.0]))
    assert np.allclose([A0[0], B0[0], C0[0]], [23 / 12 * 0.01, -16 / 12 * 0.01, 5 / 12 * 0.01], rtol=1e-14)


# ---------------------------------------------------------------- Smagorinsky
def test_smagorinsky_stress_of_shear_layer():
    """u = (sin y, 0, 0): S_xy = cos(y)/2, |S| = sqrt(2 S:S) = |cos y|, so
    tau_xy = 2 (c Delta)^2 |S| S_xy = (c Delta)^2 |cos y| cos y; other components vanish."""
    g = Grid(32, 32, 32)
    x, y, z = g.coords()
    c = 0.16
    s = OracleSim(g, 0.0, dt=0.01, les_c=c)
    s.set_velocity(np.array([np.sin(y), 0 * x, 0 * x]))
    s.initialize()
    delta = (2 * np.pi / 32) / (2 / 3)
    expect = g.myfourier((c * delta) ** 2 * np.abs(np.cos(y)) * np.cos(y))
    assert np.abs(s.tauh[1] - expect).max() < 1e-15
    for i in (0, 2, 3, 4):
        assert np.abs(s.tauh[i]).max() < 1e-16


def test_smagorinsky_drains_energy():
    g = Grid(16, 16, 16)
    from ffs.initial import random_spectrum

    s = OracleSim(g, 0.0, dt=0.001, les_c=0.17, dt_stat=1)
    s.set_velocity(random_spectrum(16, seed=5))
    s.initialize()
    s.advance_in_time(2)
    st = s.records[-1][0]
    pr = st[-1]
    assert pr > 0 and abs(st[21] - (st[14] + pr)) < 1e-15  # tdiss = diss + pr


# ---------------------------------------------------------------- Boussinesq
def test_internal_gravity_wave_frequency():
    """Linear internal wave with k = (1, 0, 1): omega = N k_h / |k| = 1/sqrt(2) for
    N = sqrt(g dρ/dz) = 1 (inviscid, small amplitude)."""
    g = Grid(16, 16, 16)
    x, y, z = g.coords()
    eps = 1e-6
    # rho = eps cos(x + z), u = 0 initially: rho(t) = eps cos(omega t) cos(x + z)
    dens = {"Pr": 1.0, "gradient": 1.0, "g": 1.0}
    dt = 0.01
    s = OracleSim(g, 0.0, dt=dt, density=dens, integrator="AB3")
    s.set_velocity(np.zeros((3, 16, 16, 16)))
    s.set_density(eps * np.cos(x + z))
    s.initialize()
    steps = 300
    s.advance_in_time(steps)
    omega = 1 / math.sqrt(2)
    amp = 2 * s.rhoh[1, 0, 1].real / eps
    assert abs(amp - math.cos(omega * steps * dt)) < 2e-3
    # w and u_x are the only velocity components; div u = 0 keeps u_x = -w
    assert abs(s.uh[0][1, 0, 1] + s.uh[2][1, 0, 1]) < 1e-12 * eps


def test_boussinesq_energy_exchange():
    """The buoyancy flux bflux = -g <w rho> (buoyancy_flux, density_stratification_types.jl:196)
    moves kinetic into potential energy: d k/dt = -bflux - diss (unforced, without LES)."""
    g = Grid(16, 16, 16)
    from ffs.initial import random_scalar, random_spectrum

    dens = {"Pr": 1.0, "gradient": 1.0, "g": 1.0}
    dt = 1e-3
    s = OracleSim(g, 0.01, dt=dt, density=dens, dt_stat=1, integrator="AB3")
    s.set_velocity(random_spectrum(16, seed=3))
    s.set_density(random_scalar(16, seed=4, rms=0.3))
    s.initialize()
    s.advance_in_time(6)
    k = [r[0][11] for r in s.records]
    st = s.records[3][0]
    bflux, diss = st[29], st[14]
    dk = (k[4] - k[2]) / (2 * dt)
    assert abs(dk - (-bflux - diss)) < 2e-3 * max(abs(bflux), diss)
