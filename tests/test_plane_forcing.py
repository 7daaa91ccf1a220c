"""The oracle's restatement of the reference's kz = 0-plane forcing
(oracle/plane_forcing.py), pinned by brute force on small grids and by what the
definitions fix: QfForcing injects hp = 0.99 sum R with vp = hp 0.01/0.99; the
oscillators stay at R = 0 while the spectrum equals its target; one step from
R = 0 with Em = 2 Ef gives R = dt w^2 Ef (w = 2 pi/(Tf ts) with ts = dt/cfl,
10 dt for aForcing, Tf for the c-variants) and the matching power; the N
spectrum of a right-hand side c u + nu k^2 u is 2 c Ef; with N = 0 the N
variants reduce to their plain forms; aForcing's per-mode energies sum to the
shell spectrum."""
import math

import numpy as np
import pytest

from oracle.ffs_oracle import Grid, OracleSim
from oracle.plane_forcing import PlaneForcing, compute_shells2d
from oracle.seeded import seeded_velocity_hat


def test_shells_by_brute_force():
    g = Grid(12, 10, 8, 1.0, 1.3, 1.0)
    n, maxdk, num, kh = compute_shells2d(g.KX, g.KY, g.NX, g.NY)
    assert n == min(g.NX, g.NY // 2)
    assert maxdk == max(g.KX[1], g.KY[1])
    acc = np.zeros(n)
    cnt = np.zeros(n, dtype=int)
    for i in range(g.NX):
        for j in range(g.NY):
            K = math.hypot(g.KX[i], g.KY[j])
            s = int(np.rint(K / maxdk))
            if s < n:
                acc[s] += K
                cnt[s] += 1
    assert np.array_equal(num, cnt)
    assert np.allclose(kh, acc / cnt, rtol=1e-15, atol=0)


class _Sim:
    """the fields a forcing functor reads"""

    def __init__(self, g, uh, dt, cfl=0.5, it=10, variable_dt=True, rhs=None, nu=0.0, dt2=None):
        self.g, self.uh, self.dt, self.cfl, self.iteration, self.variable_dt = g, uh, dt, cfl, it, variable_dt
        self.rhs, self.nu, self.dt2 = rhs, nu, dt if dt2 is None else dt2


def field(n=32, seed=3):
    g = Grid(n, n, n)
    return g, np.array(seeded_velocity_hat(g, seed, k0=3.0))


def test_qf_power_closed_form():
    g, uh = field()
    f = PlaneForcing(g, "qf", kf=3.0, P=0.2, C=1.0, Kf_alv=2.0)
    forced = [i for i in range(1, f.nshells) if f.kh[i] <= 3.0]
    assert abs(sum(f.R) - 0.2) < 1e-15  # R is P times a normalised profile
    f(_Sim(g, uh, dt=1e-3))
    assert abs(f.hp - 0.99 * sum(f.R[i] for i in forced)) < 1e-12 * 0.2
    assert abs(f.vp - f.hp * 0.01 / 0.99) < 1e-12 * f.hp


def test_qf_les_filter_damps_the_spectrum():
    g, _ = field(16)
    a = PlaneForcing(g, "qf", kf=3.0, P=1.0, C=1.0, Kf_alv=2.0)
    b = PlaneForcing(g, "qf", kf=3.0, P=1.0, C=1.0, Kf_alv=2.0, les_delta=0.5)
    for i in range(1, a.nshells):
        if a.kh[i] <= 3.0:
            assert abs(b.R[i] - a.R[i] * math.exp(-0.25 / 24 * a.kh[i] ** 2) ** 2) < 1e-15


def test_rf_fixed_point_at_the_target():
    g, uh = field()
    f = PlaneForcing(g, "rf", kf=2.5, tf=1.0, alphac=0.5)
    f.initialize(_Sim(g, uh, 1e-3))
    f(_Sim(g, uh, dt=1e-3))
    assert np.all(f.R == 0) and f.hp == 0 and f.vp == 0
    assert not np.any(f.forcex) and not np.any(f.forcey)


def test_rf_one_step_by_hand():
    g, uh = field()
    f = PlaneForcing(g, "rf", kf=2.5, tf=2.0, alphac=0.5)
    dt, cfl = 2e-3, 0.4
    Ef = f.u1u2_spectrum(uh)
    f.Em = 2 * Ef
    f.init = False
    f(_Sim(g, uh, dt=dt, cfl=cfl, it=10))
    omega = 2 * math.pi / (2.0 * (dt / cfl))
    forced = [i for i in range(1, f.nshells) if f.kh[i] <= 2.5]
    assert forced
    hp = 0.0
    for i in forced:
        R = dt * omega * omega * Ef[i]
        assert abs(f.R[i] - R) <= 1e-14 * R
        fac = math.sqrt(R / Ef[i]) * dt
        hp += (fac * fac * 2 * Ef[i] * 0.5 + fac * 2 * Ef[i]) / dt
    assert abs(f.hp - hp) <= 1e-12 * hp
    # the force is fac * u on the forced elements of the kz = 0 plane only
    k = np.sqrt(g.kx**2 + g.ky**2)[:, :, 0]
    n = np.rint(k / f.maxdk).astype(int)
    for i in range(g.NX):
        for j in range(g.NY):
            if 0 < n[i, j] < f.nshells and f.kh[n[i, j]] <= 2.5 and g.RANGE[i, j, 0]:
                assert f.forcex[i, j, 0] == uh[0][i, j, 0] * f.factor[n[i, j]]
    # off the kz = 0 plane only the vertical modes kx = ky = 0, kz = +-1..+-5 carry force
    off = f.forcex[:, :, 1:].copy()
    vert = [k - 1 for k in (1, 2, 3, 4, 5, g.NZ - 5, g.NZ - 4, g.NZ - 3, g.NZ - 2, g.NZ - 1)]
    assert np.all(off[0, 0, vert] != 0)
    off[0, 0, vert] = 0
    assert not np.any(off)


def test_rf_run_injects_energy_where_the_target_is_above():
    # a run whose target spectrum exceeds the current one gains kz = 0 energy
    g = Grid(16, 16, 16)
    uh = np.array(seeded_velocity_hat(g, 4, k0=2.0))
    f = PlaneForcing(g, "rf", kf=2.5, tf=1.0, alphac=0.5)
    sim = OracleSim(g, 1e-3, dt=1e-3, variable_dt=False, cfl=0.5, plane_forcing=f)
    sim.set_velocity(np.array([g.irfft(c) for c in uh]))
    sim.initialize()
    f.Em = 4 * f.Em
    e0 = f.u1u2_spectrum(sim.uh)
    sim.advance_in_time(20)
    e1 = f.u1u2_spectrum(sim.uh)
    forced = [i for i in range(1, f.nshells) if f.kh[i] <= 2.5]
    assert sum(e1[i] for i in forced) > sum(e0[i] for i in forced)


def plane_rhs(g, uh, c, nu):
    """a right-hand side whose kz = 0 plane is (c + nu k^2) u, so rhs - nu k^2 u = c u"""
    k2 = (g.kx**2 + g.ky**2)
    return np.array([(c + nu * k2) * uh[q] for q in range(3)])


def test_nn_spectrum_closed_form():
    g, uh = field()
    f = PlaneForcing(g, "nrf", kf=2.5, tf=1.0, alphac=0.5)
    nu, c = 3e-3, -0.7
    N = f.nn_spectrum(uh, plane_rhs(g, uh, c, nu), nu)
    Ef = f.u1u2_spectrum(uh)
    assert np.allclose(N, 2 * c * Ef, rtol=1e-12, atol=1e-14 * np.abs(Ef).max())
    # the viscous term enters with its sign: rhs = 0 gives -nu sum w k^2 |u|^2 per shell
    N0 = f.nn_spectrum(uh, np.zeros_like(uh), nu)
    k2 = (g.kx**2 + g.ky**2)[:, :, 0]
    n = np.rint(np.sqrt(k2) / f.maxdk).astype(int)
    want = np.zeros(f.nshells)
    for i in g.XRANGE:
        for j in g.YRANGE:
            if n[i, j] < f.nshells:
                w = 1.0 if i == 0 else 2.0
                want[n[i, j]] -= w * nu * k2[i, j] * (abs(uh[0][i, j, 0]) ** 2 + abs(uh[1][i, j, 0]) ** 2)
    assert np.allclose(N0, want, rtol=1e-12, atol=1e-16)


@pytest.mark.parametrize("kind", ["crf", "cnrf"])
def test_c_variants_one_step_by_hand(kind):
    # ts = Tf; w = 2 pi / Tf (cRf) or 2 pi / Tf^2 (cNRf); N = 0 for the N variant
    g, uh = field()
    tf, dt = 1.5, 2e-3
    f = PlaneForcing(g, kind, kf=2.5, tf=tf, alphac=0.5)
    Ef = f.u1u2_spectrum(uh)
    f.Em = 2 * Ef
    f.init = False
    f(_Sim(g, uh, dt=dt, rhs=plane_rhs(g, uh, 0.0, 0.0), it=2))
    omega = 2 * math.pi / tf if kind == "crf" else 2 * math.pi / (tf * tf)
    for i in range(1, f.nshells):
        if f.kh[i] <= 2.5:
            R = dt * omega * omega * Ef[i]
            assert abs(f.R[i] - R) <= 1e-14 * R


@pytest.mark.parametrize("plain,withn", [("rf", "nrf"), ("crf", "cnrf")])
def test_n_variants_reduce_without_n(plain, withn):
    # cNRf's w is 2 pi / Tf^2 where cRf's is 2 pi / Tf: equal at Tf = 1
    g, uh = field()
    out = []
    for kind in (plain, withn):
        f = PlaneForcing(g, kind, kf=2.5, tf=1.0, alphac=0.5)
        f.Em = 2 * f.u1u2_spectrum(uh)
        f.init = False
        f(_Sim(g, uh, dt=2e-3, cfl=0.4, rhs=plane_rhs(g, uh, 0.0, 2e-3), nu=2e-3))
        out.append(f)
    a, b = out
    assert not np.any(b.N)
    assert np.array_equal(a.R, b.R) and np.array_equal(a.factor, b.factor)
    assert np.array_equal(a.forcex, b.forcex) and np.array_equal(a.forcey, b.forcey)
    assert a.hp == b.hp and a.vp == b.vp


def test_nrf_one_step_by_hand():
    g, uh = field()
    dt, dt2, cfl, alpha, tf, c = 2e-3, 1.7e-3, 0.4, 0.5, 2.0, -0.5
    f = PlaneForcing(g, "nrf", kf=2.5, tf=tf, alphac=alpha)
    Ef = f.u1u2_spectrum(uh)
    f.Em = 2 * Ef
    f.init = False
    prev = 0.3 * Ef
    f.N = prev.copy()
    f(_Sim(g, uh, dt=dt, cfl=cfl, it=10, rhs=plane_rhs(g, uh, c, 0.0), dt2=dt2))
    omega = 2 * math.pi / (tf * (dt / cfl))
    assert np.array_equal(f.Nm1, prev)
    assert np.allclose(f.N, 2 * c * Ef, rtol=1e-12, atol=0)
    for i in range(1, f.nshells):
        if f.kh[i] <= 2.5:
            N = 2 * c * Ef[i]
            R = dt * (-2 * alpha * omega * N - (N - prev[i]) / dt2 + omega * omega * Ef[i])
            assert R > 0
            assert abs(f.R[i] - R) <= 1e-12 * R


def test_nrf_initialize_sets_n_and_nm1():
    g, uh = field()
    f = PlaneForcing(g, "nrf", kf=2.5, tf=1.0, alphac=0.5)
    f.initialize(_Sim(g, uh, 1e-3, rhs=plane_rhs(g, uh, 0.25, 0.0)))
    assert np.array_equal(f.Em, f.u1u2_spectrum(uh))
    assert np.allclose(f.N, 0.5 * f.Em, rtol=1e-12, atol=0) and np.array_equal(f.Nm1, f.N)


def test_a_mode_energies_sum_to_the_shell_spectrum():
    g, uh = field()
    f = PlaneForcing(g, "a", kf=3.0, tf=1.0, alphac=0.5)
    f.initialize(_Sim(g, uh, 1e-3))
    Ef = f.u1u2_spectrum(uh)
    k = np.sqrt(g.kx**2 + g.ky**2)[:, :, 0]
    n = np.rint(k / f.maxdk).astype(int)
    tot = np.zeros(f.nshells)
    for i in g.XRANGE:
        for j in g.YRANGE:
            if n[i, j] < f.nshells:
                tot[n[i, j]] += f.Em[i, j]
    assert np.allclose(tot, Ef, rtol=1e-13, atol=0)
    # at its target every oscillator stays at rest and nothing is forced
    f(_Sim(g, uh, dt=1e-3))
    assert not np.any(f.R) and f.hp == 0 and not np.any(f.forcex)


def test_a_one_step_by_hand():
    g, uh = field()
    tf, dt, kf = 1.5, 2e-3, 3.0
    f = PlaneForcing(g, "a", kf=kf, tf=tf, alphac=0.5)
    f.initialize(_Sim(g, uh, dt))
    E = f.Em.copy()
    f.Em = 2 * E
    f(_Sim(g, uh, dt=dt))
    omega = 2 * math.pi / (tf * 10 * dt)
    hp = 0.0
    forced = 0
    for i in g.XRANGE:
        for j in g.YRANGE:
            k = math.hypot(g.KX[i], g.KY[j])
            if 0 < k <= kf:
                forced += 1
                R = dt * omega * omega * E[i, j]
                assert abs(f.R[i, j] - R) <= 1e-14 * R
                fac = math.sqrt(R / E[i, j]) * dt
                assert abs(f.forcex[i, j, 0] - uh[0][i, j, 0] * fac) <= 1e-14 * abs(uh[0][i, j, 0] * fac)
                hp += (0.5 * fac * fac * 2 * E[i, j] + fac * 2 * E[i, j]) / dt
            else:
                assert f.R[i, j] == 0 and f.forcex[i, j, 0] == 0
    assert forced > 10
    assert abs(f.hp - hp) <= 1e-12 * hp
    assert f.vp > 0 and abs(f.vp / f.hp - 0.01) < 0.01   # the vertical modes take ~1% of the power
