"""Pins of the oracle's set_dt! and Scales/Numbers rows by closed forms.

set_dt! (time_step_functions.jl:86-129, Adams_Bashforth.jl:166-176, ETD.jl:110-114) runs on
exactly steady states of the inviscid equations whose max|u_i| and max|rho| are known: a shear
mode u = (A sin y, 0, 0) (u x w = grad(A^2 sin^2(y)/2) is removed by the projection) and, with
the Boussinesq density, rho = B sin z (rho g e_z is a gradient, u . grad rho = 0, u_z = 0).
The expected dt of every iteration is written from the reference's rules as numbers: the CFL ramp
0.1/0.2/0.4/0.8 at iterations 0-3, AB3's halving at its first step, the x1.25 clamp in both
directions after iteration 3 and the density limit cfl sqrt(dx/(rho_max |g|))).

The Scales/Numbers rows (stats.jl:64-165,952-980) are checked against physical-space
quantities of analytic fields: the integral lengths are means of squared line averages
(Parseval on the kx = 0 plane and the ky = 0 / kz = 0 lines), the Taylor lengths and the
dissipation come from analytic derivatives evaluated on the grid; no transform is involved.
The Taylor-Green t = 0 case gives closed forms: k = 1/8, eps = 6 nu k, every Taylor length
sqrt(2) and vanishing integral lengths."""
import math

import numpy as np

from oracle.ffs_oracle import Grid, OracleSim, taylor_green


def shear(g, amp):
    x, y, z = g.coords()
    return np.array([amp * np.sin(y / g.LY), np.zeros_like(x), np.zeros_like(x)])


def test_set_dt_ab3_ramp_halving_and_clamps():
    n, cfl, amp = 16, 0.5, 2.0
    g = Grid(n, n, n)
    s = OracleSim(g, 0.0, variable_dt=True, cfl=cfl, integrator="AB3")
    s.set_velocity(shear(g, amp))
    s.initialize()
    dx = 2 * math.pi / n
    base = cfl * dx / amp  # cfl dx / max|u_i| with the ramp factor 1
    # iteration 0: 0.1 base halved by AB3's first step, dt2 = dt3 = dt
    assert s.umax == amp or abs(s.umax - amp) < 1e-14 * amp
    hist = []
    for it in range(5):
        s.advance_in_time(1)
        hist.append((s.dt, s.dt2, s.dt3))
    want = [0.1 * base / 2, 0.2 * base, 0.4 * base, 0.8 * base, 1.0 * base]
    for it, w in enumerate(want):
        assert abs(hist[it][0] - w) < 1e-13 * w, (it, hist[it][0], w)
    # the history rotates: dt2 <- dt, dt3 <- dt2 (all equal at iteration 0)
    assert abs(hist[0][1] - want[0]) < 1e-13 * want[0] and abs(hist[0][2] - want[0]) < 1e-13 * want[0]
    for it in range(1, 5):
        assert hist[it][1] == hist[it - 1][0] and hist[it][2] == hist[it - 1][1]
    # upward clamp: max|u| halves, so the CFL step doubles; dt grows by at most 1.25 per step
    s.uh = s.uh * 0.5
    got = []
    for _ in range(4):
        s.advance_in_time(1)
        got.append(s.dt)
    want = [1.25 * base, 1.5625 * base, 1.953125 * base, 2.0 * base]
    for a, w in zip(got, want):
        assert abs(a - w) < 1e-13 * w, (got, want)
    # downward clamp: max|u| x 8, CFL step base / 4; dt shrinks by at most 1/1.25 per step
    s.uh = s.uh * 8.0
    got = []
    for _ in range(3):
        s.advance_in_time(1)
        got.append(s.dt)
    want = [1.6 * base, 1.28 * base, 1.024 * base]
    for a, w in zip(got, want):
        assert abs(a - w) < 1e-13 * w, (got, want)


def test_set_dt_density_limit_etd3():
    n, cfl, amp, rho_amp, grav = 16, 0.4, 1.0, 10.0, -2.0
    g = Grid(n, n, n)
    s = OracleSim(g, 0.0, variable_dt=True, cfl=cfl, integrator="ETD3",
                  density={"Pr": 1.0, "gradient": 0.0, "g": grav})
    s.set_velocity(shear(g, amp))
    x, y, z = g.coords()
    s.set_density(rho_amp * np.sin(z))
    s.initialize()
    dx = 2 * math.pi / n
    # the density limit cfl sqrt(dx / (rho_max |g|)) is below cfl dx / max|u_i| at every ramp factor
    lim = math.sqrt(dx / (rho_amp * abs(grav)))
    assert lim < dx / amp
    got = []
    for _ in range(5):
        s.advance_in_time(1)
        got.append(s.dt)
    # ETD3 has no first-step halving
    want = [r * cfl * lim for r in (0.1, 0.2, 0.4, 0.8, 1.0)]
    for a, w in zip(got, want):
        assert abs(a - w) < 1e-12 * w, (got, want)
    # with a weak density the velocity limit is the smaller one again
    s.rhoh = s.rhoh * 1e-4
    s.advance_in_time(1)
    w = min(1.25 * want[-1], cfl * dx / amp)
    assert abs(s.dt - w) < 1e-12 * w, (s.dt, w)


def test_scales_numbers_taylor_green_t0():
    n, nu = 32, 0.02
    g = Grid(n, n, n)
    s = OracleSim(g, nu, dt=0.01, dt_stat=1)
    s.set_velocity(taylor_green(g))
    s.initialize()
    s.advance_in_time(1)
    st, sc, nb = s.records[0]
    k, eps = 0.125, 6 * nu * 0.125
    assert abs(st[11] - k) < 1e-15 and abs(st[21] - eps) < 1e-15  # k, tdiss
    urms, uh = math.sqrt((0.125 + 0.125) / 3), math.sqrt(0.125)
    Lk, Lt = (nu**3 / eps) ** 0.25, k**1.5 / eps
    lg, Tk, Tt = math.sqrt(15 * nu / eps) * urms, math.sqrt(nu / eps), k / eps
    # every squared derivative mean of u1, u2 along x and y is 1/8, so each Taylor length is sqrt 2
    lam = math.sqrt(2.0)
    want = {2: Lk, 3: Lt, 4: lg, 5: lam, 6: lam, 11: lam, 12: lam, 22: Tk, 23: Tt}
    for i, w in want.items():
        assert abs(sc[i] - w) < 1e-13 * w, (i, sc[i], w)
    # u1 and u2 have no kx = 0, ky = 0 or kz = 0 content: zero integral lengths (those of u3 are
    # ratios of roundoff, since <u3^2> is)
    for i in (7, 9, 13, 14, 15, 16, 17, 18, 24):  # Lhh, Lhv, L11..L23, Thh
        assert abs(sc[i]) < 1e-13, (i, sc[i])
    assert st[7] < 1e-30
    # Numbers: Reh, Reh1 vanish with Lhh; the Taylor Reynolds numbers are urms lambda / nu
    assert abs(nb[2]) < 1e-12 and abs(nb[3]) < 1e-12
    assert abs(nb[4] - urms * lam / nu) < 1e-12 * nb[4] and abs(nb[5] - urms * lam / nu) < 1e-12 * nb[5]
    r1 = (lam / Lk) ** (4 / 3)
    assert abs(nb[6] - r1) < 1e-12 * r1 and abs(nb[7] - r1) < 1e-12 * r1
    assert uh > 0


class Field:
    """A divergence-free sum of trigonometric terms on a (2 pi LX, 2 pi LY, 2 pi LZ) box with its
    analytic gradient.  Angles a = x/LX, b = y/LY, c = z/LZ."""

    def __init__(self, g, p):
        x, y, z = g.coords()
        L = (g.LX, g.LY, g.LZ)
        a, b, c = x / L[0], y / L[1], z / L[2]
        a1, a2, b1, b2, c1, c2, q = p
        sa, ca, sb, cb = np.sin(a), np.cos(a), np.sin(b), np.cos(b)
        # u1 = a1 cos b + a2 sin 2c + q LX cos a sin b; u2 = b1 cos a + b2 cos 3c - q LY sin a cos b;
        # u3 = c1 sin 2a + c2 cos b (the q terms form a 2D Taylor-Green cell, divergence-free)
        self.u = np.array([a1 * cb + a2 * np.sin(2 * c) + q * L[0] * ca * sb,
                           b1 * ca + b2 * np.cos(3 * c) - q * L[1] * sa * cb,
                           c1 * np.sin(2 * a) + c2 * cb])
        # grad[i][j] = d u_i / d x_j
        self.grad = np.array([
            [-q * sa * sb, (-a1 * sb + q * L[0] * ca * cb) / L[1], 2 * a2 * np.cos(2 * c) / L[2]],
            [(-b1 * sa - q * L[1] * ca * cb) / L[0], q * L[1] * sa * sb / L[1], -3 * b2 * np.sin(3 * c) / L[2]],
            [2 * c1 * np.cos(2 * a) / L[0], -c2 * sb / L[1], np.zeros_like(a)]])


def test_scales_numbers_physical_space():
    """Integral and Taylor lengths, Reynolds and Froude numbers of an asymmetric field on a
    non-cubic box with unequal domain sizes, against physical-space means."""
    nx, ny, nz, L = 24, 16, 32, (1.5, 1.0, 0.75)
    nu, grav, drhodz = 0.013, 2.0, 0.7
    g = Grid(nx, ny, nz, *L)
    f = Field(g, (0.9, 0.35, 0.6, 0.25, 0.4, 0.7, 0.3))
    d = f.grad
    assert np.abs(d[0][0] + d[1][1] + d[2][2]).max() < 1e-14  # divergence-free
    s = OracleSim(g, nu, dt=1e-3, dt_stat=1, density={"Pr": 1.0, "gradient": drhodz, "g": grav})
    s.set_velocity(f.u)
    s.initialize()
    s.advance_in_time(1)
    st, sc, nb = s.records[0]
    u = f.u
    up2 = [float(np.mean(c * c)) for c in u]
    for i in range(3):
        assert abs(st[5 + i] - up2[i]) < 1e-14, (i, st[5 + i], up2[i])
    k = sum(up2) / 2
    eps = nu * float(sum(np.mean(d[i][j] ** 2) for i in range(3) for j in range(3)))
    assert abs(st[11] - k) < 1e-14 and abs(st[14] - eps) < 1e-14 * eps
    assert st[21] == st[14]  # tdiss = diss (no LES, no hyperviscosity)

    def integral(i, axis):
        m = np.mean(u[i], axis=axis)  # the line average along `axis`: its mean square is that plane's energy
        return L[axis] * float(np.mean(m * m)) * math.pi / up2[i]

    Lij = [[integral(i, j) for j in range(3)] for i in range(3)]
    dsq = [[float(np.mean(d[i][j] ** 2)) for j in range(3)] for i in range(3)]
    l12 = math.sqrt(2 * up2[0] / dsq[0][1])
    l21 = math.sqrt(2 * up2[1] / dsq[1][0])
    ll = math.sqrt(2 * (up2[0] + up2[1]) / (dsq[0][0] + dsq[1][1]))
    lt = math.sqrt(2 * (up2[0] + up2[1]) / (dsq[0][1] + dsq[1][0]))
    urms, uh = math.sqrt(sum(up2) / 3), math.sqrt((up2[0] + up2[1]) / 2)
    Lhh, Lvv = (Lij[0][0] + Lij[1][1]) / 2, Lij[2][2]
    Lhv, Lvh = (Lij[0][2] + Lij[1][2]) / 2, (Lij[2][0] + Lij[2][1]) / 2
    Lk, Lt, lg = (nu**3 / eps) ** 0.25, math.sqrt(k**3) / eps, math.sqrt(15 * nu / eps) * urms
    Tk, Tt, Thh = math.sqrt(nu / eps), k / eps, Lhh / urms
    Nb = math.sqrt(grav * drhodz)
    Lb, Lo = 2 * math.pi * urms / Nb, math.sqrt(eps / Nb**3)
    Lom = math.sqrt(eps / (Nb**3 / (8 * math.pi**3)))
    want_sc = [Lk, Lt, lg, ll, lt, Lhh, Lvv, Lhv, Lvh, l12, l21, *Lij[0], *Lij[1], *Lij[2],
               Lb, Lo, Lom, Tk, Tt, Thh, 2 * math.pi / Nb, 1 / Nb]
    assert len(sc) == 2 + len(want_sc)
    for i, w in enumerate(want_sc):
        assert abs(sc[2 + i] - w) < 1e-12 * abs(w), (i + 2, sc[2 + i], w)
    # the pins distinguish what a slip would swap
    assert abs(l12 - l21) > 0.1 * l12 and abs(Lij[0][1] - Lij[1][0]) > 0.1 * Lij[0][1] and abs(ll - lt) > 0.05 * ll
    want_nb = [uh * Lhh / nu, (Lhh / Lk) ** (4 / 3), urms * lg / nu, urms * ll / nu, urms * lt / nu,
               (lg / Lk) ** (4 / 3), (ll / Lk) ** (4 / 3), (lt / Lk) ** (4 / 3), (Lo / Lk) ** (4 / 3),
               (Lom / Lk) ** (4 / 3), 2 * math.pi * uh / (Nb * Lhh), Lhv / Lhh]
    assert len(nb) == 2 + len(want_nb)
    for i, w in enumerate(want_nb):
        assert abs(nb[2 + i] - w) < 1e-12 * abs(w), (i + 2, nb[2 + i], w)
