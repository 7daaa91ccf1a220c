"""The reference's kz = 0-plane forcing schemes, restated loop for loop (TEST INFRASTRUCTURE ONLY).

  compute_shells2D        utils.jl:161-182
  calculate_Zf            utils.jl:184-197
  calculate_u1u2_spectrum!      Forcing_methods/spectrum.jl:1-20
  calculate_u1u2_mag_spectrum!  Forcing_methods/spectrum.jl:44-65
  RfForcing functor       Forcing_methods/forcing.jl:6-114
  CRfForcing functor      Forcing_methods/cforcing.jl:1-97
  NRfForcing functor      Forcing_methods/nrforcing.jl:1-117, calculate_nn_spectrum! cnrforcing.jl:103-126
  CNRfForcing functor     Forcing_methods/cnrforcing.jl:1-101
  AForcing functor        Forcing_methods/aforcing.jl:1-87, calculate_abs2! spectrum.jl:22-37,
                          calculate_Zf(d,kf,kx,ky) utils.jl:199-218
  initialize!(::RfForcing)      Forcing_methods/forcing_types.jl:66-74
  build_Qf_power_spectrum!      Forcing_methods/forcing_types.jl:304-331
  QfForcing functor       Forcing_methods/qfforcing.jl:1-81
  add_Rf_forcing!         Forcing_methods/forcing_types.jl:54-64
Indices below are 0-based; the reference's 1-based shell n is n0 + 1.

Parity status: pinned by hand-computed cases and by the closed forms the
definitions fix (tests/test_plane_forcing.py): the shells of a small grid by
brute force, the QfForcing powers hp = 0.99 sum R and vp = hp 0.01/0.99, the
oscillator's fixed point R = 0 when Ef = Em, one rf step by hand.
"""
import math
from fractions import Fraction

import numpy as np

from .ffs_oracle import proj

__all__ = ["compute_shells2d", "PlaneForcing"]


def compute_shells2d(kx, ky, nx, ny):
    """(nShells2D, maxdk2D, numPtsInShell2D, kh) of compute_shells2D(kx, ky, Nx, Ny)."""
    nshells = min(nx, ny // 2)
    maxdk = max(kx[1], ky[1])
    kh = np.zeros(nshells)
    num = np.zeros(nshells, dtype=int)
    for j in range(ny):
        for i in range(nx):
            K = math.sqrt(kx[i] ** 2 + ky[j] ** 2)
            ii = int(round_half_even(K / maxdk))
            if ii >= nshells:
                break
            kh[ii] += K
            num[ii] += 1
    with np.errstate(invalid="ignore", divide="ignore"):
        kh = kh / num
    return nshells, maxdk, num, kh


def fma(a, b, c):
    """a*b + c rounded once (Julia's muladd where it fuses; exact via rationals)."""
    return float(Fraction(a) * Fraction(b) + Fraction(c))


def round_half_even(x):
    """Julia's round(Int, x) (RoundNearest: ties to even)."""
    return np.rint(x)


class PlaneForcing:
    """kind "rf", "crf", "nrf", "cnrf", "a" (rfForcing, cRfForcing, nRfForcing, cNRfForcing, aForcing:
    TF, alphac, kf, Zf) or "qf" (QfForcing: alveliusP, alveliusC, alveliusKf, kf)."""

    def __init__(self, grid, kind, kf, tf=0.0, alphac=0.0, zf="cutoff", P=0.0, C=1.0, Kf_alv=0.0, les_delta=None):
        self.g = grid
        self.kind = kind
        self.kf = kf
        self.tf = tf
        self.alpha = alphac
        self.nshells, self.maxdk, self.num, self.kh = compute_shells2d(grid.KX, grid.KY, grid.NX, grid.NY)
        n = self.nshells
        self.Ef = np.zeros(n)
        self.Em = np.zeros(n)
        self.R = np.zeros(n)
        self.N = np.zeros(n)
        self.Nm1 = np.zeros(n)
        self.factor = np.zeros(n)
        if kind == "a":  # one oscillator per (i, j) of the kz = 0 plane
            shape = (grid.NX, grid.NY)
            self.Em = np.zeros(shape)
            self.R = np.zeros(shape)
            self.factor = np.zeros(shape)
        self.init = True
        self.hp = 0.0
        self.vp = 0.0
        self.Zf = np.zeros(n)
        for i, k in enumerate(self.kh):
            if zf == "cutoff":
                self.Zf[i] = float(k <= kf)
            else:
                self.Zf[i] = math.tanh((kf - k) / (0.25 * kf)) * (0.0 if (kf - k) <= 0.0 else 1.0)
        if kind == "a":
            self.Zfa = np.zeros((grid.NX, grid.NY))
            for j in range(grid.NY):
                for i in range(grid.NX):
                    k = math.sqrt(grid.KX[i] ** 2 + grid.KY[j] ** 2)
                    if zf == "tanh":
                        self.Zfa[i, j] = math.tanh((kf - k) / (0.25 * kf)) * (0.0 if (kf - k) <= 0.0 else 1.0)
                    else:
                        self.Zfa[i, j] = float(k <= kf)
        if kind == "qf":
            self.build_qf(P, C, Kf_alv, les_delta)
        self.forcex = np.zeros((grid.NX, grid.NY, grid.NZ), dtype=complex)
        self.forcey = np.zeros_like(self.forcex)

    def build_qf(self, P, C, Kf_alv, les_delta):
        integral = 0.0
        self.R[:] = 0.0
        for i in range(1, self.nshells):
            kh = self.kh[i]
            if kh <= self.kf:
                term = math.exp(-((kh - Kf_alv) / C) ** 2)
                self.R[i] = P * term
                integral += term
        self.R /= integral
        if les_delta is not None:
            d2 = les_delta**2
            for i in range(1, self.nshells):
                kh = self.kh[i]
                if kh <= self.kf:
                    G = math.exp(-d2 / 24 * (kh * kh))
                    self.R[i] *= G * G

    # ---------------------------------------------------------------- spectra of the kz = 0 plane
    def u1u2_spectrum(self, u):
        g = self.g
        Ef = np.zeros(self.nshells)
        maxdk2d = max(g.KX[1], g.KY[1])
        for j in g.YRANGE:
            cf = 1.0
            for i in g.XRANGE:
                k = math.sqrt(g.KX[i] ** 2 + g.KY[j] ** 2)
                n = int(round_half_even(k / maxdk2d))
                if n < self.nshells:
                    Ef[n] += 0.5 * cf * (abs(u[0][i, j, 0]) ** 2 + abs(u[1][i, j, 0]) ** 2)
                cf = 2.0
        return Ef

    def u1u2_mag_spectrum(self, u):
        g = self.g
        Ef = np.zeros(self.nshells)
        maxdk2d = max(g.KX[1], g.KY[1])
        for j in g.YRANGE:
            cf = 1.0
            for i in g.XRANGE:
                k = math.sqrt(g.KX[i] ** 2 + g.KY[j] ** 2)
                n = int(round_half_even(k / maxdk2d))
                if n < self.nshells:
                    Ef[n] += cf * math.sqrt(abs(u[0][i, j, 0]) ** 2 + abs(u[1][i, j, 0]) ** 2)
                cf = 2.0
        return Ef

    def nn_spectrum(self, u, rhs, nu):
        """calculate_nn_spectrum! (cnrforcing.jl:103-126) on the kz = 0 plane."""
        g = self.g
        N = np.zeros(self.nshells)
        maxdk2d = max(g.KX[1], g.KY[1])
        for j in g.YRANGE:
            cf = 1.0
            for i in g.XRANGE:
                kxy2 = g.KX[i] ** 2 + g.KY[j] ** 2
                k2 = kxy2 + g.KZ[0] ** 2
                k = math.sqrt(kxy2)
                n = int(round_half_even(k / maxdk2d))
                if n < self.nshells:
                    magsq = (proj(u[0][i, j, 0], rhs[0][i, j, 0] - nu * k2 * u[0][i, j, 0])
                             + proj(u[1][i, j, 0], rhs[1][i, j, 0] - nu * k2 * u[1][i, j, 0]))
                    N[n] += cf * magsq
                cf = 2.0
        return N

    def abs2_plane(self, u):
        """calculate_abs2! (spectrum.jl:22-37)."""
        g = self.g
        E = np.zeros((g.NX, g.NY))
        for j in g.YRANGE:
            cf = 1.0
            for i in g.XRANGE:
                E[i, j] = 0.5 * cf * (abs(u[0][i, j, 0]) ** 2 + abs(u[1][i, j, 0]) ** 2)
                cf = 2.0
        return E

    def initialize(self, sim):
        if self.kind == "a":
            self.Em = self.abs2_plane(sim.uh)
        elif self.kind != "qf" and self.init:
            self.Em = self.u1u2_spectrum(sim.uh)
            if self.kind in ("nrf", "cnrf"):
                self.N = self.nn_spectrum(sim.uh, sim.rhs, sim.nu)
                self.Nm1 = self.N.copy()

    # ---------------------------------------------------------------- the functors
    def __call__(self, sim):
        if self.kind == "qf":
            self.qf(sim)
        elif self.kind == "a":
            self.a(sim)
        else:
            self.rf(sim)

    def _elements(self):
        """(i, j, n0) of the forced kz = 0 elements in the reference's loop order."""
        g = self.g
        for j in g.YRANGE:
            first = True
            for i in g.XRANGE:
                k = math.sqrt(fma(g.KX[i], g.KX[i], g.KY[j] ** 2))
                n = int(round_half_even(k / self.maxdk))
                if 0 < n < self.nshells and self.kh[n] <= self.kf:
                    yield i, j, n, (1.0 if first else 2.0)
                else:
                    yield i, j, -1, (1.0 if first else 2.0)
                first = False

    def rf(self, sim):
        """rfForcing, cRfForcing, nRfForcing, cNRfForcing (they differ in ts and the N terms)."""
        g = self.g
        dt = sim.dt
        if self.kind in ("crf", "cnrf"):
            ts = self.tf
        else:
            ts = dt / sim.cfl
            it = sim.iteration
            if sim.variable_dt:
                if it == 0:
                    ts /= 0.1
                elif it == 1:
                    ts /= 0.2
                elif it == 2:
                    ts /= 0.4
                elif it == 3:
                    ts /= 0.8
        omega = 2 * math.pi / ts if self.kind == "crf" else 2 * math.pi / (self.tf * ts)
        eps = np.finfo(np.float64).eps
        s1 = self.forcex
        s2 = self.forcey
        s1[:] = 0
        s2[:] = 0
        u1, u2 = sim.uh[0], sim.uh[1]
        self.Ef = self.u1u2_spectrum(sim.uh)
        with_n = self.kind in ("nrf", "cnrf")
        if with_n:
            self.Nm1 = self.N.copy()
            self.N = self.nn_spectrum(sim.uh, sim.rhs, sim.nu)
            dtm2 = sim.dt2
        for i in range(1, self.nshells):
            if self.kh[i] <= self.kf:
                if with_n:
                    self.R[i] += dt * (-2 * self.alpha * omega * (self.R[i] + self.N[i])
                                       - (self.N[i] - self.Nm1[i]) / dtm2 - omega * omega * (self.Ef[i] - self.Em[i]))
                else:
                    self.R[i] += dt * (-2 * self.alpha * omega * self.R[i]
                                       - omega * omega * (self.Ef[i] - self.Em[i]))
                self.R[i] = max(0.0, self.R[i])
                self.factor[i] = math.sqrt(self.R[i] / max(self.Ef[i], eps)) * self.Zf[i] * dt
        ff = 0.0
        fv = 0.0
        for i, j, n, cf in self._elements():
            if n < 0:
                continue
            s1[i, j, 0] = u1[i, j, 0] * self.factor[n]
            s2[i, j, 0] = u2[i, j, 0] * self.factor[n]
            ff += (abs(s1[i, j, 0]) ** 2 + abs(s2[i, j, 0]) ** 2) * cf
            fv += (proj(u1[i, j, 0], s1[i, j, 0]) + proj(u2[i, j, 0], s2[i, j, 0])) * cf
        ff = ff * 0.5 / dt
        fv = fv / dt
        hp = ff + fv
        self.hp = hp
        self._vertical_rf(sim, hp)

    def a(self, sim):
        """aForcing: one oscillator per kz = 0 mode with 0 < |k_h| <= kf, ts = 10 dt."""
        g = self.g
        dt = sim.dt
        omega = 2 * math.pi / (self.tf * (10 * dt))
        eps = np.finfo(np.float64).eps
        s1 = self.forcex
        s2 = self.forcey
        s1[:] = 0
        s2[:] = 0
        u1, u2 = sim.uh[0], sim.uh[1]
        for j in g.YRANGE:
            for i in g.XRANGE:
                k = math.sqrt(fma(g.KX[i], g.KX[i], g.KY[j] ** 2))
                if 0 < k <= self.kf:
                    Ef = 0.5 * (1 + (i > 0)) * (abs(u1[i, j, 0]) ** 2 + abs(u2[i, j, 0]) ** 2)
                    self.R[i, j] += dt * (-2 * self.alpha * omega * self.R[i, j] - omega * omega * (Ef - self.Em[i, j]))
                    self.R[i, j] = max(0.0, self.R[i, j])
                    self.factor[i, j] = math.sqrt(self.R[i, j] / max(Ef, eps)) * self.Zfa[i, j] * dt
        ff = 0.0
        fv = 0.0
        for j in g.YRANGE:
            cf = 1.0
            for i in g.XRANGE:
                k = math.sqrt(fma(g.KX[i], g.KX[i], g.KY[j] ** 2))
                if 0 < k <= self.kf:
                    s1[i, j, 0] = u1[i, j, 0] * self.factor[i, j]
                    s2[i, j, 0] = u2[i, j, 0] * self.factor[i, j]
                    ff += (abs(s1[i, j, 0]) ** 2 + abs(s2[i, j, 0]) ** 2) * cf
                    fv += (proj(u1[i, j, 0], s1[i, j, 0]) + proj(u2[i, j, 0], s2[i, j, 0])) * cf
                cf = 2.0
        hp = ff * 0.5 / dt + fv / dt
        self.hp = hp
        self._vertical_rf(sim, hp)

    def _vertical_rf(self, sim, hp):
        g = self.g
        dt = sim.dt
        s1 = self.forcex
        s2 = self.forcey
        u1, u2 = sim.uh[0], sim.uh[1]
        NZ = g.NZ
        umag = 0.0
        for k in range(1, 6):  # 1-based 2:6 -> kz = 1..5 and -1..-5
            umag += math.sqrt(abs(u1[0, 0, k]) ** 2 + abs(u2[0, 0, k]) ** 2) + \
                math.sqrt(abs(u1[0, 0, NZ - k]) ** 2 + abs(u2[0, 0, NZ - k]) ** 2)
        f = dt * math.sqrt(2.0 * hp * (1.0 - 0.99)) / umag
        ff = 0.0
        fv = 0.0
        for k in range(1, 6):
            for kk in (k, NZ - k):
                s1[0, 0, kk] = u1[0, 0, kk] * f
                s2[0, 0, kk] = u2[0, 0, kk] * f
            ff += (abs(s1[0, 0, k]) ** 2 + abs(s2[0, 0, k]) ** 2 + abs(s1[0, 0, NZ - k]) ** 2
                   + abs(s2[0, 0, NZ - k]) ** 2)
            fv += (proj(u1[0, 0, k], s1[0, 0, k]) + proj(u2[0, 0, k], s2[0, 0, k])
                   + proj(u1[0, 0, NZ - k], s1[0, 0, NZ - k]) + proj(u2[0, 0, NZ - k], s2[0, 0, NZ - k]))
        self.vp = ff * 0.5 / dt + fv / dt

    def qf(self, sim):
        g = self.g
        dt = sim.dt
        s1 = self.forcex
        s2 = self.forcey
        s1[:] = 0
        s2[:] = 0
        u1, u2 = sim.uh[0], sim.uh[1]
        self.Ef = self.u1u2_mag_spectrum(sim.uh)
        fv = 0.0
        for i, j, n, cf in self._elements():
            if n < 0:
                continue
            umag = math.sqrt(abs(u1[i, j, 0]) ** 2 + abs(u2[i, j, 0]) ** 2)
            fkn = dt * self.R[n] * 0.99 / self.Ef[n]
            s1[i, j, 0] = (u1[i, j, 0] / umag) * fkn
            s2[i, j, 0] = (u2[i, j, 0] / umag) * fkn
            fv += (proj(u1[i, j, 0], s1[i, j, 0] / dt) + proj(u2[i, j, 0], s2[i, j, 0] / dt)) * cf
        hp = fv
        self.hp = hp
        NZ = g.NZ
        umagsq = 0.0
        for k in range(3, 6):  # 1-based 4:6 -> kz = 3..5 and -3..-5
            umagsq += (abs(u1[0, 0, k]) ** 2 + abs(u2[0, 0, k]) ** 2 + abs(u1[0, 0, NZ - k]) ** 2
                       + abs(u2[0, 0, NZ - k]) ** 2)
        f = dt * hp * (1 - 0.99) / (umagsq * 0.99)
        fv = 0.0
        for k in range(3, 6):
            for kk in (k, NZ - k):
                s1[0, 0, kk] = u1[0, 0, kk] * f
                s2[0, 0, kk] = u2[0, 0, kk] * f
            fv += (proj(u1[0, 0, k], s1[0, 0, k] / dt) + proj(u2[0, 0, k], s2[0, 0, k] / dt)
                   + proj(u1[0, 0, NZ - k], s1[0, 0, NZ - k] / dt) + proj(u2[0, 0, NZ - k], s2[0, 0, NZ - k] / dt))
        self.vp = fv

    def add(self, u):
        """add_Rf_forcing! on u1, u2: the XRANGE x YRANGE part of the kz = 0 plane and the modes
        kx = ky = 0, kz = +-1..+-5."""
        g = self.g
        NZ = g.NZ
        ix = np.ix_(g.XRANGE, g.YRANGE)
        for c, f in ((0, self.forcex), (1, self.forcey)):
            u[c][:, :, 0][ix] += f[:, :, 0][ix]
            for k in (1, 2, 3, 4, 5, NZ - 5, NZ - 4, NZ - 3, NZ - 2, NZ - 1):
                u[c][0, 0, k] += f[0, 0, k]
        return u
