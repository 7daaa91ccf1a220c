"""The oracle's hyperviscosity and spectral barrier (Hyperviscosity/
hyperviscosity_types.jl, init_c! ETD.jl:56-95, models.jl:298-322), pinned by
closed forms: a single shear mode (no nonlinear transfer) decays under ETD3 as
exp((-nu k^2 - nu_h k^2M) t) exactly, and under the barrier as
exp((-nu k^2 + log esp(k)) t), vanishing in one step beyond endk and feeling
only nu below initk; the AB3 right-hand side of that mode is
(-nu k^2 - nu_h k^2M) u; the energy budget dk/dt = -(diss + hdiss) holds for a
random field; the hvis spectrum sums to hdiss (to -hdiss for the barrier, whose
spectrum keeps the sign of f); the parser maps the keys."""
import math

import numpy as np
import pytest

from oracle.ffs_oracle import Grid, OracleSim
from ffs.initial import random_spectrum
from ffs import config

N = 32


def shear_mode(g, n):
    """u = (sin(n y), 0, 0): divergence free, u x w is a pure gradient, so the step is linear"""
    x, y, z = g.coords()
    return np.array([np.sin(n * y), np.zeros_like(y), np.zeros_like(y)])


def amplitude(s, n):
    return abs(s.uh[0][0, n, 0])


def esp(k, i, p):
    return 1.0 if k <= i else (math.exp(-((k - i) ** 2) / (p ** 0.25 - k ** 0.25)) if k < p else 0.0)


@pytest.mark.parametrize("hv,n", [(dict(kind="hyper", nu=2e-6, m=2), 9), (dict(kind="hyper", nu=1e-9, m=3), 7),
                                  (dict(kind="barrier", initk=5.0, endk=12.0), 9),
                                  (dict(kind="barrier", initk=5.0, endk=12.0), 4)])
def test_single_mode_etd_decay(hv, n):
    g = Grid(N, N, N)
    nu, dt = 3e-3, 0.01
    s = OracleSim(g, nu, dt=dt, integrator="ETD3", hv=hv)
    s.set_velocity(shear_mode(g, n))
    s.initialize()
    a0 = amplitude(s, n)
    s.advance_in_time(7)
    k = float(n)
    if hv["kind"] == "hyper":
        c = -nu * k * k - hv["nu"] * k ** (2 * hv["m"])
    else:
        c = -nu * k * k + math.log(esp(k, hv["initk"], hv["endk"]))
    want = a0 * math.exp(c * 7 * dt)
    assert abs(amplitude(s, n) - want) <= 1e-12 * a0, (amplitude(s, n), want)
    if hv["kind"] == "barrier" and n <= hv["initk"]:
        assert abs(c + nu * k * k) == 0.0   # below initk the barrier adds nothing


def test_barrier_removes_modes_beyond_endk():
    g = Grid(N, N, N)
    s = OracleSim(g, 3e-3, dt=0.01, integrator="ETD3", hv=dict(kind="barrier", initk=5.0, endk=8.5))
    s.set_velocity(shear_mode(g, 9) + shear_mode(g, 3))
    s.initialize()
    s.advance_in_time(1)
    assert amplitude(s, 9) == 0.0
    assert amplitude(s, 3) > 0.1


def test_ab3_rhs_of_a_single_mode():
    g = Grid(N, N, N)
    nu, nuh, m, n = 3e-3, 2e-6, 2, 9
    s = OracleSim(g, nu, dt=0.01, hv=dict(kind="hyper", nu=nuh, m=m))
    s.set_velocity(shear_mode(g, n))
    s.initialize()
    k2 = float(n * n)
    want = (-nu * k2 - nuh * k2 ** m) * s.uh[0][0, n, 0]
    assert abs(s.rhs[0][0, n, 0] - want) <= 1e-13 * abs(want)


# the barrier is stiff (-f grows without bound towards endk): a smaller step keeps the central difference accurate
@pytest.mark.parametrize("hv,integrator,dt", [(dict(kind="hyper", nu=2e-5, m=2), "ETD3", 1e-3),
                                              (dict(kind="hyper", nu=2e-5, m=2), "AB3", 1e-3),
                                              (dict(kind="barrier", initk=8.0, endk=60.0), "ETD3", 1e-4)])
def test_energy_budget_with_hyperviscosity(hv, integrator, dt):
    g = Grid(N, N, N)
    s = OracleSim(g, 0.005, dt=dt, dt_stat=1, dt_spec=1, integrator=integrator, hv=hv)
    s.set_velocity(random_spectrum(N, seed=9, k0=6.0))
    s.initialize()
    s.advance_in_time(6)
    k = np.array([r[0][11] for r in s.records])
    tdiss = np.array([r[0][21] for r in s.records])
    hd = np.array([r[0][-1] for r in s.records])
    dk = (k[4] - k[2]) / (2 * dt)
    assert hd[3] > 0.05 * tdiss[3]                      # the hyperviscous part matters here
    assert abs(dk + tdiss[3]) < 3e-3 * tdiss[3]
    for r in s.records:                                  # hdiss = hdissh + hdissv, tdiss = diss + hdiss
        st = r[0]
        assert st[-1] == st[-3] + st[-2]
        assert abs(st[21] - (st[14] + st[-1])) <= 1e-15 * st[21]
    # the hvis spectrum sums to hdiss (unit box: specV bins carry the Parseval weights); the barrier's
    # spectrum carries f(|k|) itself, its stats -f (spec.jl:270-295 against hyperviscosity_types.jl:50-79)
    it, q = s.spectra[-1]
    st = s.records[-1][0]
    assert it == s.records[-1][0][0]
    sign = 1.0 if hv["kind"] == "hyper" else -1.0
    for part, col in (("hvis_h", -3), ("hvis_v", -2)):
        assert abs(sign * q[part][1].sum() - st[col]) <= 1e-12 * abs(st[col])


def test_keys():
    base = dict(nx="16", ny="16", nz="16", xDomainSize="1.0", yDomainSize="1.0", zDomainSize="1.0",
                kinematicViscosity="0.01", dtStat="1", dtSpec="1", writeTime="1")
    p = config.parameters(dict(base, hyperViscosity="true", hyperViscosityCoefficient="1e-6"))
    c = p.to_c()
    assert (c.hyperviscosity, c.hv_nu, c.hv_m) == (1, 1e-6, 2)
    p = config.parameters(dict(base, hyperViscosity="x", hyperViscosityCoefficient="2e-9", hyperViscosityM="3"))
    assert p.to_c().hv_m == 3
    with pytest.raises(KeyError):   # the coefficient is required (simulationtypes.jl:299)
        config.parameters(dict(base, hyperViscosity="true"))
    p = config.parameters(dict(base, hyperViscosity="spectralBarrier", initk="4", endk="7.5"))
    c = p.to_c()
    assert (c.hyperviscosity, c.sb_initk, c.sb_endk) == (2, 4.0, 7.5)
    assert config.parameters(dict(base, hyperViscosity="spectralBarrier")).to_c().sb_initk == 0.0
