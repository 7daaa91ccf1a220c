"""The oracle's passive scalar (PassiveScalar, simulationtypes.jl:198-224,
320-352; models.jl:245-273, 353-357, 470), pinned by closed forms: a scalar
mode carried by a uniform flow U evolves as exp(-i k.U t - alpha k^2 t) (the
advection sign and the ETD3 diffusion), converging at the integrators' order;
the mean-gradient source -G u_dir of a decaying shear mode u_z = A sin x
gives phi = -G A sin x (e^{-nu t} - e^{-alpha t}) / (alpha - nu), and nothing
when the gradient points along x; the stats columns are Parseval means of
the real field; the parser maps the keys of the reference's Scalar_Advection
case."""
import math

import numpy as np
import pytest

from oracle.ffs_oracle import Grid, OracleSim
from ffs import config

N = 16


def uniform_flow_case(dt, integrator, alpha_pr, steps):
    g = Grid(N, N, N)
    x, y, z = g.coords()
    U = (0.3, -0.2, 0.5)
    nu = 0.01
    s = OracleSim(g, nu, dt=dt, integrator=integrator, scalar=dict(Pr=alpha_pr, gradient=0.0, dir=2))
    s.set_velocity(np.array([np.full_like(x, U[0]), np.full_like(x, U[1]), np.full_like(x, U[2])]))
    s.set_scalar(np.cos(2 * x + y - z))
    s.initialize()
    p0 = s.phih[2, 1, N - 1]
    s.advance_in_time(steps)
    k = (2.0, 1.0, -1.0)
    alpha = nu / alpha_pr
    t = s.time
    exact = p0 * np.exp(-1j * (k[0] * U[0] + k[1] * U[1] + k[2] * U[2]) * t - alpha * 6.0 * t)
    return abs(s.phih[2, 1, N - 1] - exact) / abs(p0), s


@pytest.mark.parametrize("integrator,pr", [("AB3", math.inf), ("ETD3", 0.5)])
def test_uniform_advection_converges_to_the_closed_form(integrator, pr):
    e1, s = uniform_flow_case(0.02, integrator, pr, 10)
    e2, _ = uniform_flow_case(0.01, integrator, pr, 20)
    assert e1 < 1e-4
    # the Euler / AB2 (ETD1 / ETD2) start makes the global error second order
    assert math.log2(e1 / e2) > 1.8, (e1, e2)
    # the flow stays uniform: the scalar is passive
    assert np.allclose(s.uh[:, 0, 0, 0].real, (0.3, -0.2, 0.5), rtol=0, atol=1e-15)


def gradient_case(direction, dt, steps):
    g = Grid(N, N, N)
    x, y, z = g.coords()
    nu, pr, G, A = 0.02, 0.5, 0.7, 0.3
    s = OracleSim(g, nu, dt=dt, integrator="ETD3", scalar=dict(Pr=pr, gradient=G, dir=direction))
    s.set_velocity(np.array([np.zeros_like(x), np.zeros_like(x), A * np.sin(x)]))
    s.set_scalar(np.zeros_like(x))
    s.initialize()
    s.advance_in_time(steps)
    t = s.time
    alpha = nu / pr
    phi = g.irfft(s.phih)
    want = -G * A * np.sin(x) * (math.exp(-nu * t) - math.exp(-alpha * t)) / (alpha - nu)
    return phi, want


def test_mean_gradient_source():
    phi1, want = gradient_case(2, 0.01, 20)
    phi2, want2 = gradient_case(2, 0.005, 40)
    assert np.array_equal(want, want2) or np.abs(want - want2).max() < 1e-15
    e1 = np.abs(phi1 - want).max() / np.abs(want).max()
    e2 = np.abs(phi2 - want).max() / np.abs(want).max()
    assert e1 < 1e-5 and math.log2(e1 / e2) > 1.8, (e1, e2)
    # a gradient along x meets u_x = 0: no source
    phi, _ = gradient_case(0, 0.01, 20)
    assert np.abs(phi).max() == 0.0


def test_scalar_stats_are_parseval_means():
    g = Grid(N, N, N)
    x, y, z = g.coords()
    from ffs.initial import random_spectrum, random_scalar

    s = OracleSim(g, 0.01, dt=0.005, dt_stat=1, scalar=dict(Pr=0.7, gradient=0.3, dir=1))
    s.set_velocity(random_spectrum(N, seed=4))
    s.set_scalar(random_scalar(N, N, N, seed=5))
    s.initialize()
    s.advance_in_time(2)
    s.calculate_rhs()
    row = s.stats()[0]  # of the current fields
    phi = g.irfft(s.phih)
    grads = [g.irfft(1j * kk * s.phih) for kk in (g.kx, g.ky, g.kz)]
    assert abs(row[23] - np.mean(phi * phi)) < 1e-13 * row[23]
    for c in range(3):
        assert abs(row[24 + c] - np.mean(grads[c] ** 2)) < 1e-12 * row[24 + c]
    assert abs(row[27] - (0.01 / 0.7) * sum(row[24:27])) <= 1e-15 * row[27]
    assert row[22] == s.phih[0, 0, 0].real


def test_scalar_advection_case_keys():
    # test/Scalar_Advection/global of the reference
    d = dict(kinematicViscosity="0.03333333", xDomainSize="1.0", yDomainSize="1.0", zDomainSize="1.0", nx="128",
             ny="128", nz="128", passiveScalar="true", scalarGradient="0.0", scalarPr="Inf", dtStat="125",
             nt="500", dt="0.002", variableTimestepFlag="false", velocityTimeStep="Adams_Bashforth3rdO",
             dealias="sphere", threaded="true", writeTime="125", dtSpec="1000")
    p = config.parameters(d)
    c = p.to_c()
    assert c.passive_scalar == 1 and math.isinf(c.scalar_pr) and c.scalar_gradient == 0.0 and c.scalar_direction == 2
    assert config.parameters(dict(d, scalarDirection="x")).to_c().scalar_direction == 0
    with pytest.raises(KeyError):
        config.parameters({k: v for k, v in d.items() if k != "scalarPr"})
