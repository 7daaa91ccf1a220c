"""The oracle's conservation form (equation NS: -div(u u) with u u traceless,
models.jl:176-180, 262) against its rotational form: u x w = grad(|u|^2/2) -
div(u u) for a divergence-free field, so after the projection the two
right-hand sides agree to roundoff on every mode whose |k_i| are all below the
2/3 cutoff.  The modes at the cutoff (|k_i| = N/3 is kept by the reference's
box mask) receive the 2k_c aliases of the products, which the two forms alias
differently: there they differ, as they do in the reference."""
import numpy as np

from oracle.ffs_oracle import Grid, OracleSim
from ffs.initial import random_spectrum
from ffs import config

N = 24


def pair(**kw):
    g = Grid(N, N, N)
    u0 = random_spectrum(N, seed=17, k0=4.0)
    out = []
    for eq in ("rotational", "NS"):
        s = OracleSim(g, 5e-3, equation=eq, **kw)
        s.set_velocity(u0)
        s.initialize()
        out.append(s)
    return out


def test_right_hand_sides_agree_below_the_cutoff():
    for kw in (dict(dt=2e-3), dict(dt=2e-3, les_c=0.16)):
        a, b = pair(**kw)
        g = a.g
        kc = max(abs(g.KX[i]) for i in g.XRANGE)
        inner = (np.abs(g.kx) < kc) & (np.abs(g.ky) < kc) & (np.abs(g.kz) < kc)
        d = np.abs(b.rhs - a.rhs).max(axis=0)
        scale = np.abs(a.rhs).max()
        assert d[inner].max() < 1e-13 * scale
        assert d[~inner].max() > 1e-6 * scale  # the aliased cutoff modes: not a tautology
        # the NS right-hand side is already solenoidal after fourierspacep2!
        assert np.abs(g.project(b.rhs) - b.rhs).max() < 1e-13 * scale


def test_equation_key():
    base = dict(nx="16", ny="16", nz="16", xDomainSize="1.0", yDomainSize="1.0", zDomainSize="1.0",
                kinematicViscosity="0.01", dtStat="1", dtSpec="1", writeTime="1")
    assert config.parameters(dict(base, equation="NS")).to_c().equation == 1
    assert config.parameters(dict(base, equation="rotational")).to_c().equation == 0
    assert config.parameters(base).to_c().equation == 0
