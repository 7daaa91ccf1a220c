"""The host twin of the device's seeded initial fields (oracle/seeded.py),
pinned by the published SplitMix64 sequence, by what the definition fixes
(solenoidal, real, dealiased, Parseval energy / rms) and by the statistics of
the draws."""
import numpy as np

from oracle.ffs_oracle import Grid
from oracle.seeded import gauss, mix64, seeded_scalar_hat, seeded_velocity_hat


def test_splitmix64_published_sequence():
    # SplitMix64 from state 0: the state advances by the golden gamma, the output is mix(state)
    expected = [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]
    state = 0
    for e in expected:
        assert int(mix64(np.uint64(state))) == e
        state = (state + 0x9E3779B97F4A7C15) % 2**64


def test_gaussian_draws_have_unit_complex_variance():
    i = np.arange(200000)
    re, im = gauss(12345, 0, i % 97, (i // 97) % 101, i // (97 * 101))
    for x in (re, im):
        assert abs(x.mean()) < 0.01
        assert abs(x.var() - 1) < 0.02
        assert abs(np.mean(x**4) - 3) < 0.1  # Gaussian kurtosis
    assert abs(np.mean(re * im)) < 0.01


def physical(g, uh):
    return g.irfft(uh)


def test_velocity_is_real_solenoidal_dealiased_and_normalised():
    g = Grid(32, 24, 40, 1.0, 1.5, 0.75)
    u = seeded_velocity_hat(g, 42, k0=3.0, energy=0.37)
    div = g.kx * u[0] + g.ky * u[1] + g.kz * u[2]
    assert np.abs(div).max() < 1e-15 * max(np.abs(x).max() for x in u)
    for x in u:
        assert np.all(x[g.DEALIAS] == 0)
        assert x[0, 0, 0] == 0
        # a real field: the Hermitian round trip reproduces the spectrum exactly
        back = g.rfft(physical(g, x)) * g.scale
        assert np.abs(back - x).max() < 1e-14 * np.abs(x).max()
    e = 0.5 * np.mean(sum(physical(g, x) ** 2 for x in u))
    assert abs(e - 0.37) < 1e-13


def test_scalar_rms_and_reality():
    g = Grid(32, 32, 32)
    r = seeded_scalar_hat(g, 9, k0=4.0, rms=0.01)
    p = physical(g, r)
    assert abs(np.sqrt(np.mean(p**2)) - 0.01) < 1e-15
    assert np.abs(g.rfft(p) * g.scale - r).max() < 1e-14 * np.abs(r).max()


def test_seed_determinism_and_independence():
    g = Grid(16, 16, 16)
    a = seeded_velocity_hat(g, 5)
    b = seeded_velocity_hat(g, 5)
    c = seeded_velocity_hat(g, 6)
    assert all(np.array_equal(x, y) for x, y in zip(a, b))
    corr = abs(np.vdot(a[0], c[0])) / (np.linalg.norm(a[0]) * np.linalg.norm(c[0]))
    assert corr < 0.2


def test_shell_spectrum_follows_the_prescribed_shape():
    # averaged over seeds, each shell's energy matches its expectation, the sum over the shell's lattice modes
    # of w A(|k|)^2 E|P xi|^2 (P xi keeps 2 of the 3 complex Gaussians, each of mean square 2), up to the
    # common normalisation and sampling error
    g = Grid(48, 48, 48)
    k2 = np.broadcast_to(g.K2, (g.NX, g.NY, g.NZ))
    k = np.sqrt(k2)
    bins = np.rint(k).astype(int)
    amp2 = np.where(g.DEALIAS | (k2 == 0), 0.0, (k * np.exp(-(k / 4.0) ** 2)) ** 2 / (4 * np.pi))
    expect = np.bincount(bins.ravel(), weights=(g.W * amp2 * 4).ravel(), minlength=40)[:12]
    acc = 0
    for seed in range(8):
        u = seeded_velocity_hat(g, seed, k0=4.0)
        e = g.W * sum(np.abs(x) ** 2 for x in u)
        acc = acc + np.bincount(bins.ravel(), weights=e.ravel(), minlength=40)[:12]
    sel = slice(2, 10)
    ratio = (acc[sel] / acc[sel].sum()) / (expect[sel] / expect[sel].sum())
    assert np.all(np.abs(ratio - 1) < 0.12), ratio
