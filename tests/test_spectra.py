"""The oracle's spectra (write_spec / write_spec_forcing, spec.jl:1-168),
pinned by what the definitions fix: a single mode lands in the bins of its
|kx|, |ky| and |kz| with calculate_specHV's weights; on a unit box every
spectrum sums (over specV bins) to the Parseval-weighted total that the stats
row holds (energy, dissipation, nonlinear and pressure transfer, LES transfer,
density variance, its dissipation, buoyancy flux); the nonlinear and density
transfers and the pressure work vanish in total (divergence-free u, 2/3
dealiasing), the pressure work per mode; QfForcing's forcing spectrum sums to
its injected power; the Julia float format of the written files."""
import math

import numpy as np
import pytest

from oracle.ffs_oracle import Grid, OracleSim
from oracle.plane_forcing import PlaneForcing
from ffs.initial import random_scalar, random_spectrum
from ffs.simulation import jl_float, spectra_files
import ffs


def test_single_mode_bins():
    # a 2 x 3 x 1.5 box: MAXDKH = dkx = 1/2 (finer than dky = 1/3? no: 1/2 > 1/3, KH = KX), DKZ = 2/3
    g = Grid(16, 16, 16, 2.0, 3.0, 1.5)
    s = OracleSim(g, 0.01, dt=0.01)
    assert g.MAXDKH == g.KX[1] and g.DKZ == g.KRZ[1]
    for (i, j, l) in [(3, 2, 1), (0, 4, 3), (2, 14, 13)]:
        a = np.zeros((g.NX, g.NY, g.NZ))
        a[i, j, l] = 1.0
        h, v = s.spec_hv(a)
        want_h = np.zeros_like(h)
        want_v = np.zeros_like(v)
        nx = int(round(abs(g.KX[i]) / g.MAXDKH))
        ny = int(round(abs(g.KY[j]) / g.MAXDKH))
        nz = int(round(abs(g.KZ[l]) / g.DKZ))
        if i == 0:   # h[1] and h[n(ky)] take half each, v takes it whole
            want_h[0] += 0.5
            want_h[ny] += 0.5
            want_v[nz] += 1.0
        else:        # the mode and its conjugate: weight 2, times the bin widths
            want_h[nx] += g.MAXDKH
            want_h[ny] += g.MAXDKH
            want_v[nz] += 2 * g.DKZ
        assert np.array_equal(h, want_h), (i, j, l)
        assert np.array_equal(v, want_v), (i, j, l)


@pytest.fixture(scope="module")
def les_density_run():
    g = Grid(16, 16, 16)
    s = OracleSim(g, 0.01, dt=0.0, variable_dt=True, cfl=0.3, dt_stat=1, dt_spec=1, les_c=0.16,
                  density={"Pr": 0.7, "gradient": 1.0, "g": -1.0})
    s.set_velocity(random_spectrum(16, seed=3))
    s.set_density(random_scalar(16, 16, 16, seed=4))
    s.initialize()
    s.advance_in_time(2)
    s.calculate_rhs()
    row = s.stats()[0]
    return s, s.spectra[-1][1], row


def total(spec):
    return float(spec[1].sum())  # specV: weight 1 at kx = 0, 2 DKZ = 2 elsewhere on a unit box


def test_spectra_sum_to_the_stats(les_density_run):
    s, q, row = les_density_run
    assert s.g.DKZ == 1.0 and s.g.MAXDKH == 1.0
    pairs = {"u_h": row[8], "u_v": row[9], "vis_h": row[12], "vis_v": row[13], "nl_h": row[15], "nl_v": row[16],
             "press_h": row[17], "press_v": row[18], "rho": row[23], "rhodiss": row[27], "buoyancy_v": -row[29],
             "les_h": -row[30], "les_v": -row[31]}
    diss = total(q["vis_h"]) + total(q["vis_v"])
    for name, want in pairs.items():
        got = total(q[name])
        tol = 1e-12 * max(abs(want), diss if name[:2] in ("nl", "pr") else 0.0)
        assert abs(got - want) <= tol, (name, got, want)
        # specH counts every mode in its kx bin and its ky bin: the same total on a unit box
        assert abs(float(q[name][0].sum()) - want) <= tol, name


def test_transfers_vanish(les_density_run):
    s, q, _ = les_density_run
    scale = total(q["vis_h"]) + total(q["vis_v"])
    # u_c (u x w - grad k)_c = -u . grad(u_c^2 / 2) integrates to zero for each component
    for c in ("nl_h", "nl_v"):
        assert abs(total(q[c])) < 1e-12 * scale
        assert np.abs(q[c][1]).max() > 1e-6 * scale   # while the shells exchange energy
    assert abs(total(q["press_h"]) + total(q["press_v"])) < 1e-12 * scale
    assert abs(total(q["press_h"])) > 1e-6 * scale   # the pressure moves energy between h and v
    # k . u = 0: proj(u, p1 k) = 0 for each mode, so h + v cancels in every bin
    for b in (0, 1):
        assert np.abs(q["press_h"][b] + q["press_v"][b]).max() < 1e-12 * scale
    assert abs(total(q["rhonl"])) < 1e-12 * total(q["rhodiss"])


def test_qf_forcing_spectrum_is_its_power():
    g = Grid(16, 16, 16)
    pf = PlaneForcing(g, "qf", kf=3.0, P=0.1, C=1.0, Kf_alv=2.0)
    s = OracleSim(g, 2e-3, dt=0.0, variable_dt=True, cfl=0.4, dt_stat=2, dt_spec=2, plane_forcing=pf)
    s.set_velocity(random_spectrum(16, seed=5))
    s.initialize()
    s.advance_in_time(3)
    assert [r[0] for r in s.spectra] == [0, 2]
    for (it, q), rec in zip(s.spectra, s.records):
        hp, vp = rec[0][-3], rec[0][-2]
        assert abs(total(q["forcing_h"]) - (hp + vp)) <= 1e-12 * (hp + vp)


def test_spectra_files_and_format():
    p = ffs.Params(nx=16, ny=16, nz=16, density=True, z_acceleration=-2.0, reference_density=1.0,
                   density_gradient=3.0)
    h, v = np.array([1.0, 2.0]), np.array([3.0, 4.0])
    q = {"u_h": (h, v), "u_v": (2 * h, 2 * v), "buoyancy_v": (h, v), "rho": (h, v)}
    f = spectra_files(q, p)
    assert set(f) == {"u_h.specH", "u_h.specV", "u.specH", "u_v.specH", "u_v.specV", "buoyancy_v.specH",
                      "buoyancy_v.specV", "rhosource.specH", "rhosource.specV", "rho.specH", "rho.specV"}
    assert np.array_equal(f["u.specH"][0], 3 * h)
    assert np.array_equal(f["rhosource.specV"][0], -(v / -2.0) * 3.0)
    assert [jl_float(x) for x in (1e-5, 0.0001, 100000.0, 1e6, 2.0**20, -2.5e-7, 0.0, 1 / 3)] == \
        ["1.0e-5", "0.0001", "100000.0", "1.0e6", "1.048576e6", "-2.5e-7", "0.0", "0.3333333333333333"]
    for x in (1e-5, 3.14159e-300, 6.02e23, 123.456, -7.0):
        assert float(jl_float(x)) == x
