"""The `global` file: parsing, expression values, defaults and required keys
(simulationtypes.jl:242-365, Globals.jl:106-182; format of
test/Scalar_Advection/global)."""
import math

import numpy as np
import pytest

from ffs import config, fields

# the key/value content of the reference's example case (test/Scalar_Advection/global)
EXAMPLE = """// Set physical parameters
kinematicViscosity 0.03333333

// Set domain parameters
xDomainSize 1.0
yDomainSize 1.0
zDomainSize 1.0
nx 128
ny 128
nz 128

passiveScalar true
scalarGradient 0.0
scalarPr Inf

// Set simulation control variables
dtStat 125
nt 500
dt 0.002
variableTimestepFlag false
velocityTimeStep Adams_Bashforth3rdO
dealias sphere
threaded true

// Times to write output files
writeTime 125
"""


def test_readglobal_example(tmp_path):
    f = tmp_path / "global"
    f.write_text(EXAMPLE)
    d = config.readglobal(str(f))
    assert d["nx"] == "128" and d["kinematicViscosity"] == "0.03333333" and d["scalarPr"] == "Inf"
    assert d["velocityTimeStep"] == "Adams_Bashforth3rdO" and not any(k.startswith("//") for k in d)


def test_missing_required_key_raises_keyerror(tmp_path):
    """dtSpec is required (simulationtypes.jl:262) and the example omits it."""
    f = tmp_path / "global"
    f.write_text(EXAMPLE)
    with pytest.raises(KeyError, match="dtSpec"):
        config.parameters(config.readglobal(str(f)))


def test_parameters_defaults_and_values(tmp_path):
    f = tmp_path / "global"
    f.write_text(EXAMPLE + "dtSpec 1000000\n")
    p = config.parameters(config.readglobal(str(f)))
    assert (p.nx, p.ny, p.nz) == (128, 128, 128)
    assert p.nu == 0.03333333 and p.dt == 0.002 and p.dt_stat == 125 and p.write_time == 125 and p.nt == 500
    assert p.integrator == "Adams_Bashforth3rdO" and p.variable_dt is False
    assert p.cfl == 0.18 and abs(p.cutoff - 2 / 3) < 1e-16  # defaults
    assert "dealias" in p.extra  # read nowhere (the box mask follows `cutoff`)
    assert p.passive_scalar and p.scalar_pr == math.inf and p.scalar_gradient == 0.0


@pytest.mark.parametrize("s,v", [("2π", 2 * math.pi), ("1/3", 1 / 3), ("2^-4", 2**-4), ("0.5e-3", 5e-4),
                                 ("(1+2)*3", 9.0), ("-2", -2.0), ("Inf", math.inf), ("2pi/3", 2 * math.pi / 3),
                                 ("3(1+1)", 6.0), ("1.0", 1.0)])
def test_parse_expr(s, v):
    assert config.parse_expr(s) == pytest.approx(v, rel=1e-15)


def test_parse_expr_rejects_garbage():
    for bad in ("", "2+", "(1", "abc"):
        with pytest.raises(ValueError):
            config.parse_expr(bad)


def test_field_files_padded_and_unpadded(tmp_path):
    nx, ny, nz = 6, 4, 5
    a = np.random.default_rng(0).standard_normal((nx, ny, nz))
    p = fields.pad(a)
    assert p.shape == (8, 4, 5) and np.all(p[6:] == 0)
    fields.write_field(str(tmp_path / "u1.0"), p)
    assert (tmp_path / "u1.0").stat().st_size == 8 * 8 * 4 * 5  # 8 * 2(nx/2+1) ny nz bytes
    back = fields.read_field(str(tmp_path / "u1.0"), nx, ny, nz)
    assert np.array_equal(back[:nx], a)
    a.ravel(order="F").astype("<f8").tofile(tmp_path / "u2.0")  # unpadded
    assert np.array_equal(fields.read_field(str(tmp_path / "u2.0"), nx, ny, nz)[:nx], a)
    (tmp_path / "u3.0").write_bytes(b"\0" * 16)
    with pytest.raises(ValueError):
        fields.read_field(str(tmp_path / "u3.0"), nx, ny, nz)


def test_padded_file_size_256():
    """Padded 256^3 field is 135,266,304 bytes; 64^3 is 2,162,688 bytes."""
    assert 8 * np.prod(fields.padded_shape(256, 256, 256)) == 135266304
    assert 8 * np.prod(fields.padded_shape(64, 64, 64)) == 2162688


def test_plane_forcing_keys():
    base = {"nx": "32", "ny": "32", "nz": "32", "xDomainSize": "1.0", "yDomainSize": "1.0", "zDomainSize": "1.0",
            "kinematicViscosity": "1e-3", "dtStat": "10", "dtSpec": "100", "writeTime": "100"}
    d = dict(base, forcing="rfForcing", TF="2.0", alphac="0.5", kf="2.5", Zf="tanh")
    p = config.parameters(d)
    assert (p.forcing, p.forcing_tf, p.forcing_alpha, p.forcing_kf, p.forcing_zf) == ("rfForcing", 2.0, 0.5, 2.5, "tanh")
    c = p.to_c()
    assert c.forcing == 2 and c.forcing_zf_tanh == 1 and c.forcing_tf == 2.0
    with pytest.raises(KeyError):  # TF is required, as in forcing_model (forcing_types.jl:339)
        config.parameters(dict(base, forcing="rfForcing", alphac="0.5", kf="2.5"))
    q = config.parameters(dict(base, forcing="QfForcing", alveliusP="0.1", alveliusC="1.0", alveliusKf="2", kf="3"))
    assert (q.alvelius_p, q.alvelius_c, q.alvelius_kf, q.forcing_kf) == (0.1, 1.0, 2.0, 3.0)
    assert q.to_c().forcing == 3
    codes = {"cRfForcing": 4, "nRfForcing": 5, "cNRfForcing": 6, "aForcing": 7}
    for name, code in codes.items():
        r = config.parameters(dict(base, forcing=name, TF="0.5", alphac="0.25", kf="3", Zf="other"))
        c = r.to_c()
        assert (c.forcing, c.forcing_tf, c.forcing_alpha, c.forcing_kf) == (code, 0.5, 0.25, 3.0)
        # calculate_Zf: an unknown Zf is tanh for the shell form, the cutoff for aForcing's matrix form
        assert c.forcing_zf_tanh == (0 if name == "aForcing" else 1)
        assert config.parameters(dict(base, forcing=name, TF="1", alphac="1", kf="3")).to_c().forcing_zf_tanh == 0


def test_les_model_keys():
    """les_types (les_types.jl:360-404): each model reads its own constant key; unset constants stay
    None (the library applies the model's default)."""
    base = {"nx": "32", "ny": "32", "nz": "32", "xDomainSize": "1.0", "yDomainSize": "1.0", "zDomainSize": "1.0",
            "kinematicViscosity": "1e-3", "dtStat": "10", "dtSpec": "100", "writeTime": "100"}
    from ffs import _lib
    v = config.parameters(dict(base, lesModel="Vreman", vremanConstant="0.05", smagorinskyConstant="0.3"))
    assert (v.les_model, v.smagorinsky_c) == ("Vreman", 0.05) and v.to_c().les_model == _lib.FFS_LES_VREMAN
    s = config.parameters(dict(base, lesModel="Silvis", silvisPConstant="4.0"))
    assert s.smagorinsky_c is None and s.les_cp == 4.0
    c = s.to_c()
    assert math.isnan(c.smagorinsky_c) and c.les_cp == 4.0 and c.les_model == _lib.FFS_LES_SILVIS
    q = config.parameters(dict(base, lesModel="SandP", sLesModel="productionViscosity", productionConstant="0.2",
                               pConstant="0.05", filterWidth="0.1"))
    assert (q.les_s_model, q.smagorinsky_c, q.les_cp, q.filter_width) == ("productionViscosity", 0.2, 0.05, 0.1)
    assert q.to_c().les_s_model == _lib.FFS_LES_PRODUCTION and q.to_c().les_model == _lib.FFS_LES_SANDP
    n = config.parameters(dict(base, lesModel="stableNonLinear", nlConstant="0.2"))
    assert n.smagorinsky_c == 0.2
    with pytest.raises(KeyError):
        config.parameters(dict(base, lesModel="SandP"))  # sLesModel required (les_types.jl:391)
    d = config.parameters(dict(base, lesModel="dynamicSmagorinsky", TestFilterWidth="0.3", cmin="-0.2"))
    assert (d.les_test_width, d.les_cmin) == (0.3, -0.2) and d.to_c().les_model == _lib.FFS_LES_DYNAMIC_SMAGORINSKY
    pm = config.parameters(dict(base, lesModel="piomelliSmagorinsky", cmin="0.01"))
    assert pm.les_cmin == 0.01 and pm.to_c().les_model == _lib.FFS_LES_PIOMELLI_SMAGORINSKY
    ds = config.parameters(dict(base, lesModel="piomelliSandP", sLesModel="dynamicSmagorinsky", cmin="0.02"))
    assert (ds.les_s_model, ds.les_cmin) == ("dynamicSmagorinsky", 0.02)
    assert ds.to_c().les_model == _lib.FFS_LES_PIOMELLI_SANDP and ds.to_c().les_s_model == _lib.FFS_LES_DYNAMIC_SMAGORINSKY
    assert config.parameters(dict(base, lesModel="fancyModel")).to_c().les_model == -1  # refused at create


def test_density_hyperviscosity_keys():
    """densityHyperViscosity (density_stratification_types.jl:137-149): its own coefficient and M, or
    the spectral barrier with densityInitk / densityEndk."""
    from ffs import _lib
    base = {"nx": "32", "ny": "32", "nz": "32", "xDomainSize": "1.0", "yDomainSize": "1.0", "zDomainSize": "1.0",
            "kinematicViscosity": "1e-3", "dtStat": "10", "dtSpec": "100", "writeTime": "100"}
    h = config.parameters(dict(base, densityHyperViscosity="hyperViscosity", densityHyperViscosityCoefficient="1e-8",
                               densityHyperViscosityM="3"))
    c = h.to_c()
    assert (c.density_hv, c.density_hv_nu, c.density_hv_m, c.hyperviscosity) == (_lib.FFS_HV_HYPER, 1e-8, 3,
                                                                                 _lib.FFS_HV_NONE)
    b = config.parameters(dict(base, densityHyperViscosity="spectralBarrier", densityInitk="5", densityEndk="9"))
    assert (b.to_c().density_hv, b.density_sb_initk, b.density_sb_endk) == (_lib.FFS_HV_BARRIER, 5.0, 9.0)
    with pytest.raises(KeyError):
        config.parameters(dict(base, densityHyperViscosity="hyperViscosity"))


def test_velocity_time_step_values():
    """velocityTimeStep (simulationtypes.jl:308-318): Euller, Adams_Bashforth3rdO, anything else ETD3rdO."""
    from ffs import _lib
    base = {"nx": "16", "ny": "16", "nz": "16", "xDomainSize": "1.0", "yDomainSize": "1.0", "zDomainSize": "1.0",
            "kinematicViscosity": "1e-3", "dtStat": "10", "dtSpec": "100", "writeTime": "100"}
    for v, code in (("Euller", _lib.FFS_EULER), ("Adams_Bashforth3rdO", _lib.FFS_AB3), ("ETD3rdO", _lib.FFS_ETD3),
                    ("RK4", _lib.FFS_ETD3)):
        assert config.parameters(dict(base, velocityTimeStep=v)).to_c().integrator == code
    assert config.parameters(base).to_c().integrator == _lib.FFS_ETD3


def test_gravity_direction_key():
    from ffs import _lib
    base = {"nx": "16", "ny": "16", "nz": "16", "xDomainSize": "1.0", "yDomainSize": "1.0", "zDomainSize": "1.0",
            "kinematicViscosity": "1e-3", "dtStat": "10", "dtSpec": "100", "writeTime": "100",
            "densityStratification": "true", "Pr": "1", "densityGradient": "1", "zAcceleration": "-1",
            "referenceDensity": "1"}
    assert config.parameters(base).to_c().gravity_direction == 2
    assert config.parameters(dict(base, gravityDirection="y")).to_c().gravity_direction == 1
    with pytest.raises(ValueError):
        config.parameters(dict(base, gravityDirection="w"))
    assert _lib.default_params().gravity_direction == 2


@pytest.mark.parametrize("key,value", [("specFil", "2")])
def test_unbuilt_keys_are_refused(key, value):
    base = {"nx": "16", "ny": "16", "nz": "16", "xDomainSize": "1.0", "yDomainSize": "1.0", "zDomainSize": "1.0",
            "kinematicViscosity": "1e-3", "dtStat": "10", "dtSpec": "100", "writeTime": "100"}
    with pytest.raises(NotImplementedError, match=key):
        config.parameters(dict(base, **{key: value}))


def test_scalar_les_keys():
    from ffs import _lib
    base = {"nx": "16", "ny": "16", "nz": "16", "xDomainSize": "1.0", "yDomainSize": "1.0", "zDomainSize": "1.0",
            "kinematicViscosity": "1e-3", "dtStat": "10", "dtSpec": "100", "writeTime": "100", "lesModel": "Vreman"}
    c = config.parameters(dict(base, scalarLesModel="vorticityDiffusion", scalarVorticityCoefficient="1.3",
                               densityLesModel="eddyDiffusion")).to_c()
    assert (c.scalar_les, c.scalar_les_c, c.density_les) == (2, 1.3, 1) and math.isnan(c.density_les_c)
    c = config.parameters(dict(base, densityLesModel="vorticityDiffusion")).to_c()
    assert (c.scalar_les, c.density_les) == (0, 2) and math.isnan(c.density_les_c)
    assert _lib.default_params().scalar_les == 0 and math.isnan(_lib.default_params().scalar_les_c)
