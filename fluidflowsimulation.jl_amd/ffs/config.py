"""The `global` parameter file: reader and conversion to run parameters.

Mirrors ReadGlobal.readglobal() (un-vendored; format as in
test/Scalar_Advection/global:1-26) and parameters(d) of
simulationtypes.jl:242-365 together with the module-load constants of
Globals.jl:106-182.  Required keys raise KeyError exactly where the reference
does (`d[:nx]` ...); optional keys take the reference defaults.
"""
import math
import re
from dataclasses import dataclass, field, fields

__all__ = ["readglobal", "parse_expr", "Params", "parameters"]


def readglobal(path="global"):
    """Reads `key value` lines; `//` starts a comment; returns {key: value-string}."""
    d = {}
    with open(path) as f:
        for line in f:
            line = line.split("//", 1)[0].strip()
            if not line:
                continue
            parts = line.split(None, 1)
            d[parts[0]] = parts[1].strip() if len(parts) > 1 else ""
    return d


_TOK = re.compile(r"\s*(?:(\d+\.\d*(?:[eE][-+]?\d+)?|\.\d+(?:[eE][-+]?\d+)?|\d+(?:[eE][-+]?\d+)?)|(π|pi|Inf|inf)|(.))")


def parse_expr(s):
    """Evaluates the numeric expressions the reference accepts through
    `Float64(eval(Meta.parse(v)))` (e.g. `2π`, `1/3`, `2^-4`, `0.03333333`):
    numbers, + - * / ^, parentheses, π/pi, Inf and implicit multiplication of a
    number by π or a parenthesis."""
    toks = []
    for m in _TOK.finditer(s.strip()):
        num, const, op = m.groups()
        if num is not None:
            toks.append(("n", float(num)))
        elif const is not None:
            toks.append(("n", math.pi if const in ("π", "pi") else math.inf))
        elif op is not None and not op.isspace():
            toks.append(("o", op))
    pos = 0

    def peek():
        return toks[pos] if pos < len(toks) else None

    def take():
        nonlocal pos
        if pos >= len(toks):
            raise ValueError(f"incomplete expression: {s!r}")
        t = toks[pos]
        pos += 1
        return t

    def expr():
        v = term()
        while peek() and peek() in (("o", "+"), ("o", "-")):
            op = take()[1]
            r = term()
            v = v + r if op == "+" else v - r
        return v

    def term():
        v = unary()
        while True:
            t = peek()
            if t in (("o", "*"), ("o", "/")):
                op = take()[1]
                r = unary()
                v = v * r if op == "*" else v / r
            elif t is not None and (t[0] == "n" or t == ("o", "(")):
                v = v * unary()  # juxtaposition: 2π, 2(1+x)
            else:
                return v

    def unary():
        t = peek()
        if t == ("o", "-"):
            take()
            return -unary()
        if t == ("o", "+"):
            take()
            return unary()
        return power()

    def power():
        b = atom()
        if peek() == ("o", "^"):
            take()
            return b ** unary()
        return b

    def atom():
        t = take() if peek() else None
        if t is None:
            raise ValueError(f"incomplete expression: {s!r}")
        if t[0] == "n":
            return t[1]
        if t == ("o", "("):
            v = expr()
            if take() != ("o", ")"):
                raise ValueError(f"missing ')' in {s!r}")
            return v
        raise ValueError(f"unexpected {t[1]!r} in {s!r}")

    v = expr()
    if pos != len(toks):
        raise ValueError(f"trailing input in {s!r}")
    return float(v)


def _parse_bool(s):
    s = s.strip().lower()
    if s in ("true", "1"):
        return True
    if s in ("false", "0"):
        return False
    raise ValueError(f"not a Bool: {s!r}")


@dataclass
class Params:
    """Run parameters (the `ffs_params` struct of include/ffs/ffs.h)."""

    nx: int
    ny: int
    nz: int
    x_domain: float = 1.0
    y_domain: float = 1.0
    z_domain: float = 1.0
    nu: float = 0.0
    cutoff: float = 2.0 / 3.0
    integrator: str = "ETD3rdO"
    variable_dt: bool = True
    cfl: float = 0.18
    dt: float = 0.0
    start: int = 0
    start_time: float = 0.0
    dt_stat: int = 1 << 30
    dt_spec: int = 1 << 30
    write_time: int = 1 << 30
    nt: int = None
    time_duration: float = math.inf
    threaded: bool = True
    precision: str = "fp64"
    les_model: str = None           # lesModel: Smagorinsky, Vreman, Silvis, productionViscosity,
                                    # stableNonLinear or SandP (les_types.jl:364-404)
    smagorinsky_c: float = None     # the model's constant (None: its default; SandP: the S model's)
    les_cp: float = None            # silvisPConstant / SandP's pConstant (None: 5.0 / 0.09)
    les_s_model: str = None         # sLesModel of SandP
    les_test_width: float = 0.0     # TestFilterWidth of the dynamic models (0: twice the filter width)
    les_cmin: float = 0.0           # cmin of the dynamic models
    filter_width: float = 0.0
    density: bool = False
    pr: float = 1.0
    density_gradient: float = 0.0
    z_acceleration: float = 0.0
    reference_density: float = 1.0
    forcing: str = None
    forcing_kf: float = 2.5
    forcing_power: float = 0.1
    forcing_tf: float = 0.0         # TF (rfForcing)
    forcing_alpha: float = 0.0      # alphac (rfForcing)
    forcing_zf: str = "cutoff"      # Zf: cutoff | tanh
    alvelius_p: float = 0.0         # alveliusP, alveliusC, alveliusKf (QfForcing)
    alvelius_c: float = 1.0
    alvelius_kf: float = 0.0
    hyperviscosity: str = None      # hyperViscosity: "spectralBarrier", or any other value: HyperViscosity
    scalar_les: str = None          # scalarLesModel: "vorticityDiffusion", else EddyDiffusion (with lesModel)
    scalar_les_c: float = None      # scalarVorticityCoefficient (None: 1.0)
    density_les: str = None         # densityLesModel: "vorticityDiffusion" or any other value: EddyDiffusion
    density_les_c: float = None     # densityVorticityCoefficient (None: 0.1)
    gravity_direction: str = "z"    # gravityDirection: the axis of g and of the density's mean gradient
    density_hyperviscosity: str = None  # densityHyperViscosity: the density's own, as hyperviscosity
    density_hv_nu: float = 0.0      # densityHyperViscosityCoefficient
    density_hv_m: int = 2           # densityHyperViscosityM
    density_sb_initk: float = 0.0   # densityInitk / densityEndk (0: the defaults)
    density_sb_endk: float = 0.0
    hv_nu: float = 0.0              # hyperViscosityCoefficient
    hv_m: int = 2                   # hyperViscosityM
    sb_initk: float = 0.0           # initk / endk of the spectral barrier (0: the reference's defaults)
    sb_endk: float = 0.0
    passive_scalar: bool = False    # passiveScalar (presence)
    scalar_pr: float = 1.0          # scalarPr (Inf: no diffusion)
    scalar_gradient: float = 0.0    # scalarGradient
    scalar_direction: str = "z"     # scalarDirection
    equation: str = None            # equation: "NS" is the conservation form, anything else the rotational
    extra: dict = field(default_factory=dict)

    def to_c(self):
        from . import _lib

        p = _lib.default_params()
        p.nx, p.ny, p.nz = self.nx, self.ny, self.nz
        p.x_domain, p.y_domain, p.z_domain = self.x_domain, self.y_domain, self.z_domain
        p.nu = self.nu
        p.cutoff = self.cutoff
        # velocityTimeStep (simulationtypes.jl:308-318): any value but these two selects ETD3rdO
        p.integrator = {"Adams_Bashforth3rdO": _lib.FFS_AB3, "Euller": _lib.FFS_EULER}.get(self.integrator,
                                                                                            _lib.FFS_ETD3)
        p.variable_dt = int(bool(self.variable_dt))
        p.cfl = self.cfl
        p.dt = self.dt
        p.start = self.start
        p.start_time = self.start_time
        p.dt_stat, p.dt_spec, p.write_time = self.dt_stat, self.dt_spec, self.write_time
        p.precision = _lib.FFS_FP32 if self.precision == "fp32" else _lib.FFS_FP64
        p.les_model = _lib.FFS_LES_NONE if self.les_model is None else _lib.LES_MODELS.get(self.les_model, -1)
        p.smagorinsky_c = math.nan if self.smagorinsky_c is None else self.smagorinsky_c
        p.les_cp = math.nan if self.les_cp is None else self.les_cp
        if self.les_s_model is not None:
            p.les_s_model = _lib.LES_MODELS.get(self.les_s_model, -1)
        p.les_test_width, p.les_cmin = self.les_test_width, self.les_cmin
        p.filter_width = self.filter_width
        p.density = int(bool(self.density))
        p.pr, p.density_gradient = self.pr, self.density_gradient
        p.z_acceleration, p.reference_density = self.z_acceleration, self.reference_density
        p.forcing = {None: _lib.FFS_FORCING_NONE, "shellForcing": _lib.FFS_FORCING_SHELL,
                     "rfForcing": _lib.FFS_FORCING_RF, "QfForcing": _lib.FFS_FORCING_QF,
                     "cRfForcing": _lib.FFS_FORCING_CRF, "nRfForcing": _lib.FFS_FORCING_NRF,
                     "cNRfForcing": _lib.FFS_FORCING_CNRF, "aForcing": _lib.FFS_FORCING_A}.get(self.forcing, -1)
        p.forcing_kf, p.forcing_power = self.forcing_kf, self.forcing_power
        p.forcing_tf, p.forcing_alpha = self.forcing_tf, self.forcing_alpha
        # calculate_Zf (utils.jl:184-218): the shell form takes the cutoff only for "cutoff", aForcing's
        # (NX, NY) form takes tanh only for "tanh"
        if self.forcing == "aForcing":
            p.forcing_zf_tanh = int(self.forcing_zf == "tanh")
        else:
            p.forcing_zf_tanh = int(self.forcing_zf != "cutoff")
        p.alvelius_p, p.alvelius_c, p.alvelius_kf = self.alvelius_p, self.alvelius_c, self.alvelius_kf
        def hv_code(v):
            return _lib.FFS_HV_NONE if v is None else (_lib.FFS_HV_BARRIER if v == "spectralBarrier"
                                                       else _lib.FFS_HV_HYPER)
        p.hyperviscosity = hv_code(self.hyperviscosity)
        p.hv_nu, p.hv_m = self.hv_nu, self.hv_m
        p.sb_initk, p.sb_endk = self.sb_initk, self.sb_endk
        p.density_hv = hv_code(self.density_hyperviscosity)
        p.gravity_direction = {"x": 0, "y": 1, "z": 2}[self.gravity_direction]
        p.scalar_les = 2 if self.scalar_les == "vorticityDiffusion" else (1 if self.scalar_les is not None else 0)
        p.density_les = 0 if self.density_les is None else (2 if self.density_les == "vorticityDiffusion" else 1)
        p.scalar_les_c = math.nan if self.scalar_les_c is None else self.scalar_les_c
        p.density_les_c = math.nan if self.density_les_c is None else self.density_les_c
        p.density_hv_nu, p.density_hv_m = self.density_hv_nu, self.density_hv_m
        p.density_sb_initk, p.density_sb_endk = self.density_sb_initk, self.density_sb_endk
        p.passive_scalar = int(bool(self.passive_scalar))
        p.scalar_pr, p.scalar_gradient = self.scalar_pr, self.scalar_gradient
        p.scalar_direction = {"x": 0, "y": 1, "z": 2}[self.scalar_direction]
        p.equation = _lib.FFS_EQ_NS if self.equation == "NS" else _lib.FFS_EQ_ROTATIONAL
        return p


# the oscillator schemes of forcing_model (TF, alphac, kf, Zf)
RF_FAMILY = ("rfForcing", "cRfForcing", "nRfForcing", "cNRfForcing", "aForcing")
PLANE_FORCINGS = RF_FAMILY + ("QfForcing",)


def parameters(d):
    """parameters(d) of simulationtypes.jl:242-365 (plus nt/timeDuration of
    FluidFlowSimulation.jl:26-27).  Raises KeyError for a missing required key."""
    known = set()

    def req(k):
        known.add(k)
        return d[k]

    def opt(k, conv, default):
        known.add(k)
        return conv(d[k]) if k in d else default

    p = Params(nx=int(req("nx")), ny=int(req("ny")), nz=int(req("nz")))
    p.x_domain = parse_expr(req("xDomainSize"))
    p.y_domain = parse_expr(req("yDomainSize"))
    p.z_domain = parse_expr(req("zDomainSize"))
    p.nu = parse_expr(req("kinematicViscosity"))
    p.start = opt("start", int, 0)
    p.start_time = opt("startTime", float, 0.0)
    p.dt_stat = int(req("dtStat"))
    p.dt_spec = int(req("dtSpec"))
    p.write_time = int(req("writeTime"))
    p.threaded = opt("threaded", _parse_bool, True)
    p.cutoff = opt("cutoff", parse_expr, 2.0 / 3.0)
    p.integrator = opt("velocityTimeStep", str, "ETD3rdO")
    p.variable_dt = opt("variableTimestepFlag", _parse_bool, True)
    p.cfl = opt("cfl", float, 0.18)
    p.dt = opt("dt", parse_expr, 0.0)
    p.nt = opt("nt", int, None)
    p.time_duration = opt("timeDuration", float, math.inf)
    p.precision = opt("precision", str, "fp64")
    if "lesModel" in d:  # les_types (les_types.jl:360-404): each model reads its own constant keys
        p.les_model = d["lesModel"]
        sandp = p.les_model in ("SandP", "dynamicSandP", "piomelliSandP")
        smodel = d.get("sLesModel") if sandp else None
        known.add("sLesModel")
        if sandp and smodel is None:
            raise KeyError("sLesModel")
        key = {"Smagorinsky": "smagorinskyConstant", "Vreman": "vremanConstant", "Silvis": "silvisConstant",
               "productionViscosity": "productionConstant", "stableNonLinear": "nlConstant"}
        ck = key.get(smodel if p.les_model == "SandP" else p.les_model)
        if ck is not None:
            p.smagorinsky_c = opt(ck, parse_expr, None)
        if p.les_model == "Silvis":
            p.les_cp = opt("silvisPConstant", parse_expr, None)
        elif p.les_model == "SandP":
            p.les_cp = opt("pConstant", parse_expr, None)
            p.les_s_model = smodel
        elif p.les_model in ("dynamicSmagorinsky", "piomelliSmagorinsky", "dynamicSandP", "piomelliSandP"):
            if sandp:
                p.les_s_model = smodel
            p.les_test_width = opt("TestFilterWidth", parse_expr, 0.0)
            p.les_cmin = opt("cmin", float, 0.0)
        p.filter_width = opt("filterWidth", parse_expr, 0.0)
    known.add("lesModel")
    if "densityStratification" in d:
        known.add("densityStratification")
        p.density = True
        p.pr = parse_expr(req("Pr"))
        p.density_gradient = parse_expr(req("densityGradient"))
        p.z_acceleration = parse_expr(req("zAcceleration"))
        p.reference_density = parse_expr(req("referenceDensity"))
        p.gravity_direction = opt("gravityDirection", str, "z")  # density_stratification_types.jl:152
        if p.gravity_direction not in ("x", "y", "z"):
            raise ValueError(f"gravityDirection {p.gravity_direction!r}: x, y or z")
    p.equation = opt("equation", str, None)  # simulationtypes.jl:265-273
    # keys whose features this build does not have: refused rather than ignored
    if "specFil" in d:
        raise NotImplementedError("specFil: the filtered statistics and spectra (Stats_fil.txt, *_fil spectra) "
                                  "are not built")
    # the scalars' LES fluxes (simulationtypes.jl:334-346, density_stratification_types.jl:176-184)
    if "scalarLesModel" in d:
        p.scalar_les = d["scalarLesModel"]
        if p.scalar_les == "vorticityDiffusion":
            p.scalar_les_c = opt("scalarVorticityCoefficient", float, None)
    if "densityLesModel" in d:
        p.density_les = d["densityLesModel"]
        if p.density_les == "vorticityDiffusion":
            p.density_les_c = opt("densityVorticityCoefficient", float, None)
    known.update(("scalarLesModel", "densityLesModel"))
    if "passiveScalar" in d:  # simulationtypes.jl:320-352
        known.add("passiveScalar")
        p.passive_scalar = True
        p.scalar_pr = parse_expr(req("scalarPr"))
        p.scalar_gradient = parse_expr(req("scalarGradient"))
        p.scalar_direction = opt("scalarDirection", str, "z")
        if p.scalar_direction not in ("x", "y", "z"):
            raise ValueError(f"scalarDirection {p.scalar_direction!r}: x, y or z")
    if "densityHyperViscosity" in d:  # density_stratification_types.jl:137-149
        known.add("densityHyperViscosity")
        p.density_hyperviscosity = d["densityHyperViscosity"]
        if p.density_hyperviscosity == "spectralBarrier":
            p.density_sb_initk = opt("densityInitk", float, 0.0)
            p.density_sb_endk = opt("densityEndk", float, 0.0)
        else:
            p.density_hv_nu = float(req("densityHyperViscosityCoefficient"))
            p.density_hv_m = opt("densityHyperViscosityM", int, 2)
    if "hyperViscosity" in d:  # simulationtypes.jl:293-305
        known.add("hyperViscosity")
        p.hyperviscosity = d["hyperViscosity"]
        if p.hyperviscosity == "spectralBarrier":
            p.sb_initk = opt("initk", float, 0.0)
            p.sb_endk = opt("endk", float, 0.0)
        else:
            p.hv_nu = float(req("hyperViscosityCoefficient"))
            p.hv_m = opt("hyperViscosityM", int, 2)
    if "forcing" in d:
        known.add("forcing")
        p.forcing = d["forcing"]
        if p.forcing in RF_FAMILY:  # forcing_model, forcing_types.jl:338-422
            p.forcing_tf = float(req("TF"))
            p.forcing_alpha = float(req("alphac"))
            p.forcing_kf = float(req("kf"))
            p.forcing_zf = opt("Zf", str, "cutoff")
        elif p.forcing == "QfForcing":  # forcing_types.jl:424-446
            p.alvelius_p = float(req("alveliusP"))
            p.alvelius_c = float(req("alveliusC"))
            p.alvelius_kf = float(req("alveliusKf"))
            p.forcing_kf = float(req("kf"))
        else:
            p.forcing_kf = opt("forcingKf", parse_expr, 2.5)
            p.forcing_power = opt("forcingPower", parse_expr, 0.1)
    p.extra = {k: v for k, v in d.items() if k not in known}
    return p
