"""Host-side mirror of the reference's driver API over the C ABI.

Names follow FluidFlowSimulation.jl: `parameters`, `initialize` (initialize!),
`advance_in_time` (advance_in_time!), `writestats`, `writeoutput`,
`run_simulation` (FluidFlowSimulation.jl:23-91).  All computation happens in
libffs_hip.so; this module only moves files and rows.
"""
import ctypes
import math
import os
from decimal import Decimal

import numpy as np

from . import _lib
from .config import PLANE_FORCINGS, RF_FAMILY, Params, parameters, readglobal
from .fields import padded_shape, read_field, write_field

__all__ = ["Simulation", "footprint", "run_simulation"]


def _dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class Simulation:
    """One run on the GPU (an `ffs_sim` handle)."""

    def __init__(self, params: Params, plan=None):
        """plan: None keeps the library's default kernel plan, "auto" takes the tuned plan of these
        parameters on this device (ffs.tune: timed once, then cached), a dict sets the fields it names."""
        self.params = params
        self._lib = _lib.load()
        cp = params.to_c()
        h = ctypes.c_void_p()
        _lib.check(self._lib.ffs_create(ctypes.byref(cp), ctypes.byref(h)))
        self._h = h
        self.nx, self.ny, self.nz = params.nx, params.ny, params.nz
        self.stats_rows = []
        if plan is not None:
            if plan == "auto":
                from .tune import tuned_plan

                plan = tuned_plan(params)
            self.set_plan(plan)

    # ---------------------------------------------------------------- lifetime
    def close(self):
        if getattr(self, "_h", None):
            self._lib.ffs_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc):
        return _lib.check(rc, self._h)

    # ---------------------------------------------------------------- fields
    @property
    def padded_size(self):
        return 2 * (self.nx // 2 + 1) * self.ny * self.nz

    def set_field(self, field, data, space=_lib.FFS_SPACE_REAL):
        """data: real (nx,ny,nz) / padded real, or complex (nx//2+1, ny, nz) half spectrum."""
        a = np.asarray(data)
        if np.iscomplexobj(a):
            space = _lib.FFS_SPACE_FOURIER
            buf = np.empty(a.size * 2, dtype=np.float64)
            flat = np.asfortranarray(a).ravel(order="F")
            buf[0::2] = flat.real
            buf[1::2] = flat.imag
        else:
            buf = np.ascontiguousarray(np.asarray(a, dtype=np.float64).ravel(order="F"))
        self._check(self._lib.ffs_set_field(self._h, field, space, _dptr(buf), buf.size))

    def set_velocity(self, u1, u2, u3):
        for f, a in enumerate((u1, u2, u3)):
            self.set_field(f, a)

    def get_field(self, field, space=_lib.FFS_SPACE_REAL):
        """Real space (FFS_SPACE_REAL, FFS_SPACE_FOURIER_TO_REAL): padded Fortran array (2*(nx//2+1), ny, nz).
        Fourier space: complex (nx//2+1, ny, nz) (the retained modes; zeros elsewhere)."""
        buf = np.empty(self.padded_size, dtype=np.float64)
        self._check(self._lib.ffs_get_field(self._h, field, space, _dptr(buf), buf.size))
        if space in (_lib.FFS_SPACE_REAL, _lib.FFS_SPACE_FOURIER_TO_REAL):
            return buf.reshape(padded_shape(self.nx, self.ny, self.nz), order="F")
        c = buf[0::2] + 1j * buf[1::2]
        return c.reshape((self.nx // 2 + 1, self.ny, self.nz), order="F")

    def get_field_noalloc(self, field, out=None):
        """The real-space read-out of a state field (u1 u2 u3 rho scalar) through
        ffs_get_field_noalloc, which allocates no device memory; `out` (float64, the padded size)
        is filled in place when given.  Returns the padded Fortran array (2*(nx//2+1), ny, nz)."""
        buf = np.empty(self.padded_size, dtype=np.float64) if out is None else out
        if buf.dtype != np.float64 or buf.size != self.padded_size or not buf.flags.c_contiguous:
            raise ValueError("out must be a contiguous float64 array of the padded size")
        self._check(self._lib.ffs_get_field_noalloc(self._h, field, _dptr(buf), buf.size))
        return buf.reshape(padded_shape(self.nx, self.ny, self.nz), order="F")

    def seed_velocity(self, seed, k0=4.0, energy=0.5):
        """Seeded random-phase solenoidal velocity generated on the device (see ffs_seed_velocity)."""
        self._check(self._lib.ffs_seed_velocity(self._h, int(seed) & (2**64 - 1), float(k0), float(energy)))

    def seed_density(self, seed, k0=4.0, rms=0.01):
        """Seeded Gaussian random density field generated on the device (see ffs_seed_density)."""
        self._check(self._lib.ffs_seed_density(self._h, int(seed) & (2**64 - 1), float(k0), float(rms)))

    def set_density(self, rho):
        """rho.$start of the Boussinesq approximation (density_stratification_types.jl:165-166)."""
        self.set_field(_lib.FFS_FIELD_RHO, rho)

    def density_hat(self):
        return self.get_field(_lib.FFS_FIELD_RHO, _lib.FFS_SPACE_FOURIER)

    def velocity_hat(self):
        return [self.get_field(f, _lib.FFS_SPACE_FOURIER) for f in range(3)]

    def velocity(self):
        return [self.get_field(f, _lib.FFS_SPACE_REAL)[: self.nx] for f in range(3)]

    # ---------------------------------------------------------------- the step
    def initialize(self):
        """initialize!(s) (FluidFlowSimulation.jl:67-91)."""
        self._check(self._lib.ffs_initialize(self._h))

    def advance(self, nsteps=1):
        """advance_in_time! nsteps times; returns (steps done, event)."""
        done = ctypes.c_int64()
        ev = ctypes.c_int()
        self._check(self._lib.ffs_advance(self._h, nsteps, ctypes.byref(done), ctypes.byref(ev)))
        return done.value, ev.value

    def advance_in_time(self, nsteps=1, on_output=None):
        """Runs nsteps steps, calling on_output(self) at output iterations."""
        left = nsteps
        while left > 0:
            done, ev = self.advance(left)
            left -= done
            if ev == _lib.FFS_EVENT_OUTPUT and on_output is not None:
                on_output(self)
        return self

    def forcing_state(self, what=_lib.FFS_FSTATE_R):
        """State of the plane forcing (R, Em, N, mean |k_h|, current spectrum): one value per 2D shell;
        for aForcing one per mode of the reference's (NX, NY) matrices, x fastest."""
        n = ctypes.c_int()
        self._check(self._lib.ffs_forcing_state(self._h, what, 0, None, 0, ctypes.byref(n)))
        a = np.zeros(n.value)
        self._check(self._lib.ffs_forcing_state(self._h, what, 0, _dptr(a), a.size, ctypes.byref(n)))
        return a

    def set_forcing_state(self, what, values):
        """Sets R (an R.N file), N (an N.N file) or the target spectrum Em (targSpectrum.dat; before
        initialize, so the target is not taken from the initial field)."""
        a = np.ascontiguousarray(values, dtype=np.float64)
        n = ctypes.c_int()
        self._check(self._lib.ffs_forcing_state(self._h, what, 1, _dptr(a), a.size, ctypes.byref(n)))

    def checkpoint(self):
        """The run's full integrator state as bytes (ffs_checkpoint_save): a run restored from it
        continues bitwise as this one would."""
        n = ctypes.c_size_t()
        self._check(self._lib.ffs_checkpoint_size(self._h, ctypes.byref(n)))
        buf = ctypes.create_string_buffer(n.value)
        self._check(self._lib.ffs_checkpoint_save(self._h, buf, n.value))
        return buf.raw

    def restore(self, blob):
        """Loads a checkpoint into this (not yet initialized) run; replaces set_field + initialize."""
        buf = ctypes.create_string_buffer(bytes(blob), len(blob))
        self._check(self._lib.ffs_checkpoint_load(self._h, buf, len(blob)))

    def write_checkpoint(self, path):
        with open(path, "wb") as f:
            f.write(self.checkpoint())

    def read_checkpoint(self, path):
        with open(path, "rb") as f:
            self.restore(f.read())

    def writestats_now(self):
        self._check(self._lib.ffs_writestats_now(self._h))

    def synchronize(self):
        self._check(self._lib.ffs_synchronize(self._h))

    def set_graph(self, on=True):
        """Replay steady steps as HIP graphs (default on; bitwise identical to eager)."""
        self._check(self._lib.ffs_set_graph(self._h, int(on)))

    def set_fusion(self, on=True):
        """Run the forward z pass and the update of plain steps as one kernel (default on for FP64, off
        for FP32; bitwise identical to the two kernels)."""
        self._check(self._lib.ffs_set_fusion(self._h, int(on)))

    PLAN_FIELDS = ("x_pair", "x_waves", "x_split", "fusion", "z_slab")

    @property
    def plan(self):
        """The run's kernel plan (ffs_plan) as a dict."""
        p = _lib.FfsPlan()
        self._check(self._lib.ffs_get_plan(self._h, ctypes.byref(p)))
        return {k: getattr(p, k) for k in self.PLAN_FIELDS}

    def set_plan(self, plan):
        """Applies a plan (a dict with any of x_pair, x_waves, x_split, fusion, z_slab; the rest kept)."""
        cur = self.plan
        unknown = set(plan) - set(cur)
        if unknown:
            raise KeyError(f"unknown plan fields {sorted(unknown)}")
        cur.update(plan)
        p = _lib.FfsPlan(**{k: int(v) for k, v in cur.items()})
        self._check(self._lib.ffs_set_plan(self._h, ctypes.byref(p)))

    def plan_options(self):
        """Which plan fields change this run's kernels: dict of field -> bool; x_pair -> the largest value that
        applies (0: none, 1: the mirror-paired pass, 2: its looping form too); z_slab -> the run's default planes
        per z slab (0: slabs do not apply)."""
        p = _lib.FfsPlan()
        self._check(self._lib.ffs_plan_options(self._h, ctypes.byref(p)))
        return {k: (int(getattr(p, k)) if k in ("z_slab", "x_pair") else bool(getattr(p, k)))
                for k in self.PLAN_FIELDS}

    # ---------------------------------------------------------------- diagnostics
    def pop_stats(self):
        """Returns the queued stats records as a list of (stats, scales, numbers) rows."""
        out = []
        a = np.zeros(64)
        b = np.zeros(64)
        c = np.zeros(64)
        na, nb, nc = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        while True:
            r = self._check(self._lib.ffs_stats_pop(self._h, _dptr(a), 64, ctypes.byref(na), _dptr(b), 64,
                                                    ctypes.byref(nb), _dptr(c), 64, ctypes.byref(nc)))
            if r == 0:
                return out
            out.append((a[: na.value].copy(), b[: nb.value].copy(), c[: nc.value].copy()))

    def header(self, table):
        buf = ctypes.create_string_buffer(4096)
        self._check(self._lib.ffs_stats_header(self._h, table, buf, 4096))
        return buf.value.decode()

    def spectrum(self):
        """spectrum_u + calculate_specHV (spec.jl:316-327,424-450):
        returns ((u_h.specH, u_h.specV), (u_v.specH, u_v.specV))."""
        h = np.zeros(8192)
        v = np.zeros(8192)
        nh, nv = ctypes.c_int(), ctypes.c_int()
        self._check(self._lib.ffs_spectrum(self._h, _dptr(h), h.size, ctypes.byref(nh), _dptr(v), v.size,
                                           ctypes.byref(nv)))
        a, b = nh.value, nv.value
        return (h[:a].copy(), v[:b].copy()), (h[a:2 * a].copy(), v[b:2 * b].copy())

    def spectra_axes(self):
        """(KH, KRZ) of Globals.jl:148-152: the bins of the specH and specV files."""
        kh = np.zeros(8192)
        kz = np.zeros(8192)
        nh, nv = ctypes.c_int(), ctypes.c_int()
        self._check(self._lib.ffs_spectra_axes(self._h, _dptr(kh), kh.size, ctypes.byref(nh), _dptr(kz), kz.size,
                                               ctypes.byref(nv)))
        return kh[: nh.value].copy(), kz[: nv.value].copy()

    def pop_spectra(self):
        """The queued spectra records (write_spec / write_spec_forcing, spec.jl:1-168, one per spectra
        iteration) as a list of (iteration, {quantity: (specH, specV)}), quantities named as in
        _lib.SPEC_QUANTITIES; the ones this run does not have are left out."""
        kh, kz = self.spectra_axes()
        nh, nv = kh.size, kz.size
        buf = np.zeros(_lib.FFS_SPEC_Q * (nh + nv))
        it = ctypes.c_int64()
        n = ctypes.c_size_t()
        have = set(_lib.SPEC_QUANTITIES[:8])
        if self.params.les_model not in (None, "", "noLesModel"):
            have |= {"les_h", "les_v"}
        if self.params.density:
            have |= {"buoyancy_v", "rho", "rhodiss", "rhonl"}
        if self.params.forcing in PLANE_FORCINGS:
            have.add("forcing_h")
        if self.params.hyperviscosity is not None:
            have |= {"hvis_h", "hvis_v"}
        out = []
        while True:
            r = self._check(self._lib.ffs_spectra_pop(self._h, ctypes.byref(it), _dptr(buf), buf.size,
                                                      ctypes.byref(n)))
            if r == 0:
                return out
            rec = buf.reshape(_lib.FFS_SPEC_Q, nh + nv)
            out.append((int(it.value), {q: (rec[i, :nh].copy(), rec[i, nh:].copy())
                                        for i, q in enumerate(_lib.SPEC_QUANTITIES) if q in have}))

    def writespectra(self, directory="."):
        """Writes every queued spectra record as write_spec / write_spec_forcing do (writespectrum,
        spec.jl:171-194; writedlm of (KH, specH) and (KRZ, specV) rows)."""
        kh, kz = self.spectra_axes()
        recs = self.pop_spectra()
        for it, q in recs:
            for name, (h, v) in spectra_files(q, self.params).items():
                axis = kh if name.endswith("specH") else kz
                _writedlm(os.path.join(directory, f"{name}.{it}"), axis, h if v is None else v)
        return recs

    def scalar(self, what):
        out = ctypes.c_double()
        self._check(self._lib.ffs_get_scalar(self._h, what, ctypes.byref(out)))
        return out.value

    @property
    def iteration(self):
        return int(self.scalar(_lib.FFS_SCALAR_ITERATION))

    @property
    def time(self):
        return self.scalar(_lib.FFS_SCALAR_TIME)

    @property
    def dt(self):
        return self.scalar(_lib.FFS_SCALAR_DT)

    def device_bytes(self):
        b = ctypes.c_size_t()
        self._check(self._lib.ffs_device_bytes(self._h, ctypes.byref(b)))
        return b.value

    # ---------------------------------------------------------------- profiling
    def profile(self, on=True):
        self._check(self._lib.ffs_profile_enable(self._h, int(on)))

    def profile_reset(self):
        self._check(self._lib.ffs_profile_reset(self._h))

    def profile_get(self):
        cap = 64
        names = (ctypes.c_char_p * cap)()
        ms = np.zeros(cap)
        cnt = np.zeros(cap, dtype=np.int64)
        n = ctypes.c_int()
        self._check(self._lib.ffs_profile_get(self._h, names, _dptr(ms),
                                              cnt.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), cap,
                                              ctypes.byref(n)))
        return {names[i].decode(): (ms[i], int(cnt[i])) for i in range(min(n.value, cap))}

    # ---------------------------------------------------------------- files
    def writeheader(self, directory="."):
        """writeheader (stats.jl:9-37)."""
        for name, t in (("Stats.txt", _lib.FFS_TABLE_STATS), ("Scales.txt", _lib.FFS_TABLE_SCALES),
                        ("Numbers.txt", _lib.FFS_TABLE_NUMBERS)):
            with open(os.path.join(directory, name), "w") as f:
                f.write(self.header(t) + "\n")

    def writestats(self, directory="."):
        """Appends every queued record (writestats, stats.jl:39-62)."""
        recs = self.pop_stats()
        for st, sc, nu in recs:
            for name, row in (("Stats.txt", st), ("Scales.txt", sc), ("Numbers.txt", nu)):
                with open(os.path.join(directory, name), "a+") as f:
                    f.write(",".join(_fmt(v, i) for i, v in enumerate(row)) + "\n")
        self.stats_rows.extend(recs)
        return recs

    def writeoutput(self, directory=".", closing=False):
        """writeoutput (output.jl:1-67): u1.$it u2.$it u3.$it (padded real), scalar.$it with the passive
        scalar, rho.$it with density, dflux1..3.$it with a density LES model (its real-space LES flux
        nu_t grad rho (+ the vorticity term) of the current state), cs.$it (and cp.$it) with a dynamic
        model (the averaged coefficient the current velocity gives, with Piomelli's extrapolation), and
        the forcing state: R.$it (the oscillators), N.$it (nRf, cNRf), Eh.$it = (mean |k_h|, Ef)
        rows (rf, cRf, nRf, cNRf); with Smagorinsky the stress T11 T12 T13 T22 T23: the real-space
        product of the current velocity, or with `closing` the stored transform back in real space
        (the real!(tau) of run_simulation's closing output, FluidFlowSimulation.jl:53-57)."""
        it = self.iteration
        for f in range(3):
            write_field(os.path.join(directory, f"u{f + 1}.{it}"), self.get_field(f, _lib.FFS_SPACE_REAL))
        if self.params.passive_scalar:
            write_field(os.path.join(directory, f"scalar.{it}"),
                        self.get_field(_lib.FFS_FIELD_SCALAR, _lib.FFS_SPACE_REAL))
        if self.params.density:
            write_field(os.path.join(directory, f"rho.{it}"), self.get_field(_lib.FFS_FIELD_RHO, _lib.FFS_SPACE_REAL))
            if self.params.density_les is not None and self.params.les_model not in (None, "", "noLesModel"):
                for c in range(3):
                    write_field(os.path.join(directory, f"dflux{c + 1}.{it}"),
                                self.get_field(_lib.FFS_FIELD_DFLUX1 + c, _lib.FFS_SPACE_REAL))
        if self.params.les_model not in (None, "", "noLesModel"):
            space = _lib.FFS_SPACE_FOURIER_TO_REAL if closing else _lib.FFS_SPACE_REAL
            for name, f in (("T11", _lib.FFS_FIELD_T11), ("T12", _lib.FFS_FIELD_T12), ("T13", _lib.FFS_FIELD_T13),
                            ("T22", _lib.FFS_FIELD_T22), ("T23", _lib.FFS_FIELD_T23)):
                write_field(os.path.join(directory, f"{name}.{it}"), self.get_field(f, space))
            lm = self.params.les_model
            if lm in ("dynamicSmagorinsky", "piomelliSmagorinsky", "dynamicSandP", "piomelliSandP"):
                write_field(os.path.join(directory, f"cs.{it}"), self.get_field(_lib.FFS_FIELD_CS, _lib.FFS_SPACE_REAL))
                if lm in ("dynamicSandP", "piomelliSandP"):
                    write_field(os.path.join(directory, f"cp.{it}"),
                                self.get_field(_lib.FFS_FIELD_CP, _lib.FFS_SPACE_REAL))
        forcing = self.params.forcing
        if forcing in RF_FAMILY:
            R = self.forcing_state(_lib.FFS_FSTATE_R)
            if forcing == "aForcing":  # the (NX, NY) matrix, one row per kx
                R = R.reshape(self.params.ny, -1).T
            _writedlm_values(os.path.join(directory, f"R.{it}"), R)
        if forcing in ("nRfForcing", "cNRfForcing"):
            _writedlm_values(os.path.join(directory, f"N.{it}"), self.forcing_state(_lib.FFS_FSTATE_N))
        if forcing in ("rfForcing", "cRfForcing", "nRfForcing", "cNRfForcing"):
            _writedlm(os.path.join(directory, f"Eh.{it}"), self.forcing_state(_lib.FFS_FSTATE_KH),
                      self.forcing_state(_lib.FFS_FSTATE_EF))


def footprint(params: Params):
    """Device bytes a run with these parameters needs (no device work)."""
    lib = _lib.load()
    cp = params.to_c()
    b = ctypes.c_size_t()
    _lib.check(lib.ffs_footprint(ctypes.byref(cp), ctypes.byref(b)))
    return b.value


def spectra_files(q, params):
    """{file stem: (values, None)} of one spectra record: writespectrum's five files for u, vis, nl,
    press (and les, hvis), the specH/specV pair of the single spectra, rhosource from buoyancy_v
    (spec.jl:1-168)."""
    out = {}
    for n in ("u", "vis", "nl", "press", "les", "hvis"):
        if n + "_h" not in q:
            continue
        (hH, hV), (vH, vV) = q[n + "_h"], q[n + "_v"]
        out[f"{n}_h.specH"] = (hH, None)
        out[f"{n}_h.specV"] = (hV, None)
        out[f"{n}.specH"] = (hH + vH, None)
        out[f"{n}_v.specH"] = (vH, None)
        out[f"{n}_v.specV"] = (vV, None)
    if "buoyancy_v" in q:
        bH, bV = q["buoyancy_v"]
        out["buoyancy_v.specH"] = (bH, None)
        out["buoyancy_v.specV"] = (bV, None)
        gz = params.z_acceleration / params.reference_density
        with np.errstate(divide="ignore", invalid="ignore"):
            out["rhosource.specH"] = (-(bH / gz) * params.density_gradient, None)
            out["rhosource.specV"] = (-(bV / gz) * params.density_gradient, None)
    for n in ("rho", "rhodiss", "rhonl", "forcing_h"):
        if n in q:
            out[f"{n}.specH"] = (q[n][0], None)
            out[f"{n}.specV"] = (q[n][1], None)
    return out


def _writedlm_values(path, a):
    """writedlm of a vector (one value per line) or a matrix (one row per line, tab separated)."""
    a = np.asarray(a, dtype=np.float64)
    with open(path, "w") as f:
        for row in (a.reshape(-1, 1) if a.ndim == 1 else a):
            f.write("\t".join(jl_float(x) for x in row) + "\n")


def _writedlm(path, k, v):
    with open(path, "w") as f:
        for a, b in zip(k, v):
            f.write(f"{jl_float(a)}\t{jl_float(b)}\n")


def jl_float(x):
    """A Float64 as Julia prints it (shortest round-trip digits; scientific notation below 1e-4 and
    from 1e6 on, as 1.0e-5 / 1.048576e6)."""
    x = float(x)
    if x != x:
        return "NaN"
    if math.isinf(x):
        return "Inf" if x > 0 else "-Inf"
    if x == 0.0:
        return "-0.0" if math.copysign(1.0, x) < 0 else "0.0"
    sign, digits, exp = Decimal(repr(x)).as_tuple()
    d = "".join(map(str, digits))
    stripped = d.rstrip("0")
    exp += len(d) - len(stripped)
    d = stripped or "0"
    e10 = len(d) - 1 + exp
    if e10 < -4 or e10 >= 6:
        body = f"{d[0]}.{d[1:] or '0'}e{e10}"
    elif exp >= 0:
        body = d + "0" * exp + ".0"
    else:
        point = len(d) + exp
        body = (d[:point] + "." + d[point:]) if point > 0 else ("0." + "0" * (-point) + d)
    return ("-" if sign else "") + body


def _fmt(v, i):
    if i == 0:
        return str(int(v))
    return jl_float(v)


def run_simulation(directory=".", params=None, nt=None):
    """run_simulation() of FluidFlowSimulation.jl:23-65: reads `global` and
    u1.$start..u3.$start from `directory`, runs, and writes Stats/Scales/Numbers
    rows and field dumps there."""
    if params is None:
        params = parameters(readglobal(os.path.join(directory, "global")))
    if nt is None:
        nt = params.nt if params.nt is not None else (1 << 62) - params.start
    s = Simulation(params)
    nx, ny, nz = params.nx, params.ny, params.nz
    for f in range(3):
        s.set_field(f, read_field(os.path.join(directory, f"u{f + 1}.{params.start}"), nx, ny, nz))
    scalar_file = os.path.join(directory, f"scalar.{params.start}")
    if params.passive_scalar and os.path.exists(scalar_file):  # zero scalar without one (simulationtypes.jl:324)
        s.set_field(_lib.FFS_FIELD_SCALAR, read_field(scalar_file, nx, ny, nz))
    rho_file = os.path.join(directory, f"rho.{params.start}")
    if params.density and os.path.exists(rho_file):  # zero density without one (density_stratification_types.jl:166)
        s.set_field(_lib.FFS_FIELD_RHO, read_field(rho_file, nx, ny, nz))
    _load_forcing_files(s, directory, params)
    if params.start == 0:
        s.writeheader(directory)
    s.initialize()
    assert nt + params.start >= params.write_time
    assert nt + params.start >= params.dt_stat
    final_step = params.start + nt
    final_time = params.start_time + params.time_duration
    while s.iteration < final_step and (math.isinf(final_time) or s.time < final_time):
        left = final_step - s.iteration
        if not math.isinf(final_time):
            left = 1
        done, ev = s.advance(left)
        if ev == _lib.FFS_EVENT_OUTPUT:
            s.writeoutput(directory)
        s.writestats(directory)
        s.writespectra(directory)
    # the closing calculate_rhs! writes the output files when the last iteration is an output
    # iteration (realspace!, models.jl:157), else run_simulation writes them after the forcing
    # (FluidFlowSimulation.jl:53-61): the forcing state files differ by that forcing call
    due = s.iteration % params.write_time == 0
    if due:
        s.writeoutput(directory)
    s.writestats_now()
    s.writestats(directory)
    s.writespectra(directory)
    if not due:
        s.writeoutput(directory, closing=True)
    if params.forcing == "rfForcing":  # writedlm("R.$it", s.forcing.R), FluidFlowSimulation.jl:59-61
        _writedlm_values(os.path.join(directory, f"R.{s.iteration}"), s.forcing_state(_lib.FFS_FSTATE_R))
    return s


def _load_forcing_files(s, directory, params):
    """The files the plane forcings read at initialize!: targSpectrum.dat for the shell oscillators
    (calculate_Em!, forcing_types.jl:76-83: skip the header line, linear interpolation at the shells'
    mean |k_h|; np.interp holds the end values outside the table), R.$start (forcing_types.jl:66-74,
    112-118, 151-164, 207-216, 251-264; aForcing's is the (NX, NY) matrix, read column-major as vec
    does) and N.$start for nRfForcing / cNRfForcing."""
    if params.forcing not in RF_FAMILY:
        return
    targ = os.path.join(directory, "targSpectrum.dat")
    if params.forcing != "aForcing" and os.path.exists(targ):
        data = np.loadtxt(targ, skiprows=1)
        kh = s.forcing_state(_lib.FFS_FSTATE_KH)
        s.set_forcing_state(_lib.FFS_FSTATE_EM, np.interp(np.nan_to_num(kh), data[:, 0], data[:, 1]))
    rfile = os.path.join(directory, f"R.{params.start}")
    if os.path.exists(rfile):
        s.set_forcing_state(_lib.FFS_FSTATE_R, np.loadtxt(rfile).ravel(order="F"))
    nfile = os.path.join(directory, f"N.{params.start}")
    if params.forcing in ("nRfForcing", "cNRfForcing") and os.path.exists(nfile):
        s.set_forcing_state(_lib.FFS_FSTATE_N, np.loadtxt(nfile).ravel(order="F"))
