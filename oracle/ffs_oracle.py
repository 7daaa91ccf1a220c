"""CPU restatement of FluidFlowSimulation.jl's time step (FP64, numpy/scipy).

TEST INFRASTRUCTURE ONLY.  Nothing in the product path imports this module:
it is the parity checker of tests/, the checker of __graft_entry__.smoke()
and the CPU baseline leg of bench.py.

It follows the reference's structure literally, on full FFTW-layout arrays
(half spectrum (nx/2+1, ny, nz), x fastest), with the 12 real transforms of a
plain step, separate dealias/scale sweeps, the u round trip and the
projection after it:
  advance_in_time!      models.jl:3-17
  calculate_rhs!        models.jl:19-51
  fourierspacep1!       models.jl:53-119       (w = ik x u)
  realspace!            models.jl:121-214      (u x w, 1/2|u|^2, u round trip)
  fourierspacep2_velocity! models.jl:292-341   (-nu k^2 u, Leray)
  pressure_projection!  models.jl:443-456
  myfourier!/dealias!/myscale!  utils.jl:41-51,101-137
  AB3                   Adams_Bashforth.jl:142-256, fix_fm2! ETD.jl:130-140
  set_dt!               time_step_functions.jl:86-129
  stats                 stats.jl:64-219,289-331,608-635,750-814,897-917,952-980
  spectrum_u + calculate_specHV  spec.jl:316-327,424-450
The FFT is scipy.fft (pocketfft), standing in for FFTW with the same
definitions (unnormalised forward e^{-ikx}, unnormalised backward).

Parity status of each function is recorded in its docstring ("pinned by"):
the checks live in tests/test_oracle.py.
"""
import math
import os

import numpy as np
import scipy.fft as sfft

from . import les_models

__all__ = ["Grid", "OracleSim", "taylor_green", "workers"]


def workers():
    """FFT threads: FFS_ORACLE_WORKERS, else the host's cores capped at 16."""
    return int(os.environ.get("FFS_ORACLE_WORKERS", min(16, os.cpu_count() or 1)))


def rfftfreq(n, s):
    """Globals.jl:95-97."""
    return np.array([(n / 2 - i) / s for i in np.arange(n / 2, -1, -1)], dtype=np.float64)


def fftfreq(n, s):
    """Globals.jl:99-104."""
    if n % 2 == 0:
        a = [(n / 2 - i) / s for i in np.arange(n / 2, 0, -1)]
        b = [-i / s for i in np.arange(n / 2, 0, -1)]
    else:
        a = [(n / 2 - i) / s for i in np.arange(n / 2, -0.5, -1)]
        b = [-i / s for i in np.arange((n - 1) / 2, 0, -1)]
    return np.array(a + b, dtype=np.float64)


class Grid:
    """Module-load constants of Globals.jl:106-182.

    pinned by: closed forms of the wavenumbers, mask and loop ranges
    (tests/test_oracle.py::test_grid_*)."""

    def __init__(self, nx, ny, nz, lx=1.0, ly=1.0, lz=1.0, cutoff=2 / 3):
        self.NRX, self.NY, self.NZ = nx, ny, nz
        self.NX = nx // 2 + 1
        self.LX, self.LY, self.LZ = lx, ly, lz
        kxp = rfftfreq(nx, lx)
        kyp = fftfreq(ny, ly)
        kzp = fftfreq(nz, lz)
        self.KX, self.KY, self.KZ = kxp, kyp, kzp
        cutoffx = (cutoff * kxp[-1]) ** 2
        cutoffy = (cutoff * np.max(np.abs(kyp))) ** 2
        cutoffz = (cutoff * np.max(np.abs(kzp))) ** 2
        KXa = kxp[:, None, None]
        KYa = kyp[None, :, None]
        KZa = kzp[None, None, :]
        self.DEALIAS = (KXa**2 > cutoffx) | (KYa**2 > cutoffy) | (KZa**2 > cutoffz)
        self.kx, self.ky, self.kz = KXa, KYa, KZa
        self.K2 = KXa * KXa + KYa * KYa + KZa * KZa
        # loop ranges (Globals.jl:163-169), 0-based
        fx = int(np.argmax(self.DEALIAS[:, 0, 0])) if self.DEALIAS[:, 0, 0].any() else self.NX
        fy = int(np.argmax(self.DEALIAS[0, :, 0])) if self.DEALIAS[0, :, 0].any() else ny
        fz = int(np.argmax(self.DEALIAS[0, 0, :])) if self.DEALIAS[0, 0, :].any() else nz
        self.XRANGE = np.arange(0, fx)
        self.YRANGE = np.concatenate([np.arange(0, fy), np.arange(ny - fy - 2, ny)]) if fy < ny else np.arange(ny)
        self.ZRANGE = np.concatenate([np.arange(0, fz), np.arange(nz - fz - 2, nz)]) if fz < nz else np.arange(nz)
        # retained loop set as a mask (the ranges cover every unmasked mode)
        R = np.zeros((self.NX, ny, nz), dtype=bool)
        R[np.ix_(self.XRANGE, self.YRANGE, self.ZRANGE)] = True
        self.RANGE = R
        self.W = np.where(np.arange(self.NX) > 0, 2.0, 1.0)[:, None, None]  # (1 + (i>1))
        self.scale = 1.0 / (nx * ny * nz)
        # spectra bins (Globals.jl:148-152)
        self.KH = kxp if kxp[1] >= rfftfreq(ny, ly)[1] else rfftfreq(ny, ly)
        self.MAXDKH = self.KH[1]
        self.KRZ = rfftfreq(nz, lz)
        self.DKZ = self.KRZ[1]
        self.cutoff = cutoff

    def coords(self):
        """Physical grid points x_i = 2 pi L i / n."""
        x = 2 * np.pi * self.LX * np.arange(self.NRX) / self.NRX
        y = 2 * np.pi * self.LY * np.arange(self.NY) / self.NY
        z = 2 * np.pi * self.LZ * np.arange(self.NZ) / self.NZ
        return np.meshgrid(x, y, z, indexing="ij")

    # transforms on (nx, ny, nz) real arrays, x = axis 0
    def rfft(self, a):
        """Unnormalised R2C (FFTW forward).  pinned by: brute-force DFT (test_oracle)."""
        return sfft.rfftn(a, axes=(2, 1, 0), workers=workers())

    def irfft(self, c):
        """Unnormalised C2R (FFTW backward).  pinned by: brute-force DFT (test_oracle)."""
        return sfft.irfftn(c, s=(self.NZ, self.NY, self.NRX), axes=(2, 1, 0), norm="forward", workers=workers())

    def myfourier(self, a):
        """rfft! + dealias! + myscale! (utils.jl:101-106)."""
        c = self.rfft(a)
        c[self.DEALIAS] = 0
        c[self.RANGE] *= self.scale
        return c

    def project(self, v):
        """pressure_projection! (models.jl:443-456) on the retained range."""
        kx, ky, kz = self.kx, self.ky, self.kz
        with np.errstate(invalid="ignore", divide="ignore"):
            p1 = -(kx * v[0] + ky * v[1] + kz * v[2]) / self.K2
        p1[0, 0, 0] = 0
        p1 = np.where(self.RANGE, p1, 0)
        return np.array([p1 * kx + v[0], p1 * ky + v[1], p1 * kz + v[2]])


def taylor_green(grid):
    """u1 = cos x sin y sin z, u2 = -sin x cos y sin z, u3 = 0 (the Taylor-Green case)."""
    x, y, z = grid.coords()
    return np.array([np.cos(x) * np.sin(y) * np.sin(z), -np.sin(x) * np.cos(y) * np.sin(z), np.zeros_like(x)])


def proj(a, b):
    """proj(a,b) = re(a) re(b) + im(a) im(b) (Forcing_methods/forcing.jl:4)."""
    return a.real * b.real + a.imag * b.imag


class OracleSim:
    """Plain vorticity-form run with AB3 (fixed or variable dt) and optional
    isotropic constant-power shell forcing.

    pinned by: Taylor-Green t=0 closed forms (k=1/8, eps=6 nu k, pressure),
    single-mode curl/projection closed forms, divergence-free state, energy
    budget dk/dt = -eps, AB3 constant-step limit and third-order convergence
    (tests/test_oracle.py)."""

    def __init__(self, grid, nu, dt=0.0, variable_dt=False, cfl=0.18, dt_stat=1 << 30, dt_spec=1 << 30,
                 forcing_kf=None, forcing_power=0.1, start=0, start_time=0.0, integrator="AB3", les_c=None,
                 les_delta=None, density=None, plane_forcing=None, hv=None, scalar=None, equation="rotational",
                 les_model="Smagorinsky", les_cp=None, les_smodel=None, les_test_width=None, les_cmin=0.0,
                 density_les=None, scalar_les=None):
        """integrator: "AB3" (Adams_Bashforth3rdO), "Euler" (Euller) or "ETD3" (ETD3rdO);
        les_c: the LES model's constant (None: no LES), les_delta: filter width
        (default (2 pi L/nx)/cutoff, les_types.jl:366); les_model: Smagorinsky or one of
        oracle.les_models.MODELS, with les_cp (Silvis P constant, SandP pConstant) and les_smodel
        (SandP's sLesModel); dynamicSmagorinsky takes les_test_width (default twice the filter
        width) and les_cmin; density_les / scalar_les: None, "eddy" or ("vorticity", c), the scalars'
        LES fluxes (EddyDiffusion, VorticityDiffusion(c); les_real_calc.jl:120-136);
        density: dict(Pr=, gradient=, g=[, dir=][, hv=]) for the Boussinesq approximation, dir the
        gravity's axis (gravityDirection, 0 1 2, default 2), hv its own
        hyperviscosity (densityHyperViscosity, as hv below; none by default);
        hv: dict(kind="hyper", nu=, m=) (HyperViscosity{nu_h, M}) or dict(kind="barrier",
        initk=, endk=) (SpectralBarrier), Hyperviscosity/hyperviscosity_types.jl;
        scalar: dict(Pr=, gradient=, dir=0|1|2) for the passive scalar (PassiveScalar,
        simulationtypes.jl:198-224, 320-352); equation: "rotational" (u x w) or "NS"
        (the conservation form, -div(uu) with uu = traceless(symouter(u, -u)), models.jl:176-180, 262)."""
        self.g = grid
        self.nu = nu
        self.integrator = integrator
        self.les_c = les_c
        self.density_les, self.scalar_les = density_les, scalar_les
        self.les_model, self.les_cp, self.les_smodel = les_model, les_cp, les_smodel
        if les_c is not None:
            delta = les_delta if les_delta else les_models.default_delta(les_model, les_smodel, grid.LX, grid.NRX,
                                                                          grid.cutoff)
            self.les_alpha = les_c * les_c * delta * delta
            self.les_d2 = delta * delta
            td = les_test_width if les_test_width else 2 * delta
            self.les_td2, self.les_cmin = td * td, les_cmin
            # dynamicSandP / piomelliSandP carry their eddy-viscosity model in les_smodel
            dmodel = les_smodel if les_model in ("dynamicSandP", "piomelliSandP") else les_model
            shape = (grid.NRX, grid.NY, grid.NZ)
            if dmodel == "piomelliSmagorinsky":  # c, cm1 start at 0.0289, dtm1 at 0 (les_types.jl:296-300)
                self.pio_c, self.pio_cm1, self.pio_dtm1 = np.full(shape, 0.0289), np.full(shape, 0.0289), 0.0
            if les_model == "piomelliSandP":  # cp, cpm1 start at 0.09, dtpm1 at 0 (les_types.jl:320-329)
                self.pio_cp, self.pio_cpm1, self.pio_dtpm1 = np.full(shape, 0.09), np.full(shape, 0.09), 0.0
            self.les_dmodel = dmodel
        self.density = density
        if density is not None:
            self.alpha = nu / density["Pr"]
            self.drhodz = density["gradient"]
            # gravityDirection (density_stratification_types.jl:152-162): g = g e_dir, the mean gradient's
            # direction too; calculate_N and spectrum_buoyancy read g.z, which is 0 off the z axis
            self.gdir = density.get("dir", 2)
            self.gval = density["g"]
            self.gz = self.gval if self.gdir == 2 else 0.0
        self.variable_dt = variable_dt
        self.cfl = cfl
        self.dt = self.dt2 = self.dt3 = dt
        self.dt_stat = dt_stat
        self.dt_spec = dt_spec
        self.iteration = start
        self.time = start_time
        self.tit = 0
        self.forcing_kf = forcing_kf
        self.forcing_power = forcing_power
        self.force_factor = 0.0
        self.plane = plane_forcing  # oracle.plane_forcing.PlaneForcing (one of the kz = 0 plane forcings) or None
        self.hv = hv
        self.equation = equation
        self.scalar = scalar
        if scalar is not None:
            self.palpha = nu / scalar["Pr"]
        self.umax = 0.0
        self.records = []
        self.spectra = []

    # -------------------------------------------------------------- setup
    def set_velocity(self, u):
        """Real (3, nx, ny, nz) field -> myfourier! of initialize! (FluidFlowSimulation.jl:69)."""
        self.uh = np.array([self.g.myfourier(np.asarray(c, dtype=np.float64)) for c in u])

    def set_scalar(self, phi):
        """myfourier!(phi) + dealias! of initialize! (FluidFlowSimulation.jl:72-75)."""
        self.phih = self.g.myfourier(np.asarray(phi, dtype=np.float64))

    def set_density(self, rho):
        """myfourier!(rho) + dealias! of initialize! (FluidFlowSimulation.jl:77-80)."""
        self.rhoh = self.g.myfourier(np.asarray(rho, dtype=np.float64))

    def initialize(self):
        """initialize!(s) (FluidFlowSimulation.jl:67-91)."""
        if self.density is not None and not hasattr(self, "rhoh"):
            self.rhoh = np.zeros_like(self.uh[0])
        if self.scalar is not None and not hasattr(self, "phih"):
            self.phih = np.zeros_like(self.uh[0])
        self.calculate_rhs(collect=False)
        self.fm1 = self.rhs.copy()
        self.fm2 = self.fm1.copy()
        self.dt2 = self.dt
        self.dt3 = self.dt
        if self.density is not None:
            self.rfm1 = self.rhorhs.copy()
            self.rfm2 = self.rfm1.copy()
        if self.scalar is not None:
            self.pfm1 = self.phirhs.copy()
            self.pfm2 = self.pfm1.copy()
        if self.plane is not None:
            self.plane.initialize(self)

    # -------------------------------------------------------------- the step
    def is_stats_time(self):
        return self.iteration % self.dt_stat == 0

    def is_spec_time(self):
        return self.iteration % self.dt_spec == 0

    def calculate_rhs(self, collect=True):
        g = self.g
        u = self.uh
        kx, ky, kz = g.kx, g.ky, g.kz
        kv = (kx, ky, kz)
        # fourierspacep1!: rhs = (ik) x u; LES: tau = S = symouter(ik, u) (xx, xy, xz, yy, yz)
        w = np.array([1j * (ky * u[2] - kz * u[1]), 1j * (kz * u[0] - kx * u[2]), 1j * (kx * u[1] - ky * u[0])])
        w = np.where(g.RANGE, w, 0)
        if self.les_c is not None:
            pairs = ((0, 0), (0, 1), (0, 2), (1, 1), (1, 2))
            S = np.array([np.where(g.RANGE, 0.5 * (1j * kv[a] * u[b] + 1j * kv[b] * u[a]), 0) for a, b in pairs])
        # realspace!
        ur = np.array([g.irfft(c) for c in u])
        wr = np.array([g.irfft(c) for c in w])
        need_k = self.is_spec_time() or self.is_stats_time()
        if need_k:
            self.kspec = g.myfourier(0.5 * (ur[0] * ur[0] + ur[1] * ur[1] + ur[2] * ur[2]))
        if self.density is not None:
            rhor = g.irfft(self.rhoh)
        if self.scalar is not None:
            phir = g.irfft(self.phih)
        if self.les_c is not None:
            Sr = np.array([g.irfft(c) for c in S])
            # calc_les! Smagorinsky branch (les_real_calc.jl:77-79): tau = 2 (c Delta)^2 |S| S,
            # |S| = sqrt(2 S:S) with Szz = -Sxx - Syy
            szz = -Sr[0] - Sr[3]
            ss = Sr[0] ** 2 + Sr[3] ** 2 + szz**2 + 2 * (Sr[1] ** 2 + Sr[2] ** 2 + Sr[4] ** 2)
            if self.les_model == "Smagorinsky":
                taur = 2 * (self.les_alpha * np.sqrt(2 * ss)) * Sr
            elif self.les_model in ("dynamicSmagorinsky", "piomelliSmagorinsky", "dynamicSandP", "piomelliSandP"):
                # realspace_dynamic_les_calculation! (dynamic_smagorinsky.jl:3-41): the Piomelli
                # coefficients are extrapolated first, c then cp, each with its own previous dt
                cstar = cpstar = None
                if self.les_dmodel == "piomelliSmagorinsky":
                    cstar, self.pio_cm1 = les_models.piomelli_extrapolate(self.pio_c, self.pio_cm1, self.dt,
                                                                          self.pio_dtm1, self.iteration)
                    self.pio_dtm1 = self.dt
                if self.les_model == "piomelliSandP":
                    cpstar, self.pio_cpm1 = les_models.piomelli_extrapolate(self.pio_cp, self.pio_cpm1, self.dt,
                                                                            self.pio_dtpm1, self.iteration)
                    self.pio_dtpm1 = self.dt
                pmode = {"dynamicSandP": "dyn", "piomelliSandP": "pio"}.get(self.les_model)
                taur, c, cp = les_models.dynamic_stress(g, u, S, Sr, ur, self.les_d2, self.les_td2, self.les_cmin,
                                                        cstar=cstar, wh=w, wr=wr, pmode=pmode, cpstar=cpstar)
                self.les_coef, self.les_pcoef = c, cp  # cs and cp of writeoutput (output.jl:36-41)
                if cstar is not None:
                    self.pio_c = c
                if cpstar is not None:
                    self.pio_cp = cp
            else:  # the other pointwise models (les_real_calc.jl:80-114), from S and the vorticity
                taur = les_models.stress(self.les_model, Sr, wr, self.les_c, self.les_d2, self.les_cp,
                                         self.les_smodel)
            self.taur = taur  # the real-space stress writeoutput writes (T11 ... T23, output.jl:28-35)
        if self.equation == "NS":
            # traceless(symouter(u, -u)): xx xy xz yy yz, then div!(rhs, uu) = (i k) . uu
            xx, yy, zz = -(ur[0] * ur[0]), -(ur[1] * ur[1]), -(ur[2] * ur[2])
            tr = (xx + yy + zz) / 3
            uur = np.array([xx - tr, -(ur[0] * ur[1]), -(ur[0] * ur[2]), yy - tr, -(ur[1] * ur[2])])
        else:
            cross = np.array([ur[1] * wr[2] - ur[2] * wr[1], ur[2] * wr[0] - ur[0] * wr[2],
                              ur[0] * wr[1] - ur[1] * wr[0]])
        def les_flux(spec, hat):  # the scalar's LES flux of calc_les! (les_real_calc.jl:120-136)
            grad = [g.irfft(np.where(g.RANGE, 1j * kk * hat, 0)) for kk in (kx, ky, kz)]
            nut = les_models.eddy_nu(self.les_model, Sr, wr, self.les_c, self.les_d2, self.les_smodel)
            cv = None if spec == "eddy" else spec[1] * self.les_d2
            return les_models.scalar_les_flux(nut, grad, wr, cv)
        if self.density is not None:
            fluxr = np.array([(-rhor) * ur[c] for c in range(3)])
            if self.density_les is not None:
                self.dfluxr = les_flux(self.density_les, self.rhoh)  # dflux1..3 (output.jl:22-26)
                fluxr = fluxr + self.dfluxr
        if self.scalar is not None:  # realspacecalculation!: flux = -phi u (models.jl:245-273)
            pfluxr = np.array([(-phir) * ur[c] for c in range(3)])
            if self.scalar_les is not None:
                pfluxr = pfluxr + les_flux(self.scalar_les, self.phih)
        if self.variable_dt:
            self.umax = float(np.max(np.abs(ur)))
            if self.density is not None:
                self.rhomax = float(np.max(np.abs(rhor)))
        if self.equation == "NS":
            t = [g.myfourier(c) for c in uur]
            txx, txy, txz, tyy, tyz = t
            tzz = -txx - tyy
            rhs = np.array([1j * (kx * txx + ky * txy + kz * txz), 1j * (kx * txy + ky * tyy + kz * tyz),
                            1j * (kx * txz + ky * tyz + kz * tzz)])
        else:
            rhs = np.array([g.myfourier(c) for c in cross])
            rhs[:, 0, 0, 0] = 0
        self.uh = g.project(np.array([g.myfourier(c) for c in ur]))
        if self.density is not None:
            fluxh = np.array([g.myfourier(c) for c in fluxr])
            self.rhoh = g.myfourier(rhor)
        if self.scalar is not None:
            pfluxh = np.array([g.myfourier(c) for c in pfluxr])
            self.phih = g.myfourier(phir)
        if self.les_c is not None:
            self.tauh = np.array([g.myfourier(c) for c in taur])
        if self.variable_dt:
            self.set_dt()
        if self.density is not None:
            self.fluxh = fluxh
        if collect and self.is_spec_time():
            # write_spec (models.jl:37): after set_dt!, before fourierspacep2!
            self.spectra.append((self.iteration, self.all_spectra(rhs)))
        if collect and self.is_stats_time():
            self.nlstats = self.non_linear_stats(rhs)
            self.pressstats = self.pressure_stats(rhs)
        # fourierspacep2_velocity! (models.jl:292-341): explicit viscosity unless the integrator is implicit
        if self.integrator in ("AB3", "Euler"):
            if self.hv is not None:  # (K2 mnu + K2^M mnuh) v, models.jl:320-322
                rhs = rhs + (g.K2 * (-self.nu) + self.ipow(g.K2, self.hv["m"]) * (-self.hv["nu"])) * self.uh
            else:
                rhs = rhs + (g.K2 * (-self.nu)) * self.uh
        if self.les_c is not None:
            rhs = rhs + self.div_tau()
        if self.density is not None:
            rhs = rhs.copy()
            rhs[self.gdir] = rhs[self.gdir] + self.rhoh * self.gval
        rhs = np.where(g.RANGE, rhs, 0)
        self.rhs = g.project(rhs)
        if self.scalar is not None:
            # div!(rhs, flux, u_dir, -scalarGradient) (models.jl:353-355); no explicit diffusion (AB3)
            r = 1j * (kx * pfluxh[0] + (ky * pfluxh[1] + kz * pfluxh[2])) + \
                (-self.scalar["gradient"]) * self.uh[self.scalar["dir"]]
            self.phirhs = np.where(g.RANGE, r, 0)
        if self.density is not None:
            # div!(rhorhs, flux, u_z, -drhodz) (vectorfunctions.jl:93-103); explicit diffusion is
            # absent in the reference (models.jl:361), so AB3 density has none
            r = 1j * (kx * fluxh[0] + (ky * fluxh[1] + kz * fluxh[2])) + (-self.drhodz) * self.uh[self.gdir]
            self.rhorhs = np.where(g.RANGE, r, 0)

    def div_tau(self):
        """(i k) . tau with tau symmetric traceless (models.jl:330)."""
        g = self.g
        t = self.tauh
        txx, txy, txz, tyy, tyz = t
        tzz = -txx - tyy
        return np.array([1j * (g.kx * txx + g.ky * txy + g.kz * txz), 1j * (g.kx * txy + g.ky * tyy + g.kz * tyz),
                         1j * (g.kx * txz + g.ky * tyz + g.kz * tzz)])

    def set_dt(self):
        """set_dt! (time_step_functions.jl:86-129) with the intended fresh max|u_i|.

        pinned by: steady inviscid states with known max|u_i| and max|rho| stepped through the CFL
        ramp, AB3's first-step halving, both x1.25 clamps and the density limit
        (tests/test_oracle_pins.py::test_set_dt_*)."""
        g = self.g
        dx = 2 * math.pi * g.LY / g.NY
        cfl = self.cfl
        it = self.iteration
        cfl *= {0: 0.1, 1: 0.2, 2: 0.4, 3: 0.8}.get(it, 1.0)
        newdt = cfl * dx / self.umax
        if self.density is not None:
            newdt = min(newdt, cfl * math.sqrt(dx / (self.rhomax * abs(self.gval))))
        if it > 3:
            dt = self.dt
            newdt = max(dt / 1.25, newdt) if newdt < dt else min(1.25 * dt, newdt)
        if self.integrator == "AB3" and self.tit == 0:
            # Adams_Bashforth3rdO set_dt! (Adams_Bashforth.jl:166-176)
            self.dt3 = self.dt2 = self.dt = newdt / 2
        else:
            # ETD3rdO set_dt! (ETD.jl:110-114) and AB3 after its first step
            self.dt3, self.dt2, self.dt = self.dt2, self.dt, newdt

    def ab3_coefficients(self):
        """get_At/get_Bt/get_Ct, Adams_Bashforth.jl:178-218 (variable-step form)."""
        dt, dt2, dt3 = self.dt, self.dt2, self.dt3
        At = 1.0 + (dt * (2 * dt + 6 * dt2 + 3 * dt3)) / (6 * dt2 * (dt2 + dt3))
        At *= dt
        Bt = -dt * dt * (2 * dt + 3 * dt2 + 3 * dt3) / (6 * dt2 * dt3 * (dt + dt2))
        Bt *= dt + dt2
        Ct = dt * dt * (2 * dt + 3 * dt2) / (6 * dt3 * (dt + dt2 + dt3) * (dt2 + dt3))
        Ct *= dt + dt2 + dt3
        return At, Bt, Ct

    def forcing(self):
        """Constant-power isotropic shell forcing (no reference counterpart):
        factor = P / sum_{0<|k|<=kf} w |u_k|^2."""
        if self.plane is not None:
            self.plane(self)
        if self.forcing_kf is None:
            return
        band = (self.g.K2 <= self.forcing_kf**2) & (self.g.K2 > 0) & self.g.RANGE
        e = np.sum(self.g.W * band * (np.abs(self.uh) ** 2).sum(axis=0))
        self.force_factor = self.forcing_power / e if e > 0 else 0.0

    def time_step(self):
        """VectorTimeStep functor (time_step_functions.jl:72-80) and the density
        integrator (models.jl:471)."""
        g = self.g
        self.tit += 1
        if self.integrator == "AB3":
            u = self.ab3(self.uh, self.rhs, "fm1", "fm2")
        elif self.integrator == "Euler":  # Euller functor (Euller.jl:14-25): u = muladd(dt, rhs, u)
            u = self.dt * self.rhs + self.uh
        else:
            u = self.etd(self.uh, self.rhs, "fm1", "fm2", -self.nu)
        u = np.where(g.RANGE, u, 0)
        if self.forcing_kf is not None:
            band = (g.K2 <= self.forcing_kf**2) & (g.K2 > 0) & g.RANGE
            u = u + np.where(band, self.dt * self.force_factor, 0.0) * u
        if self.plane is not None:
            u = self.plane.add(np.array(u))
        self.uh = g.project(u)
        if self.scalar is not None:  # the passive scalar's integrator (models.jl:470)
            if self.integrator == "AB3":
                r = self.ab3(self.phih, self.phirhs, "pfm1", "pfm2")
            elif self.integrator == "Euler":
                r = self.dt * self.phirhs + self.phih
            else:
                r = self.etd(self.phih, self.phirhs, "pfm1", "pfm2", -self.palpha)
            self.phih = np.where(g.RANGE, r, 0)
        if self.density is not None:
            if self.integrator == "AB3":
                r = self.ab3(self.rhoh, self.rhorhs, "rfm1", "rfm2")
            elif self.integrator == "Euler":
                r = self.dt * self.rhorhs + self.rhoh
            else:
                r = self.etd(self.rhoh, self.rhorhs, "rfm1", "rfm2", -self.alpha, self.density.get("hv"))
            self.rhoh = np.where(g.RANGE, r, 0)

    def ab3(self, u, rhs, n1, n2):
        """_tAdams_Bashforth3rdO! (Adams_Bashforth.jl:231-256) with fix_fm2! at step 2."""
        fm1, fm2 = getattr(self, n1), getattr(self, n2)
        if self.tit == 2:
            fm2 = -(self.dt3 + self.dt2) * (rhs - fm1) / self.dt2 + rhs
        At, Bt, Ct = self.ab3_coefficients()
        out = At * rhs + (Bt * fm1 + (Ct * fm2 + u))
        setattr(self, n2, fm1)
        setattr(self, n1, rhs)
        return out

    def etd_coefficients(self, c):
        """set_ABCt! (ETD.jl:142-196) with get_ats/bts/cts (:198-227)."""
        dt, dt2, dt3 = self.dt, self.dt2, self.dt3
        a0 = (dt * (2 * dt * dt + 6 * dt2 * (dt2 + dt3) + 3 * dt * (2 * dt2 + dt3))) / (6 * dt2 * (dt2 + dt3))
        a1 = (dt * dt * (dt**2 + 6 * dt2 * (dt2 + dt3) + 2 * dt * (2 * dt2 + dt3))) / (12 * dt2 * (dt2 + dt3))
        a2 = (dt * dt * dt * (4 * dt2 * (dt2 + dt3) + dt * (2 * dt2 + dt3))) / (24 * dt2 * (dt2 + dt3))
        a3 = dt * dt * dt * dt / 24
        b0 = -((dt**2 * (2 * dt + 3 * (dt2 + dt3))) / (6 * (dt2 * dt3)))
        b1 = -(((dt * dt * dt * (dt + 2 * (dt2 + dt3)))) / (12 * (dt2 * dt3)))
        b2 = -((dt * dt * dt * dt * (dt2 + dt3)) / (24 * (dt2 * dt3)))
        c0 = (dt * dt * (2 * dt + 3 * dt2)) / (6 * dt3 * (dt2 + dt3))
        c1 = (dt * dt * dt * (dt + 2 * dt2)) / (12 * dt3 * (dt2 + dt3))
        c2 = (dt * dt * dt * dt * dt2) / (24 * dt3 * (dt2 + dt3))
        e1 = dt**2 + dt * (2 * dt2 + dt3)
        e2 = dt2 * (dt2 + dt3)
        e3 = 2 * dt
        e4 = 2 * dt2 + dt3
        e5 = dt * (dt + dt2 + dt3)
        e6 = dt2 + dt3
        e7 = dt2 * dt3
        e8 = dt * (dt + dt2)
        e9 = dt3 * (dt2 + dt3)
        l1 = c
        l2 = l1 * l1
        l3 = l2 * l1
        ldt = l1 * dt
        test = -l1 * dt <= 1e-4
        em = np.expm1(ldt)
        with np.errstate(all="ignore"):
            At = np.where(test, a1 * l1 + (a2 * l2 + (a3 * l3 + a0)),
                          (2 * em - l2 * (e1 - e2 * em) - l1 * (e3 - e4 * em)) / (l3 * e2))
            Bt = np.where(test, b1 * l1 + (b2 * l2 + b0), (e5 * l2 - 2 * em + l1 * (e3 - e6 * em)) / (l3 * e7))
            Ct = np.where(test, c1 * l1 + (c2 * l2 + c0), (2 * em - e8 * l2 - l1 * (e3 - dt2 * em)) / (l3 * e9))
        inf = l1 == -np.inf
        return np.where(inf, 0.0, At), np.where(inf, 0.0, Bt), np.where(inf, 0.0, Ct)

    def modk(self):
        """muladd(KX, KX, muladd(KY, KY, KZ^2)) of init_c! and the hyperviscosity loops (unfused here
        and on the device, so both round alike)."""
        g = self.g
        return g.kx * g.kx + (g.ky * g.ky + g.kz * g.kz)

    @staticmethod
    def ipow(x, m):
        """x^m as m - 1 products (as the device forms k^2M)."""
        r = x
        for _ in range(1, m):
            r = r * x
        return r

    def barrier(self, k, hv=None):
        """SpectralBarrier's f(k) = log(esp(k, initk, endk, 1/4)) (hyperviscosity_types.jl:44-48)."""
        hv = self.hv if hv is None else hv
        i, p = hv["initk"], hv["endk"]
        ak = np.abs(k)
        with np.errstate(all="ignore"):
            e = np.where(ak <= i, 1.0, np.where(ak < p, np.exp(-((ak - i) ** 2) / (p ** 0.25 - ak ** 0.25)), 0.0))
            return np.log(e)

    def etd_c(self, mnu, hv="velocity"):
        """init_c! (ETD.jl:56-95): c = -nu k^2, plus -nu_h k^2M (HyperViscosity) or f(|k|)
        (SpectralBarrier, -Inf beyond endk).  hv: the velocity's model (also the passive scalar's,
        simulationtypes.jl:331) or a field's own (the density's densityHyperViscosity,
        density_stratification_types.jl:137-149)."""
        g = self.g
        hv = self.hv if hv == "velocity" else hv
        if hv is None:
            c = g.K2 * mnu
        elif hv["kind"] == "hyper":
            mk = self.modk()
            c = mk * mnu + self.ipow(mk, hv["m"]) * (-hv["nu"])
        else:
            mk = self.modk()
            c = mk * mnu + self.barrier(np.sqrt(mk), hv)
        return np.where(g.RANGE, c, 0.0)

    def hv_weight(self, hv=None):
        """Per-mode factor of the hyperviscous dissipation: nu_h k^2M (hyperviscosity_stats,
        stats.jl:532-561) or -f(|k|) with f = -Inf read as 0 (spectral_barrier_stats,
        hyperviscosity_types.jl:50-79)."""
        hv = self.hv if hv is None else hv
        mk = self.modk()
        if hv["kind"] == "hyper":
            return hv["nu"] * self.ipow(mk, hv["m"])
        f = self.barrier(np.sqrt(mk), hv)
        return -np.where(f == -np.inf, 0.0, f)

    def etd(self, u, rhs, n1, n2, mnu, hv="velocity"):
        """ETD3rdO functor (ETD.jl:229-247) with ETD1stO!/ETD2ndO! (ETD1stand2ndO.jl)
        for the first two steps; c from init_c! (ETD.jl:56-95); c = -Inf gives zero
        coefficients (the test2 branches)."""
        g = self.g
        c = self.etd_c(mnu, hv)
        inf = c == -np.inf
        dt = self.dt
        fm1, fm2 = getattr(self, n1), getattr(self, n2)
        with np.errstate(all="ignore"):
            if self.tit == 1:
                cdt = c * dt
                cdt2 = cdt * dt
                c2dt3 = cdt2 * cdt
                c3dt4 = c2dt3 * cdt
                test = -cdt <= 1e-4
                f = np.where(inf, 0.0, np.where(test, dt + 0.5 * cdt2 + c2dt3 / 6 + c3dt4 / 24, np.expm1(cdt) / c))
                out = np.exp(cdt) * u + f * rhs
            elif self.tit == 2:
                dt2 = self.dt2
                cp2 = c * c
                cdt = c * dt
                cdtp2 = cdt * dt
                c2dtp3 = cdtp2 * cdt
                c3dtp4 = c2dtp3 * cdt
                em = np.expm1(cdt)
                c2dt2 = c * c * dt2
                dtp2 = dt * dt
                dtp3 = dtp2 * dt
                dtp4 = dtp3 * dt
                test = -cdt <= 1e-4
                At = np.where(test, dt + cdtp2 / 2 + c2dtp3 / 6 + c3dtp4 / 24 + dtp2 / (2 * dt2) + (c * dtp3) / (6 * dt2)
                              + (cp2 * dtp4) / (24 * dt2), (em - c * (dt - dt2 * em)) / c2dt2)
                Bt = np.where(test, (-(dtp2 / (2 * dt2)) - (c * dtp3) / (6 * dt2) - (cp2 * dtp4) / (24 * dt2)),
                              (cdt - em) / c2dt2)
                At = np.where(inf, 0.0, At)
                Bt = np.where(inf, 0.0, Bt)
                out = np.exp(cdt) * u + (At * rhs + Bt * fm1)
            else:
                At, Bt, Ct = self.etd_coefficients(c)
                out = At * rhs + (Bt * fm1 + (Ct * fm2 + np.exp(c * dt) * u))
        setattr(self, n2, fm1)
        setattr(self, n1, rhs)
        return out

    def advance_in_time(self, n=1):
        for _ in range(n):
            self.calculate_rhs()
            self.forcing()
            if self.plane is not None and self.is_spec_time():  # write_spec_forcing (models.jl:8)
                self.spectra[-1][1]["forcing_h"] = self.forcing_spectrum()
            if self.is_stats_time():
                self.records.append(self.stats())
            self.time_step()
            self.iteration += 1
            self.time += self.dt

    # -------------------------------------------------------------- stats
    def non_linear_stats(self, nl):
        """non_linear_stats (stats.jl:897-917)."""
        g = self.g
        kv = (g.kx, g.ky, g.kz)
        out = [proj(nl[c] - 1j * kv[c] * self.kspec, self.uh[c]) for c in range(3)]
        nlh = np.sum(np.where(g.RANGE, g.W * (out[0] + out[1]), 0))
        nlv = np.sum(np.where(g.RANGE, g.W * out[2], 0))
        return nlh, nlv

    def pressure_stats(self, nl):
        """pressure_stats (stats.jl:750-814)."""
        g = self.g
        kv = (g.kx, g.ky, g.kz)
        r = [nl[c] - (1j * self.kspec) * kv[c] for c in range(3)]
        if self.les_c is not None:
            dt_ = self.div_tau()
            r = [r[c] + dt_[c] for c in range(3)]
        if self.density is not None:
            r[self.gdir] = r[self.gdir] + self.rhoh * self.gval
        with np.errstate(invalid="ignore", divide="ignore"):
            p1 = -(g.kx * r[0] + g.ky * r[1] + g.kz * r[2]) / g.K2
        p1[0, 0, 0] = 0
        out = [proj(self.uh[c], p1 * kv[c]) for c in range(3)]
        ph = np.sum(np.where(g.RANGE, g.W * (out[0] + out[1]), 0))
        pv = np.sum(np.where(g.RANGE, g.W * out[2], 0))
        return ph, pv

    def stats(self):
        """stats(s) (stats.jl:64-165): returns (Stats row, Scales row, Numbers row).

        pinned by: Taylor-Green t = 0 closed forms and the physical-space means of an asymmetric
        analytic field on a non-cubic box for every Scales and Numbers entry
        (tests/test_oracle_pins.py::test_scales_numbers_*); the Stats columns by the tests of
        tests/test_oracle.py and tests/test_spectra.py."""
        g = self.g
        u = self.uh
        nu = self.nu
        W = np.where(g.RANGE, g.W, 0.0)
        e = [np.abs(c) ** 2 for c in u]
        u12, u22, u32 = (float(np.sum(W * c)) for c in e)
        u1, u2, u3 = (float(c[0, 0, 0].real) for c in u)
        kh = (u12 + u22) / 2
        kv = u32 / 2
        k = kh + kv
        with np.errstate(all="ignore"):
            kaniso = np.float64(2 * kv) / np.float64(kh)
        eh = nu * float(np.sum(W * g.K2 * (e[0] + e[1])))
        ev = nu * float(np.sum(W * g.K2 * e[2]))
        eps = eh + ev
        nlh, nlv = self.nlstats
        pressh, pressv = self.pressstats
        vstats = [u1, u2, u3, u12, u22, u32, kh, kv, kaniso, k, eh, ev, eps, nlh, nlv, pressh, pressv,
                  -(nlh + pressh), (nlv + pressv)]
        tdiss = eps
        R = g.RANGE

        def lx(c, p2):
            return g.LX * np.float64(np.sum(np.where(R[0], e[c][0], 0))) * math.pi / p2

        def ly(c, p2):
            return g.LY * np.float64(np.sum(np.where(R[:, 0, :], (g.W[:, 0, :] * e[c][:, 0, :]), 0))) * math.pi / p2

        def lz(c, p2):
            return g.LZ * np.float64(np.sum(np.where(R[:, :, 0], (g.W[:, :, 0] * e[c][:, :, 0]), 0))) * math.pi / p2

        def dsq(c, kk):
            return float(np.sum(W * kk * kk * e[c]))

        with np.errstate(all="ignore"):  # Julia float semantics: x/0 -> Inf, 0/0 -> NaN
            u12, u22, u32, k, tdiss, nu = (np.float64(v) for v in (u12, u22, u32, k, tdiss, nu))
            L11, L12, L13 = lx(0, u12), ly(0, u12), lz(0, u12)
            du1dyp2, du1dxp2 = dsq(0, g.ky), dsq(0, g.kx)
            l12 = np.sqrt(2 * u12 / du1dyp2)
            L21, L22, L23 = lx(1, u22), ly(1, u22), lz(1, u22)
            du2dyp2, du2dxp2 = dsq(1, g.ky), dsq(1, g.kx)
            l21 = np.sqrt(2 * u22 / du2dxp2)
            L31, L32, L33 = lx(2, u32), ly(2, u32), lz(2, u32)
            uh = np.sqrt((u12 + u22) / 2)
            urms = np.sqrt((u12 + u22 + u32) / 3)
            Lhh = (L11 + L22) / 2
            Thh = Lhh / urms
            ll = np.sqrt(2 * (u12 + u22) / (du1dxp2 + du2dyp2))
            lt = np.sqrt(2 * (u12 + u22) / (du1dyp2 + du2dxp2))
            Lhv = (L13 + L23) / 2
            Lvv = L33
            Lvh = (L31 + L32) / 2
        other = []
        if self.scalar is not None:
            # scalar_stats (stats.jl:221-248), before the density's (sim_fields, macros.jl:48)
            ep = np.abs(self.phih) ** 2
            sd = [float(np.sum(W * kk * kk * ep)) for kk in (g.kx, g.ky, g.kz)]
            other += [float(self.phih[0, 0, 0].real), float(np.sum(W * ep)), *sd,
                      self.palpha * (sd[0] + sd[1] + sd[2])]
        hvstats = []
        if self.hv is not None:  # hyperviscosity_stats / spectral_barrier_stats: hdissh, hdissv, hdiss
            # stats.jl:118: tdiss takes the hyperviscous part before the LES part
            wgt = np.where(g.RANGE, g.W * self.hv_weight(), 0.0)
            hh = float(np.sum(wgt * (e[0] + e[1])))
            hv_ = float(np.sum(wgt * e[2]))
            hvstats = [hh, hv_, hh + hv_]
            tdiss = tdiss + (hh + hv_)
        if self.density is not None:
            # stats(::AbstractDensityStratification) (density_stratification_types.jl:46-87)
            r = self.rhoh
            er = np.abs(r) ** 2
            rho = float(r[0, 0, 0].real)
            rho2 = float(np.sum(W * er))
            drd = [float(np.sum(W * kk * kk * er)) for kk in (g.kx, g.ky, g.kz)]
            rhodiss = self.alpha * (drd[0] + drd[1] + drd[2])
            bf = -self.gval * float(np.sum(W * proj(u[self.gdir], r)))  # buoyancy_flux (:192-211)
            trhodiss = rhodiss
            extra = []
            if self.density_les is not None:  # scalar_les_stats (stats.jl:1032-1053) of the total flux
                f = self.fluxh
                dv = 1j * (g.kx * f[0] + g.ky * f[1] + g.kz * f[2])
                rhopr = -float(np.sum(np.where(g.RANGE, W * proj(dv, r), 0.0)))
                trhodiss += rhopr
                extra.append(rhopr)
            if self.density.get("hv") is not None:  # scalar_hvis_stats (stats.jl:1086-1170): rhohdiss
                wgt = np.where(g.RANGE, g.W * self.hv_weight(self.density["hv"]), 0.0)
                rhohdiss = float(np.sum(wgt * np.abs(self.rhoh) ** 2))
                trhodiss += rhohdiss
                extra.append(rhohdiss)
            other += [rho, rho2, *drd, rhodiss, trhodiss, bf, *extra]
        if self.les_c is not None:
            # les_stats (stats.jl:679-705)
            d = self.div_tau()
            prh = -float(np.sum(W * (proj(d[0], u[0]) + proj(d[1], u[1]))))
            prv = -float(np.sum(W * proj(d[2], u[2])))
            with np.errstate(all="ignore"):
                other += [prh, prv, np.float64(2 * prv) / np.float64(prh), prh + prv]
            tdiss = tdiss + (prh + prv)
        with np.errstate(all="ignore"):
            tdiss = np.float64(tdiss)
            Lt = np.sqrt(k**3) / tdiss
            Tt = k / tdiss
            Lk = np.sqrt(np.sqrt(nu**3 / tdiss))
            Tk = np.sqrt(nu / tdiss)
            lg = np.sqrt(15 * nu / tdiss) * urms
        if self.plane is not None:  # stats(::RfForcing ...): hp, vp, tp; stats(::AForcing): hp, vp
            other += [self.plane.hp, self.plane.vp]
            if self.plane.kind != "a":
                other.append(self.plane.hp + self.plane.vp)
        other += hvstats
        stats_row = [self.iteration, self.time, *vstats, tdiss, *other]
        scales = [self.iteration, self.time, Lk, Lt, lg, ll, lt, Lhh, Lvv, Lhv, Lvh, l12, l21, L11, L12, L13, L21,
                  L22, L23, L31, L32, L33]
        with np.errstate(all="ignore"):
            Reh, Reh1 = uh * Lhh / nu, (Lhh / Lk) ** (4 / 3)
            Rel, Rel1 = urms * ll / nu, (ll / Lk) ** (4 / 3)
            Ret, Ret1 = urms * lt / nu, (lt / Lk) ** (4 / 3)
            if self.density is not None:
                Nb = np.sqrt(np.float64(self.gz * self.drhodz))  # calculate_N
                Lb = 2 * np.pi * urms / Nb
                Lo = np.sqrt(tdiss / (Nb * Nb * Nb))
                Lom = np.sqrt(tdiss / (Nb * Nb * Nb / (8 * np.pi * np.pi * np.pi)))
                Tb = 2 * np.pi / Nb
                Tbm = 1 / Nb
                Rlg, Rlg1 = urms * lg / nu, (lg / Lk) ** (4 / 3)
                scales += [Lb, Lo, Lom, Tk, Tt, Thh, Tb, Tbm]
                numbers = [self.iteration, self.time, Reh, Reh1, Rlg, Rel, Ret, Rlg1, Rel1, Ret1, (Lo / Lk) ** (4 / 3),
                           (Lom / Lk) ** (4 / 3), 2 * np.pi * uh / (Nb * Lhh), Lhv / Lhh]
            else:
                scales += [Tk, Tt, Thh]
                numbers = [self.iteration, self.time, Reh, Reh1, Rel, Ret, Rel1, Ret1]
        return np.array(stats_row), np.array(scales), np.array(numbers)

    def spec_hv(self, arr):
        """calculate_specHV (spec.jl:424-450), loop for loop: (specH over KH, specV over KRZ)."""
        g = self.g
        h = np.zeros(len(g.KH))
        v = np.zeros(len(g.KRZ))
        nx = np.rint(np.abs(g.KX[1:]) / g.MAXDKH).astype(int)
        for l in g.ZRANGE:
            nk = int(round(abs(g.KZ[l]) / g.DKZ))
            for j in g.YRANGE:
                nj = int(round(abs(g.KY[j]) / g.MAXDKH))
                ee = arr[0, j, l]
                h[0] += 0.5 * ee
                h[nj] += 0.5 * ee
                v[nk] += ee
                ee2 = 2 * arr[1:, j, l]
                np.add.at(h, nx, 0.5 * ee2 * g.MAXDKH)
                h[nj] += float(np.sum(0.5 * ee2 * g.MAXDKH))
                v[nk] += float(np.sum(ee2 * g.DKZ))
        return h, v

    def spectrum(self):
        """spectrum_u + calculate_specHV (spec.jl:316-327,424-450)."""
        g = self.g
        u = self.uh
        hs = np.where(g.RANGE, (proj(u[0], u[0]) + proj(u[1], u[1])) / 2, 0.0)
        vs = np.where(g.RANGE, proj(u[2], u[2]) / 2, 0.0)
        return [self.spec_hv(hs), self.spec_hv(vs)]

    def all_spectra(self, nl):
        """The per-mode arrays of write_spec (spec.jl:1-168) binned by calculate_specHV:
        {quantity: (specH, specV)} for u, vis, nl, press (h and v parts), les, buoyancy_v, rho,
        rhodiss, rhonl.  nl is the nonlinear term F(u x w) with its k = 0 mode zero."""
        g = self.g
        u = self.uh
        kv = (g.kx, g.ky, g.kz)
        R = g.RANGE
        out = {}

        def put(name, arr):
            out[name] = self.spec_hv(np.where(R, arr, 0.0))

        # spectrum_u (:316-327)
        hu = (proj(u[0], u[0]) + proj(u[1], u[1])) / 2
        vu = proj(u[2], u[2]) / 2
        put("u_h", hu)
        put("u_v", vu)
        # spectrum_viscosity (:340-354): 2 nu k^2 times the u spectrum, k^2 = muladd(kx, kx, muladd(ky, ky, kz^2))
        k2 = g.kx * g.kx + (g.ky * g.ky + g.kz * g.kz)
        put("vis_h", 2 * self.nu * k2 * hu)
        put("vis_v", 2 * self.nu * k2 * vu)
        # spectrum_non_linear (:210-222): vecouterproj(nl - (i k) ke, u)
        o = [proj(nl[c] - 1j * kv[c] * self.kspec, u[c]) for c in range(3)]
        put("nl_h", o[0] + o[1])
        put("nl_v", o[2])
        # spectrum_pressure (:370-422)
        r = [nl[c] - (1j * self.kspec) * kv[c] for c in range(3)]
        if self.les_c is not None:
            d = self.div_tau()
            r = [r[c] + d[c] for c in range(3)]
        if self.density is not None:
            r[self.gdir] = r[self.gdir] + self.rhoh * self.gval
        with np.errstate(invalid="ignore", divide="ignore"):
            p1 = -(g.kx * r[0] + g.ky * r[1] + g.kz * r[2]) / g.K2
        p1[0, 0, 0] = 0
        o = [proj(u[c], p1 * kv[c]) for c in range(3)]
        put("press_h", o[0] + o[1])
        put("press_v", o[2])
        if self.les_c is not None:  # spectrum_les (:224-235): vecouterproj((i k).tau, u)
            d = self.div_tau()
            o = [proj(d[c], u[c]) for c in range(3)]
            put("les_h", o[0] + o[1])
            put("les_v", o[2])
        if self.hv is not None:
            # spectrum_hyperviscosity (:297-314): nu_h k^2M vecouterproj(u, u); spectrum_spectral_barrier
            # (:270-295): f(|k|) vecouterproj(u, u), f = -Inf read as 0
            if self.hv["kind"] == "hyper":
                fk = self.hv["nu"] * self.ipow(self.modk(), self.hv["m"])
            else:
                fk = -self.hv_weight()
            put("hvis_h", fk * (proj(u[0], u[0]) + proj(u[1], u[1])))
            put("hvis_v", fk * proj(u[2], u[2]))
        if self.density is not None:
            rho = self.rhoh
            put("buoyancy_v", proj(u[2], self.gz * rho))                     # spectrum_buoyancy (:249-257)
            put("rho", proj(rho, rho))                                        # spectrum_rho (:329-338)
            put("rhodiss", self.alpha * k2 * proj(rho, rho))                  # spectrum_rhodiss (:356-368)
            div = 1j * (g.kx * self.fluxh[0] + g.ky * self.fluxh[1] + g.kz * self.fluxh[2])
            put("rhonl", proj(div, rho))                                      # spectrum_rhonl (:237-247)
        return out

    def forcing_spectrum(self):
        """spectrum_forcing (spec.jl:259-268) of the plane forcing: (proj(ux, fx) + proj(uy, fy)) / dt."""
        u = self.uh
        f = self.plane
        arr = (proj(u[0], f.forcex) + proj(u[1], f.forcey)) / self.dt
        return self.spec_hv(np.where(self.g.RANGE, arr, 0.0))
