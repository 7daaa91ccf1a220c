/*
 * ffs.h -- C ABI of libffs_hip.so, the MI355X (gfx950) time-step engine that
 * stands in for the hot path of favba/FluidFlowSimulation.jl.
 *
 * Every entry point takes plain pointers and sizes; no C++ or torch types
 * cross this boundary.  All functions return 0 on success or a negative
 * FFS_ERR_* code; the message of the last failure of a handle is available
 * from ffs_last_error(handle) (or ffs_last_error(NULL) for ffs_create).
 * Nothing aborts or throws across the ABI.
 *
 * Threading: a handle is not re-entrant and owns one HIP stream; distinct
 * handles may be driven from distinct host threads.  The library never
 * retains a host pointer beyond the call that received it.
 *
 * Each declaration names the reference function it replaces (paths relative
 * to the reference's src/ directory).
 */
#ifndef FFS_FFS_H
#define FFS_FFS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FFS_ABI_VERSION 4

/* error codes */
#define FFS_OK 0
#define FFS_ERR_ARG -1        /* bad argument / unsupported configuration */
#define FFS_ERR_MEMORY -2     /* device allocation failed or budget exceeded */
#define FFS_ERR_IO -3         /* file or size mismatch */
#define FFS_ERR_NONFINITE -4  /* the state contains NaN/Inf */
#define FFS_ERR_HIP -5        /* HIP runtime error */
#define FFS_ERR_STATE -6      /* call out of order (e.g. advance before initialize) */

/* integrators: `velocityTimeStep` key (simulationtypes.jl:308) */
#define FFS_AB3 0  /* Adams_Bashforth3rdO */
#define FFS_ETD3 1 /* ETD3rdO (reference default) */
#define FFS_EULER 2 /* Euller (time_step_functions/Euller.jl): u += dt rhs, explicit viscosity */

/* precision of the bulk arithmetic */
#define FFS_FP64 0
#define FFS_FP32 1 /* FP32 fields/FFTs; FP64 coefficients and reductions */

/* LES models (`lesModel` key, LESmodels/les_types.jl:360-404).  The pointwise
 * models of calc_les! (les_real_calc.jl:74-114) and the dynamic ones
 * (dynamic_smagorinsky.jl, dynamic_piomelli.jl).  A standalone fakeSmagorinsky is
 * not built and is refused by ffs_create. */
#define FFS_LES_NONE 0
#define FFS_LES_SMAGORINSKY 1
#define FFS_LES_VREMAN 2           /* VremanLESModel.jl */
#define FFS_LES_SILVIS 3           /* silvismodel.jl */
#define FFS_LES_PRODUCTION 4       /* productionViscosity, ProductionViscosityLESModel.jl */
#define FFS_LES_STABLE_NL 5        /* stableNonLinear, stablenonlinearmodel.jl */
#define FFS_LES_FAKE_SMAGORINSKY 6 /* fakeSmagorinsky: only as the S model of SandP */
#define FFS_LES_SANDP 7            /* SandP: an S model (les_s_model) + cb Delta^2 Lie(S, W) */
#define FFS_LES_DYNAMIC_SMAGORINSKY 8 /* dynamicSmagorinsky: Gaussian test filter, averaged coefficient */
#define FFS_LES_PIOMELLI_SMAGORINSKY 9 /* piomelliSmagorinsky: the coefficient extrapolated in time */
#define FFS_LES_DYNAMIC_SANDP 10  /* dynamicSandP: a dynamic S model (les_s_model) + a dynamic P part */
#define FFS_LES_PIOMELLI_SANDP 11 /* piomelliSandP: the same with the P coefficient extrapolated in time */

/* forcing (`forcing` key); FFS_FORCING_SHELL is the constant-power isotropic
 * low-wavenumber forcing f = (P / sum_{0<|k|<=kf} |u_k|^2) u_k, which has no
 * counterpart in the reference's kz=0-plane schemes */
#define FFS_FORCING_NONE 0
#define FFS_FORCING_SHELL 1
/* the reference's kz = 0-plane schemes (Forcing_methods/forcing_types.jl): the 2D shells of the
 * horizontal velocity on the kz = 0 plane driven toward a target spectrum Em by a damped
 * oscillator R per shell (rfForcing, forcing.jl:6-114), or with a fixed power spectrum
 * (QfForcing, qfforcing.jl), plus 1 % of the power on kx = ky = 0, |kz| <= 5 */
#define FFS_FORCING_RF 2   /* rfForcing   (forcing.jl) */
#define FFS_FORCING_QF 3   /* QfForcing   (qfforcing.jl) */
#define FFS_FORCING_CRF 4  /* cRfForcing  (cforcing.jl: ts = TF) */
#define FFS_FORCING_NRF 5  /* nRfForcing  (nrforcing.jl: with the nonlinear transfer spectrum N) */
#define FFS_FORCING_CNRF 6 /* cNRfForcing (cnrforcing.jl) */
#define FFS_FORCING_A 7    /* aForcing    (aforcing.jl: one oscillator per kz = 0 mode, |k_h| <= kf) */

/* equation: the nonlinear term's form (simulationtypes.jl:108-116, 265-273) */
#define FFS_EQ_ROTATIONAL 0 /* VorticityEquation: u x w (the default) */
#define FFS_EQ_NS 1         /* NavierStokesEquation ("NS"): -div(u u), u u traceless (models.jl:176-180, 262) */

/* hyperviscosity (Hyperviscosity/hyperviscosity_types.jl) */
#define FFS_HV_NONE 0
#define FFS_HV_HYPER 1   /* HyperViscosity{nu_h, M}: -nu_h k^2M beside -nu k^2 (AB3 rhs, ETD c) */
#define FFS_HV_BARRIER 2 /* SpectralBarrier: ETD c += log(esp(|k|, initk, endk, 1/4)); ETD3rdO only */

/* fields */
#define FFS_FIELD_U1 0
#define FFS_FIELD_U2 1
#define FFS_FIELD_U3 2
#define FFS_FIELD_RHO 3
#define FFS_FIELD_SCALAR 4 /* the passive scalar (scalar.N files, output.jl:10-15) */
/* the Smagorinsky stress xx xy xz yy yz (the T11 T12 T13 T22 T23 files, output.jl:28-35; get only):
 * FFS_SPACE_REAL gives the real-space product 2 (c Delta)^2 |S| S of the current velocity as
 * realspace! forms it, FFS_SPACE_FOURIER the stored transform of the last nonlinear pass,
 * FFS_SPACE_FOURIER_TO_REAL that transform back in real space (real!(tau), FluidFlowSimulation.jl:53-57) */
#define FFS_FIELD_T11 5
#define FFS_FIELD_T12 6
#define FFS_FIELD_T13 7
#define FFS_FIELD_T22 8
#define FFS_FIELD_T23 9
/* real space only, formed from the current state as the next step forms them (writeoutput,
 * output.jl:22-40): the density's LES flux (densityLesModel) and the dynamic models' coefficients */
#define FFS_FIELD_DFLUX1 10
#define FFS_FIELD_DFLUX2 11
#define FFS_FIELD_DFLUX3 12
#define FFS_FIELD_CS 13 /* the dynamic models' averaged coefficient c */
#define FFS_FIELD_CP 14 /* dynamicSandP / piomelliSandP: the P coefficient */

/* spaces for get/set_field */
#define FFS_SPACE_REAL 0    /* padded real (2*(nx/2+1), ny, nz), column major; or unpadded (nx,ny,nz) on input */
#define FFS_SPACE_FOURIER 1 /* half spectrum (nx/2+1, ny, nz) complex, interleaved re/im */
#define FFS_SPACE_FOURIER_TO_REAL 2 /* get only: the stored transform back in real space (= REAL but for the stress) */

/* scalars for ffs_get_scalar */
#define FFS_SCALAR_ITERATION 0
#define FFS_SCALAR_TIME 1
#define FFS_SCALAR_DT 2
#define FFS_SCALAR_UMAX 3

/* stats tables (stats.jl:1-62) */
#define FFS_TABLE_STATS 0   /* Stats.txt   */
#define FFS_TABLE_SCALES 1  /* Scales.txt  */
#define FFS_TABLE_NUMBERS 2 /* Numbers.txt */

/* events returned through ffs_advance's *event */
#define FFS_EVENT_NONE 0
#define FFS_EVENT_OUTPUT 1 /* an output step (writeoutput, models.jl:158) is due at the current iteration */

/* Parameters: the `global` file keys of simulationtypes.jl:242-365 and
 * Globals.jl:106-182 as plain fields.  `size` must be sizeof(ffs_params). */
typedef struct ffs_params {
  uint32_t size;
  int32_t nx, ny, nz;                       /* nx ny nz */
  double x_domain, y_domain, z_domain;      /* xDomainSize ...: the box is 2 pi L per side */
  double nu;                                /* kinematicViscosity */
  double cutoff;                            /* cutoff (2/3): box dealias mask, Globals.jl:132-142 */
  int32_t integrator;                       /* FFS_AB3 / FFS_ETD3 */
  int32_t variable_dt;                      /* variableTimestepFlag */
  double cfl;                               /* cfl (0.18) */
  double dt;                                /* dt (initial / fixed step) */
  int64_t start;                            /* start: initial iteration number */
  double start_time;                        /* startTime */
  int32_t dt_stat, dt_spec, write_time;     /* dtStat, dtSpec, writeTime */
  int32_t precision;                        /* FFS_FP64 / FFS_FP32 */
  int32_t les_model;                        /* FFS_LES_* */
  double smagorinsky_c;                     /* the model's constant: smagorinskyConstant (0.17),
                                               vremanConstant (0.032), silvisConstant (sqrt 0.17),
                                               productionConstant (0.4; 0.1 under SandP) or
                                               nlConstant (0.1); SandP: its S model's.  NaN (the
                                               ffs_params_default value) selects that default */
  double filter_width;                      /* filterWidth; <= 0 selects (2 pi L/nx)/cutoff
                                               (twice that for SandP's Smagorinsky) */
  int32_t density;                          /* densityStratification present */
  double pr, density_gradient, z_acceleration, reference_density; /* Pr, densityGradient, zAcceleration, referenceDensity */
  int32_t forcing;                          /* FFS_FORCING_* */
  double forcing_kf;                        /* forced band |k| <= kf (`kf` of rf/Qf forcing) */
  double forcing_power;                     /* injected power P (shell forcing) */
  double forcing_tf;                        /* TF (rfForcing) */
  double forcing_alpha;                     /* alphac (rfForcing) */
  int32_t forcing_zf_tanh;                  /* Zf tanh (1) or cutoff (0, default) */
  double alvelius_p, alvelius_c, alvelius_kf; /* alveliusP, alveliusC, alveliusKf (QfForcing) */
  int32_t hyperviscosity;                   /* FFS_HV_*: hyperViscosity (simulationtypes.jl:293-305) */
  double hv_nu;                             /* hyperViscosityCoefficient nu_h */
  int32_t hv_m;                             /* hyperViscosityM (2) */
  double sb_initk, sb_endk;                 /* initk, endk of spectralBarrier; <= 0 selects the defaults
                                               cutoff 2 KX[end] / 3 and sqrt(3 cutoff) KX[end] */
  int32_t passive_scalar;                   /* passiveScalar present (simulationtypes.jl:320-352) */
  double scalar_pr;                         /* scalarPr: diffusivity nu / scalarPr (Inf: none) */
  double scalar_gradient;                   /* scalarGradient: the mean gradient's source -G u_dir */
  int32_t scalar_direction;                 /* scalarDirection: 0 x, 1 y, 2 z (default z) */
  int32_t equation;                         /* FFS_EQ_*: equation (simulationtypes.jl:265-273) */
  double les_cp;                            /* silvisPConstant (5.0) or SandP's pConstant cb (0.09);
                                               NaN selects the default */
  int32_t les_s_model;                      /* sLesModel: SandP's SMAGORINSKY, VREMAN, PRODUCTION or
                                               FAKE_SMAGORINSKY; dynamicSandP's and piomelliSandP's
                                               DYNAMIC_SMAGORINSKY or PIOMELLI_SMAGORINSKY */
  double les_test_width;                    /* TestFilterWidth of the dynamic models; <= 0 selects
                                               twice the filter width */
  double les_cmin;                          /* cmin of the dynamic models (0): the coefficient's floor */
  /* the density's own hyperviscosity (densityHyperViscosity, density_stratification_types.jl:137-149;
   * the velocity's hyperViscosity does not reach the density, while the passive scalar shares it):
   * FFS_HV_*, densityHyperViscosityCoefficient, densityHyperViscosityM (2), densityInitk,
   * densityEndk (<= 0: the defaults of initk / endk).  It acts through the density's ETD3 operator
   * and adds the rhohdiss stats column. */
  /* the scalars' LES fluxes (scalarLesModel / densityLesModel, les_real_calc.jl:120-136): 0 none,
   * 1 EddyDiffusion (nu_t grad s), 2 VorticityDiffusion (+ (c Delta^2 / 2) w x grad s).  With an
   * LES model the passive scalar always has one (0 reads as 1, as in the reference); the density's
   * needs an LES model.  The coefficients: NaN selects scalarVorticityCoefficient's 1.0 and
   * densityVorticityCoefficient's 0.1.  Pointwise LES models and the rotational form only. */
  int32_t scalar_les;
  double scalar_les_c;
  int32_t density_les;
  double density_les_c;
  int32_t gravity_direction;                /* gravityDirection: 0 x, 1 y, 2 z (default): the axis of g and
                                               of the density's mean gradient */
  int32_t density_hv;
  double density_hv_nu;
  int32_t density_hv_m;
  double density_sb_initk, density_sb_endk;
} ffs_params;

typedef struct ffs_sim ffs_sim;

/* Fills p with the reference defaults (parameters(), simulationtypes.jl:242-365). */
int ffs_params_default(ffs_params* p);

/* Replaces parameters(d) (simulationtypes.jl:242-365) + the module-load grid
 * constants (Globals.jl:106-182): allocates the device state of one run. */
int ffs_create(const ffs_params* p, ffs_sim** out);

/* Replaces VectorField{Float64}("u1.$start",...) (simulationtypes.jl:276): loads
 * one component.  n_doubles selects the layout: 2*(nx/2+1)*ny*nz (padded real or
 * half spectrum) or nx*ny*nz (unpadded real, space must be REAL). */
int ffs_set_field(ffs_sim* s, int field, int space, const double* host, size_t n_doubles);

/* Seeded random-phase initial fields generated on the device (no host field, no
 * transfer).  The reference has no built-in generator (its runs read u1.0 ...
 * produced elsewhere); these produce the `u?.0` / `rho.0` a case would load:
 * velocity u_hat = A(|k|) (I - k k/|k|^2) xi with A = |k| exp(-(|k|/k0)^2)/sqrt(4 pi)
 * (shell energy ~ k^4 exp(-2 (k/k0)^2)), scaled to 1/2 <|u|^2> = energy; density
 * rho_hat = A(|k|) xi scaled to <rho^2> = rms^2.  xi is a complex Gaussian per
 * (component, mode) from a counter-based generator keyed on the seed (two
 * SplitMix64 hashes -> Box-Muller), conjugate-symmetric on the kx = 0 plane, so
 * the field is independent of the grid's launch shape and reproducible on the host. */
int ffs_seed_velocity(ffs_sim* s, uint64_t seed, double k0, double energy);
int ffs_seed_density(ffs_sim* s, uint64_t seed, double k0, double rms);

/* Replaces real!(s.u) + write("u1.$it", s.u.rr.x) (output.jl:6-8) and
 * complex views of the state: copies one component out in the padded layout. */
int ffs_get_field(ffs_sim* s, int field, int space, double* host, size_t n_doubles);

/* No counterpart in the reference.  For callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): the real-space
 * read-out of ffs_get_field(s, field, FFS_SPACE_REAL, host, n) for the state fields u1 u2 u3 rho
 * scalar, with the same result, calling no hipMalloc / hipFree.  The inverse transforms use the
 * run's own pencil workspace and the last pass stages slabs of whole z planes through one
 * library-wide 32 MiB device array per GPU; a z plane (2*(nx/2+1)*ny doubles) larger than that is
 * rejected with FFS_ERR_ARG (no compiled size reaches it).  Calls that overlap on one GPU (two
 * host threads, any handles) race on that array: serialise them.  Synchronous, like ffs_get_field. */
int ffs_get_field_noalloc(ffs_sim* s, int field, double* host, size_t n_doubles);

/* Replaces initialize!(s) (FluidFlowSimulation.jl:67-91). */
int ffs_initialize(ffs_sim* s);

/* Replaces nsteps calls of advance_in_time!(s) (models.jl:3-17).  Stops early
 * with *event = FFS_EVENT_OUTPUT before a step whose iteration is an output
 * time; the next call performs that step.  *done receives the steps taken. */
int ffs_advance(ffs_sim* s, int64_t nsteps, int64_t* done, int* event);

/* Exact-resume checkpoints.  The reference restarts from u?.N alone and starts
 * its integrator over (Adams_Bashforth.jl:28-35, ETD.jl:99-108: history and dt
 * history are not restored); a run restored from a checkpoint instead continues
 * bitwise as the original would have.  The blob holds the velocity (and density)
 * with the integrator history, dt history, time and iteration in this build's
 * layout; it is loaded into a run created with the same parameters and not yet
 * initialized (ffs_checkpoint_load replaces ffs_set_field + ffs_initialize). */
int ffs_checkpoint_size(const ffs_sim* s, size_t* bytes);
int ffs_checkpoint_save(ffs_sim* s, void* host, size_t cap);
int ffs_checkpoint_load(ffs_sim* s, const void* host, size_t n);

/* State of the plane forcing schemes, per 2D shell (compute_shells2D, utils.jl:161-182):
 * FFS_FSTATE_R the oscillator state R (rfForcing; the `R.N` file of FluidFlowSimulation.jl:59-61
 * and forcing_types.jl:66-74) or the fixed power spectrum (QfForcing), FFS_FSTATE_EM the target
 * spectrum (taken from the initial field at ffs_initialize unless set before, as from
 * targSpectrum.dat, forcing_types.jl:76-83), FFS_FSTATE_KH the shells' mean |k_h| (read only),
 * FFS_FSTATE_EF the spectrum of the last forcing call over every shell (calculate_u1u2_spectrum!,
 * u1u2_mag for QfForcing; the values of the `Eh.N` file, output.jl:44-66; read only).
 * aForcing keeps one value per kz = 0 mode instead: an (nx/2+1) x ny array, x fastest (the
 * reference's R matrix), and has no FFS_FSTATE_KH.
 * set = 1 writes data[0..n), set = 0 reads; *n receives the number of values. */
#define FFS_FSTATE_R 0
#define FFS_FSTATE_EM 1
#define FFS_FSTATE_KH 2
#define FFS_FSTATE_EF 3
#define FFS_FSTATE_N 4 /* nRf / cNRf: the transfer spectrum N (the `N.N` file, forcing_types.jl:142-144) */
int ffs_forcing_state(ffs_sim* s, int what, int set, double* data, int cap, int* n);

/* Replaces the tail of run_simulation (FluidFlowSimulation.jl:46-49):
 * calculate_rhs!(s) (with set_dt! on variable-dt runs, and the spectra record
 * when the iteration is a spectra iteration), the forcing, writestats(s) at the
 * current iteration. */
int ffs_writestats_now(ffs_sim* s);

/* Pops the oldest queued statistics record (one per stats step, written by
 * writestats, stats.jl:39-62).  Returns 1 and fills the three rows if a record
 * was pending, 0 if none, <0 on error.  Rows start with iteration, time. */
int ffs_stats_pop(ffs_sim* s, double* stats, int cap_stats, int* n_stats, double* scales, int cap_scales,
                  int* n_scales, double* numbers, int cap_numbers, int* n_numbers);

/* Header line of a stats table (statsheader / writeheader, stats.jl:1-37). */
int ffs_stats_header(const ffs_sim* s, int table, char* buf, size_t cap);

/* Replaces spectrum_u + calculate_specHV (spec.jl:316-327,424-450) for the
 * current state.  *n_h = length(KH), *n_v = length(KRZ) = nz/2+1; h receives
 * 2 * n_h values (specH of (|u1|^2+|u2|^2)/2, then specH of |u3|^2/2) and v
 * receives 2 * n_v values (the matching specV arrays): the u_h.specH,
 * u_v.specH, u_h.specV and u_v.specV files of writespectrum (spec.jl:173-194). */
int ffs_spectrum(ffs_sim* s, double* h, int cap_h, int* n_h, double* v, int cap_v, int* n_v);

/* The spectra files of write_spec and write_spec_forcing (spec.jl:1-168),
 * computed on the device at every spectra iteration (is_spec_time,
 * simulationtypes.jl:102-104: iteration % dtSpec == 0) at the point the
 * reference computes them (models.jl:37, models.jl:8) and queued like the
 * stats records.  A record holds FFS_SPEC_Q per-mode quantities, each binned by
 * calculate_specHV (spec.jl:424-450) into specH (n_h values over KH) then specV
 * (n_v values over KRZ), quantity after quantity in the order of FFS_SPEC_*.
 * Quantities a run does not have (LES, density, plane forcing) read zero.
 * The files: for u, vis, nl, press, les, hvis: n_h.specH/V (the _H quantity),
 * n_v.specH/V (the _V one) and n.specH (their specH sum); buoyancy_v,
 * rhosource (= -specH,V(buoyancy_v) / g_z * densityGradient), rho, rhodiss,
 * rhonl and forcing_h: .specH and .specV. */
#define FFS_SPEC_U_H 0
#define FFS_SPEC_U_V 1
#define FFS_SPEC_VIS_H 2
#define FFS_SPEC_VIS_V 3
#define FFS_SPEC_NL_H 4
#define FFS_SPEC_NL_V 5
#define FFS_SPEC_PRESS_H 6
#define FFS_SPEC_PRESS_V 7
#define FFS_SPEC_LES_H 8
#define FFS_SPEC_LES_V 9
#define FFS_SPEC_BUOYANCY 10
#define FFS_SPEC_RHO 11
#define FFS_SPEC_RHODISS 12
#define FFS_SPEC_RHONL 13
#define FFS_SPEC_FORCING 14
#define FFS_SPEC_HVIS_H 15
#define FFS_SPEC_HVIS_V 16
#define FFS_SPEC_Q 17
/* KH (n_h values) and KRZ (n_v values) of Globals.jl:148-152. */
int ffs_spectra_axes(ffs_sim* s, double* kh, int cap_h, int* n_h, double* krz, int cap_v, int* n_v);
/* Pops the oldest spectra record (waiting for it if it is still in flight):
 * returns 1 and fills *iteration and FFS_SPEC_Q * (n_h + n_v) values (*n), 0 if
 * no record is pending, <0 on error. */
int ffs_spectra_pop(ffs_sim* s, int64_t* iteration, double* data, size_t cap, size_t* n);

/* Scalars: s.iteration[], s.time[], get_dt(s), last max|u_i|. */
int ffs_get_scalar(const ffs_sim* s, int what, double* out);

/* Per-kernel device time accumulation with HIP events on the handle's stream. */
int ffs_profile_enable(ffs_sim* s, int on);
/* Fills up to cap entries; names point into storage owned by the handle. */
int ffs_profile_get(ffs_sim* s, const char** names, double* total_ms, int64_t* launches, int cap, int* n);
int ffs_profile_reset(ffs_sim* s);

/* Device bytes held by the handle. */
int ffs_device_bytes(const ffs_sim* s, size_t* bytes);

/* Blocks until all queued work of the handle is done. */
int ffs_synchronize(ffs_sim* s);

/* Device bytes ffs_create would allocate for these parameters (state, pencil
 * work buffers, tables, and the transient staging of ffs_set_field), computed
 * without touching the device.  No counterpart in the reference (one Julia
 * process per run, FluidFlowSimulation.jl:23); used for HBM admission control
 * of many runs on one GPU. */
int ffs_footprint(const ffs_params* p, size_t* bytes);

/* Free and total bytes of the current device (hipMemGetInfo), for the same
 * admission control.  No counterpart in the reference. */
int ffs_device_memory(size_t* free_bytes, size_t* total_bytes);

/* Replays steady steps (no stats, spectrum or output at that iteration, after
 * the integrator's start-up steps) as HIP graphs, one per rotation of the
 * three-level history: the same kernels with the same arguments as the eager
 * path, so results are bitwise identical.  on = 0 turns it off.  Default on.
 * Steps under ffs_profile_enable always run eagerly.  No counterpart in the
 * reference (the Julia step is host code). */
int ffs_set_graph(ffs_sim* s, int on);

/* Plain steps (the rotational form without LES, density, passive scalar,
 * plane forcing or hyperviscosity, and no stats or spectra at that iteration)
 * run the forward z transform of the nonlinear term and the time-step update
 * as one kernel, so the nonlinear term is not written to and read back from
 * device memory; the results are bitwise those of the two kernels.  on = 0
 * runs them as two kernels.  Default on for FP64, off for FP32 (whose fused
 * tile holds too many registers to be faster).  No counterpart in the
 * reference. */
int ffs_set_fusion(ffs_sim* s, int on);

/* A run's plan: the kernel choices that do not change what is computed, only
 * how (results agree with the oracle to the same tolerance under every plan,
 * and are bitwise reproducible under one plan).  An autotuner times the
 * candidates (ffs/tune.py) and keeps the fastest per device, grid, precision
 * and model.  No counterpart in the reference (FFTW picks its own plans). */
typedef struct ffs_plan {
  int x_pair;  /* 1: power-of-two x passes (up to 512) use the mirror-paired kernel;
                * 2 (the default): FP64 u x w passes its looping form, which copies
                * each next row group's first field pair into LDS ahead (other
                * passes as 1); ffs_plan_options reports the largest that applies */
  int x_waves; /* wavefronts per block of the register x passes: 1 or 2 */
  int x_split; /* 1: density / scalar fluxes and the Smagorinsky stress in
                * register x launches of their own; 0: one general x pass */
  int fusion;  /* as ffs_set_fusion */
  int z_slab;  /* planes per z slab of the nonlinear term's inverse y, x and
                * forward y passes (0: whole grid per launch); ffs_plan_options
                * reports the default slab size of the run here (0: no slabs).
                * Slabs thicker than the default may allocate device memory
                * that ffs_footprint does not count */
  int reserved[3];
} ffs_plan;
int ffs_get_plan(const ffs_sim* s, ffs_plan* out);
/* The current device as "<arch> <CUs>CU <memory>GiB" (the tuner's cache key: fields every process sees the same). */
int ffs_device_name(char* buf, size_t n);
/* Applies a plan (FFS_ERR_ARG, the plan unchanged, if the x pass cannot hold
 * the run's rows under it); waits for the run's queued work. */
int ffs_set_plan(ffs_sim* s, const ffs_plan* plan);
/* Which fields of a plan change this run's kernels (1) and which do not (0). */
int ffs_plan_options(const ffs_sim* s, ffs_plan* matters);

/* An ensemble: up to FFS_MAX_BATCH initialized handles on the current device
 * advanced in lockstep.  A step in which every run is steady is one graph
 * launch whose branches are independent, so they overlap on the device: runs
 * whose steps launch the same kernels with the same shapes (same transform
 * lengths, precision and model) are merged into one group, whose step is one
 * launch per pass over all its runs; any other run is a branch of its own.
 * Other steps run each handle's own path.  Each run's results are identical
 * to running it alone.  The batch does not own the handles: destroy it before
 * any of them.  No counterpart in the reference (one simulation per
 * process). */
#define FFS_MAX_BATCH 256
typedef struct ffs_batch ffs_batch;
int ffs_batch_create(ffs_sim* const* sims, int n, ffs_batch** out);
/* Advances every run by up to nsteps steps.  Stops early, with *event =
 * FFS_EVENT_OUTPUT and *who = the first such run's index, before a step at
 * which some run has an output iteration; the next call performs that step.
 * *done receives the lockstep steps taken. */
int ffs_batch_advance(ffs_batch* b, int64_t nsteps, int64_t* done, int* event, int* who);
/* Blocks until every run's queued work is done. */
int ffs_batch_synchronize(ffs_batch* b);
/* Without blocking: 1 when every run's queued work is done, 0 while some is
 * still running, or an FFS_ERR_* code. */
int ffs_batch_query(ffs_batch* b);
/* Device time of the last ffs_batch_advance call that took a step, from a
 * timestamp taken after every run's earlier work to one after all of the
 * call's work (so it counts what the runs waited for each other and for
 * other work on the device), and the steps it took.  Returns 1 while that
 * work is still running, FFS_ERR_STATE before any step. */
int ffs_batch_last_advance_ms(ffs_batch* b, double* ms, int64_t* steps);
/* on = 0: every run keeps a branch of its own (no merged launches); on = 1
 * (default) merges runs of the same launch shapes.  Bitwise the same results
 * either way. */
int ffs_batch_set_merge(ffs_batch* b, int on);
/* The last steady-step graph launched: its kernel launches per step, its
 * merged groups and the runs inside them.  FFS_ERR_STATE before the first
 * steady step. */
int ffs_batch_launch_info(const ffs_batch* b, int* launches, int* groups, int* merged_runs);
const char* ffs_batch_last_error(const ffs_batch* b);
void ffs_batch_destroy(ffs_batch* b);

const char* ffs_last_error(const ffs_sim* s);
void ffs_destroy(ffs_sim* s);
int ffs_abi_version(void);

#ifdef __cplusplus
}
#endif
#endif /* FFS_FFS_H */
