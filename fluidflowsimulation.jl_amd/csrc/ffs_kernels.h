// Kernel launch interface of the FFS HIP solver (internal to libffs_hip.so).
#pragma once
#include <string>
#include <vector>
#include <hip/hip_runtime.h>
#include <cstdint>
#include <type_traits>

namespace fsk {

// Geometry of one run, as seen by the kernels.  Spectral fields are stored
// "compact": only the retained (2/3-rule) modes, mx x my x mz, x fastest,
// with kx = 0..mx-1, and the kept ky / kz listed non-negative first.
struct Geo {
  int nrx, ny, nz;   // physical grid (nx, ny, nz of the `global` file)
  int nxh;           // nrx/2 + 1  (NX of Globals.jl:111)
  int mx, my, mz;    // retained modes per direction
  int sx;            // x stride of spectral/pencil arrays: mx rounded up to a 128-byte line
  const int* ymap;   // compact y -> FFT index       [my]
  const int* zmap;   // compact z -> FFT index       [mz]
  const int* yinv;   // FFT y index -> compact or -1 [ny]
  const int* zinv;   // FFT z index -> compact or -1 [nz]
  const double* kx;  // wavenumbers per compact index
  const double* ky;
  const double* kz;
  // The retained set along y (z) is two contiguous blocks of FFT indices: [0, fy) and
  // [ny - (my - fy), ny).  Kernels map indices arithmetically, so no table load sits in
  // front of a data load; wavenumbers come from the tables (their loads depend only on the
  // index and overlap the data loads; a double division per element costs more).
  int fy, fz;
  double lx, ly, lz;  // domain factors L (the box is 2 pi L per side): k = signed index / L
};

// FFT index -> compact index, or -1 for a dealiased row
__host__ __device__ __forceinline__ int y_compact(const Geo& g, int p) {
  return p < g.fy ? p : (p >= g.ny - (g.my - g.fy) ? p - g.ny + g.my : -1);
}
__host__ __device__ __forceinline__ int z_compact(const Geo& g, int p) {
  return p < g.fz ? p : (p >= g.nz - (g.mz - g.fz) ? p - g.nz + g.mz : -1);
}
// compact index -> FFT index
__host__ __device__ __forceinline__ int y_fft(const Geo& g, int c) { return c < g.fy ? c : c + g.ny - g.my; }
__host__ __device__ __forceinline__ int z_fft(const Geo& g, int c) { return c < g.fz ? c : c + g.nz - g.mz; }
// wavenumbers exactly as rfftfreq / fftfreq (Globals.jl:95-104): signed index / L
__host__ __device__ __forceinline__ double kx_of(const Geo& g, int cx) { return (double)cx / g.lx; }
__host__ __device__ __forceinline__ double ky_of(const Geo& g, int cy) {
  const int j = y_fft(g, cy);
  return (double)(j < (g.ny + 1) / 2 ? j : j - g.ny) / g.ly;
}
__host__ __device__ __forceinline__ double kz_of(const Geo& g, int cz) {
  const int j = z_fft(g, cz);
  return (double)(j < (g.nz + 1) / 2 ? j : j - g.nz) / g.lz;
}

// The x pass's maxima (max|u_i| for the CFL, max|rho|) are gathered per block and combined with one
// atomicMax per block on the bit pattern of a non-negative double.  The atomics go to kMaxSlots slots, one
// 128-byte line apart (block b uses slot b mod kMaxSlots), so thousands of blocks do not queue on one
// address; set_dt takes the maximum over the slots and clears them.
constexpr int kMaxSlots = 32;
constexpr int kMaxSlotStride = 16;  // unsigned long longs between slots
struct alignas(128) MaxSlots {
  unsigned long long v[kMaxSlots * kMaxSlotStride];
};
__device__ __forceinline__ unsigned long long* max_slot(unsigned long long* slots, int block) {
  return slots + (block & (kMaxSlots - 1)) * kMaxSlotStride;
}

// Device-resident scalar state of a run (so variable-dt runs never stall the
// host): mirrors the Refs of Adams_Bashforth3rdO / ETD3rdO and Simulation.
struct DevState {
  double dt, dt2, dt3;     // current and two previous steps
  double At, Bt, Ct;       // AB3 coefficients for the current step
  double time;             // s.time[]
  double umax_d;           // last max|u_i| (decoded)
  MaxSlots umax;           // bit patterns of max|u_i| per slot
  MaxSlots rhomax;         // ... of max|rho|
  double force_factor;     // isotropic forcing: P / sum |u_lowk|^2
  double rhomax_d;         // last max|rho|
  double etd[20];          // ETD3 scalars a0..a3 b0..b2 c0..c2 e1..e9 of set_ABCt! (ETD.jl:142-227)
  double force_hp, force_vp;  // plane forcing: horizontal and vertical power (the hp, vp stats)
  double force_vert;          // plane forcing: factor of the kx = ky = 0 modes
};

// kz = 0-plane forcing, the six schemes of Forcing_methods/ (forcing_types.jl): kind 2 rfForcing
// (forcing.jl:6-114), 3 QfForcing (qfforcing.jl), 4 cRfForcing (cforcing.jl), 5 nRfForcing
// (nrforcing.jl), 6 cNRfForcing (cnrforcing.jl), 7 aForcing (aforcing.jl).  A slot is a forced 2D
// shell (or, for aForcing, a forced mode); the forced elements are the retained kz = 0 modes of the
// forced slots, listed in the reference's loop order (y outer, x inner).
enum : int { PF_RF = 2, PF_QF = 3, PF_CRF = 4, PF_NRF = 5, PF_CNRF = 6, PF_A = 7 };
struct PlaneForcing {
  int kind;              // PF_*; negated: compute the initial target (Em = Ef, and N, Nm1) only
  int nel;               // forced elements
  const int* elem;       // compact index of each (kz = 0 plane)
  const int* eslot;      // its slot
  const double* ew;      // its weight (1 at kx = 0, else 2)
  const double* ek2;     // its |k_h|^2 (the nu k^2 of calculate_nn_spectrum!)
  const int* slot_ptr;   // elements of slot s: slot_el[slot_ptr[s] .. slot_ptr[s + 1]) (list order)
  const int* slot_el;
  // every 2D shell (the Ef of the reference, written as Eh.N): its retained kz = 0 modes
  int nshell;            // 0 for aForcing
  const int* shell_ptr;  // modes of shell n: shell_el[shell_ptr[n] .. shell_ptr[n + 1]) (y outer, x inner)
  const int* shell_el;   // compact index
  double* Ef_all;        // per shell
  int nslot;
  double* R;             // per slot: oscillator state, or the power spectrum (QfForcing)
  double* Em;            // per slot: target spectrum
  double* Ef;            // per slot: current spectrum
  double* N;             // per slot: nonlinear transfer spectrum (nRf, cNRf) and its previous value
  double* Nm1;
  const double* Zf;      // per slot
  double* slot_factor;   // per slot: factor (rf family) / fkn (QfForcing)
  double* plane_factor;  // per element of the compact kz = 0 plane (sx x my): the force is u * factor
  int nvert;             // vertical modes kx = ky = 0 in the reference's order (+m, -m, ...)
  const int* vert;
  double tf, alpha, cfl, nu;
  int variable_dt;
};
enum : int { kStatsSlots = 40 };

struct Tables {  // per-length twiddle tables (device pointers), by precision
  const void* tw;
};

// forward-z outputs at most: uu (5) + density and scalar fluxes (3 + 3) + tau (5) + 1/2|u|^2
enum : int { kMaxFields = 17 };

// extra operands of the column kernels
template <typename T> struct ColAux {
  const typename std::conditional<sizeof(T) == 8, double2, float2>::type* rho = nullptr;  // inverse z: density
  const typename std::conditional<sizeof(T) == 8, double2, float2>::type* phi = nullptr;  // and passive scalar
  int nsform = 0;  // conservation form: the inputs have no vorticity (u, then rho, phi, S)
  // the scalars' LES fluxes: i k s of these scalars follow S (3 fields each), in this order
  const typename std::conditional<sizeof(T) == 8, double2, float2>::type* grad1 = nullptr;
  const typename std::conditional<sizeof(T) == 8, double2, float2>::type* grad2 = nullptr;
  typename std::conditional<sizeof(T) == 8, double2, float2>::type* dst[kMaxFields] = {};  // forward z outputs
};

// ---- launch records: batching the launches of several runs' steps ----
// A recorded launch of a batchable kernel: its batch form (an array of call records with the run index
// folded into grid z), the launch shape and the record's bytes.
struct KLaunch {
  const void* batch = nullptr;
  dim3 grid, block;
  size_t lds = 0;
  const char* name = "";
  std::vector<unsigned char> call;
};
struct KRecorder {
  std::vector<KLaunch> launches;
};
// set on a thread, batchable launches are recorded there instead of issued
extern thread_local KRecorder* g_recorder;
// one launch of k.batch over nrun runs whose call records (k.call's layout) are at `calls` (device memory)
void issue_batch(const KLaunch& k, const void* calls, int nrun, hipStream_t s);

// ---- launchers (T = double or float) ----
// inverse z pass; curl = true: the nf physical-space inputs of the nonlinear term
// (u, w, [rho], [S]) are formed on the fly from u (and rho)
template <typename T>
void launch_inv_z(const Geo& g, const void* u, void* a1, int nf, bool curl, const void* tw, hipStream_t s,
                  const void* rho = nullptr, const void* phi = nullptr, int nsform = 0,
                  const void* grad1 = nullptr, const void* grad2 = nullptr);  // i k s after S, per scalar
template <typename T>
// z0, nzc: only the planes [z0, z0 + nzc) (default: all), written as a slab of nzc planes
void launch_inv_y(const Geo& g, const void* a1, void* a2, int nf, const void* tw, hipStream_t s, int z0 = 0,
                  int nzc = -1);
// field set of the fused x pass
// pointwise LES models of calc_les! (les_real_calc.jl:74-114); codes as FFS_LES_* of the C-ABI
enum : int { LES_SMAGORINSKY = 1, LES_VREMAN = 2, LES_SILVIS = 3, LES_PRODUCTION = 4, LES_STABLE_NL = 5,
             LES_FAKE_SMAGORINSKY = 6, LES_SANDP = 7, LES_DYNAMIC = 8 };
struct LesArgs {
  int model;   // LES_*
  int smodel;  // SandP's S model
  double c;    // the model constant (SandP: its S model's)
  double cp;   // Silvis P constant or SandP's cb
  double d2;   // Delta^2
};
// whether the model reads the vorticity (every one but Smagorinsky)
inline bool les_needs_w(const LesArgs& l) { return l.model != LES_SMAGORINSKY && l.model != LES_DYNAMIC; }

struct XOps {
  int nin, nout;      // physical inputs (6 [+1 rho] [+1 phi] [+5 S]) and forward outputs (3 [+3] [+3] [+5] [+1 ke])
  int rho, les, ke;   // flags: rho, phi and les in that order after u x w, ke last
  int phi;
  int ns;             // conservation form: inputs u (no w), outputs uu = traceless(-u u) (5) in place of u x w
  double les_2alpha;  // 2 (c Delta)^2 of the Smagorinsky stress
  LesArgs lm;         // the LES model (lm.model != LES_SMAGORINSKY: the general x pass, with w at slots 3..5)
  // the scalars' LES fluxes (les_real_calc.jl:120-136): 0 none, 1 EddyDiffusion nu_t grad s, 2 also
  // (c Delta^2 / 2) w x grad s (VorticityDiffusion); their gradients are inputs after S, rho's first
  int rho_les = 0, phi_les = 0;
  double rho_cv = 0.0, phi_cv = 0.0;  // c Delta^2 of VorticityDiffusion
  double les_alpha = 0.0;             // c^2 Delta^2 (Smagorinsky's nu_t = alpha |S|)
  // the run's plan (ffs_set_plan).  pair: power-of-two rows may take the mirror-paired pass (2: FP64 u x w
  // rows its looping form, which copies each next row group's first field pair into LDS ahead); waves:
  // wavefronts per block of the register passes (1 or 2); split: with a register-capable length, the
  // density / scalar fluxes and the Smagorinsky stress run as register launches of their own (0: all
  // products in one general LDS pass)
  int pair = 1, waves = 2, split = 1;
};
// whether the plan's choices change anything for a run of x length n with these products (pair: 1 the
// mirror-paired pass applies, 2 its looping form too; split: both the register launches and the general
// pass can hold the row)
template <typename T>
void x_plan_options(int n, int mx, const XOps& ops, int* pair, int* waves, int* split);
// whether n is a compiled transform length, and (empty: yes) whether the x pass of length n can run
// with these options -- checked when a run is created, before anything is allocated
// the compiled transform lengths, space-separated
std::string compiled_lengths();
bool length_supported(int n);
template <typename T>
std::string x_pass_check(int n, const XOps& ops);
template <typename T>
void launch_x_nl(const Geo& g, const void* a2, void* b2, const XOps& ops, unsigned long long* umax,
                 unsigned long long* rhomax, const void* tw, hipStream_t s);
template <typename T>
void launch_x_c2r(const Geo& g, const void* a2, T* real, int nf, const void* tw, hipStream_t s);
// the same C2R for the rows [row0, row0 + nrow) of one field (row0 even), written as doubles at
// real[(row - row0) * 2 nxh + x]
template <typename T>
void launch_x_c2r_rows(const Geo& g, const void* a2, double* real, int row0, int nrow, const void* tw, hipStream_t s);
template <typename T>
void launch_x_r2c(const Geo& g, const T* real, void* b2, int nf, const void* tw, hipStream_t s);
template <typename T>
// z0, nzc: the input is a slab of nzc planes, written to the planes [z0, z0 + nzc) (default: all)
void launch_fwd_y(const Geo& g, const void* b2, void* b1, int nf, const void* tw, hipStream_t s, int z0 = 0,
                  int nzc = -1);
// forward z pass: field f goes to dst[f] (compact spectral), scaled; the k=0 mode
// of the first nzero_origin fields is zeroed
template <typename T>
void launch_fwd_z(const Geo& g, const void* b1, void* const* dst, int nf, int nzero_origin, double scale,
                  const void* tw, hipStream_t s);

// hyperviscosity (Hyperviscosity/hyperviscosity_types.jl): kind 0 none, 1 HyperViscosity{nu_h, M}
// (-nu_h k^2M in the AB3 rhs and the ETD c), 2 SpectralBarrier (ETD c += log(esp(|k|, initk, endk, 1/4)))
struct HvArgs {
  int kind;
  int m;
  double nuh;
  double sb_i, sb_p;  // initk, endk
};

struct UpdateArgs {
  int integrator;      // 0 = Adams_Bashforth3rdO (explicit viscosity), 1 = ETD3rdO (ETD1/ETD2 start)
  int step;            // integrator iteration after this step's increment (1, 2, ...); 0 = final rhs only
  double nu;           // kinematic viscosity
  int forcing;         // 0 none, 1 isotropic constant-power shell forcing, 2 rf / 3 qf plane forcing
  const double* pf_factor;  // plane forcing: per-element factor on the kz = 0 plane (sx x my, 0 if not forced)
  int pf_kz0;               // smallest |kz| of the vertical modes (qf 3, the others 1; the largest is 5)
  double kf2;          // forced shell |k|^2 bound
  const void* tau;     // Smagorinsky stress, 5 compact components (or null)
  void* rho;           // density (updated in place) or null
  void* rrhs;          // density rhs out
  const void* rfm1;    // density history
  const void* rfm2;
  const void* flux;    // F(-rho u), 3 compact components
  double alpha;        // density diffusivity nu / Pr (ETD only, as in the reference)
  double drhodz;       // mean density gradient
  double gz;           // unused by the update (see gval)
  int gdir;            // gravityDirection: the axis of g and of the density's mean gradient (0 x, 1 y, 2 z)
  double gval;         // zAcceleration / referenceDensity, the component of g along gdir
  double* time;        // s.time[] (DevState::time): advanced by dt once per step, or null
  HvArgs hv;
  HvArgs hvd;          // the density's own hyperviscosity (its ETD3 operator)
  // passive scalar (PassiveScalar, simulationtypes.jl:198-224; models.jl:353-357, 470)
  void* phi;           // updated in place, or null
  void* prhs;          // its rhs out
  const void* pfm1;    // its history
  const void* pfm2;
  const void* pflux;   // F(-phi u), 3 compact components
  double palpha;       // diffusivity nu / scalarPr (ETD only)
  double pgrad;        // scalarGradient
  int pdir;            // scalarDirection: 0 x, 1 y, 2 z
};
template <typename T>
void launch_update(const Geo& g, void* u, void* rhs, const void* fm1, const void* fm2, const DevState* st,
                   const UpdateArgs& a, hipStream_t s);
// whether an update needs the optional paths (LES, density, passive scalar, plane forcing, hyperviscosity)
inline bool update_full(const UpdateArgs& a) {
  return a.rho || a.phi || a.tau || a.forcing >= 2 || a.hv.kind != 0;
}
// the forward z pass of the 3 nonlinear components (b1, as launch_fwd_z with nzero 3) fused with the update
// of a step for which update_full is false; rhs receives the final rhs, as launch_fwd_z + launch_update
template <typename T>
void launch_fwd_z_update(const Geo& g, const void* b1, void* u, void* rhs, const void* fm1, const void* fm2,
                         const DevState* st, const UpdateArgs& a, double scale, const void* tw, hipStream_t s);
// the Smagorinsky stress in real space: s[5] padded real S fields (xx xy xz yy yz) -> c2a |S| S in place
template <typename T>
void launch_les_tau_real(T* s5, size_t n_padded, double c2a, hipStream_t s);
// the density's LES flux alone (its real-space field at writeoutput, output.jl:22-26): nu_t grad rho
// [+ (cv/2) w x grad rho] from S (5), w (3) and grad rho (3) padded real fields into out3
template <typename T>
void launch_les_scalar_flux_real(const T* s5, const T* w3, const T* g3, T* out3, size_t n_padded, const LesArgs& l,
                                 double alpha, int les, double cv, hipStream_t s);
// the same for a model that reads the vorticity: w3[3] padded real w fields
template <typename T>
void launch_les_model_real(T* s5, const T* w3, size_t n_padded, const LesArgs& l, hipStream_t s);

// Dynamic Smagorinsky (LESmodels/dynamic_smagorinsky.jl).  Gaussian test filter G = exp(-td2 k^2 / 24)
// (Gaussfilter, utils.jl:220-225).
// fourierspacep1! (models.jl:98-108) on the compact modes: out = S (5), M = G S (5), uf = G u (3)
template <typename T>
void launch_dyn_prep(const Geo& g, const void* u, void* out, double td2, hipStream_t s, int with_w = 0);
// with_w (dynamicSandP, piomelliSandP): also w = ik x u (3) and G w (3) at out slots 13..18
// realspace_dynamic_les_calculation! (dynamic_smagorinsky.jl:44-76) on padded real fields:
// sa = Delta^2 |S| S (5), l = symouter(u, u) (6: xx xy xz yy yz zz)
// The P part of dynamicSandP / piomelliSandP: w3 the vorticity, pa5 out P = [cpstar] Delta^2 Lie(S, AntiSymTen(w/2))
struct DynP {
  const void* w3 = nullptr;     // real vorticity (null: no P part)
  const void* wf3 = nullptr;    // its test-filtered field
  void* pa5 = nullptr;          // P (filtered between the products and the coefficient)
  const void* cpstar = nullptr; // piomelliSandP's extrapolated coefficient (null: dynamicSandP)
  void* cp = nullptr;           // the coefficient cp out
};
template <typename T>
void launch_dyn_products(const T* s5, const T* u3, T* sa5, T* l6, size_t n_padded, double d2, hipStream_t s,
                         const T* cstar = nullptr,  // piomelliSmagorinsky: sa = 2 cstar Delta^2 |S| S
                         const DynP& dp = DynP{});
// x G of a spectrum stored on the full mode set of g (fourierspace_dynamic_les_calculation!, average_c!)
template <typename T>
void launch_gauss(const Geo& g, void* spec, double td2, hipStream_t s);
// coefficient_les_calculation! (dynamic_smagorinsky.jl:131-200, dynamicSmagorinsky branch): from the test-
// filtered uf (3), M (5), L (6), sa (5): c = clamp(0.5 (L:M)/(M:M), -1, 1), L = traceless(uf uf - L),
// M = td2e |M| M - sa
template <typename T>
void launch_dyn_coef(const T* uf3, const T* m5, const T* l6, const T* sa5, T* c, size_t n_padded, double td2e,
                     hipStream_t s, int piomelli = 0,  // piomelli: c = clamp((sa + L):M/(M:M)), M = 2 td2e |M| M
                     const DynP& dp = DynP{});          // and cp of the SandP variants
// piomelliSmagorinsky's calc_cstar! (dynamic_piomelli.jl:1-11): c <- c + r (c - cm1), cm1 <- c, with
// r = dt / dtm1 from the device dt (0 read as 1) and *dtm1 (0 read as dt); then *dtm1 <- dt
template <typename T>
void launch_piomelli_cstar(T* c, T* cm1, const DevState* st, double* dtm1, size_t n_padded, hipStream_t s);
template <typename T>
void launch_fill_real(T* a, size_t n, double v, hipStream_t s);
// calc_les! dynamic branch (les_real_calc.jl:75-77): t = 2 max(cmin, min(0.4, c)) Delta^2 |S| S
template <typename T>
void launch_dyn_tau(const T* s5, const T* c, T* t5, size_t n_padded, double d2, double cmin, hipStream_t s,
                    const T* w3 = nullptr, const T* cp = nullptr);  // + clamp(cp, +-0.8) Delta^2 Lie(S, AntiSymTen(w/2))
// conservation form: rhs = (i k) . uu (div!, vectorfunctions.jl:107-117 on the SymTrTen field)
template <typename T>
void launch_div_uu(const Geo& g, const void* uu, void* rhs, hipStream_t s);
// the plane forcing of one step (rhs_final = 0: rhs holds the nonlinear term, formed per element as
// the update kernel does; 1: rhs is final, as after initialization)
template <typename T>
void launch_plane_forcing(const Geo& g, const void* u, const void* rhs, DevState* st, const PlaneForcing& pf,
                          const UpdateArgs& ua, int rhs_final, int it, hipStream_t s);
// the u (rho, phi) round trip of the reference on the kx = 0 and kx = N/2 planes: Hermitian part, then
// the projection of u; needed only when a Nyquist mode is retained
template <typename T>
void launch_hermitian_planes(const Geo& g, void* u, void* rho, void* phi, hipStream_t s);
template <typename T>
void launch_project(const Geo& g, void* u, hipStream_t s);

// stats: partial sums per block then a deterministic final reduction
struct StatsAux {
  const void* tau;  // Smagorinsky stress (5 compact components) or null
  const void* rho;  // density or null
  double gz;        // unused (see gval)
  int gdir;         // the axis of g (buoyancy in the pressure statistics, the buoyancy flux)
  double gval;      // g along gdir
  HvArgs hv;        // hyperviscous dissipation sums (slots 30, 31)
  HvArgs hvd;       // the density's (slot 37)
  const void* phi;  // passive scalar or null (slots 32-36)
  const void* rflux = nullptr;  // the density's total flux, with an LES flux (rhopr, slot 38)
};
template <typename T>
int launch_stats(const Geo& g, const void* u, const void* nl, const void* ke, const StatsAux& aux, double* partials,
                 double* out, hipStream_t s);
template <typename T>
void launch_spectrum(const Geo& g, const void* u, double* out_x, double* out_y0, double* out_y, double* out_z0,
                     double* out_z, double* partials, hipStream_t s);
// the spectra of write_spec and write_spec_forcing (spec.jl:1-168) at once: per mode, the kSpecQ
// quantities below, reduced to the marginals calculate_specHV bins (spec.jl:424-450)
enum : int {
  SQ_U_H, SQ_U_V, SQ_VIS_H, SQ_VIS_V, SQ_NL_H, SQ_NL_V, SQ_PRESS_H, SQ_PRESS_V, SQ_LES_H, SQ_LES_V,
  SQ_BUOY, SQ_RHO, SQ_RHODISS, SQ_RHONL, SQ_FORCING, SQ_HVIS_H, SQ_HVIS_V, kSpecQ
};
struct SpecArgs {
  const void* nl;     // the nonlinear term F(u x w) (rhs before fourierspacep2!)
  const void* ke;     // F(|u|^2 / 2)
  const void* tau;    // Smagorinsky stress (5 components) or null
  const void* rho;    // density or null
  const void* flux;   // F(-rho u) (3 components) or null
  const double* pf_factor;  // plane forcing: per kz = 0 element factor, or null (no forcing_h)
  int pf_kz0;               // plane forcing: the vertical modes are pf_kz0 <= |kz| <= 5
  double nu, alpha, gz;  // gz = g.z (spectrum_buoyancy reads g.z only)
  int gdir;              // the axis of g for the pressure's buoyancy term
  double gval;
  const DevState* st;
  HvArgs hv;  // hvis: spectrum_hyperviscosity / spectrum_spectral_barrier
};
// per quantity q: X[mx] | Y0[my] | Y[my] | Z0[mz] | Z[mz]  (spectra_out_len doubles each)
__host__ __device__ inline size_t spectra_out_len(const Geo& g) { return (size_t)g.mx + 2 * (size_t)g.my + 2 * (size_t)g.mz; }
inline size_t spectra_partials_len(const Geo& g) {
  return (size_t)kSpecQ * g.mz * ((size_t)g.mx + 2 * (size_t)g.my + 2);
}
template <typename T>
void launch_spectra(const Geo& g, const void* u, const SpecArgs& a, double* partials, double* out, hipStream_t s);
template <typename T>
void launch_forcing_sum(const Geo& g, const void* u, const int* band, int nband, DevState* st, double power,
                        hipStream_t s);

// set_dt! on device; etd selects the ETD3rdO history rule; gnorm > 0 adds the density limit
void launch_set_dt(DevState* st, int it, int tit, double cfl, double dx, int etd, double gnorm, hipStream_t s);
// set_dt (if do_dt) and the forcing band sum (if nband > 0) in one single-block launch
struct ScalarArgs {
  int do_dt, it, tit, etd;
  double cfl, dx, gnorm;
  const int* band;
  int nband;
  double power;
};
template <typename T>
void launch_step_scalars(const Geo& g, const void* u, DevState* st, const ScalarArgs& a, hipStream_t s);
// AB3 coefficients and ETD3 scalars of the current dt, dt2, dt3 (fixed-step runs)
void host_step_coefficients(DevState* st);
void launch_end_step(DevState* st, hipStream_t s);

template <typename T>
void launch_cast_copy(const void* src, void* dst, size_t n, hipStream_t s);

// seeded random-phase field in the compact layout: nf = 3 solenoidal velocity scaled to energy `target`,
// nf = 1 scalar scaled to rms `target` (partials: >= 512 * kStatsSlots doubles; red: >= 1 double)
template <typename T>
void launch_seed(const Geo& g, void* out, int nf, unsigned long long seed, double k0, double target, double* partials,
                 double* red, hipStream_t s);

}  // namespace fsk
