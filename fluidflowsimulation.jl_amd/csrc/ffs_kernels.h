// Kernel launch interface of the FFS HIP solver (internal to libffs_hip.so).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

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
};

// Device-resident scalar state of a run (so variable-dt runs never stall the
// host): mirrors the Refs of Adams_Bashforth3rdO / ETD3rdO and Simulation.
struct DevState {
  double dt, dt2, dt3;     // current and two previous steps
  double At, Bt, Ct;       // AB3 coefficients for the current step
  double time;             // s.time[]
  double umax_d;           // last max|u_i| (decoded)
  unsigned long long umax; // bit pattern of max |u_i| (atomicMax on non-negative doubles)
  unsigned long long rhomax;
  double force_factor;     // isotropic forcing: P / sum |u_lowk|^2
  double pad[4];
};

enum : int { kStatsSlots = 32 };

struct Tables {  // per-length twiddle tables (device pointers), by precision
  const void* tw;
};

// ---- launchers (T = double or float) ----
template <typename T>
void launch_inv_z(const Geo& g, const void* u, void* a1, int nf, bool curl, const void* tw, hipStream_t s);
template <typename T>
void launch_inv_y(const Geo& g, const void* a1, void* a2, int nf, const void* tw, hipStream_t s);
template <typename T>
void launch_x_nl(const Geo& g, const void* a2, void* b2, bool with_ke, unsigned long long* umax, const void* tw,
                 hipStream_t s);
template <typename T>
void launch_x_c2r(const Geo& g, const void* a2, T* real, int nf, const void* tw, hipStream_t s);
template <typename T>
void launch_x_r2c(const Geo& g, const T* real, void* b2, int nf, const void* tw, hipStream_t s);
template <typename T>
void launch_fwd_y(const Geo& g, const void* b2, void* b1, int nf, const void* tw, hipStream_t s);
template <typename T>
void launch_fwd_z(const Geo& g, const void* b1, void* out, void* out2, int nf, int nzero_origin, double scale,
                  const void* tw, hipStream_t s);

struct UpdateArgs {
  int mode;         // 0 = AB3 step, 1 = AB3 step with fm2 extrapolation (fix_fm2!), 3 = final rhs only
  double nu;        // kinematic viscosity (explicit term, AB3)
  int forcing;      // 0 none, 1 isotropic constant-power shell forcing
  double kf2;       // forced shell |k|^2 bound
};
template <typename T>
void launch_update_ab3(const Geo& g, void* u, void* rhs, const void* fm1, const void* fm2, const DevState* st,
                       const UpdateArgs& a, hipStream_t s);
template <typename T>
void launch_project(const Geo& g, void* u, hipStream_t s);

// stats: partial sums per block then a deterministic final reduction
template <typename T>
int launch_stats(const Geo& g, const void* u, const void* nl, const void* ke, double* partials, double* out,
                 hipStream_t s);
template <typename T>
void launch_spectrum(const Geo& g, const void* u, double* out_x, double* out_y0, double* out_y, double* out_z0,
                     double* out_z, double* partials, hipStream_t s);
template <typename T>
void launch_forcing_sum(const Geo& g, const void* u, const int* band, int nband, DevState* st, double power,
                        hipStream_t s);

void launch_set_dt_ab3(DevState* st, int it, int tstep_it, double cfl, double dx, hipStream_t s);
void launch_end_step(DevState* st, hipStream_t s);

template <typename T>
void launch_cast_copy(const void* src, void* dst, size_t n, hipStream_t s);

}  // namespace fsk
