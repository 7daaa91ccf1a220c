// Host driver of libffs_hip.so: one `ffs_sim` per run, owning its device
// buffers, twiddle tables, HIP stream and a queue of statistics records.
//
// The step follows advance_in_time! (models.jl:3-17):
//   calculate_rhs!  -> inv_z (+curl) / inv_y / fused x pass / fwd_y / fwd_z
//   set_dt!         -> one-thread kernel on device-resident dt (variable dt)
//   forcing         -> band reduction (isotropic shell forcing)
//   writestats      -> deterministic reductions, copied asynchronously
//   time_step!      -> fused update kernel, history rotated by pointer
#include "ffs_kernels.h"
#include "ffs/ffs.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

using namespace fsk;

struct FfsError : std::runtime_error {
  int code;
  FfsError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define HIPCHECK(x)                                                                                     \
  do {                                                                                                  \
    hipError_t e_ = (x);                                                                                \
    if (e_ != hipSuccess)                                                                               \
      throw FfsError(FFS_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_) + " (" + __FILE__ + ":" + \
                                      std::to_string(__LINE__) + ")");                                  \
  } while (0)

std::mutex g_err_mu;
std::string g_last_create_error;

// rfftfreq / fftfreq of Globals.jl:95-104 (index -> wavenumber)
std::vector<double> rfftfreq(int n, double s) {
  std::vector<double> v(n / 2 + 1);
  for (int i = 0; i <= n / 2; ++i) v[i] = i / s;
  return v;
}
std::vector<double> fftfreq(int n, double s) {
  std::vector<double> v(n);
  const int h = (n % 2 == 0) ? n / 2 : (n + 1) / 2;  // non-negative count
  for (int i = 0; i < n; ++i) v[i] = (i < h ? i : i - n) / s;
  return v;
}

template <typename T>
std::vector<T> twiddles(int n) {
  std::vector<T> t(2 * (size_t)n);
  const long double tau = 6.283185307179586476925286766559005768L;
  for (int m = 0; m < n; ++m) {
    const long double a = -tau * (long double)m / (long double)n;
    t[2 * m] = (T)cosl(a);
    t[2 * m + 1] = (T)sinl(a);
  }
  return t;
}

struct StatsSlot {
  double* red = nullptr;   // pinned: reduction results
  DevState* st = nullptr;  // pinned: state snapshot
  hipEvent_t ev = nullptr;
  int64_t iteration = 0;
  bool busy = false;
};

struct Record {
  std::vector<double> stats, scales, numbers;
};

}  // namespace

// for launchers whose pointers are the precision's real type
#define FFS_DISPATCH_R(fn, ptr, ...)                       \
  do {                                                     \
    if (fp32) fn<float>((float*)(ptr), __VA_ARGS__);       \
    else fn<double>((double*)(ptr), __VA_ARGS__);          \
  } while (0)

struct ffs_sim {
  ffs_params p{};
  bool fp32 = false;
  size_t csz = 16;  // bytes per complex element
  Geo g{};
  hipStream_t stream = nullptr;
  int device = 0;

  // grid (host)
  std::vector<int> ymap, zmap, yinv, zinv;
  std::vector<double> KX, KY, KZ;  // full FFT-order wavenumbers
  std::vector<double> kxc, kyc, kzc;
  double scale = 1.0;

  // device buffers
  std::vector<void*> allocs;
  size_t dev_bytes = 0;
  size_t S = 0;  // compact spectral elements per component
  void* u = nullptr;
  void* rhs = nullptr;
  void* fm1 = nullptr;
  void* fm2 = nullptr;
  void* ke = nullptr;
  void* tau = nullptr;   // Smagorinsky stress, 5 components
  void* flux = nullptr;  // F(-rho u), 3 components
  void* rho = nullptr;   // density
  void* rrhs = nullptr;  // density rhs / history (rotated like the velocity ones)
  void* rfm1 = nullptr;
  void* rfm2 = nullptr;
  void* phi = nullptr;    // passive scalar, its rhs / history and F(-phi u)
  void* prhs = nullptr;
  void* pfm1 = nullptr;
  void* pfm2 = nullptr;
  void* pflux = nullptr;
  void* uu = nullptr;     // conservation form: F(traceless(-u u)), 5 components
  bool les = false, dens = false, etd = false, pscal = false, nsform = false;
  int sles = 0, dles = 0;   // the passive scalar's and the density's LES flux (0 none, 1 eddy, 2 vorticity)
  double sles_cv = 0.0, dles_cv = 0.0;  // c Delta^2 of VorticityDiffusion
  int nin = 6, nfw = 3;     // physical inputs / forward outputs of the nonlinear term (without 1/2|u|^2)
  double les_alpha = 0.0;   // (c Delta)^2
  LesArgs lesm{};           // the LES model (lesm.model = p.les_model)
  // dynamicSmagorinsky: the stress is formed before the nonlinear pass (dynamic_les), through real
  // fields and Gaussian test filters on the full (not dealiased) spectrum
  bool dyn = false;
  bool pio = false;           // piomelliSmagorinsky: c and cm1 (DR_PC, DR_PCM1) and dtm1 persist across steps
  double* pio_dtm1 = nullptr;
  bool dynp = false;          // dynamicSandP / piomelliSandP: the P part (w, G w, P, cp)
  bool piop = false;          // piomelliSandP: cp and cpm1 (DR_CP, DR_CPM1) and dtpm1 persist
  double* pio_dtpm1 = nullptr;
  double dyn_td2 = 0.0, dyn_cmin = 0.0;
  Geo gf{};                   // every mode of the grid (full spectrum layout)
  void* dspec = nullptr;      // compact S (5), G S (5), G u (3)
  void* dreal = nullptr;      // padded real fields, kDynReal of them
  void* fspec = nullptr;      // one full spectrum
  void* fw1 = nullptr;        // its pencil scratch
  void* fw2 = nullptr;
  // slots of dreal: u, then S, M = G S, uf = G u in dspec's order, then Delta^2 |S| S (later tau), L, c
  // (Piomelli: its state c, cm1 and their copies around a stress read-out)
  // (the SandP variants: w, G w, P, cp, cpm1 and the copies of cp, cpm1)
  enum { DR_U = 0, DR_S = 3, DR_M = 8, DR_UF = 13, DR_SA = 16, DR_L = 21, DR_C = 27, DR_PC = 28, DR_PCM1 = 29,
         DR_SAVE = 30, DR_W = 32, DR_WF = 35, DR_PA = 38, DR_CP = 43, DR_CPM1 = 44, DR_SAVEP = 45, kDynReal = 47 };
  // gravity g = gval e_gdir (gravityDirection); gz = g.z, what calculate_N and spectrum_buoyancy read
  double alpha = 0.0, gz = 0.0, gval = 0.0, drhodz = 0.0;
  int gdir = 2;
  double palpha = 0.0;  // passive scalar diffusivity nu / scalarPr
  void* w1 = nullptr;
  void* w2 = nullptr;
  void *tw_x = nullptr, *tw_y = nullptr, *tw_z = nullptr;
  DevState* st = nullptr;
  double* partials = nullptr;
  double* red = nullptr;
  int* band = nullptr;
  int nband = 0;
  double kf2 = 0.0;
  // kz = 0-plane forcing (rf / qf)
  bool plane = false;
  PlaneForcing pfd{};
  std::vector<double> pf_avgk;    // mean |k_h| per 2D shell
  int pf_nsh = 0;
  std::vector<int> pf_slot_shell; // shell of each forced slot
  std::vector<int> pf_slot_mode;  // aForcing: (NX, NY) position of each forced mode
  bool pf_em_set = false;         // Em given before initialize (targSpectrum.dat)
  std::vector<double> pf_n_pending;  // N given before initialize (the N.N file), applied after its init

  // run state (host mirror of s.iteration[] and the integrator's iteration)
  int64_t iteration = 0;
  int64_t tit = 0;
  bool initialized = false;
  int64_t output_served = -1;
  double host_time = 0.0;

  // stats
  std::vector<StatsSlot> slots;
  std::deque<int> pending;
  std::deque<Record> ready;
  int nstats = 0;

  // spectra (write_spec / write_spec_forcing at the spectra iterations)
  struct SpecSlot {
    double* host = nullptr;  // pinned: kSpecQ marginal sets
    hipEvent_t ev = nullptr;
    int64_t iteration = 0;
    bool busy = false;
  };
  double* spec_part = nullptr;  // device: plane partials (the transform scratch when it is large enough)
  double* spec_out = nullptr;   // device: kSpecQ x spectra_out_len marginals
  std::vector<SpecSlot> spec_slots;
  std::deque<int> spec_pending;
  std::deque<std::pair<int64_t, std::vector<double>>> spec_ready;  // iteration, kSpecQ x (H[nh], V[nv])
  std::vector<double> spec_kh, spec_krz;

  // plain steps run the forward z pass and the update as one kernel (ffs_set_fusion; default FP64 only)
  bool fuse_ok = true;
  // the x pass's plan (ffs_set_plan): mirror-paired rows, wavefronts per block, split product launches
  int x_pair = 2, x_waves = 2, x_split = 1;
  // z slabs of the nonlinear term's inverse y, x and forward y passes (ffs_set_plan): planes per slab (0: the
  // whole grid in one launch each) and the slab buffers (their capacity in planes)
  int z_slab = 0, slab_cap = 0;
  void* sa2 = nullptr;  // inputs of the x pass for one slab: nin fields of (sx, ny, z_slab)
  void* sb2 = nullptr;  // its outputs: nfw + 1 fields of (sx, ny, z_slab)
  bool slab_own = false;  // sa2/sb2 allocated for themselves (else they lie in w2, after the forward y output)
  size_t w2_bytes = 0;
  // steady-state graphs: one per history-rotation phase
  bool graph_on = true;
  int phase = 0;  // history rotations mod 3
  hipGraph_t graphs[3] = {};
  hipGraphExec_t gexec[3] = {};
  // dry run: sizes only (ffs_footprint)
  bool dry = false;

  // profiling
  bool prof = false;
  struct ProfEv {
    std::string name;
    hipEvent_t a, b;
  };
  std::vector<ProfEv> prof_pending;
  std::vector<hipEvent_t> ev_pool;
  std::map<std::string, std::pair<double, int64_t>> prof_acc;
  std::vector<std::string> prof_names;

  std::string err;

  void* dalloc(size_t bytes) {
    void* ptr = nullptr;
    if (bytes == 0) bytes = 256;
    if (dry) {  // footprint only: distinct non-null placeholders, nothing allocated
      dev_bytes += bytes;
      return (void*)(uintptr_t)(dev_bytes);
    }
    hipError_t e = hipMalloc(&ptr, bytes);
    if (e != hipSuccess)
      throw FfsError(FFS_ERR_MEMORY, "hipMalloc of " + std::to_string(bytes) + " bytes failed: " + hipGetErrorString(e));
    allocs.push_back(ptr);
    dev_bytes += bytes;
    return ptr;
  }

  void dfree(void* ptr) {
    if (!ptr || dry) return;
    for (size_t i = 0; i < allocs.size(); ++i)
      if (allocs[i] == ptr) {
        hipFree(ptr);
        allocs.erase(allocs.begin() + i);
        return;
      }
  }

  template <typename V>
  V* upload(const std::vector<V>& v) {
    V* d = (V*)dalloc(v.size() * sizeof(V));
    if (dry) return d;
    HIPCHECK(hipMemcpy(d, v.data(), v.size() * sizeof(V), hipMemcpyHostToDevice));
    return d;
  }

  hipEvent_t get_ev() {
    if (!ev_pool.empty()) {
      hipEvent_t e = ev_pool.back();
      ev_pool.pop_back();
      return e;
    }
    hipEvent_t e;
    HIPCHECK(hipEventCreate(&e));
    return e;
  }

  // a launch that the runtime rejects (bad configuration) surfaces here, not as silently skipped work
  static void check_launch(const char* name) {
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) throw FfsError(FFS_ERR_HIP, std::string("launch of ") + name + ": " + hipGetErrorString(e));
  }

  template <class F>
  void launch(const char* name, F&& f) {
    if (!prof) {
      f();
      check_launch(name);
      return;
    }
    hipEvent_t a = get_ev(), b = get_ev();
    HIPCHECK(hipEventRecord(a, stream));
    f();
    check_launch(name);
    HIPCHECK(hipEventRecord(b, stream));
    prof_pending.push_back({name, a, b});
    if (prof_pending.size() > 4096) drain_prof();
  }

  void drain_prof() {
    for (auto& e : prof_pending) {
      HIPCHECK(hipEventSynchronize(e.b));
      float ms = 0;
      HIPCHECK(hipEventElapsedTime(&ms, e.a, e.b));
      auto& acc = prof_acc[e.name];
      acc.first += ms;
      acc.second += 1;
      ev_pool.push_back(e.a);
      ev_pool.push_back(e.b);
    }
    prof_pending.clear();
  }

  // ---------------------------------------------------------------- setup
  void setup_grid() {
    const int nrx = p.nx, ny = p.ny, nz = p.nz;
    if (nrx <= 0 || ny <= 0 || nz <= 0) throw FfsError(FFS_ERR_ARG, "nx, ny, nz must be positive");
    if (nrx % 2) throw FfsError(FFS_ERR_ARG, "nx must be even");
    const int nxh = nrx / 2 + 1;
    KX = rfftfreq(nrx, p.x_domain);
    KY = fftfreq(ny, p.y_domain);
    KZ = fftfreq(nz, p.z_domain);
    // DEALIAS = (kx^2 > (c kx_max)^2) | (ky^2 > (c max|ky|)^2) | (kz^2 > (c max|kz|)^2)   Globals.jl:132-142
    double mky = 0, mkz = 0;
    for (double v : KY) mky = std::max(mky, std::fabs(v));
    for (double v : KZ) mkz = std::max(mkz, std::fabs(v));
    const double cx2 = std::pow(p.cutoff * KX.back(), 2);
    const double cy2 = std::pow(p.cutoff * mky, 2);
    const double cz2 = std::pow(p.cutoff * mkz, 2);
    int mx = 0;
    while (mx < nxh && !(KX[mx] * KX[mx] > cx2)) ++mx;
    ymap.clear();
    zmap.clear();
    yinv.assign(ny, -1);
    zinv.assign(nz, -1);
    for (int j = 0; j < ny; ++j)
      if (!(KY[j] * KY[j] > cy2)) {
        yinv[j] = (int)ymap.size();
        ymap.push_back(j);
      }
    for (int k = 0; k < nz; ++k)
      if (!(KZ[k] * KZ[k] > cz2)) {
        zinv[k] = (int)zmap.size();
        zmap.push_back(k);
      }
    if (mx == 0 || ymap.empty() || zmap.empty()) throw FfsError(FFS_ERR_ARG, "dealiasing removes every mode");
    kxc.assign(KX.begin(), KX.begin() + mx);
    kyc.resize(ymap.size());
    kzc.resize(zmap.size());
    for (size_t j = 0; j < ymap.size(); ++j) kyc[j] = KY[ymap[j]];
    for (size_t k = 0; k < zmap.size(); ++k) kzc[k] = KZ[zmap[k]];
    g.nrx = nrx;
    g.ny = ny;
    g.nz = nz;
    g.nxh = nxh;
    g.mx = mx;
    g.my = (int)ymap.size();
    g.mz = (int)zmap.size();
    // the kernels' arithmetic maps (two contiguous retained blocks per axis, k = index / L)
    g.fy = 0;
    while (g.fy < g.my && ymap[g.fy] == g.fy) ++g.fy;
    g.fz = 0;
    while (g.fz < g.mz && zmap[g.fz] == g.fz) ++g.fz;
    g.lx = p.x_domain;
    g.ly = p.y_domain;
    g.lz = p.z_domain;
    for (int j = 0; j < ny; ++j)
      if (y_compact(g, j) != yinv[j]) throw FfsError(FFS_ERR_ARG, "retained y modes are not two contiguous blocks");
    for (int k = 0; k < nz; ++k)
      if (z_compact(g, k) != zinv[k]) throw FfsError(FFS_ERR_ARG, "retained z modes are not two contiguous blocks");
    for (int c = 0; c < mx; ++c)
      if (kx_of(g, c) != kxc[c]) throw FfsError(FFS_ERR_ARG, "kx map mismatch");
    for (int c = 0; c < g.my; ++c)
      if (ky_of(g, c) != kyc[c]) throw FfsError(FFS_ERR_ARG, "ky map mismatch");
    for (int c = 0; c < g.mz; ++c)
      if (kz_of(g, c) != kzc[c]) throw FfsError(FFS_ERR_ARG, "kz map mismatch");
    const int line = fp32 ? 16 : 8;  // complex values per 128-byte line
    g.sx = (mx + line - 1) / line * line;
    g.ymap = upload(ymap);
    g.zmap = upload(zmap);
    g.yinv = upload(yinv);
    g.zinv = upload(zinv);
    g.kx = upload(kxc);
    g.ky = upload(kyc);
    g.kz = upload(kzc);
    scale = 1.0 / ((double)nrx * ny * nz);
    S = (size_t)g.sx * g.my * g.mz;
    if (dyn) {  // every mode kept: identity maps, the full rfftfreq / fftfreq wavenumbers
      gf = g;
      gf.mx = nxh;
      gf.my = gf.fy = ny;
      gf.mz = gf.fz = nz;
      gf.sx = (nxh + line - 1) / line * line;
      std::vector<int> iy(ny), iz(nz);
      for (int j = 0; j < ny; ++j) iy[j] = j;
      for (int k = 0; k < nz; ++k) iz[k] = k;
      gf.ymap = gf.yinv = upload(iy);
      gf.zmap = gf.zinv = upload(iz);
      gf.kx = upload(KX);
      gf.ky = upload(KY);
      gf.kz = upload(KZ);
    }
  }

  template <typename T>
  void* tw_table(int n) {
    return upload(twiddles<T>(n));
  }

  void setup_buffers() {
    csz = fp32 ? 8 : 16;
    u = dalloc(3 * S * csz);
    rhs = dalloc(3 * S * csz);
    fm1 = dalloc(3 * S * csz);
    fm2 = dalloc(3 * S * csz);
    ke = dalloc(S * csz);
    if (!dry) {
      HIPCHECK(hipMemset(u, 0, 3 * S * csz));
      HIPCHECK(hipMemset(rhs, 0, 3 * S * csz));
      HIPCHECK(hipMemset(fm1, 0, 3 * S * csz));
      HIPCHECK(hipMemset(fm2, 0, 3 * S * csz));
      HIPCHECK(hipMemset(ke, 0, S * csz));
    }
    if (les) {
      tau = dalloc(5 * S * csz);
      if (!dry) HIPCHECK(hipMemset(tau, 0, 5 * S * csz));
    }
    if (dyn) {
      const size_t nf = (size_t)gf.sx * gf.ny * gf.nz;
      dspec = dalloc(dyn_spec_fields() * S * csz);
      dreal = dalloc(dyn_real_fields() * padded_len() * (csz / 2));
      fspec = dalloc(nf * csz);
      fw1 = dalloc(nf * csz);
      fw2 = dalloc(nf * csz);
      if (!dry) HIPCHECK(hipMemset(dspec, 0, dyn_spec_fields() * S * csz));
      if (pio) {  // c = cm1 = 0.0289, dtm1 = 0 (PiomelliSmagorinsky, les_types.jl:296-300)
        pio_dtm1 = (double*)dalloc(sizeof(double));
        if (!dry) {
          HIPCHECK(hipMemset(pio_dtm1, 0, sizeof(double)));
          FFS_DISPATCH_R(launch_fill_real, rptr(dreal, DR_PC), 2 * padded_len(), 0.0289, nullptr);
          HIPCHECK(hipDeviceSynchronize());
        }
      }
      if (piop) {  // cp = cpm1 = 0.09, dtpm1 = 0 (PiomelliSandP, les_types.jl:320-329)
        pio_dtpm1 = (double*)dalloc(sizeof(double));
        if (!dry) {
          HIPCHECK(hipMemset(pio_dtpm1, 0, sizeof(double)));
          FFS_DISPATCH_R(launch_fill_real, rptr(dreal, DR_CP), 2 * padded_len(), 0.09, nullptr);
          HIPCHECK(hipDeviceSynchronize());
        }
      }
    }
    if (dens) {
      flux = dalloc(3 * S * csz);
      rho = dalloc(S * csz);
      rrhs = dalloc(S * csz);
      rfm1 = dalloc(S * csz);
      rfm2 = dalloc(S * csz);
      if (!dry)
        for (void* b : {flux, rho, rrhs, rfm1, rfm2}) HIPCHECK(hipMemset(b, 0, (b == flux ? 3 : 1) * S * csz));
    }
    if (nsform) {
      uu = dalloc(5 * S * csz);
      if (!dry) HIPCHECK(hipMemset(uu, 0, 5 * S * csz));
    }
    if (pscal) {
      pflux = dalloc(3 * S * csz);
      phi = dalloc(S * csz);
      prhs = dalloc(S * csz);
      pfm1 = dalloc(S * csz);
      pfm2 = dalloc(S * csz);
      if (!dry)
        for (void* b : {pflux, phi, prhs, pfm1, pfm2}) HIPCHECK(hipMemset(b, 0, (b == pflux ? 3 : 1) * S * csz));
    }
    const size_t nmax = (size_t)std::max(nin, nfw + 1);
    const size_t a1 = nmax * g.sx * g.my * g.nz, b2 = nmax * g.sx * g.ny * g.nz;
    const size_t a2 = nmax * g.sx * g.ny * g.nz, b1 = nmax * g.sx * g.my * g.nz;
    w1 = dalloc(std::max(a1, b2) * csz);
    w2 = dalloc(std::max(a2, b1) * csz);
    w2_bytes = std::max(a2, b1) * csz;
    z_slab = default_slab();
    alloc_slabs(z_slab);
    // the spectra partials live in w1 (free between the nonlinear term and the next step) when it fits
    const size_t spec_bytes = spectra_partials_len(g) * sizeof(double);
    spec_part = std::max(a1, b2) * csz >= spec_bytes ? (double*)w1 : (double*)dalloc(spec_bytes);
    spec_out = (double*)dalloc((size_t)kSpecQ * spectra_out_len(g) * sizeof(double));
    if (fp32) {
      tw_x = tw_table<float>(g.nrx);
      tw_y = tw_table<float>(g.ny);
      tw_z = tw_table<float>(g.nz);
    } else {
      tw_x = tw_table<double>(g.nrx);
      tw_y = tw_table<double>(g.ny);
      tw_z = tw_table<double>(g.nz);
    }
    st = (DevState*)dalloc(sizeof(DevState));
    const size_t npart = std::max<size_t>((size_t)512 * kStatsSlots,
                                          (size_t)g.mz * 2 * (g.mx + g.my) + (size_t)g.mz * 4 + 64);
    partials = (double*)dalloc(npart * sizeof(double));
    red = (double*)dalloc(64 * sizeof(double) + 4 * (size_t)(g.mx + g.my) * sizeof(double));
    // forcing band (0 < |k|^2 <= kf^2), listed in index order
    if (p.forcing == FFS_FORCING_SHELL) {
      kf2 = p.forcing_kf * p.forcing_kf;
      std::vector<int> b;
      for (int cz = 0; cz < g.mz; ++cz)
        for (int cy = 0; cy < g.my; ++cy)
          for (int cx = 0; cx < g.mx; ++cx) {
            const double K2 = kxc[cx] * kxc[cx] + kyc[cy] * kyc[cy] + kzc[cz] * kzc[cz];
            if (K2 > 0 && K2 <= kf2) b.push_back(cx + g.sx * (cy + g.my * cz));
          }
      nband = (int)b.size();
      if (nband) band = upload(b);
    }
    if (plane) setup_plane_forcing();
  }

  // compute_shells2D (utils.jl:161-182), calculate_Zf (utils.jl:184-197 and 199-218 for aForcing),
  // the forced elements of the kz = 0 plane and the vertical modes, and for QfForcing
  // build_Qf_power_spectrum! (forcing_types.jl:304-331)
  void setup_plane_forcing() {
    const int NX = g.nxh, NY = g.ny;
    const int nsh = std::min(NX, NY / 2);
    const double maxdk = std::max(KX[1], KY[1]);
    std::vector<double> kh(nsh, 0.0);
    std::vector<int> cnt(nsh, 0);
    for (int j = 0; j < NY; ++j)
      for (int i = 0; i < NX; ++i) {
        const double K = std::sqrt(KX[i] * KX[i] + KY[j] * KY[j]);
        const int ii = (int)std::nearbyint(K / maxdk);
        if (ii >= nsh) break;
        kh[ii] += K;
        cnt[ii] += 1;
      }
    for (int i = 0; i < nsh; ++i) kh[i] = kh[i] / cnt[i];  // 0/0 = NaN for an empty shell, as the reference
    pf_avgk = kh;
    pf_nsh = nsh;
    const double kf = p.forcing_kf;
    auto zf_of = [&](double k) {
      return p.forcing_zf_tanh ? std::tanh((kf - k) / (0.25 * kf)) * (((kf - k) <= 0.0) ? 0.0 : 1.0)
                               : (k <= kf ? 1.0 : 0.0);
    };
    const bool per_mode = p.forcing == FFS_FORCING_A;
    std::vector<int> elem, eslot;
    std::vector<double> ew, ek2, zf;
    pf_slot_shell.clear();
    pf_slot_mode.clear();
    if (!per_mode) {
      // forced shells: 1-based n in 2..nsh with mean |k_h| <= kf
      std::vector<int> slot_of(nsh, -1);
      for (int nn = 1; nn < nsh; ++nn)
        if (kh[nn] <= kf) {
          slot_of[nn] = (int)pf_slot_shell.size();
          pf_slot_shell.push_back(nn);
          zf.push_back(zf_of(kh[nn]));
        }
      for (int cy = 0; cy < g.my; ++cy)
        for (int cx = 0; cx < g.mx; ++cx) {
          const double k = std::sqrt(std::fma(kxc[cx], kxc[cx], kyc[cy] * kyc[cy]));
          const int nn = (int)std::nearbyint(k / maxdk);
          if (nn >= 1 && nn < nsh && slot_of[nn] >= 0) {
            elem.push_back(cx + g.sx * cy);
            eslot.push_back(slot_of[nn]);
            ew.push_back(cx == 0 ? 1.0 : 2.0);
            ek2.push_back(kxc[cx] * kxc[cx] + kyc[cy] * kyc[cy]);
          }
        }
    } else {
      // one oscillator per retained kz = 0 mode with 0 < |k_h| <= kf (aforcing.jl:25-36)
      for (int cy = 0; cy < g.my; ++cy)
        for (int cx = 0; cx < g.mx; ++cx) {
          const double k = std::sqrt(std::fma(kxc[cx], kxc[cx], kyc[cy] * kyc[cy]));
          if (0 < k && k <= kf) {
            elem.push_back(cx + g.sx * cy);
            eslot.push_back((int)pf_slot_mode.size());
            pf_slot_mode.push_back(cx + NX * ymap[cy]);  // position in the reference's (NX, NY) matrices
            ew.push_back(cx == 0 ? 1.0 : 2.0);
            ek2.push_back(kxc[cx] * kxc[cx] + kyc[cy] * kyc[cy]);
            // calculate_Zf(d, kf, KX, KY): tanh only when asked, else the cutoff
            zf.push_back(p.forcing_zf_tanh ? std::tanh((kf - k) / (0.25 * kf)) * (((kf - k) <= 0.0) ? 0.0 : 1.0)
                                           : (k <= kf ? 1.0 : 0.0));
          }
        }
    }
    const int nslot = (int)zf.size();
    if (nslot == 0) throw FfsError(FFS_ERR_ARG, "no kz = 0 mode or 2D shell is forced (kf too small)");
    // vertical modes kx = ky = 0: (kz = +m, kz = -m) for m = m0..5
    const int m0 = p.forcing == FFS_FORCING_QF ? 3 : 1;
    std::vector<int> vert;
    for (int m = m0; m <= 5; ++m)
      for (int sgn : {1, -1}) {
        const int cz = z_compact(g, sgn > 0 ? m : g.nz - m);
        if (cz < 0) throw FfsError(FFS_ERR_ARG, "plane forcing needs the modes |kz| <= 5 (nz >= 16)");
        vert.push_back(g.sx * g.my * cz);
      }
    std::vector<double> R(nslot, 0.0), zeros(nslot, 0.0);
    if (p.forcing == FFS_FORCING_QF) {
      double integral = 0.0;
      for (int sl = 0; sl < nslot; ++sl) {
        const double k = kh[pf_slot_shell[sl]];
        const double term = std::exp(-std::pow((k - p.alvelius_kf) / p.alvelius_c, 2));
        R[sl] = p.alvelius_p * term;
        integral += term;
      }
      for (double& r : R) r /= integral;
      if (les) {
        const double delta = p.filter_width > 0 ? p.filter_width : (p.x_domain * 2 * M_PI / p.nx) / p.cutoff;
        for (int sl = 0; sl < nslot; ++sl) {
          const double k = kh[pf_slot_shell[sl]];
          const double G = std::exp(-(delta * delta) / 24 * (k * k));  // Gaussfilter, utils.jl:220-223
          R[sl] *= G * G;
        }
      }
    }
    auto one = [](auto v) { if (v.empty()) v.resize(1); return v; };
    pfd.kind = p.forcing;
    pfd.nel = (int)elem.size();
    pfd.elem = upload(one(elem));
    pfd.eslot = upload(one(eslot));
    pfd.ew = upload(one(ew));
    pfd.ek2 = upload(one(ek2));
    // per-slot element index (counting sort keeps the list order within a slot)
    std::vector<int> sptr(nslot + 1, 0), sel(elem.size());
    for (int sl : eslot) sptr[sl + 1] += 1;
    for (int i = 0; i < nslot; ++i) sptr[i + 1] += sptr[i];
    {
      std::vector<int> fill(sptr.begin(), sptr.end() - 1);
      for (int i = 0; i < (int)eslot.size(); ++i) sel[fill[eslot[i]]++] = i;
    }
    pfd.slot_ptr = upload(sptr);
    pfd.slot_el = upload(one(sel));
    // every shell's retained kz = 0 modes, in the reference's loop order (y outer, x inner)
    if (!per_mode) {
      std::vector<std::vector<int>> by(nsh);
      for (int cy = 0; cy < g.my; ++cy)
        for (int cx = 0; cx < g.mx; ++cx) {
          const double k = std::sqrt(kxc[cx] * kxc[cx] + kyc[cy] * kyc[cy]);
          const int nn = (int)std::nearbyint(k / maxdk);
          if (nn < nsh) by[nn].push_back(cx + g.sx * cy);
        }
      std::vector<int> hptr(nsh + 1, 0), hel;
      for (int nn = 0; nn < nsh; ++nn) {
        hel.insert(hel.end(), by[nn].begin(), by[nn].end());
        hptr[nn + 1] = (int)hel.size();
      }
      pfd.nshell = nsh;
      pfd.shell_ptr = upload(hptr);
      pfd.shell_el = upload(one(hel));
      pfd.Ef_all = upload(std::vector<double>(nsh, 0.0));
    } else {
      pfd.nshell = 0;
    }
    pfd.nslot = nslot;
    pfd.R = upload(R);
    pfd.Em = upload(zeros);
    pfd.Ef = upload(zeros);
    pfd.N = upload(zeros);
    pfd.Nm1 = upload(zeros);
    pfd.Zf = upload(zf);
    pfd.slot_factor = upload(zeros);
    pfd.plane_factor = upload(std::vector<double>((size_t)g.sx * g.my, 0.0));
    pfd.nvert = (int)vert.size();
    pfd.vert = upload(vert);
    pfd.tf = p.forcing_tf;
    pfd.alpha = p.forcing_alpha;
    pfd.cfl = p.cfl;
    pfd.nu = p.nu;
    pfd.variable_dt = p.variable_dt;
  }

  // per-slot state <-> the reference's arrays (per 2D shell; aForcing: per (NX, NY) mode, x fastest)
  int pf_state_len() const { return p.forcing == FFS_FORCING_A ? g.nxh * g.ny : pf_nsh; }
  int pf_state_pos(int sl) const { return p.forcing == FFS_FORCING_A ? pf_slot_mode[sl] : pf_slot_shell[sl]; }

  void setup_state() {
    DevState h{};
    h.dt = h.dt2 = h.dt3 = p.dt;
    h.time = p.start_time;
    host_step_coefficients(&h);
    HIPCHECK(hipMemcpy(st, &h, sizeof(h), hipMemcpyHostToDevice));
    host_time = p.start_time;
    iteration = p.start;
    tit = 0;
  }

  void create(const ffs_params& in) {
    p = in;
    if (p.size != sizeof(ffs_params)) throw FfsError(FFS_ERR_ARG, "ffs_params.size mismatch (ABI version)");
    if (p.integrator != FFS_AB3 && p.integrator != FFS_ETD3 && p.integrator != FFS_EULER)
      throw FfsError(FFS_ERR_ARG, "unknown integrator");
    if (p.les_model < FFS_LES_NONE || p.les_model > FFS_LES_PIOMELLI_SANDP)
      throw FfsError(FFS_ERR_ARG, "unknown LES model");
    if (p.les_model == FFS_LES_FAKE_SMAGORINSKY)
      throw FfsError(FFS_ERR_ARG, "fakeSmagorinsky is built only as the S model of SandP");
    if (p.les_model == FFS_LES_SANDP &&
        p.les_s_model != FFS_LES_SMAGORINSKY && p.les_s_model != FFS_LES_VREMAN &&
        p.les_s_model != FFS_LES_PRODUCTION && p.les_s_model != FFS_LES_FAKE_SMAGORINSKY)
      throw FfsError(FFS_ERR_ARG, "SandP's sLesModel: Smagorinsky, Vreman, productionViscosity or fakeSmagorinsky");
    if ((p.les_model == FFS_LES_DYNAMIC_SANDP || p.les_model == FFS_LES_PIOMELLI_SANDP) &&
        p.les_s_model != FFS_LES_DYNAMIC_SMAGORINSKY && p.les_s_model != FFS_LES_PIOMELLI_SMAGORINSKY)
      throw FfsError(FFS_ERR_ARG, "the dynamic SandP models' sLesModel: dynamicSmagorinsky or piomelliSmagorinsky");
    if (p.forcing < FFS_FORCING_NONE || p.forcing > FFS_FORCING_A) throw FfsError(FFS_ERR_ARG, "unsupported forcing");
    plane = p.forcing >= FFS_FORCING_RF && p.forcing <= FFS_FORCING_A;
    if (plane && p.forcing != FFS_FORCING_QF && !(p.forcing_tf > 0))
      throw FfsError(FFS_ERR_ARG, "the oscillator forcing schemes need TF > 0");
    if (p.forcing == FFS_FORCING_QF && !(p.alvelius_c != 0))
      throw FfsError(FFS_ERR_ARG, "QfForcing needs alveliusC != 0");
    if (p.precision != FFS_FP64 && p.precision != FFS_FP32) throw FfsError(FFS_ERR_ARG, "unknown precision");
    if (p.density_hv < FFS_HV_NONE || p.density_hv > FFS_HV_BARRIER)
      throw FfsError(FFS_ERR_ARG, "unknown densityHyperViscosity");
    if (p.density_hv == FFS_HV_HYPER && p.density_hv_m < 1)
      throw FfsError(FFS_ERR_ARG, "densityHyperViscosityM must be >= 1");
    if (!p.density) p.density_hv = FFS_HV_NONE;  // no density, nothing to act on
    if (p.hyperviscosity < FFS_HV_NONE || p.hyperviscosity > FFS_HV_BARRIER)
      throw FfsError(FFS_ERR_ARG, "unknown hyperviscosity");
    if (p.hyperviscosity == FFS_HV_HYPER && p.hv_m < 1) throw FfsError(FFS_ERR_ARG, "hyperViscosityM must be >= 1");
    if (p.hyperviscosity == FFS_HV_BARRIER && p.integrator != FFS_ETD3)
      // the AB3 right-hand side asks nuh() of the hyperviscosity type, which SpectralBarrier lacks
      // (models.jl:298-300, hyperviscosity_types.jl:13-14)
      throw FfsError(FFS_ERR_ARG, "spectralBarrier needs velocityTimeStep ETD3rdO");
    etd = (p.integrator == FFS_ETD3);
    les = (p.les_model != FFS_LES_NONE);
    if (std::isnan(p.smagorinsky_c)) {  // the constants' defaults (les_types.jl:364-404)
      const int m = p.les_model == FFS_LES_SANDP ? p.les_s_model : p.les_model;
      p.smagorinsky_c = m == FFS_LES_VREMAN ? 0.032
                        : m == FFS_LES_SILVIS ? std::sqrt(0.17)
                        : m == FFS_LES_PRODUCTION ? (p.les_model == FFS_LES_SANDP ? 0.1 : 0.4)
                        : m == FFS_LES_STABLE_NL ? 0.1
                        : m == FFS_LES_FAKE_SMAGORINSKY ? 0.0
                                                        : 0.17;
    }
    if (std::isnan(p.les_cp)) p.les_cp = p.les_model == FFS_LES_SILVIS ? 5.0 : (p.les_model == FFS_LES_SANDP ? 0.09 : 0.0);
    dens = (p.density != 0);
    pscal = (p.passive_scalar != 0);
    if (p.equation != FFS_EQ_ROTATIONAL && p.equation != FFS_EQ_NS) throw FfsError(FFS_ERR_ARG, "unknown equation");
    nsform = (p.equation == FFS_EQ_NS);
    if (pscal) {
      // with an LES model the reference gives the scalar an eddy-diffusion flux (simulationtypes.jl:335-348)
      // with an LES model the reference gives the scalar an LES flux, EddyDiffusion unless
      // scalarLesModel is vorticityDiffusion (simulationtypes.jl:334-346)
      if (les) sles = p.scalar_les == 2 ? 2 : 1;
      if (!(p.scalar_pr > 0)) throw FfsError(FFS_ERR_ARG, "scalarPr must be > 0 (Inf: no diffusion)");
      if (p.scalar_direction < 0 || p.scalar_direction > 2) throw FfsError(FFS_ERR_ARG, "scalarDirection is x, y or z");
      palpha = p.nu / p.scalar_pr;  // nu / Inf = 0
    }
    if (les) {
      // Delta default (2 pi lx / nx) / cutoff, twice that for SandP's Smagorinsky (les_types.jl:366-401)
      const bool twice = p.les_model == FFS_LES_SANDP && p.les_s_model == FFS_LES_SMAGORINSKY;
      const double delta = p.filter_width > 0 ? p.filter_width
                                              : (twice ? 2.0 : 1.0) * (p.x_domain * 2 * M_PI / p.nx) / p.cutoff;
      les_alpha = p.smagorinsky_c * p.smagorinsky_c * delta * delta;
      lesm.model = p.les_model;
      lesm.smodel = p.les_s_model;
      lesm.c = p.smagorinsky_c;
      lesm.cp = p.les_cp;
      lesm.d2 = delta * delta;
      dynp = p.les_model == FFS_LES_DYNAMIC_SANDP || p.les_model == FFS_LES_PIOMELLI_SANDP;
      piop = p.les_model == FFS_LES_PIOMELLI_SANDP;
      pio = p.les_model == FFS_LES_PIOMELLI_SMAGORINSKY || (dynp && p.les_s_model == FFS_LES_PIOMELLI_SMAGORINSKY);
      dyn = p.les_model == FFS_LES_DYNAMIC_SMAGORINSKY || pio || dynp;
      if (dynp && p.equation == FFS_EQ_NS)  // P reads the vorticity the rotational form's rhs holds
        throw FfsError(FFS_ERR_ARG, "the dynamic SandP models need equation = rotational here");
      if (dyn) {  // DynamicSmagorinsky(Delta, TestFilterWidth = 2 Delta, dims, cmin) (les_types.jl:383-387)
        const double td = p.les_test_width > 0 ? p.les_test_width : 2 * delta;
        dyn_td2 = td * td;
        dyn_cmin = p.les_cmin;
      }
      // the models that read w take it from the rotational form's curl pass; the reference also forms w
      // for NS with LES (need_vorticity, simulationtypes.jl:41-42), which this build does not implement
      if (les_needs_w(lesm) && p.equation == FFS_EQ_NS)
        throw FfsError(FFS_ERR_ARG, "LES models other than Smagorinsky with equation = NS are not implemented");
    }
    if (dens) {
      if (!(p.pr > 0) || !(p.reference_density != 0))
        throw FfsError(FFS_ERR_ARG, "density needs Pr > 0 and referenceDensity != 0");
      alpha = p.nu / p.pr;
      if (p.gravity_direction < 0 || p.gravity_direction > 2)
        throw FfsError(FFS_ERR_ARG, "gravityDirection is x, y or z");
      gdir = p.gravity_direction;
      gval = p.z_acceleration / p.reference_density;
      gz = gdir == 2 ? gval : 0.0;
      drhodz = p.density_gradient;
    }
    if (dens && p.density_les) {  // densityLesModel (density_stratification_types.jl:176-184)
      if (p.density_les < 0 || p.density_les > 2) throw FfsError(FFS_ERR_ARG, "unknown densityLesModel");
      if (!les) throw FfsError(FFS_ERR_ARG, "densityLesModel needs an LES model");
      dles = p.density_les;
    }
    if (sles || dles) {
      if (dyn) throw FfsError(FFS_ERR_ARG, "the scalars' LES fluxes are built for the pointwise LES models only");
      // this build forms w only in the rotational form's curl pass (the reference also forms it for NS
      // with LES, need_vorticity, simulationtypes.jl:41-42)
      if (nsform && (sles == 2 || dles == 2))
        throw FfsError(FFS_ERR_ARG, "vorticityDiffusion with equation = NS is not implemented (use rotational)");
      sles_cv = (std::isnan(p.scalar_les_c) ? 1.0 : p.scalar_les_c) * lesm.d2;
      dles_cv = (std::isnan(p.density_les_c) ? 0.1 : p.density_les_c) * lesm.d2;
    }
    // the x pass carries S and tau for the pointwise models; the dynamic model forms tau beforehand
    const bool xles = les && !dyn;
    nin = (nsform ? 3 : 6) + (dens ? 1 : 0) + (pscal ? 1 : 0) + (xles ? 5 : 0) + 3 * ((dles ? 1 : 0) + (sles ? 1 : 0));
    nfw = (nsform ? 5 : 3) + (dens ? 3 : 0) + (pscal ? 3 : 0) + (xles ? 5 : 0);
    if (p.dt_stat <= 0 || p.dt_spec <= 0 || p.write_time <= 0)
      throw FfsError(FFS_ERR_ARG, "dtStat, dtSpec and writeTime must be positive");
    if (!p.variable_dt && !(p.dt > 0)) throw FfsError(FFS_ERR_ARG, "a fixed time step needs dt > 0");
    fp32 = (p.precision == FFS_FP32);
    // the FP32 tile (16 columns x 32 threads) holds too many registers to fuse profitably, and the fused tile
    // of a z length that is not a power of two, or longer than 512, is not register-resident (private memory or
    // spills): those runs default to the two kernels (ffs_set_fusion / the plan can still choose fusion)
    fuse_ok = !fp32 && (p.nz & (p.nz - 1)) == 0 && p.nz <= 512;
    check_lengths();
    if (dry) {  // footprint: the same allocations, counted, plus the transient staging of set_field
      setup_grid();
      setup_buffers();
      const size_t padded = 2 * (size_t)(p.nx / 2 + 1) * p.ny * p.nz;
      dev_bytes += padded * (sizeof(double) + (fp32 ? sizeof(float) : 0));
      return;
    }
    HIPCHECK(hipGetDevice(&device));
    HIPCHECK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    setup_grid();
    setup_buffers();
    setup_state();
    slots.resize(64);
    for (auto& sl : slots) {
      HIPCHECK(hipHostMalloc((void**)&sl.red, 64 * sizeof(double), hipHostMallocDefault));
      HIPCHECK(hipHostMalloc((void**)&sl.st, sizeof(DevState), hipHostMallocDefault));
      HIPCHECK(hipEventCreateWithFlags(&sl.ev, hipEventDisableTiming));
    }
  }

  void drop_graphs() {
    for (int i = 0; i < 3; ++i) {
      if (gexec[i]) hipGraphExecDestroy(gexec[i]);
      if (graphs[i]) hipGraphDestroy(graphs[i]);
      gexec[i] = nullptr;
      graphs[i] = nullptr;
    }
  }

  void destroy() {
    if (dry) {
      allocs.clear();
      return;
    }
    if (stream) hipStreamSynchronize(stream);
    drop_graphs();
    for (auto& e : prof_pending) {
      hipEventDestroy(e.a);
      hipEventDestroy(e.b);
    }
    for (auto e : ev_pool) hipEventDestroy(e);
    for (auto& sl : slots) {
      if (sl.red) hipHostFree(sl.red);
      if (sl.st) hipHostFree(sl.st);
      if (sl.ev) hipEventDestroy(sl.ev);
    }
    for (auto& sl : spec_slots) {
      if (sl.host) hipHostFree(sl.host);
      if (sl.ev) hipEventDestroy(sl.ev);
    }
    for (void* a : allocs) hipFree(a);
    allocs.clear();
    if (stream) hipStreamDestroy(stream);
    stream = nullptr;
  }

  // ---------------------------------------------------------------- transforms
#define FFS_DISPATCH(fn, ...)           \
  do {                                  \
    if (fp32) fn<float>(__VA_ARGS__);   \
    else fn<double>(__VA_ARGS__);       \
  } while (0)

  char* cptr(void* base, size_t elems) { return (char*)base + elems * csz; }

  // compact spectral (nf components at `spec`) -> padded real (nf fields at `real`, precision T)
  void inverse_to_real(void* spec, int nf, void* real) {
    launch("inv_z", [&] { FFS_DISPATCH(launch_inv_z, g, spec, w1, nf, false, tw_z, stream); });
    launch("inv_y", [&] { FFS_DISPATCH(launch_inv_y, g, w1, w2, nf, tw_y, stream); });
    if (fp32) launch("x_c2r", [&] { launch_x_c2r<float>(g, w2, (float*)real, nf, tw_x, stream); });
    else launch("x_c2r", [&] { launch_x_c2r<double>(g, w2, (double*)real, nf, tw_x, stream); });
  }

  // padded real -> compact spectral, dealiased and scaled (myfourier!, utils.jl:101-106)
  void forward_from_real(void* real, int nf, void* spec) {
    if (fp32) launch("x_r2c", [&] { launch_x_r2c<float>(g, (const float*)real, w1, nf, tw_x, stream); });
    else launch("x_r2c", [&] { launch_x_r2c<double>(g, (const double*)real, w1, nf, tw_x, stream); });
    launch("fwd_y", [&] { FFS_DISPATCH(launch_fwd_y, g, w1, w2, nf, tw_y, stream); });
    void* dst[kMaxFields] = {};
    for (int f = 0; f < nf && f < kMaxFields; ++f) dst[f] = cptr(spec, (size_t)f * S);
    launch("fwd_z", [&] { FFS_DISPATCH(launch_fwd_z, g, w2, dst, nf, 0, scale, tw_z, stream); });
  }

  size_t padded_len() const { return 2 * (size_t)(p.nx / 2 + 1) * p.ny * p.nz; }
  // the scalars whose i k s follow S in the curl pass's inputs: the density's first
  const void* grad1() const { return dles ? rho : (sles ? phi : nullptr); }
  const void* grad2() const { return (dles && sles) ? phi : nullptr; }
  int dyn_spec_fields() const { return dynp ? 19 : 13; }
  int dyn_real_fields() const { return dynp ? kDynReal : DR_W; }
  char* rptr(void* base, size_t fields) { return (char*)base + fields * padded_len() * (csz / 2); }

  // compact spectra -> padded real, in groups the pencil scratch holds
  void inverse_many(void* spec, int nf, void* real) {
    const int chunk = std::max(nin, nfw + 1);
    for (int f0 = 0; f0 < nf; f0 += chunk) {
      const int k = std::min(chunk, nf - f0);
      inverse_to_real(cptr(spec, (size_t)f0 * S), k, rptr(real, f0));
    }
  }

  // fullfourier! + x G + real! of one padded real field in place (dynamic_smagorinsky.jl:79-117,
  // average_c! dynamic_piomelli.jl:13-26): the full spectrum, scaled by 1/N^3, is never dealiased
  void filter_real(void* real) {
    if (fp32) {
      launch("dyn_filter", [&] { launch_x_r2c<float>(gf, (const float*)real, fw1, 1, tw_x, stream); });
    } else {
      launch("dyn_filter", [&] { launch_x_r2c<double>(gf, (const double*)real, fw1, 1, tw_x, stream); });
    }
    launch("dyn_filter", [&] { FFS_DISPATCH(launch_fwd_y, gf, fw1, fw2, 1, tw_y, stream); });
    void* dst[kMaxFields] = {fspec};
    launch("dyn_filter", [&] { FFS_DISPATCH(launch_fwd_z, gf, fw2, dst, 1, 0, scale, tw_z, stream); });
    launch("dyn_filter", [&] { FFS_DISPATCH(launch_gauss, gf, fspec, dyn_td2, stream); });
    launch("dyn_filter", [&] { FFS_DISPATCH(launch_inv_z, gf, fspec, fw1, 1, false, tw_z, stream); });
    launch("dyn_filter", [&] { FFS_DISPATCH(launch_inv_y, gf, fw1, fw2, 1, tw_y, stream); });
    if (fp32) launch("dyn_filter", [&] { launch_x_c2r<float>(gf, fw2, (float*)real, 1, tw_x, stream); });
    else launch("dyn_filter", [&] { launch_x_c2r<double>(gf, fw2, (double*)real, 1, tw_x, stream); });
  }

  // the dynamic Smagorinsky stress of the current velocity (realspace_LES_calculation!, les_real_calc.jl:3-11)
  // into dreal's slots DR_SA..DR_SA+4, and with store its dealiased transform into tau
  void dynamic_les(bool store) {
    const size_t rb = padded_len() * (csz / 2);
    if (!store) {  // a read-out leaves the coefficient state as the next step finds it
      if (pio) {
        HIPCHECK(hipMemcpyAsync(rptr(dreal, DR_SAVE), rptr(dreal, DR_PC), 2 * rb, hipMemcpyDeviceToDevice, stream));
        HIPCHECK(hipMemcpyAsync(&sd_save_dtm1, pio_dtm1, sizeof(double), hipMemcpyDeviceToHost, stream));
      }
      if (piop) {
        HIPCHECK(hipMemcpyAsync(rptr(dreal, DR_SAVEP), rptr(dreal, DR_CP), 2 * rb, hipMemcpyDeviceToDevice, stream));
        HIPCHECK(hipMemcpyAsync(&sd_save_dtpm1, pio_dtpm1, sizeof(double), hipMemcpyDeviceToHost, stream));
      }
      HIPCHECK(hipStreamSynchronize(stream));
    }
    // realspace_dynamic_les_calculation!: cm1 = c at iteration 0, then calc_cstar!; the same for cp
    if (pio) {
      if (iteration == 0)
        HIPCHECK(hipMemcpyAsync(rptr(dreal, DR_PCM1), rptr(dreal, DR_PC), rb, hipMemcpyDeviceToDevice, stream));
      launch("dyn_cstar", [&] { piomelli_cstar(DR_PC, DR_PCM1, pio_dtm1); });
    }
    if (piop) {
      if (iteration == 0)
        HIPCHECK(hipMemcpyAsync(rptr(dreal, DR_CPM1), rptr(dreal, DR_CP), rb, hipMemcpyDeviceToDevice, stream));
      launch("dyn_cstar", [&] { piomelli_cstar(DR_CP, DR_CPM1, pio_dtpm1); });
    }
    launch("dyn_prep", [&] { FFS_DISPATCH(launch_dyn_prep, g, u, dspec, dyn_td2, stream, dynp ? 1 : 0); });
    inverse_many(u, 3, rptr(dreal, DR_U));
    inverse_many(dspec, 13, rptr(dreal, DR_S));  // S, M, uf
    if (dynp) inverse_many(cptr(dspec, 13 * S), 6, rptr(dreal, DR_W));  // w, G w
    if (fp32) dyn_products<float>();
    else dyn_products<double>();
    for (int f = 0; f < 5; ++f) filter_real(rptr(dreal, DR_SA + f));
    for (int f = 0; f < 6; ++f) filter_real(rptr(dreal, DR_L + f));
    if (dynp)
      for (int f = 0; f < 5; ++f) filter_real(rptr(dreal, DR_PA + f));
    if (fp32) dyn_coef<float>();
    else dyn_coef<double>();
    filter_real(rptr(dreal, coef_slot()));  // average_c! (avg = true)
    if (dynp) filter_real(rptr(dreal, DR_CP));
    if (fp32) dyn_tau<float>();
    else dyn_tau<double>();
    if (store) forward_from_real(rptr(dreal, DR_SA), 5, tau);
    if (!store) {
      // the read-out's coefficients: c at DR_C, cp at DR_M (free once the coefficient is formed)
      if (pio) HIPCHECK(hipMemcpyAsync(rptr(dreal, DR_C), rptr(dreal, DR_PC), rb, hipMemcpyDeviceToDevice, stream));
      if (dynp) HIPCHECK(hipMemcpyAsync(rptr(dreal, DR_M), rptr(dreal, DR_CP), rb, hipMemcpyDeviceToDevice, stream));
      if (pio) {
        HIPCHECK(hipMemcpyAsync(rptr(dreal, DR_PC), rptr(dreal, DR_SAVE), 2 * rb, hipMemcpyDeviceToDevice, stream));
        HIPCHECK(hipMemcpyAsync(pio_dtm1, &sd_save_dtm1, sizeof(double), hipMemcpyHostToDevice, stream));
      }
      if (piop) {
        HIPCHECK(hipMemcpyAsync(rptr(dreal, DR_CP), rptr(dreal, DR_SAVEP), 2 * rb, hipMemcpyDeviceToDevice, stream));
        HIPCHECK(hipMemcpyAsync(pio_dtpm1, &sd_save_dtpm1, sizeof(double), hipMemcpyHostToDevice, stream));
      }
      HIPCHECK(hipStreamSynchronize(stream));
    }
  }
  void piomelli_cstar(int c, int cm1, double* dtm1) {
    if (fp32) launch_piomelli_cstar<float>(rslot<float>(c), rslot<float>(cm1), st, dtm1, padded_len(), stream);
    else launch_piomelli_cstar<double>(rslot<double>(c), rslot<double>(cm1), st, dtm1, padded_len(), stream);
  }
  DynP dyn_p() {
    DynP d;
    if (!dynp) return d;
    d.w3 = rptr(dreal, DR_W);
    d.wf3 = rptr(dreal, DR_WF);
    d.pa5 = rptr(dreal, DR_PA);
    d.cpstar = piop ? rptr(dreal, DR_CP) : nullptr;
    d.cp = rptr(dreal, DR_CP);
    return d;
  }
  double sd_save_dtpm1 = 0.0;
  double sd_save_dtm1 = 0.0;
  // the coefficient field: the step's own (dynamicSmagorinsky) or the persistent one (Piomelli)
  int coef_slot() const { return pio ? DR_PC : DR_C; }
  template <typename T> T* rslot(int slot) { return (T*)rptr(dreal, slot); }
  template <typename T> void dyn_products() {
    launch("dyn_products", [&] {
      launch_dyn_products<T>(rslot<T>(DR_S), rslot<T>(DR_U), rslot<T>(DR_SA), rslot<T>(DR_L), padded_len(), lesm.d2,
                             stream, pio ? rslot<T>(DR_PC) : nullptr, dyn_p());
    });
  }
  template <typename T> void dyn_coef() {
    // the filter widths add up in the coefficient (dynamic_smagorinsky.jl:155)
    launch("dyn_coef", [&] {
      launch_dyn_coef<T>(rslot<T>(DR_UF), rslot<T>(DR_M), rslot<T>(DR_L), rslot<T>(DR_SA), rslot<T>(coef_slot()),
                         padded_len(), dyn_td2 + lesm.d2, stream, pio ? 1 : 0, dyn_p());
    });
  }
  template <typename T> void dyn_tau() {
    launch("dyn_tau", [&] {
      launch_dyn_tau<T>(rslot<T>(DR_S), rslot<T>(coef_slot()), rslot<T>(DR_SA), padded_len(), lesm.d2, dyn_cmin,
                        stream, dynp ? rslot<T>(DR_W) : nullptr, dynp ? rslot<T>(DR_CP) : nullptr);
    });
  }

  // ---------------------------------------------------------------- z slabs
  // Device bytes of one z plane of the slab buffers (the x pass's inputs and outputs)
  size_t slab_plane_bytes() const { return (size_t)(nin + nfw + 1) * g.sx * g.ny * csz; }
  // Planes per slab by default: the slab buffers of the three passes take at most about kSlabBytes, most of the
  // device's 256 MB last-level cache, so what one pass writes is still there when the next pass reads it; 0 (no
  // slabs) when the whole grid's buffers fit that already or a slab would be thinner than 8 planes.  Thinner
  // slabs lose more to their launches' shorter grids than they gain (at 256^3: 64-plane slabs 0.858 ms per
  // FP64 step, 32-plane slabs 0.917, none 0.869).
  static constexpr size_t kSlabBytes = (size_t)240 << 20;
  int default_slab() const {
    const size_t per = slab_plane_bytes();
    if ((size_t)g.nz * per <= kSlabBytes) return 0;
    int zs = (int)(kSlabBytes / per);
    if (zs < 8) return 0;
    for (int k = 1; k <= zs; k *= 2)
      if (g.nz % k == 0 && 2 * k > zs) return k;  // the largest power of two that divides nz and fits
    return zs;
  }
  // slab buffers for slabs of up to `planes` planes (0: none).  With slabs, w2 holds only the forward y
  // output (nfw + 1 fields of (sx, my, nz)) during the nonlinear term, so the slab buffers go after it when
  // they fit there; otherwise they get allocations of their own.
  void alloc_slabs(int planes) {
    if (planes <= slab_cap) return;
    if (slab_own) {
      dfree(sa2);
      dfree(sb2);
    }
    sa2 = sb2 = nullptr;
    slab_own = false;
    slab_cap = 0;
    const size_t a = (size_t)nin * g.sx * g.ny * planes * csz, b = (size_t)(nfw + 1) * g.sx * g.ny * planes * csz;
    const size_t b1 = ((size_t)(nfw + 1) * g.sx * g.my * g.nz * csz + 255) / 256 * 256;
    if (b1 + a + b <= w2_bytes) {
      sa2 = (char*)w2 + b1;
      sb2 = (char*)w2 + b1 + a;
    } else {
      sa2 = dalloc(a);
      sb2 = dalloc(b);
      slab_own = true;
    }
    slab_cap = planes;
  }
  int slab_planes() const { return (z_slab > 0 && z_slab < g.nz) ? z_slab : 0; }

  // ---------------------------------------------------------------- the step
  bool is_stats_time(int64_t it) const { return it % p.dt_stat == 0; }
  bool is_spec_time(int64_t it) const { return it % p.dt_spec == 0; }
  // every transform length must be compiled and the x pass must hold a row's fields with these options;
  // checked before anything is allocated, so an unsupported run fails at creation, not mid-step
  void check_lengths() const {
    const char* ax[3] = {"nx", "ny", "nz"};
    const int n[3] = {p.nx, p.ny, p.nz};
    for (int i = 0; i < 3; ++i)
      if (!length_supported(n[i]))
        throw FfsError(FFS_ERR_ARG, std::string(ax[i]) + " = " + std::to_string(n[i]) +
                                        " is not a compiled transform length (" + compiled_lengths() + ")");
    XOps ops{};
    ops.nin = nin;
    ops.nout = nfw + 1;  // a stats step adds 1/2|u|^2
    ops.rho = dens ? 1 : 0;
    ops.phi = pscal ? 1 : 0;
    ops.ns = nsform ? 1 : 0;
    ops.les = (les && !dyn) ? 1 : 0;
    ops.lm = lesm;
    ops.rho_les = dles;
    ops.phi_les = sles;
    ops.pair = x_pair;
    ops.waves = x_waves;
    ops.split = x_split;
    const std::string e = fp32 ? x_pass_check<float>(p.nx, ops) : x_pass_check<double>(p.nx, ops);
    if (!e.empty()) throw FfsError(FFS_ERR_ARG, e);
  }

  // which of the plan's choices change this run's kernels
  void plan_options(int* pair, int* waves, int* split, int* fusion) const {
    XOps ops{};
    ops.nin = nin;
    ops.nout = nfw;
    ops.rho = dens ? 1 : 0;
    ops.phi = pscal ? 1 : 0;
    ops.ns = nsform ? 1 : 0;
    ops.les = (les && !dyn) ? 1 : 0;
    ops.lm = lesm;
    ops.rho_les = dles;
    ops.phi_les = sles;
    ops.pair = x_pair;
    ops.waves = x_waves;
    ops.split = x_split;
    if (fp32) x_plan_options<float>(p.nx, g.mx, ops, pair, waves, split);
    else x_plan_options<double>(p.nx, g.mx, ops, pair, waves, split);
    *fusion = (!nsform && !dens && !pscal && !les && !plane && p.hyperviscosity == FFS_HV_NONE) ? 1 : 0;
  }

  // a Nyquist index is retained in some direction (cutoff >= 1): the Hermitian planes need restoring
  bool nyquist_kept() const {
    return g.mx == g.nxh || (g.ny % 2 == 0 && g.my == g.ny) || (g.nz % 2 == 0 && g.mz == g.nz);
  }

  bool is_output_time(int64_t it) const { return tit != 0 && it % p.write_time == 0; }

  // calculate_rhs! up to (not including) fourierspacep2!: rhs <- dealias(F(u x w))/N^3,
  // flux <- F(-rho u), tau <- F(2 (c Delta)^2 |S| S), ke <- F(1/2|u|^2) on stats/spectra steps,
  // max|u_i| / max|rho| into st (variable dt)
  // fused = true (a plain step, fuse_update()): set_dt and the forcing sum run before the forward y pass,
  // and the forward z pass ends in the update, so the step is complete when this returns
  void nonlinear(bool need_ke, bool fused = false) {
    if (dyn) dynamic_les(true);
    launch("inv_z_curl", [&] {
      FFS_DISPATCH(launch_inv_z, g, u, w1, nin, true, tw_z, stream, rho, phi, nsform ? 1 : 0, grad1(), grad2());
    });
    XOps ops{};
    ops.nin = nin;
    ops.nout = nfw + (need_ke ? 1 : 0);
    ops.rho = dens ? 1 : 0;
    ops.phi = pscal ? 1 : 0;
    ops.ns = nsform ? 1 : 0;
    ops.les = (les && !dyn) ? 1 : 0;
    ops.ke = need_ke ? 1 : 0;
    ops.les_2alpha = 2 * les_alpha;
    ops.lm = lesm;
    ops.rho_les = dles;
    ops.phi_les = sles;
    ops.rho_cv = dles_cv;
    ops.phi_cv = sles_cv;
    ops.les_alpha = les_alpha;
    ops.pair = x_pair;
    ops.waves = x_waves;
    ops.split = x_split;
    unsigned long long* um = p.variable_dt ? st->umax.v : nullptr;
    unsigned long long* rm = (p.variable_dt && dens) ? st->rhomax.v : nullptr;
    const int nf = ops.nout;
    const int zs = slab_planes();
    if (zs > 0) {
      // the inverse y, x and forward y passes slab by slab along z: the slab buffers are the same for every
      // slab, so what one pass writes is read by the next from the last-level cache, not from HBM
      for (int z0 = 0; z0 < g.nz; z0 += zs) {
        const int nzc = std::min(zs, g.nz - z0);
        Geo gs = g;
        gs.nz = nzc;  // the x pass sees the slab's rows only
        launch("inv_y", [&] { FFS_DISPATCH(launch_inv_y, g, w1, sa2, nin, tw_y, stream, z0, nzc); });
        launch("x_nl", [&] { FFS_DISPATCH(launch_x_nl, gs, sa2, sb2, ops, um, rm, tw_x, stream); });
        launch("fwd_y", [&] { FFS_DISPATCH(launch_fwd_y, g, sb2, w2, nf, tw_y, stream, z0, nzc); });
      }
    } else {
      launch("inv_y", [&] { FFS_DISPATCH(launch_inv_y, g, w1, w2, nin, tw_y, stream); });
      launch("x_nl", [&] { FFS_DISPATCH(launch_x_nl, g, w2, w1, ops, um, rm, tw_x, stream); });
    }
    // the state's round trip through real space (its inverse transforms above are done)
    if (nyquist_kept())
      launch("hermitian", [&] { FFS_DISPATCH(launch_hermitian_planes, g, u, dens ? rho : nullptr,
                                             pscal ? phi : nullptr, stream); });
    if (fused) {
      step_scalars(true, true);  // reads u and the maximum the x pass left, as after the forward passes
      if (zs == 0) launch("fwd_y", [&] { FFS_DISPATCH(launch_fwd_y, g, w1, w2, nf, tw_y, stream); });
      UpdateArgs a = update_args((int)std::min<int64_t>(tit + 1, 3));
      a.time = &st->time;
      launch("fwd_z_update", [&] {
        FFS_DISPATCH(launch_fwd_z_update, g, w2, u, rhs, fm1, fm2, st, a, scale, tw_z, stream);
      });
      return;
    }
    if (zs == 0) launch("fwd_y", [&] { FFS_DISPATCH(launch_fwd_y, g, w1, w2, nf, tw_y, stream); });
    void* dst[kMaxFields] = {};
    int f = 0;
    if (nsform)
      for (int c = 0; c < 5; ++c) dst[f++] = cptr(uu, c * S);
    else
      for (int c = 0; c < 3; ++c) dst[f++] = cptr(rhs, c * S);
    if (dens)
      for (int c = 0; c < 3; ++c) dst[f++] = cptr(flux, c * S);
    if (pscal)
      for (int c = 0; c < 3; ++c) dst[f++] = cptr(pflux, c * S);
    if (les && !dyn)
      for (int c = 0; c < 5; ++c) dst[f++] = cptr(tau, c * S);
    if (need_ke) dst[f++] = ke;
    if (f != nf || nf > kMaxFields) throw FfsError(FFS_ERR_ARG, "forward z outputs do not match the x pass");
    launch("fwd_z", [&] { FFS_DISPATCH(launch_fwd_z, g, w2, dst, nf, nsform ? 0 : 3, scale, tw_z, stream); });
    // conservation form: rhs = div!(uu) (models.jl:176-180)
    if (nsform) launch("div_uu", [&] { FFS_DISPATCH(launch_div_uu, g, uu, rhs, stream); });
  }

  // set_dt! (variable dt) and the forcing band sum, as one single-block kernel
  void step_scalars(bool do_dt, bool do_force) {
    ScalarArgs a{};
    a.do_dt = (do_dt && p.variable_dt) ? 1 : 0;
    // it matters only through it <= 3 (cfl ramp) and tit through tit == 0 (AB3 halving)
    a.it = (int)std::min<int64_t>(iteration, 4);
    a.tit = (int)std::min<int64_t>(tit, 1);
    a.etd = (etd || p.integrator == FFS_EULER) ? 1 : 0;  // no first-step halving (Euller set_dt!, Euller.jl:12)
    a.cfl = p.cfl;
    a.dx = 2 * M_PI * p.y_domain / p.ny;
    a.gnorm = dens ? std::fabs(gval) : 0.0;  // norm(gravity(s))
    const bool force = do_force && p.forcing == FFS_FORCING_SHELL && nband > 0;
    a.band = band;
    a.nband = force ? nband : 0;
    a.power = p.forcing_power;
    if (a.do_dt || a.nband) launch("step_scalars", [&] { FFS_DISPATCH(launch_step_scalars, g, u, st, a, stream); });
    if (do_force && plane) {
      const int it = (int)std::min<int64_t>(iteration, 4);  // matters only through it <= 3 (the ramp)
      const UpdateArgs ua = update_args(3);
      launch("plane_forcing", [&] { FFS_DISPATCH(launch_plane_forcing, g, u, rhs, st, pfd, ua, 0, it, stream); });
    }
  }
  void set_dt() { step_scalars(true, false); }
  void forcing() { step_scalars(false, true); }

  // KH, KRZ of Globals.jl:148-152: the horizontal bins follow the finer of the x and y wavenumber steps
  void spec_axes() {
    if (!spec_kh.empty()) return;
    spec_kh = (KX.size() > 1 && KY.size() > 1 && KX[1] >= KY[1]) ? KX : rfftfreq(g.ny, p.y_domain);
    spec_krz = rfftfreq(g.nz, p.z_domain);
  }

  // calculate_specHV (spec.jl:424-450) from the marginals of one per-mode quantity:
  // h[1] += 1/2 sum e(i=1); h[n(kx)] += MAXDKH X[cx] (cx > 0); h[n(ky)] += 1/2 Y0 + MAXDKH Y;
  // v[n(kz)] += Z0 + 2 DKZ Z.  Bins past the end of KH / KRZ are dropped.
  void bin_spectrum(const double* X, const double* Y0, const double* Y, const double* Z0, const double* Z,
                    double* H, double* V) const {
    const int nh = (int)spec_kh.size(), nv = (int)spec_krz.size();
    const double dkh = nh > 1 ? spec_kh[1] : 1.0, dkz = nv > 1 ? spec_krz[1] : 1.0;
    auto bin = [](double k, double dk) { return (int)std::nearbyint(std::fabs(k) / dk); };
    for (int i = 0; i < nh; ++i) H[i] = 0.0;
    for (int i = 0; i < nv; ++i) V[i] = 0.0;
    H[0] += 0.5 * X[0];
    for (int cx = 1; cx < g.mx; ++cx) {
      const int b = bin(kxc[cx], dkh);
      if (b < nh) H[b] += dkh * X[cx];
    }
    for (int cy = 0; cy < g.my; ++cy) {
      const int b = bin(kyc[cy], dkh);
      if (b < nh) H[b] += 0.5 * Y0[cy] + dkh * Y[cy];
    }
    for (int cz = 0; cz < g.mz; ++cz) {
      const int b = bin(kzc[cz], dkz);
      if (b < nv) V[b] += Z0[cz] + 2 * dkz * Z[cz];
    }
  }

  void queue_spectra() {
    spec_axes();
    if (spec_slots.empty()) {
      spec_slots.resize(4);
      for (auto& sl : spec_slots) {
        HIPCHECK(hipHostMalloc((void**)&sl.host, (size_t)kSpecQ * spectra_out_len(g) * sizeof(double),
                               hipHostMallocDefault));
        HIPCHECK(hipEventCreateWithFlags(&sl.ev, hipEventDisableTiming));
      }
    }
    int slot = -1;
    for (int pass = 0; pass < 2 && slot < 0; ++pass) {
      if (pass == 1) finish_oldest_spectra();
      for (size_t i = 0; i < spec_slots.size(); ++i)
        if (!spec_slots[i].busy) {
          slot = (int)i;
          break;
        }
    }
    SpecArgs a{};
    a.nl = rhs;
    a.ke = ke;
    a.tau = les ? tau : nullptr;
    a.rho = dens ? rho : nullptr;
    a.flux = dens ? flux : nullptr;
    a.pf_factor = plane ? pfd.plane_factor : nullptr;
    a.pf_kz0 = p.forcing == FFS_FORCING_QF ? 3 : 1;
    a.nu = p.nu;
    a.alpha = dens ? alpha : 0.0;
    a.gz = dens ? gz : 0.0;
    a.gdir = gdir;
    a.gval = dens ? gval : 0.0;
    a.st = st;
    a.hv = hv_args();
    launch("spectra", [&] { FFS_DISPATCH(launch_spectra, g, u, a, spec_part, spec_out, stream); });
    auto& sl = spec_slots[slot];
    HIPCHECK(hipMemcpyAsync(sl.host, spec_out, (size_t)kSpecQ * spectra_out_len(g) * sizeof(double),
                            hipMemcpyDeviceToHost, stream));
    HIPCHECK(hipEventRecord(sl.ev, stream));
    sl.iteration = iteration;
    sl.busy = true;
    spec_pending.push_back(slot);
  }

  void finish_oldest_spectra() {
    if (spec_pending.empty()) return;
    const int slot = spec_pending.front();
    spec_pending.pop_front();
    auto& sl = spec_slots[slot];
    HIPCHECK(hipEventSynchronize(sl.ev));
    const size_t L = spectra_out_len(g);
    const int nh = (int)spec_kh.size(), nv = (int)spec_krz.size();
    std::vector<double> rec((size_t)kSpecQ * (nh + nv));
    for (int q = 0; q < kSpecQ; ++q) {
      const double* m = sl.host + (size_t)q * L;
      double* H = rec.data() + (size_t)q * (nh + nv);
      bin_spectrum(m, m + g.mx, m + g.mx + g.my, m + g.mx + 2 * g.my, m + g.mx + 2 * g.my + g.mz, H, H + nh);
    }
    spec_ready.emplace_back(sl.iteration, std::move(rec));
    sl.busy = false;
  }

  void queue_stats() {
    // find a free slot (complete the oldest pending record if all are busy)
    int slot = -1;
    for (size_t i = 0; i < slots.size(); ++i)
      if (!slots[i].busy) {
        slot = (int)i;
        break;
      }
    if (slot < 0) {
      finish_oldest();
      for (size_t i = 0; i < slots.size(); ++i)
        if (!slots[i].busy) {
          slot = (int)i;
          break;
        }
    }
    auto& sl = slots[slot];
    StatsAux aux{les ? tau : nullptr, dens ? rho : nullptr, gz, gdir, dens ? gval : 0.0, hv_args(),
                 dens ? hvd_args() : HvArgs{},
                 pscal ? phi : nullptr, dles ? flux : nullptr};
    launch("stats", [&] { FFS_DISPATCH(launch_stats, g, u, rhs, ke, aux, partials, red, stream); });
    HIPCHECK(hipMemcpyAsync(sl.red, red, kStatsSlots * sizeof(double), hipMemcpyDeviceToHost, stream));
    HIPCHECK(hipMemcpyAsync(sl.st, st, sizeof(DevState), hipMemcpyDeviceToHost, stream));
    HIPCHECK(hipEventRecord(sl.ev, stream));
    sl.iteration = iteration;
    sl.busy = true;
    pending.push_back(slot);
  }

  void finish_oldest() {
    if (pending.empty()) return;
    const int slot = pending.front();
    pending.pop_front();
    auto& sl = slots[slot];
    HIPCHECK(hipEventSynchronize(sl.ev));
    ready.push_back(make_record(sl.red, *sl.st, sl.iteration));
    sl.busy = false;
  }

  // stats(s) of stats.jl:64-165 from the device sums (IEEE division semantics: x/0 = Inf as in Julia)
  Record make_record(const double* a, const DevState& sd, int64_t it) const {
    const double nu = p.nu;
    const double u1 = 0, u2 = 0, u3 = 0;  // real(u[1,1,1]): the k=0 mode is zero (dealiased mean flow)
    const double u12 = a[0], u22 = a[1], u32 = a[2];
    const double kh = (u12 + u22) / 2, kv = u32 / 2, k = kh + kv, kaniso = 2 * kv / kh;
    const double eh = nu * a[3], ev = nu * a[4], e = eh + ev;
    const double nlh = a[5], nlv = a[6], pressh = a[7], pressv = a[8];
    Record r;
    const double time = sd.time;
    r.stats = {(double)it, time, u1, u2, u3, u12, u22, u32, kh, kv, kaniso, k, eh, ev, e,
               nlh, nlv, pressh, pressv, -(nlh + pressh), (nlv + pressv)};
    double tdiss = e;
    std::vector<double> other;
    double hvs[3] = {0, 0, 0};
    if (p.hyperviscosity) {  // hyperviscosity_stats (stats.jl:532-561), spectral_barrier_stats: hdissh, hdissv, hdiss
      const double f = p.hyperviscosity == FFS_HV_HYPER ? p.hv_nu : -1.0;
      hvs[0] = f * a[30];
      hvs[1] = f * a[31];
      hvs[2] = hvs[0] + hvs[1];
      tdiss += hvs[2];  // before the LES part (stats.jl:118-121)
    }
    if (pscal) {  // scalar_stats (stats.jl:221-248): the passive scalar comes first (sim_fields, macros.jl:48)
      const double sdiss = palpha * (a[33] + a[34] + a[35]);
      other = {a[36], a[32], a[33], a[34], a[35], sdiss};
    }
    if (dens) {  // stats(::AbstractDensityStratification), density_stratification_types.jl:46-87
      const double rhodiss = alpha * (a[23] + a[24] + a[25]);
      double trhodiss = rhodiss;
      double rhohdiss = 0.0, rhopr = 0.0;
      if (dles) {  // scalar_les_stats of the total flux (density_stratification_types.jl:52-56)
        rhopr = -a[38];
        trhodiss += rhopr;
      }
      if (p.density_hv) {  // scalar_hvis_stats (stats.jl:1086-1170): + rhohdiss
        rhohdiss = (p.density_hv == FFS_HV_HYPER ? p.density_hv_nu : -1.0) * a[37];
        trhodiss += rhohdiss;
      }
      other.insert(other.end(), {a[29], a[22], a[23], a[24], a[25], rhodiss, trhodiss, -gval * a[26]});
      if (dles) other.push_back(rhopr);
      if (p.density_hv) other.push_back(rhohdiss);
    }
    if (les) {  // les_stats, stats.jl:679-705
      const double prh = -a[27], prv = -a[28];
      other.insert(other.end(), {prh, prv, (2 * prv) / prh, prh + prv});
      tdiss += prh + prv;
    }
    r.stats.push_back(tdiss);
    r.stats.insert(r.stats.end(), other.begin(), other.end());
    if (plane) {  // stats(forcing): hp, vp, tp (forcing_types.jl:33, 291); aForcing hp, vp (:105)
      r.stats.insert(r.stats.end(), {sd.force_hp, sd.force_vp});
      if (p.forcing != FFS_FORCING_A) r.stats.push_back(sd.force_hp + sd.force_vp);
    }
    if (p.hyperviscosity) r.stats.insert(r.stats.end(), {hvs[0], hvs[1], hvs[2]});
    const double u1p2 = u12, u2p2 = u22, u3p2 = u32;
    const double uh = std::sqrt((u1p2 + u2p2) / 2);
    const double urms = std::sqrt((u1p2 + u2p2 + u3p2) / 3);
    const double LX = p.x_domain, LY = p.y_domain, LZ = p.z_domain;
    const double L11 = LX * a[13] * M_PI / u1p2, L12 = LY * a[14] * M_PI / u1p2, L13 = LZ * a[15] * M_PI / u1p2;
    const double du1dyp2 = a[10], du1dxp2 = a[9], du2dyp2 = a[12], du2dxp2 = a[11];
    const double l12 = std::sqrt(2 * u1p2 / du1dyp2);
    const double L21 = LX * a[16] * M_PI / u2p2, L22 = LY * a[17] * M_PI / u2p2, L23 = LZ * a[18] * M_PI / u2p2;
    const double l21 = std::sqrt(2 * u2p2 / du2dxp2);
    const double L31 = LX * a[19] * M_PI / u3p2, L32 = LY * a[20] * M_PI / u3p2, L33 = LZ * a[21] * M_PI / u3p2;
    const double Lhh = (L11 + L22) / 2, Thh = Lhh / urms;
    const double ll = std::sqrt(2 * (u1p2 + u2p2) / (du1dxp2 + du2dyp2));
    const double lt = std::sqrt(2 * (u1p2 + u2p2) / (du1dyp2 + du2dxp2));
    const double Lhv = (L13 + L23) / 2, Lvv = L33, Lvh = (L31 + L32) / 2;
    const double Lt = std::sqrt(k * k * k) / tdiss, Tt = k / tdiss;
    const double Lk = std::sqrt(std::sqrt(nu * nu * nu / tdiss)), Tk = std::sqrt(nu / tdiss);
    const double lg = std::sqrt(15 * nu / tdiss) * urms;
    r.scales = {(double)it, time, Lk, Lt, lg, ll, lt, Lhh, Lvv, Lhv, Lvh, l12, l21, L11, L12, L13, L21, L22, L23,
                L31, L32, L33};
    const double Reh = uh * Lhh / nu, Reh1 = std::pow(Lhh / Lk, 4.0 / 3);
    const double Rel = urms * ll / nu, Rel1 = std::pow(ll / Lk, 4.0 / 3);
    const double Ret = urms * lt / nu, Ret1 = std::pow(lt / Lk, 4.0 / 3);
    if (dens) {
      const double Nb = std::sqrt(gz * drhodz);  // calculate_N
      const double Lb = 2 * M_PI * urms / Nb, Lo = std::sqrt(tdiss / (Nb * Nb * Nb));
      const double Lom = std::sqrt(tdiss / (Nb * Nb * Nb / (8 * M_PI * M_PI * M_PI)));
      const double Tb = 2 * M_PI / Nb, Tbm = 1 / Nb;
      const double Rlg = urms * lg / nu, Rlg1 = std::pow(lg / Lk, 4.0 / 3);
      r.scales.insert(r.scales.end(), {Lb, Lo, Lom, Tk, Tt, Thh, Tb, Tbm});
      r.numbers = {(double)it, time, Reh, Reh1, Rlg, Rel, Ret, Rlg1, Rel1, Ret1, std::pow(Lo / Lk, 4.0 / 3),
                   std::pow(Lom / Lk, 4.0 / 3), 2 * M_PI * uh / (Nb * Lhh), Lhv / Lhh};
    } else {
      r.scales.insert(r.scales.end(), {Tk, Tt, Thh});
      r.numbers = {(double)it, time, Reh, Reh1, Rel, Ret, Rel1, Ret1};
    }
    return r;
  }

  // SpectralBarrier defaults (simulationtypes.jl:294-296): initk = cutoff 2 KX[end] / 3, endk = sqrt(3 cutoff) KX[end]
  HvArgs hv_args() const { return make_hv(p.hyperviscosity, p.hv_m, p.hv_nu, p.sb_initk, p.sb_endk); }
  // the density's own (densityHyperViscosity, density_stratification_types.jl:137-149; same defaults)
  HvArgs hvd_args() const {
    return make_hv(p.density_hv, p.density_hv_m, p.density_hv_nu, p.density_sb_initk, p.density_sb_endk);
  }
  HvArgs make_hv(int kind, int m, double nuh, double initk, double endk) const {
    HvArgs h{};
    h.kind = kind;
    h.m = m;
    h.nuh = nuh;
    const double kxe = KX.empty() ? 0.0 : KX.back();
    h.sb_i = initk > 0 ? initk : p.cutoff * 2 * kxe / 3;
    h.sb_p = endk > 0 ? endk : std::sqrt(p.cutoff * 3) * kxe;
    return h;
  }

  UpdateArgs update_args(int step) const {
    UpdateArgs a{};
    a.hv = hv_args();
    a.hvd = hvd_args();
    a.integrator = etd ? 1 : (p.integrator == FFS_EULER ? 2 : 0);
    a.step = step;
    a.nu = p.nu;
    a.forcing = (p.forcing == FFS_FORCING_SHELL && nband > 0) ? 1 : 0;
    if (plane) {
      a.forcing = p.forcing;
      a.pf_factor = pfd.plane_factor;
      a.pf_kz0 = p.forcing == FFS_FORCING_QF ? 3 : 1;
    }
    a.kf2 = kf2;
    a.tau = les ? tau : nullptr;
    if (dens) {
      a.rho = rho;
      a.rrhs = rrhs;
      a.rfm1 = rfm1;
      a.rfm2 = rfm2;
      a.flux = flux;
      a.alpha = alpha;
      a.drhodz = drhodz;
      a.gz = gz;
      a.gdir = gdir;
      a.gval = gval;
    }
    if (pscal) {
      a.phi = phi;
      a.prhs = prhs;
      a.pfm1 = pfm1;
      a.pfm2 = pfm2;
      a.pflux = pflux;
      a.palpha = palpha;
      a.pgrad = p.scalar_gradient;
      a.pdir = p.scalar_direction;
    }
    return a;
  }

  // the update kernel of one step; the integrator step enters as min(tit + 1, 3) (AB3 distinguishes
  // step 2, ETD steps 1, 2 and >= 3)
  void enqueue_update() {
    const UpdateArgs a = update_args((int)std::min<int64_t>(tit + 1, 3));
    UpdateArgs b = a;
    b.time = &st->time;  // the update advances s.time[] by dt
    launch("update", [&] { FFS_DISPATCH(launch_update, g, u, rhs, fm1, fm2, st, b, stream); });
  }

  // host bookkeeping after a step: fm2 <- fm1, fm1 <- rhs, old fm2 becomes the next rhs buffer
  void rotate() {
    void* old2 = fm2;
    fm2 = fm1;
    fm1 = rhs;
    rhs = old2;
    if (dens) {
      void* r2 = rfm2;
      rfm2 = rfm1;
      rfm1 = rrhs;
      rrhs = r2;
    }
    if (pscal) {
      void* q2 = pfm2;
      pfm2 = pfm1;
      pfm1 = prhs;
      prhs = q2;
    }
    tit += 1;
    iteration += 1;
    phase = (phase + 1) % 3;
    if (!p.variable_dt) host_time += p.dt;
  }

  void time_step() {
    enqueue_update();
    rotate();
  }

  // a step whose kernel arguments equal those of every later non-stats step of the same phase
  bool steady() const {
    return tit >= 2 && iteration >= 4 && !is_stats_time(iteration) && !is_spec_time(iteration) && !prof;
  }

  // the plain step: the rotational form without LES, density, passive scalar, plane forcing or
  // hyperviscosity, whose update the forward z pass carries out itself
  bool fuse_update() const {
    return fuse_ok && !nsform && !dens && !pscal && !les && !plane && p.hyperviscosity == FFS_HV_NONE;
  }

  // kernels of one steady step (no host bookkeeping), for capture
  void enqueue_steady_step() {
    if (fuse_update()) {
      nonlinear(false, true);
      return;
    }
    nonlinear(false);
    step_scalars(true, true);
    enqueue_update();
  }

  hipGraphExec_t graph_for_phase() {
    if (!gexec[phase]) {
      HIPCHECK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
      enqueue_steady_step();
      HIPCHECK(hipStreamEndCapture(stream, &graphs[phase]));
      HIPCHECK(hipGraphInstantiate(&gexec[phase], graphs[phase], nullptr, nullptr, 0));
    }
    return gexec[phase];
  }

  void advance_one() {
    if (graph_on && steady()) {
      HIPCHECK(hipGraphLaunch(graph_for_phase(), stream));
      rotate();
      return;
    }
    const bool stats_t = is_stats_time(iteration);
    const bool spec_t = is_spec_time(iteration);
    if (!stats_t && !spec_t && fuse_update()) {
      nonlinear(false, true);
      rotate();
      return;
    }
    nonlinear(stats_t || spec_t);
    step_scalars(true, true);
    if (spec_t) queue_spectra();
    if (stats_t) queue_stats();
    time_step();
  }

  void initialize() {
    if (initialized) throw FfsError(FFS_ERR_STATE, "already initialized");
    // calculate_rhs!(s) at iteration `start`: nonlinear term from the (dealiased) initial field,
    // then the projection of u done by realspace! (models.jl:192-193), then fourierspacep2!
    nonlinear(false);
    set_dt();
    launch("project", [&] { FFS_DISPATCH(launch_project, g, u, stream); });
    const UpdateArgs a = update_args(0);
    launch("update_rhs", [&] { FFS_DISPATCH(launch_update, g, u, rhs, fm1, fm2, st, a, stream); });
    // initialize!(timestep): fm1 = fm2 = rhs; dt2 = dt3 = dt (Adams_Bashforth.jl:157-164, ETD.jl:99-108)
    HIPCHECK(hipMemcpyAsync(fm1, rhs, 3 * S * csz, hipMemcpyDeviceToDevice, stream));
    HIPCHECK(hipMemcpyAsync(fm2, rhs, 3 * S * csz, hipMemcpyDeviceToDevice, stream));
    if (dens) {
      HIPCHECK(hipMemcpyAsync(rfm1, rrhs, S * csz, hipMemcpyDeviceToDevice, stream));
      HIPCHECK(hipMemcpyAsync(rfm2, rrhs, S * csz, hipMemcpyDeviceToDevice, stream));
    }
    if (pscal) {
      HIPCHECK(hipMemcpyAsync(pfm1, prhs, S * csz, hipMemcpyDeviceToDevice, stream));
      HIPCHECK(hipMemcpyAsync(pfm2, prhs, S * csz, hipMemcpyDeviceToDevice, stream));
    }
    // initialize! of the oscillator schemes (forcing_types.jl:66-74, 112-118, 151-164, 207-216,
    // 251-264): without a given target spectrum the target is the initial field's, and the N variants
    // take N = Nm1 from the initial right-hand side; with a target given (targSpectrum.dat) `init` is
    // false: Em stays as given and N = Nm1 = 0.  An N given (the N.N file) replaces N afterwards.
    // aForcing always takes the initial field's per-mode energies (forcing_types.jl:112-118); its
    // target can be replaced after initialize.
    if (plane && p.forcing != FFS_FORCING_QF && (!pf_em_set || p.forcing == FFS_FORCING_A)) {
      PlaneForcing e = pfd;
      e.kind = -p.forcing;
      const UpdateArgs ua = update_args(0);
      launch("plane_forcing_init", [&] { FFS_DISPATCH(launch_plane_forcing, g, u, rhs, st, e, ua, 1, 0, stream); });
    }
    if (plane && !pf_n_pending.empty())
      HIPCHECK(hipMemcpyAsync(pfd.N, pf_n_pending.data(), pfd.nslot * sizeof(double), hipMemcpyHostToDevice, stream));
    HIPCHECK(hipStreamSynchronize(stream));
    pf_n_pending.clear();
    initialized = true;
  }

  // the closing calculate_rhs! (with set_dt! and write_spec when due), forcing and writestats of
  // run_simulation (FluidFlowSimulation.jl:46-49); the step's dt becomes the one set_dt! computes
  void writestats_now() {
    nonlinear(true);
    step_scalars(true, true);
    if (is_spec_time(iteration)) queue_spectra();
    queue_stats();
  }

  // ---------------------------------------------------------------- checkpoints
  // Blob: header, DevState, then u, fm1, fm2 (3 components each) and, with density, rho, rfm1, rfm2,
  // each in the compact layout of this precision.  That is the whole integrator state (the rhs
  // buffer is rewritten before it is read), so a run restored from it continues bitwise as the
  // original would have.
  struct CkptHeader {
    char magic[8];
    uint32_t version, precision;
    int32_t nx, ny, nz, integrator, les, dens, variable_dt, pscal;
    int64_t iteration, tit;
    double host_time;
    uint64_t S, csz;
  };

  std::vector<void*> ckpt_buffers() const {
    std::vector<void*> b = {u, fm1, fm2};
    if (dens) b.insert(b.end(), {rho, rfm1, rfm2});
    if (pscal) b.insert(b.end(), {phi, pfm1, pfm2});
    return b;
  }
  size_t ckpt_buffer_bytes(const void* b) const { return (b == u || b == fm1 || b == fm2 ? 3 : 1) * S * csz; }

  size_t checkpoint_bytes() const {
    size_t n = sizeof(CkptHeader) + sizeof(DevState);
    for (void* b : ckpt_buffers()) n += ckpt_buffer_bytes(b);
    if (plane) n += 4 * pfd.nslot * sizeof(double);  // the forcing's R, Em, N, Nm1
    if (pio) n += 2 * padded_len() * (csz / 2) + sizeof(double);  // Piomelli's c, cm1 and dtm1
    if (piop) n += 2 * padded_len() * (csz / 2) + sizeof(double);  // and cp, cpm1, dtpm1
    return n;
  }

  CkptHeader ckpt_header() const {
    CkptHeader h{};
    std::memcpy(h.magic, "FFSCKPT1", 8);
    h.version = 2;  // 2: the state carries the per-slot maxima
    h.precision = fp32 ? 1 : 0;
    h.nx = p.nx;
    h.ny = p.ny;
    h.nz = p.nz;
    h.integrator = p.integrator;
    h.les = les ? p.les_model : 0;
    h.dens = dens ? 1 : 0;
    h.pscal = pscal ? 1 : 0;
    h.variable_dt = p.variable_dt;
    h.iteration = iteration;
    h.tit = tit;
    h.host_time = host_time;
    h.S = S;
    h.csz = csz;
    return h;
  }

  void save_checkpoint(unsigned char* out, size_t cap) {
    if (!initialized) throw FfsError(FFS_ERR_STATE, "ffs_initialize must be called first");
    if (cap < checkpoint_bytes()) throw FfsError(FFS_ERR_ARG, "checkpoint buffer too small");
    const CkptHeader h = ckpt_header();
    std::memcpy(out, &h, sizeof h);
    size_t off = sizeof h;
    HIPCHECK(hipMemcpyAsync(out + off, st, sizeof(DevState), hipMemcpyDeviceToHost, stream));
    off += sizeof(DevState);
    for (void* b : ckpt_buffers()) {
      HIPCHECK(hipMemcpyAsync(out + off, b, ckpt_buffer_bytes(b), hipMemcpyDeviceToHost, stream));
      off += ckpt_buffer_bytes(b);
    }
    if (plane)
      for (double* b : {pfd.R, pfd.Em, pfd.N, pfd.Nm1}) {
        HIPCHECK(hipMemcpyAsync(out + off, b, pfd.nslot * sizeof(double), hipMemcpyDeviceToHost, stream));
        off += pfd.nslot * sizeof(double);
      }
    if (pio) {
      const size_t rb = 2 * padded_len() * (csz / 2);
      HIPCHECK(hipMemcpyAsync(out + off, rptr(dreal, DR_PC), rb, hipMemcpyDeviceToHost, stream));
      off += rb;
      HIPCHECK(hipMemcpyAsync(out + off, pio_dtm1, sizeof(double), hipMemcpyDeviceToHost, stream));
      off += sizeof(double);
    }
    if (piop) {
      const size_t rb = 2 * padded_len() * (csz / 2);
      HIPCHECK(hipMemcpyAsync(out + off, rptr(dreal, DR_CP), rb, hipMemcpyDeviceToHost, stream));
      off += rb;
      HIPCHECK(hipMemcpyAsync(out + off, pio_dtpm1, sizeof(double), hipMemcpyDeviceToHost, stream));
    }
    HIPCHECK(hipStreamSynchronize(stream));
  }

  void load_checkpoint(const unsigned char* in, size_t n) {
    if (initialized) throw FfsError(FFS_ERR_STATE, "a checkpoint is loaded into a run that is not initialized");
    if (n < sizeof(CkptHeader)) throw FfsError(FFS_ERR_IO, "checkpoint too short");
    CkptHeader h;
    std::memcpy(&h, in, sizeof h);
    const CkptHeader mine = ckpt_header();
    if (std::memcmp(h.magic, mine.magic, 8) != 0 || h.version != mine.version)
      throw FfsError(FFS_ERR_IO, "not a checkpoint of this library version");
    if (h.precision != mine.precision || h.nx != mine.nx || h.ny != mine.ny || h.nz != mine.nz ||
        h.integrator != mine.integrator || h.les != mine.les || h.dens != mine.dens || h.pscal != mine.pscal ||
        h.variable_dt != mine.variable_dt || h.S != mine.S || h.csz != mine.csz)
      throw FfsError(FFS_ERR_ARG, "checkpoint was written by a run with other parameters");
    if (n != checkpoint_bytes()) throw FfsError(FFS_ERR_IO, "checkpoint size mismatch");
    size_t off = sizeof h;
    HIPCHECK(hipMemcpy(st, in + off, sizeof(DevState), hipMemcpyHostToDevice));
    off += sizeof(DevState);
    for (void* b : ckpt_buffers()) {
      HIPCHECK(hipMemcpy(b, in + off, ckpt_buffer_bytes(b), hipMemcpyHostToDevice));
      off += ckpt_buffer_bytes(b);
    }
    if (plane)
      for (double* b : {pfd.R, pfd.Em, pfd.N, pfd.Nm1}) {
        HIPCHECK(hipMemcpy(b, in + off, pfd.nslot * sizeof(double), hipMemcpyHostToDevice));
        off += pfd.nslot * sizeof(double);
      }
    if (pio) {
      const size_t rb = 2 * padded_len() * (csz / 2);
      HIPCHECK(hipMemcpy(rptr(dreal, DR_PC), in + off, rb, hipMemcpyHostToDevice));
      off += rb;
      HIPCHECK(hipMemcpy(pio_dtm1, in + off, sizeof(double), hipMemcpyHostToDevice));
      off += sizeof(double);
    }
    if (piop) {
      const size_t rb = 2 * padded_len() * (csz / 2);
      HIPCHECK(hipMemcpy(rptr(dreal, DR_CP), in + off, rb, hipMemcpyHostToDevice));
      off += rb;
      HIPCHECK(hipMemcpy(pio_dtpm1, in + off, sizeof(double), hipMemcpyHostToDevice));
    }
    iteration = h.iteration;
    tit = h.tit;
    host_time = h.host_time;
    phase = 0;  // graph cache is empty; the buffer roles were restored as saved
    initialized = true;
  }
};

// An ensemble advanced in lockstep.  Steps in which every run is steady replay one graph whose
// branches (forked from `lead` by an event, joined back by one event each) are the runs' steady
// steps; the branch structure is fixed by the runs' rotation phases, so graphs are cached per phase
// signature.  Runs whose steady steps launch the same kernels with the same launch shapes (same
// transform lengths, precision and model) form a group, and a group's step is one launch per pass of
// the kernels' batch forms over a device table of the runs' call records (the run index in grid z),
// so a group costs as many launches as one run.  A run whose step holds work that has no batch form
// keeps a branch of its own.  Ordering against each run's own stream: before a graph launch `lead`
// waits for every run's stream, and before a run's stream is used again it waits for `lead`.
struct ffs_batch {
  std::vector<ffs_sim*> runs;
  hipStream_t lead = nullptr;
  hipEvent_t fork = nullptr, join_lead = nullptr;
  std::vector<hipEvent_t> joins;
  struct Cached {
    std::vector<int> sig;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    void* calls = nullptr;  // the groups' call-record tables (device memory)
    int launches = 0;       // kernel launches per step
    int groups = 0;         // merged groups (of two runs or more)
    int merged_runs = 0;    // runs inside those groups
  };
  std::vector<Cached> cache;
  bool merge = true;
  int last_graph = -1;  // cache index of the last graph launched
  bool lead_behind = true;    // lead must wait for the runs' streams before the next graph
  bool runs_behind = false;   // the runs' streams must wait for lead before their next work
  std::string err;
  // device timestamps around the work of the last advance call that took a step: t0 after every run's
  // earlier work, t1 after all of this call's work (both on lead)
  hipEvent_t t0 = nullptr, t1 = nullptr;
  int64_t timed_steps = 0;

  void create(ffs_sim* const* sims, int n) {
    if (n <= 0 || n > FFS_MAX_BATCH) throw FfsError(FFS_ERR_ARG, "batch size must be 1.." + std::to_string(FFS_MAX_BATCH));
    int dev = 0;
    HIPCHECK(hipGetDevice(&dev));
    for (int i = 0; i < n; ++i) {
      ffs_sim* s = sims[i];
      if (!s) throw FfsError(FFS_ERR_ARG, "null handle in batch");
      if (!s->initialized) throw FfsError(FFS_ERR_STATE, "every run must be initialized before batching");
      if (s->device != dev) throw FfsError(FFS_ERR_ARG, "every run of a batch must live on the current device");
      for (int j = 0; j < i; ++j)
        if (sims[j] == s) throw FfsError(FFS_ERR_ARG, "a handle appears twice in the batch");
      runs.push_back(s);
    }
    HIPCHECK(hipStreamCreateWithFlags(&lead, hipStreamNonBlocking));
    HIPCHECK(hipEventCreateWithFlags(&fork, hipEventDisableTiming));
    HIPCHECK(hipEventCreateWithFlags(&join_lead, hipEventDisableTiming));
    joins.resize(n);
    for (auto& e : joins) HIPCHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    HIPCHECK(hipEventCreate(&t0));
    HIPCHECK(hipEventCreate(&t1));
  }

  void destroy() {
    if (lead) hipStreamSynchronize(lead);
    drop_cache();
    if (lead) hipStreamSynchronize(lead);  // the tables' release
    for (auto e : joins)
      if (e) hipEventDestroy(e);
    if (fork) hipEventDestroy(fork);
    if (join_lead) hipEventDestroy(join_lead);
    if (t0) hipEventDestroy(t0);
    if (t1) hipEventDestroy(t1);
    if (lead) hipStreamDestroy(lead);
    lead = nullptr;
  }

  std::vector<int> signature() const {
    std::vector<int> sig(runs.size());
    for (size_t i = 0; i < runs.size(); ++i) sig[i] = runs[i]->phase;
    return sig;
  }

  void drop_cache() {
    for (auto& c : cache) {
      if (c.exec) hipGraphExecDestroy(c.exec);
      if (c.graph) hipGraphDestroy(c.graph);
      if (c.calls) hipFreeAsync(c.calls, lead);
    }
    cache.clear();
    last_graph = -1;
  }

  // The launches of run r's steady step, recorded instead of issued.  The step is enqueued on r's
  // stream under capture with the recorder set: a launch with a batch form is recorded, anything else
  // lands in the captured graph, so an empty graph means the whole step was recorded.
  static bool record_step(ffs_sim* r, KRecorder& rec) {
    hipGraph_t gr = nullptr;
    HIPCHECK(hipStreamBeginCapture(r->stream, hipStreamCaptureModeThreadLocal));
    g_recorder = &rec;
    try {
      r->enqueue_steady_step();
    } catch (...) {
      g_recorder = nullptr;
      if (hipStreamEndCapture(r->stream, &gr) == hipSuccess && gr) hipGraphDestroy(gr);
      throw;
    }
    g_recorder = nullptr;
    HIPCHECK(hipStreamEndCapture(r->stream, &gr));
    size_t nodes = 0;
    const hipError_t e = gr ? hipGraphGetNodes(gr, nullptr, &nodes) : hipSuccess;
    if (gr) HIPCHECK(hipGraphDestroy(gr));
    HIPCHECK(e);
    return nodes == 0 && !rec.launches.empty();
  }

  // two recorded steps can share launches when every launch has the same kernel and shape
  static bool same_shape(const KRecorder& a, const KRecorder& b) {
    if (a.launches.size() != b.launches.size()) return false;
    for (size_t j = 0; j < a.launches.size(); ++j) {
      const KLaunch &x = a.launches[j], &y = b.launches[j];
      if (x.batch != y.batch || x.lds != y.lds || x.call.size() != y.call.size() || x.grid.x != y.grid.x ||
          x.grid.y != y.grid.y || x.grid.z != y.grid.z || x.block.x != y.block.x || x.block.y != y.block.y ||
          x.block.z != y.block.z)
        return false;
    }
    return true;
  }

  hipGraphExec_t graph_for(const std::vector<int>& sig) {
    for (size_t k = 0; k < cache.size(); ++k)
      if (cache[k].sig == sig) {
        last_graph = (int)k;
        return cache[k].exec;
      }
    Cached c;
    c.sig = sig;
    const size_t n = runs.size();
    // groups of runs whose steps were recorded whole and share their launch shapes
    std::vector<KRecorder> rec(n);
    std::vector<std::vector<int>> groups;
    std::vector<int> alone;
    for (size_t i = 0; i < n; ++i) {
      if (!merge || !record_step(runs[i], rec[i])) {
        alone.push_back((int)i);
        continue;
      }
      bool placed = false;
      for (auto& gr : groups)
        if (same_shape(rec[gr[0]], rec[i])) {
          gr.push_back((int)i);
          placed = true;
          break;
        }
      if (!placed) groups.push_back({(int)i});
    }
    // a group of one runs its own step; larger groups are cut so that grid z stays within 65535
    std::vector<std::vector<int>> merged;
    for (auto& gr : groups) {
      if (gr.size() == 1) {
        alone.push_back(gr[0]);
        continue;
      }
      unsigned zmax = 1;
      for (const KLaunch& k : rec[gr[0]].launches) zmax = std::max(zmax, k.grid.z * k.block.z);
      const size_t per = std::max<size_t>(1, 65535 / zmax);
      for (size_t a = 0; a < gr.size(); a += per) {
        std::vector<int> part(gr.begin() + a, gr.begin() + std::min(gr.size(), a + per));
        if (part.size() == 1)
          alone.push_back(part[0]);
        else
          merged.push_back(std::move(part));
      }
    }
    std::sort(alone.begin(), alone.end());
    // the call-record tables: for each merged group and launch, the group's records back to back
    std::vector<std::vector<size_t>> offs(merged.size());
    size_t bytes = 0;
    for (size_t m = 0; m < merged.size(); ++m)
      for (const KLaunch& k : rec[merged[m][0]].launches) {
        offs[m].push_back(bytes);
        bytes += (k.call.size() * merged[m].size() + 255) / 256 * 256;
      }
    if (bytes) {
      std::vector<unsigned char> host(bytes, 0);
      for (size_t m = 0; m < merged.size(); ++m)
        for (size_t j = 0; j < offs[m].size(); ++j)
          for (size_t r = 0; r < merged[m].size(); ++r) {
            const std::vector<unsigned char>& call = rec[merged[m][r]].launches[j].call;
            std::memcpy(host.data() + offs[m][j] + r * call.size(), call.data(), call.size());
          }
      // stream-ordered, so that neither this allocation nor its release (drop_cache) waits for the device
      // as a whole: a serving loop rebuilds ensembles while others run
      HIPCHECK(hipMallocAsync(&c.calls, bytes, lead));
      HIPCHECK(hipMemcpyAsync(c.calls, host.data(), bytes, hipMemcpyHostToDevice, lead));
      HIPCHECK(hipStreamSynchronize(lead));  // `host` goes out of scope
    }
    try {
      HIPCHECK(hipStreamBeginCapture(lead, hipStreamCaptureModeThreadLocal));
      HIPCHECK(hipEventRecord(fork, lead));
      for (size_t m = 0; m < merged.size(); ++m) {
        const int first = merged[m][0];
        hipStream_t st = runs[first]->stream;  // the group's branch
        HIPCHECK(hipStreamWaitEvent(st, fork, 0));
        const auto& ls = rec[first].launches;
        for (size_t j = 0; j < ls.size(); ++j)
          issue_batch(ls[j], (const char*)c.calls + offs[m][j], (int)merged[m].size(), st);
        HIPCHECK(hipEventRecord(joins[first], st));
        HIPCHECK(hipStreamWaitEvent(lead, joins[first], 0));
        c.merged_runs += (int)merged[m].size();
      }
      c.groups = (int)merged.size();
      for (int i : alone) {
        HIPCHECK(hipStreamWaitEvent(runs[i]->stream, fork, 0));
        runs[i]->enqueue_steady_step();
        HIPCHECK(hipEventRecord(joins[i], runs[i]->stream));
        HIPCHECK(hipStreamWaitEvent(lead, joins[i], 0));
      }
      HIPCHECK(hipStreamEndCapture(lead, &c.graph));
      HIPCHECK(hipGraphInstantiate(&c.exec, c.graph, nullptr, nullptr, 0));
    } catch (...) {
      hipGraph_t gr = nullptr;
      hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
      if (hipStreamIsCapturing(lead, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone &&
          hipStreamEndCapture(lead, &gr) == hipSuccess && gr)
        hipGraphDestroy(gr);
      if (c.graph) hipGraphDestroy(c.graph);
      if (c.calls) hipFreeAsync(c.calls, lead);
      throw;
    }
    {  // kernel launches per step, counted in the graph
      size_t nodes = 0;
      HIPCHECK(hipGraphGetNodes(c.graph, nullptr, &nodes));
      std::vector<hipGraphNode_t> nv(nodes);
      if (nodes) HIPCHECK(hipGraphGetNodes(c.graph, nv.data(), &nodes));
      c.launches = 0;
      for (auto nd : nv) {
        hipGraphNodeType t;
        HIPCHECK(hipGraphNodeGetType(nd, &t));
        if (t == hipGraphNodeTypeKernel) ++c.launches;
      }
    }
    cache.push_back(c);
    last_graph = (int)cache.size() - 1;
    return cache.back().exec;
  }

  // lead waits for the work already queued on every run's stream
  void lead_waits_runs() {
    for (size_t i = 0; i < runs.size(); ++i) {
      HIPCHECK(hipEventRecord(joins[i], runs[i]->stream));
      HIPCHECK(hipStreamWaitEvent(lead, joins[i], 0));
    }
    lead_behind = false;
  }

  // every run's stream waits for the graphs already queued on lead
  void runs_wait_lead() {
    HIPCHECK(hipEventRecord(join_lead, lead));
    for (auto* r : runs) HIPCHECK(hipStreamWaitEvent(r->stream, join_lead, 0));
    runs_behind = false;
  }

  void advance(int64_t nsteps, int64_t* done, int* event, int* who) {
    lead_behind = true;  // the runs may have been used on their own since the last call
    int64_t taken = 0;
    bool marked = false;
    // the start mark goes in before the first step, after every run's earlier work
    auto mark_start = [&] {
      if (marked) return;
      lead_waits_runs();
      HIPCHECK(hipEventRecord(t0, lead));
      runs_behind = true;
      marked = true;
    };
    // the end mark, after every run's work of this call
    auto mark_end = [&] {
      if (!marked) return;
      lead_waits_runs();
      HIPCHECK(hipEventRecord(t1, lead));
      runs_behind = true;
      timed_steps = taken;
    };
    for (int64_t k = 0; k < nsteps; ++k) {
      for (size_t i = 0; i < runs.size(); ++i) {
        ffs_sim* r = runs[i];
        if (r->is_output_time(r->iteration) && r->output_served != r->iteration) {
          for (auto* q : runs)  // every run due at this step is served by the one event
            if (q->is_output_time(q->iteration)) q->output_served = q->iteration;
          mark_end();
          if (runs_behind) runs_wait_lead();
          if (event) *event = FFS_EVENT_OUTPUT;
          if (who) *who = (int)i;
          return;
        }
      }
      mark_start();
      bool all_steady = true;
      for (auto* r : runs) all_steady = all_steady && r->steady();
      if (all_steady) {
        if (lead_behind) lead_waits_runs();
        HIPCHECK(hipGraphLaunch(graph_for(signature()), lead));
        for (auto* r : runs) r->rotate();
        runs_behind = true;
      } else {
        if (runs_behind) runs_wait_lead();
        for (auto* r : runs) r->advance_one();
        lead_behind = true;
      }
      if (done) *done += 1;
      ++taken;
    }
    mark_end();
    if (runs_behind) runs_wait_lead();
  }

  // 1: the work of every advance call so far has finished on the device; 0: some is still queued
  int idle() const {
    hipError_t e = hipStreamQuery(lead);
    if (e == hipErrorNotReady) return 0;
    HIPCHECK(e);
    for (auto* r : runs) {
      e = hipStreamQuery(r->stream);
      if (e == hipErrorNotReady) return 0;
      HIPCHECK(e);
    }
    return 1;
  }
};

#define FFS_DISPATCH_S(fn, ...)            \
  do {                                     \
    if (s->fp32) fn<float>(__VA_ARGS__);   \
    else fn<double>(__VA_ARGS__);          \
  } while (0)

namespace {

int fail(ffs_sim* s, const FfsError& e) {
  if (s) s->err = e.what();
  return e.code;
}

template <class F>
int guard(ffs_sim* s, F&& f) {
  try {
    f();
    return FFS_OK;
  } catch (const FfsError& e) {
    return fail(s, e);
  } catch (const std::exception& e) {
    if (s) s->err = e.what();
    return FFS_ERR_ARG;
  }
}

}  // namespace

// ---- the allocation-free real-space read-out (ffs_get_field_noalloc)
// One file-scope staging array per GPU, of kReadoutScratchBytes: the x pass writes slabs of whole z
// planes of the padded real field into it as doubles and each slab is copied to the caller's buffer.
// Host code reaches a device's copy only through the address hipGetSymbolAddress returns for it
// (`scratch` itself names a host-side stand-in).  Every read-out on a GPU shares that GPU's array.
constexpr size_t kReadoutScratchBytes = (size_t)32 << 20;
static __device__ __align__(256) double scratch[kReadoutScratchBytes / sizeof(double)];

namespace {

std::mutex g_scratch_mu;
std::map<int, double*> g_scratch_addr;  // device -> its copy of `scratch`

// the device's copy of `scratch`, looked up once per device with that device current
double* readout_scratch(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  const auto it = g_scratch_addr.find(device);
  if (it != g_scratch_addr.end()) return it->second;
  int cur = -1;
  HIPCHECK(hipGetDevice(&cur));
  if (cur != device) HIPCHECK(hipSetDevice(device));
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (cur != device) HIPCHECK(hipSetDevice(cur));
  if (e != hipSuccess || !ptr)
    throw FfsError(FFS_ERR_HIP, std::string("hipGetSymbolAddress of the read-out scratch: ") + hipGetErrorString(e));
  g_scratch_addr[device] = (double*)ptr;
  return (double*)ptr;
}

}  // namespace

extern "C" {

int ffs_abi_version(void) { return FFS_ABI_VERSION; }

int ffs_params_default(ffs_params* p) {
  if (!p) return FFS_ERR_ARG;
  std::memset(p, 0, sizeof(*p));
  p->size = sizeof(ffs_params);
  p->x_domain = p->y_domain = p->z_domain = 1.0;
  p->cutoff = 2.0 / 3.0;
  p->integrator = FFS_ETD3;
  p->variable_dt = 1;
  p->cfl = 0.18;
  p->dt = 0.0;
  p->dt_stat = p->dt_spec = p->write_time = 1 << 30;
  p->precision = FFS_FP64;
  p->smagorinsky_c = NAN;  // the model's own default (les_types.jl:364-404)
  p->les_cp = NAN;
  p->les_s_model = FFS_LES_SMAGORINSKY;
  p->filter_width = 0.0;
  p->pr = 1.0;
  p->reference_density = 1.0;
  p->forcing_kf = 2.5;
  p->forcing_power = 0.1;
  p->forcing_tf = 0.0;
  p->forcing_alpha = 0.0;
  p->forcing_zf_tanh = 0;
  p->alvelius_p = 0.0;
  p->alvelius_c = 1.0;
  p->alvelius_kf = 0.0;
  p->hyperviscosity = FFS_HV_NONE;
  p->hv_nu = 0.0;
  p->hv_m = 2;
  p->sb_initk = p->sb_endk = 0.0;
  p->passive_scalar = 0;
  p->scalar_pr = 1.0;
  p->scalar_gradient = 0.0;
  p->scalar_direction = 2;
  p->equation = FFS_EQ_ROTATIONAL;
  p->density_hv = FFS_HV_NONE;
  p->density_hv_m = 2;
  p->gravity_direction = 2;
  p->scalar_les_c = NAN;
  p->density_les_c = NAN;
  return FFS_OK;
}

int ffs_create(const ffs_params* p, ffs_sim** out) {
  if (!p || !out) return FFS_ERR_ARG;
  *out = nullptr;
  ffs_sim* s = new ffs_sim();
  try {
    s->create(*p);
  } catch (const FfsError& e) {
    {
      std::lock_guard<std::mutex> lk(g_err_mu);
      g_last_create_error = e.what();
    }
    s->destroy();
    delete s;
    return e.code;
  } catch (const std::exception& e) {
    {
      std::lock_guard<std::mutex> lk(g_err_mu);
      g_last_create_error = e.what();
    }
    s->destroy();
    delete s;
    return FFS_ERR_ARG;
  }
  *out = s;
  return FFS_OK;
}

int ffs_set_field(ffs_sim* s, int field, int space, const double* host, size_t n) {
  if (!s || !host) return FFS_ERR_ARG;
  return guard(s, [&] {
    if (field < FFS_FIELD_U1 || field > FFS_FIELD_SCALAR) throw FfsError(FFS_ERR_ARG, "unknown field");
    if (field == FFS_FIELD_RHO && !s->dens) throw FfsError(FFS_ERR_ARG, "the run has no density field");
    if (field == FFS_FIELD_SCALAR && !s->pscal) throw FfsError(FFS_ERR_ARG, "the run has no passive scalar");
    if (s->initialized) throw FfsError(FFS_ERR_STATE, "fields are set before ffs_initialize");
    const size_t nxh = s->g.nxh, ny = s->g.ny, nz = s->g.nz, nrx = s->g.nrx;
    const size_t padded = 2 * nxh * ny * nz, unpadded = nrx * ny * nz;
    void* dst = field == FFS_FIELD_RHO ? s->rho : field == FFS_FIELD_SCALAR ? s->phi : s->cptr(s->u, (size_t)field * s->S);
    if (space == FFS_SPACE_REAL) {
      std::vector<double> buf;
      const double* src = host;
      if (n == unpadded && n != padded) {  // FluidFields reader: accept the unpadded layout too
        buf.assign(padded, 0.0);
        for (size_t r = 0; r < ny * nz; ++r) std::memcpy(&buf[r * 2 * nxh], host + r * nrx, nrx * sizeof(double));
        src = buf.data();
      } else if (n != padded) {
        throw FfsError(FFS_ERR_IO, "field size " + std::to_string(n) + " matches neither the padded (" +
                                       std::to_string(padded) + ") nor the unpadded (" + std::to_string(unpadded) +
                                       ") layout");
      }
      double* d64 = (double*)s->dalloc(padded * sizeof(double));
      HIPCHECK(hipMemcpy(d64, src, padded * sizeof(double), hipMemcpyHostToDevice));
      void* real = d64;
      if (s->fp32) {
        real = s->dalloc(padded * sizeof(float));
        launch_cast_copy<float>(d64, real, padded, s->stream);
      }
      s->forward_from_real(real, 1, dst);
      HIPCHECK(hipStreamSynchronize(s->stream));
      // release the staging buffers
      for (void* b : {real, (void*)d64}) {
        for (size_t i = 0; i < s->allocs.size(); ++i)
          if (s->allocs[i] == b) {
            hipFree(b);
            s->allocs.erase(s->allocs.begin() + i);
            break;
          }
      }
      s->dev_bytes -= padded * sizeof(double) + (s->fp32 ? padded * sizeof(float) : 0);
    } else if (space == FFS_SPACE_FOURIER) {
      if (n != padded) throw FfsError(FFS_ERR_IO, "half-spectrum size mismatch");
      // gather the retained modes
      std::vector<double> c(2 * s->S);
      for (int cz = 0; cz < s->g.mz; ++cz)
        for (int cy = 0; cy < s->g.my; ++cy)
          for (int cx = 0; cx < s->g.mx; ++cx) {
            const size_t src = cx + nxh * (s->ymap[cy] + ny * (size_t)s->zmap[cz]);
            const size_t d = cx + (size_t)s->g.sx * (cy + (size_t)s->g.my * cz);
            c[2 * d] = host[2 * src];
            c[2 * d + 1] = host[2 * src + 1];
          }
      if (s->fp32) {
        std::vector<float> f(c.begin(), c.end());
        HIPCHECK(hipMemcpy(dst, f.data(), f.size() * sizeof(float), hipMemcpyHostToDevice));
      } else {
        HIPCHECK(hipMemcpy(dst, c.data(), c.size() * sizeof(double), hipMemcpyHostToDevice));
      }
    } else {
      throw FfsError(FFS_ERR_ARG, "unknown space");
    }
  });
}

int ffs_seed_velocity(ffs_sim* s, uint64_t seed, double k0, double energy) {
  if (!s) return FFS_ERR_ARG;
  return guard(s, [&] {
    if (s->initialized) throw FfsError(FFS_ERR_STATE, "fields are set before ffs_initialize");
    if (!(k0 > 0) || !(energy > 0)) throw FfsError(FFS_ERR_ARG, "k0 and energy must be positive");
    FFS_DISPATCH_S(launch_seed, s->g, s->u, 3, (unsigned long long)seed, k0, energy, s->partials, s->red, s->stream);
    HIPCHECK(hipStreamSynchronize(s->stream));
  });
}

int ffs_seed_density(ffs_sim* s, uint64_t seed, double k0, double rms) {
  if (!s) return FFS_ERR_ARG;
  return guard(s, [&] {
    if (!s->dens) throw FfsError(FFS_ERR_ARG, "the run has no density field");
    if (s->initialized) throw FfsError(FFS_ERR_STATE, "fields are set before ffs_initialize");
    if (!(k0 > 0) || !(rms > 0)) throw FfsError(FFS_ERR_ARG, "k0 and rms must be positive");
    FFS_DISPATCH_S(launch_seed, s->g, s->rho, 1, (unsigned long long)seed, k0, rms, s->partials, s->red, s->stream);
    HIPCHECK(hipStreamSynchronize(s->stream));
  });
}

int ffs_get_field(ffs_sim* s, int field, int space, double* host, size_t n) {
  if (!s || !host) return FFS_ERR_ARG;
  return guard(s, [&] {
    if (field < FFS_FIELD_U1 || field > FFS_FIELD_CP) throw FfsError(FFS_ERR_ARG, "unknown field");
    if (field >= FFS_FIELD_DFLUX1) {  // real-space read-outs formed from the current state
      if (space != FFS_SPACE_REAL) throw FfsError(FFS_ERR_ARG, "dflux, cs and cp are real-space fields");
      if (field <= FFS_FIELD_DFLUX3 && !s->dles) throw FfsError(FFS_ERR_ARG, "the run has no density LES flux");
      if (field == FFS_FIELD_CS && !s->dyn) throw FfsError(FFS_ERR_ARG, "the run has no dynamic LES model");
      if (field == FFS_FIELD_CP && !s->dynp) throw FfsError(FFS_ERR_ARG, "the run has no dynamic P coefficient");
    }
    if (field == FFS_FIELD_RHO && !s->dens) throw FfsError(FFS_ERR_ARG, "the run has no density field");
    if (field == FFS_FIELD_SCALAR && !s->pscal) throw FfsError(FFS_ERR_ARG, "the run has no passive scalar");
    const bool stress = field >= FFS_FIELD_T11 && field <= FFS_FIELD_T23;
    if (stress && !s->les) throw FfsError(FFS_ERR_ARG, "the run has no LES stress");
    if (space != FFS_SPACE_REAL && space != FFS_SPACE_FOURIER && space != FFS_SPACE_FOURIER_TO_REAL)
      throw FfsError(FFS_ERR_ARG, "unknown space");
    const size_t nxh = s->g.nxh, ny = s->g.ny, nz = s->g.nz;
    const size_t padded = 2 * nxh * ny * nz;
    if (n != padded) throw FfsError(FFS_ERR_IO, "output buffer must hold 2*(nx/2+1)*ny*nz doubles");
    void* src = field == FFS_FIELD_RHO      ? s->rho
                : field == FFS_FIELD_SCALAR ? s->phi
                : stress                    ? s->cptr(s->tau, (size_t)(field - FFS_FIELD_T11) * s->S)
                                            : s->cptr(s->u, (size_t)field * s->S);
    const size_t tb = s->fp32 ? sizeof(float) : sizeof(double);
    // copies one padded real field of the run's precision from the device to host doubles
    auto fetch_real = [&](const void* real) {
      if (s->fp32) {
        std::vector<float> f(padded);
        HIPCHECK(hipMemcpyAsync(f.data(), real, padded * sizeof(float), hipMemcpyDeviceToHost, s->stream));
        HIPCHECK(hipStreamSynchronize(s->stream));
        for (size_t i = 0; i < padded; ++i) host[i] = f[i];
      } else {
        HIPCHECK(hipMemcpyAsync(host, real, padded * sizeof(double), hipMemcpyDeviceToHost, s->stream));
        HIPCHECK(hipStreamSynchronize(s->stream));
      }
    };
    auto release = [&](void* real, size_t bytes) {
      for (size_t i = 0; i < s->allocs.size(); ++i)
        if (s->allocs[i] == real) {
          hipFree(real);
          s->allocs.erase(s->allocs.begin() + i);
          break;
        }
      s->dev_bytes -= bytes;
    };
    if (space == FFS_SPACE_FOURIER) {
      std::vector<double> c(2 * s->S);
      if (s->fp32) {
        std::vector<float> f(2 * s->S);
        HIPCHECK(hipMemcpyAsync(f.data(), src, f.size() * sizeof(float), hipMemcpyDeviceToHost, s->stream));
        HIPCHECK(hipStreamSynchronize(s->stream));
        for (size_t i = 0; i < f.size(); ++i) c[i] = f[i];
      } else {
        HIPCHECK(hipMemcpyAsync(c.data(), src, c.size() * sizeof(double), hipMemcpyDeviceToHost, s->stream));
        HIPCHECK(hipStreamSynchronize(s->stream));
      }
      std::memset(host, 0, n * sizeof(double));
      for (int cz = 0; cz < s->g.mz; ++cz)
        for (int cy = 0; cy < s->g.my; ++cy)
          for (int cx = 0; cx < s->g.mx; ++cx) {
            const size_t dsti = cx + nxh * (s->ymap[cy] + ny * (size_t)s->zmap[cz]);
            const size_t d = cx + (size_t)s->g.sx * (cy + (size_t)s->g.my * cz);
            host[2 * dsti] = c[2 * d];
            host[2 * dsti + 1] = c[2 * d + 1];
          }
    } else if (field == FFS_FIELD_CS || field == FFS_FIELD_CP) {
      s->dynamic_les(false);
      fetch_real(s->rptr(s->dreal, field == FFS_FIELD_CS ? ffs_sim::DR_C : ffs_sim::DR_M));
    } else if (field >= FFS_FIELD_DFLUX1) {
      // realspace!: S, w and i k rho to real space, then the LES flux alone (les_real_calc.jl:120-127)
      s->launch("inv_z_curl", [&] {
        FFS_DISPATCH_S(launch_inv_z, s->g, s->u, s->w1, s->nin, true, s->tw_z, s->stream, s->rho, s->phi,
                       s->nsform ? 1 : 0, s->grad1(), s->grad2());
      });
      s->launch("inv_y", [&] { FFS_DISPATCH_S(launch_inv_y, s->g, s->w1, s->w2, s->nin, s->tw_y, s->stream); });
      // S, then i k rho (the density's gradient comes first); w (slots 3..5) only in the rotational form,
      // and only the vorticity term reads it
      const int s0 = (s->nsform ? 3 : 6) + 1 + (s->pscal ? 1 : 0);
      const size_t S2 = (size_t)s->g.sx * s->g.ny * s->g.nz;
      void* real = s->dalloc(14 * padded * tb);  // S (5), w (3), grad rho (3), the flux (3)
      auto R = [&](int k) { return (char*)real + (size_t)k * padded * tb; };
      if (s->fp32) {
        s->launch("x_c2r", [&] { launch_x_c2r<float>(s->g, s->cptr(s->w2, s0 * S2), (float*)R(0), 5, s->tw_x, s->stream); });
        s->launch("x_c2r", [&] { launch_x_c2r<float>(s->g, s->cptr(s->w2, 3 * S2), (float*)R(5), 3, s->tw_x, s->stream); });
        s->launch("x_c2r", [&] {
          launch_x_c2r<float>(s->g, s->cptr(s->w2, (s0 + 5) * S2), (float*)R(8), 3, s->tw_x, s->stream);
        });
        s->launch("les_flux_real", [&] {
          launch_les_scalar_flux_real<float>((const float*)R(0), (const float*)R(5), (const float*)R(8), (float*)R(11),
                                             padded, s->lesm, s->les_alpha, s->dles, s->dles_cv, s->stream);
        });
      } else {
        s->launch("x_c2r", [&] { launch_x_c2r<double>(s->g, s->cptr(s->w2, s0 * S2), (double*)R(0), 5, s->tw_x, s->stream); });
        s->launch("x_c2r", [&] { launch_x_c2r<double>(s->g, s->cptr(s->w2, 3 * S2), (double*)R(5), 3, s->tw_x, s->stream); });
        s->launch("x_c2r", [&] {
          launch_x_c2r<double>(s->g, s->cptr(s->w2, (s0 + 5) * S2), (double*)R(8), 3, s->tw_x, s->stream);
        });
        s->launch("les_flux_real", [&] {
          launch_les_scalar_flux_real<double>((const double*)R(0), (const double*)R(5), (const double*)R(8),
                                              (double*)R(11), padded, s->lesm, s->les_alpha, s->dles, s->dles_cv,
                                              s->stream);
        });
      }
      fetch_real(R(11 + (field - FFS_FIELD_DFLUX1)));
      release(real, 14 * padded * tb);
    } else if (stress && space == FFS_SPACE_REAL && s->dyn) {
      // the dynamic model's stress of the current velocity, as the step forms it
      s->dynamic_les(false);
      fetch_real(s->rptr(s->dreal, ffs_sim::DR_SA + (field - FFS_FIELD_T11)));
    } else if (stress && space == FFS_SPACE_REAL) {
      // realspace!: the curl pass forms S from the current velocity, the x pass's inverse transform
      // takes the five S fields to real space, then the pointwise product (les_real_calc.jl:77-79)
      void* real = s->dalloc(5 * padded * tb);
      s->launch("inv_z_curl", [&] {
        FFS_DISPATCH_S(launch_inv_z, s->g, s->u, s->w1, s->nin, true, s->tw_z, s->stream, s->rho, s->phi,
                       s->nsform ? 1 : 0, s->grad1(), s->grad2());
      });
      s->launch("inv_y", [&] { FFS_DISPATCH_S(launch_inv_y, s->g, s->w1, s->w2, s->nin, s->tw_y, s->stream); });
      const int s0 = (s->nsform ? 3 : 6) + (s->dens ? 1 : 0) + (s->pscal ? 1 : 0);
      const size_t S2 = (size_t)s->g.sx * s->g.ny * s->g.nz;
      void* a2 = s->cptr(s->w2, (size_t)s0 * S2);
      const bool w_les = les_needs_w(s->lesm);
      void* wreal = w_les ? s->dalloc(3 * padded * tb) : nullptr;  // w, for the models that read it
      void* a2w = s->cptr(s->w2, (size_t)3 * S2);                   // w at slots 3..5 (rotational form)
      if (s->fp32) {
        s->launch("x_c2r", [&] { launch_x_c2r<float>(s->g, a2, (float*)real, 5, s->tw_x, s->stream); });
        if (w_les) {
          s->launch("x_c2r", [&] { launch_x_c2r<float>(s->g, a2w, (float*)wreal, 3, s->tw_x, s->stream); });
          s->launch("les_model_real", [&] {
            launch_les_model_real<float>((float*)real, (const float*)wreal, padded, s->lesm, s->stream);
          });
        } else {
          s->launch("les_tau_real", [&] { launch_les_tau_real<float>((float*)real, padded, 2 * s->les_alpha, s->stream); });
        }
      } else {
        s->launch("x_c2r", [&] { launch_x_c2r<double>(s->g, a2, (double*)real, 5, s->tw_x, s->stream); });
        if (w_les) {
          s->launch("x_c2r", [&] { launch_x_c2r<double>(s->g, a2w, (double*)wreal, 3, s->tw_x, s->stream); });
          s->launch("les_model_real", [&] {
            launch_les_model_real<double>((double*)real, (const double*)wreal, padded, s->lesm, s->stream);
          });
        } else {
          s->launch("les_tau_real", [&] { launch_les_tau_real<double>((double*)real, padded, 2 * s->les_alpha, s->stream); });
        }
      }
      fetch_real((const char*)real + (size_t)(field - FFS_FIELD_T11) * padded * tb);
      if (wreal) release(wreal, 3 * padded * tb);
      release(real, 5 * padded * tb);
    } else {  // REAL (a state field) or FOURIER_TO_REAL: the stored transform back in real space
      void* real = s->dalloc(padded * tb);
      s->inverse_to_real(src, 1, real);
      fetch_real(real);
      release(real, padded * tb);
    }
  });
}

int ffs_get_field_noalloc(ffs_sim* s, int field, double* host, size_t n) {
  if (!s || !host) return FFS_ERR_ARG;
  return guard(s, [&] {
    if (field < FFS_FIELD_U1 || field > FFS_FIELD_SCALAR)
      throw FfsError(FFS_ERR_ARG, "the allocation-free read-out serves the state fields u1 u2 u3 rho scalar");
    if (field == FFS_FIELD_RHO && !s->dens) throw FfsError(FFS_ERR_ARG, "the run has no density field");
    if (field == FFS_FIELD_SCALAR && !s->pscal) throw FfsError(FFS_ERR_ARG, "the run has no passive scalar");
    const size_t nxh = s->g.nxh, ny = s->g.ny, nz = s->g.nz;
    const size_t plane = 2 * nxh * ny;  // doubles per z plane of the padded real layout
    if (n != plane * nz) throw FfsError(FFS_ERR_IO, "output buffer must hold 2*(nx/2+1)*ny*nz doubles");
    const size_t cap = kReadoutScratchBytes / sizeof(double);
    if (plane > cap)
      throw FfsError(FFS_ERR_ARG, "one z plane of the padded real field (" + std::to_string(plane * sizeof(double)) +
                                      " bytes) exceeds the read-out scratch (" +
                                      std::to_string(kReadoutScratchBytes) + " bytes)");
    double* buf = readout_scratch(s->device);
    void* src = field == FFS_FIELD_RHO ? s->rho : field == FFS_FIELD_SCALAR ? s->phi : s->cptr(s->u, (size_t)field * s->S);
    // the inverse z and y passes into the run's own pencil workspace, then the x pass slab by slab
    FFS_DISPATCH_S(launch_inv_z, s->g, src, s->w1, 1, false, s->tw_z, s->stream);
    ffs_sim::check_launch("inv_z");
    FFS_DISPATCH_S(launch_inv_y, s->g, s->w1, s->w2, 1, s->tw_y, s->stream);
    ffs_sim::check_launch("inv_y");
    const int zper = (int)std::min<size_t>(nz, cap / plane);
    for (int z0 = 0; z0 < (int)nz; z0 += zper) {
      const int nzs = std::min(zper, (int)nz - z0);
      FFS_DISPATCH_S(launch_x_c2r_rows, s->g, s->w2, buf, z0 * (int)ny, nzs * (int)ny, s->tw_x, s->stream);
      ffs_sim::check_launch("x_c2r");
      // stream order keeps the next slab's x pass behind this copy
      HIPCHECK(hipMemcpyAsync(host + (size_t)z0 * plane, buf, (size_t)nzs * plane * sizeof(double),
                              hipMemcpyDeviceToHost, s->stream));
    }
    HIPCHECK(hipStreamSynchronize(s->stream));
  });
}

int ffs_initialize(ffs_sim* s) {
  if (!s) return FFS_ERR_ARG;
  return guard(s, [&] { s->initialize(); });
}

int ffs_advance(ffs_sim* s, int64_t nsteps, int64_t* done, int* event) {
  if (!s) return FFS_ERR_ARG;
  if (done) *done = 0;
  if (event) *event = FFS_EVENT_NONE;
  return guard(s, [&] {
    if (!s->initialized) throw FfsError(FFS_ERR_STATE, "ffs_initialize must be called first");
    for (int64_t i = 0; i < nsteps; ++i) {
      if (s->is_output_time(s->iteration) && s->output_served != s->iteration) {
        s->output_served = s->iteration;
        if (event) *event = FFS_EVENT_OUTPUT;
        return;
      }
      s->advance_one();
      if (done) *done += 1;
    }
  });
}

int ffs_forcing_state(ffs_sim* s, int what, int set, double* data, int cap, int* n) {
  if (!s) return FFS_ERR_ARG;
  return guard(s, [&] {
    if (!s->plane) throw FfsError(FFS_ERR_ARG, "the run has no plane forcing (Forcing_methods schemes)");
    const int len = s->pf_state_len(), nslot = s->pfd.nslot;
    if (n) *n = len;
    if (!data) return;
    if (what == FFS_FSTATE_KH) {
      if (set) throw FfsError(FFS_ERR_ARG, "the shell wavenumbers are read only");
      if (s->p.forcing == FFS_FORCING_A) throw FfsError(FFS_ERR_ARG, "aForcing has no 2D shells");
      for (int i = 0; i < len && i < cap; ++i) data[i] = s->pf_avgk[i];
      return;
    }
    if (what == FFS_FSTATE_EF) {  // the last forcing call's spectrum over every shell
      if (set) throw FfsError(FFS_ERR_ARG, "the current spectrum is read only");
      if (s->p.forcing == FFS_FORCING_A) throw FfsError(FFS_ERR_ARG, "aForcing keeps no shell spectrum");
      std::vector<double> e(len);
      HIPCHECK(hipStreamSynchronize(s->stream));
      HIPCHECK(hipMemcpy(e.data(), s->pfd.Ef_all, len * sizeof(double), hipMemcpyDeviceToHost));
      for (int i = 0; i < len && i < cap; ++i) data[i] = e[i];
      return;
    }
    double* dev = what == FFS_FSTATE_R ? s->pfd.R : what == FFS_FSTATE_EM ? s->pfd.Em
                : what == FFS_FSTATE_N ? s->pfd.N : nullptr;
    if (!dev) throw FfsError(FFS_ERR_ARG, "unknown forcing state");
    std::vector<double> slots(nslot);
    if (set) {
      if (cap < len) throw FfsError(FFS_ERR_ARG, "the state needs one value per shell (aForcing: per mode)");
      for (int sl = 0; sl < nslot; ++sl) slots[sl] = data[s->pf_state_pos(sl)];
      if (what == FFS_FSTATE_N && !s->initialized) {
        s->pf_n_pending = slots;
        return;
      }
      HIPCHECK(hipStreamSynchronize(s->stream));
      HIPCHECK(hipMemcpy(dev, slots.data(), nslot * sizeof(double), hipMemcpyHostToDevice));
      if (what == FFS_FSTATE_EM) s->pf_em_set = true;
    } else {
      HIPCHECK(hipStreamSynchronize(s->stream));
      HIPCHECK(hipMemcpy(slots.data(), dev, nslot * sizeof(double), hipMemcpyDeviceToHost));
      for (int i = 0; i < len && i < cap; ++i) data[i] = 0.0;
      for (int sl = 0; sl < nslot; ++sl)
        if (s->pf_state_pos(sl) < cap) data[s->pf_state_pos(sl)] = slots[sl];
    }
  });
}

int ffs_checkpoint_size(const ffs_sim* s, size_t* bytes) {
  if (!s || !bytes) return FFS_ERR_ARG;
  *bytes = s->checkpoint_bytes();
  return FFS_OK;
}

int ffs_checkpoint_save(ffs_sim* s, void* host, size_t cap) {
  if (!s || !host) return FFS_ERR_ARG;
  return guard(s, [&] { s->save_checkpoint((unsigned char*)host, cap); });
}

int ffs_checkpoint_load(ffs_sim* s, const void* host, size_t n) {
  if (!s || !host) return FFS_ERR_ARG;
  return guard(s, [&] { s->load_checkpoint((const unsigned char*)host, n); });
}

int ffs_writestats_now(ffs_sim* s) {
  if (!s) return FFS_ERR_ARG;
  return guard(s, [&] {
    if (!s->initialized) throw FfsError(FFS_ERR_STATE, "ffs_initialize must be called first");
    s->writestats_now();
  });
}

int ffs_stats_pop(ffs_sim* s, double* stats, int cap_stats, int* n_stats, double* scales, int cap_scales,
                  int* n_scales, double* numbers, int cap_numbers, int* n_numbers) {
  if (!s) return FFS_ERR_ARG;
  int got = 0;
  int rc = guard(s, [&] {
    if (s->ready.empty()) s->finish_oldest();
    if (s->ready.empty()) return;
    Record r = std::move(s->ready.front());
    s->ready.pop_front();
    auto put = [](const std::vector<double>& v, double* dst, int cap, int* n) {
      if (n) *n = (int)v.size();
      if (dst)
        for (int i = 0; i < (int)v.size() && i < cap; ++i) dst[i] = v[i];
    };
    put(r.stats, stats, cap_stats, n_stats);
    put(r.scales, scales, cap_scales, n_scales);
    put(r.numbers, numbers, cap_numbers, n_numbers);
    got = 1;
  });
  return rc < 0 ? rc : got;
}

int ffs_stats_header(const ffs_sim* s, int table, char* buf, size_t cap) {
  if (!s || !buf || cap == 0) return FFS_ERR_ARG;
  std::string h;
  if (table == FFS_TABLE_STATS) {
    h = "iteration,time,u1,u2,u3,u1p2,u2p2,u3p2,kh,kv,kaniso,k,dissh,dissv,diss,nlh,nlv,pressh,pressv,tnlh,tnlv,tdiss";
    if (s->pscal) h += ",scalar,scalarp2,dscalardxp2,dscalardyp2,dscalardzp2,scalardiss";
    if (s->dens) h += ",rho,rhop2,drhodxp2,drhodyp2,drhodzp2,rhodiss,trhodiss,bflux";
    if (s->dens && s->dles) h += ",rhopr";
    if (s->dens && s->p.density_hv) h += ",rhohdiss";
    if (s->les) h += ",prh,prv,praniso,pr";
    if (s->plane) h += s->p.forcing == FFS_FORCING_A ? ",hp,vp" : ",hp,vp,tp";
    if (s->p.hyperviscosity) h += ",hdissh,hdissv,hdiss";
  } else if (table == FFS_TABLE_SCALES) {
    h = "iteration,time,Lk,Lt,lamg,laml,lamt,Lhh,Lvv,Lhv,Lvh,lam12,lam21,L11,L12,L13,L21,L22,L23,L31,L32,L33";
    h += s->dens ? ",Lb,Lo,Lom,Tk,Tt,Thh,Tb,Tbm" : ",Tk,Tt,Thh";
  } else if (table == FFS_TABLE_NUMBERS) {
    h = "iteration,time,Reh,Reh1,Relamg,Relaml,Relamt,Relamg1,Relaml1,Relamt1";
    if (s->dens) h += ",Reb,Rebm,Frh,Frh1";
  } else {
    return FFS_ERR_ARG;
  }
  std::snprintf(buf, cap, "%s", h.c_str());
  return h.size() < cap ? FFS_OK : FFS_ERR_ARG;
}

int ffs_spectrum(ffs_sim* s, double* h, int cap_h, int* n_h, double* v, int cap_v, int* n_v) {
  if (!s) return FFS_ERR_ARG;
  return guard(s, [&] {
    const Geo& g = s->g;
    double* ox = s->red + 64;
    double* oy0 = ox + 2 * g.mx;
    double* oy = oy0 + 2 * g.my;
    s->launch("spectrum", [&] {
      if (s->fp32) launch_spectrum<float>(g, s->u, ox, oy0, oy, nullptr, nullptr, s->partials, s->stream);
      else launch_spectrum<double>(g, s->u, ox, oy0, oy, nullptr, nullptr, s->partials, s->stream);
    });
    std::vector<double> X(2 * g.mx), Y0(2 * g.my), Y(2 * g.my), PZ(4 * (size_t)g.mz);
    HIPCHECK(hipMemcpyAsync(X.data(), ox, X.size() * 8, hipMemcpyDeviceToHost, s->stream));
    HIPCHECK(hipMemcpyAsync(Y0.data(), oy0, Y0.size() * 8, hipMemcpyDeviceToHost, s->stream));
    HIPCHECK(hipMemcpyAsync(Y.data(), oy, Y.size() * 8, hipMemcpyDeviceToHost, s->stream));
    const double* pz = s->partials + (size_t)g.mz * 2 * g.mx + (size_t)g.mz * 2 * g.my;
    HIPCHECK(hipMemcpyAsync(PZ.data(), pz, PZ.size() * 8, hipMemcpyDeviceToHost, s->stream));
    HIPCHECK(hipStreamSynchronize(s->stream));
    s->spec_axes();
    const int nh = (int)s->spec_kh.size(), nv = (int)s->spec_krz.size();
    // hv = 0: (|ux|^2+|uy|^2)/2, hv = 1: |uz|^2/2
    std::vector<double> H(2 * (size_t)nh), V(2 * (size_t)nv), Z0(g.mz), Z(g.mz);
    for (int hv = 0; hv < 2; ++hv) {
      for (int cz = 0; cz < g.mz; ++cz) {
        Z0[cz] = PZ[4 * cz + hv];
        Z[cz] = PZ[4 * cz + 2 + hv];
      }
      s->bin_spectrum(X.data() + (size_t)hv * g.mx, Y0.data() + (size_t)hv * g.my, Y.data() + (size_t)hv * g.my,
                      Z0.data(), Z.data(), H.data() + (size_t)hv * nh, V.data() + (size_t)hv * nv);
    }
    if (n_h) *n_h = nh;
    if (n_v) *n_v = nv;
    if (h)
      for (int i = 0; i < 2 * nh && i < cap_h; ++i) h[i] = H[i];
    if (v)
      for (int i = 0; i < 2 * nv && i < cap_v; ++i) v[i] = V[i];
  });
}

static_assert(kSpecQ == FFS_SPEC_Q && SQ_U_H == FFS_SPEC_U_H && SQ_PRESS_V == FFS_SPEC_PRESS_V &&
                  SQ_LES_H == FFS_SPEC_LES_H && SQ_BUOY == FFS_SPEC_BUOYANCY && SQ_RHONL == FFS_SPEC_RHONL &&
                  SQ_FORCING == FFS_SPEC_FORCING && SQ_HVIS_H == FFS_SPEC_HVIS_H && SQ_HVIS_V == FFS_SPEC_HVIS_V,
              "the device's spectra quantities follow the FFS_SPEC_* order");

int ffs_spectra_axes(ffs_sim* s, double* kh, int cap_h, int* n_h, double* krz, int cap_v, int* n_v) {
  if (!s) return FFS_ERR_ARG;
  return guard(s, [&] {
    s->spec_axes();
    if (n_h) *n_h = (int)s->spec_kh.size();
    if (n_v) *n_v = (int)s->spec_krz.size();
    if (kh)
      for (int i = 0; i < (int)s->spec_kh.size() && i < cap_h; ++i) kh[i] = s->spec_kh[i];
    if (krz)
      for (int i = 0; i < (int)s->spec_krz.size() && i < cap_v; ++i) krz[i] = s->spec_krz[i];
  });
}

int ffs_spectra_pop(ffs_sim* s, int64_t* iteration, double* data, size_t cap, size_t* n) {
  if (!s) return FFS_ERR_ARG;
  int got = 0;
  int rc = guard(s, [&] {
    if (s->spec_ready.empty()) s->finish_oldest_spectra();
    if (s->spec_ready.empty()) return;
    auto r = std::move(s->spec_ready.front());
    s->spec_ready.pop_front();
    if (iteration) *iteration = r.first;
    if (n) *n = r.second.size();
    if (data)
      for (size_t i = 0; i < r.second.size() && i < cap; ++i) data[i] = r.second[i];
    got = 1;
  });
  return rc < 0 ? rc : got;
}

int ffs_get_scalar(const ffs_sim* cs, int what, double* out) {
  ffs_sim* s = const_cast<ffs_sim*>(cs);
  if (!s || !out) return FFS_ERR_ARG;
  return guard(s, [&] {
    if (what == FFS_SCALAR_ITERATION) {
      *out = (double)s->iteration;
      return;
    }
    DevState h;
    HIPCHECK(hipMemcpyAsync(&h, s->st, sizeof(h), hipMemcpyDeviceToHost, s->stream));
    HIPCHECK(hipStreamSynchronize(s->stream));
    if (what == FFS_SCALAR_TIME) *out = h.time;
    else if (what == FFS_SCALAR_DT) *out = h.dt;
    else if (what == FFS_SCALAR_UMAX) *out = h.umax_d;
    else throw FfsError(FFS_ERR_ARG, "unknown scalar");
  });
}

int ffs_profile_enable(ffs_sim* s, int on) {
  if (!s) return FFS_ERR_ARG;
  return guard(s, [&] {
    if (!on) s->drain_prof();
    s->prof = on != 0;
  });
}

int ffs_profile_reset(ffs_sim* s) {
  if (!s) return FFS_ERR_ARG;
  return guard(s, [&] {
    s->drain_prof();
    s->prof_acc.clear();
  });
}

int ffs_profile_get(ffs_sim* s, const char** names, double* total_ms, int64_t* launches, int cap, int* n) {
  if (!s) return FFS_ERR_ARG;
  return guard(s, [&] {
    s->drain_prof();
    s->prof_names.clear();
    for (auto& kv : s->prof_acc) s->prof_names.push_back(kv.first);
    int i = 0;
    for (auto& kv : s->prof_acc) {
      if (i >= cap) break;
      if (names) names[i] = s->prof_names[i].c_str();
      if (total_ms) total_ms[i] = kv.second.first;
      if (launches) launches[i] = kv.second.second;
      ++i;
    }
    if (n) *n = (int)s->prof_acc.size();
  });
}

int ffs_device_bytes(const ffs_sim* s, size_t* bytes) {
  if (!s || !bytes) return FFS_ERR_ARG;
  *bytes = s->dev_bytes;
  return FFS_OK;
}

int ffs_synchronize(ffs_sim* s) {
  if (!s) return FFS_ERR_ARG;
  return guard(s, [&] { HIPCHECK(hipStreamSynchronize(s->stream)); });
}

int ffs_footprint(const ffs_params* p, size_t* bytes) {
  if (!p || !bytes) return FFS_ERR_ARG;
  ffs_sim tmp;
  tmp.dry = true;
  const int rc = guard(&tmp, [&] { tmp.create(*p); });
  if (rc != FFS_OK) {
    std::lock_guard<std::mutex> lk(g_err_mu);
    g_last_create_error = tmp.err;
    return rc;
  }
  *bytes = tmp.dev_bytes;
  tmp.destroy();
  return FFS_OK;
}

int ffs_device_memory(size_t* free_bytes, size_t* total_bytes) {
  size_t f = 0, t = 0;
  const hipError_t e = hipMemGetInfo(&f, &t);
  if (e != hipSuccess) {
    std::lock_guard<std::mutex> lk(g_err_mu);
    g_last_create_error = std::string("hipMemGetInfo: ") + hipGetErrorString(e);
    return FFS_ERR_HIP;
  }
  if (free_bytes) *free_bytes = f;
  if (total_bytes) *total_bytes = t;
  return FFS_OK;
}

int ffs_set_graph(ffs_sim* s, int on) {
  if (!s) return FFS_ERR_ARG;
  return guard(s, [&] {
    s->graph_on = on != 0;
    if (!s->graph_on) {
      HIPCHECK(hipStreamSynchronize(s->stream));
      s->drop_graphs();
    }
  });
}

int ffs_set_fusion(ffs_sim* s, int on) {
  if (!s) return FFS_ERR_ARG;
  return guard(s, [&] {
    if (s->fuse_ok == (on != 0)) return;
    HIPCHECK(hipStreamSynchronize(s->stream));
    s->drop_graphs();  // the captured steps hold the other kernel sequence
    s->fuse_ok = on != 0;
  });
}

int ffs_device_name(char* buf, size_t n) {
  if (!buf || n == 0) return FFS_ERR_ARG;
  int dev = 0;
  hipDeviceProp_t prop{};
  if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) return FFS_ERR_HIP;
  // only what the device reports the same in every process: its marketing name comes from a lookup table
  // that some environments lack (the name is then empty), so it is left out
  const std::string name = std::string(prop.gcnArchName) + " " + std::to_string(prop.multiProcessorCount) + "CU " +
                           std::to_string((prop.totalGlobalMem + (1ull << 29)) >> 30) + "GiB";
  std::snprintf(buf, n, "%s", name.c_str());
  return FFS_OK;
}

int ffs_get_plan(const ffs_sim* s, ffs_plan* out) {
  if (!s || !out) return FFS_ERR_ARG;
  *out = ffs_plan{};
  out->x_pair = s->x_pair;
  out->x_waves = s->x_waves;
  out->x_split = s->x_split;
  out->fusion = s->fuse_ok ? 1 : 0;
  out->z_slab = s->z_slab;
  return FFS_OK;
}

int ffs_set_plan(ffs_sim* s, const ffs_plan* plan) {
  if (!s || !plan) return FFS_ERR_ARG;
  return guard(s, [&] {
    if (plan->x_waves != 1 && plan->x_waves != 2) throw FfsError(FFS_ERR_ARG, "x_waves must be 1 or 2");
    if (plan->x_pair < 0 || plan->x_pair > 2) throw FfsError(FFS_ERR_ARG, "x_pair must be 0, 1 or 2");
    if (plan->z_slab < 0) throw FfsError(FFS_ERR_ARG, "z_slab must be >= 0");
    const int old[3] = {s->x_pair, s->x_waves, s->x_split};
    s->x_pair = plan->x_pair;
    s->x_waves = plan->x_waves;
    s->x_split = plan->x_split != 0;
    try {
      s->check_lengths();  // the general pass must hold the row if the products are not split
    } catch (...) {
      s->x_pair = old[0];
      s->x_waves = old[1];
      s->x_split = old[2];
      throw;
    }
    HIPCHECK(hipStreamSynchronize(s->stream));
    s->drop_graphs();  // the captured steps hold the old plan's launches
    s->fuse_ok = plan->fusion != 0;
    const int zs = plan->z_slab < s->g.nz ? plan->z_slab : 0;
    s->alloc_slabs(zs);
    s->z_slab = zs;
  });
}

int ffs_plan_options(const ffs_sim* s, ffs_plan* matters) {
  if (!s || !matters) return FFS_ERR_ARG;
  *matters = ffs_plan{};
  s->plan_options(&matters->x_pair, &matters->x_waves, &matters->x_split, &matters->fusion);
  matters->z_slab = s->default_slab();  // > 0: slabs change this run's launches (the default slab size)
  return FFS_OK;
}

int ffs_batch_create(ffs_sim* const* sims, int n, ffs_batch** out) {
  if (!sims || !out) return FFS_ERR_ARG;
  *out = nullptr;
  auto* b = new ffs_batch();
  try {
    b->create(sims, n);
  } catch (const FfsError& e) {
    {
      std::lock_guard<std::mutex> lk(g_err_mu);
      g_last_create_error = e.what();
    }
    b->destroy();
    delete b;
    return e.code;
  }
  *out = b;
  return FFS_OK;
}

int ffs_batch_advance(ffs_batch* b, int64_t nsteps, int64_t* done, int* event, int* who) {
  if (!b) return FFS_ERR_ARG;
  if (done) *done = 0;
  if (event) *event = FFS_EVENT_NONE;
  if (who) *who = -1;
  try {
    b->advance(nsteps, done, event, who);
    return FFS_OK;
  } catch (const FfsError& e) {
    b->err = e.what();
    return e.code;
  } catch (const std::exception& e) {
    b->err = e.what();
    return FFS_ERR_ARG;
  }
}

int ffs_batch_synchronize(ffs_batch* b) {
  if (!b) return FFS_ERR_ARG;
  try {
    HIPCHECK(hipStreamSynchronize(b->lead));
    for (auto* r : b->runs) HIPCHECK(hipStreamSynchronize(r->stream));
    return FFS_OK;
  } catch (const FfsError& e) {
    b->err = e.what();
    return e.code;
  }
}

int ffs_batch_query(ffs_batch* b) {
  if (!b) return FFS_ERR_ARG;
  try {
    return b->idle();
  } catch (const FfsError& e) {
    b->err = e.what();
    return e.code;
  }
}

int ffs_batch_last_advance_ms(ffs_batch* b, double* ms, int64_t* steps) {
  if (!b || !ms) return FFS_ERR_ARG;
  try {
    if (b->timed_steps <= 0) throw FfsError(FFS_ERR_STATE, "no advance call has taken a step yet");
    const hipError_t e = hipEventQuery(b->t1);
    if (e == hipErrorNotReady) return 1;
    HIPCHECK(e);
    float f = 0;
    HIPCHECK(hipEventElapsedTime(&f, b->t0, b->t1));
    *ms = f;
    if (steps) *steps = b->timed_steps;
    return FFS_OK;
  } catch (const FfsError& e) {
    b->err = e.what();
    return e.code;
  }
}

int ffs_batch_set_merge(ffs_batch* b, int on) {
  if (!b) return FFS_ERR_ARG;
  try {
    if (b->merge == (on != 0)) return FFS_OK;
    HIPCHECK(hipStreamSynchronize(b->lead));
    b->drop_cache();  // the cached graphs hold the other launch structure
    b->merge = on != 0;
    return FFS_OK;
  } catch (const FfsError& e) {
    b->err = e.what();
    return e.code;
  }
}

int ffs_batch_launch_info(const ffs_batch* b, int* launches, int* groups, int* merged_runs) {
  if (!b) return FFS_ERR_ARG;
  const bool any = b->last_graph >= 0;
  const auto* c = any ? &b->cache[b->last_graph] : nullptr;
  if (launches) *launches = any ? c->launches : 0;
  if (groups) *groups = any ? c->groups : 0;
  if (merged_runs) *merged_runs = any ? c->merged_runs : 0;
  return any ? FFS_OK : FFS_ERR_STATE;
}

const char* ffs_batch_last_error(const ffs_batch* b) {
  if (!b) {
    std::lock_guard<std::mutex> lk(g_err_mu);
    return g_last_create_error.c_str();
  }
  return b->err.c_str();
}

void ffs_batch_destroy(ffs_batch* b) {
  if (!b) return;
  b->destroy();
  delete b;
}

const char* ffs_last_error(const ffs_sim* s) {
  if (!s) {
    std::lock_guard<std::mutex> lk(g_err_mu);
    return g_last_create_error.c_str();
  }
  return s->err.c_str();
}

void ffs_destroy(ffs_sim* s) {
  if (!s) return;
  s->destroy();
  delete s;
}

}  // extern "C"
