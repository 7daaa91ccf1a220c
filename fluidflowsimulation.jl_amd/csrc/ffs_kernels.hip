// HIP kernels of the FFS pseudo-spectral time step (gfx950).
//
// One 3D real transform is three pencil passes.  For the vorticity-form step
// (models.jl:121-214) the passes are fused with the pointwise work around them:
//   inverse z pass  : reads the retained modes of u, forms w = ik x u on the fly
//                     (fourierspacep1!, models.jl:53-119), writes 6 fields
//   inverse y pass  : plain column transform, zero-padding the dealiased rows
//   x pass          : per (y,z) row, C2R of u and w, u x w (realspacecalculation!,
//                     models.jl:236-280), 1/2|u|^2 on stats steps (calc_k!,
//                     models.jl:216-226), max|u_i| for the CFL (find_max,
//                     utils.jl:247-258), R2C of the products -- physical-space
//                     fields never reach HBM
//   forward y pass  : keeps only the retained rows (dealias!, utils.jl:114-120)
//   forward z pass  : keeps retained rows, scales by 1/(nx ny nz) (myscale!,
//                     utils.jl:41-51), zeroes the k=0 mode (models.jl:174)
// The spectral update (fourierspacep2_velocity! models.jl:292-341, AB3
// Adams_Bashforth.jl:242-256, projection models.jl:443-456) is one
// elementwise kernel over the retained modes.
#include "ffs_kernels.h"
#include "fft_device.h"

#include <algorithm>
#include <atomic>
#include <cstdio>
#include <mutex>
#include <stdexcept>
#include <string>

// The library compiles this file several times in parallel, once per part (FFS_PART = 0 .. 7): every part
// sees all of the templates, but each instantiates only its own share of the launch functions below, and part
// 0 alone holds the definitions that are not templates.  Without FFS_PART one object holds everything.
#ifdef FFS_PART
#define FFS_IN_PART(p) (FFS_PART == (p))
#else
#define FFS_IN_PART(p) 1
#endif

namespace fsk {

// columns per tile of the y/z passes: one 128-byte line of x-consecutive values
template <int N> __host__ __device__ constexpr int tper() { return N / default_ppt(N); }
// the column kernels run TX x TP threads per block, at most kColThreads (their launch bound)
constexpr int kColThreads = 512;
// columns per tile: one 128-byte line of x-consecutive values, halved while a tile of columns of length n (at
// 16 points per thread) would exceed the block (FP32 columns of 640 and 768)
template <typename T> __host__ __device__ constexpr int col_tx(int n, int line = 128) {
  int tx = line / (2 * (int)sizeof(T));
  while (tx > 1 && tx * (n / default_ppt(n)) > kColThreads) tx /= 2;
  return tx;
}
// the inverse-z pass that forms w: bytes per tile line (each block writes one line per z position).  Lines of
// 256 bytes (twice the columns per block, half the blocks per CU) took 161.4 us per 256^3 FP64 step against
// 152.8, FP32 89.1 against 81.1
#ifndef FFS_CURL_LINE
#define FFS_CURL_LINE 128
#endif
template <typename T, int N, bool CURL> __host__ __device__ constexpr int col_txk() {
  return CURL ? col_tx<T>(N, FFS_CURL_LINE) : col_tx<T>(N);
}
// The inverse-z pass that forms w = ik x u (and rho, S) on load issues two loads per point in its
// first stage; 8 points per thread (radix <= 8) instead of 16 halves the registers those loads hold.
// FFS_COL_PPT8: 0 = that variant only, 1 = every column pass, -1 = none.
#ifndef FFS_COL_PPT8
#define FFS_COL_PPT8 1
#endif
// FFS_ZMAP_ARITH: the inverse-z load maps z indices arithmetically (1) or through the table (0)
// XCDs of the device (MI355X: 8); workgroups are dispatched to them round-robin
constexpr int kXcd = 8;
// FFS_CURL_SPLIT: the inverse-z pass that forms w (and rho, S) runs one field per block (1) or
// loops over the fields in each block (0)
#ifndef FFS_CURL_SPLIT
#define FFS_CURL_SPLIT 1
#endif
// FFS_COL_BRANCHFREE: the inverse column passes load every position and zero by select (1)
#ifndef FFS_COL_BRANCHFREE
#define FFS_COL_BRANCHFREE 0
#endif
#ifndef FFS_ZMAP_ARITH
#define FFS_ZMAP_ARITH 1
#endif
template <int N> __host__ __device__ constexpr bool ppt8_ok() { return N % 8 == 0 && plan_even(N, N / 8, 8); }
template <typename T, int N, bool CURL> __host__ __device__ constexpr bool col_ppt8() {
  return ppt8_ok<N>() && ((FFS_COL_PPT8 == 0 && CURL) || FFS_COL_PPT8 == 1) &&
         col_txk<T, N, CURL>() * (N / 8) <= kColThreads;
}
template <typename T, int N, bool CURL> __host__ __device__ constexpr int col_tper() {
  return col_ppt8<T, N, CURL>() ? N / 8 : tper<N>();
}
template <typename T, int N, bool CURL> __host__ __device__ constexpr int col_maxr() {
  return col_ppt8<T, N, CURL>() ? 8 : 16;
}
template <int N> __host__ __device__ constexpr int xrows() { return (64 / tper<N>()) > 0 ? (64 / tper<N>()) : 1; }
// rows per block of the general x pass: at most 256 threads
template <int N, int NSL> __host__ __device__ constexpr int xgen_rows() {
  return (256 / (NSL * tper<N>())) > 0 ? (256 / (NSL * tper<N>())) : 1;
}

// Tunables of the fused x pass: points per lane, minimum waves per SIMD the
// register allocation must allow, and the LDS padding of the exchange.
#ifndef FFS_XNL_PPT64
#define FFS_XNL_PPT64 8
#endif
#ifndef FFS_XNL_PPT32
#define FFS_XNL_PPT32 16
#endif
#ifndef FFS_XNL_WAVES
#define FFS_XNL_WAVES 1
#endif
#ifndef FFS_XNL_BRANCHFREE
#define FFS_XNL_BRANCHFREE 1  // 1: the x pass loads its inputs without branches
#endif
#ifndef FFS_XNL_PREFETCH
#define FFS_XNL_PREFETCH 1  // 1: request every input of a row before its first transform
#endif
#ifndef FFS_XNL_PAD_SHIFT
#define FFS_XNL_PAD_SHIFT 3  // 0: no padding; s > 0: one slot per 2^s values
#endif
#ifndef FFS_COL_WAVES
#define FFS_COL_WAVES 1
#endif

// points per lane of the register-resident x pass and whether it applies
template <typename T, int N> __host__ __device__ constexpr int xreg_ppt() {
  return sizeof(T) == 8 ? FFS_XNL_PPT64 : FFS_XNL_PPT32;
}
template <typename T, int N> __host__ __device__ constexpr int xreg_ldx() {
  return FFS_XNL_PAD_SHIFT > 0 ? N + (N >> FFS_XNL_PAD_SHIFT) + 1 : N + 1;
}
template <typename T, int N> __host__ __device__ constexpr bool xreg_ok() {
  constexpr int P = xreg_ppt<T, N>();
  return N % P == 0 && 64 % (N / P) == 0 && plan_even(N, N / P, P);
}

template <typename T> __device__ __forceinline__ T nanmax(T a, T b) { return (a > b || a != a) ? a : b; }

// Copies the length-N twiddle table into LDS (one copy per block): the FFT
// stages then read their twiddles with LDS latency instead of an L2 round
// trip inside the dependent butterfly chain.  The caller synchronises.
#ifndef FFS_LDS_TW64
#define FFS_LDS_TW64 1
#endif
#ifndef FFS_LDS_TW32
#define FFS_LDS_TW32 1
#endif
template <typename T> __host__ __device__ constexpr bool lds_twiddles() {
  return sizeof(T) == 8 ? FFS_LDS_TW64 : FFS_LDS_TW32;
}
// LDS values a block reserves for the staged twiddle table (none when the table stays in global memory).
// Staged, at 256^3: FP64 step 0.816 -> 0.807 ms (inverse z pass 156 -> 150 us), FP32 0.464 -> 0.459, measured
// side by side, although an FP64 column block then holds 36 KB and one fewer fits per CU.
template <typename T> __host__ __device__ constexpr int tw_lds(int n) { return lds_twiddles<T>() ? n : 0; }
template <typename T, int N>
__device__ __forceinline__ const C<T>* stage_twiddles(C<T>* __restrict__ twl, const C<T>* __restrict__ tw) {
  if constexpr (!lds_twiddles<T>()) return tw;
  for (int i = threadIdx.x; i < N; i += blockDim.x) twl[i] = tw[i];
  return twl;
}

// wave64 + cross-wave max, then one atomic per block on the IEEE bit pattern
// (monotone for non-negative values; NaN sorts above +inf so it propagates), into the block's slot.
__device__ __forceinline__ void block_max_to(double v, unsigned long long* dst) {
  __shared__ double red[16];
  for (int off = 32; off > 0; off >>= 1) v = nanmax(v, __shfl_xor(v, off, 64));
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) red[w] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = red[0];
    const int nw = (blockDim.x + 63) >> 6;
    for (int i = 1; i < nw; ++i) m = nanmax(m, red[i]);
    atomicMax(max_slot(dst, blockIdx.x), (unsigned long long)__double_as_longlong(m));
  }
  // `red` is one array per kernel: a following call must not overwrite it before thread 0 has read it
  __syncthreads();
}

// A store marked non-temporal (streamed past the caches: the value is not read again soon).  The outputs read
// back only a whole pass later are written so: the inverse z pass's (the inverse y pass reads them slab by
// slab), the forward y pass's (read by the forward z pass after the last slab) and the update's right-hand
// side (read at the next step's update), while the slab buffers, which the next pass reads at once from the
// last-level cache, and u, which the next step's inverse z pass reads first, keep ordinary stores.  At 256^3, measured side by side: FP64 step 0.823 -> 0.758 ms (the
// inverse z pass 153 -> 134 us and the inverse y pass 193 -> 178 with the first alone), FP32 0.482 -> 0.462.
#ifndef FFS_CURL_NT
#define FFS_CURL_NT 1  // the inverse-z pass that forms w
#endif
#ifndef FFS_NT_FWDY
#define FFS_NT_FWDY 1  // the forward y pass
#endif
#ifndef FFS_NT_UPDATE
#define FFS_NT_UPDATE 1  // the update's right-hand side (and u with FFS_NT_U)
#endif
#ifndef FFS_NT_U
#define FFS_NT_U 0  // 0: u with ordinary stores (the next step's inverse z pass reads it first; FP32 62 vs 72 us)
#endif
template <typename T> __device__ __forceinline__ void store_nt(C<T>* p, C<T> v) {
  typedef T V2 __attribute__((ext_vector_type(2)));
  const V2 w = {v.x, v.y};
  __builtin_nontemporal_store(w, reinterpret_cast<V2*>(p));
}

// Compute units of the current device (read once per device).
static int device_cus() {
  static std::atomic<int> cus[64];
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) throw std::runtime_error("hipGetDevice failed");
  int v = cus[dev & 63].load(std::memory_order_relaxed);
  if (v == 0) {
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || v <= 0)
      throw std::runtime_error("reading the device's compute unit count failed");
    cus[dev & 63].store(v, std::memory_order_relaxed);
  }
  return v;
}

// Blocks of `threads` threads of a kernel (no dynamic LDS) that one compute unit holds at once.
static int resident_blocks(const void* fn, int threads) {
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, threads, 0) != hipSuccess || nb <= 0)
    throw std::runtime_error("reading a kernel's occupancy failed");
  return nb;
}

// Raises a kernel's dynamic LDS limit to `bytes`, once per process and kernel.  Handles may be
// driven from several host threads, so the first call is serialised by the kernel's once_flag; a
// failure is reported (and retried by the next call) rather than surfacing as a failed launch.
static void lds_limit_once(std::once_flag& once, const void* fn, size_t bytes, const char* what) {
  std::call_once(once, [&] {
    const hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    if (e != hipSuccess)
      throw std::runtime_error(std::string("raising the LDS limit of ") + what + " to " + std::to_string(bytes) +
                               " bytes: " + hipGetErrorString(e));
  });
}

// Launches of the batchable kernels go through launch_call: a kernel's single form takes its arguments as
// before, its batch form an array of call records (one per run, all of one launch shape) with the run index
// folded into grid z; both run the same body.  With a recorder set on the calling thread the launch is recorded, not issued, so
// the launches of several runs' steps can be merged into one launch per pass (issue_batch).
#if FFS_IN_PART(0)
thread_local KRecorder* g_recorder = nullptr;
#endif
// a batch form's call records: read-only, the same for every thread of a block
#define FFS_CALLS(T) const T* __restrict__

// A struct that no thread of the kernel writes, at an address that is the same for the whole block, read
// word by word through the constant address space: every word comes in with a scalar load and the copy
// lives in SGPRs, as kernel arguments do.  (Read through a pointer the compiler cannot prove unwritten,
// fields used late would be fetched with vector loads into VGPRs.)
template <class S> __device__ __forceinline__ S load_const(const S* p) {
#if defined(__HIP_DEVICE_COMPILE__)
  static_assert(sizeof(S) % 4 == 0 && alignof(S) >= 4, "read as whole words");
  typedef const __attribute__((address_space(4))) unsigned cword;
  cword* src = (cword*)p;
  S c;
  unsigned* d = reinterpret_cast<unsigned*>(&c);
#pragma unroll
  for (int k = 0; k < (int)(sizeof(S) / 4); ++k) d[k] = src[k];
  return c;
#else
  return *p;
#endif
}
// run i's call record
template <class Call> __device__ __forceinline__ Call load_call(const Call* calls, int i) { return load_const(calls + i); }

// Every buffer the kernels touch is global memory.  The compiler knows it of pointer kernel arguments but
// not of pointers read from memory (a call record, a struct argument), whose loads and stores it would
// otherwise issue as flat instructions; glb() states it for a pointer that is the same across the block
// (it passes through an SGPR pair).  The empty asm hands the compiler a global-address-space pointer it
// cannot fold back into the generic one, so the accesses through it become global instructions.
template <class P> __device__ __forceinline__ P* glb(P* p) {
  __attribute__((address_space(1))) P* q;
  asm("" : "=s"(q) : "0"(p));
  return (P*)q;
}
__device__ __forceinline__ Geo glb(Geo g) {
  g.ymap = glb(g.ymap);
  g.zmap = glb(g.zmap);
  g.yinv = glb(g.yinv);
  g.zinv = glb(g.zinv);
  g.kx = glb(g.kx);
  g.ky = glb(g.ky);
  g.kz = glb(g.kz);
  return g;
}
__device__ __forceinline__ UpdateArgs glb(UpdateArgs a) {
  a.pf_factor = glb(a.pf_factor);
  a.tau = glb(a.tau);
  a.rho = glb(a.rho);
  a.rrhs = glb(a.rrhs);
  a.rfm1 = glb(a.rfm1);
  a.rfm2 = glb(a.rfm2);
  a.flux = glb(a.flux);
  a.time = glb(a.time);
  a.phi = glb(a.phi);
  a.prhs = glb(a.prhs);
  a.pfm1 = glb(a.pfm1);
  a.pfm2 = glb(a.pfm2);
  a.pflux = glb(a.pflux);
  return a;
}

template <class Call, class F>
static void launch_call(const void* batch, dim3 grid, dim3 block, size_t lds, const Call& c, const char* name,
                        F&& single) {
  if (g_recorder) {
    KLaunch k;
    k.batch = batch;
    k.grid = grid;
    k.block = block;
    k.lds = lds;
    k.name = name;
    k.call.assign(reinterpret_cast<const unsigned char*>(&c), reinterpret_cast<const unsigned char*>(&c) + sizeof(Call));
    g_recorder->launches.push_back(std::move(k));
    return;
  }
  single();
}

#if FFS_IN_PART(0)
void issue_batch(const KLaunch& k, const void* calls, int nrun, hipStream_t s) {
  dim3 grid = k.grid;
  grid.z *= (unsigned)nrun;
  void* args[] = {const_cast<void**>(&calls), &nrun};
  const hipError_t e = hipLaunchKernel(k.batch, grid, k.block, args, k.lds, s);
  if (e != hipSuccess) throw std::runtime_error(std::string("batched launch of ") + k.name + ": " + hipGetErrorString(e));
}
#endif

// ===========================================================================
// Column (y / z) passes.  Block = TX consecutive x columns x TPER threads per
// column; LDS holds one [N][TX] tile.
enum ColKind { INV_Z = 0, INV_Y = 1, FWD_Y = 2, FWD_Z = 3 };

// GRAD: the curl pass also forms i k s of the scalars with an LES flux (a separate instantiation,
// so the plain step's curl pass carries none of that selection in its loads)
template <typename T> struct ColCall {
  Geo g;
  const C<T>* in;
  C<T>* out;
  ColAux<T> aux;
  const C<T>* tw;
  int nzero;
  T scale;
  int z0, nzc;  // y passes: the z planes [z0, z0 + nzc), their (sx, ny) side in a buffer of nzc planes
};

template <typename T, int N, int KIND, bool CURL, bool GRAD>
__device__ __forceinline__ void col_body(Geo g, const C<T>* __restrict__ in, C<T>* __restrict__ out, const ColAux<T>& aux,
                                         const C<T>* __restrict__ tw_g, int nzero, T scale, int bX, int bY, int bZ,
                                         int z0, int nzc) {
  g = glb(g);
  constexpr int TX = col_txk<T, N, CURL>();
  constexpr int TP = col_tper<T, N, CURL>();
  constexpr int MAXR = col_maxr<T, N, CURL>();
  static_assert(TX * TP <= kColThreads, "column block exceeds its launch bound");
  // The split variant of the inverse-z pass runs one field per block.  Its 1D grid is laid out for
  // the round-robin dispatch of workgroups over the XCDs: block L goes to XCD L mod kXcd, so the
  // nzero field blocks of tile t are given the ids with L mod kXcd = t mod kXcd and consecutive
  // L / kXcd; they run on one XCD, back to back, and read the u tile once from HBM into its L2.
  constexpr bool SPLIT = CURL && FFS_CURL_SPLIT;
  int bx, bb, bf = 0;
  if constexpr (SPLIT) {
    const int ntx = (g.mx + TX - 1) / TX;
    const int L = bX, sq = L / kXcd;
    bf = sq % nzero;
    const int t = (sq / nzero) * kXcd + L % kXcd;
    if (t >= ntx * g.my) return;  // whole block: the grid is padded to a multiple of kXcd tiles
    bx = t % ntx;
    bb = t / ntx;
  } else {
    bx = bX;
    bb = bY;
  }
  extern __shared__ unsigned char smem_raw[];
  C<T>* lds = reinterpret_cast<C<T>*>(smem_raw);
  const C<T>* tw = stage_twiddles<T, N>(lds + N * TX, tw_g);
  __syncthreads();
  const int c = threadIdx.x % TX;
  const int q = threadIdx.x / TX;
  const int cx = bx * TX + c;
  const bool valid = cx < g.mx;
  const int b = bb;
  const size_t sx = g.sx;
  const size_t Ssp = sx * g.my * g.mz;  // compact spectral
  const size_t S1 = sx * g.my * g.nz;   // (sx, my, nz)
  const size_t S2 = sx * g.ny * g.nz;   // (sx, ny, nz)
  auto ldsf = [&](int p) -> C<T>& { return lds[p * TX + c]; };

  if constexpr (KIND == INV_Z) {
    const int cy = b;
    // SRC variant: the block loops over the nin physical-space inputs of the
    // nonlinear term -- u (0..2), w = ik x u (3..5) (fourierspacep1!,
    // models.jl:88-95), rho (if aux.rho), phi (if aux.phi), S = sym(ik (x) u) (xx,xy,xz,yy,yz; if
    // aux.les) -- so the u tile is fetched from HBM once and re-read from L2
    const int f0 = SPLIT ? bf : (CURL ? 0 : bZ);
    const int f1 = SPLIT ? bf + 1 : (CURL ? nzero : bZ + 1);
    const T kx = valid ? (T)g.kx[cx] : (T)0, ky = (T)g.ky[cy];
    for (int f = f0; f < f1; ++f) {
      if (f != f0) __syncthreads();
      auto ldv = [&](int p) -> C<T> {
        const int cz0 = FFS_ZMAP_ARITH ? z_compact(g, p) : g.zinv[p];
#if FFS_COL_BRANCHFREE
        const int cz = cz0 < 0 ? 0 : cz0;  // dealiased rows load row 0 and are zeroed by the caller
#else
        const int cz = cz0;
        if (!valid || cz < 0) return mk<T>(0, 0);
#endif
        const size_t idx = cx + sx * (cy + (size_t)g.my * cz);
        if (!CURL || f < 3) return in[f * Ssp + idx];
        const T kz = (T)g.kz[cz];
        if (!aux.nsform && f < 6) {
          C<T> a;  // w = i k x u
          if (f == 3) {
            const C<T> uy = in[Ssp + idx], uz = in[2 * Ssp + idx];
            a = mk<T>(ky * uz.x - kz * uy.x, ky * uz.y - kz * uy.y);
          } else if (f == 4) {
            const C<T> ux = in[idx], uz = in[2 * Ssp + idx];
            a = mk<T>(kz * ux.x - kx * uz.x, kz * ux.y - kx * uz.y);
          } else {
            const C<T> ux = in[idx], uy = in[Ssp + idx];
            a = mk<T>(kx * uy.x - ky * ux.x, kx * uy.y - ky * ux.y);
          }
          return mk<T>(-a.y, a.x);
        }
        int e = f - (aux.nsform ? 3 : 6);
        if (aux.rho) {
          if (e == 0) return glb(aux.rho)[idx];
          --e;
        }
        if (aux.phi) {
          if (e == 0) return glb(aux.phi)[idx];
          --e;
        }
        if (GRAD && e >= 5) {  // i k s of the scalars with an LES flux (fourierspacep1!, models.jl:110-116)
          e -= 5;
          const C<T> sv = e < 3 ? glb(aux.grad1)[idx] : glb(aux.grad2)[idx];
          const int d = e < 3 ? e : e - 3;
          const T kd = d == 0 ? kx : (d == 1 ? ky : kz);
          return mk<T>(-kd * sv.y, kd * sv.x);
        }
        // S_ab = (i k_a u_b + i k_b u_a) / 2 for ab = xx, xy, xz, yy, yz (symouter, models.jl:98)
        const int a_ = (e < 3) ? 0 : 1;
        const int b_ = (e < 3) ? e : e - 2;
        const T ka = a_ == 0 ? kx : ky;
        const T kb = b_ == 0 ? kx : (b_ == 1 ? ky : kz);
        const C<T> ua = in[a_ * Ssp + idx], ub = in[b_ * Ssp + idx];
        const C<T> sum = mk<T>((T)0.5 * (ka * ub.x + kb * ua.x), (T)0.5 * (ka * ub.y + kb * ua.y));
        return mk<T>(-sum.y, sum.x);
      };
      // branch-free: every lane loads (columns past mx read the in-bounds row padding), so the
      // first stage's loads issue back to back; the zero of dealiased / padding positions is a select
      auto ld = [&](int p, int, int) -> C<T> {
#if FFS_COL_BRANCHFREE
        const C<T> v = ldv(p);
        const bool keep = valid && (FFS_ZMAP_ARITH ? z_compact(g, p) : g.zinv[p]) >= 0;
        return keep ? v : mk<T>(0, 0);
#else
        return ldv(p);
#endif
      };
      auto st = [&](int p, int, int, C<T> v) {
        if (valid) {
          if constexpr (FFS_CURL_NT && CURL) store_nt<T>(&out[f * S1 + cx + sx * (cy + (size_t)g.my * p)], v);
          else out[f * S1 + cx + sx * (cy + (size_t)g.my * p)] = v;
        }
      };
      Fft<T, N, +1, TP, false, false, SYNC_BLOCK, MAXR>::run(q, tw, ld, st, ldsf);
    }
  } else if constexpr (KIND == INV_Y) {
    // plane z0 + z of the (sx, my, nz) input into plane z of an output of nzc planes (a z slab)
    const int z = b, zg = z0 + b, f = bZ;
    const size_t S2c = sx * g.ny * nzc;
    auto ld = [&](int p, int, int) -> C<T> {
      const int cy = y_compact(g, p);
#if FFS_COL_BRANCHFREE
      const C<T> v = in[f * S1 + cx + sx * ((cy < 0 ? 0 : cy) + (size_t)g.my * zg)];
      return (valid && cy >= 0) ? v : mk<T>(0, 0);
#else
      if (!valid || cy < 0) return mk<T>(0, 0);
      return in[f * S1 + cx + sx * (cy + (size_t)g.my * zg)];
#endif
    };
    auto st = [&](int p, int, int, C<T> v) {
      if (valid) out[f * S2c + cx + sx * (p + (size_t)g.ny * z)] = v;
    };
    Fft<T, N, +1, TP, false, false, SYNC_BLOCK, MAXR>::run(q, tw, ld, st, ldsf);
  } else if constexpr (KIND == FWD_Y) {
    // plane z of an input of nzc planes (a z slab) into plane z0 + z of the (sx, my, nz) output
    const int z = b, zg = z0 + b, f = bZ;
    const size_t S2c = sx * g.ny * nzc;
    auto ld = [&](int p, int, int) -> C<T> {
      if (!valid) return mk<T>(0, 0);
      return in[f * S2c + cx + sx * (p + (size_t)g.ny * z)];
    };
    auto st = [&](int p, int, int, C<T> v) {
      const int cy = y_compact(g, p);
      if (valid && cy >= 0) {
        if constexpr (FFS_NT_FWDY) store_nt<T>(&out[f * S1 + cx + sx * (cy + (size_t)g.my * zg)], v);
        else out[f * S1 + cx + sx * (cy + (size_t)g.my * zg)] = v;
      }
    };
    Fft<T, N, -1, TP, false, false, SYNC_BLOCK, MAXR>::run(q, tw, ld, st, ldsf);
  } else {  // FWD_Z
    const int cy = b, f = bZ;
    auto ld = [&](int p, int, int) -> C<T> {
      if (!valid) return mk<T>(0, 0);
      return in[f * S1 + cx + sx * (cy + (size_t)g.my * p)];
    };
    auto st = [&](int p, int, int, C<T> v) {
      const int cz = z_compact(g, p);
      if (valid && cz >= 0) {
        C<T> r = cscale<T>(v, scale);
        if (f < nzero && cx == 0 && cy == 0 && cz == 0) r = mk<T>(0, 0);
        const size_t o = cx + sx * (cy + (size_t)g.my * cz);
        glb(aux.dst[f])[o] = r;  // rhs components, density flux, tau, 1/2|u|^2 ...
      }
    };
    Fft<T, N, -1, TP, false, false, SYNC_BLOCK, MAXR>::run(q, tw, ld, st, ldsf);
  }
}

template <typename T, int N, int KIND, bool CURL, bool GRAD = false>
__global__ void __launch_bounds__(kColThreads, FFS_COL_WAVES) k_col(Geo g, const C<T>* __restrict__ in, C<T>* __restrict__ out,
                                            ColAux<T> aux, const C<T>* __restrict__ tw_g, int nzero, T scale,
                                            int z0, int nzc) {
  col_body<T, N, KIND, CURL, GRAD>(g, in, out, aux, tw_g, nzero, scale, blockIdx.x, blockIdx.y, blockIdx.z, z0, nzc);
}
template <typename T, int N, int KIND, bool CURL, bool GRAD = false>
__global__ void __launch_bounds__(kColThreads, FFS_COL_WAVES) k_col_b(FFS_CALLS(ColCall<T>) calls, int nrun) {
  const int nz = gridDim.z / nrun;
  const ColCall<T>& c = calls[blockIdx.z / nz];  // (aux.dst is indexed by the field: read in place)
  col_body<T, N, KIND, CURL, GRAD>(c.g, glb(c.in), glb(c.out), c.aux, glb(c.tw), c.nzero, c.scale, blockIdx.x,
                                   blockIdx.y, blockIdx.z % nz, c.z0, c.nzc);
}

template <typename T, int N, int KIND, bool CURL, bool GRAD = false>
static void run_col(const Geo& g, const void* in, void* out, int nf, int nb, const void* tw, int nzero, double scale,
                    hipStream_t s, const ColAux<T>& aux = ColAux<T>{}, int z0 = 0) {
  constexpr int TX = col_txk<T, N, CURL>();
  constexpr int TP = col_tper<T, N, CURL>();
  static_assert(TX * TP <= kColThreads, "column block exceeds its launch bound");
  const size_t lds = (size_t)(N * TX + tw_lds<T>(N)) * sizeof(C<T>);
  static std::once_flag once;
  lds_limit_once(once, (const void*)&k_col<T, N, KIND, CURL, GRAD>, lds, "a column pass");
  static std::once_flag once_b;
  lds_limit_once(once_b, (const void*)&k_col_b<T, N, KIND, CURL, GRAD>, lds, "a batched column pass");
  const int ntile = (g.mx + TX - 1) / TX * nb;
  const dim3 grid = (CURL && FFS_CURL_SPLIT) ? dim3((unsigned)(nzero * ((ntile + kXcd - 1) / kXcd * kXcd)))
                                             : dim3((g.mx + TX - 1) / TX, nb, CURL ? 1 : nf);
  // the y passes' nb planes from z0 (a slab of nb planes, or all of them)
  const ColCall<T> c{g, (const C<T>*)in, (C<T>*)out, aux, (const C<T>*)tw, nzero, (T)scale, z0, nb};
  launch_call((const void*)&k_col_b<T, N, KIND, CURL, GRAD>, grid, dim3(TX * TP), lds, c, "column pass", [&] {
    k_col<T, N, KIND, CURL, GRAD><<<grid, TX * TP, lds, s>>>(c.g, c.in, c.out, c.aux, c.tw, c.nzero, c.scale, c.z0,
                                                               c.nzc);
  });
}

#ifndef FFS_SIZES
#define FFS_SIZES(X)                                                                                      \
  X(16) X(24) X(32) X(40) X(48) X(64) X(80) X(96) X(128) X(160) X(192) X(256) X(320) X(384) X(512) X(640) \
  X(768) X(1024)
#endif

[[noreturn]] static void bad_size(const char* what, int n) {
  throw std::runtime_error(std::string("unsupported transform length ") + std::to_string(n) + " in " + what);
}

template <typename T>
void launch_inv_z(const Geo& g, const void* u, void* a1, int nf, bool curl, const void* tw, hipStream_t s,
                  const void* rho, const void* phi, int nsform, const void* grad1, const void* grad2) {
  ColAux<T> aux{};
  aux.rho = (const C<T>*)rho;
  aux.phi = (const C<T>*)phi;
  aux.nsform = nsform;
  aux.grad1 = (const C<T>*)grad1;
  aux.grad2 = (const C<T>*)grad2;
  switch (g.nz) {
#define X(n)                                                                                    \
  case n:                                                                                       \
    if (curl && grad1) run_col<T, n, INV_Z, true, true>(g, u, a1, 1, g.my, tw, nf, 1.0, s, aux); \
    else if (curl) run_col<T, n, INV_Z, true>(g, u, a1, 1, g.my, tw, nf, 1.0, s, aux);          \
    else run_col<T, n, INV_Z, false>(g, u, a1, nf, g.my, tw, 0, 1.0, s, aux);                   \
    break;
    FFS_SIZES(X)
#undef X
    default: bad_size("inverse z pass", g.nz);
  }
}

template <typename T>
void launch_inv_y(const Geo& g, const void* a1, void* a2, int nf, const void* tw, hipStream_t s, int z0, int nzc) {
  if (nzc < 0) nzc = g.nz - z0;
  if (z0 < 0 || nzc <= 0 || z0 + nzc > g.nz) throw std::runtime_error("inverse y pass: z slab out of range");
  switch (g.ny) {
#define X(n)                                                            \
  case n: run_col<T, n, INV_Y, false>(g, a1, a2, nf, nzc, tw, 0, 1.0, s, ColAux<T>{}, z0); break;
    FFS_SIZES(X)
#undef X
    default: bad_size("inverse y pass", g.ny);
  }
}

template <typename T>
void launch_fwd_y(const Geo& g, const void* b2, void* b1, int nf, const void* tw, hipStream_t s, int z0, int nzc) {
  if (nzc < 0) nzc = g.nz - z0;
  if (z0 < 0 || nzc <= 0 || z0 + nzc > g.nz) throw std::runtime_error("forward y pass: z slab out of range");
  switch (g.ny) {
#define X(n)                                                            \
  case n: run_col<T, n, FWD_Y, false>(g, b2, b1, nf, nzc, tw, 0, 1.0, s, ColAux<T>{}, z0); break;
    FFS_SIZES(X)
#undef X
    default: bad_size("forward y pass", g.ny);
  }
}

template <typename T>
void launch_fwd_z(const Geo& g, const void* b1, void* const* dst, int nf, int nzero, double scale, const void* tw,
                  hipStream_t s) {
  ColAux<T> aux{};
  if (nf > kMaxFields) throw std::runtime_error("forward z pass: more outputs than kMaxFields");
  for (int f = 0; f < nf; ++f) aux.dst[f] = (C<T>*)dst[f];
  switch (g.nz) {
#define X(n)                                                                        \
  case n: run_col<T, n, FWD_Z, false>(g, b1, nullptr, nf, g.my, tw, nzero, scale, s, aux); break;
    FFS_SIZES(X)
#undef X
    default: bad_size("forward z pass", g.nz);
  }
}

// ===========================================================================
// x passes.  LDS slots of N+1 complex values; TPER threads per slot.  Two real
// rows travel as one complex sequence z = a + i b: the inverse builds
// Z[k] = A[k] + i B[k], Z[N-k] = conj(A[k]) + i conj(B[k]) from the retained
// half spectra, the forward separates A = (Z[k] + conj Z[N-k])/2,
// B = (Z[k] - conj Z[N-k])/(2i).

template <typename T, int N>
__device__ __forceinline__ void build_pair(C<T>* Z, int k, C<T> A, C<T> B) {
  if (k == 0 || k == N / 2) {
    Z[k] = mk<T>(A.x, B.x);  // DC / Nyquist of a real row: imaginary parts are not representable
  } else {
    Z[k] = mk<T>(A.x - B.y, A.y + B.x);
    Z[N - k] = mk<T>(A.x + B.y, B.x - A.y);
  }
}

template <typename T, int N>
__device__ __forceinline__ void split_pair(const C<T>* Z, int k, C<T>& A, C<T>& B) {
  const C<T> zk = Z[k], zn = Z[(N - k) % N];
  A = mk<T>((T)0.5 * (zk.x + zn.x), (T)0.5 * (zk.y - zn.y));
  B = mk<T>((T)0.5 * (zk.y + zn.y), (T)0.5 * (zn.x - zk.x));
}

// calc_les! for the models that read the vorticity w (les_real_calc.jl:80-114): S (xx xy xz yy yz, zz =
// -xx - yy) and w -> the five stored components of t.  AntiSymTen(v) is the cross-product matrix of v
// (AntiSymTen(v) x = v x x), so A = AntiSymTen(-w/2) has A_xy = w_z/2, A_xz = -w_y/2, A_yz = w_x/2 and
// S + A is the transposed velocity gradient; Lie(a, b) = ab - ba; a:b sums every product a_ij b_ij.
template <typename T>
__device__ __forceinline__ void les_stress(const LesArgs& l, const T s[5], const T w[3], T t[5], T* nut_out = nullptr) {
  const T S[3][3] = {{s[0], s[1], s[2]}, {s[1], s[3], s[4]}, {s[2], s[4], -s[0] - s[3]}};
  const T hx = (T)0.5 * w[0], hy = (T)0.5 * w[1], hz = (T)0.5 * w[2];
  const T A[3][3] = {{0, hz, -hy}, {-hz, 0, hx}, {hy, -hx, 0}};  // AntiSymTen(-w/2)
  const T d2 = (T)l.d2;
  T ss = 0;  // S:S
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) ss += S[i][j] * S[i][j];
  // P = Lie(S, AntiSymTen(w/2)) = A S - S A (Silvis, SandP)
  auto lie_p = [&](T P[3][3]) {
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        T v = 0;
#pragma unroll
        for (int k = 0; k < 3; ++k) v += A[i][k] * S[k][j] - S[i][k] * A[k][j];
        P[i][j] = v;
      }
  };
  // c Delta^2 (square(W):S - square(S):S) / (2 S:S), W = A (production_eddy_viscosity, stable_nl_eddy_viscosity)
  auto prod_x = [&](T c) {
    T ws = 0, s3 = 0;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        T w2 = 0, q2 = 0;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          w2 += A[i][k] * A[k][j];
          q2 += S[i][k] * S[k][j];
        }
        ws += w2 * S[i][j];
        s3 += q2 * S[i][j];
      }
    return c * d2 * (ws - s3) / ((T)2 * ss);
  };
  auto eddy_nu = [&](int model, T c) -> T {  // nu_t of an eddy-viscosity model, t = 2 nu_t S
    if (model == LES_SMAGORINSKY) return c * c * d2 * sqrt((T)2 * ss);
    if (model == LES_VREMAN) {  // Vreman_eddy_viscosity: alpha = S + A, beta = Delta^2 alpha' alpha
      T a[3][3], aa = 0;
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          a[i][j] = S[i][j] + A[i][j];
          aa += a[i][j] * a[i][j];
        }
      auto beta = [&](int i, int j) {
        T v = 0;
#pragma unroll
        for (int m = 0; m < 3; ++m) v += a[m][i] * a[m][j];
        return d2 * v;
      };
      const T b11 = beta(0, 0), b22 = beta(1, 1), b33 = beta(2, 2), b12 = beta(0, 1), b13 = beta(0, 2),
              b23 = beta(1, 2);
      T B = b11 * (b22 + b33) + b22 * b33 - (b12 * b12 + b13 * b13 + b23 * b23);
      B = B < (T)0 ? (T)0 : B;
      return c * sqrt(B / aa);
    }
    if (model == LES_PRODUCTION) return ss == (T)0 ? (T)0 : fmax((T)0, prod_x(c));
    return (T)0;  // fakeSmagorinsky: nu_t = 0
  };
  T R[3][3];
  T nut = 0;  // the eddy viscosity the scalars' LES fluxes read
  if (l.model == LES_SANDP) {
    const T nu = eddy_nu(l.smodel, (T)l.c);
    nut = nu;
    T P[3][3];
    lie_p(P);
    const T beta = (T)l.cp * d2;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) R[i][j] = (l.smodel == LES_FAKE_SMAGORINSKY ? (T)0 : (T)2 * nu * S[i][j]) + beta * P[i][j];
  } else if (l.model == LES_SILVIS) {
    T sw[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) sw[i] = S[i][0] * w[0] + S[i][1] * w[1] + S[i][2] * w[2];
    const T f = sqrt(sw[0] * sw[0] + sw[1] * sw[1] + sw[2] * sw[2]) /
                (sqrt(ss) * sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]));  // fvs
    const T c = (T)l.c, cp = (T)l.cp;
    nut = c * c * d2 * f * f * f * sqrt(ss) / (T)2;
    const T nup = cp * d2 * f * f * f * f / (T)4;
    T P[3][3];
    lie_p(P);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) R[i][j] = (T)2 * nut * S[i][j] + nup * P[i][j];
  } else if (l.model == LES_STABLE_NL) {  // c Delta^2 traceless(W^2 - Lie(S, W) - S^2) + 2 mnu S, W = A
    const T c = (T)l.c;
    const T x = prod_x(c);
    const T mnu = ss == (T)0 ? (T)0 : -fmin((T)0, x);
    nut = ss == (T)0 ? (T)0 : fmax((T)0, x);
    T M[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        T v = 0;
#pragma unroll
        for (int k = 0; k < 3; ++k)
          v += A[i][k] * A[k][j] - (S[i][k] * A[k][j] - A[i][k] * S[k][j]) - S[i][k] * S[k][j];
        M[i][j] = v;
      }
    const T tr = (M[0][0] + M[1][1] + M[2][2]) / (T)3;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) R[i][j] = c * d2 * (M[i][j] - (i == j ? tr : (T)0)) + (T)2 * mnu * S[i][j];
  } else {  // Vreman, productionViscosity, Smagorinsky
    const T nu = eddy_nu(l.model, (T)l.c);
    nut = nu;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) R[i][j] = (T)2 * nu * S[i][j];
  }
  t[0] = R[0][0];
  t[1] = R[0][1];
  t[2] = R[0][2];
  t[3] = R[1][1];
  t[4] = R[1][2];
  if (nut_out) *nut_out = nut;
}

// a scalar's flux with its LES part: -s u + nu_t grad s [+ (cv/2) w x grad s] (les_real_calc.jl:120-136,
// models.jl:264-273)
template <typename T>
__device__ __forceinline__ void scalar_flux(T s, T ux, T uy, T uz, int les, T nut, const T* g, const T* w, T cv,
                                            T* out) {
  T f[3] = {(-s) * ux, (-s) * uy, (-s) * uz};
  if (les) {
    T l[3] = {nut * g[0], nut * g[1], nut * g[2]};
    if (les == 2) {
      const T h = cv / (T)2;
      l[0] += h * (w[1] * g[2] - w[2] * g[1]);
      l[1] += h * (w[2] * g[0] - w[0] * g[2]);
      l[2] += h * (w[0] * g[1] - w[1] * g[0]);
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) f[c] += l[c];
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) out[c] = f[c];
}

// Fused nonlinear x pass, general field set (LDS slots; any transform length).
// Inputs (a2, y/z-transformed): u (0..2), w (3..5), [rho], [S xx xy xz yy yz].
// Outputs (b2, x-forward transformed): u x w (0..2), [-rho u (3)], [tau =
// 2 (c Delta)^2 |S| S (5), calc_les! les_real_calc.jl:77-79 with
// |S| = sqrt(2 S:S)], [1/2|u|^2].  Each row holds NSL slots of two real
// sequences; outputs overwrite inputs slot by slot after the pointwise phase.
template <typename T> struct XGCall {
  Geo g;
  const C<T>* a2;
  C<T>* b2;
  const C<T>* tw;
  XOps ops;
  unsigned long long* umax;
  unsigned long long* rhomax;
};
template <typename T, int N, int NSL>
__device__ __forceinline__ void x_gen_body(Geo g, const C<T>* __restrict__ a2, C<T>* __restrict__ b2,
                                           const C<T>* __restrict__ tw_g, XOps ops, unsigned long long* __restrict__ umax,
                                           unsigned long long* __restrict__ rhomax, int bX) {
  g = glb(g);
  constexpr int TP = tper<N>();
  constexpr int ROWS = xgen_rows<N, NSL>();
  constexpr int LD = N + 1;
  constexpr int NTHR = ROWS * NSL * TP;
  extern __shared__ unsigned char smem_raw[];
  C<T>* lds = reinterpret_cast<C<T>*>(smem_raw);
  const C<T>* tw = stage_twiddles<T, N>(lds + ROWS * NSL * LD, tw_g);  // published by the barrier below
  const int tid = threadIdx.x;
  const int nrows = g.ny * g.nz;
  const int row0 = bX * ROWS;
  const int mx = g.mx;
  const size_t S2 = (size_t)g.sx * nrows;
  const int nin = ops.nin, nout = ops.nout;

  // A: gather retained half spectra into Hermitian-packed pairs
  for (int w = tid; w < ROWS * NSL * (N / 2 + 1); w += NTHR) {
    const int sl = w / (N / 2 + 1), k = w % (N / 2 + 1);
    const int r = sl / NSL, pr = sl % NSL;
    const int row = row0 + r;
    C<T> A = mk<T>(0, 0), B = mk<T>(0, 0);
    if (row < nrows && k < mx) {
      if (2 * pr < nin) A = a2[(size_t)(2 * pr) * S2 + (size_t)row * g.sx + k];
      if (2 * pr + 1 < nin) B = a2[(size_t)(2 * pr + 1) * S2 + (size_t)row * g.sx + k];
    }
    build_pair<T, N>(lds + sl * LD, k, A, B);
  }
  __syncthreads();
  const int sl = tid / TP, q = tid % TP;
  C<T>* Zs = lds + sl * LD;
  Fft<T, N, +1, TP, true, true>::run(
      q, tw, [&](int p, int, int) { return Zs[p]; }, [&](int p, int, int, C<T> v) { Zs[p] = v; },
      [&](int p) -> C<T>& { return Zs[p]; });
  __syncthreads();
  // C: pointwise
  double lmax = 0.0, rmax = 0.0;
  for (int w = tid; w < ROWS * N; w += NTHR) {
    const int r = w / N, n = w % N;
    C<T>* Z = lds + (r * NSL) * LD;
    // real slot k of this point is component k & 1 of Z[(k >> 1) * LD + n]; the slots a field occupies
    // depend on the run's options, so values are read from and written to LDS at those slots directly
    // (a register array indexed at run time would live in private memory).  Every input is read
    // before the first output is written: inputs and outputs share the slots.
    T* zr = reinterpret_cast<T*>(Z);
    auto IN = [&](int k) -> T { return zr[2 * ((k >> 1) * LD + n) + (k & 1)]; };
    auto OUT = [&](int k, T v) { zr[2 * ((k >> 1) * LD + n) + (k & 1)] = v; };
    const T ux = IN(0), uy = IN(1), uz = IN(2);
    const int i0 = ops.ns ? 3 : 6;  // rho, phi, then S, then the scalars' gradients
    const T wv3[3] = {ops.ns ? (T)0 : IN(3), ops.ns ? (T)0 : IN(4), ops.ns ? (T)0 : IN(5)};
    const int i_rho = i0, i_phi = i0 + (ops.rho ? 1 : 0);
    const int i_s = i_phi + (ops.phi ? 1 : 0);
    const T rho = ops.rho ? IN(i_rho) : (T)0;
    const T phi = ops.phi ? IN(i_phi) : (T)0;
    const bool need_s = ops.les || ops.rho_les || ops.phi_les;
    T sv[5];
#pragma unroll
    for (int e = 0; e < 5; ++e) sv[e] = need_s ? IN(i_s + e) : (T)0;
    T g_rho[3], g_phi[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      g_rho[c] = ops.rho_les ? IN(i_s + 5 + c) : (T)0;
      g_phi[c] = ops.phi_les ? IN(i_s + 5 + (ops.rho_les ? 3 : 0) + c) : (T)0;
    }
    int o;
    if (ops.ns) {  // traceless(symouter(u, -u)) (models.jl:262): xx xy xz yy yz
      const T xx = -(ux * ux), yy = -(uy * uy), zz = -(uz * uz);
      const T t = (xx + yy + zz) / (T)3;
      OUT(0, xx - t);
      OUT(1, -(ux * uy));
      OUT(2, -(ux * uz));
      OUT(3, yy - t);
      OUT(4, -(uy * uz));
      o = 5;
    } else {
      OUT(0, uy * wv3[2] - uz * wv3[1]);
      OUT(1, uz * wv3[0] - ux * wv3[2]);
      OUT(2, ux * wv3[1] - uy * wv3[0]);
      o = 3;
    }
    // the eddy viscosity of the scalars' LES fluxes
    T nut = 0;
    if (ops.rho_les || ops.phi_les) {
      if (ops.lm.model == LES_SMAGORINSKY) {  // nu_t = alpha |S| (les_real_calc.jl:77-79)
        const T szz = -sv[0] - sv[3];
        const T ss = sv[0] * sv[0] + sv[3] * sv[3] + szz * szz + (T)2 * (sv[1] * sv[1] + sv[2] * sv[2] + sv[4] * sv[4]);
        nut = (T)ops.les_alpha * sqrt((T)2 * ss);
      } else {
        T tv[5];
        les_stress<T>(ops.lm, sv, wv3, tv, &nut);
      }
    }
    if (ops.rho) {
      T f[3];
      scalar_flux<T>(rho, ux, uy, uz, ops.rho_les, nut, g_rho, wv3, (T)ops.rho_cv, f);
#pragma unroll
      for (int c = 0; c < 3; ++c) OUT(o + c, f[c]);
      o += 3;
      if (row0 + r < nrows) rmax = nanmax(rmax, (double)fabs(rho));
    }
    if (ops.phi) {  // -phi u (the passive scalar enters no time-step bound)
      T f[3];
      scalar_flux<T>(phi, ux, uy, uz, ops.phi_les, nut, g_phi, wv3, (T)ops.phi_cv, f);
#pragma unroll
      for (int c = 0; c < 3; ++c) OUT(o + c, f[c]);
      o += 3;
    }
    if (ops.les && ops.lm.model != LES_SMAGORINSKY) {  // the models that read w (slots 3..5)
      T tv[5];
      les_stress<T>(ops.lm, sv, wv3, tv);
#pragma unroll
      for (int e = 0; e < 5; ++e) OUT(o + e, tv[e]);
      o += 5;
    } else if (ops.les) {
      const T szz = -sv[0] - sv[3];
      const T ss = sv[0] * sv[0] + sv[3] * sv[3] + szz * szz + (T)2 * (sv[1] * sv[1] + sv[2] * sv[2] + sv[4] * sv[4]);
      const T f = (T)ops.les_2alpha * sqrt((T)2 * ss);
#pragma unroll
      for (int e = 0; e < 5; ++e) OUT(o + e, f * sv[e]);
      o += 5;
    }
    if (ops.ke) OUT(o++, (T)0.5 * (ux * ux + uy * uy + uz * uz));
    for (int k = o; k < 2 * NSL; ++k) OUT(k, (T)0);  // the unused slots transform as zeros
    if (row0 + r < nrows) lmax = nanmax(lmax, (double)nanmax(fabs(ux), nanmax(fabs(uy), fabs(uz))));
  }
  __syncthreads();
  // D: forward transforms
  Fft<T, N, -1, TP, true, true>::run(
      q, tw, [&](int p, int, int) { return Zs[p]; }, [&](int p, int, int, C<T> v) { Zs[p] = v; },
      [&](int p) -> C<T>& { return Zs[p]; });
  __syncthreads();
  // E: separate the pairs, keep k < mx
  for (int w = tid; w < ROWS * NSL * mx; w += NTHR) {
    const int k = w % mx, rj = w / mx, r = rj / NSL, j = rj % NSL;
    const int row = row0 + r;
    if (row >= nrows || 2 * j >= nout) continue;
    C<T> A, B;
    split_pair<T, N>(lds + (r * NSL + j) * LD, k, A, B);
    b2[(size_t)(2 * j) * S2 + (size_t)row * g.sx + k] = A;
    if (2 * j + 1 < nout) b2[(size_t)(2 * j + 1) * S2 + (size_t)row * g.sx + k] = B;
  }
  if (umax) block_max_to(lmax, umax);
  if (rhomax) block_max_to(rmax, rhomax);
}
template <typename T, int N, int NSL>
__global__ void __launch_bounds__(256) k_x_gen(Geo g, const C<T>* __restrict__ a2, C<T>* __restrict__ b2,
                                              const C<T>* __restrict__ tw_g, XOps ops,
                                              unsigned long long* __restrict__ umax,
                                              unsigned long long* __restrict__ rhomax) {
  x_gen_body<T, N, NSL>(g, a2, b2, tw_g, ops, umax, rhomax, blockIdx.x);
}
template <typename T, int N, int NSL>
__global__ void __launch_bounds__(256) k_x_gen_b(FFS_CALLS(XGCall<T>) calls, int nrun) {
  const XGCall<T> c = load_call(calls, blockIdx.z);
  x_gen_body<T, N, NSL>(c.g, glb(c.a2), glb(c.b2), glb(c.tw), c.ops, glb(c.umax), glb(c.rhomax), blockIdx.x);
}

// Register-resident variant of the fused x pass (used when every FFT stage
// splits evenly over the TPER lanes of a row, i.e. power-of-two nx): TPER
// consecutive lanes own one row; lane q keeps the row's values at the
// positions p = q (mod TPER) in registers from the first load to the last
// store, so LDS carries only the one exchange inside each transform (plus the
// k <-> N-k pairing at the end), synchronised per wavefront.
// What one launch of the register-resident x pass computes (MODE):
//   0  u x w (+ 1/2|u|^2 into slot ke_slot) from u, w (slots 0..5), max|u_i|
//   1  the Smagorinsky stress tau = c2a |S| S (c2a = 2 (c Delta)^2, |S| = sqrt(2 S:S),
//      calc_les! les_real_calc.jl:77-79) from S xx xy xz yy yz (slots in0..in0+4) into out0..out0+4
//   2  the density flux -rho u from u (slots 0..2) and rho (slot in0) into out0..out0+2, max|rho|
struct XRegArgs {
  int in0, out0, ke_slot;
  double c2a;
};
template <typename T> struct XCall {
  Geo g;
  const C<T>* a2;
  C<T>* b2;
  const C<T>* tw;
  unsigned long long* vmax;
  XRegArgs xa;
};
// MODE 0: u x w (+ 1/2|u|^2); 1: the Smagorinsky stress; 2: -s u of a scalar s; 3: the conservation
// form's traceless(-u u) (+ 1/2|u|^2)
template <int MODE> __host__ __device__ constexpr int xreg_npin() { return MODE == 2 || MODE == 3 ? 2 : 3; }
template <int MODE, bool KE> __host__ __device__ constexpr int xreg_npout() {
  return MODE == 1 || MODE == 3 ? 3 : 2;
}

template <typename T, int N, int MODE, bool KE>
__device__ __forceinline__ void x_reg_body(Geo g, const C<T>* __restrict__ a2, C<T>* __restrict__ b2, const C<T>* __restrict__ tw_g,
                                           unsigned long long* __restrict__ vmax, XRegArgs xa, int bX) {
  g = glb(g);
  constexpr int PPT = xreg_ppt<T, N>();
  constexpr int TP = N / PPT;
  constexpr int GPW = 64 / TP;  // rows per wavefront
  constexpr int LDX = xreg_ldx<T, N>();
  constexpr int NSTG = plan_ns(N, PPT);
  constexpr int M0 = (N / plan_r(N, 0, PPT)) / TP;         // register stride of first-stage legs
  constexpr int ML = (N / plan_r(N, NSTG - 1, PPT)) / TP;  // register stride of last-stage legs
  constexpr int NPI = xreg_npin<MODE>();
  constexpr int NPO = xreg_npout<MODE, KE>();
  static_assert(xreg_ok<T, N>(), "register-resident x pass needs an even plan");
  extern __shared__ unsigned char smem_raw[];
  C<T>* twl = reinterpret_cast<C<T>*>(smem_raw);
  C<T>* lds = twl + tw_lds<T>(N);
  const C<T>* tw = stage_twiddles<T, N>(twl, tw_g);
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int grp = lane / TP, q = lane % TP;
  const int nrows = g.ny * g.nz;
  const int row = (bX * (int)(blockDim.x >> 6) + wave) * GPW + grp;
  const bool row_ok = row < nrows;
  const int mx = g.mx;
  const size_t S2 = (size_t)g.sx * nrows;
  const size_t roff = (size_t)(row_ok ? row : 0) * g.sx;
  C<T>* X = lds + (size_t)(wave * GPW + grp) * LDX;
  auto xref = [&](int p) -> C<T>& {
    if constexpr (FFS_XNL_PAD_SHIFT > 0) return X[p + (p >> FFS_XNL_PAD_SHIFT)];
    else return X[p];
  };
  // input slots of pair j (B < 0: the pair's imaginary row is zero)
  auto slot_a = [&](int j) -> int {
    if constexpr (MODE == 0) return 2 * j;
    else if constexpr (MODE == 1) return xa.in0 + 2 * j;
    else return 2 * j;
  };
  auto slot_b = [&](int j) -> int {
    if constexpr (MODE == 0) return 2 * j + 1;
    else if constexpr (MODE == 1) return j < 2 ? xa.in0 + 2 * j + 1 : -1;
    else if constexpr (MODE == 3) return j == 0 ? 1 : -1;
    else return j == 0 ? 1 : xa.in0;
  };

  C<T> R[3][PPT];
  // Z = A + iB of a pair of real rows at position p, from their retained half spectra
  auto zval = [&](const C<T>* A, const C<T>* B, int p) -> C<T> {
#if FFS_XNL_BRANCHFREE
    // straight-line: every lane loads (index 0 where the value is zero), so no branch
    // separates the loads and they issue back to back
    const int k = N - p;
    const bool lo = p < mx;
    const bool hi = !lo && k < mx && 2 * p != N;
    const int idx = lo ? p : (hi ? k : 0);
    const C<T> a = A[idx];
    const C<T> b = B ? B[idx] : mk<T>(0, 0);
    const bool real = p == 0 || 2 * p == N;
    C<T> z = lo ? (real ? mk<T>(a.x, b.x) : mk<T>(a.x - b.y, a.y + b.x)) : mk<T>(a.x + b.y, b.x - a.y);
    return (row_ok && (lo || hi)) ? z : mk<T>(0, 0);
#else
    if (!row_ok) return mk<T>(0, 0);
    if (p < mx) {
      const C<T> a = A[p], b = B ? B[p] : mk<T>(0, 0);
      if (p == 0 || 2 * p == N) return mk<T>(a.x, b.x);
      return mk<T>(a.x - b.y, a.y + b.x);
    }
    const int k = N - p;
    if (k < mx && 2 * p != N) {
      const C<T> a = A[k], b = B ? B[k] : mk<T>(0, 0);
      return mk<T>(a.x + b.y, b.x - a.y);
    }
    return mk<T>(0, 0);
#endif
  };
  auto row_a = [&](int j) { return a2 + (size_t)slot_a(j) * S2 + roff; };
  auto row_b = [&](int j) -> const C<T>* { return slot_b(j) < 0 ? nullptr : a2 + (size_t)slot_b(j) * S2 + roff; };
#if FFS_XNL_PREFETCH
  // every input of the row is requested before the first transform (lane q's first-stage legs
  // are its register positions p = q + TP m), so the loads of all pairs are in flight together
#pragma unroll
  for (int j = 0; j < NPI; ++j)
#pragma unroll
    for (int m = 0; m < PPT; ++m) R[j][m] = zval(row_a(j), row_b(j), q + TP * m);
#endif
  // inverse transforms of the input pairs; Z built from the half spectra on load
#pragma unroll
  for (int j = 0; j < NPI; ++j) {
    auto ld = [&](int p, int t, int r) -> C<T> {
      if constexpr (FFS_XNL_PREFETCH) return R[j][t + r * M0];
      else return zval(row_a(j), row_b(j), p);
    };
    auto st = [&](int, int t, int r, C<T> v) { R[j][t + r * ML] = v; };
    if (j > 0) __syncwarp();
    Fft<T, N, +1, TP, false, false, SYNC_WAVE, PPT>::run(q, tw, ld, st, xref);
  }
  // pointwise products
  double lmax = 0.0;
#pragma unroll
  for (int m = 0; m < PPT; ++m) {
    if constexpr (MODE == 0) {  // u x w, 1/2|u|^2, max|u_i|
      const T ux = R[0][m].x, uy = R[0][m].y, uz = R[1][m].x, wx = R[1][m].y, wy = R[2][m].x, wz = R[2][m].y;
      T ke = 0;
      if constexpr (KE) ke = (T)0.5 * (ux * ux + uy * uy + uz * uz);
      R[0][m] = mk<T>(uy * wz - uz * wy, uz * wx - ux * wz);
      R[1][m] = mk<T>(ux * wy - uy * wx, ke);
      if (row_ok) lmax = nanmax(lmax, (double)nanmax(fabs(ux), nanmax(fabs(uy), fabs(uz))));
    } else if constexpr (MODE == 1) {  // tau = c2a |S| S
      const T sxx = R[0][m].x, sxy = R[0][m].y, sxz = R[1][m].x, syy = R[1][m].y, syz = R[2][m].x;
      const T szz = -sxx - syy;
      const T ss = sxx * sxx + syy * syy + szz * szz + (T)2 * (sxy * sxy + sxz * sxz + syz * syz);
      const T f = (T)xa.c2a * sqrt((T)2 * ss);
      R[0][m] = mk<T>(f * sxx, f * sxy);
      R[1][m] = mk<T>(f * sxz, f * syy);
      R[2][m] = mk<T>(f * syz, (T)0);
    } else if constexpr (MODE == 3) {  // traceless(symouter(u, -u)) (models.jl:262), 1/2|u|^2, max|u_i|
      const T ux = R[0][m].x, uy = R[0][m].y, uz = R[1][m].x;
      const T xx = -(ux * ux), yy = -(uy * uy), zz = -(uz * uz);
      const T t = (xx + yy + zz) / (T)3;
      T ke = 0;
      if constexpr (KE) ke = (T)0.5 * (ux * ux + uy * uy + uz * uz);
      R[0][m] = mk<T>(xx - t, -(ux * uy));
      R[1][m] = mk<T>(-(ux * uz), yy - t);
      R[2][m] = mk<T>(-(uy * uz), ke);
      if (row_ok) lmax = nanmax(lmax, (double)nanmax(fabs(ux), nanmax(fabs(uy), fabs(uz))));
    } else {  // -rho u, max|rho|
      const T ux = R[0][m].x, uy = R[0][m].y, uz = R[1][m].x, rho = R[1][m].y;
      R[0][m] = mk<T>((-rho) * ux, (-rho) * uy);
      R[1][m] = mk<T>((-rho) * uz, (T)0);
      if (row_ok) lmax = nanmax(lmax, (double)fabs(rho));
    }
  }
  // forward transforms of the output pairs, split into the half spectra k < mx
#pragma unroll
  for (int j = 0; j < NPO; ++j) {
    auto ld = [&](int, int t, int r) { return R[j][t + r * M0]; };
    auto st = [&](int, int t, int r, C<T> v) { R[j][t + r * ML] = v; };
    __syncwarp();
    Fft<T, N, -1, TP, false, false, SYNC_WAVE, PPT>::run(q, tw, ld, st, xref);
    __syncwarp();
#pragma unroll
    for (int m = 0; m < PPT; ++m) xref(q + TP * m) = R[j][m];
    __syncwarp();
    // output slots of the pair's halves (-1: not written)
    int oa, ob;
    if constexpr (MODE == 0) {
      oa = 2 * j;
      ob = j == 0 ? 1 : (KE ? xa.ke_slot : -1);
    } else if constexpr (MODE == 1) {
      oa = xa.out0 + 2 * j;
      ob = j < 2 ? xa.out0 + 2 * j + 1 : -1;
    } else if constexpr (MODE == 3) {
      oa = 2 * j;
      ob = j < 2 ? 2 * j + 1 : (KE ? xa.ke_slot : -1);
    } else {
      oa = xa.out0 + 2 * j;
      ob = j == 0 ? xa.out0 + 1 : -1;
    }
#pragma unroll
    for (int m = 0; m < PPT; ++m) {
      const int k = q + TP * m;
      if (row_ok && k < mx) {
        const C<T> zk = R[j][m], zn = xref((N - k) % N);
        b2[(size_t)oa * S2 + roff + k] = mk<T>((T)0.5 * (zk.x + zn.x), (T)0.5 * (zk.y - zn.y));
        if (ob >= 0) b2[(size_t)ob * S2 + roff + k] = mk<T>((T)0.5 * (zk.y + zn.y), (T)0.5 * (zn.x - zk.x));
      }
    }
  }
  if (MODE != 1 && vmax) block_max_to(lmax, vmax);
}

template <typename T, int N, int MODE, bool KE>
__global__ void __launch_bounds__(128, FFS_XNL_WAVES) k_x_nl_reg(Geo g, const C<T>* __restrict__ a2, C<T>* __restrict__ b2,
                                                   const C<T>* __restrict__ tw_g, unsigned long long* __restrict__ vmax,
                                                   XRegArgs xa) {
  x_reg_body<T, N, MODE, KE>(g, a2, b2, tw_g, vmax, xa, blockIdx.x);
}
template <typename T, int N, int MODE, bool KE>
__global__ void __launch_bounds__(128, FFS_XNL_WAVES) k_x_nl_reg_b(FFS_CALLS(XCall<T>) calls, int nrun) {
  const XCall<T> c = load_call(calls, blockIdx.z);
  x_reg_body<T, N, MODE, KE>(c.g, glb(c.a2), glb(c.b2), glb(c.tw), glb(c.vmax), c.xa, blockIdx.x);
}

// ---------------------------------------------------------------------------
// The x pass with mirror-paired lanes (FP64, 8 points per lane, TP = N/8 lanes per row).  The
// inverse transforms start with a radix-4 stage and the forward transforms end with one (the same
// plan run in reverse order), and those stages deal their N/4 butterflies to the lanes in mirror
// pairs: lane q owns butterflies q and N/4 - q (lane 0: 0 and N/8), i.e. the positions
// p(t, r) = b_t + r N/4, which hold Z[p] and Z[N - p] together.  So each lane loads just the
// retained modes k < mx among its own positions -- every mode of a row is read once, no load for the
// 2/3 zero band -- and builds both Z[k] and Z[N - k] from them, and after the forward transforms it
// splits Z[k], Z[N - k] into the two half spectra in registers, without the LDS round trip of a
// split by position.  Lo legs r < 2 lie below N/2; the hi legs are their mirrors:
//   q > 0: (t, r) <-> (1 - t, 3 - r);  q = 0: (0, r) <-> (0, (4 - r) % 4), (1, r) <-> (1, 3 - r).
#ifndef FFS_XNL_PAIR
#define FFS_XNL_PAIR 1
#endif
template <int N> __host__ __device__ constexpr int xpair_maxr() { return plan_r(N, 0, 8) == 4 ? 8 : 4; }
// FP32 too: at 256^3 the mirror-paired pass (102 VGPRs, four waves per SIMD) takes 62.8 us per slab where the
// register pass (16 points per lane, 231 VGPRs, two waves) takes 87.9 (FP32 step 0.515 -> 0.476 ms)
#ifndef FFS_XPAIR_F32
#define FFS_XPAIR_F32 1
#endif
template <typename T, int N> __host__ __device__ constexpr bool xpair_ok() {
  return FFS_XNL_PAIR && (sizeof(T) == 8 || FFS_XPAIR_F32) && N % 8 == 0 && 64 % (N / 8) == 0 &&
         plan_r(N, 0, xpair_maxr<N>()) == 4 && plan_even(N, N / 8, xpair_maxr<N>());
}

#ifndef FFS_XPAIR_WAVES
#define FFS_XPAIR_WAVES 2  // minimum waves per SIMD the mirror-paired pass's registers must allow (3 spills)
#endif
// The twiddle table's pointer, opaque to the optimiser: each transform of a row then reads its own
// twiddles instead of one set being computed once and kept live, in registers, across all the row's
// transforms.  A table left in global memory comes back as a global pointer (as glb()).
template <typename T> __device__ __forceinline__ const C<T>* opaque_tw(const C<T>* p) {
  if constexpr (lds_twiddles<T>()) {
    asm volatile("" : "+s"(p));
    return p;
  } else {
    __attribute__((address_space(1))) const C<T>* q;
    asm volatile("" : "=s"(q) : "0"(p));
    return (const C<T>*)q;
  }
}
// The same for a table in a static LDS array: the pointer stays visibly derived from the array (an opaque zero
// offset is added instead), so the compiler still tells its reads apart from LDS written by asynchronous
// global-to-LDS loads and sets no wait for those in front of them.
template <typename T> __device__ __forceinline__ const C<T>* opaque_tw_static(const C<T>* p) {
  int z = 0;
  asm volatile("" : "+s"(z));
  return p + z;
}

// The looping form (plan x_pair 2; FP64 u x w rows; DW wavefronts per block): each block runs row groups bX,
// bX + gridDim.x, ... over a grid of resident blocks and, while a group's last inverse transform, products and
// forward transforms run, its wavefronts copy the next group's first field pair from global memory straight
// into LDS (global_load_lds: no registers held), so that pair is there when the next group starts.  Its LDS
// is static, so the compiler keeps the exchange and twiddle reads apart from the copies in flight.  At 256^3
// (64-plane slabs) the pass took 208 us per step where the one-group form took 225 (198 with the buffer loads
// below).
template <typename T, int N> __host__ __device__ constexpr bool xdma_ok() {
  return sizeof(T) == 8 && xpair_ok<T, N>();
}
constexpr int kXLoopWaves = 2;  // wavefronts per block of the looping form

// Exchange layout of the paired pass: position p of a row at p plus pad slots, added (not XOR-ed), so the
// compiler folds a stage's per-position constants into the LDS instructions' offsets.  Measured on the device
// (SQ_LDS_BANK_CONFLICT of the 256-point row's exchanges replayed per candidate, then the x pass itself): FP64
// keeps one pad per 8 values (2.0 conflict cycles per access; no two-term pad did better; one per 16: 5.3, the
// x pass +3 %; XOR swizzles halve the conflicts but spill the looping form's registers, +12 %); FP32 takes
// three per 32 (fewer LDS instructions and registers, five waves per SIMD: x pass 125.3 -> 119-120 us per
// 256^3 step, although its accesses conflict more).
template <typename T> __host__ __device__ constexpr int xpair_slot(int p) {
  return sizeof(C<T>) == 16 ? p + (p >> 3) : p + 3 * (p >> 5);
}
// values per row: the slot of the last position, plus one
template <typename T, int N> __host__ __device__ constexpr int xpair_ldx() { return xpair_slot<T>(N - 1) + 1; }

// A wavefront's rows of a field as a buffer resource (raw, stride 0): loads through it check the offset
// against the range and read zeros past it, so positions outside the retained modes and rows past the grid
// need no select on the loaded values.  kBufOut is an offset past every range.  The descriptor's last word,
// 0x00020000, is the gfx9-family data format of 32-bit raw buffer accesses (checked on the device: in-range
// lanes read their values, out-of-range lanes zeros, in the register and the LDS-copy forms).
constexpr unsigned kBufOut = 0xFFFFFFF0u;
#if defined(__HIP_DEVICE_COMPILE__)
using BufRsrc = __amdgpu_buffer_rsrc_t;
template <typename T> __device__ __forceinline__ BufRsrc rows_rsrc(const C<T>* rows, int nvals) {
  return __builtin_amdgcn_make_buffer_rsrc((void*)rows, (short)0, nvals * (int)sizeof(C<T>), 0x00020000);
}
template <typename T> __device__ __forceinline__ C<T> buf_load(BufRsrc r, unsigned off) {
  if constexpr (sizeof(C<T>) == 16) return __builtin_bit_cast(C<T>, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0));
  else return __builtin_bit_cast(C<T>, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
}
// 16 bytes per lane into LDS at dst + 16 * lane (dst the same for the whole wavefront)
__device__ __forceinline__ void buf_load_lds16(BufRsrc r, void* dst, unsigned off) {
  __builtin_amdgcn_raw_ptr_buffer_load_lds(r, (__attribute__((address_space(3))) void*)dst, 16, off, 0, 0, 0);
}
#else  // the host pass parses the device bodies but never runs them
struct BufRsrc {};
template <typename T> BufRsrc rows_rsrc(const C<T>*, int) { return {}; }
template <typename T> C<T> buf_load(BufRsrc, unsigned) { return {}; }
inline void buf_load_lds16(BufRsrc, void*, unsigned) {}
#endif

template <typename T, int N, int MODE, bool KE, int DW = 0>
__device__ __forceinline__ void x_pair_body(Geo g, const C<T>* __restrict__ a2, C<T>* __restrict__ b2, const C<T>* __restrict__ tw_g,
                                           unsigned long long* __restrict__ vmax, XRegArgs xa, int bX) {
  g = glb(g);
  constexpr int TP = N / 8;
  constexpr int GPW = 64 / TP;  // rows per wavefront
  constexpr int MAXR = xpair_maxr<N>();
  constexpr int LDX = xpair_ldx<T, N>();
  constexpr int NSTG = plan_ns(N, MAXR);
  constexpr int M = N / 4;  // butterflies of the radix-4 end stages
  // register stride of the inverse's last-stage legs = of the forward's first-stage legs
  constexpr int ML = (N / plan_r(N, NSTG - 1, MAXR)) / TP;
  static_assert(ML == (N / plan_r(N, 0, MAXR, true)) / TP, "the two plans meet at the same positions");
  constexpr int NPI = xreg_npin<MODE>();
  constexpr int NPO = xreg_npout<MODE, KE>();
  static_assert(xpair_ok<T, N>(), "mirror-paired x pass: a radix-4 end stage, an even plan");
  constexpr bool DMA = DW > 0;
  static_assert(!DMA || sizeof(C<T>) == 16, "the staged copies move one complex FP64 value per lane");
  // staged field pair of a wavefront: [field][chunk c][lane], chunk c holding modes c TP + (lane % TP) of
  // row lane / TP; modes k < N/2 take 4 chunks
  constexpr int SCH = 4 * 64;
  C<T>* twl;
  C<T>* lds;
  C<T>* stg = nullptr;
  const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  if constexpr (DMA) {
    __shared__ C<T> s_tw[tw_lds<T>(N) > 0 ? tw_lds<T>(N) : 1];
    __shared__ C<T> s_x[DW * GPW * LDX];
    __shared__ C<T> s_stage[DW * 2 * SCH];
    twl = s_tw;
    lds = s_x;
    stg = s_stage + wave * (2 * SCH);
  } else {
    extern __shared__ unsigned char smem_raw[];
    twl = reinterpret_cast<C<T>*>(smem_raw);
    lds = twl + tw_lds<T>(N);
  }
  const C<T>* tw = stage_twiddles<T, N>(twl, tw_g);
  __syncthreads();
  const int grp = lane / TP, q = lane % TP;
  const int nrows = g.ny * g.nz;
  const int wpb = DMA ? DW : (int)(blockDim.x >> 6);
  const int nbg = (nrows + wpb * GPW - 1) / (wpb * GPW);  // row groups of wpb GPW rows
  const int mx = g.mx;
  const size_t S2 = (size_t)g.sx * nrows;
  const unsigned lro = (unsigned)(grp * g.sx);  // the lane's row within its wavefront's GPW rows
  C<T>* X = lds + (size_t)(wave * GPW + grp) * LDX;
  auto xref = [&](int p) -> C<T>& {
    return X[xpair_slot<T>(p)];
  };
  auto twp = [&]() {
    if constexpr (DMA && lds_twiddles<T>()) return opaque_tw_static<T>(tw);
    else return opaque_tw<T>(tw);
  };
  auto slot_a = [&](int j) -> int {
    if constexpr (MODE == 1) return xa.in0 + 2 * j;
    else return 2 * j;
  };
  auto slot_b = [&](int j) -> int {
    if constexpr (MODE == 0) return 2 * j + 1;
    else if constexpr (MODE == 1) return j < 2 ? xa.in0 + 2 * j + 1 : -1;
    else if constexpr (MODE == 3) return j == 0 ? 1 : -1;
    else return j == 0 ? 1 : xa.in0;
  };
  const bool q0 = q == 0;
  const int bt1 = q0 ? M / 2 : M - q;
  auto pos = [&](int t, int r) { return (t == 0 ? q : bt1) + r * M; };
  const C<T> zero = mk<T>(0, 0);
  // a group's first field pair into this wavefront's stage: all 4 chunks of both fields (modes past mx and
  // rows past the grid read as zeros through the range check; a missing second field repeats the first, never
  // used), so that the number of copies in flight is fixed and the compiler's waits for older loads can count
  // past them
  auto stage_pair0 = [&](int bgn) {
    if constexpr (DMA) {
      const int row0n = (bgn * DW + wave) * GPW;
      const bool okn = row0n + grp < nrows;
#pragma unroll
      for (int f = 0; f < 2; ++f) {
        const int sl = f == 0 || slot_b(0) < 0 ? slot_a(0) : slot_b(0);
        const auto RF = rows_rsrc<T>(a2 + (size_t)sl * S2 + (size_t)row0n * g.sx, GPW * g.sx);
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          const int k = c * TP + q;
          const unsigned o = (okn && k < mx) ? (lro + (unsigned)k) * (unsigned)sizeof(C<T>) : kBufOut;
          buf_load_lds16(RF, stg + f * SCH + c * 64, o);
        }
      }
    }
  };
  if constexpr (DMA) {
    if (bX < nbg) stage_pair0(bX);
  }
  // max|.| of the lane's values (fmax drops NaN) and whether any was NaN: together the NaN-propagating maximum
  T vmax_l = 0;
  bool nan_l = false;
  auto track = [&](T a, T b, T c) {
    vmax_l = fmax(vmax_l, fmax(fabs(a), fmax(fabs(b), fabs(c))));
    nan_l = nan_l || a != a || b != b || c != c;
  };
  auto group = [&](int bg) {  // the rows of row group bg
  const int row = (bg * wpb + wave) * GPW + grp;
  const bool row_ok = row < nrows;

  // the modes a pair needs: A[k], B[k] at the lo legs, read through buffer resources over the wavefront's
  // rows (zeros for k >= mx and rows past the grid); pair j + 1's loads are issued before pair j's transform
  const int row0 = (bg * wpb + wave) * GPW;
  const unsigned rbase = (unsigned)(row_ok ? row : 0) * (unsigned)g.sx;
  C<T> la[2][2][2], lb[2][2][2];  // [buffer][t][r]
  auto fetch = [&](int j, C<T> (&fa)[2][2], C<T> (&fb)[2][2]) {
    const int sb = slot_b(j);
    const auto RA = rows_rsrc<T>(a2 + (size_t)slot_a(j) * S2 + (size_t)row0 * g.sx, GPW * g.sx);
    const auto RB = rows_rsrc<T>(a2 + (size_t)(sb < 0 ? 0 : sb) * S2 + (size_t)row0 * g.sx, GPW * g.sx);
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
      for (int r = 0; r < 2; ++r) {
        const int k = pos(t, r);
        const unsigned o = (row_ok && k < mx) ? (lro + (unsigned)k) * (unsigned)sizeof(C<T>) : kBufOut;
        fa[t][r] = buf_load<T>(RA, o);
        fb[t][r] = sb < 0 ? zero : buf_load<T>(RB, o);
      }
  };
  C<T> R[3][8];
  if constexpr (DMA) {  // the first pair from the stage
    __syncwarp();
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
      for (int r = 0; r < 2; ++r) {
        const int k = pos(t, r);  // < N/2: within the 4 chunks
        const int si = (k / TP) * 64 + grp * TP + k % TP;
        la[0][t][r] = stg[si];
        lb[0][t][r] = slot_b(0) < 0 ? zero : stg[SCH + si];
      }
  } else {
    fetch(0, la[0], lb[0]);
  }
#pragma unroll
  for (int j = 0; j < NPI; ++j) {
    if (j + 1 < NPI) fetch(j + 1, la[(j + 1) & 1], lb[(j + 1) & 1]);
    // the next group's first pair (the last group: its own again, unused), once the group's register loads have
    // all arrived (vmcnt(0), which the last pair's products wait for anyway): the compiler does not count later
    // waits past a copy into LDS, so a register load still outstanding behind one would wait for the copy too
    if constexpr (DMA) {
      if (j == NPI - 1) {
        __builtin_amdgcn_s_waitcnt(0x0F70);
        stage_pair0(bg + (int)gridDim.x < nbg ? bg + (int)gridDim.x : bg);
      }
    }
    // Z of the pair at the lane's 8 positions: A + iB at the lo legs, conj(A) + i conj(B) mirrored
    C<T> Z[2][4], mh[2][2];
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
      for (int r = 0; r < 2; ++r) {
        // positions outside the retained modes hold loaded zeros, so Z and mh are +0 there without a select;
        // the mean mode (lane 0's first position) takes the real parts alone
        const C<T> a = la[j & 1][t][r], b = lb[j & 1][t][r];
        Z[t][r] = (t == 0 && r == 0 && q0) ? mk<T>(a.x, b.x) : mk<T>(a.x - b.y, a.y + b.x);
        mh[t][r] = mk<T>(a.x + b.y, b.x - a.y);
      }
    Z[0][2] = q0 ? zero : mh[1][1];  // lane 0: the Nyquist position, zero (N/2 >= mx)
    Z[0][3] = q0 ? mh[0][1] : mh[1][0];
    Z[1][2] = q0 ? mh[1][1] : mh[0][1];
    Z[1][3] = q0 ? mh[1][0] : mh[0][0];
    auto ld = [&](int, int t, int r) -> C<T> { return Z[t][r]; };
    auto st = [&](int, int t, int r, C<T> v) { R[j][t + r * ML] = v; };
    if (j > 0 || DMA) __syncwarp();
    Fft<T, N, +1, TP, false, false, SYNC_WAVE, MAXR, 0, false, 0>::run(q, twp(), ld, st, xref);
  }
  // pointwise products at positions q + TP m
#pragma unroll
  for (int m = 0; m < 8; ++m) {
    if constexpr (MODE == 0) {  // u x w, 1/2|u|^2, max|u_i|
      const T ux = R[0][m].x, uy = R[0][m].y, uz = R[1][m].x, wx = R[1][m].y, wy = R[2][m].x, wz = R[2][m].y;
      T ke = 0;
      if constexpr (KE) ke = (T)0.5 * (ux * ux + uy * uy + uz * uz);
      R[0][m] = mk<T>(uy * wz - uz * wy, uz * wx - ux * wz);
      R[1][m] = mk<T>(ux * wy - uy * wx, ke);
      if (row_ok) track(ux, uy, uz);
    } else if constexpr (MODE == 1) {  // tau = c2a |S| S
      const T sxx = R[0][m].x, sxy = R[0][m].y, sxz = R[1][m].x, syy = R[1][m].y, syz = R[2][m].x;
      const T szz = -sxx - syy;
      const T ss = sxx * sxx + syy * syy + szz * szz + (T)2 * (sxy * sxy + sxz * sxz + syz * syz);
      const T f = (T)xa.c2a * sqrt((T)2 * ss);
      R[0][m] = mk<T>(f * sxx, f * sxy);
      R[1][m] = mk<T>(f * sxz, f * syy);
      R[2][m] = mk<T>(f * syz, (T)0);
    } else if constexpr (MODE == 3) {  // traceless(symouter(u, -u)) (models.jl:262), 1/2|u|^2, max|u_i|
      const T ux = R[0][m].x, uy = R[0][m].y, uz = R[1][m].x;
      const T xx = -(ux * ux), yy = -(uy * uy), zz = -(uz * uz);
      const T t = (xx + yy + zz) / (T)3;
      T ke = 0;
      if constexpr (KE) ke = (T)0.5 * (ux * ux + uy * uy + uz * uz);
      R[0][m] = mk<T>(xx - t, -(ux * uy));
      R[1][m] = mk<T>(-(ux * uz), yy - t);
      R[2][m] = mk<T>(-(uy * uz), ke);
      if (row_ok) track(ux, uy, uz);
    } else {  // -rho u, max|rho|
      const T ux = R[0][m].x, uy = R[0][m].y, uz = R[1][m].x, rho = R[1][m].y;
      R[0][m] = mk<T>((-rho) * ux, (-rho) * uy);
      R[1][m] = mk<T>((-rho) * uz, (T)0);
      if (row_ok) track(rho, rho, rho);
    }
  }
  // forward transforms (the plan reversed, mirror pairs in its last stage), split in registers
#pragma unroll
  for (int j = 0; j < NPO; ++j) {
    C<T> W[2][4];
    auto ld = [&](int, int t, int r) { return R[j][t + r * ML]; };
    auto st = [&](int, int t, int r, C<T> v) { W[t][r] = v; };
    __syncwarp();
    Fft<T, N, -1, TP, false, false, SYNC_WAVE, MAXR, 0, true, NSTG - 1>::run(q, twp(), ld, st, xref);
    int oa, ob;
    if constexpr (MODE == 0) {
      oa = 2 * j;
      ob = j == 0 ? 1 : (KE ? xa.ke_slot : -1);
    } else if constexpr (MODE == 1) {
      oa = xa.out0 + 2 * j;
      ob = j < 2 ? xa.out0 + 2 * j + 1 : -1;
    } else if constexpr (MODE == 3) {
      oa = 2 * j;
      ob = j < 2 ? 2 * j + 1 : (KE ? xa.ke_slot : -1);
    } else {
      oa = xa.out0 + 2 * j;
      ob = j == 0 ? xa.out0 + 1 : -1;
    }
    C<T> zn[2][2];  // Z[N - k] of the lo legs
    zn[0][0] = q0 ? W[0][0] : W[1][3];
    zn[0][1] = q0 ? W[0][3] : W[1][2];
    zn[1][0] = q0 ? W[1][3] : W[0][3];
    zn[1][1] = q0 ? W[1][2] : W[0][2];
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
      for (int r = 0; r < 2; ++r) {
        const int k = pos(t, r);
        if (row_ok && k < mx) {
          const C<T> zk = W[t][r], zm = zn[t][r];
          const unsigned o = rbase + (unsigned)k;
          (b2 + (size_t)oa * S2)[o] = mk<T>((T)0.5 * (zk.x + zm.x), (T)0.5 * (zk.y - zm.y));
          if (ob >= 0) (b2 + (size_t)ob * S2)[o] = mk<T>((T)0.5 * (zk.y + zm.y), (T)0.5 * (zm.x - zk.x));
        }
      }
  }
  };
  if constexpr (DMA) {
    for (int bg = bX; bg < nbg; bg += (int)gridDim.x) group(bg);
    __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): no copy into LDS outlives the block
  } else {
    group(bX);
  }
  if (MODE != 1 && vmax) block_max_to(nan_l ? __builtin_nan("") : (double)vmax_l, vmax);
}

template <typename T, int N, int MODE, bool KE>
__global__ void __launch_bounds__(128, FFS_XPAIR_WAVES) k_x_pair(Geo g, const C<T>* __restrict__ a2, C<T>* __restrict__ b2,
                                                   const C<T>* __restrict__ tw_g, unsigned long long* __restrict__ vmax,
                                                   XRegArgs xa) {
  x_pair_body<T, N, MODE, KE>(g, a2, b2, tw_g, vmax, xa, blockIdx.x);
}
template <typename T, int N, int MODE, bool KE>
__global__ void __launch_bounds__(128, FFS_XPAIR_WAVES) k_x_pair_b(FFS_CALLS(XCall<T>) calls, int nrun) {
  const XCall<T> c = load_call(calls, blockIdx.z);
  x_pair_body<T, N, MODE, KE>(c.g, glb(c.a2), glb(c.b2), glb(c.tw), glb(c.vmax), c.xa, blockIdx.x);
}
// the looping form: DW wavefronts per block, static LDS
template <typename T, int N, int MODE, bool KE, int DW>
__global__ void __launch_bounds__(64 * DW, FFS_XPAIR_WAVES) k_x_pair_l(Geo g, const C<T>* __restrict__ a2, C<T>* __restrict__ b2,
                                                        const C<T>* __restrict__ tw_g, unsigned long long* __restrict__ vmax,
                                                        XRegArgs xa) {
  if constexpr (xdma_ok<T, N>()) x_pair_body<T, N, MODE, KE, DW>(g, a2, b2, tw_g, vmax, xa, blockIdx.x);
}
template <typename T, int N, int MODE, bool KE, int DW>
__global__ void __launch_bounds__(64 * DW, FFS_XPAIR_WAVES) k_x_pair_lb(FFS_CALLS(XCall<T>) calls, int nrun) {
  if constexpr (xdma_ok<T, N>()) {
    const XCall<T> c = load_call(calls, blockIdx.z);
    x_pair_body<T, N, MODE, KE, DW>(c.g, glb(c.a2), glb(c.b2), glb(c.tw), glb(c.vmax), c.xa, blockIdx.x);
  }
}

// C2R of nf fields (rows paired 2m, 2m+1) into the padded real layout
// (2*nxh, ny, nz), padding entries zeroed.  Only the rows [row0, row0 + nrow) of each field are
// transformed and written (row0 even; row = y + ny z), at row - row0 of an output of nrow rows per
// field, as values of type O.
template <typename T, int N, typename O = T>
__global__ void __launch_bounds__(256) k_x_c2r(Geo g, const C<T>* __restrict__ a2, O* __restrict__ real,
                                              const C<T>* __restrict__ tw_g, int nf, int row0, int nrow) {
  constexpr int TP = tper<N>();
  constexpr int NSLOT = xrows<N>() * 3;
  constexpr int LD = N + 1;
  constexpr int NTHR = NSLOT * TP;
  extern __shared__ unsigned char smem_raw[];
  C<T>* lds = reinterpret_cast<C<T>*>(smem_raw);
  const C<T>* tw = stage_twiddles<T, N>(lds + NSLOT * LD, tw_g);  // published by the first __syncthreads below
  const int tid = threadIdx.x;
  const int nrows = g.ny * g.nz;
  const int npairs = (nrow + 1) / 2;
  const int mx = g.mx;
  const size_t S2 = (size_t)g.sx * nrows;
  const int rl = 2 * g.nxh;
  const size_t RS = (size_t)rl * nrow;
  const int p0 = blockIdx.x * NSLOT;

  for (int w = tid; w < NSLOT * (N / 2 + 1); w += NTHR) {
    const int sl = w / (N / 2 + 1), k = w % (N / 2 + 1);
    const int pidx = p0 + sl;
    C<T> A = mk<T>(0, 0), B = mk<T>(0, 0);
    if (pidx < nf * npairs && k < mx) {
      const int f = pidx / npairs, row = row0 + 2 * (pidx % npairs);
      A = a2[f * S2 + (size_t)row * g.sx + k];
      if (row + 1 < row0 + nrow) B = a2[f * S2 + (size_t)(row + 1) * g.sx + k];
    }
    build_pair<T, N>(lds + sl * LD, k, A, B);
  }
  __syncthreads();
  const int sl = tid / TP, q = tid % TP;
  C<T>* Zs = lds + sl * LD;
  Fft<T, N, +1, TP, true, true>::run(
      q, tw, [&](int p, int, int) { return Zs[p]; }, [&](int p, int, int, C<T> v) { Zs[p] = v; },
      [&](int p) -> C<T>& { return Zs[p]; });
  __syncthreads();
  for (int w = tid; w < NSLOT * rl; w += NTHR) {
    const int sl2 = w / rl, n = w % rl;
    const int pidx = p0 + sl2;
    if (pidx >= nf * npairs) continue;
    const int f = pidx / npairs, row = 2 * (pidx % npairs);  // relative to row0
    const C<T> z = (n < N) ? lds[sl2 * LD + n] : mk<T>(0, 0);
    real[f * RS + (size_t)row * rl + n] = (O)z.x;
    if (row + 1 < nrow) real[f * RS + (size_t)(row + 1) * rl + n] = (O)z.y;
  }
}

// R2C of nf padded real fields into (mx, ny, nz) half spectra (k < mx kept).
template <typename T, int N>
__global__ void __launch_bounds__(256) k_x_r2c(Geo g, const T* __restrict__ real, C<T>* __restrict__ b2,
                                              const C<T>* __restrict__ tw_g, int nf) {
  constexpr int TP = tper<N>();
  constexpr int NSLOT = xrows<N>() * 3;
  constexpr int LD = N + 1;
  constexpr int NTHR = NSLOT * TP;
  extern __shared__ unsigned char smem_raw[];
  C<T>* lds = reinterpret_cast<C<T>*>(smem_raw);
  const C<T>* tw = stage_twiddles<T, N>(lds + NSLOT * LD, tw_g);  // published by the first __syncthreads below
  const int tid = threadIdx.x;
  const int nrows = g.ny * g.nz;
  const int npairs = (nrows + 1) / 2;
  const int mx = g.mx;
  const size_t S2 = (size_t)g.sx * nrows;
  const int rl = 2 * g.nxh;
  const size_t RS = (size_t)rl * nrows;
  const int p0 = blockIdx.x * NSLOT;

  for (int w = tid; w < NSLOT * N; w += NTHR) {
    const int sl = w / N, n = w % N;
    const int pidx = p0 + sl;
    T a = 0, b = 0;
    if (pidx < nf * npairs) {
      const int f = pidx / npairs, row = 2 * (pidx % npairs);
      a = real[f * RS + (size_t)row * rl + n];
      if (row + 1 < nrows) b = real[f * RS + (size_t)(row + 1) * rl + n];
    }
    lds[sl * LD + n] = mk<T>(a, b);
  }
  __syncthreads();
  const int sl = tid / TP, q = tid % TP;
  C<T>* Zs = lds + sl * LD;
  Fft<T, N, -1, TP, true, true>::run(
      q, tw, [&](int p, int, int) { return Zs[p]; }, [&](int p, int, int, C<T> v) { Zs[p] = v; },
      [&](int p) -> C<T>& { return Zs[p]; });
  __syncthreads();
  for (int w = tid; w < NSLOT * mx; w += NTHR) {
    const int sl2 = w / mx, k = w % mx;
    const int pidx = p0 + sl2;
    if (pidx >= nf * npairs) continue;
    const int f = pidx / npairs, row = 2 * (pidx % npairs);
    C<T> A, B;
    split_pair<T, N>(lds + sl2 * LD, k, A, B);
    b2[f * S2 + (size_t)row * g.sx + k] = A;
    if (row + 1 < nrows) b2[f * S2 + (size_t)(row + 1) * g.sx + k] = B;
  }
}

template <typename T, int N>
static size_t x_lds_bytes() {
  return ((size_t)3 * xrows<N>() * (N + 1) + tw_lds<T>(N)) * sizeof(C<T>);
}

template <typename T, int N, int NSL>
static void run_x_gen(const Geo& g, const void* a2, void* b2, const XOps& ops, unsigned long long* umax,
                      unsigned long long* rhomax, const void* tw, hipStream_t s) {
  constexpr int ROWS = xgen_rows<N, NSL>();
  if (ROWS * NSL * tper<N>() > 256)  // the kernel's launch bound
    throw std::runtime_error("nonlinear x pass: " + std::to_string(NSL) + " field pairs of length " + std::to_string(N) +
                             " exceed one block");
  const size_t lds = ((size_t)ROWS * NSL * (N + 1) + tw_lds<T>(N)) * sizeof(C<T>);
  static std::once_flag once;
  lds_limit_once(once, (const void*)&k_x_gen<T, N, NSL>, lds, "the general x pass");
  static std::once_flag once_b;
  lds_limit_once(once_b, (const void*)&k_x_gen_b<T, N, NSL>, lds, "the batched general x pass");
  const int nrows = g.ny * g.nz;
  const XGCall<T> c{g, (const C<T>*)a2, (C<T>*)b2, (const C<T>*)tw, ops, umax, rhomax};
  const dim3 grid((nrows + ROWS - 1) / ROWS), block(ROWS * NSL * tper<N>());
  launch_call((const void*)&k_x_gen_b<T, N, NSL>, grid, block, lds, c, "general x pass", [&] {
    k_x_gen<T, N, NSL><<<grid, block, lds, s>>>(c.g, c.a2, c.b2, c.tw, c.ops, c.umax, c.rhomax);
  });
}

template <typename T, int N, int MODE, bool KE>
static void launch_xreg_plain(const Geo& g, const void* a2, void* b2, const void* tw, unsigned long long* vmax,
                              const XRegArgs& xa, hipStream_t s, int waves) {
  constexpr int TP = N / xreg_ppt<T, N>();
  if (waves != 1 && waves != 2) throw std::runtime_error("x pass: 1 or 2 wavefronts per block");
  const int RPB = waves * (64 / TP);
  const size_t lds = ((size_t)RPB * xreg_ldx<T, N>() + tw_lds<T>(N)) * sizeof(C<T>);
  static std::once_flag once;
  lds_limit_once(once, (const void*)&k_x_nl_reg<T, N, MODE, KE>, lds, "the x pass");
  static std::once_flag once_b;
  lds_limit_once(once_b, (const void*)&k_x_nl_reg_b<T, N, MODE, KE>, lds, "the batched x pass");
  const int nrows = g.ny * g.nz;
  const XCall<T> c{g, (const C<T>*)a2, (C<T>*)b2, (const C<T>*)tw, vmax, xa};
  const dim3 grid((nrows + RPB - 1) / RPB);
  launch_call((const void*)&k_x_nl_reg_b<T, N, MODE, KE>, grid, dim3(64 * waves), lds, c, "x pass", [&] {
    k_x_nl_reg<T, N, MODE, KE><<<grid, 64 * waves, lds, s>>>(c.g, c.a2, c.b2, c.tw, c.vmax, c.xa);
  });
}

template <typename T, int N, int MODE, bool KE>
static void launch_xreg(const Geo& g, const void* a2, void* b2, const void* tw, unsigned long long* vmax,
                        const XRegArgs& xa, hipStream_t s, const XOps& ops) {
  constexpr bool PAIR = xpair_ok<T, N>();
  // the mirror-paired pass assumes the x Nyquist mode is dealiased (mx <= N/2; cutoff < 1)
  if (g.mx > N / 2 || !PAIR || !ops.pair) {
    if constexpr (xreg_ok<T, N>()) {
      launch_xreg_plain<T, N, MODE, KE>(g, a2, b2, tw, vmax, xa, s, ops.waves);
      return;
    } else {
      throw std::runtime_error("x pass: no register variant for this length");
    }
  }
  constexpr int TP = N / 8;
  const int WPB = ops.waves;  // wavefronts per block
  if (WPB != 1 && WPB != 2) throw std::runtime_error("x pass: 1 or 2 wavefronts per block");
  const int RPB = WPB * (64 / TP);
  const size_t lds = ((size_t)RPB * xpair_ldx<T, N>() + tw_lds<T>(N)) * sizeof(C<T>);
  const int nrows = g.ny * g.nz;
  if constexpr (MODE == 0 && xdma_ok<T, N>()) {
    if (ops.pair == 2) {  // the looping form: static LDS, a grid of resident blocks (x_waves does not apply)
      constexpr int DW = kXLoopWaves;
      const int nbg = (nrows + DW * (64 / TP) - 1) / (DW * (64 / TP));
      const XCall<T> c{g, (const C<T>*)a2, (C<T>*)b2, (const C<T>*)tw, vmax, xa};
      static const int resident = resident_blocks((const void*)&k_x_pair_l<T, N, MODE, KE, DW>, 64 * DW);
      const dim3 grid(std::max(1, std::min(nbg, device_cus() * resident)));
      launch_call((const void*)&k_x_pair_lb<T, N, MODE, KE, DW>, grid, dim3(64 * DW), 0, c, "x pass", [&] {
        k_x_pair_l<T, N, MODE, KE, DW><<<grid, 64 * DW, 0, s>>>(c.g, c.a2, c.b2, c.tw, c.vmax, c.xa);
      });
      return;
    }
  }
  if constexpr (PAIR) {
    static std::once_flag once;
    lds_limit_once(once, (const void*)&k_x_pair<T, N, MODE, KE>, lds, "the x pass");
    static std::once_flag once_b;
    lds_limit_once(once_b, (const void*)&k_x_pair_b<T, N, MODE, KE>, lds, "the batched x pass");
    const XCall<T> c{g, (const C<T>*)a2, (C<T>*)b2, (const C<T>*)tw, vmax, xa};
    const dim3 grid((nrows + RPB - 1) / RPB);
    launch_call((const void*)&k_x_pair_b<T, N, MODE, KE>, grid, dim3(64 * WPB), lds, c, "x pass", [&] {
      k_x_pair<T, N, MODE, KE><<<grid, 64 * WPB, lds, s>>>(c.g, c.a2, c.b2, c.tw, c.vmax, c.xa);
    });
  }
}

// The x pass a run's options select: 0 for the register pass (one launch per product mode), else
// the field-pair slots per row of the general LDS pass, which holds u, w, S and the scalars'
// gradients of a row together
template <typename T, int N>
static int x_general_slots(const XOps& ops) {
  const bool plain = !ops.rho && !ops.les && !ops.phi;
  const bool w_les = (ops.les && les_needs_w(ops.lm)) || ops.rho_les || ops.phi_les;
  if constexpr (xreg_ok<T, N>()) {
    if ((ops.split || plain) && !w_les) return 0;
  }
  return (std::max(ops.nin, ops.nout) + 1) / 2;
}

// the general pass's slot counts are instantiated as 3, 4, 6, 8 and 10; one row's slots must fit the
// kernel's 256-thread launch bound
static int x_gen_bucket(int nsl) {
  return nsl <= 3 ? 3 : nsl <= 4 ? 4 : nsl <= 6 ? 6 : nsl <= 8 ? 8 : nsl <= 10 ? 10 : -1;
}

template <typename T, int N>
static std::string x_check(const XOps& ops) {
  const int nsl = x_general_slots<T, N>(ops);
  if (nsl == 0) return "";
  const int b = x_gen_bucket(nsl);
  if (b < 0) return "the x pass holds at most 10 field pairs per row (" + std::to_string(nsl) + " needed)";
  if (b * tper<N>() > 256)
    return "the x pass of length " + std::to_string(N) + " holds at most " + std::to_string(256 / tper<N>()) +
           " field pairs per row (" + std::to_string(nsl) + " needed with these options)";
  return "";
}

#if FFS_IN_PART(0)
bool length_supported(int n) {
  switch (n) {
#define X(m) case m:
    FFS_SIZES(X)
#undef X
    return true;
    default: return false;
  }
}
std::string compiled_lengths() {
  std::string s;
#define X(m) s += (s.empty() ? "" : " ") + std::to_string(m);
  FFS_SIZES(X)
#undef X
  return s;
}
#endif

template <typename T>
std::string x_pass_check(int n, const XOps& ops) {
  switch (n) {
#define X(m) \
  case m: return x_check<T, m>(ops);
    FFS_SIZES(X)
#undef X
    default: return "unsupported transform length " + std::to_string(n);
  }
}
#if FFS_IN_PART(0)
template std::string x_pass_check<float>(int, const XOps&);
template std::string x_pass_check<double>(int, const XOps&);
#endif

template <typename T, int N>
static void x_plan_options_n(int mx, const XOps& ops, int* pair, int* waves, int* split) {
  XOps a = ops, b = ops;
  a.split = 1;
  b.split = 0;
  const bool reg_split = x_general_slots<T, N>(a) == 0, reg_joint = x_general_slots<T, N>(b) == 0;
  *split = (reg_split != reg_joint && x_check<T, N>(b).empty()) ? 1 : 0;
  const bool reg = x_general_slots<T, N>(ops) == 0;
  *waves = reg ? 1 : 0;
  *pair = (reg && xpair_ok<T, N>() && mx <= N / 2) ? ((xdma_ok<T, N>() && !ops.ns) ? 2 : 1) : 0;
}
template <typename T>
void x_plan_options(int n, int mx, const XOps& ops, int* pair, int* waves, int* split) {
  switch (n) {
#define X(m) \
  case m: return x_plan_options_n<T, m>(mx, ops, pair, waves, split);
    FFS_SIZES(X)
#undef X
    default: *pair = *waves = *split = 0;
  }
}
#if FFS_IN_PART(0)
template void x_plan_options<float>(int, int, const XOps&, int*, int*, int*);
template void x_plan_options<double>(int, int, const XOps&, int*, int*, int*);
#endif

template <typename T, int N>
static void run_x_nl(const Geo& g, const void* a2, void* b2, const XOps& ops, unsigned long long* umax,
                     unsigned long long* rhomax, const void* tw, hipStream_t s) {
  if constexpr (xreg_ok<T, N>()) {
    if (x_general_slots<T, N>(ops) == 0) {
      // u x w or traceless(-u u) (+ 1/2|u|^2, always the last output slot), then -rho u, -phi u and
      // tau as launches of their own
      XRegArgs xa{0, 0, ops.nout - 1, 0.0};
      if (ops.ns) {
        if (ops.ke) launch_xreg<T, N, 3, true>(g, a2, b2, tw, umax, xa, s, ops);
        else launch_xreg<T, N, 3, false>(g, a2, b2, tw, umax, xa, s, ops);
      } else {
        if (ops.ke) launch_xreg<T, N, 0, true>(g, a2, b2, tw, umax, xa, s, ops);
        else launch_xreg<T, N, 0, false>(g, a2, b2, tw, umax, xa, s, ops);
      }
      const int ns = (ops.rho ? 1 : 0) + (ops.phi ? 1 : 0);
      const int i0 = ops.ns ? 3 : 6, o0 = ops.ns ? 5 : 3;  // first scalar input / output
      if (ops.rho) {
        XRegArgs xr{i0, o0, 0, 0.0};
        launch_xreg<T, N, 2, false>(g, a2, b2, tw, rhomax, xr, s, ops);
      }
      if (ops.phi) {  // -phi u through the same kernel, without a maximum
        XRegArgs xp{i0 + (ops.rho ? 1 : 0), o0 + (ops.rho ? 3 : 0), 0, 0.0};
        launch_xreg<T, N, 2, false>(g, a2, b2, tw, nullptr, xp, s, ops);
      }
      if (ops.les) {
        XRegArgs xl{i0 + ns, o0 + 3 * ns, 0, ops.les_2alpha};
        launch_xreg<T, N, 1, false>(g, a2, b2, tw, nullptr, xl, s, ops);
      }
      return;
    }
  }
  const int nsl = x_general_slots<T, N>(ops);
  if (nsl > 10) throw std::runtime_error("nonlinear x pass: more fields than the general pass holds");
  if (nsl <= 3) run_x_gen<T, N, 3>(g, a2, b2, ops, umax, rhomax, tw, s);
  else if (nsl <= 4) run_x_gen<T, N, 4>(g, a2, b2, ops, umax, rhomax, tw, s);
  else if (nsl <= 6) run_x_gen<T, N, 6>(g, a2, b2, ops, umax, rhomax, tw, s);
  else if (nsl <= 8) run_x_gen<T, N, 8>(g, a2, b2, ops, umax, rhomax, tw, s);
  else run_x_gen<T, N, 10>(g, a2, b2, ops, umax, rhomax, tw, s);
}

template <typename T>
void launch_x_nl(const Geo& g, const void* a2, void* b2, const XOps& ops, unsigned long long* umax,
                 unsigned long long* rhomax, const void* tw, hipStream_t s) {
  switch (g.nrx) {
#define X(n)                                                     \
  case n: run_x_nl<T, n>(g, a2, b2, ops, umax, rhomax, tw, s); break;
    FFS_SIZES(X)
#undef X
    default: bad_size("nonlinear x pass", g.nrx);
  }
}

template <typename T>
void launch_x_c2r(const Geo& g, const void* a2, T* real, int nf, const void* tw, hipStream_t s) {
  const int npairs = (g.ny * g.nz + 1) / 2;
  switch (g.nrx) {
#define X(n)                                                                                      \
  case n: {                                                                                       \
    constexpr int NS = 3 * xrows<n>();                                                            \
    dim3 grid((nf * npairs + NS - 1) / NS);                                                       \
    k_x_c2r<T, n><<<grid, NS * tper<n>(), x_lds_bytes<T, n>(), s>>>(g, (const C<T>*)a2, real, (const C<T>*)tw, nf, \
                                                                      0, g.ny * g.nz);                               \
    break;                                                                                        \
  }
    FFS_SIZES(X)
#undef X
    default: bad_size("x C2R pass", g.nrx);
  }
}

template <typename T>
void launch_x_c2r_rows(const Geo& g, const void* a2, double* real, int row0, int nrow, const void* tw, hipStream_t s) {
  if (row0 < 0 || row0 % 2 || nrow < 1 || row0 + nrow > g.ny * g.nz)
    throw std::runtime_error("x C2R pass: row range outside the field or not starting at an even row");
  const int npairs = (nrow + 1) / 2;
  switch (g.nrx) {
#define X(n)                                                                                            \
  case n: {                                                                                             \
    constexpr int NS = 3 * xrows<n>();                                                                  \
    dim3 grid((npairs + NS - 1) / NS);                                                                  \
    k_x_c2r<T, n, double><<<grid, NS * tper<n>(), x_lds_bytes<T, n>(), s>>>(g, (const C<T>*)a2, real,    \
                                                                           (const C<T>*)tw, 1, row0, nrow); \
    break;                                                                                              \
  }
    FFS_SIZES(X)
#undef X
    default: bad_size("x C2R pass", g.nrx);
  }
}

template <typename T>
void launch_x_r2c(const Geo& g, const T* real, void* b2, int nf, const void* tw, hipStream_t s) {
  const int npairs = (g.ny * g.nz + 1) / 2;
  switch (g.nrx) {
#define X(n)                                                                                      \
  case n: {                                                                                       \
    constexpr int NS = 3 * xrows<n>();                                                            \
    dim3 grid((nf * npairs + NS - 1) / NS);                                                       \
    k_x_r2c<T, n><<<grid, NS * tper<n>(), x_lds_bytes<T, n>(), s>>>(g, real, (C<T>*)b2, (const C<T>*)tw, nf); \
    break;                                                                                        \
  }
    FFS_SIZES(X)
#undef X
    default: bad_size("x R2C pass", g.nrx);
  }
}

// ===========================================================================
// Spectral update of one mode, fused: fourierspacep2! (models.jl:282-364:
// -nu k^2 u for AB3, + (ik).tau, + rho g, Leray; rhs.c[1] = 0 was applied by
// the forward z pass) then VectorTimeStep (time_step_functions.jl:72-80: AB3
// Adams_Bashforth.jl:231-256 with fix_fm2! ETD.jl:130-140, or ETD1/ETD2/ETD3
// ETD1stand2ndO.jl:264-327 / ETD.jl:229-262, forcing, projection) and the
// density integrator (models.jl:471) with rho rhs = ik.F(-rho u) - drhodz w
// (div!, vectorfunctions.jl:93-103) from the pre-step w.  rhs / rho rhs are
// written back: they are the next step's fm1 (history rotated by pointer).

// ---- hyperviscosity (Hyperviscosity/hyperviscosity_types.jl, ETD.jl:56-95) ----
// k^2M as repeated products (M >= 1)
__device__ __forceinline__ double ipow_m(double x, int m) {
  double r = x;
  for (int i = 1; i < m; ++i) r *= x;
  return r;
}
// SpectralBarrier's f(k) = log(esp(k, initk, endk, 1/4)): 0 up to initk, -Inf from endk on
__device__ __forceinline__ double barrier_f(double k, double ki, double kp) {
  const double ak = fabs(k);
  const double e = ak <= ki ? 1.0 : (ak < kp ? exp(-((ak - ki) * (ak - ki)) / (pow(kp, 0.25) - pow(ak, 0.25))) : 0.0);
  return log(e);
}
// muladd(KX, KX, muladd(KY, KY, KZ^2)) of init_c! and the hyperviscosity loops.  Evaluated unfused: the
// ETD3 closed forms cancel near their series switch, so c must round exactly as the host restatement's
// (a reference that fuses muladd differs from both by an ulp of c)
__device__ __forceinline__ double modk2(double kx, double ky, double kz) {
#pragma clang fp contract(off)
  return kx * kx + (ky * ky + kz * kz);
}
// init_c! (ETD.jl:56-95): the linear operator c of one mode for diffusivity -mnu
__device__ __forceinline__ double etd_c(const HvArgs& hv, double K2, const double kv3[3], double mnu) {
#pragma clang fp contract(off)
  if (hv.kind == 0) return K2 * mnu;
  const double mk = modk2(kv3[0], kv3[1], kv3[2]);
  if (hv.kind == 1) return mk * mnu + ipow_m(mk, hv.m) * (-hv.nuh);
  return mk * mnu + barrier_f(sqrt(mk), hv.sb_i, hv.sb_p);
}
// per-mode weight of the hyperviscous dissipation sums: k^2M (hyperviscosity_stats, stats.jl:532-561; nu_h
// multiplies the sum) or f(|k|) with -Inf read as 0 (spectral_barrier_stats, hyperviscosity_types.jl:50-79)
__device__ __forceinline__ double hv_weight(const HvArgs& hv, const double kv3[3]) {
  const double mk = modk2(kv3[0], kv3[1], kv3[2]);
  if (hv.kind == 1) return ipow_m(mk, hv.m);
  const double f = barrier_f(sqrt(mk), hv.sb_i, hv.sb_p);
  return f == -INFINITY ? 0.0 : f;
}

// ETD coefficients of one mode (FP64): the closed forms with expm1 and the
// series below |c dt| = 1e-4 exactly as set_ABCt! / ETD1stO! / ETD2ndO!; c = -Inf
// (beyond a spectral barrier) gives zero multipliers (their test2 branches).
struct EtdCoef {
  double E, A, B, Cc;  // exp(c dt) and the multipliers of rhs, fm1, fm2
};
__device__ __forceinline__ EtdCoef etd_coef(double c, int step, const DevState* __restrict__ st) {
  // no FMA contraction: the closed forms cancel to O((c dt)^3), so every rounding step has to be
  // the one the reference's plain * and + perform
#pragma clang fp contract(off)
  const double dt = st->dt;
  EtdCoef r;
  r.E = exp(c * dt);
  r.A = r.B = r.Cc = 0.0;
  if (c == -INFINITY) return r;
  if (step == 1) {
    const double cdt = c * dt, cdt2 = cdt * dt, c2dt3 = cdt2 * cdt, c3dt4 = c2dt3 * cdt;
    r.A = (-cdt <= 1e-4) ? dt + 0.5 * cdt2 + c2dt3 / 6 + c3dt4 / 24 : expm1(cdt) / c;
  } else if (step == 2) {
    const double dt2 = st->dt2;
    const double cp2 = c * c, cdt = c * dt, cdtp2 = cdt * dt, c2dtp3 = cdtp2 * cdt, c3dtp4 = c2dtp3 * cdt;
    const double em = expm1(cdt), c2dt2 = c * c * dt2, dtp2 = dt * dt, dtp3 = dtp2 * dt, dtp4 = dtp3 * dt;
    if (-cdt <= 1e-4) {
      r.A = dt + cdtp2 / 2 + (c2dtp3) / 6 + (c3dtp4) / 24 + dtp2 / (2 * dt2) + (c * dtp3) / (6 * dt2) +
            (cp2 * dtp4) / (24 * dt2);
      r.B = (-(dtp2 / (2 * dt2)) - (c * dtp3) / (6 * dt2) - (cp2 * dtp4) / (24 * dt2));
    } else {
      r.A = (em - c * (dt - dt2 * em)) / c2dt2;
      r.B = (cdt - em) / c2dt2;
    }
  } else {
    const double* e = st->etd;  // a0 a1 a2 a3 b0 b1 b2 c0 c1 c2 e1 .. e9
    const double l1 = c, l2 = l1 * l1, l3 = l2 * l1, ldt = l1 * dt;
    if (-l1 * dt <= 1e-4) {
      r.A = fma(e[1], l1, fma(e[2], l2, fma(e[3], l3, e[0])));
      r.B = fma(e[5], l1, fma(e[6], l2, e[4]));
      r.Cc = fma(e[8], l1, fma(e[9], l2, e[7]));
    } else {
      const double em = expm1(ldt), dt2 = st->dt2;
      const double e1 = e[10], e2 = e[11], e3 = e[12], e4 = e[13], e5 = e[14], e6 = e[15], e7 = e[16], e8 = e[17],
                   e9 = e[18];
      r.A = (2 * em - l2 * (e1 - e2 * em) - l1 * (e3 - e4 * em)) / (l3 * e2);
      r.B = (e5 * l2 - 2 * em + l1 * (e3 - e6 * em)) / (l3 * e7);
      r.Cc = (2 * em - e8 * l2 - l1 * (e3 - dt2 * em)) / (l3 * e9);
    }
  }
  return r;
}

// fourierspacep2_velocity! (models.jl:282-341) on one mode: rhs += -nu k^2 u (AB3: explicit
// viscosity), + (i k) . tau (LES), + rho g (Boussinesq), then the Leray projection.  RH receives rho.
// FULL = false: the plain step (no LES, density, passive scalar, plane forcing or hyperviscosity), whose
// optional paths are compiled out so the update kernel holds fewer registers
template <typename T, bool FULL = true>
__device__ __forceinline__ void form_rhs(const UpdateArgs& a, size_t idx, size_t n, const double kv3[3], double K2,
                                         bool origin, const double U[3][2], double R[3][2], double RH[2]) {
  if (a.integrator != 1) {  // explicit viscosity (is_implicit false: AB3, Euler): (K2 mnu + K2^M mnuh) v
    double kvis;
    {
#pragma clang fp contract(off)
      kvis = (FULL && a.hv.kind == 1) ? K2 * (-a.nu) + ipow_m(K2, a.hv.m) * (-a.hv.nuh) : K2 * (-a.nu);
    }
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int p = 0; p < 2; ++p) R[c][p] = fma(kvis, U[c][p], R[c][p]);
  }
  if (FULL && a.tau) {  // rhs += (i k) . tau, tau symmetric traceless (xx xy xz yy yz)
    const C<T>* t = (const C<T>*)a.tau;
    const C<T> txx = t[idx], txy = t[n + idx], txz = t[2 * n + idx], tyy = t[3 * n + idx], tyz = t[4 * n + idx];
    const double T6[6][2] = {{txx.x, txx.y}, {txy.x, txy.y}, {txz.x, txz.y},
                             {tyy.x, tyy.y}, {tyz.x, tyz.y}, {-(double)txx.x - tyy.x, -(double)txx.y - tyy.y}};
    // rows of the symmetric tensor: x (xx xy xz), y (xy yy yz), z (xz yz zz)
    const int row[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const double sr = fma(kv3[2], T6[row[c][2]][0], fma(kv3[1], T6[row[c][1]][0], kv3[0] * T6[row[c][0]][0]));
      const double si = fma(kv3[2], T6[row[c][2]][1], fma(kv3[1], T6[row[c][1]][1], kv3[0] * T6[row[c][0]][1]));
      R[c][0] += -si;
      R[c][1] += sr;
    }
  }
  if (FULL && a.rho) {  // buoyancy rhs += rho g, g = g e_gdir (models.jl:334-336)
    const C<T> r0 = ((const C<T>*)a.rho)[idx];
    RH[0] = r0.x;
    RH[1] = r0.y;
    R[a.gdir][0] = fma(RH[0], a.gval, R[a.gdir][0]);
    R[a.gdir][1] = fma(RH[1], a.gval, R[a.gdir][1]);
  }
#pragma unroll
  for (int p = 0; p < 2; ++p) {
    const double p1 = origin ? 0.0 : -fma(kv3[2], R[2][p], fma(kv3[1], R[1][p], kv3[0] * R[0][p])) / K2;
#pragma unroll
    for (int c = 0; c < 3; ++c) R[c][p] = fma(p1, kv3[c], R[c][p]);
  }
}

// The update of one compact mode idx = (cx, cy, cz) from its nonlinear term R (the dealiased, scaled
// F(u x w) as the forward z pass leaves it in rhs): the final rhs is stored to rhs, the new u to u.
template <typename T, bool FULL>
__device__ __forceinline__ void update_point(const Geo& g, C<T>* __restrict__ u, C<T>* __restrict__ rhs,
                                             const C<T>* __restrict__ fm1, const C<T>* __restrict__ fm2,
                                             const DevState* __restrict__ st, const UpdateArgs& a, size_t idx,
                                             size_t n, int cx, int cy, int cz, double (&R)[3][2]) {
  const double kv3[3] = {g.kx[cx], g.ky[cy], g.kz[cz]};
  // the sums of products are written as explicit fma chains, so every kernel that inlines the update
  // rounds them alike whatever its compiler contracts around them
  const double K2 = fma(kv3[2], kv3[2], fma(kv3[1], kv3[1], kv3[0] * kv3[0]));
  const bool origin = (idx == 0);
  double U[3][2];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const C<T> v = u[c * n + idx];
    U[c][0] = v.x; U[c][1] = v.y;
  }
  // plane forcing (rf / qf): the force is built from the pre-step u1, u2 (forcing.jl:85-90)
  double F0[2][2] = {{0, 0}, {0, 0}};
  if (FULL && a.forcing >= 2 && a.step > 0) {
    double f = 0.0;
    if (cz == 0) {
      f = a.pf_factor[cx + (size_t)g.sx * cy];  // per-element factor of this step (0 where not forced)
    } else if (cx == 0 && cy == 0) {
      const int j = z_fft(g, cz), kz = j < (g.nz + 1) / 2 ? j : j - g.nz;
      const int akz = kz < 0 ? -kz : kz;
      if (akz >= a.pf_kz0 && akz <= 5) f = st->force_vert;
    }
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
      for (int p = 0; p < 2; ++p) F0[c][p] = U[c][p] * f;
  }
  double RH[2] = {0, 0};
  form_rhs<T, FULL>(a, idx, n, kv3, K2, origin, U, R, RH);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    if constexpr (FFS_NT_UPDATE) store_nt<T>(&rhs[c * n + idx], mk<T>((T)R[c][0], (T)R[c][1]));
    else rhs[c * n + idx] = mk<T>((T)R[c][0], (T)R[c][1]);
  }
  // density rhs from the pre-step vertical velocity
  double RR[2] = {0, 0};
  if (FULL && a.rho) {
    const C<T>* f = (const C<T>*)a.flux;
    const C<T> fx = f[idx], fy = f[n + idx], fz = f[2 * n + idx];
    const double kf_r = fma(kv3[0], (double)fx.x, fma(kv3[1], (double)fy.x, kv3[2] * fz.x));
    const double kf_i = fma(kv3[0], (double)fx.y, fma(kv3[1], (double)fy.y, kv3[2] * fz.y));
    RR[0] = -kf_i + (-a.drhodz) * U[a.gdir][0];  // div!(rhs, flux, u_gdir, -drhodz) (models.jl:358-360)
    RR[1] = kf_r + (-a.drhodz) * U[a.gdir][1];
    ((C<T>*)a.rrhs)[idx] = mk<T>((T)RR[0], (T)RR[1]);
  }
  // passive scalar rhs: div!(rhs, flux, u_dir, -scalarGradient) (models.jl:353-355) from the pre-step u
  double PR[2] = {0, 0}, PH[2] = {0, 0};
  if (FULL && a.phi) {
    const C<T> p0 = ((const C<T>*)a.phi)[idx];
    PH[0] = p0.x;
    PH[1] = p0.y;
    const C<T>* f = (const C<T>*)a.pflux;
    const C<T> fx = f[idx], fy = f[n + idx], fz = f[2 * n + idx];
    const double kf_r = fma(kv3[0], (double)fx.x, fma(kv3[1], (double)fy.x, kv3[2] * fz.x));
    const double kf_i = fma(kv3[0], (double)fx.y, fma(kv3[1], (double)fy.y, kv3[2] * fz.y));
    const double wr = a.pdir == 0 ? U[0][0] : (a.pdir == 1 ? U[1][0] : U[2][0]);
    const double wi = a.pdir == 0 ? U[0][1] : (a.pdir == 1 ? U[1][1] : U[2][1]);
    PR[0] = -kf_i + (-a.pgrad) * wr;
    PR[1] = kf_r + (-a.pgrad) * wi;
    ((C<T>*)a.prhs)[idx] = mk<T>((T)PR[0], (T)PR[1]);
  }
  if (a.step == 0) return;  // initialisation: final rhs only (initialize!, FluidFlowSimulation.jl:82)
  if (a.integrator == 2) {  // Euller (Euller.jl:14-25): u = muladd(dt, rhs, u), no history
    const double dt = st->dt;
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int p = 0; p < 2; ++p) U[c][p] = fma(dt, R[c][p], U[c][p]);
    if constexpr (FULL) {
#pragma unroll
      for (int p = 0; p < 2; ++p) RH[p] = fma(dt, RR[p], RH[p]);
#pragma unroll
      for (int p = 0; p < 2; ++p) PH[p] = fma(dt, PR[p], PH[p]);
    }
  } else {
    double M1[3][2], M2[3][2];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const C<T> h1 = fm1[c * n + idx], h2 = fm2[c * n + idx];
      M1[c][0] = h1.x; M1[c][1] = h1.y;
      M2[c][0] = h2.x; M2[c][1] = h2.y;
    }
    double HR1[2] = {0, 0}, HR2[2] = {0, 0};
    if (FULL && a.rho) {
      const C<T> h1 = ((const C<T>*)a.rfm1)[idx], h2 = ((const C<T>*)a.rfm2)[idx];
      HR1[0] = h1.x; HR1[1] = h1.y;
      HR2[0] = h2.x; HR2[1] = h2.y;
    }
    double HP1[2] = {0, 0}, HP2[2] = {0, 0};
    if (FULL && a.phi) {
      const C<T> h1 = ((const C<T>*)a.pfm1)[idx], h2 = ((const C<T>*)a.pfm2)[idx];
      HP1[0] = h1.x; HP1[1] = h1.y;
      HP2[0] = h2.x; HP2[1] = h2.y;
    }
    if (a.integrator == 0) {
      const double At = st->At, Bt = st->Bt, Ct = st->Ct;
      if (a.step == 2) {  // fix_fm2!: fm2 = -(dt3+dt2)(rhs - fm1)/dt2 + rhs
        const double dt2 = st->dt2, dt3 = st->dt3;
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
          for (int p = 0; p < 2; ++p) M2[c][p] = -(dt3 + dt2) * (R[c][p] - M1[c][p]) / dt2 + R[c][p];
        if constexpr (FULL) {
#pragma unroll
          for (int p = 0; p < 2; ++p) HR2[p] = -(dt3 + dt2) * (RR[p] - HR1[p]) / dt2 + RR[p];
#pragma unroll
          for (int p = 0; p < 2; ++p) HP2[p] = -(dt3 + dt2) * (PR[p] - HP1[p]) / dt2 + PR[p];
        }
      }
#pragma unroll
      for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int p = 0; p < 2; ++p) U[c][p] = fma(At, R[c][p], fma(Bt, M1[c][p], fma(Ct, M2[c][p], U[c][p])));
      if constexpr (FULL) {
#pragma unroll
        for (int p = 0; p < 2; ++p) RH[p] = fma(At, RR[p], fma(Bt, HR1[p], fma(Ct, HR2[p], RH[p])));
#pragma unroll
        for (int p = 0; p < 2; ++p) PH[p] = fma(At, PR[p], fma(Bt, HP1[p], fma(Ct, HP2[p], PH[p])));
      }
    } else {
      const EtdCoef e = etd_coef(FULL ? etd_c(a.hv, K2, kv3, -a.nu) : K2 * (-a.nu), a.step, st);
#pragma unroll
      for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int p = 0; p < 2; ++p) {
          if (a.step == 1) U[c][p] = fma(e.E, U[c][p], e.A * R[c][p]);
          else if (a.step == 2) U[c][p] = fma(e.E, U[c][p], fma(e.A, R[c][p], e.B * M1[c][p]));
          else U[c][p] = fma(e.A, R[c][p], fma(e.B, M1[c][p], fma(e.Cc, M2[c][p], e.E * U[c][p])));
        }
      if (FULL && a.rho) {  // the density's own hyperviscosity (density_stratification_types.jl:168-174)
        const EtdCoef d = etd_coef(etd_c(a.hvd, K2, kv3, -a.alpha), a.step, st);
#pragma unroll
        for (int p = 0; p < 2; ++p) {
          if (a.step == 1) RH[p] = fma(d.E, RH[p], d.A * RR[p]);
          else if (a.step == 2) RH[p] = fma(d.E, RH[p], fma(d.A, RR[p], d.B * HR1[p]));
          else RH[p] = fma(d.A, RR[p], fma(d.B, HR1[p], fma(d.Cc, HR2[p], d.E * RH[p])));
        }
      }
      if (FULL && a.phi) {
        const EtdCoef d = etd_coef(etd_c(a.hv, K2, kv3, -a.palpha), a.step, st);
#pragma unroll
        for (int p = 0; p < 2; ++p) {
          if (a.step == 1) PH[p] = fma(d.E, PH[p], d.A * PR[p]);
          else if (a.step == 2) PH[p] = fma(d.E, PH[p], fma(d.A, PR[p], d.B * HP1[p]));
          else PH[p] = fma(d.A, PR[p], fma(d.B, HP1[p], fma(d.Cc, HP2[p], d.E * PH[p])));
        }
      }
    }
  }
  // constant-power isotropic shell forcing, added after the integrator like add_forcing!
  // (time_step_functions.jl:77): u += dt * (P / sum_band w|u|^2) u on 0 < |k|^2 <= kf2
  if (a.forcing == 1 && K2 <= a.kf2 && !origin) {
    const double fac = st->dt * st->force_factor;
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int p = 0; p < 2; ++p) U[c][p] = fma(fac, U[c][p], U[c][p]);
  }
  // plane forcing: u_h += force (add_Rf_forcing!, forcing_types.jl:54-64)
  if constexpr (FULL) {
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
      for (int p = 0; p < 2; ++p) U[c][p] += F0[c][p];
  }
#pragma unroll
  for (int p = 0; p < 2; ++p) {
    const double p1 = origin ? 0.0 : -fma(kv3[2], U[2][p], fma(kv3[1], U[1][p], kv3[0] * U[0][p])) / K2;
#pragma unroll
    for (int c = 0; c < 3; ++c) U[c][p] = fma(p1, kv3[c], U[c][p]);
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    if constexpr (FFS_NT_UPDATE && FFS_NT_U) store_nt<T>(&u[c * n + idx], mk<T>((T)U[c][0], (T)U[c][1]));
    else u[c * n + idx] = mk<T>((T)U[c][0], (T)U[c][1]);
  }
  if (FULL && a.rho) ((C<T>*)a.rho)[idx] = mk<T>((T)RH[0], (T)RH[1]);
  if (FULL && a.phi) ((C<T>*)a.phi)[idx] = mk<T>((T)PH[0], (T)PH[1]);
}

template <typename T> struct UCall {
  Geo g;
  C<T>* u;
  C<T>* rhs;
  const C<T>* fm1;
  const C<T>* fm2;
  const DevState* st;
  UpdateArgs a;
};

template <typename T, bool FULL>
__device__ __forceinline__ void update_body(Geo g, C<T>* __restrict__ u, C<T>* __restrict__ rhs,
                                            const C<T>* __restrict__ fm1, const C<T>* __restrict__ fm2,
                                            const DevState* __restrict__ st, UpdateArgs a, int bX) {
  g = glb(g);
  a = glb(a);
  // the step's scalars (no thread reads the time this kernel advances)
  const DevState sv = load_const(st);
  const size_t n = (size_t)g.sx * g.my * g.mz;
  const size_t idx = (size_t)bX * blockDim.x + threadIdx.x;
  // s.time[] += dt, once per step (no kernel reads the time during the update)
  if (a.time && idx == 0) *a.time += sv.dt;
  if (idx >= n) return;
  const int cx = idx % g.sx;
  if (cx >= g.mx) return;  // row padding
  const int cy = (idx / g.sx) % g.my;
  const int cz = idx / ((size_t)g.sx * g.my);
  double R[3][2];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const C<T> r = rhs[c * n + idx];
    R[c][0] = r.x; R[c][1] = r.y;
  }
  update_point<T, FULL>(g, u, rhs, fm1, fm2, &sv, a, idx, n, cx, cy, cz, R);
}
template <typename T, bool FULL>
__global__ void __launch_bounds__(256) k_update(Geo g, C<T>* __restrict__ u, C<T>* __restrict__ rhs,
                                               const C<T>* __restrict__ fm1, const C<T>* __restrict__ fm2,
                                               const DevState* __restrict__ st, UpdateArgs a) {
  update_body<T, FULL>(g, u, rhs, fm1, fm2, st, a, blockIdx.x);
}
template <typename T, bool FULL>
__global__ void __launch_bounds__(256) k_update_b(FFS_CALLS(UCall<T>) calls, int nrun) {
  const UCall<T> c = load_call(calls, blockIdx.z);
  update_body<T, FULL>(c.g, glb(c.u), glb(c.rhs), glb(c.fm1), glb(c.fm2), glb(c.st), c.a, blockIdx.x);
}

// The forward z pass of the three nonlinear components fused with the update of the plain step (no
// LES, density, passive scalar, plane forcing or hyperviscosity): a block transforms the columns of
// its TX x-columns at one cy for all three components, keeps the results in registers and updates
// those modes, so the nonlinear term never goes through rhs in HBM.  The values are those of the
// forward z pass followed by the update kernel, bit for bit.
template <typename T> struct FzuCall {
  Geo g;
  const C<T>* in;
  C<T>* u;
  C<T>* rhs;
  const C<T>* fm1;
  const C<T>* fm2;
  const DevState* st;
  UpdateArgs a;
  const C<T>* tw;
  T scale;
};
template <typename T, int N>
__device__ __forceinline__ void fwd_z_update_body(Geo g, const C<T>* __restrict__ in, C<T>* __restrict__ u,
                                                  C<T>* __restrict__ rhs, const C<T>* __restrict__ fm1,
                                                  const C<T>* __restrict__ fm2, const DevState* __restrict__ st,
                                                  UpdateArgs a, const C<T>* __restrict__ tw_g, T scale, int bX, int bY) {
  g = glb(g);
  a = glb(a);
  constexpr int TX = col_tx<T>(N);
  constexpr int TP = col_tper<T, N, false>();
  constexpr int MAXR = col_maxr<T, N, false>();
  using Last = Stage<T, N, -1, plan_ns(N, MAXR) - 1, TP, MAXR>;
  constexpr int NSL = Last::NB * Last::R;  // outputs per thread
  extern __shared__ unsigned char smem_raw[];
  C<T>* lds = reinterpret_cast<C<T>*>(smem_raw);
  const C<T>* tw = stage_twiddles<T, N>(lds + N * TX, tw_g);
  if (a.time && bX == 0 && bY == 0 && threadIdx.x == 0) *a.time += st->dt;
  __syncthreads();
  const int c = threadIdx.x % TX;
  const int q = threadIdx.x / TX;
  const int cx = bX * TX + c;
  const int cy = bY;
  const bool valid = cx < g.mx;
  const size_t sx = g.sx;
  const size_t n = sx * g.my * g.mz;
  const size_t S1 = sx * g.my * g.nz;
  auto ldsf = [&](int p) -> C<T>& { return lds[p * TX + c]; };
  C<T> V[3][NSL];
  int P[NSL];
#pragma unroll
  for (int m = 0; m < NSL; ++m) P[m] = -1;
#pragma unroll
  for (int f = 0; f < 3; ++f) {
    if (f > 0) __syncthreads();
    auto ld = [&](int p, int, int) -> C<T> {
      if (!valid) return mk<T>(0, 0);
      return in[f * S1 + cx + sx * (cy + (size_t)g.my * p)];
    };
    auto stv = [&](int p, int t, int r, C<T> v) {
      V[f][t * Last::R + r] = v;
      P[t * Last::R + r] = p;
    };
    Fft<T, N, -1, TP, false, false, SYNC_BLOCK, MAXR>::run(q, tw, ld, stv, ldsf);
  }
  if (!valid) return;
#pragma unroll
  for (int m = 0; m < NSL; ++m) {
    const int cz = P[m] < 0 ? -1 : z_compact(g, P[m]);
    if (cz < 0) continue;
    const size_t idx = cx + sx * (cy + (size_t)g.my * cz);
    const bool zero = cx == 0 && cy == 0 && cz == 0;  // the mean of the nonlinear term
    double R[3][2];
#pragma unroll
    for (int f = 0; f < 3; ++f) {
      // rounded as the forward z pass stores it: the product is not contracted into the update's sums
#pragma clang fp contract(off)
      const T rx = V[f][m].x * scale, ry = V[f][m].y * scale;
      R[f][0] = zero ? 0.0 : (double)rx;
      R[f][1] = zero ? 0.0 : (double)ry;
    }
    update_point<T, false>(g, u, rhs, fm1, fm2, st, a, idx, n, cx, cy, cz, R);
  }
}

template <typename T, int N>
__global__ void __launch_bounds__(kColThreads, FFS_COL_WAVES)
    k_fwd_z_update(Geo g, const C<T>* __restrict__ in, C<T>* __restrict__ u, C<T>* __restrict__ rhs,
                   const C<T>* __restrict__ fm1, const C<T>* __restrict__ fm2, const DevState* __restrict__ st,
                   UpdateArgs a, const C<T>* __restrict__ tw_g, T scale) {
  fwd_z_update_body<T, N>(g, in, u, rhs, fm1, fm2, st, a, tw_g, scale, blockIdx.x, blockIdx.y);
}
template <typename T, int N>
__global__ void __launch_bounds__(kColThreads, FFS_COL_WAVES) k_fwd_z_update_b(FFS_CALLS(FzuCall<T>) calls,
                                                                              int nrun) {
  const FzuCall<T> c = load_call(calls, blockIdx.z);
  fwd_z_update_body<T, N>(c.g, glb(c.in), glb(c.u), glb(c.rhs), glb(c.fm1), glb(c.fm2), glb(c.st), c.a, glb(c.tw),
                          c.scale, blockIdx.x, blockIdx.y);
}

// calc_les! Smagorinsky branch (les_real_calc.jl:77-79) on padded real S fields: tau = 2 (c Delta)^2
// sqrt(2 S:S) S with Szz = -Sxx - Syy, as the x pass forms it
template <typename T>
__global__ void __launch_bounds__(256) k_les_tau_real(T* __restrict__ s5, size_t n, double c2a) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const T sxx = s5[i], sxy = s5[n + i], sxz = s5[2 * n + i], syy = s5[3 * n + i], syz = s5[4 * n + i];
  const T szz = -sxx - syy;
  const T ss = sxx * sxx + syy * syy + szz * szz + (T)2 * (sxy * sxy + sxz * sxz + syz * syz);
  const T f = (T)c2a * sqrt((T)2 * ss);
  s5[i] = f * sxx;
  s5[n + i] = f * sxy;
  s5[2 * n + i] = f * sxz;
  s5[3 * n + i] = f * syy;
  s5[4 * n + i] = f * syz;
}

template <typename T>
void launch_les_tau_real(T* s5, size_t n_padded, double c2a, hipStream_t s) {
  k_les_tau_real<T><<<(unsigned)((n_padded + 255) / 256), 256, 0, s>>>(s5, n_padded, c2a);
}

template <typename T>
__global__ void __launch_bounds__(256) k_les_scalar_flux_real(const T* __restrict__ s5, const T* __restrict__ w3,
                                                              const T* __restrict__ g3, T* __restrict__ out3,
                                                              size_t n, LesArgs l, double alpha, int les, double cv) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const T sv[5] = {s5[i], s5[n + i], s5[2 * n + i], s5[3 * n + i], s5[4 * n + i]};
  const T wv[3] = {w3[i], w3[n + i], w3[2 * n + i]};
  const T gv[3] = {g3[i], g3[n + i], g3[2 * n + i]};
  T nut;
  if (l.model == LES_SMAGORINSKY) {
    const T szz = -sv[0] - sv[3];
    const T ss = sv[0] * sv[0] + sv[3] * sv[3] + szz * szz + (T)2 * (sv[1] * sv[1] + sv[2] * sv[2] + sv[4] * sv[4]);
    nut = (T)alpha * sqrt((T)2 * ss);
  } else {
    T tv[5];
    les_stress<T>(l, sv, wv, tv, &nut);
  }
  T f[3];
  scalar_flux<T>((T)0, (T)0, (T)0, (T)0, les, nut, gv, wv, (T)cv, f);
#pragma unroll
  for (int c = 0; c < 3; ++c) out3[c * n + i] = f[c];
}

template <typename T>
void launch_les_scalar_flux_real(const T* s5, const T* w3, const T* g3, T* out3, size_t n_padded, const LesArgs& l,
                                 double alpha, int les, double cv, hipStream_t s) {
  k_les_scalar_flux_real<T><<<(unsigned)((n_padded + 255) / 256), 256, 0, s>>>(s5, w3, g3, out3, n_padded, l, alpha,
                                                                               les, cv);
}

// calc_les! of a model that reads w, on padded real S and w fields (S overwritten by t)
template <typename T>
__global__ void __launch_bounds__(256) k_les_model_real(T* __restrict__ s5, const T* __restrict__ w3, size_t n,
                                                        LesArgs l) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const T sv[5] = {s5[i], s5[n + i], s5[2 * n + i], s5[3 * n + i], s5[4 * n + i]};
  const T wv[3] = {w3[i], w3[n + i], w3[2 * n + i]};
  T tv[5];
  les_stress<T>(l, sv, wv, tv);
#pragma unroll
  for (int e = 0; e < 5; ++e) s5[e * n + i] = tv[e];
}

template <typename T>
void launch_les_model_real(T* s5, const T* w3, size_t n_padded, const LesArgs& l, hipStream_t s) {
  k_les_model_real<T><<<(unsigned)((n_padded + 255) / 256), 256, 0, s>>>(s5, w3, n_padded, l);
}

// ---- dynamic Smagorinsky ---------------------------------------------------------------------------
__device__ __forceinline__ double gauss_filter(double td2, double k2) { return exp((-td2 / 24) * k2); }

template <typename T>
__global__ void __launch_bounds__(256) k_dyn_prep(Geo g, const C<T>* __restrict__ u, C<T>* __restrict__ out,
                                                  double td2, int with_w) {
  const size_t n = (size_t)g.sx * g.my * g.mz;
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n) return;
  const int cx = idx % g.sx;
  if (cx >= g.mx) return;
  const int cy = (idx / g.sx) % g.my;
  const int cz = idx / ((size_t)g.sx * g.my);
  const double kxd = g.kx[cx], kyd = g.ky[cy], kzd = g.kz[cz];
  const T kx = (T)kxd, ky = (T)kyd, kz = (T)kzd;
  const C<T> v[3] = {u[idx], u[n + idx], u[2 * n + idx]};
  const T gf = (T)gauss_filter(td2, kxd * kxd + kyd * kyd + kzd * kzd);
  // S_ab = (i k_a u_b + i k_b u_a) / 2 for ab = xx, xy, xz, yy, yz, as the curl pass forms it
  const int A_[5] = {0, 0, 0, 1, 1}, B_[5] = {0, 1, 2, 1, 2};
#pragma unroll
  for (int e = 0; e < 5; ++e) {
    const T ka = A_[e] == 0 ? kx : ky;
    const T kb = B_[e] == 0 ? kx : (B_[e] == 1 ? ky : kz);
    const C<T> ua = v[A_[e]], ub = v[B_[e]];
    const C<T> sum = mk<T>((T)0.5 * (ka * ub.x + kb * ua.x), (T)0.5 * (ka * ub.y + kb * ua.y));
    const C<T> S = mk<T>(-sum.y, sum.x);
    out[e * n + idx] = S;
    out[(5 + e) * n + idx] = mk<T>(gf * S.x, gf * S.y);
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) out[(10 + c) * n + idx] = mk<T>(gf * v[c].x, gf * v[c].y);
  if (with_w) {  // w = i k x u as the curl pass forms it, then G w (models.jl:89-108)
    const C<T> a[3] = {mk<T>(ky * v[2].x - kz * v[1].x, ky * v[2].y - kz * v[1].y),
                       mk<T>(kz * v[0].x - kx * v[2].x, kz * v[0].y - kx * v[2].y),
                       mk<T>(kx * v[1].x - ky * v[0].x, kx * v[1].y - ky * v[0].y)};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const C<T> w = mk<T>(-a[c].y, a[c].x);
      out[(13 + c) * n + idx] = w;
      out[(16 + c) * n + idx] = mk<T>(gf * w.x, gf * w.y);
    }
  }
}

template <typename T>
void launch_dyn_prep(const Geo& g, const void* u, void* out, double td2, hipStream_t s, int with_w) {
  const size_t n = (size_t)g.sx * g.my * g.mz;
  k_dyn_prep<T><<<(unsigned)((n + 255) / 256), 256, 0, s>>>(g, (const C<T>*)u, (C<T>*)out, td2, with_w);
}

// Lie(S, AntiSymTen(w/2)) = A S - S A, A = AntiSymTen(-w/2), of S stored as xx xy xz yy yz (zz = -xx - yy);
// the result is symmetric and traceless, returned in the same five components
template <typename T> __device__ __forceinline__ void lie_s_w(const T s[5], T wx, T wy, T wz, T P[5]) {
  const T S[3][3] = {{s[0], s[1], s[2]}, {s[1], s[3], s[4]}, {s[2], s[4], -s[0] - s[3]}};
  const T hx = (T)0.5 * wx, hy = (T)0.5 * wy, hz = (T)0.5 * wz;
  const T A[3][3] = {{0, hz, -hy}, {-hz, 0, hx}, {hy, -hx, 0}};
  auto e = [&](int i, int j) {
    T v = 0;
#pragma unroll
    for (int k = 0; k < 3; ++k) v += A[i][k] * S[k][j] - S[i][k] * A[k][j];
    return v;
  };
  P[0] = e(0, 0);
  P[1] = e(0, 1);
  P[2] = e(0, 2);
  P[3] = e(1, 1);
  P[4] = e(1, 2);
}

// |S| = sqrt(2 S:S) of a traceless symmetric tensor stored as xx xy xz yy yz
template <typename T> __device__ __forceinline__ T norm_s5(T xx, T xy, T xz, T yy, T yz) {
  const T zz = -xx - yy;
  return sqrt((T)2 * (xx * xx + yy * yy + zz * zz + (T)2 * (xy * xy + xz * xz + yz * yz)));
}

template <typename T>
__global__ void __launch_bounds__(256) k_dyn_products(const T* __restrict__ s5, const T* __restrict__ u3,
                                                      T* __restrict__ sa5, T* __restrict__ l6, size_t n, double d2,
                                                      const T* __restrict__ cstar, const T* __restrict__ w3,
                                                      const T* __restrict__ cpstar, T* __restrict__ pa5) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const T S[5] = {s5[i], s5[n + i], s5[2 * n + i], s5[3 * n + i], s5[4 * n + i]};
  const T f = (cstar ? (T)2 * cstar[i] * (T)d2 : (T)d2) * norm_s5<T>(S[0], S[1], S[2], S[3], S[4]);
#pragma unroll
  for (int e = 0; e < 5; ++e) sa5[e * n + i] = f * S[e];
  const T ux = u3[i], uy = u3[n + i], uz = u3[2 * n + i];
  l6[i] = ux * ux;
  l6[n + i] = ux * uy;
  l6[2 * n + i] = ux * uz;
  l6[3 * n + i] = uy * uy;
  l6[4 * n + i] = uy * uz;
  l6[5 * n + i] = uz * uz;
  if (w3) {  // P = [cp] Delta^2 Lie(S, AntiSymTen(w/2)) (dynamic_smagorinsky.jl:71-75)
    T P[5];
    lie_s_w<T>(S, w3[i], w3[n + i], w3[2 * n + i], P);
    const T fp = cpstar ? cpstar[i] * (T)d2 : (T)d2;
#pragma unroll
    for (int e = 0; e < 5; ++e) pa5[e * n + i] = fp * P[e];
  }
}

template <typename T>
void launch_dyn_products(const T* s5, const T* u3, T* sa5, T* l6, size_t n_padded, double d2, hipStream_t s,
                         const T* cstar, const DynP& dp) {
  k_dyn_products<T><<<(unsigned)((n_padded + 255) / 256), 256, 0, s>>>(
      s5, u3, sa5, l6, n_padded, d2, cstar, (const T*)dp.w3, (const T*)dp.cpstar, (T*)dp.pa5);
}

template <typename T>
__global__ void __launch_bounds__(256) k_gauss(Geo g, C<T>* __restrict__ f, double td2) {
  const size_t n = (size_t)g.sx * g.my * g.mz;
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n) return;
  const int cx = idx % g.sx;
  if (cx >= g.mx) return;
  const int cy = (idx / g.sx) % g.my;
  const int cz = idx / ((size_t)g.sx * g.my);
  const double kx = g.kx[cx], ky = g.ky[cy], kz = g.kz[cz];
  const T gf = (T)gauss_filter(td2, kx * kx + ky * ky + kz * kz);
  const C<T> v = f[idx];
  f[idx] = mk<T>(gf * v.x, gf * v.y);
}

template <typename T>
void launch_gauss(const Geo& g, void* spec, double td2, hipStream_t s) {
  const size_t n = (size_t)g.sx * g.my * g.mz;
  k_gauss<T><<<(unsigned)((n + 255) / 256), 256, 0, s>>>(g, (C<T>*)spec, td2);
}

template <typename T>
__global__ void __launch_bounds__(256) k_dyn_coef(const T* __restrict__ uf3, const T* __restrict__ m5,
                                                  const T* __restrict__ l6, const T* __restrict__ sa5,
                                                  T* __restrict__ c, size_t n, double td2e, int piomelli,
                                                  const T* __restrict__ wf3, const T* __restrict__ pa5,
                                                  T* __restrict__ cp, int ppio) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const T ux = uf3[i], uy = uf3[n + i], uz = uf3[2 * n + i];
  // L = traceless(symouter(uf, uf) - L)
  const T lxx = ux * ux - l6[i], lxy = ux * uy - l6[n + i], lxz = ux * uz - l6[2 * n + i];
  const T lyy = uy * uy - l6[3 * n + i], lyz = uy * uz - l6[4 * n + i], lzz = uz * uz - l6[5 * n + i];
  const T tr = (lxx + lyy + lzz) / (T)3;
  const T Lxx = lxx - tr, Lyy = lyy - tr, Lzz = lzz - tr;
  const T h[5] = {m5[i], m5[n + i], m5[2 * n + i], m5[3 * n + i], m5[4 * n + i]};
  const T a[5] = {sa5[i], sa5[n + i], sa5[2 * n + i], sa5[3 * n + i], sa5[4 * n + i]};
  T pa[5] = {0, 0, 0, 0, 0};
  if (pa5) {
#pragma unroll
    for (int e = 0; e < 5; ++e) pa[e] = pa5[e * n + i];
  }
  // X:Y of a traceless symmetric tau (5 stored) + L (traceless, full) with a traceless symmetric Y (5 stored)
  auto tl_dot = [&](const T t[5], const T y[5]) {
    const T yzz = -y[0] - y[3];
    return (t[0] + Lxx) * y[0] + (t[3] + Lyy) * y[3] + ((-t[0] - t[3]) + Lzz) * yzz +
           (T)2 * ((t[1] + lxy) * y[1] + (t[2] + lxz) * y[2] + (t[4] + lyz) * y[4]);
  };
  auto dot5 = [&](const T x[5], const T y[5]) {
    const T xzz = -x[0] - x[3], yzz = -y[0] - y[3];
    return x[0] * y[0] + x[3] * y[3] + xzz * yzz + (T)2 * (x[1] * y[1] + x[2] * y[2] + x[4] * y[4]);
  };
  if (piomelli) {  // M = 2 td2e |Sh| Sh, Mn = M / (M:M), c = (tau + L):Mn, tau = sa (+ P under piomelliSandP)
    const T f = (T)2 * (T)td2e * norm_s5<T>(h[0], h[1], h[2], h[3], h[4]);
    T M[5], Mn[5], tau[5];
#pragma unroll
    for (int e = 0; e < 5; ++e) M[e] = f * h[e];
    const T MM = dot5(M, M);
#pragma unroll
    for (int e = 0; e < 5; ++e) {
      Mn[e] = M[e] / MM;
      tau[e] = ppio == 2 ? a[e] + pa[e] : a[e];
    }
    c[i] = fmax((T)-1, fmin((T)1, tl_dot(tau, Mn)));
  } else {  // M = td2e |Sh| Sh - sa, c = 0.5 (L:M)/(M:M)
    const T f = (T)td2e * norm_s5<T>(h[0], h[1], h[2], h[3], h[4]);
    T M[5];
#pragma unroll
    for (int e = 0; e < 5; ++e) M[e] = f * h[e] - a[e];
    const T Mzz = -M[0] - M[3];
    const T LM = Lxx * M[0] + Lyy * M[3] + Lzz * Mzz + (T)2 * (lxy * M[1] + lxz * M[2] + lyz * M[4]);
    const T MM = M[0] * M[0] + M[3] * M[3] + Mzz * Mzz + (T)2 * (M[1] * M[1] + M[2] * M[2] + M[4] * M[4]);
    c[i] = fmax((T)-1, fmin((T)1, (T)0.5 * (LM / MM)));
  }
  if (ppio) {  // the SandP coefficient from Ph = Lie(Sh, AntiSymTen(wf/2)) (dynamic_smagorinsky.jl:185-199)
    T Ph[5];
    lie_s_w<T>(h, wf3[i], wf3[n + i], wf3[2 * n + i], Ph);
    const T zero[5] = {0, 0, 0, 0, 0};
    T Mp[5];
    if (ppio == 1) {  // dynamicSandP: Mp = td2e Ph - P, cp = L:Mp / Mp:Mp
#pragma unroll
      for (int e = 0; e < 5; ++e) Mp[e] = (T)td2e * Ph[e] - pa[e];
      cp[i] = fmax((T)-1, fmin((T)1, tl_dot(zero, Mp) / dot5(Mp, Mp)));
    } else {  // piomelliSandP: Mp = td2e Ph, cp = (tau + L):(Mp / Mp:Mp), tau = sa + P or P
#pragma unroll
      for (int e = 0; e < 5; ++e) Mp[e] = (T)td2e * Ph[e];
      const T MM = dot5(Mp, Mp);
      T Mpn[5], tau[5];
#pragma unroll
      for (int e = 0; e < 5; ++e) {
        Mpn[e] = Mp[e] / MM;
        tau[e] = piomelli ? a[e] + pa[e] : pa[e];
      }
      cp[i] = fmax((T)-1, fmin((T)1, tl_dot(tau, Mpn)));
    }
  }
}

template <typename T>
void launch_dyn_coef(const T* uf3, const T* m5, const T* l6, const T* sa5, T* c, size_t n_padded, double td2e,
                     hipStream_t s, int piomelli, const DynP& dp) {
  const int ppio = dp.wf3 ? (dp.cpstar ? 2 : 1) : 0;
  k_dyn_coef<T><<<(unsigned)((n_padded + 255) / 256), 256, 0, s>>>(uf3, m5, l6, sa5, c, n_padded, td2e, piomelli,
                                                                   (const T*)dp.wf3, (const T*)dp.pa5, (T*)dp.cp,
                                                                   ppio);
}

template <typename T>
__global__ void __launch_bounds__(256) k_piomelli_cstar(T* __restrict__ c, T* __restrict__ cm1,
                                                        const DevState* __restrict__ st,
                                                        const double* __restrict__ dtm1, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double dt = st->dt, dtm = *dtm1;
  const double dtf = dt == 0.0 ? 1.0 : dt;
  const double dtm1f = dtm == 0.0 ? dtf : dtm;
  const T r = (T)(dtf / dtm1f);
  const T ci = c[i], pi = cm1[i];
  cm1[i] = ci;
  c[i] = fma(r, ci - pi, ci);
}

#if FFS_IN_PART(0)
__global__ void k_store_dt(const DevState* __restrict__ st, double* __restrict__ dst) { *dst = st->dt; }
#else
__global__ void k_store_dt(const DevState* __restrict__ st, double* __restrict__ dst);
#endif

template <typename T>
void launch_piomelli_cstar(T* c, T* cm1, const DevState* st, double* dtm1, size_t n_padded, hipStream_t s) {
  k_piomelli_cstar<T><<<(unsigned)((n_padded + 255) / 256), 256, 0, s>>>(c, cm1, st, dtm1, n_padded);
  k_store_dt<<<1, 1, 0, s>>>(st, dtm1);
}

template <typename T>
__global__ void __launch_bounds__(256) k_fill_real(T* __restrict__ a, size_t n, T v) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

template <typename T>
void launch_fill_real(T* a, size_t n, double v, hipStream_t s) {
  k_fill_real<T><<<(unsigned)((n + 255) / 256), 256, 0, s>>>(a, n, (T)v);
}

template <typename T>
__global__ void __launch_bounds__(256) k_dyn_tau(const T* __restrict__ s5, const T* __restrict__ c,
                                                 T* __restrict__ t5, size_t n, double d2, double cmin,
                                                 const T* __restrict__ w3, const T* __restrict__ cp) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const T S[5] = {s5[i], s5[n + i], s5[2 * n + i], s5[3 * n + i], s5[4 * n + i]};
  const T nut = fmax((T)cmin, fmin((T)0.4, c[i])) * (T)d2 * norm_s5<T>(S[0], S[1], S[2], S[3], S[4]);
  T t[5];
#pragma unroll
  for (int e = 0; e < 5; ++e) t[e] = (T)2 * nut * S[e];
  if (cp) {  // + clamp(cp, -0.8, 0.8) Delta^2 Lie(S, AntiSymTen(w/2)) (les_real_calc.jl:116-118)
    T P[5];
    lie_s_w<T>(S, w3[i], w3[n + i], w3[2 * n + i], P);
    const T fp = fmax((T)-0.8, fmin((T)0.8, cp[i])) * (T)d2;
#pragma unroll
    for (int e = 0; e < 5; ++e) t[e] += fp * P[e];
  }
#pragma unroll
  for (int e = 0; e < 5; ++e) t5[e * n + i] = t[e];
}

template <typename T>
void launch_dyn_tau(const T* s5, const T* c, T* t5, size_t n_padded, double d2, double cmin, hipStream_t s,
                    const T* w3, const T* cp) {
  k_dyn_tau<T><<<(unsigned)((n_padded + 255) / 256), 256, 0, s>>>(s5, c, t5, n_padded, d2, cmin, w3, cp);
}

// div!(rhs, uu) of the conservation form (models.jl:176-180): rhs_c = (i k) . uu, uu symmetric
// traceless (xx xy xz yy yz), as the LES (i k) . tau
template <typename T>
__global__ void __launch_bounds__(256) k_div_uu(Geo g, const C<T>* __restrict__ uu, C<T>* __restrict__ rhs) {
  const size_t n = (size_t)g.sx * g.my * g.mz;
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n) return;
  const int cx = idx % g.sx;
  if (cx >= g.mx) return;
  const int cy = (idx / g.sx) % g.my;
  const int cz = idx / ((size_t)g.sx * g.my);
  const double kv3[3] = {g.kx[cx], g.ky[cy], g.kz[cz]};
  const C<T> txx = uu[idx], txy = uu[n + idx], txz = uu[2 * n + idx], tyy = uu[3 * n + idx], tyz = uu[4 * n + idx];
  const double T6[6][2] = {{txx.x, txx.y}, {txy.x, txy.y}, {txz.x, txz.y},
                           {tyy.x, tyy.y}, {tyz.x, tyz.y}, {-(double)txx.x - tyy.x, -(double)txx.y - tyy.y}};
  const int row[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const double sr = kv3[0] * T6[row[c][0]][0] + kv3[1] * T6[row[c][1]][0] + kv3[2] * T6[row[c][2]][0];
    const double si = kv3[0] * T6[row[c][0]][1] + kv3[1] * T6[row[c][1]][1] + kv3[2] * T6[row[c][2]][1];
    rhs[c * n + idx] = mk<T>((T)(-si), (T)sr);
  }
}

template <typename T>
void launch_div_uu(const Geo& g, const void* uu, void* rhs, hipStream_t s) {
  const size_t n = (size_t)g.sx * g.my * g.mz;
  k_div_uu<T><<<(unsigned)((n + 255) / 256), 256, 0, s>>>(g, (const C<T>*)uu, (C<T>*)rhs);
}

template <typename T>
void launch_update(const Geo& g, void* u, void* rhs, const void* fm1, const void* fm2, const DevState* st,
                   const UpdateArgs& a, hipStream_t s) {
  const size_t n = (size_t)g.sx * g.my * g.mz;
  const UCall<T> c{g, (C<T>*)u, (C<T>*)rhs, (const C<T>*)fm1, (const C<T>*)fm2, st, a};
  const dim3 grid((unsigned)((n + 255) / 256));
  if (update_full(a))
    launch_call((const void*)&k_update_b<T, true>, grid, dim3(256), 0, c, "update", [&] {
      k_update<T, true><<<grid, 256, 0, s>>>(c.g, c.u, c.rhs, c.fm1, c.fm2, c.st, c.a);
    });
  else
    launch_call((const void*)&k_update_b<T, false>, grid, dim3(256), 0, c, "update", [&] {
      k_update<T, false><<<grid, 256, 0, s>>>(c.g, c.u, c.rhs, c.fm1, c.fm2, c.st, c.a);
    });
}

template <typename T, int N>
static void run_fwd_z_update(const Geo& g, const void* b1, void* u, void* rhs, const void* fm1, const void* fm2,
                             const DevState* st, const UpdateArgs& a, double scale, const void* tw, hipStream_t s) {
  constexpr int TX = col_tx<T>(N);
  constexpr int TP = col_tper<T, N, false>();
  const size_t lds = (size_t)(N * TX + tw_lds<T>(N)) * sizeof(C<T>);
  static std::once_flag once;
  lds_limit_once(once, (const void*)&k_fwd_z_update<T, N>, lds, "the fused forward z + update");
  static std::once_flag once_b;
  lds_limit_once(once_b, (const void*)&k_fwd_z_update_b<T, N>, lds, "the batched fused forward z + update");
  const dim3 grid((g.mx + TX - 1) / TX, g.my);
  const FzuCall<T> c{g, (const C<T>*)b1, (C<T>*)u, (C<T>*)rhs, (const C<T>*)fm1, (const C<T>*)fm2, st, a,
                     (const C<T>*)tw, (T)scale};
  launch_call((const void*)&k_fwd_z_update_b<T, N>, grid, dim3(TX * TP), lds, c, "fwd_z_update", [&] {
    k_fwd_z_update<T, N><<<grid, TX * TP, lds, s>>>(c.g, c.in, c.u, c.rhs, c.fm1, c.fm2, c.st, c.a, c.tw, c.scale);
  });
}

template <typename T>
void launch_fwd_z_update(const Geo& g, const void* b1, void* u, void* rhs, const void* fm1, const void* fm2,
                         const DevState* st, const UpdateArgs& a, double scale, const void* tw, hipStream_t s) {
  if (update_full(a) || a.step < 1) throw std::runtime_error("fused forward z + update: not a plain step");
  switch (g.nz) {
#define X(n)                                                                                 \
  case n: run_fwd_z_update<T, n>(g, b1, u, rhs, fm1, fm2, st, a, scale, tw, s); break;
    FFS_SIZES(X)
#undef X
    default: bad_size("forward z pass", g.nz);
  }
}

// The reference carries the velocity (and rho, phi) through real space every step (real!(u), then
// myfourier!(u) and pressure_projection!(u), models.jl:122-193).  That round trip is the identity except on
// the kx = 0 and kx = N/2 planes, where the C2R keeps only the Hermitian part of each pair of modes
// k, -k: f(k) <- (f(k) + conj f(-k)) / 2.  This kernel applies it, and the projection after it to the
// velocity (its first three fields), once per step.  It matters only when a Nyquist mode is retained
// (cutoff >= 1): the curl, the divergence and the projection at ky or kz = -N/2, whose partner index is
// itself, are not Hermitian there; with the Nyquist rows dealiased the planes already are.
// Thread (plane, cy, cz) with its partner (cy', cz') = (-ky, -kz): the lower index writes both.
template <typename T>
__global__ void __launch_bounds__(256) k_hermitian_planes(Geo g, C<T>* __restrict__ u, C<T>* __restrict__ rho,
                                                          C<T>* __restrict__ phi, int nplane) {
  const size_t n = (size_t)g.sx * g.my * g.mz;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int per = g.my * g.mz;
  if (t >= nplane * per) return;
  const int plane = t / per, cy = (t % per) % g.my, cz = (t % per) / g.my;
  const int cx = plane == 0 ? 0 : g.mx - 1;  // the second plane is kx = N/2 (retained only when mx = N/2 + 1)
  const int py = g.yinv[(g.ny - g.ymap[cy]) % g.ny], pz = g.zinv[(g.nz - g.zmap[cz]) % g.nz];
  if (py < 0 || pz < 0) return;
  const size_t a = cx + (size_t)g.sx * (cy + (size_t)g.my * cz);
  const size_t b = cx + (size_t)g.sx * (py + (size_t)g.my * pz);
  if (b < a) return;
  auto herm = [&](C<T>* f) {
    const C<T> fa = f[a], fb = f[b];
    const C<T> h = mk<T>((T)0.5 * (fa.x + fb.x), (T)0.5 * (fa.y - fb.y));
    f[a] = h;
    f[b] = mk<T>(h.x, -h.y);
  };
  for (int c = 0; c < 3; ++c) herm(u + c * n);
  if (rho) herm(rho);
  if (phi) herm(phi);
  const size_t mode[2] = {a, b};
  const int my_[2] = {cy, py}, mz_[2] = {cz, pz};
  for (int m = 0; m < (b == a ? 1 : 2); ++m) {
    const size_t idx = mode[m];
    if (idx == 0) continue;
    const double kx = g.kx[cx], ky = g.ky[my_[m]], kz = g.kz[mz_[m]];
    const double K2 = kx * kx + ky * ky + kz * kz;
    const C<T> v0 = u[idx], v1 = u[n + idx], v2 = u[2 * n + idx];
    double ux[2] = {v0.x, v0.y}, uy[2] = {v1.x, v1.y}, uz[2] = {v2.x, v2.y};
    for (int c = 0; c < 2; ++c) {
      const double p1 = -(kx * ux[c] + ky * uy[c] + kz * uz[c]) / K2;
      ux[c] = p1 * kx + ux[c];
      uy[c] = p1 * ky + uy[c];
      uz[c] = p1 * kz + uz[c];
    }
    u[idx] = mk<T>((T)ux[0], (T)ux[1]);
    u[n + idx] = mk<T>((T)uy[0], (T)uy[1]);
    u[2 * n + idx] = mk<T>((T)uz[0], (T)uz[1]);
  }
}

template <typename T>
void launch_hermitian_planes(const Geo& g, void* u, void* rho, void* phi, hipStream_t s) {
  const int nplane = g.mx == g.nxh ? 2 : 1;
  const int nt = nplane * g.my * g.mz;
  k_hermitian_planes<T><<<(nt + 255) / 256, 256, 0, s>>>(g, (C<T>*)u, (C<T>*)rho, (C<T>*)phi, nplane);
}

// Leray projection of a spectral vector field (pressure_projection!, models.jl:443-456)
template <typename T>
__global__ void k_project(Geo g, C<T>* __restrict__ u) {
  const size_t n = (size_t)g.sx * g.my * g.mz;
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n || idx == 0) return;
  const int cx = idx % g.sx;
  if (cx >= g.mx) return;
  const int cy = (idx / g.sx) % g.my;
  const int cz = idx / ((size_t)g.sx * g.my);
  const double kx = g.kx[cx], ky = g.ky[cy], kz = g.kz[cz];
  const double K2 = kx * kx + ky * ky + kz * kz;
  const C<T> v0 = u[idx], v1 = u[n + idx], v2 = u[2 * n + idx];
  double ux[2] = {v0.x, v0.y}, uy[2] = {v1.x, v1.y}, uz[2] = {v2.x, v2.y};
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    const double p1 = -(kx * ux[c] + ky * uy[c] + kz * uz[c]) / K2;
    ux[c] = p1 * kx + ux[c];
    uy[c] = p1 * ky + uy[c];
    uz[c] = p1 * kz + uz[c];
  }
  u[idx] = mk<T>((T)ux[0], (T)ux[1]);
  u[n + idx] = mk<T>((T)uy[0], (T)uy[1]);
  u[2 * n + idx] = mk<T>((T)uz[0], (T)uz[1]);
}

template <typename T>
void launch_project(const Geo& g, void* u, hipStream_t s) {
  const size_t n = (size_t)g.sx * g.my * g.mz;
  k_project<T><<<(n + 255) / 256, 256, 0, s>>>(g, (C<T>*)u);
}

// ===========================================================================
// Statistics (stats.jl velocity_stats :168-219, squared_mean :289-307,
// viscosity_stats :608-635, d?_squared_mean :334-473, pressure_stats
// :750-814, non_linear_stats :897-917, integral_lenght_? :952-980).
// Weight (1 + (i>1)) = 2 for kx > 0.  Each block writes kStatsSlots partial
// sums; one block then adds the partials in a fixed order, so results are
// bit-reproducible run to run.
//  0..2  sum w|u_c|^2            3  sum w k^2 (|ux|^2+|uy|^2)   4  sum w k^2 |uz|^2
//  5,6   nlh, nlv                7,8 pressh, pressv
//  9 sum w kx^2|ux|^2  10 sum w ky^2|ux|^2  11 sum w kx^2|uy|^2  12 sum w ky^2|uy|^2
//  13+3c+{0,1,2}: integral-length plane sums of component c over kx=0 (no weight), ky=0, kz=0
//  density (scalar_stats :221-248, buoyancy_flux density_stratification_types.jl:192-211):
//  22 sum w|rho|^2  23..25 sum w k_i^2 |rho|^2  26 sum w proj(uz, rho)  29 Re rho(k=0)
//  LES (les_stats :679-705): 27 sum w proj((ik.tau)_{x,y}, u_{x,y})  28 the same for z
//  the density's hyperviscous dissipation (scalar_hvis_stats :1086-1170): 37 sum w f_hvd |rho|^2
//  the density's LES flux (scalar_les_stats :1032-1053): 38 sum w proj((i k).F, rho)
enum { kStatsN = 39, kStatsBlocks = 512, kStatsThreads = 128 };
static_assert(kStatsN <= kStatsSlots, "a stats row of partial sums fits its slot");

__device__ __forceinline__ double proj2(double ax, double ay, double bx, double by) { return ax * bx + ay * by; }

template <typename T>
__global__ void __launch_bounds__(kStatsThreads) k_stats(Geo g, const C<T>* __restrict__ u,
                                                        const C<T>* __restrict__ nl, const C<T>* __restrict__ ke,
                                                        StatsAux aux, double* __restrict__ partials) {
  const size_t n = (size_t)g.sx * g.my * g.mz;
  double acc[kStatsN];
#pragma unroll
  for (int i = 0; i < kStatsN; ++i) acc[i] = 0.0;
  const C<T>* tau = (const C<T>*)aux.tau;
  const C<T>* rho = (const C<T>*)aux.rho;
  for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < n; idx += (size_t)gridDim.x * blockDim.x) {
    const int cx = idx % g.sx;
    if (cx >= g.mx) continue;
    const int cy = (idx / g.sx) % g.my;
    const int cz = idx / ((size_t)g.sx * g.my);
    const double kx = g.kx[cx], ky = g.ky[cy], kz = g.kz[cz];
    const double w = cx > 0 ? 2.0 : 1.0;
    const double K2 = kx * kx + ky * ky + kz * kz;
    const C<T> a0 = u[idx], a1 = u[n + idx], a2 = u[2 * n + idx];
    const double ux[2] = {a0.x, a0.y}, uy[2] = {a1.x, a1.y}, uz[2] = {a2.x, a2.y};
    const double ex = ux[0] * ux[0] + ux[1] * ux[1];
    const double ey = uy[0] * uy[0] + uy[1] * uy[1];
    const double ez = uz[0] * uz[0] + uz[1] * uz[1];
    acc[0] += w * ex;
    acc[1] += w * ey;
    acc[2] += w * ez;
    acc[3] += w * K2 * (ex + ey);
    acc[4] += w * K2 * ez;
    if (aux.phi) {  // scalar_stats (stats.jl:221-248)
      const C<T> p0 = ((const C<T>*)aux.phi)[idx];
      const double ep = (double)p0.x * p0.x + (double)p0.y * p0.y;
      acc[32] += w * ep;
      acc[33] += w * kx * kx * ep;
      acc[34] += w * ky * ky * ep;
      acc[35] += w * kz * kz * ep;
      if (idx == 0) acc[36] += p0.x;
    }
    if (aux.hv.kind) {
      const double kv[3] = {kx, ky, kz};
      const double f = hv_weight(aux.hv, kv);
      acc[30] += w * f * (ex + ey);
      acc[31] += w * f * ez;
    }
    acc[9] += w * kx * kx * ex;
    acc[10] += w * ky * ky * ex;
    acc[11] += w * kx * kx * ey;
    acc[12] += w * ky * ky * ey;
    const double e3[3] = {ex, ey, ez};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      if (cx == 0) acc[13 + 3 * c] += e3[c];
      if (ky == 0.0) acc[14 + 3 * c] += w * e3[c];
      if (kz == 0.0) acc[15 + 3 * c] += w * e3[c];
    }
    // (i k) . tau per component (les_stats and the pressure rhs)
    double dtau[3][2] = {{0, 0}, {0, 0}, {0, 0}};
    if (tau) {
      const C<T> txx = tau[idx], txy = tau[n + idx], txz = tau[2 * n + idx], tyy = tau[3 * n + idx],
                 tyz = tau[4 * n + idx];
      const double T6[6][2] = {{txx.x, txx.y}, {txy.x, txy.y}, {txz.x, txz.y},
                               {tyy.x, tyy.y}, {tyz.x, tyz.y}, {-(double)txx.x - tyy.x, -(double)txx.y - tyy.y}};
      const int row[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
      const double kv3[3] = {kx, ky, kz};
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const double sr = kv3[0] * T6[row[c][0]][0] + kv3[1] * T6[row[c][1]][0] + kv3[2] * T6[row[c][2]][0];
        const double si = kv3[0] * T6[row[c][0]][1] + kv3[1] * T6[row[c][1]][1] + kv3[2] * T6[row[c][2]][1];
        dtau[c][0] = -si;
        dtau[c][1] = sr;
      }
      acc[27] += w * (proj2(dtau[0][0], dtau[0][1], ux[0], ux[1]) + proj2(dtau[1][0], dtau[1][1], uy[0], uy[1]));
      acc[28] += w * proj2(dtau[2][0], dtau[2][1], uz[0], uz[1]);
    }
    double rr[2] = {0, 0};
    if (rho) {
      const C<T> r0 = rho[idx];
      rr[0] = r0.x;
      rr[1] = r0.y;
      const double er = rr[0] * rr[0] + rr[1] * rr[1];
      acc[22] += w * er;
      acc[23] += w * kx * kx * er;
      acc[24] += w * ky * ky * er;
      acc[25] += w * kz * kz * er;
      const double* ug = aux.gdir == 0 ? ux : (aux.gdir == 1 ? uy : uz);  // buoyancy_flux's u_gdir
      acc[26] += w * proj2(ug[0], ug[1], rr[0], rr[1]);
      if (idx == 0) acc[29] += rr[0];
      if (aux.hvd.kind) {
        const double kv[3] = {kx, ky, kz};
        acc[37] += w * hv_weight(aux.hvd, kv) * er;
      }
      if (aux.rflux) {  // scalar_les_stats (stats.jl:1032-1053): proj((i k).F, rho) of the total flux F
        const C<T>* fl = (const C<T>*)aux.rflux;
        const C<T> fx = fl[idx], fy = fl[n + idx], fz = fl[2 * n + idx];
        const double dr = kx * fx.x + ky * fy.x + kz * fz.x, di = kx * fx.y + ky * fy.y + kz * fz.y;
        acc[38] += w * proj2(-di, dr, rr[0], rr[1]);
      }
    }
    if (nl) {
      const C<T> n0 = nl[idx], n1 = nl[n + idx], n2 = nl[2 * n + idx];
      const C<T> kk = ke[idx];
      const double kr = kk.x, ki = kk.y;
      // nl - (i k) ke  (== rhs - (i ke) k of pressure_stats)
      const double qx[2] = {n0.x + kx * ki, n0.y - kx * kr};
      const double qy[2] = {n1.x + ky * ki, n1.y - ky * kr};
      const double qz[2] = {n2.x + kz * ki, n2.y - kz * kr};
      acc[5] += w * (proj2(qx[0], qx[1], ux[0], ux[1]) + proj2(qy[0], qy[1], uy[0], uy[1]));
      acc[6] += w * proj2(qz[0], qz[1], uz[0], uz[1]);
      if (idx != 0) {
        // pressure rhs: + (i k).tau (LES) + rho g (density)
        const double gx = aux.gdir == 0 ? aux.gval : 0.0, gy = aux.gdir == 1 ? aux.gval : 0.0,
                     gzz = aux.gdir == 2 ? aux.gval : 0.0;
        const double px[2] = {qx[0] + dtau[0][0] + rr[0] * gx, qx[1] + dtau[0][1] + rr[1] * gx};
        const double py[2] = {qy[0] + dtau[1][0] + rr[0] * gy, qy[1] + dtau[1][1] + rr[1] * gy};
        const double pz[2] = {qz[0] + dtau[2][0] + rr[0] * gzz, qz[1] + dtau[2][1] + rr[1] * gzz};
        const double p1r = -(kx * px[0] + ky * py[0] + kz * pz[0]) / K2;
        const double p1i = -(kx * px[1] + ky * py[1] + kz * pz[1]) / K2;
        acc[7] += w * (proj2(ux[0], ux[1], p1r * kx, p1i * kx) + proj2(uy[0], uy[1], p1r * ky, p1i * ky));
        acc[8] += w * proj2(uz[0], uz[1], p1r * kz, p1i * kz);
      }
    }
  }
  // block reduction in a fixed tree
  __shared__ double red[kStatsN][kStatsThreads];
#pragma unroll
  for (int i = 0; i < kStatsN; ++i) red[i][threadIdx.x] = acc[i];
  __syncthreads();
  for (int off = kStatsThreads / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) {
#pragma unroll
      for (int i = 0; i < kStatsN; ++i) red[i][threadIdx.x] += red[i][threadIdx.x + off];
    }
    __syncthreads();
  }
  if (threadIdx.x < kStatsN) partials[(size_t)blockIdx.x * kStatsSlots + threadIdx.x] = red[threadIdx.x][0];
}

// sums nblk rows of `partials` (each kStatsSlots wide) column-wise in a fixed order
#if !FFS_IN_PART(0)
__global__ void k_reduce_partials(const double* __restrict__ partials, int nblk, int ncols, double* __restrict__ out);
#else
__global__ void k_reduce_partials(const double* __restrict__ partials, int nblk, int ncols, double* __restrict__ out) {
  __shared__ double red[256];
  for (int col = 0; col < ncols; ++col) {
    double a = 0.0;
    for (int b = threadIdx.x; b < nblk; b += blockDim.x) a += partials[(size_t)b * kStatsSlots + col];
    red[threadIdx.x] = a;
    __syncthreads();
    for (int off = blockDim.x / 2; off > 0; off >>= 1) {
      if (threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
      __syncthreads();
    }
    if (threadIdx.x == 0) out[col] = red[0];
    __syncthreads();
  }
}
#endif

template <typename T>
int launch_stats(const Geo& g, const void* u, const void* nl, const void* ke, const StatsAux& aux, double* partials,
                 double* out, hipStream_t s) {
  const size_t n = (size_t)g.sx * g.my * g.mz;
  int nblk = (int)((n + kStatsThreads - 1) / kStatsThreads);
  if (nblk > kStatsBlocks) nblk = kStatsBlocks;
  k_stats<T><<<nblk, kStatsThreads, 0, s>>>(g, (const C<T>*)u, (const C<T>*)nl, (const C<T>*)ke, aux, partials);
  k_reduce_partials<<<1, 256, 0, s>>>(partials, nblk, kStatsN, out);
  return kStatsN;
}

// ===========================================================================
// Spectrum marginals for calculate_specHV (spec.jl:424-450) of spectrum_u
// (spec.jl:316-327): per mode h = (|ux|^2+|uy|^2)/2, v = |uz|^2/2.  The
// binned spectra are sums of marginals along one axis, so the device
// produces, for h and v alike,
//   X[cx]   = sum_{cy,cz} e           Y0[cy] = sum_cz e(cx=0)
//   Y[cy]   = sum_{cz,cx>0} e         Z0[cz] = sum_cy e(cx=0)
//   Z[cz]   = sum_{cy,cx>0} e
// and the host bins them (a few hundred values).  Plane partials go through
// `partials` (mz x (mx + my) x 2) and are reduced in a fixed order.
template <typename T>
__global__ void k_spec_planes(Geo g, const C<T>* __restrict__ u, double* __restrict__ px, double* __restrict__ py,
                              double* __restrict__ pz) {
  // one block per z plane; layouts: px[(cz*2 + hv)*mx + cx], py[(cz*2+hv)*my + cy] (cx>0 part),
  // y0 values are read directly by the second kernel; pz[cz*4 + {h0,v0,h,v}]
  const int cz = blockIdx.x;
  const size_t n = (size_t)g.sx * g.my * g.mz;
  __shared__ double sh[4][256];
  double z0h = 0, z0v = 0, zh = 0, zv = 0;
  // X partials: each thread owns cx values, loops over cy
  for (int cx = threadIdx.x; cx < g.mx; cx += blockDim.x) {
    double sh_ = 0, sv_ = 0;
    for (int cy = 0; cy < g.my; ++cy) {
      const size_t idx = cx + (size_t)g.sx * (cy + (size_t)g.my * cz);
      const C<T> a0 = u[idx], a1 = u[n + idx], a2 = u[2 * n + idx];
      const double h = ((double)a0.x * a0.x + (double)a0.y * a0.y + (double)a1.x * a1.x + (double)a1.y * a1.y) / 2;
      const double v = ((double)a2.x * a2.x + (double)a2.y * a2.y) / 2;
      sh_ += h;
      sv_ += v;
    }
    px[((size_t)cz * 2 + 0) * g.mx + cx] = sh_;
    px[((size_t)cz * 2 + 1) * g.mx + cx] = sv_;
  }
  // Y (cx > 0) partials and Z sums: each thread owns cy values, loops over cx
  for (int cy = threadIdx.x; cy < g.my; cy += blockDim.x) {
    double sh_ = 0, sv_ = 0;
    for (int cx = 0; cx < g.mx; ++cx) {
      const size_t idx = cx + (size_t)g.sx * (cy + (size_t)g.my * cz);
      const C<T> a0 = u[idx], a1 = u[n + idx], a2 = u[2 * n + idx];
      const double h = ((double)a0.x * a0.x + (double)a0.y * a0.y + (double)a1.x * a1.x + (double)a1.y * a1.y) / 2;
      const double v = ((double)a2.x * a2.x + (double)a2.y * a2.y) / 2;
      if (cx == 0) {
        z0h += h;
        z0v += v;
      } else {
        sh_ += h;
        sv_ += v;
      }
    }
    py[((size_t)cz * 2 + 0) * g.my + cy] = sh_;
    py[((size_t)cz * 2 + 1) * g.my + cy] = sv_;
    zh += sh_;
    zv += sv_;
  }
  sh[0][threadIdx.x] = z0h;
  sh[1][threadIdx.x] = z0v;
  sh[2][threadIdx.x] = zh;
  sh[3][threadIdx.x] = zv;
  __syncthreads();
  for (int off = blockDim.x / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off)
      for (int i = 0; i < 4; ++i) sh[i][threadIdx.x] += sh[i][threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x < 4) pz[(size_t)cz * 4 + threadIdx.x] = sh[threadIdx.x][0];
}

template <typename T>
__global__ void k_spec_finish(Geo g, const C<T>* __restrict__ u, const double* __restrict__ px,
                              const double* __restrict__ py, double* __restrict__ ox, double* __restrict__ oy0,
                              double* __restrict__ oy) {
  // ox[hv*mx + cx] = sum_cz px ; oy[hv*my + cy] = sum_cz py ; oy0[hv*my+cy] = sum_cz e(0,cy,cz)
  const size_t n = (size_t)g.sx * g.my * g.mz;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < 2 * g.mx) {
    const int hv = t / g.mx, cx = t % g.mx;
    double a = 0;
    for (int cz = 0; cz < g.mz; ++cz) a += px[((size_t)cz * 2 + hv) * g.mx + cx];
    ox[t] = a;
  } else if (t < 2 * g.mx + 2 * g.my) {
    const int tt = t - 2 * g.mx;
    const int hv = tt / g.my, cy = tt % g.my;
    double a = 0, a0 = 0;
    for (int cz = 0; cz < g.mz; ++cz) {
      a += py[((size_t)cz * 2 + hv) * g.my + cy];
      const size_t idx = (size_t)g.sx * (cy + (size_t)g.my * cz);
      double e;
      if (hv == 0) {
        const C<T> a0v = u[idx], a1v = u[n + idx];
        e = ((double)a0v.x * a0v.x + (double)a0v.y * a0v.y + (double)a1v.x * a1v.x + (double)a1v.y * a1v.y) / 2;
      } else {
        const C<T> a2v = u[2 * n + idx];
        e = ((double)a2v.x * a2v.x + (double)a2v.y * a2v.y) / 2;
      }
      a0 += e;
    }
    oy[tt] = a;
    oy0[tt] = a0;
  }
}

template <typename T>
void launch_spectrum(const Geo& g, const void* u, double* out_x, double* out_y0, double* out_y, double* out_z0,
                     double* out_z, double* partials, hipStream_t s) {
  // partials layout: px [mz*2*mx] | py [mz*2*my] | pz [mz*4]
  double* px = partials;
  double* py = px + (size_t)g.mz * 2 * g.mx;
  double* pz = py + (size_t)g.mz * 2 * g.my;
  k_spec_planes<T><<<g.mz, 256, 0, s>>>(g, (const C<T>*)u, px, py, pz);
  const int nt = 2 * g.mx + 2 * g.my;
  k_spec_finish<T><<<(nt + 255) / 256, 256, 0, s>>>(g, (const C<T>*)u, px, py, out_x, out_y0, out_y);
  (void)out_z0;
  (void)out_z;  // z marginals are the pz plane sums (copied by the host)
}

// ===========================================================================
// The spectra of write_spec / write_spec_forcing (spec.jl:1-168), all kSpecQ
// per-mode quantities of one mode at once (proj(a, b) = muladd(a.re, b.re,
// a.im b.im), forcing.jl:4):
//   u      spectrum_u (spec.jl:316-327): h = (|ux|^2+|uy|^2)/2, v = |uz|^2/2
//   vis    spectrum_viscosity (:340-354): 2 nu k^2 times the u quantities
//   nl     spectrum_non_linear (:210-222): proj(nl - i k ke, u), h = x + y, v = z
//   press  spectrum_pressure (:370-422): proj(u, p1 k), p1 = -(k . r)/k^2 with
//          r = nl - i ke k (+ (i k).tau) (+ rho g), 0 at k = 0
//   les    spectrum_les (:224-235): proj((i k).tau, u)
//   buoy   spectrum_buoyancy (:249-257): proj(uz, g_z rho)
//   rho    spectrum_rho (:329-338): |rho|^2;  rhodiss (:356-368): alpha k^2 |rho|^2
//   rhonl  spectrum_rhonl (:237-247): proj((i k).flux, rho)
//   forcing spectrum_forcing (:259-268): (proj(ux, fx) + proj(uy, fy)) / dt with
//          f = factor u on the forced modes
//   hvis   spectrum_hyperviscosity (:297-314): nu_h k^2M proj(u, u) (h = x + y, v = z), or
//          spectrum_spectral_barrier (:270-295) with f(|k|) in place of nu_h k^2M
// The nonlinear term is zero at k = 0 (rhs.c[1] = 0, models.jl:174).
__device__ __forceinline__ double sproj(double ar, double ai, double br, double bi) { return fma(ar, br, ai * bi); }

template <typename T>
__device__ __forceinline__ void spec_values(const Geo& g, const C<T>* __restrict__ u, const SpecArgs& a, int cx,
                                            int cy, int cz, double v[kSpecQ]) {
  const size_t n = (size_t)g.sx * g.my * g.mz;
  const size_t idx = cx + (size_t)g.sx * (cy + (size_t)g.my * cz);
  const double K[3] = {g.kx[cx], g.ky[cy], g.kz[cz]};
  const bool origin = idx == 0;
  double U[3][2], N[3][2];
  const C<T>* nl = (const C<T>*)a.nl;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const C<T> x = u[c * n + idx];
    U[c][0] = x.x;
    U[c][1] = x.y;
    const C<T> y = nl[c * n + idx];
    N[c][0] = origin ? 0.0 : (double)y.x;
    N[c][1] = origin ? 0.0 : (double)y.y;
  }
  const C<T> ke = ((const C<T>*)a.ke)[idx];
  const double k2 = fma(K[0], K[0], fma(K[1], K[1], K[2] * K[2]));
  const double K2 = K[0] * K[0] + K[1] * K[1] + K[2] * K[2];
  const double uh = (sproj(U[0][0], U[0][1], U[0][0], U[0][1]) + sproj(U[1][0], U[1][1], U[1][0], U[1][1])) / 2;
  const double uv = sproj(U[2][0], U[2][1], U[2][0], U[2][1]) / 2;
  v[SQ_U_H] = uh;
  v[SQ_U_V] = uv;
  v[SQ_VIS_H] = 2 * a.nu * k2 * uh;
  v[SQ_VIS_V] = 2 * a.nu * k2 * uv;
  // r = nl - (i ke) k
  double R[3][2];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    R[c][0] = N[c][0] + (double)ke.y * K[c];
    R[c][1] = N[c][1] - (double)ke.x * K[c];
  }
  v[SQ_NL_H] = sproj(R[0][0], R[0][1], U[0][0], U[0][1]) + sproj(R[1][0], R[1][1], U[1][0], U[1][1]);
  v[SQ_NL_V] = sproj(R[2][0], R[2][1], U[2][0], U[2][1]);
  double D[3][2] = {{0, 0}, {0, 0}, {0, 0}};
  if (a.tau) {  // (i k) . tau, tau symmetric traceless (xx xy xz yy yz)
    const C<T>* t = (const C<T>*)a.tau;
    const C<T> txx = t[idx], txy = t[n + idx], txz = t[2 * n + idx], tyy = t[3 * n + idx], tyz = t[4 * n + idx];
    const double T6[6][2] = {{txx.x, txx.y}, {txy.x, txy.y}, {txz.x, txz.y},
                             {tyy.x, tyy.y}, {tyz.x, tyz.y}, {-(double)txx.x - tyy.x, -(double)txx.y - tyy.y}};
    const int row[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const double sr = K[0] * T6[row[c][0]][0] + K[1] * T6[row[c][1]][0] + K[2] * T6[row[c][2]][0];
      const double si = K[0] * T6[row[c][0]][1] + K[1] * T6[row[c][1]][1] + K[2] * T6[row[c][2]][1];
      D[c][0] = -si;
      D[c][1] = sr;
      R[c][0] += D[c][0];
      R[c][1] += D[c][1];
    }
  }
  v[SQ_LES_H] = sproj(D[0][0], D[0][1], U[0][0], U[0][1]) + sproj(D[1][0], D[1][1], U[1][0], U[1][1]);
  v[SQ_LES_V] = sproj(D[2][0], D[2][1], U[2][0], U[2][1]);
  double rr = 0, ri = 0;
  if (a.rho) {
    const C<T> r0 = ((const C<T>*)a.rho)[idx];
    rr = r0.x;
    ri = r0.y;
    R[a.gdir][0] += rr * a.gval;
    R[a.gdir][1] += ri * a.gval;
  }
  double P[2];
#pragma unroll
  for (int p = 0; p < 2; ++p) P[p] = origin ? 0.0 : -(K[0] * R[0][p] + K[1] * R[1][p] + K[2] * R[2][p]) / K2;
  // pressure = p1 k: proj(u_c, p1 k_c)
  v[SQ_PRESS_H] = sproj(U[0][0], U[0][1], P[0] * K[0], P[1] * K[0]) + sproj(U[1][0], U[1][1], P[0] * K[1], P[1] * K[1]);
  v[SQ_PRESS_V] = sproj(U[2][0], U[2][1], P[0] * K[2], P[1] * K[2]);
  v[SQ_BUOY] = sproj(U[2][0], U[2][1], a.gz * rr, a.gz * ri);
  const double rho2 = sproj(rr, ri, rr, ri);
  v[SQ_RHO] = rho2;
  v[SQ_RHODISS] = a.alpha * k2 * rho2;
  double nr = 0, ni = 0;
  if (a.flux) {  // (i k) . flux
    const C<T>* f = (const C<T>*)a.flux;
    const C<T> fx = f[idx], fy = f[n + idx], fz = f[2 * n + idx];
    const double sr = K[0] * fx.x + K[1] * fy.x + K[2] * fz.x;
    const double si = K[0] * fx.y + K[1] * fy.y + K[2] * fz.y;
    nr = -si;
    ni = sr;
  }
  v[SQ_RHONL] = sproj(nr, ni, rr, ri);
  double f = 0.0;
  if (a.pf_factor) {
    if (cz == 0) {
      f = a.pf_factor[cx + (size_t)g.sx * cy];
    } else if (cx == 0 && cy == 0) {
      const int j = z_fft(g, cz), kz = j < (g.nz + 1) / 2 ? j : j - g.nz;
      const int akz = kz < 0 ? -kz : kz;
      if (akz >= a.pf_kz0 && akz <= 5) f = a.st->force_vert;
    }
  }
  double fk = 0.0;
  if (a.hv.kind == 1) fk = a.hv.nuh * hv_weight(a.hv, K);
  else if (a.hv.kind == 2) fk = hv_weight(a.hv, K);
  v[SQ_HVIS_H] = fk * (sproj(U[0][0], U[0][1], U[0][0], U[0][1]) + sproj(U[1][0], U[1][1], U[1][0], U[1][1]));
  v[SQ_HVIS_V] = fk * sproj(U[2][0], U[2][1], U[2][0], U[2][1]);
  v[SQ_FORCING] = f == 0.0 ? 0.0
                           : (sproj(U[0][0], U[0][1], U[0][0] * f, U[0][1] * f) +
                              sproj(U[1][0], U[1][1], U[1][0] * f, U[1][1] * f)) / a.st->dt;
}

// one block per z plane: px[q][cz][cx] = sum_cy, py[q][cz][cy] = sum_{cx>0}, py0[q][cz][cy] = value at cx = 0,
// pz[q][cz][0|1] = sums over the plane at cx = 0 | cx > 0; every sum in a fixed order
template <typename T>
__global__ void __launch_bounds__(256) k_spectra_planes(Geo g, const C<T>* __restrict__ u, SpecArgs a,
                                                       double* __restrict__ part) {
  const int cz = blockIdx.x;
  const size_t per = (size_t)g.mz;
  double* px = part;
  double* py = px + kSpecQ * per * g.mx;
  double* py0 = py + kSpecQ * per * g.my;
  double* pz = py0 + kSpecQ * per * g.my;
  double v[kSpecQ];
  for (int cx = threadIdx.x; cx < g.mx; cx += blockDim.x) {
    double acc[kSpecQ];
#pragma unroll
    for (int q = 0; q < kSpecQ; ++q) acc[q] = 0;
    for (int cy = 0; cy < g.my; ++cy) {
      spec_values<T>(g, u, a, cx, cy, cz, v);
#pragma unroll
      for (int q = 0; q < kSpecQ; ++q) acc[q] += v[q];
    }
#pragma unroll
    for (int q = 0; q < kSpecQ; ++q) px[((size_t)q * per + cz) * g.mx + cx] = acc[q];
  }
  double z0[kSpecQ], z1[kSpecQ];
#pragma unroll
  for (int q = 0; q < kSpecQ; ++q) z0[q] = z1[q] = 0;
  for (int cy = threadIdx.x; cy < g.my; cy += blockDim.x) {
    double acc[kSpecQ];
#pragma unroll
    for (int q = 0; q < kSpecQ; ++q) acc[q] = 0;
    for (int cx = 0; cx < g.mx; ++cx) {
      spec_values<T>(g, u, a, cx, cy, cz, v);
      if (cx == 0) {
#pragma unroll
        for (int q = 0; q < kSpecQ; ++q) {
          py0[((size_t)q * per + cz) * g.my + cy] = v[q];
          z0[q] += v[q];
        }
      } else {
#pragma unroll
        for (int q = 0; q < kSpecQ; ++q) acc[q] += v[q];
      }
    }
#pragma unroll
    for (int q = 0; q < kSpecQ; ++q) {
      py[((size_t)q * per + cz) * g.my + cy] = acc[q];
      z1[q] += acc[q];
    }
  }
  __shared__ double sh[256];
  for (int q = 0; q < 2 * kSpecQ; ++q) {
    double mine = 0;
#pragma unroll
    for (int r = 0; r < kSpecQ; ++r)
      if (r == q % kSpecQ) mine = q < kSpecQ ? z0[r] : z1[r];
    sh[threadIdx.x] = mine;
    __syncthreads();
    for (int off = blockDim.x / 2; off > 0; off >>= 1) {
      if ((int)threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
      __syncthreads();
    }
    if (threadIdx.x == 0) pz[((size_t)(q % kSpecQ) * per + cz) * 2 + (q < kSpecQ ? 0 : 1)] = sh[0];
    __syncthreads();
  }
}

// out[q] = X[mx] | Y0[my] | Y[my] | Z0[mz] | Z[mz], the plane partials summed over cz in order
#if !FFS_IN_PART(0)
__global__ void k_spectra_finish(Geo g, const double* __restrict__ part, double* __restrict__ out);
#else
__global__ void k_spectra_finish(Geo g, const double* __restrict__ part, double* __restrict__ out) {
  const size_t per = (size_t)g.mz;
  const double* px = part;
  const double* py = px + kSpecQ * per * g.mx;
  const double* py0 = py + kSpecQ * per * g.my;
  const double* pz = py0 + kSpecQ * per * g.my;
  const size_t L = spectra_out_len(g);
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= kSpecQ * L) return;
  const int q = (int)(t / L);
  const size_t r = t % L;
  double acc = 0;
  if (r < (size_t)g.mx) {
    for (int cz = 0; cz < g.mz; ++cz) acc += px[((size_t)q * per + cz) * g.mx + r];
  } else if (r < (size_t)g.mx + g.my) {
    const size_t cy = r - g.mx;
    for (int cz = 0; cz < g.mz; ++cz) acc += py0[((size_t)q * per + cz) * g.my + cy];
  } else if (r < (size_t)g.mx + 2 * g.my) {
    const size_t cy = r - g.mx - g.my;
    for (int cz = 0; cz < g.mz; ++cz) acc += py[((size_t)q * per + cz) * g.my + cy];
  } else if (r < (size_t)g.mx + 2 * g.my + g.mz) {
    acc = pz[((size_t)q * per + (r - g.mx - 2 * g.my)) * 2 + 0];
  } else {
    acc = pz[((size_t)q * per + (r - g.mx - 2 * g.my - g.mz)) * 2 + 1];
  }
  out[t] = acc;
}
#endif

template <typename T>
void launch_spectra(const Geo& g, const void* u, const SpecArgs& a, double* partials, double* out, hipStream_t s) {
  k_spectra_planes<T><<<g.mz, 256, 0, s>>>(g, (const C<T>*)u, a, partials);
  const size_t nt = kSpecQ * spectra_out_len(g);
  k_spectra_finish<<<(unsigned)((nt + 255) / 256), 256, 0, s>>>(g, partials, out);
}

// ===========================================================================
// isotropic constant-power forcing: factor = P / sum_{0<|k|^2<=kf2} w |u|^2 over
// the (few) band modes listed by the host, summed in list order by one block.
template <typename T>
__global__ void k_forcing_sum(Geo g, const C<T>* __restrict__ u, const int* __restrict__ band, int nband,
                              DevState* st, double power) {
  const size_t n = (size_t)g.sx * g.my * g.mz;
  double a = 0;
  for (int b = threadIdx.x; b < nband; b += blockDim.x) {
    const int idx = band[b];
    const double w = (idx % g.sx) > 0 ? 2.0 : 1.0;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const C<T> v = u[c * n + idx];
      a += w * ((double)v.x * v.x + (double)v.y * v.y);
    }
  }
  __shared__ double red[64];
  red[threadIdx.x] = a;
  __syncthreads();
  for (int off = 32; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) st->force_factor = red[0] > 0 ? power / red[0] : 0.0;
}

template <typename T>
void launch_forcing_sum(const Geo& g, const void* u, const int* band, int nband, DevState* st, double power,
                        hipStream_t s) {
  k_forcing_sum<T><<<1, 64, 0, s>>>(g, (const C<T>*)u, band, nband, st, power);
}

// ===========================================================================
// Variable time step (set_dt!, time_step_functions.jl:86-129) with the history
// shift of the integrator (AB3 Adams_Bashforth.jl:166-176 halves the first
// step; ETD3rdO ETD.jl:110-114 shifts) and the coefficients of the step
// (AB3 :178-218; ETD3 scalars of set_ABCt! ETD.jl:142-227).  `it` is
// s.iteration[], `tit` the integrator's own iteration counter.
__host__ __device__ inline void ab3_coefficients(DevState* st) {
#pragma clang fp contract(off)
  const double dt = st->dt, dt2 = st->dt2, dt3 = st->dt3;
  double At = 1. + (dt * (2 * dt + 6 * dt2 + 3 * dt3)) / (6 * dt2 * (dt2 + dt3));
  At *= dt;
  double Bt = -dt * dt * (2 * dt + 3 * dt2 + 3 * dt3) / (6 * dt2 * dt3 * (dt + dt2));
  Bt *= (dt + dt2);
  double Ct = dt * dt * (2 * dt + 3 * dt2) / (6 * dt3 * (dt + dt2 + dt3) * (dt2 + dt3));
  Ct *= (dt + dt2 + dt3);
  st->At = At;
  st->Bt = Bt;
  st->Ct = Ct;
}

__host__ __device__ inline void etd_scalars(DevState* st) {
#pragma clang fp contract(off)
  const double dt = st->dt, dt2 = st->dt2, dt3 = st->dt3;
  double* e = st->etd;
  e[0] = (dt * (2 * dt * dt + 6 * dt2 * (dt2 + dt3) + 3 * dt * (2 * dt2 + dt3))) / (6 * dt2 * (dt2 + dt3));
  e[1] = (dt * dt * (dt * dt + 6 * dt2 * (dt2 + dt3) + 2 * dt * (2 * dt2 + dt3))) / (12 * dt2 * (dt2 + dt3));
  e[2] = (dt * dt * dt * (4 * dt2 * (dt2 + dt3) + dt * (2 * dt2 + dt3))) / (24 * dt2 * (dt2 + dt3));
  e[3] = dt * dt * dt * dt / 24;
  e[4] = -((dt * dt * (2 * dt + 3 * (dt2 + dt3))) / (6 * (dt2 * dt3)));
  e[5] = -(((dt * dt * dt * (dt + 2 * (dt2 + dt3)))) / (12 * (dt2 * dt3)));
  e[6] = -((dt * dt * dt * dt * (dt2 + dt3)) / (24 * (dt2 * dt3)));
  e[7] = (dt * dt * (2 * dt + 3 * dt2)) / (6 * dt3 * (dt2 + dt3));
  e[8] = (dt * dt * dt * (dt + 2 * dt2)) / (12 * dt3 * (dt2 + dt3));
  e[9] = (dt * dt * dt * dt * dt2) / (24 * dt3 * (dt2 + dt3));
  e[10] = dt * dt + dt * (2 * dt2 + dt3);  // e1
  e[11] = dt2 * (dt2 + dt3);               // e2
  e[12] = 2 * dt;                          // e3
  e[13] = 2 * dt2 + dt3;                   // e4
  e[14] = dt * (dt + dt2 + dt3);           // e5
  e[15] = dt2 + dt3;                       // e6
  e[16] = dt2 * dt3;                       // e7
  e[17] = dt * (dt + dt2);                 // e8
  e[18] = dt3 * (dt2 + dt3);               // e9
}

__device__ void set_dt_body(DevState* st, int it, int tit, double cfl, double dx, int etd, double gnorm) {
  unsigned long long ub = 0ull;
#pragma unroll
  for (int k = 0; k < kMaxSlots; ++k) {
    unsigned long long* slot = max_slot(st->umax.v, k);
    ub = *slot > ub ? *slot : ub;
    *slot = 0ull;  // cleared for the next step's atomicMax
  }
  const double umax = __longlong_as_double((long long)ub);
  st->umax_d = umax;
  if (it == 0) cfl *= 0.1;
  else if (it == 1) cfl *= 0.2;
  else if (it == 2) cfl *= 0.4;
  else if (it == 3) cfl *= 0.8;
  double newdt = cfl * dx / umax;
  if (gnorm > 0) {  // density: newdt = min(newdt, cfl sqrt(dx / (rho_max g)))
    unsigned long long rb = 0ull;
#pragma unroll
    for (int k = 0; k < kMaxSlots; ++k) {
      unsigned long long* slot = max_slot(st->rhomax.v, k);
      rb = *slot > rb ? *slot : rb;
      *slot = 0ull;
    }
    const double rmax = __longlong_as_double((long long)rb);
    st->rhomax_d = rmax;
    newdt = fmin(newdt, cfl * sqrt(dx / (rmax * gnorm)));
  }
  if (it > 3) {
    const double dt = st->dt;
    if (newdt < dt) newdt = fmax(dt / 1.25, newdt);
    else newdt = fmin(1.25 * dt, newdt);
  }
  if (!etd && tit == 0) {
    st->dt3 = newdt / 2;
    st->dt2 = newdt / 2;
    st->dt = newdt / 2;
  } else {
    st->dt3 = st->dt2;
    st->dt2 = st->dt;
    st->dt = newdt;
  }
  ab3_coefficients(st);
  etd_scalars(st);
}

#if FFS_IN_PART(0)
__global__ void k_set_dt(DevState* st, int it, int tit, double cfl, double dx, int etd, double gnorm) {
  set_dt_body(st, it, tit, cfl, dx, etd, gnorm);
}
#endif

// kz = 0-plane forcing, one block (Forcing_methods/forcing.jl, cforcing.jl, nrforcing.jl,
// cnrforcing.jl, aforcing.jl, qfforcing.jl).  Thread s < nslot sums the spectra of its slot over its
// elements (in list order, from the per-slot index): Ef (and N for nRf / cNRf, with the final rhs formed per element as
// the update kernel forms it); thread 0 then runs the reference's serial part: the oscillator R, the
// factors, the horizontal power hp, the vertical factor and power vp.  Runs after set_dt.
__device__ __forceinline__ double proj_c(double ar, double ai, double br, double bi) { return fma(ar, br, ai * bi); }

template <typename T>
__global__ void k_plane_forcing(Geo g, const C<T>* __restrict__ u, const C<T>* __restrict__ rhs, DevState* st,
                                PlaneForcing pf, UpdateArgs ua, int rhs_final, int it) {
  const size_t n = (size_t)g.sx * g.my * g.mz;
  const C<T>* u1 = u;
  const C<T>* u2 = u + n;
  const int kind = pf.kind < 0 ? -pf.kind : pf.kind;
  const bool with_n = kind == PF_NRF || kind == PF_CNRF;
  for (int sl = threadIdx.x; sl < pf.nslot; sl += blockDim.x) {
    double e = 0.0, nn = 0.0;
    for (int jj = pf.slot_ptr[sl]; jj < pf.slot_ptr[sl + 1]; ++jj) {
      const int i = pf.slot_el[jj];
      const int idx = pf.elem[i];
      const C<T> a = u1[idx], b = u2[idx];
      const double m2 = (double)a.x * a.x + (double)a.y * a.y + ((double)b.x * b.x + (double)b.y * b.y);
      // calculate_u1u2_spectrum! (spectrum.jl:1-20), calculate_abs2! (:22-37), _mag_ (:44-65)
      e += kind == PF_QF ? pf.ew[i] * sqrt(m2) : 0.5 * pf.ew[i] * m2;
      if (with_n) {  // calculate_nn_spectrum! (cnrforcing.jl:103-126) on the kz = 0 plane
        double U[3][2], R[3][2], RH[2] = {0, 0};
        for (int c = 0; c < 3; ++c) {
          const C<T> v = u[c * n + idx], r = rhs[c * n + idx];
          U[c][0] = v.x; U[c][1] = v.y;
          R[c][0] = r.x; R[c][1] = r.y;
        }
        if (!rhs_final) {
          const int cx = idx % g.sx, cy = idx / g.sx;
          const double kv3[3] = {g.kx[cx], g.ky[cy], g.kz[0]};
          const double K2 = kv3[0] * kv3[0] + kv3[1] * kv3[1] + kv3[2] * kv3[2];
          form_rhs<T>(ua, idx, n, kv3, K2, idx == 0, U, R, RH);
        }
        const double k2 = pf.ek2[i];
        nn += pf.ew[i] * (proj_c(U[0][0], U[0][1], R[0][0] - pf.nu * k2 * U[0][0], R[0][1] - pf.nu * k2 * U[0][1]) +
                          proj_c(U[1][0], U[1][1], R[1][0] - pf.nu * k2 * U[1][0], R[1][1] - pf.nu * k2 * U[1][1]));
      }
    }
    pf.Ef[sl] = e;
    if (with_n) {
      if (pf.kind < 0) pf.Nm1[sl] = nn;
      else pf.Nm1[sl] = pf.N[sl];
      pf.N[sl] = nn;
    }
    if (pf.kind < 0) pf.Em[sl] = e;  // initialize!: the target is the initial spectrum
  }
  __syncthreads();
  if (threadIdx.x != 0 || pf.kind < 0) return;
  const double dt = st->dt;
  double hp = 0.0, vp = 0.0, fvert = 0.0;
  if (kind != PF_QF) {
    double omega;
    if (kind == PF_CRF) {
      omega = 2 * 3.141592653589793 / pf.tf;  // ts = Tf
    } else if (kind == PF_CNRF) {
      omega = 2 * 3.141592653589793 / (pf.tf * pf.tf);  // ts = Tf, omega = 2 pi / (Tf ts)
    } else if (kind == PF_A) {
      omega = 2 * 3.141592653589793 / (pf.tf * (10 * dt));  // ts = 10 dt
    } else {
      double ts = dt / pf.cfl;
      if (pf.variable_dt) {
        if (it == 0) ts /= 0.1;
        else if (it == 1) ts /= 0.2;
        else if (it == 2) ts /= 0.4;
        else if (it == 3) ts /= 0.8;
      }
      omega = 2 * 3.141592653589793 / (pf.tf * ts);
    }
    const double dtm2 = st->dt2;  // s.timestep.x.dt2[]
    for (int sl = 0; sl < pf.nslot; ++sl) {
      double R = pf.R[sl];
      if (with_n)
        R += dt * (-2 * pf.alpha * omega * (R + pf.N[sl]) - (pf.N[sl] - pf.Nm1[sl]) / dtm2 -
                   omega * omega * (pf.Ef[sl] - pf.Em[sl]));
      else
        R += dt * (-2 * pf.alpha * omega * R - omega * omega * (pf.Ef[sl] - pf.Em[sl]));
      R = fmax(0.0, R);
      pf.R[sl] = R;
      pf.slot_factor[sl] = sqrt(R / fmax(pf.Ef[sl], 2.220446049250313e-16)) * pf.Zf[sl] * dt;
    }
    double ff = 0.0, fv = 0.0;
    for (int i = 0; i < pf.nel; ++i) {
      const double f = pf.slot_factor[pf.eslot[i]];
      pf.plane_factor[pf.elem[i]] = f;
      const C<T> a = u1[pf.elem[i]], b = u2[pf.elem[i]];
      const double s1r = a.x * f, s1i = a.y * f, s2r = b.x * f, s2i = b.y * f;
      ff += (s1r * s1r + s1i * s1i + (s2r * s2r + s2i * s2i)) * pf.ew[i];
      fv += (proj_c(a.x, a.y, s1r, s1i) + proj_c(b.x, b.y, s2r, s2i)) * pf.ew[i];
    }
    hp = ff * 0.5 / dt + fv / dt;
    // vertical: kx = ky = 0, kz = 1..5 and -1..-5, 1 % of the power
    double umag = 0.0;
    for (int k = 0; k < 5; ++k) {
      const C<T> a = u1[pf.vert[2 * k]], b = u2[pf.vert[2 * k]];
      const C<T> c = u1[pf.vert[2 * k + 1]], d = u2[pf.vert[2 * k + 1]];
      umag += sqrt((double)a.x * a.x + (double)a.y * a.y + ((double)b.x * b.x + (double)b.y * b.y)) +
              sqrt((double)c.x * c.x + (double)c.y * c.y + ((double)d.x * d.x + (double)d.y * d.y));
    }
    fvert = dt * sqrt(2.0 * hp * (1.0 - 0.99)) / umag;
    ff = 0.0;
    fv = 0.0;
    for (int k = 0; k < 5; ++k)
      for (int h = 0; h < 2; ++h) {
        const C<T> a = u1[pf.vert[2 * k + h]], b = u2[pf.vert[2 * k + h]];
        const double s1r = a.x * fvert, s1i = a.y * fvert, s2r = b.x * fvert, s2i = b.y * fvert;
        ff += s1r * s1r + s1i * s1i + (s2r * s2r + s2i * s2i);
        fv += proj_c(a.x, a.y, s1r, s1i) + proj_c(b.x, b.y, s2r, s2i);
      }
    vp = ff * 0.5 / dt + fv / dt;
  } else {
    for (int sl = 0; sl < pf.nslot; ++sl) pf.slot_factor[sl] = dt * pf.R[sl] * 0.99 / pf.Ef[sl];
    double fv = 0.0;
    for (int i = 0; i < pf.nel; ++i) {
      const double fkn = pf.slot_factor[pf.eslot[i]];
      const C<T> a = u1[pf.elem[i]], b = u2[pf.elem[i]];
      const double umag = sqrt((double)a.x * a.x + (double)a.y * a.y + ((double)b.x * b.x + (double)b.y * b.y));
      pf.plane_factor[pf.elem[i]] = fkn / umag;  // force = (u / |u|) fkn
      const double s1r = a.x / umag * fkn, s1i = a.y / umag * fkn, s2r = b.x / umag * fkn, s2i = b.y / umag * fkn;
      fv += (proj_c(a.x, a.y, s1r / dt, s1i / dt) + proj_c(b.x, b.y, s2r / dt, s2i / dt)) * pf.ew[i];
    }
    hp = fv;
    // vertical: kx = ky = 0, kz = 3..5 and -3..-5
    double umagsq = 0.0;
    for (int k = 0; k < 3; ++k) {
      const C<T> a = u1[pf.vert[2 * k]], b = u2[pf.vert[2 * k]];
      const C<T> c = u1[pf.vert[2 * k + 1]], d = u2[pf.vert[2 * k + 1]];
      umagsq += (double)a.x * a.x + (double)a.y * a.y + ((double)b.x * b.x + (double)b.y * b.y) +
                ((double)c.x * c.x + (double)c.y * c.y) + ((double)d.x * d.x + (double)d.y * d.y);
    }
    fvert = dt * hp * (1 - 0.99) / (umagsq * 0.99);
    for (int k = 0; k < 3; ++k)
      for (int h = 0; h < 2; ++h) {
        const C<T> a = u1[pf.vert[2 * k + h]], b = u2[pf.vert[2 * k + h]];
        vp += proj_c(a.x, a.y, a.x * fvert / dt, a.y * fvert / dt) + proj_c(b.x, b.y, b.x * fvert / dt, b.y * fvert / dt);
      }
  }
  st->force_hp = hp;
  st->force_vp = vp;
  st->force_vert = fvert;
}

// calculate_u1u2_spectrum! (u1u2_mag for QfForcing) over every shell of the kz = 0 plane, the Ef the
// reference keeps (spectrum.jl:1-20, :44-65): one block per shell, a tree sum over its elements
template <typename T>
__global__ void __launch_bounds__(256) k_plane_shells(Geo g, const C<T>* __restrict__ u, PlaneForcing pf) {
  __shared__ double part[256];
  const size_t n = (size_t)g.sx * g.my * g.mz;
  const int kind = pf.kind < 0 ? -pf.kind : pf.kind;
  const int sh = blockIdx.x;
  double e = 0.0;
  for (int jj = pf.shell_ptr[sh] + threadIdx.x; jj < pf.shell_ptr[sh + 1]; jj += blockDim.x) {
    const int idx = pf.shell_el[jj];
    const double w = idx % g.sx == 0 ? 1.0 : 2.0;
    const C<T> a = u[idx], b = u[n + idx];
    const double m2 = (double)a.x * a.x + (double)a.y * a.y + ((double)b.x * b.x + (double)b.y * b.y);
    e += kind == PF_QF ? w * sqrt(m2) : 0.5 * w * m2;
  }
  part[threadIdx.x] = e;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) pf.Ef_all[sh] = part[0];
}

template <typename T>
void launch_plane_forcing(const Geo& g, const void* u, const void* rhs, DevState* st, const PlaneForcing& pf,
                          const UpdateArgs& ua, int rhs_final, int it, hipStream_t s) {
  k_plane_forcing<T><<<1, 64, 0, s>>>(g, (const C<T>*)u, (const C<T>*)rhs, st, pf, ua, rhs_final, it);
  if (pf.nshell > 0) k_plane_shells<T><<<pf.nshell, 256, 0, s>>>(g, (const C<T>*)u, pf);
}

// one block of 64 threads: the forcing band sum (as k_forcing_sum) and, on thread 0, set_dt
template <typename T> struct SCall {
  Geo g;
  const C<T>* u;
  DevState* st;
  ScalarArgs a;
};
template <typename T>
__device__ __forceinline__ void step_scalars_body(Geo g, const C<T>* __restrict__ u, DevState* st, ScalarArgs a) {
  g = glb(g);
  a.band = glb(a.band);
  if (a.nband > 0) {
    const size_t n = (size_t)g.sx * g.my * g.mz;
    double acc = 0;
    for (int b = threadIdx.x; b < a.nband; b += blockDim.x) {
      const int idx = a.band[b];
      const double w = (idx % g.sx) > 0 ? 2.0 : 1.0;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const C<T> v = u[c * n + idx];
        acc += w * ((double)v.x * v.x + (double)v.y * v.y);
      }
    }
    __shared__ double red[64];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int off = 32; off > 0; off >>= 1) {
      if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
      __syncthreads();
    }
    if (threadIdx.x == 0) st->force_factor = red[0] > 0 ? a.power / red[0] : 0.0;
  }
  if (a.do_dt && threadIdx.x == 0) set_dt_body(st, a.it, a.tit, a.cfl, a.dx, a.etd, a.gnorm);
}
template <typename T>
__global__ void __launch_bounds__(64) k_step_scalars(Geo g, const C<T>* __restrict__ u, DevState* st, ScalarArgs a) {
  step_scalars_body<T>(g, u, st, a);
}
template <typename T>
__global__ void __launch_bounds__(64) k_step_scalars_b(FFS_CALLS(SCall<T>) calls, int nrun) {
  const SCall<T> c = load_call(calls, blockIdx.z);
  step_scalars_body<T>(c.g, glb(c.u), glb(c.st), c.a);
}

template <typename T>
void launch_step_scalars(const Geo& g, const void* u, DevState* st, const ScalarArgs& a, hipStream_t s) {
  const SCall<T> c{g, (const C<T>*)u, st, a};
  launch_call((const void*)&k_step_scalars_b<T>, dim3(1), dim3(64), 0, c, "step_scalars",
              [&] { k_step_scalars<T><<<1, 64, 0, s>>>(c.g, c.u, c.st, c.a); });
}

#if FFS_IN_PART(0)
void launch_set_dt(DevState* st, int it, int tit, double cfl, double dx, int etd, double gnorm, hipStream_t s) {
  k_set_dt<<<1, 1, 0, s>>>(st, it, tit, cfl, dx, etd, gnorm);
}

void host_step_coefficients(DevState* st) {
  ab3_coefficients(st);
  etd_scalars(st);
}

__global__ void k_end_step(DevState* st) { st->time += st->dt; }
void launch_end_step(DevState* st, hipStream_t s) { k_end_step<<<1, 1, 0, s>>>(st); }
#endif

template <typename S, typename D>
__global__ void k_cast_copy(const S* __restrict__ a, D* __restrict__ b, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (D)a[i];
}


// ===========================================================================
// Seeded random-phase initial fields, generated in place in the compact
// spectral layout.  Mode (kx, ky, kz) (signed integer indices) of component c
// draws one complex Gaussian xi from a counter-based generator: two SplitMix64
// hashes of (seed, c, kx, ky, kz) -> two uniforms in (0,1) -> Box-Muller.  The
// kx = 0 plane takes xi(0,ky,kz) = conj(xi(0,-ky,-kz)) for the non-canonical
// half (ky < 0, or ky = 0 and kz < 0) so the field is real.  Velocity:
// u_hat = A(|k|) (I - k k / |k|^2) xi, A = |k| exp(-(|k|/k0)^2) / sqrt(4 pi)
// (shell energy ~ k^4 exp(-2 (k/k0)^2)); scalar: A(|k|) xi.  The k = 0 mode is
// zero.  A second kernel scales the result to the requested energy / rms from
// deterministic block sums of w |u_hat|^2 (w = 2 for kx > 0).

__device__ __forceinline__ unsigned long long seed_mix64(unsigned long long z) {
  z += 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

__device__ __forceinline__ void seed_gauss(unsigned long long seed, int c, int ix, int iy, int iz, double& re,
                                           double& im) {
  const unsigned long long ctr = ((((unsigned long long)(c + 1) << 16 | (unsigned long long)(ix & 0xffff)) << 16 |
                                   (unsigned long long)(iy & 0xffff))
                                  << 16) |
                                 (unsigned long long)(iz & 0xffff);
  const unsigned long long s = seed_mix64(seed);
  const unsigned long long h1 = seed_mix64(s ^ seed_mix64(2 * ctr));
  const unsigned long long h2 = seed_mix64(s ^ seed_mix64(2 * ctr + 1));
  const double u1 = ((double)(h1 >> 11) + 0.5) * 0x1.0p-53;
  const double u2 = ((double)(h2 >> 11) + 0.5) * 0x1.0p-53;
  const double r = sqrt(-2.0 * log(u1));
  const double th = 6.283185307179586 * u2;
  re = r * cos(th);
  im = r * sin(th);
}

__device__ __forceinline__ int signed_index(int j, int n) { return j < (n + 1) / 2 ? j : j - n; }

// nf = 3: solenoidal velocity into comps 0..2 of `out`; nf = 1: scalar (generator component 3)
template <typename T>
__global__ void __launch_bounds__(256) k_seed(Geo g, C<T>* __restrict__ out, int nf, unsigned long long seed,
                                              double k0, double* __restrict__ partials) {
  const size_t n = (size_t)g.sx * g.my * g.mz;
  double acc = 0.0;
  for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < n; idx += (size_t)gridDim.x * blockDim.x) {
    const int cx = idx % g.sx;
    const int cy = (idx / g.sx) % g.my;
    const int cz = idx / ((size_t)g.sx * g.my);
    if (cx >= g.mx) {
      for (int c = 0; c < nf; ++c) out[c * n + idx] = mk<T>(0, 0);
      continue;
    }
    const double kx = g.kx[cx], ky = g.ky[cy], kz = g.kz[cz];
    const double K2 = kx * kx + ky * ky + kz * kz;
    int iy = signed_index(y_fft(g, cy), g.ny), iz = signed_index(z_fft(g, cz), g.nz);
    bool conj = false;
    if (cx == 0 && (iy < 0 || (iy == 0 && iz < 0))) {
      iy = -iy;
      iz = -iz;
      conj = true;
    }
    double xr[3] = {0, 0, 0}, xi[3] = {0, 0, 0};
    if (K2 > 0) {
      for (int c = 0; c < nf; ++c) {
        seed_gauss(seed, nf == 3 ? c : 3, cx, iy, iz, xr[c], xi[c]);
        if (conj) xi[c] = -xi[c];
      }
    }
    const double kk = sqrt(K2);
    const double amp = K2 > 0 ? kk * exp(-(kk / k0) * (kk / k0)) / sqrt(4.0 * 3.141592653589793) : 0.0;
    if (nf == 3 && K2 > 0) {
      const double pr = (kx * xr[0] + ky * xr[1] + kz * xr[2]) / K2;
      const double pi = (kx * xi[0] + ky * xi[1] + kz * xi[2]) / K2;
      const double kv[3] = {kx, ky, kz};
      for (int c = 0; c < 3; ++c) {
        xr[c] -= pr * kv[c];
        xi[c] -= pi * kv[c];
      }
    }
    const double w = cx > 0 ? 2.0 : 1.0;
    for (int c = 0; c < nf; ++c) {
      const double a = amp * xr[c], b = amp * xi[c];
      acc += w * (a * a + b * b);
      out[c * n + idx] = mk<T>((T)a, (T)b);
    }
  }
  __shared__ double red[256];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) partials[(size_t)blockIdx.x * kStatsSlots] = red[0];
}

// out *= sqrt(target / (norm * sum)), sum = sum of the seed kernel's partials (read on the device)
template <typename T>
__global__ void k_seed_scale(C<T>* __restrict__ out, size_t n, const double* __restrict__ sum, double target,
                             double norm) {
  const double f = sqrt(target / (norm * sum[0]));
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const C<T> v = out[i];
    out[i] = mk<T>((T)(f * v.x), (T)(f * v.y));
  }
}

template <typename T>
void launch_seed(const Geo& g, void* out, int nf, unsigned long long seed, double k0, double target, double* partials,
                 double* red, hipStream_t s) {
  const size_t n = (size_t)g.sx * g.my * g.mz;
  const int nblk = (int)std::min<size_t>((n + 255) / 256, 512);
  k_seed<T><<<nblk, 256, 0, s>>>(g, (C<T>*)out, nf, seed, k0, partials);
  k_reduce_partials<<<1, 256, 0, s>>>(partials, nblk, 1, red);
  // velocity: energy = 1/2 sum w |u|^2; scalar: rms^2 = sum w |rho|^2
  k_seed_scale<T><<<1024, 256, 0, s>>>((C<T>*)out, nf * n, red, target * (nf == 3 ? 1.0 : target),
                                       nf == 3 ? 0.5 : 1.0);
}

// copies n reals from double to T
template <typename T>
void launch_cast_copy(const void* src, void* dst, size_t n, hipStream_t s) {
  k_cast_copy<double, T><<<(n + 255) / 256, 256, 0, s>>>((const double*)src, (T*)dst, n);
}

// ---- explicit instantiations, by part ----
#define FFS_INST_INV_Z(T)                                                                                         \
  template void launch_inv_z<T>(const Geo&, const void*, void*, int, bool, const void*, hipStream_t, const void*,  \
                                const void*, int, const void*, const void*);
#define FFS_INST_COLS(T)                                                                                            \
  template void launch_inv_y<T>(const Geo&, const void*, void*, int, const void*, hipStream_t, int, int);         \
  template void launch_fwd_y<T>(const Geo&, const void*, void*, int, const void*, hipStream_t, int, int);         \
  template void launch_fwd_z<T>(const Geo&, const void*, void* const*, int, int, double, const void*, hipStream_t);
#define FFS_INST_X(T)                                                                                               \
  template void launch_x_nl<T>(const Geo&, const void*, void*, const XOps&, unsigned long long*, unsigned long long*, \
                               const void*, hipStream_t);
#define FFS_INST_XIO(T)                                                                                             \
  template void launch_x_c2r<T>(const Geo&, const void*, T*, int, const void*, hipStream_t);                          \
  template void launch_x_c2r_rows<T>(const Geo&, const void*, double*, int, int, const void*, hipStream_t);         \
  template void launch_x_r2c<T>(const Geo&, const T*, void*, int, const void*, hipStream_t);
#define FFS_INST_FZU(T)                                                                                             \
  template void launch_fwd_z_update<T>(const Geo&, const void*, void*, void*, const void*, const void*,           \
                                       const DevState*, const UpdateArgs&, double, const void*, hipStream_t);
#define FFS_INST_REST(T)                                                                                              \
  template void launch_update<T>(const Geo&, void*, void*, const void*, const void*, const DevState*,                 \
                                 const UpdateArgs&, hipStream_t);                                                     \
  template void launch_project<T>(const Geo&, void*, hipStream_t);                                                   \
  template void launch_hermitian_planes<T>(const Geo&, void*, void*, void*, hipStream_t);                           \
  template void launch_div_uu<T>(const Geo&, const void*, void*, hipStream_t);                                     \
  template void launch_les_tau_real<T>(T*, size_t, double, hipStream_t);                                          \
  template void launch_les_model_real<T>(T*, const T*, size_t, const LesArgs&, hipStream_t);                      \
  template void launch_les_scalar_flux_real<T>(const T*, const T*, const T*, T*, size_t, const LesArgs&, double,   \
                                               int, double, hipStream_t);                                          \
  template void launch_dyn_prep<T>(const Geo&, const void*, void*, double, hipStream_t, int);                      \
  template void launch_dyn_products<T>(const T*, const T*, T*, T*, size_t, double, hipStream_t, const T*,          \
                                       const DynP&);                                                               \
  template void launch_gauss<T>(const Geo&, void*, double, hipStream_t);                                           \
  template void launch_dyn_coef<T>(const T*, const T*, const T*, const T*, T*, size_t, double, hipStream_t, int,   \
                                   const DynP&);                                                                   \
  template void launch_piomelli_cstar<T>(T*, T*, const DevState*, double*, size_t, hipStream_t);                  \
  template void launch_fill_real<T>(T*, size_t, double, hipStream_t);                                             \
  template void launch_dyn_tau<T>(const T*, const T*, T*, size_t, double, double, hipStream_t, const T*, const T*); \
  template int launch_stats<T>(const Geo&, const void*, const void*, const void*, const StatsAux&, double*, double*,  \
                               hipStream_t);                                                                       \
  template void launch_spectrum<T>(const Geo&, const void*, double*, double*, double*, double*, double*, double*,     \
                                   hipStream_t);                                                                      \
  template void launch_spectra<T>(const Geo&, const void*, const SpecArgs&, double*, double*, hipStream_t);       \
  template void launch_forcing_sum<T>(const Geo&, const void*, const int*, int, DevState*, double, hipStream_t);  \
  template void launch_cast_copy<T>(const void*, void*, size_t, hipStream_t);                                       \
  template void launch_step_scalars<T>(const Geo&, const void*, DevState*, const ScalarArgs&, hipStream_t);          \
  template void launch_plane_forcing<T>(const Geo&, const void*, const void*, DevState*, const PlaneForcing&,          \
                                        const UpdateArgs&, int, int, hipStream_t);                                    \
  template void launch_seed<T>(const Geo&, void*, int, unsigned long long, double, double, double*, double*,           \
                               hipStream_t);

#if FFS_IN_PART(0)
FFS_INST_REST(double)
FFS_INST_REST(float)
FFS_INST_XIO(double)
FFS_INST_XIO(float)
#endif
#if FFS_IN_PART(1)
FFS_INST_INV_Z(double)
#endif
#if FFS_IN_PART(2)
FFS_INST_INV_Z(float)
#endif
#if FFS_IN_PART(3)
FFS_INST_COLS(double)
#endif
#if FFS_IN_PART(4)
FFS_INST_COLS(float)
#endif
#if FFS_IN_PART(5)
FFS_INST_X(double)
#endif
#if FFS_IN_PART(6)
FFS_INST_X(float)
#endif
#if FFS_IN_PART(7)
FFS_INST_FZU(double)
FFS_INST_FZU(float)
#endif

}  // namespace fsk
