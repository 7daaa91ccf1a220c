// Device-side FFT building blocks for gfx950 (CDNA4, wave64).
//
// Everything here is written for the batched pencil transforms of a
// triply periodic pseudo-spectral solver.  A length-N complex transform is
// carried out by TPER cooperating threads as a Stockham autosort sequence of
// radix-R stages; the first stage reads its operands through a caller
// supplied functor (global memory, possibly with a fused pointwise prologue)
// and the last stage writes through another functor (global memory, possibly
// with a fused epilogue), so a two-stage plan touches LDS exactly once.
//
// Sign convention: DIR = -1 is the forward transform X[k] = sum x[n] e^{-2 pi i nk/N}
// (unnormalised, FFTW_FORWARD), DIR = +1 the backward one (unnormalised,
// FFTW_BACKWARD), which is what FluidFields' rfft!/real! call
// (utils.jl:101-112, models.jl:122-211).
//
// Rounding: the arithmetic below is compiled without FMA contraction, and every fused multiply-add is
// written out (fmaT), so a transform rounds the same way in every kernel that inlines it, whatever
// the compiler could contract around it there.  Two kernels that transform the same data therefore
// agree bit for bit (a fused pass and its unfused pair, a batched launch and a single one).
#pragma once
#include <hip/hip_runtime.h>

#pragma clang fp contract(off)

namespace fsk {

template <typename T> struct cvec;
template <> struct cvec<double> { using type = double2; };
template <> struct cvec<float> { using type = float2; };
template <typename T> using C = typename cvec<T>::type;

template <typename T> __device__ __forceinline__ C<T> mk(T x, T y) { C<T> r; r.x = x; r.y = y; return r; }
template <typename T> __device__ __forceinline__ T fmaT(T a, T b, T c);
template <> __device__ __forceinline__ double fmaT<double>(double a, double b, double c) { return __builtin_fma(a, b, c); }
template <> __device__ __forceinline__ float fmaT<float>(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
template <typename T> __device__ __forceinline__ C<T> cadd(C<T> a, C<T> b) { return mk<T>(a.x + b.x, a.y + b.y); }
template <typename T> __device__ __forceinline__ C<T> csub(C<T> a, C<T> b) { return mk<T>(a.x - b.x, a.y - b.y); }
template <typename T> __device__ __forceinline__ C<T> cmul(C<T> a, C<T> b) {
  return mk<T>(fmaT<T>(a.x, b.x, -(a.y * b.y)), fmaT<T>(a.x, b.y, a.y * b.x));
}
// a * conj(b)
template <typename T> __device__ __forceinline__ C<T> cmulc(C<T> a, C<T> b) {
  return mk<T>(fmaT<T>(a.x, b.x, a.y * b.y), fmaT<T>(a.y, b.x, -(a.x * b.y)));
}
template <typename T> __device__ __forceinline__ C<T> cconj(C<T> a) { return mk<T>(a.x, -a.y); }
template <typename T> __device__ __forceinline__ C<T> cscale(C<T> a, T s) { return mk<T>(a.x * s, a.y * s); }
// multiply by S*i (S = +1 or -1)
template <int S, typename T> __device__ __forceinline__ C<T> mul_si(C<T> a) {
  return S > 0 ? mk<T>(-a.y, a.x) : mk<T>(a.y, -a.x);
}

// cos/sin(2 pi m / 16), m = 0..15 (exact decimal expansions)
__host__ __device__ constexpr double cos16(int m) {
  constexpr double c[16] = {1.0,
                            0.92387953251128675613,
                            0.70710678118654752440,
                            0.38268343236508977173,
                            0.0,
                            -0.38268343236508977173,
                            -0.70710678118654752440,
                            -0.92387953251128675613,
                            -1.0,
                            -0.92387953251128675613,
                            -0.70710678118654752440,
                            -0.38268343236508977173,
                            0.0,
                            0.38268343236508977173,
                            0.70710678118654752440,
                            0.92387953251128675613};
  return c[m & 15];
}
__host__ __device__ constexpr double sin16(int m) { return cos16(m - 4); }

// multiply by W_16^m = cos(2 pi m/16) + S i sin(2 pi m/16)   (S = DIR)
template <int S, int M, typename T> __device__ __forceinline__ C<T> tw16(C<T> a) {
  constexpr int m = M & 15;
  if constexpr (m == 0) {
    return a;
  } else if constexpr (m == 4) {
    return mul_si<S, T>(a);
  } else if constexpr (m == 8) {
    return mk<T>(-a.x, -a.y);
  } else if constexpr (m == 12) {
    return mul_si<-S, T>(a);
  } else {
    const T c = (T)cos16(m);
    const T s = (T)(S * sin16(m));
    return mk<T>(fmaT<T>(a.x, c, -(a.y * s)), fmaT<T>(a.x, s, a.y * c));
  }
}

// ---------------------------------------------------------------------------
// In-register DFTs of size R with sign S.
template <int S, typename T> __device__ __forceinline__ void dft2(C<T>& a0, C<T>& a1) {
  C<T> t = a0;
  a0 = cadd<T>(t, a1);
  a1 = csub<T>(t, a1);
}

template <int S, typename T> __device__ __forceinline__ void dft4(C<T>& a0, C<T>& a1, C<T>& a2, C<T>& a3) {
  C<T> t0 = cadd<T>(a0, a2), t1 = csub<T>(a0, a2);
  C<T> t2 = cadd<T>(a1, a3), t3 = mul_si<S, T>(csub<T>(a1, a3));
  a0 = cadd<T>(t0, t2);
  a2 = csub<T>(t0, t2);
  a1 = cadd<T>(t1, t3);
  a3 = csub<T>(t1, t3);
}

template <int S, typename T> __device__ __forceinline__ void dft3(C<T>& a0, C<T>& a1, C<T>& a2) {
  const T h = (T)(S * 0.86602540378443864676);  // S sin(2 pi / 3)
  C<T> s = cadd<T>(a1, a2);
  C<T> e = csub<T>(a1, a2);  // a1, a2 = m +- S i h e
  C<T> m = mk<T>(fmaT<T>((T)-0.5, s.x, a0.x), fmaT<T>((T)-0.5, s.y, a0.y));
  a0 = cadd<T>(a0, s);
  a1 = mk<T>(fmaT<T>(-h, e.y, m.x), fmaT<T>(h, e.x, m.y));
  a2 = mk<T>(fmaT<T>(h, e.y, m.x), fmaT<T>(-h, e.x, m.y));
}

template <int S, typename T>
__device__ __forceinline__ void dft5(C<T>& a0, C<T>& a1, C<T>& a2, C<T>& a3, C<T>& a4) {
  const T c1 = (T)0.30901699437494742410, c2 = (T)-0.80901699437494742410;
  const T s1 = (T)0.95105651629515357212, s2 = (T)0.58778525229247312917;
  C<T> b1 = cadd<T>(a1, a4), b2 = cadd<T>(a2, a3);
  C<T> d1 = csub<T>(a1, a4), d2 = csub<T>(a2, a3);
  C<T> r1 = mk<T>(fmaT<T>(c2, b2.x, fmaT<T>(c1, b1.x, a0.x)), fmaT<T>(c2, b2.y, fmaT<T>(c1, b1.y, a0.y)));
  C<T> r2 = mk<T>(fmaT<T>(c1, b2.x, fmaT<T>(c2, b1.x, a0.x)), fmaT<T>(c1, b2.y, fmaT<T>(c2, b1.y, a0.y)));
  C<T> i1 = mul_si<S, T>(mk<T>(fmaT<T>(s2, d2.x, s1 * d1.x), fmaT<T>(s2, d2.y, s1 * d1.y)));
  C<T> i2 = mul_si<S, T>(mk<T>(fmaT<T>(-s1, d2.x, s2 * d1.x), fmaT<T>(-s1, d2.y, s2 * d1.y)));
  a0 = mk<T>(a0.x + b1.x + b2.x, a0.y + b1.y + b2.y);
  a1 = cadd<T>(r1, i1);
  a4 = csub<T>(r1, i1);
  a2 = cadd<T>(r2, i2);
  a3 = csub<T>(r2, i2);
}

// size 8 = 2 x 4 (n = 4 n1 + n2, k = k1 + 2 k2)
template <int S, typename T> __device__ __forceinline__ void dft8(C<T>* a) {
  C<T> y0[4], y1[4];
#pragma unroll
  for (int n2 = 0; n2 < 4; ++n2) {
    y0[n2] = cadd<T>(a[n2], a[n2 + 4]);
    y1[n2] = csub<T>(a[n2], a[n2 + 4]);
  }
  y1[1] = tw16<S, 2, T>(y1[1]);
  y1[2] = tw16<S, 4, T>(y1[2]);
  y1[3] = tw16<S, 6, T>(y1[3]);
  dft4<S, T>(y0[0], y0[1], y0[2], y0[3]);
  dft4<S, T>(y1[0], y1[1], y1[2], y1[3]);
#pragma unroll
  for (int k2 = 0; k2 < 4; ++k2) {
    a[2 * k2] = y0[k2];
    a[2 * k2 + 1] = y1[k2];
  }
}

// size 16 = 4 x 4 (n = 4 n1 + n2, k = k1 + 4 k2)
template <int S, typename T> __device__ __forceinline__ void dft16(C<T>* a) {
  C<T> y[4][4];  // y[n2][k1]
#pragma unroll
  for (int n2 = 0; n2 < 4; ++n2) {
    y[n2][0] = a[n2];
    y[n2][1] = a[n2 + 4];
    y[n2][2] = a[n2 + 8];
    y[n2][3] = a[n2 + 12];
    dft4<S, T>(y[n2][0], y[n2][1], y[n2][2], y[n2][3]);
  }
  y[1][1] = tw16<S, 1, T>(y[1][1]);
  y[1][2] = tw16<S, 2, T>(y[1][2]);
  y[1][3] = tw16<S, 3, T>(y[1][3]);
  y[2][1] = tw16<S, 2, T>(y[2][1]);
  y[2][2] = tw16<S, 4, T>(y[2][2]);
  y[2][3] = tw16<S, 6, T>(y[2][3]);
  y[3][1] = tw16<S, 3, T>(y[3][1]);
  y[3][2] = tw16<S, 6, T>(y[3][2]);
  y[3][3] = tw16<S, 9, T>(y[3][3]);
#pragma unroll
  for (int k1 = 0; k1 < 4; ++k1) {
    dft4<S, T>(y[0][k1], y[1][k1], y[2][k1], y[3][k1]);
    a[k1] = y[0][k1];
    a[k1 + 4] = y[1][k1];
    a[k1 + 8] = y[2][k1];
    a[k1 + 12] = y[3][k1];
  }
}

template <int R, int S, typename T> __device__ __forceinline__ void dftR(C<T>* a) {
  if constexpr (R == 2) dft2<S, T>(a[0], a[1]);
  else if constexpr (R == 3) dft3<S, T>(a[0], a[1], a[2]);
  else if constexpr (R == 4) dft4<S, T>(a[0], a[1], a[2], a[3]);
  else if constexpr (R == 5) dft5<S, T>(a[0], a[1], a[2], a[3], a[4]);
  else if constexpr (R == 8) dft8<S, T>(a);
  else if constexpr (R == 16) dft16<S, T>(a);
  else static_assert(R == 2, "unsupported radix");
}

// ---------------------------------------------------------------------------
// Radix plans.  N = 2^a 3^b 5^c; powers of two are grouped into radix-MAXR
// stages (MAXR = 16 or 8; a remainder of one turns a MAXR*2 into (MAXR/2)*4),
// then 5s and 3s.
struct Plan {
  int n;
  int ns;  // number of stages
  int r[10];
};

__host__ __device__ constexpr Plan make_plan(int N, int maxr = 16) {
  Plan p{N, 0, {0, 0, 0, 0, 0, 0, 0, 0, 0, 0}};
  int L = 0;
  while ((1 << (L + 1)) <= maxr) ++L;  // log2(maxr)
  int m = N, a = 0;
  while (m % 2 == 0) { m /= 2; ++a; }
  int nbig = a / L, rem = a % L;
  if (rem == 1 && nbig > 0) {
    --nbig;
    p.r[p.ns++] = 1 << (L - 1);
    p.r[p.ns++] = 4;
  } else if (rem > 0) {
    p.r[p.ns++] = 1 << rem;
  }
  for (int i = 0; i < nbig; ++i) p.r[p.ns++] = 1 << L;
  while (m % 5 == 0) { m /= 5; p.r[p.ns++] = 5; }
  while (m % 3 == 0) { m /= 3; p.r[p.ns++] = 3; }
  if (m != 1) p.ns = -1;  // unsupported length
  return p;
}

__host__ __device__ constexpr int plan_ns(int N, int maxr = 16) { return make_plan(N, maxr).ns; }
// REV: the plan's stages in reverse order (the small radix last)
__host__ __device__ constexpr int plan_r(int N, int s, int maxr = 16, bool rev = false) {
  return make_plan(N, maxr).r[rev ? make_plan(N, maxr).ns - 1 - s : s];
}
__host__ __device__ constexpr int plan_nsprev(int N, int s, int maxr = 16, bool rev = false) {
  int v = 1;
  for (int i = 0; i < s; ++i) v *= plan_r(N, i, maxr, rev);
  return v;
}
// points handled per thread: 16 if it divides N, else 8, 12, ... (always >= the largest radix of
// the plan's mixed part, so every stage fits in registers)
__host__ __device__ constexpr int default_ppt(int N) {
  return (N % 16 == 0) ? 16 : (N % 8 == 0) ? 8 : (N % 12 == 0) ? 12 : (N % 6 == 0) ? 6 : (N % 5 == 0) ? 5 : (N % 4 == 0) ? 4 : (N % 3 == 0) ? 3 : 2;
}

// ---------------------------------------------------------------------------
// One Stockham stage for all butterflies owned by thread q (of TPER).
//   stage s: radix R, Ns = product of the previous radices
//   butterfly b: inputs x[b + r N/R], twiddle W_{Ns R}^{r (b mod Ns)}, outputs
//   y[(b/Ns) Ns R + (b mod Ns) + r Ns]
// Load/Store functors receive (position, t, r): t indexes the thread's
// butterflies, r the radix leg; both are compile-time after unrolling, which
// lets register-resident callers index register arrays with them.
#ifndef FFS_TWIDDLE_POWERS
#define FFS_TWIDDLE_POWERS 1
#endif
// TWIDDLE_POWERS: derive W^r (r >= 2) from W^1 instead of reading every leg's twiddle
//
// MIR: the stage's 2 TPER butterflies are dealt to the threads in mirror pairs, thread q owning b = q
// and b = NBF - q (thread 0: 0 and NBF/2).  A first stage (NS = 1) reads positions b + r NBF, a last
// stage (NS = NBF) writes them, and the mirror of b + r NBF is (NBF - b) + (R - 1 - r) NBF: each
// thread then holds x[p] and x[N - p] together, which is what a pair of real rows packed into one
// complex sequence needs to be built from, or split into, its half spectra.
template <typename T, int N, int DIR, int STAGE, int TPER, int MAXR = 16, bool TWIDDLE_POWERS = FFS_TWIDDLE_POWERS,
          bool REV = false, bool MIR = false>
struct Stage {
  static constexpr int R = plan_r(N, STAGE, MAXR, REV);
  static constexpr int NS = plan_nsprev(N, STAGE, MAXR, REV);
  static constexpr int NBF = N / R;                       // butterflies
  static constexpr int NB = (NBF + TPER - 1) / TPER;      // per thread
  static_assert(!MIR || (NB == 2 && NBF == 2 * TPER && (NS == 1 || NS == NBF)),
                "mirror-paired butterflies: two per thread, in a first or last stage");

  static __device__ __forceinline__ int bfly(int q, int t) {
    if constexpr (MIR) return t == 0 ? q : (q == 0 ? NBF / 2 : NBF - q);
    else return q + t * TPER;
  }

  template <class Load>
  static __device__ __forceinline__ void load(C<T> (&v)[NB][R], int q, const Load& ld) {
#pragma unroll
    for (int t = 0; t < NB; ++t) {
      const int b = bfly(q, t);
      if (NB * TPER == NBF || b < NBF) {
#pragma unroll
        for (int r = 0; r < R; ++r) v[t][r] = ld(b + r * NBF, t, r);
      }
    }
  }

  static __device__ __forceinline__ void compute(C<T> (&v)[NB][R], int q, const C<T>* __restrict__ tw) {
#pragma unroll
    for (int t = 0; t < NB; ++t) {
      const int b = bfly(q, t);
      if (NB * TPER == NBF || b < NBF) {
        if constexpr (NS > 1) {
          const int j = b % NS;
          if constexpr (TWIDDLE_POWERS) {
            // one table read per butterfly; W^r by multiplication (at most 3 rounding steps deep)
            const C<T> w1 = tw[j * (N / (NS * R))];
            C<T> w[R];
            w[1] = w1;
#pragma unroll
            for (int r = 2; r < R; ++r) w[r] = (r & 1) ? cmul<T>(w[r - 1], w1) : cmul<T>(w[r / 2], w[r / 2]);
#pragma unroll
            for (int r = 1; r < R; ++r) v[t][r] = (DIR < 0) ? cmul<T>(v[t][r], w[r]) : cmulc<T>(v[t][r], w[r]);
          } else {
#pragma unroll
            for (int r = 1; r < R; ++r) {
              const C<T> w = tw[(r * j) * (N / (NS * R))];
              v[t][r] = (DIR < 0) ? cmul<T>(v[t][r], w) : cmulc<T>(v[t][r], w);
            }
          }
        }
        dftR<R, DIR, T>(v[t]);
      }
    }
  }

  template <class Store>
  static __device__ __forceinline__ void store(C<T> (&v)[NB][R], int q, const Store& st) {
#pragma unroll
    for (int t = 0; t < NB; ++t) {
      const int b = bfly(q, t);
      if (NB * TPER == NBF || b < NBF) {
        const int base = (b / NS) * NS * R + (b % NS);
#pragma unroll
        for (int r = 0; r < R; ++r) st(base + r * NS, t, r, v[t][r]);
      }
    }
  }
};

// Exchange synchronisation: the whole block, or only the wavefront (when all
// TPER threads of a transform live in one wavefront).
enum { SYNC_BLOCK = 0, SYNC_WAVE = 1 };
template <int SYNC> __device__ __forceinline__ void fft_sync() {
  if constexpr (SYNC == SYNC_BLOCK) __syncthreads();
  else __syncwarp();
}

// Runs stages [STAGE, ns) with stage STAGE reading through `ld` and the last
// stage writing through `st`; intermediate exchanges go through LDS via
// `lds(pos) -> C<T>&`.
//   IN_LDS : `ld` reads the same LDS region as `lds` (in-place transform), so
//            every thread must finish its first loads before anyone stores.
//   OUT_LDS: `st` writes the same LDS region, so the last stage must wait for
//            every thread's loads before storing.
// The caller synchronises before the call if the LDS region was in use by a
// previous transform, and after it before reading OUT_LDS results.  Every
// thread of the synchronisation scope must call run().
// REV runs the plan's stages in reverse order; stage MIRS (if >= 0) deals its butterflies in mirror
// pairs (see Stage).
template <typename T, int N, int DIR, int TPER, bool IN_LDS = false, bool OUT_LDS = false, int SYNC = SYNC_BLOCK,
          int MAXR = 16, int STAGE = 0, bool REV = false, int MIRS = -1>
struct Fft {
  static constexpr int NSTG = plan_ns(N, MAXR);
  static_assert(NSTG > 0, "FFT length must be 2^a 3^b 5^c");

  template <class Load, class Store, class Lds>
  static __device__ __forceinline__ void run(int q, const C<T>* __restrict__ tw, const Load& ld, const Store& st,
                                             const Lds& lds) {
    using S = Stage<T, N, DIR, STAGE, TPER, MAXR, FFS_TWIDDLE_POWERS, REV, STAGE == MIRS>;
    C<T> v[S::NB][S::R];
    S::load(v, q, ld);
    S::compute(v, q, tw);
    constexpr bool reads_lds = (STAGE > 0) || IN_LDS;
    if constexpr (STAGE + 1 == NSTG) {
      if constexpr (OUT_LDS && reads_lds) fft_sync<SYNC>();
      S::store(v, q, st);
    } else {
      if constexpr (reads_lds) fft_sync<SYNC>();  // everyone has read the previous contents
      S::store(v, q, [&](int p, int, int, C<T> x) { lds(p) = x; });
      fft_sync<SYNC>();
      Fft<T, N, DIR, TPER, IN_LDS, OUT_LDS, SYNC, MAXR, STAGE + 1, REV, MIRS>::run(
          q, tw, [&](int p, int, int) { return lds(p); }, st, lds);
    }
  }
};

// position sets of the first and the last stage coincide ({p = q mod TPER})
// when every stage's butterflies divide evenly over the TPER threads; register
// resident callers rely on it and check it with this predicate.
__host__ __device__ constexpr bool plan_even(int N, int tper, int maxr = 16) {
  for (int s = 0; s < plan_ns(N, maxr); ++s)
    if ((N / plan_r(N, s, maxr)) % tper != 0) return false;
  return true;
}

}  // namespace fsk

#pragma clang fp contract(fast)
