// Batched Stockham FFT through LDS for gfx950.
//
// A work-group holds a tile of L lines of length N in LDS, point-major:
// tile[p * LP + l], with the point stride LP >= L padded against LDS bank
// conflicts. Each radix-r stage (sub-length Ns = product of the earlier
// radices) reads the r inputs j + m*N/r of butterfly j, twiddles them by
// W_{Ns r}^{m k} (k = j mod Ns), applies an r-point DFT and writes them to
// (j / Ns) * Ns * r + k + m * Ns, which leaves the result in natural order after
// the last stage (self-sorting, no bit reversal). Stages run in place: every
// thread loads all of its butterflies' inputs into registers before a barrier,
// then writes. Threads take butterflies b = tid + i * blockDim with
// l = b mod L, so consecutive lanes touch consecutive LDS words.
//
// Forward is exp(-2 pi i jk/N), inverse exp(+...), both unnormalised, as the
// reference's FFT plans (LinAlg/FFTWPlan.cpp:82, lib/apple-opencl-fft clFFT.h:60-64).
// Twiddles come from a table of N roots of unity computed in long double on
// the host, never from device sin/cos.
#pragma once

#include "common.hpp"

namespace dda {

constexpr int kMaxStages = 12;

struct FftPlan {
  int n;
  int nst;
  int radix[kMaxStages];
  int ns[kMaxStages];
  int twmul[kMaxStages];     // n / (ns * radix): stride into the n-point twiddle table
  FastDiv nsdiv[kMaxStages];
};

// cos/sin(2 pi t / r) for the odd radices.
template <int R> struct OddRoots;
template <> struct OddRoots<3> {
  static __device__ inline double c(int t) { return t == 1 ? -0.5 : -0.5; }
  static __device__ inline double s(int t) { return t == 1 ? 0.86602540378443864676 : -0.86602540378443864676; }
};
template <> struct OddRoots<5> {
  static __device__ inline double c(int t) {
    const double v[5] = {1.0, 0.30901699437494742410, -0.80901699437494742410, -0.80901699437494742410,
                         0.30901699437494742410};
    return v[t];
  }
  static __device__ inline double s(int t) {
    const double v[5] = {0.0, 0.95105651629515357212, 0.58778525229247312917, -0.58778525229247312917,
                         -0.95105651629515357212};
    return v[t];
  }
};
template <> struct OddRoots<7> {
  static __device__ inline double c(int t) {
    const double v[7] = {1.0, 0.62348980185873353053, -0.22252093395631440429, -0.90096886790241912624,
                         -0.90096886790241912624, -0.22252093395631440429, 0.62348980185873353053};
    return v[t];
  }
  static __device__ inline double s(int t) {
    const double v[7] = {0.0, 0.78183148246802980871, 0.97492791218182360702, 0.43388373911755812048,
                         -0.43388373911755812048, -0.97492791218182360702, -0.78183148246802980871};
    return v[t];
  }
};

// In-register r-point DFT (forward unless INV).
template <int RAD, bool INV, typename C> struct Dft;

template <bool INV, typename C> struct Dft<2, INV, C> {
  static __device__ inline void run(C* a) {
    C t = a[0];
    a[0] = cadd(t, a[1]);
    a[1] = csub(t, a[1]);
  }
};

template <bool INV, typename C> struct Dft<4, INV, C> {
  static __device__ inline void run(C* a) {
    C t0 = cadd(a[0], a[2]), t1 = csub(a[0], a[2]);
    C t2 = cadd(a[1], a[3]), d = csub(a[1], a[3]);
    C t3 = INV ? cmul_pi(d) : cmul_mi(d);
    a[0] = cadd(t0, t2);
    a[2] = csub(t0, t2);
    a[1] = cadd(t1, t3);
    a[3] = csub(t1, t3);
  }
};

template <bool INV, typename C> struct Dft<8, INV, C> {
  static __device__ inline void run(C* a) {
    typedef decltype(C().x) R;
    C e[4] = {a[0], a[2], a[4], a[6]};
    C o[4] = {a[1], a[3], a[5], a[7]};
    Dft<4, INV, C>::run(e);
    Dft<4, INV, C>::run(o);
    const R h = (R)0.70710678118654752440;
    // W8^k, k = 1..3, for exp(-+ i pi k / 4)
    C w1 = INV ? cmake<C>(h, h) : cmake<C>(h, -h);
    C w3 = INV ? cmake<C>(-h, h) : cmake<C>(-h, -h);
    C o1 = cmul(o[1], w1);
    C o2 = INV ? cmul_pi(o[2]) : cmul_mi(o[2]);
    C o3 = cmul(o[3], w3);
    a[0] = cadd(e[0], o[0]);
    a[4] = csub(e[0], o[0]);
    a[1] = cadd(e[1], o1);
    a[5] = csub(e[1], o1);
    a[2] = cadd(e[2], o2);
    a[6] = csub(e[2], o2);
    a[3] = cadd(e[3], o3);
    a[7] = csub(e[3], o3);
  }
};

// Odd radices by symmetric pairs: with s_m = a_m + a_{r-m}, d_m = a_m - a_{r-m},
// X_k = a0 + sum_m cos(2 pi mk/r) s_m -+ i sin(2 pi mk/r) d_m and X_{r-k} with the
// sign of the sine term flipped.
template <int RAD, bool INV, typename C> struct DftOdd {
  static __device__ inline void run(C* a) {
    typedef decltype(C().x) R;
    constexpr int H = (RAD - 1) / 2;
    C s[H + 1], d[H + 1];
#pragma unroll
    for (int m = 1; m <= H; m++) {
      s[m] = cadd(a[m], a[RAD - m]);
      d[m] = csub(a[m], a[RAD - m]);
    }
    C x0 = a[0];
#pragma unroll
    for (int m = 1; m <= H; m++) x0 = cadd(x0, s[m]);
#pragma unroll
    for (int k = 1; k <= H; k++) {
      C cs = a[0];
      C sn = czero<C>();
#pragma unroll
      for (int m = 1; m <= H; m++) {
        const int t = (m * k) % RAD;
        const R cv = (R)OddRoots<RAD>::c(t);
        const R sv = (R)OddRoots<RAD>::s(t);
        cs.x += cv * s[m].x;
        cs.y += cv * s[m].y;
        sn.x += sv * d[m].x;
        sn.y += sv * d[m].y;
      }
      // forward: X_k = cs - i sn, X_{r-k} = cs + i sn
      C isn = cmul_pi(sn);
      if (INV) {
        a[k] = cadd(cs, isn);
        a[RAD - k] = csub(cs, isn);
      } else {
        a[k] = csub(cs, isn);
        a[RAD - k] = cadd(cs, isn);
      }
    }
    a[0] = x0;
  }
};
template <bool INV, typename C> struct Dft<3, INV, C> : DftOdd<3, INV, C> {};
template <bool INV, typename C> struct Dft<5, INV, C> : DftOdd<5, INV, C> {};
template <bool INV, typename C> struct Dft<7, INV, C> : DftOdd<7, INV, C> {};

// One Stockham stage over `total` = N * L points of the tile. PPT bounds the
// points per thread: total <= PPT * blockDim.x. The twiddles W^(m k) of a
// butterfly come from one table load, W^k, and its powers.
template <int RAD, int PPT, bool INV, typename C>
__device__ inline void fft_stage(C* tile, int total, int nr /* N / RAD */, FastDiv ldiv, int lp, int ns,
                                 FastDiv nsdiv, int twmul, const C* __restrict__ tw) {
  constexpr int NB = (PPT + RAD - 1) / RAD;
  const int nbf = total / RAD;
  const int L = (int)ldiv.d;
  C v[NB][RAD];
#pragma unroll
  for (int i = 0; i < NB; i++) {
    const int b = threadIdx.x + i * blockDim.x;
    if (b < nbf) {
      const int j = ldiv.div(b);
      const int l = b - j * L;
#pragma unroll
      for (int m = 0; m < RAD; m++) v[i][m] = tile[(j + m * nr) * lp + l];
    }
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < NB; i++) {
    const int b = threadIdx.x + i * blockDim.x;
    if (b < nbf) {
      const int j = ldiv.div(b);
      const int l = b - j * L;
      const int q = nsdiv.div(j);
      const int k = j - q * ns;
      if (ns > 1) {
        const C w1 = tw[k * twmul];
        C w = w1;
#pragma unroll
        for (int m = 1; m < RAD; m++) {
          v[i][m] = INV ? cmulc(v[i][m], w) : cmul(v[i][m], w);
          if (m + 1 < RAD) w = cmul(w, w1);
        }
      }
      Dft<RAD, INV, C>::run(v[i]);
      const int base = q * ns * RAD + k;
#pragma unroll
      for (int m = 0; m < RAD; m++) tile[(base + m * ns) * lp + l] = v[i][m];
    }
  }
  __syncthreads();
}

// Full transform of the tile (point stride lp). Callers must __syncthreads()
// after filling it.
template <int PPT, bool INV, typename C>
__device__ inline void fft_tile(C* tile, const FftPlan& p, FastDiv ldiv, int lp, const C* __restrict__ tw) {
  const int total = p.n * (int)ldiv.d;
  for (int s = 0; s < p.nst; s++) {
    const int r = p.radix[s];
    const int nr = p.n / r;
    switch (r) {
      case 8: fft_stage<8, PPT, INV>(tile, total, nr, ldiv, lp, p.ns[s], p.nsdiv[s], p.twmul[s], tw); break;
      case 4: fft_stage<4, PPT, INV>(tile, total, nr, ldiv, lp, p.ns[s], p.nsdiv[s], p.twmul[s], tw); break;
      case 2: fft_stage<2, PPT, INV>(tile, total, nr, ldiv, lp, p.ns[s], p.nsdiv[s], p.twmul[s], tw); break;
      case 3: fft_stage<3, PPT, INV>(tile, total, nr, ldiv, lp, p.ns[s], p.nsdiv[s], p.twmul[s], tw); break;
      case 5: fft_stage<5, PPT, INV>(tile, total, nr, ldiv, lp, p.ns[s], p.nsdiv[s], p.twmul[s], tw); break;
      case 7: fft_stage<7, PPT, INV>(tile, total, nr, ldiv, lp, p.ns[s], p.nsdiv[s], p.twmul[s], tw); break;
      default: break;
    }
  }
}

// Point stride of an LDS tile with L lines: L plus one element when L is even,
// so that consecutive points do not start on the same bank.
__host__ __device__ inline int lds_stride(int L) { return (L & 1) ? L : L + 1; }

}  // namespace dda
