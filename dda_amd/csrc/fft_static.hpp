// Compile-time planned Stockham FFT for the column passes.
//
// A transform of length N = R0 * R1 * ... over a tile of T adjacent columns.
// Every thread owns PPT points of one column in every stage (PPT divisible by
// each radix), so a stage is PPT / R butterflies per thread: butterfly
// j_i = t + i * TPC, TPC = N / PPT threads per column. The first stage takes its
// inputs x[j + m N / R0] from registers the caller filled (straight from HBM),
// the last stage leaves X[j + m N / R_last] in registers, and only the stages
// in between pass through LDS. Running the inverse with the radices reversed
// makes its first stage consume exactly the points the forward transform left
// in a thread's registers, so FFT -> pointwise product -> inverse FFT needs no
// exchange in the middle. Twiddles W_{Ns R}^{m k} come from the host's table
// of N roots of unity (one load per butterfly, then powers); the constants of
// the in-register DFTs are evaluated at compile time.
#pragma once

#include "common.hpp"

namespace dda {

// ------------------------------------------------------------------------------------
// Compile-time cos/sin(2 pi v / M) by exact octant reduction of v/M and a Taylor
// series on [0, pi/4].
// ------------------------------------------------------------------------------------
constexpr double kPiD = 3.14159265358979323846264338327950288;

constexpr double taylor_sin(double x) {
  double term = x, sum = x;
  for (int n = 1; n < 14; n++) {
    term *= -x * x / ((2 * n) * (2 * n + 1));
    sum += term;
  }
  return sum;
}
constexpr double taylor_cos(double x) {
  double term = 1, sum = 1;
  for (int n = 1; n < 14; n++) {
    term *= -x * x / ((2 * n - 1) * (2 * n));
    sum += term;
  }
  return sum;
}
// cos(2 pi t / n) (which = 0) or sin(2 pi t / n) (which = 1)
constexpr double cexpr_trig(long long t, long long n, int which) {
  long long M = 8 * n;
  long long v = 8 * (((t % n) + n) % n);  // angle 2 pi v / M, v in [0, M)
  double sgn_s = 1, sgn_c = 1;
  if (2 * v > M) {  // theta > pi: sin(-), cos(+) with theta -> 2 pi - theta
    v = M - v;
    sgn_s = -1;
  }
  if (4 * v > M) {  // theta > pi/2: theta -> pi - theta, cos(-)
    v = M / 2 - v;
    sgn_c = -sgn_c;
  }
  bool swap = false;
  if (8 * v > M) {  // theta > pi/4: use pi/2 - theta
    v = M / 4 - v;
    swap = true;
  }
  const double x = 2 * kPiD * (double)v / (double)M;
  double s = taylor_sin(x), c = taylor_cos(x);
  if (swap) {
    double tmp = s;
    s = c;
    c = tmp;
  }
  return which ? sgn_s * s : sgn_c * c;
}

// ------------------------------------------------------------------------------------
// In-register DFT of length R (forward exp(-2 pi i jk/R) unless INV).
// Composite lengths R = P * Q by one Cooley-Tukey split with constant twiddles:
// X[k1 + P k2] = sum_q W_Q^{q k2} W_R^{q k1} sum_p a[q + Q p] W_P^{p k1}.
// ------------------------------------------------------------------------------------
template <int R> struct SmallFactor {
  static constexpr int value = (R % 4 == 0 && R > 4) ? 4 : (R % 2 == 0 && R > 2) ? 2 : (R % 3 == 0 && R > 3) ? 3 : R;
};

template <int R, bool INV, typename C> struct RDft;

// a * W_R^k (W_R = exp(-2 pi i / R), conjugated when INV). k and R are
// constants once the calling loops are unrolled, so the branches fold away and
// multiples of a quarter turn become swaps and sign changes.
template <bool INV, typename C>
__device__ __forceinline__ C cmul_w(C a, int k, int R) {
  typedef decltype(C().x) F;
  k = ((k % R) + R) % R;
  if (k == 0) return a;
  if (2 * k == R) return cmake<C>(-a.x, -a.y);
  if (4 * k == R) return INV ? cmul_pi(a) : cmul_mi(a);
  if (4 * k == 3 * R) return INV ? cmul_mi(a) : cmul_pi(a);
  const F c = (F)cexpr_trig(k, R, 0);
  const F s = (F)cexpr_trig(k, R, 1);
  return cmul(a, cmake<C>(c, INV ? s : -s));
}

template <bool INV, typename C> struct RDft<1, INV, C> {
  static __device__ inline void run(C*) {}
};
template <bool INV, typename C> struct RDft<2, INV, C> {
  static __device__ inline void run(C* a) {
    const C t = a[0];
    a[0] = cadd(t, a[1]);
    a[1] = csub(t, a[1]);
  }
};
template <bool INV, typename C> struct RDft<4, INV, C> {
  static __device__ inline void run(C* a) {
    const C t0 = cadd(a[0], a[2]), t1 = csub(a[0], a[2]);
    const C t2 = cadd(a[1], a[3]), d = csub(a[1], a[3]);
    const C t3 = INV ? cmul_pi(d) : cmul_mi(d);
    a[0] = cadd(t0, t2);
    a[2] = csub(t0, t2);
    a[1] = cadd(t1, t3);
    a[3] = csub(t1, t3);
  }
};

// Odd primes by symmetric pairs.
template <int R, bool INV, typename C> struct RDftOdd {
  static __device__ inline void run(C* a) {
    typedef decltype(C().x) F;
    constexpr int H = (R - 1) / 2;
    C s[H + 1], d[H + 1];
#pragma unroll
    for (int m = 1; m <= H; m++) {
      s[m] = cadd(a[m], a[R - m]);
      d[m] = csub(a[m], a[R - m]);
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
        const F cv = (F)cexpr_trig(m * k, R, 0);
        const F sv = (F)cexpr_trig(m * k, R, 1);
        cs.x += cv * s[m].x;
        cs.y += cv * s[m].y;
        sn.x += sv * d[m].x;
        sn.y += sv * d[m].y;
      }
      const C isn = cmul_pi(sn);
      if (INV) {
        a[k] = cadd(cs, isn);
        a[R - k] = csub(cs, isn);
      } else {
        a[k] = csub(cs, isn);
        a[R - k] = cadd(cs, isn);
      }
    }
    a[0] = x0;
  }
};
template <bool INV, typename C> struct RDft<3, INV, C> : RDftOdd<3, INV, C> {};
template <bool INV, typename C> struct RDft<5, INV, C> : RDftOdd<5, INV, C> {};
template <bool INV, typename C> struct RDft<7, INV, C> : RDftOdd<7, INV, C> {};

template <int R, bool INV, typename C> struct RDft {
  static_assert(R == 6 || R == 8 || R == 10 || R == 12 || R == 16 || R == 20 || R == 24 || R == 32,
                "unsupported in-register DFT length");
  static __device__ inline void run(C* a) {
    constexpr int P = SmallFactor<R>::value;
    constexpr int Q = R / P;
    C b[R];
#pragma unroll
    for (int q = 0; q < Q; q++) {
      C t[P];
#pragma unroll
      for (int p = 0; p < P; p++) t[p] = a[q + Q * p];
      RDft<P, INV, C>::run(t);
#pragma unroll
      for (int k1 = 0; k1 < P; k1++) b[q + Q * k1] = cmul_w<INV>(t[k1], q * k1, R);
    }
#pragma unroll
    for (int k1 = 0; k1 < P; k1++) {
      C t[Q];
#pragma unroll
      for (int q = 0; q < Q; q++) t[q] = b[q + Q * k1];
      RDft<Q, INV, C>::run(t);
#pragma unroll
      for (int k2 = 0; k2 < Q; k2++) a[k1 + P * k2] = t[k2];
    }
  }
};

// DFT of length R (even) whose inputs a[R/2 .. R-1] are zero; those slots are
// not read. X[2j] = DFT_{R/2}(a)[j], X[2j+1] = DFT_{R/2}(a[m] W_R^m)[j].
template <int R, bool INV, typename C> struct RDftHalfIn {
  static_assert(R % 2 == 0, "half-input DFT needs an even length");
  static __device__ inline void run(C* a) {
    constexpr int H = R / 2;
    C e[H], o[H];
#pragma unroll
    for (int m = 0; m < H; m++) {
      e[m] = a[m];
      o[m] = cmul_w<INV>(a[m], m, R);
    }
    RDft<H, INV, C>::run(e);
    RDft<H, INV, C>::run(o);
#pragma unroll
    for (int j = 0; j < H; j++) {
      a[2 * j] = e[j];
      a[2 * j + 1] = o[j];
    }
  }
};

// ------------------------------------------------------------------------------------
// Plans
// ------------------------------------------------------------------------------------
template <int... Rs> struct RadixList {};

// StaticPlan<N, WIDE>: PPT and the forward radix order (Fwd). WIDE plans (used
// for complex float) hold 16 points per thread and need one LDS exchange fewer;
// narrow plans (complex double) hold 8 to keep three components in registers.
template <int N, bool WIDE> struct StaticPlan { static constexpr bool exists = false; };
#define DDA_PLAN(N, W, P, ...)                            \
  template <> struct StaticPlan<N, W> {                   \
    static constexpr bool exists = true;                  \
    static constexpr int PPT = P;                         \
    typedef RadixList<__VA_ARGS__> Fwd;                   \
  };
DDA_PLAN(32, false, 8, 8, 4)
DDA_PLAN(32, true, 16, 16, 2)
DDA_PLAN(48, false, 12, 12, 4)
DDA_PLAN(48, true, 12, 12, 4)
DDA_PLAN(64, false, 8, 8, 8)
DDA_PLAN(64, true, 8, 8, 8)
DDA_PLAN(96, false, 12, 6, 4, 4)
DDA_PLAN(96, true, 12, 6, 4, 4)
DDA_PLAN(128, false, 8, 8, 4, 4)
DDA_PLAN(128, true, 16, 16, 8)
DDA_PLAN(192, false, 12, 12, 4, 4)
DDA_PLAN(192, true, 12, 12, 4, 4)
DDA_PLAN(256, false, 8, 8, 8, 4)
DDA_PLAN(256, true, 16, 16, 16)
DDA_PLAN(512, false, 8, 8, 8, 8)
DDA_PLAN(512, true, 16, 16, 16, 2)
DDA_PLAN(1024, false, 8, 8, 8, 4, 4)
DDA_PLAN(1024, true, 16, 16, 16, 4)
#undef DDA_PLAN

// A plan given by its points per thread and radix order directly (the Z
// pass's long-column variants: 1024 = 32 x 32 holds 32 points per thread and
// needs a single LDS exchange per transform).
template <int N, int P, typename L> struct PlanOf {
  static constexpr bool exists = true;
  static constexpr int PPT = P;
  typedef L Fwd;
};

template <typename L> struct Reverse;
template <> struct Reverse<RadixList<>> { typedef RadixList<> type; };
template <int R, int... Rs> struct Reverse<RadixList<R, Rs...>> {
  template <typename A, int X> struct Append;
  template <int... As, int X> struct Append<RadixList<As...>, X> { typedef RadixList<As..., X> type; };
  typedef typename Append<typename Reverse<RadixList<Rs...>>::type, R>::type type;
};

// Stage helpers. v[i * R + m] is input m of butterfly j_i = t + i * TPC. The
// twiddle W_{NS R}^k comes from the table of N roots of unity, its powers by
// repeated multiplication. HALF (first stage only): inputs m >= R/2 are zero
// and not read.
// PRE: tw holds the R - 1 twiddle powers of the thread's one butterfly of the
// twiddled stage (formed once by the caller and shared by all its transforms).
template <int N, int PPT, int R, int NS, bool INV, bool HALF, typename C, bool PRE = false>
__device__ inline void stage_compute(C* v, int t, const C* __restrict__ tw) {
  constexpr int TPC = N / PPT;
#pragma unroll
  for (int i = 0; i < PPT / R; i++) {
    C* a = v + i * R;
    if (NS > 1 && PRE) {
#pragma unroll
      for (int m = 1; m < R; m++) a[m] = INV ? cmulc(a[m], tw[m - 1]) : cmul(a[m], tw[m - 1]);
    } else if (NS > 1) {
      const int j = t + i * TPC;
      const int k = j % NS;  // NS is a compile-time constant
      const C w1 = tw[k * (N / (NS * R))];
      C w = w1;
#pragma unroll
      for (int m = 1; m < R; m++) {
        a[m] = INV ? cmulc(a[m], w) : cmul(a[m], w);
        if (m + 1 < R) w = cmul(w, w1);
      }
    }
    if (HALF && NS == 1)
      RDftHalfIn<R, INV, C>::run(a);
    else
      RDft<R, INV, C>::run(a);
  }
}

// Writes the stage outputs to the tile at their Stockham positions.
template <int N, int PPT, int R, int NS, typename C>
__device__ inline void stage_store(const C* v, int t, C* tile, int lp, int l) {
  constexpr int TPC = N / PPT;
#pragma unroll
  for (int i = 0; i < PPT / R; i++) {
    const int j = t + i * TPC;
    const int q = j / NS, k = j % NS;
    const int base = q * NS * R + k;
#pragma unroll
    for (int m = 0; m < R; m++) tile[(base + m * NS) * lp + l] = v[i * R + m];
  }
}

// Reads the inputs of a stage of radix R from the tile.
template <int N, int PPT, int R, typename C>
__device__ inline void stage_load(C* v, int t, const C* tile, int lp, int l) {
  constexpr int TPC = N / PPT;
#pragma unroll
  for (int i = 0; i < PPT / R; i++) {
    const int j = t + i * TPC;
#pragma unroll
    for (int m = 0; m < R; m++) v[i * R + m] = tile[(j + m * (N / R)) * lp + l];
  }
}

// Runs all stages of RadixList from registers to registers with LDS exchanges
// in between. On entry v holds the first stage's inputs, on exit the last
// stage's outputs (both in the layout described above). Ends with a barrier if
// any exchange happened, so the caller may reuse the tile. HALF: the upper half
// of every first-stage butterfly's inputs is zero (a zero-padded transform of
// at most N/2 points) and is neither read nor computed with.
// PRE (two-stage plans with one butterfly per thread per stage, see
// TwoStage): tw is the thread's R2 - 1 second-stage twiddle powers.
template <int N, int PPT, bool INV, int NS, typename C, typename L, bool HALF = false, bool PRE = false>
struct RunStages;

template <int N, int PPT, bool INV, int NS, typename C, int R, bool HALF, bool PRE>
struct RunStages<N, PPT, INV, NS, C, RadixList<R>, HALF, PRE> {
  static __device__ inline void run(C* v, int t, C* tile, int lp, int l, const C* tw) {
    stage_compute<N, PPT, R, NS, INV, HALF, C, PRE>(v, t, tw);
  }
};

template <int N, int PPT, bool INV, int NS, typename C, int R, int R2, int... Rs, bool HALF, bool PRE>
struct RunStages<N, PPT, INV, NS, C, RadixList<R, R2, Rs...>, HALF, PRE> {
  static __device__ inline void run(C* v, int t, C* tile, int lp, int l, const C* tw) {
    stage_compute<N, PPT, R, NS, INV, HALF, C, PRE>(v, t, tw);
    stage_store<N, PPT, R, NS>(v, t, tile, lp, l);
    __syncthreads();
    stage_load<N, PPT, R2>(v, t, tile, lp, l);
    __syncthreads();
    RunStages<N, PPT, INV, NS * R, C, RadixList<R2, Rs...>, false, PRE>::run(v, t, tile, lp, l, tw);
  }
};

// Plans of two stages with PPT = R0 = R1 (a thread holds one butterfly of
// each stage; its second-stage twiddles are W_N^(t m), m = 1 .. R1 - 1, the
// same for the forward and the conjugated inverse transform).
template <int N, int PPT, typename L> struct TwoStage { static constexpr bool value = false; };
template <int N, int PPT, int R0, int R1> struct TwoStage<N, PPT, RadixList<R0, R1>> {
  static constexpr bool value = PPT == R0 && PPT == R1 && R0 * R1 == N;
};
template <int R, typename C>
__device__ inline void twiddle_powers(C* wp, int t, const C* __restrict__ tw) {
  const C w1 = tw[t % R];
  C w = w1;
#pragma unroll
  for (int m = 0; m < R - 1; m++) {
    wp[m] = w;
    if (m + 2 < R) w = cmul(w, w1);
  }
}

template <typename L> struct FirstRadix;
template <int R, int... Rs> struct FirstRadix<RadixList<R, Rs...>> { static constexpr int value = R; };

// Point index held in v[e] for the first stage of plan L (e = i * R0 + m).
template <int N, int PPT, typename L>
__device__ inline int first_point(int t, int e) {
  constexpr int R0 = FirstRadix<L>::value;
  return t + (e / R0) * (N / PPT) + (e % R0) * (N / R0);
}

// Whether v[e] of that layout lies in the upper half of the transform (point
// >= N/2) for every thread: m = e mod R0 >= R0 / 2 with R0 even. A transform of
// at most N/2 points neither reads (first stage) nor writes (last stage) them.
template <typename L>
__device__ __forceinline__ constexpr bool upper_half(int e) {
  return FirstRadix<L>::value % 2 == 0 && (e % FirstRadix<L>::value) >= FirstRadix<L>::value / 2;
}

}  // namespace dda
