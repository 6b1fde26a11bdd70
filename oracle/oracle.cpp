// CPU oracle for the DDA FFT-convolution matvec and the Krylov solvers around it.
//
// TEST INFRASTRUCTURE ONLY. This file is the checker, never the product: only the
// test suite, the smoke check and the CPU-baseline leg of the benchmark load the
// library built from it. The HIP product path never links or calls it.
//
// It restates, in plain C++ on the CPU, the reference's own CPU code path:
//   * FFT:            LinAlg/FFTWPlan.cpp:82,113 (forward sign -1, unnormalised,
//                     contiguous batches).  FFTW itself is absent here, so the
//                     transform is our own mixed-radix recursion, computed in
//                     long double with twiddles from long double cos/sin.
//   * grid fit:       DDA/DDAParams.cpp:45-90 (fftFit), stride :113
//   * geometry:       DDA/Shapes.cpp:53-143, DDA/DipoleGeometry.cpp:75-107
//   * polarizability: DDA/PolarizabilityDescription.cpp:65-110, DDA/FPConst.hpp:48-50,
//                     DDA/DDAParams.cpp:315-331
//   * beam:           DDA/Beam.cpp:34-51,110-133
//   * D matrix:       DDA/DMatrixCpu.cpp:49-74,124-205
//   * matvec:         DDA/MatVecCpu.cpp:67-150
//   * solvers:        DDA/IterativeSolverBase.cpp:40-103, DDA/CpuIterativeSolver.cpp:45-99,
//                     DDA/QmrCs.cpp:58-168, DDA/BicgStab.cpp:59-112, DDA/Cgnr.cpp:39-73,
//                     DDA/BicgCs.cpp:56-84
//   * observables:    DDA/Beam.cpp:58-68 (Cext sum), DDA/AbsCross.cpp:27-48,
//                     DDA/CpuFieldCalculator.cpp:48-98, DDA/FarFieldCalc.cpp:119-160,
//                     EMSim/JonesCalculus.cpp:74-86, EMSim/Mie.cpp:40-174
//
// What pins each function (the checks live in tests/test_oracle.py):
//   orc_fft            brute-force O(N^2) DFT, numpy.fft, Parseval, round trip
//   orc_dmatrix        numpy.fft.fftn of the directly evaluated Green tensor
//   orc_matvec         dense O(nv^2) direct sum (orc_matvec_direct), grid-size
//                      invariance, complex symmetry x^T A y = y^T A x
//   orc_matvec_direct  Green tensor closed form checked at hand-computed points
//   orc_solve          numpy.linalg.solve on the dense matrix; all four solvers agree;
//                      BREAKDOWN forced by a right-hand side with b^T b = 0,
//                      STAGNATION checked against the returned residual history
//   orc_cc             Clausius-Mossotti closed form; LDR -> CM as kd -> 0; the
//                      finite-kd terms (b1..b3, S) by the lattice sphere's
//                      absorption against Mie at oblique incidence, which
//                      planted wrong constants miss; LDR-avgpol is the mean of
//                      1/alpha over both polarisations
//   orc_mie            optical theorem / known Qext values (textbook)
//   orc_ext_cross_*    the two Cext formulas agree (optical theorem)
//   orc_abs_cross,     energy conservation Cext = Cabs + Csca (Csca by quadrature
//   orc_calc_field     of |E_far|^2); Cabs = 0 for real m
//   orc_jones          integral of sum |J_ij|^2 = Csca(pol 1) + Csca(pol 2); zero
//                      off-diagonals on the lattice sphere's mirror planes; forward
//                      amplitude gives Cext; symmetric = 1 from one solve equals
//                      the two-solve result on a chiral quarter-turn-symmetric
//                      target, where planted sign or axis mistakes differ
//   orc_beam_pols,     orthonormal frame, pol1 horizontal, pol2 = pol1 x prop;
//   orc_einc_plane     unit amplitude, phase k d per step along prop
// Exact iteration counts and residual histories of the reference (built with
// -ffast-math and FFTW_ESTIMATE plans) are parity unpinned to the bit; they are
// pinned to tolerance only.

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef long double ld;

namespace {

const ld kPi = 3.1415926535897932384626433832795028L;

// ------------------------------------------------------------------------------------
// FFT: recursive mixed-radix decimation in time over the prime factors of n.
// exp(-2 pi i k/N) for the forward direction, exp(+...) for the inverse, both
// unnormalised (FFTWPlan.cpp:82 uses FFTW_FORWARD / FFTW_BACKWARD).
// ------------------------------------------------------------------------------------
template <typename T>
struct Fft {
  typedef std::complex<T> C;
  int n;
  std::vector<int> fac;
  std::vector<C> tw;  // exp(-2 pi i k / n)
  std::vector<C> buf;

  explicit Fft(int n_) : n(n_), tw(n_), buf(n_) {
    int m = n;
    while (m % 4 == 0) { fac.push_back(4); m /= 4; }
    for (int p = 2; m > 1; p++)
      while (m % p == 0) { fac.push_back(p); m /= p; }
    for (int k = 0; k < n; k++) {
      ld a = -2 * kPi * (ld)k / (ld)n;
      tw[k] = C((T)std::cos(a), (T)std::sin(a));
    }
  }

  C w(long t, int len, bool inv) const {
    long idx = (t % len) * (n / len);
    return inv ? std::conj(tw[idx]) : tw[idx];
  }

  void rec(int len, const C* in, long is, C* out, size_t fi, bool inv) {
    if (len == 1) { out[0] = in[0]; return; }
    int p = fac[fi];
    int m = len / p;
    for (int q = 0; q < p; q++) rec(m, in + q * is, is * p, out + q * m, fi + 1, inv);
    C tmp[16];
    std::vector<C> big;
    C* t = tmp;
    if (p > 16) { big.resize(p); t = big.data(); }
    for (int k = 0; k < m; k++) {
      for (int q = 0; q < p; q++) t[q] = out[q * m + k] * w((long)q * k, len, inv);
      for (int s = 0; s < p; s++) {
        C acc = 0;
        for (int q = 0; q < p; q++) acc += t[q] * w((long)q * s * m, len, inv);
        out[k + s * m] = acc;
      }
    }
  }

  // In-place transform of one line with element stride `stride`.
  void run(C* data, long stride, bool inv) {
    std::vector<C> in(n);
    for (int i = 0; i < n; i++) in[i] = data[i * stride];
    rec(n, in.data(), 1, buf.data(), 0, inv);
    for (int i = 0; i < n; i++) data[i * stride] = buf[i];
  }
};

// ------------------------------------------------------------------------------------
// Problem description shared by the matvec and the solvers.
// ------------------------------------------------------------------------------------
struct Prob {
  uint32_t box[3];
  uint32_t grid[3];
  uint32_t nv;
  uint32_t vec_stride;
  const uint32_t* pos;     // [nv][3], grid coordinates
  const uint8_t* mat;      // [nv]
  uint32_t nmat;
  const double* cc_sqrt;   // [nmat][3] complex (interleaved)
  const double* dmatrix;   // reference layout: ((j + Gy*(k + Gz*i))*6 + comp) complex
};

template <typename T>
std::complex<T> cplx(const double* p, size_t i) { return std::complex<T>((T)p[2 * i], (T)p[2 * i + 1]); }

// Green tensor element, DMatrixCpu.cpp:49-63 (r in metres, k = wave number).
template <typename T>
std::complex<T> green(T rx, T ry, T rz, T k, int mu, int nu) {
  typedef std::complex<T> C;
  T r[3] = {rx, ry, rz};
  T rr = std::sqrt(rx * rx + ry * ry + rz * rz);
  T q[3] = {rx / rr, ry / rr, rz / rr};
  T kr = k * rr;
  T kr2 = kr * kr;
  T qmunu = q[mu] * q[nu];
  C expval = std::exp(C(0, kr)) / C(rr * rr * rr);
  C br = C(3 - kr2, -3 * kr) * qmunu;
  if (mu == nu) br += C(kr2 - 1, kr);
  (void)r;
  return expval * br;
}

const int kMu[6] = {0, 0, 0, 1, 1, 2};
const int kNu[6] = {0, 1, 2, 1, 2, 2};

uint32_t clip(int32_t n, uint32_t M) {  // DMatrixCpu.cpp:124-131
  if (n < 0) return n + M;
  if ((uint32_t)n >= M) return n - M;
  return n;
}

// Matvec following MatVecCpu.cpp:67-150 (one x-plane at a time).
template <typename T>
void matvec(const Prob& g, const std::complex<T>* arg, std::complex<T>* result, bool conj) {
  typedef std::complex<T> C;
  const size_t Gx = g.grid[0], Gy = g.grid[1], Gz = g.grid[2];
  const size_t By = g.box[1], Bz = g.box[2];
  const size_t S = g.vec_stride;
  // Xmatrix[gx][by][bz][3], Fortran order: x fastest.
  std::vector<C> X(Gx * By * Bz * 3, C(0));
  auto xi = [&](size_t x, size_t y, size_t z, size_t c) -> C& { return X[x + Gx * (y + By * (z + Bz * c))]; };
  auto mc = [&](C v) { return conj ? std::conj(v) : v; };
  for (uint32_t i = 0; i < g.nv; i++) {
    const uint32_t* p = g.pos + 3 * i;
    uint8_t m = g.mat[i];
    for (int c = 0; c < 3; c++) xi(p[0], p[1], p[2], c) = cplx<T>(g.cc_sqrt, 3 * m + c) * mc(arg[i + c * S]);
  }
  // Lines and x-planes are independent: with OpenMP each thread takes whole
  // lines / planes with its own FFT work space, so the result does not depend
  // on the thread count.
#pragma omp parallel
  {
    Fft<T> fx((int)Gx);
#pragma omp for schedule(static)
    for (size_t l = 0; l < By * Bz * 3; l++) fx.run(&X[l * Gx], 1, false);
  }
#pragma omp parallel
  {
  Fft<T> fy((int)Gy), fz((int)Gz);
  // slices[gz][gy][3] Fortran (z fastest); slices_tr[gy][gz][3] (y fastest).
  std::vector<C> sl(Gz * Gy * 3), st(Gy * Gz * 3);
#pragma omp for schedule(static)
  for (size_t i = 0; i < Gx; i++) {
    std::fill(sl.begin(), sl.end(), C(0));
    for (size_t j = 0; j < By; j++)
      for (size_t k = 0; k < Bz; k++)
        for (int c = 0; c < 3; c++) sl[k + Gz * (j + Gy * c)] = xi(i, j, k, c);
    for (int c = 0; c < 3; c++)
      for (size_t j = 0; j < By; j++) fz.run(&sl[Gz * (j + Gy * c)], 1, false);
    for (int c = 0; c < 3; c++)
      for (size_t k = 0; k < Gz; k++)
        for (size_t j = 0; j < Gy; j++) st[j + Gy * (k + Gz * c)] = sl[k + Gz * (j + Gy * c)];
    for (size_t l = 0; l < Gz * 3; l++) fy.run(&st[l * Gy], 1, false);
    for (size_t k = 0; k < Gz; k++)
      for (size_t j = 0; j < Gy; j++) {
        const size_t d = (j + Gy * (k + Gz * i)) * 6;
        C D[6];
        for (int e = 0; e < 6; e++) D[e] = cplx<T>(g.dmatrix, d + e);
        C x0 = st[j + Gy * (k + Gz * 0)], x1 = st[j + Gy * (k + Gz * 1)], x2 = st[j + Gy * (k + Gz * 2)];
        // SymMatrix3 (aa,ab,ac,bb,bc,cc) times vector, Math/SymMatrix3.hpp:37-41
        st[j + Gy * (k + Gz * 0)] = D[0] * x0 + D[1] * x1 + D[2] * x2;
        st[j + Gy * (k + Gz * 1)] = D[1] * x0 + D[3] * x1 + D[4] * x2;
        st[j + Gy * (k + Gz * 2)] = D[2] * x0 + D[4] * x1 + D[5] * x2;
      }
    for (size_t l = 0; l < Gz * 3; l++) fy.run(&st[l * Gy], 1, true);
    for (int c = 0; c < 3; c++)
      for (size_t k = 0; k < Gz; k++)
        for (size_t j = 0; j < Gy; j++) sl[k + Gz * (j + Gy * c)] = st[j + Gy * (k + Gz * c)];
    for (int c = 0; c < 3; c++)
      for (size_t j = 0; j < By; j++) fz.run(&sl[Gz * (j + Gy * c)], 1, true);
    for (size_t j = 0; j < By; j++)
      for (size_t k = 0; k < Bz; k++)
        for (int c = 0; c < 3; c++) xi(i, j, k, c) = sl[k + Gz * (j + Gy * c)];
  }
  }
#pragma omp parallel
  {
    Fft<T> fx((int)Gx);
#pragma omp for schedule(static)
    for (size_t l = 0; l < By * Bz * 3; l++) fx.run(&X[l * Gx], 1, true);
  }
  for (uint32_t i = g.nv; i < S; i++)
    for (int c = 0; c < 3; c++) result[i + c * S] = 0;
  for (uint32_t i = 0; i < g.nv; i++) {
    const uint32_t* p = g.pos + 3 * i;
    uint8_t m = g.mat[i];
    for (int c = 0; c < 3; c++) {
      C r = xi(p[0], p[1], p[2], c);
      C a = mc(arg[i + c * S]);
      result[i + c * S] = mc(cplx<T>(g.cc_sqrt, 3 * m + c) * r + a);
    }
  }
}

// ------------------------------------------------------------------------------------
// Krylov solvers.
// ------------------------------------------------------------------------------------
template <typename T>
struct Vec {
  typedef std::complex<T> C;
  std::vector<C> v;
  explicit Vec(size_t n = 0) : v(n, C(0)) {}
  C& operator[](size_t i) { return v[i]; }
  const C& operator[](size_t i) const { return v[i]; }
  size_t size() const { return v.size(); }
};

template <typename T> T norm2(const Vec<T>& a) {  // LinAlg/LinComb.hpp norm(): sum |v|^2
  T s = 0;
  for (size_t i = 0; i < a.size(); i++) s += std::norm(a[i]);
  return s;
}
template <typename T> std::complex<T> dotu(const Vec<T>& a, const Vec<T>& b) {  // sum a*b
  std::complex<T> s = 0;
  for (size_t i = 0; i < a.size(); i++) s += a[i] * b[i];
  return s;
}
template <typename T> std::complex<T> dotc(const Vec<T>& a, const Vec<T>& b) {  // sum a*conj(b)
  std::complex<T> s = 0;
  for (size_t i = 0; i < a.size(); i++) s += a[i] * std::conj(b[i]);
  return s;
}

enum { ST_OK = 0, ST_INVALID = 1, ST_BREAKDOWN = 4, ST_MAXITER = 5, ST_STAGNATION = 6 };
enum { K_QMR = 0, K_BICGSTAB = 1, K_CGNR = 2, K_BICG = 3 };

template <typename T>
struct Solver {
  typedef std::complex<T> C;
  const Prob& g;
  int kind;
  size_t n;
  Vec<T> r, x, Av, t1, t2, t3, t4;
  T residScale = 0, epsB = 0, inprodR = 0;
  int status = ST_OK;
  // QMR-CS scalars (QmrCs.hpp VarType)
  T omega_old = 0, omega_new = 0, c_old = 0, c_new = 0;
  C beta = 0, mBeta = 0, tautilda = 0, s_old = 0, s_new = 0;
  // BiCGStab / CGNR / BiCG scalars
  C ro_new = 0, ro_old = 0, alpha = 0, omega = 0, bbeta = 0;
  T denumOmega = 0, abs_ro_new = 0, cg_ro_new = 0, cg_ro_old = 0;

  Solver(const Prob& g_, int kind_)
      : g(g_), kind(kind_), n(3 * (size_t)g_.vec_stride), r(n), x(n), Av(n), t1(n), t2(n), t3(n), t4(n) {}

  void apply(const Vec<T>& a, Vec<T>& b, bool conj) { matvec<T>(g, a.v.data(), b.v.data(), conj); }

  C ccs(size_t i, int c) const { return cplx<T>(g.cc_sqrt, 3 * g.mat[i] + c); }

  // CpuIterativeSolver::initGeneral, CpuIterativeSolver.cpp:45-90
  T initGeneral(const Vec<T>& einc, const Vec<T>* start) {
    Vec<T>& pvec = t1;
    const size_t S = g.vec_stride;
    for (size_t i = 0; i < n; i++) pvec[i] = 0;
    for (uint32_t i = 0; i < g.nv; i++)
      for (int c = 0; c < 3; c++) pvec[i + c * S] = ccs(i, c) * einc[i + c * S];
    T temp = norm2(pvec);
    residScale = 1 / temp;
    T ip;
    if (start) {
      for (size_t i = 0; i < n; i++) x[i] = 0;
      for (uint32_t i = 0; i < g.nv; i++)
        for (int c = 0; c < 3; c++) x[i + c * S] = (*start)[i + c * S] / ccs(i, c);
      apply(x, Av, false);
      for (size_t i = 0; i < n; i++) r[i] = Av[i] * C(-1) + pvec[i];
      ip = norm2(r);
    } else {
      apply(pvec, Av, false);
      for (size_t i = 0; i < n; i++) r[i] = Av[i] * C(-1) + pvec[i];
      ip = norm2(r);
      if (temp < ip) {
        for (size_t i = 0; i < n; i++) x[i] = 0;
        std::swap(r.v, pvec.v);
        ip = temp;
      } else {
        std::swap(x.v, pvec.v);
      }
    }
    return ip;
  }

  void init() {
    if (kind == K_QMR) {  // QmrCs.cpp:58-76
      Vec<T>& v = t1;
      C rvec2 = dotu(r, r);
      omega_old = 0;
      beta = std::sqrt(rvec2);
      mBeta = -beta;
      omega_new = std::sqrt(inprodR) / std::abs(beta);
      tautilda = omega_new * beta;
      c_old = 1; c_new = 1; s_old = 0; s_new = 0;
      C ib = C(1) / beta;
      for (size_t i = 0; i < n; i++) v[i] = r[i] * ib;
    } else if (kind == K_BICGSTAB) {  // BicgStab.cpp:59-64
      t4 = r;
    } else if (kind == K_CGNR) {  // Cgnr.cpp:39-41
      cg_ro_old = 0;
    }
  }

  // Returns the new squared residual norm; sets status on breakdown.
  T iteration(size_t nr) {
    if (kind == K_QMR) return itQmr(nr);
    if (kind == K_BICGSTAB) return itBicgStab(nr);
    if (kind == K_CGNR) return itCgnr(nr);
    return itBicg(nr);
  }

  T itQmr(size_t nr) {  // QmrCs.cpp:78-168
    Vec<T>& v = t1; Vec<T>& vtilda = t2; Vec<T>& p_old = t3; Vec<T>& p_new = t4;
    T rtmp1 = std::norm(beta) * residScale;
    if (nr == 0) {
      if (rtmp1 > 1e+38f) { status = ST_BREAKDOWN; return 0; }
    } else if (rtmp1 < 1e-10f || rtmp1 > 1e+38f) { status = ST_BREAKDOWN; return 0; }
    apply(v, Av, false);
    C a = dotu(v, Av);
    if (nr == 0)
      for (size_t i = 0; i < n; i++) vtilda[i] = v[i] * (-a) + Av[i];
    else
      for (size_t i = 0; i < n; i++) vtilda[i] = vtilda[i] * mBeta + v[i] * (-a) + Av[i];
    C ctmp3 = dotu(vtilda, vtilda);
    T rtmp2 = norm2(vtilda);
    C ctmp1 = omega_old * beta;
    C ctmp2 = omega_new * a;
    C theta = std::conj(s_old) * ctmp1;
    C eta = c_old * c_new * ctmp1 + std::conj(s_new) * ctmp2;
    C zetatilda = c_new * ctmp2 - c_old * s_new * ctmp1;
    beta = std::sqrt(ctmp3);
    mBeta = -beta;
    omega_old = omega_new;
    omega_new = std::sqrt(rtmp2) / std::abs(beta);
    T zetaabs = std::sqrt(std::norm(zetatilda) + rtmp2);
    T rtmp3 = std::sqrt(std::norm(zetatilda));
    C zeta = rtmp3 < 1e-40f ? C(zetaabs) : zetaabs / rtmp3 * zetatilda;
    c_old = c_new;
    c_new = rtmp3 / zetaabs;
    s_old = s_new;
    s_new = omega_new * (beta / zeta);
    C tau = c_new * tautilda;
    tautilda = -s_new * tautilda;
    C zetaInv = C(1) / zeta;
    C mEtaZeta = -eta / zeta;
    C mThetaZeta = -theta / zeta;
    C betaInv = C(1) / beta;
    T normSNew = std::norm(s_new);
    C cNewOmegaNewTautilda = (c_new / omega_new) * tautilda;
    if (nr == 0) {
      for (size_t i = 0; i < n; i++) p_new[i] = v[i] * zetaInv;
    } else {
      if (nr == 1)
        for (size_t i = 0; i < n; i++) p_old[i] = p_new[i] * mEtaZeta + v[i] * zetaInv;
      else
        for (size_t i = 0; i < n; i++) p_old[i] = p_old[i] * mThetaZeta + p_new[i] * mEtaZeta + v[i] * zetaInv;
      std::swap(p_old.v, p_new.v);
    }
    for (size_t i = 0; i < n; i++) x[i] = p_new[i] * tau + x[i];
    for (size_t i = 0; i < n; i++) vtilda[i] = vtilda[i] * betaInv;
    std::swap(v.v, vtilda.v);
    for (size_t i = 0; i < n; i++) r[i] = r[i] * C(normSNew) + v[i] * cNewOmegaNewTautilda;
    return norm2(r);
  }

  T itBicgStab(size_t nr) {  // BicgStab.cpp:66-112
    Vec<T>& pvec = t1; Vec<T>& v = t2; Vec<T>& s = t3; Vec<T>& rtilda = t4;
    typedef std::complex<ld> CL;
    ro_new = dotc(r, rtilda);
    T dtmp = (T)(std::abs((CL)ro_new) / (ld)inprodR);
    if (!(dtmp >= 1e-16)) { status = ST_BREAKDOWN; return 0; }
    if (nr == 0) {
      pvec = r;
    } else {
      CL ro_new_alpha = (CL)ro_new * (CL)alpha;
      CL ro_old_omega = (CL)ro_old * (CL)omega;
      if (!(std::abs(ro_old_omega) / std::abs(ro_new_alpha) >= 10e-10)) { status = ST_BREAKDOWN; return 0; }
      bbeta = (C)(ro_new_alpha / ro_old_omega);
      C f2 = -bbeta * omega;
      for (size_t i = 0; i < n; i++) pvec[i] = pvec[i] * bbeta + v[i] * f2 + r[i];
    }
    apply(pvec, v, false);
    CL vR = (CL)dotc(v, rtilda);
    alpha = (C)((CL)ro_new / vR);
    for (size_t i = 0; i < n; i++) s[i] = v[i] * (-alpha) + r[i];
    T ip1 = norm2(s);
    if (ip1 < epsB) {
      for (size_t i = 0; i < n; i++) x[i] = pvec[i] * alpha + x[i];
    } else {
      apply(s, Av, false);
      denumOmega = norm2(Av);
      omega = (C)((CL)dotc(s, Av) / (ld)denumOmega);
      for (size_t i = 0; i < n; i++) x[i] = pvec[i] * alpha + s[i] * omega + x[i];
      for (size_t i = 0; i < n; i++) r[i] = Av[i] * (-omega) + s[i];
      ip1 = norm2(r);
      ro_old = ro_new;
    }
    return ip1;
  }

  T itCgnr(size_t nr) {  // Cgnr.cpp:43-73
    Vec<T>& pvec = t1;
    if (nr == 0) {
      apply(r, pvec, true);
      cg_ro_new = norm2(pvec);
    } else {
      apply(r, Av, true);
      cg_ro_new = norm2(Av);
      T b = cg_ro_new / cg_ro_old;
      for (size_t i = 0; i < n; i++) pvec[i] = pvec[i] * b + Av[i];
    }
    apply(pvec, Av, false);
    C a = cg_ro_new / norm2(Av);
    for (size_t i = 0; i < n; i++) x[i] = pvec[i] * a + x[i];
    for (size_t i = 0; i < n; i++) r[i] = Av[i] * (-a) + r[i];
    cg_ro_old = cg_ro_new;
    return norm2(r);
  }

  T itBicg(size_t nr) {  // BicgCs.cpp:58-84
    Vec<T>& pvec = t1;
    ro_new = dotu(r, r);
    abs_ro_new = std::abs(ro_new);
    T dtmp = abs_ro_new / inprodR;
    if (dtmp < 1e-10 || dtmp > 1e+10) { status = ST_BREAKDOWN; return 0; }
    if (nr == 0) {
      pvec = r;
    } else {
      bbeta = ro_new / ro_old;
      for (size_t i = 0; i < n; i++) pvec[i] = pvec[i] * bbeta + r[i];
    }
    apply(pvec, Av, false);
    C mu = dotu(pvec, Av);
    T dtmp2 = std::abs(mu) / abs_ro_new;
    if (dtmp2 < 10e-10) { status = ST_BREAKDOWN; return 0; }
    alpha = ro_new / mu;
    for (size_t i = 0; i < n; i++) x[i] = pvec[i] * alpha + x[i];
    for (size_t i = 0; i < n; i++) r[i] = Av[i] * (-alpha) + r[i];
    ro_old = ro_new;
    return norm2(r);
  }
};

size_t maxResIncreaseFor(int kind) {  // QmrCs.cpp:39, BicgStab.cpp:51, Cgnr.cpp:34, BicgCs.cpp:48
  switch (kind) {
    case K_QMR: return 50000;
    case K_BICGSTAB: return 30000;
    case K_CGNR: return 10;
    default: return 50000;
  }
}

template <typename T>
std::vector<std::complex<T>> loadVec(const double* p, size_t n) {
  std::vector<std::complex<T>> v(n);
  for (size_t i = 0; i < n; i++) v[i] = cplx<T>(p, i);
  return v;
}

}  // namespace

extern "C" {

struct orc_prob {
  uint32_t box[3];
  uint32_t grid[3];
  uint32_t nv;
  uint32_t vec_stride;
  const uint32_t* pos;
  const uint8_t* mat;
  uint32_t nmat;
  const double* cc_sqrt;
  const double* dmatrix;
};

static Prob toProb(const orc_prob* p) {
  Prob g;
  std::memcpy(g.box, p->box, sizeof g.box);
  std::memcpy(g.grid, p->grid, sizeof g.grid);
  g.nv = p->nv; g.vec_stride = p->vec_stride; g.pos = p->pos; g.mat = p->mat;
  g.nmat = p->nmat; g.cc_sqrt = p->cc_sqrt; g.dmatrix = p->dmatrix;
  return g;
}

// Batched in-place FFT of `howmany` contiguous lines of length n (long double inside).
int orc_fft(int n, long howmany, double* data, int inverse) {
  if (n <= 0 || howmany < 0) return ST_INVALID;
  Fft<ld> f(n);
  std::vector<std::complex<ld>> line(n);
  for (long h = 0; h < howmany; h++) {
    double* d = data + 2 * (size_t)n * h;
    for (int i = 0; i < n; i++) line[i] = std::complex<ld>(d[2 * i], d[2 * i + 1]);
    f.run(line.data(), 1, inverse != 0);
    for (int i = 0; i < n; i++) { d[2 * i] = (double)line[i].real(); d[2 * i + 1] = (double)line[i].imag(); }
  }
  return ST_OK;
}

// fftFit, DDAParams.cpp:45-90.
uint32_t orc_fft_fit(uint32_t x, uint32_t override_, int supportNonPot) {
  if (override_ != 0) return x <= override_ ? override_ : 0;
  if (x == 0) return 0;
  if (!supportNonPot) {
    uint32_t ret = 1;
    while (x > 1) { ret *= 2; x = (x + 1) / 2; }
    return ret;
  }
  uint32_t ret = x;
  if (ret % 2 == 1) ret++;
  for (;;) {
    uint32_t y = ret;
    while (y % 2 == 0) y /= 2;
    while (y % 3 == 0) y /= 3;
    while (y % 5 == 0) y /= 5;
    while (y % 7 == 0) y /= 7;
    if (y == 1) return ret;
    ret += 2;
  }
}

// Geometry, Shapes.cpp:53-73 (sphere), :88-106 (cylinder), :130-143 (box).
// Dipoles are emitted in (z, y, x) loop order. `out` may be null to count only;
// returns the number of dipoles and writes the box into box_out.
// The ellipsoid has no counterpart in the reference: it applies the sphere's
// cell-centre rule per axis with semi-axes (a, b, c) in grid units.
long orc_geometry(int shape, double p0, double p1, double p2, uint32_t* out, uint32_t* box_out) {
  long count = 0;
  uint32_t box[3] = {0, 0, 0};
  auto add = [&](uint32_t x, uint32_t y, uint32_t z) {
    if (out) { out[3 * count] = x; out[3 * count + 1] = y; out[3 * count + 2] = z; }
    count++;
    if (x + 1 > box[0]) box[0] = x + 1;
    if (y + 1 > box[1]) box[1] = y + 1;
    if (z + 1 > box[2]) box[2] = z + 1;
  };
  if (shape == 0) {  // sphere, p0 = radius / gridUnit
    double radius = p0, diameter = radius * 2, r2 = radius * radius;
    for (uint32_t z = 0; z < diameter; z++)
      for (uint32_t y = 0; y < diameter; y++)
        for (uint32_t x = 0; x < diameter; x++) {
          double xd = x - radius + 0.5, yd = y - radius + 0.5, zd = z - radius + 0.5;
          if (xd * xd + yd * yd + zd * zd <= r2) add(x, y, z);
        }
  } else if (shape == 1) {  // box, p = size / gridUnit
    for (uint32_t z = 0; z < p2; z++)
      for (uint32_t y = 0; y < p1; y++)
        for (uint32_t x = 0; x < p0; x++) add(x, y, z);
  } else if (shape == 2) {  // cylinder along x, p0 = length, p1 = radius
    double length = p0, radius = p1, diameter = radius * 2, r2 = radius * radius;
    for (uint32_t z = 0; z < diameter; z++)
      for (uint32_t y = 0; y < diameter; y++) {
        double yd = y - radius + 0.5, zd = z - radius + 0.5;
        if (yd * yd + zd * zd <= r2)
          for (uint32_t x = 0; x < length; x++) add(x, y, z);
      }
  } else if (shape == 3) {  // ellipsoid, semi-axes p0, p1, p2
    for (uint32_t z = 0; z < 2 * p2; z++)
      for (uint32_t y = 0; y < 2 * p1; y++)
        for (uint32_t x = 0; x < 2 * p0; x++) {
          double xd = (x - p0 + 0.5) / p0, yd = (y - p1 + 0.5) / p1, zd = (z - p2 + 0.5) / p2;
          if (xd * xd + yd * yd + zd * zd <= 1.0) add(x, y, z);
        }
  } else {
    return -1;
  }
  if (box_out) std::memcpy(box_out, box, sizeof box);
  return count;
}

// Incident polarisations for a propagation direction, Beam.cpp:34-51.
// pol1 is the "y polarisation", pol2 the "x polarisation".
void orc_beam_pols(const double* prop_in, double* prop_out, double* pol1, double* pol2) {
  ld p[3] = {prop_in[0], prop_in[1], prop_in[2]};
  ld nrm = std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
  for (int i = 0; i < 3; i++) p[i] /= nrm;
  ld px[3], py[3];
  if (std::fabs(p[2]) >= (1 - 1e-6)) {
    px[0] = p[2]; px[1] = 0; px[2] = 0;
    py[0] = 0; py[1] = 1; py[2] = 0;
  } else {
    ld t = std::sqrt(1 - p[2] * p[2]);
    px[0] = p[0] * p[2] / t; px[1] = p[1] * p[2] / t; px[2] = -t;
    py[0] = -p[1] / t; py[1] = p[0] / t; py[2] = 0;
  }
  for (int i = 0; i < 3; i++) {
    prop_out[i] = (double)p[i];
    pol1[i] = (double)py[i];
    pol2[i] = (double)px[i];
  }
}

// Couple constants (polarisabilities, m^3) and their square roots for each
// material, PolarizabilityDescription.cpp:65-110 and DDAParams.cpp:315-331.
// kind: 0 = CM, 1 = LDR, 2 = LDR with avgpol. m: [nmat][3] complex refractive
// index per axis. prop/incpol: unit vectors in the particle frame.
// orc_cc_planted takes the three LDR constants and a factor on S as arguments,
// so that tests can plant a wrong constant and show that their physical checks
// reject it; orc_cc passes the reference's values.
void orc_cc_planted(int kind, uint32_t nmat, const double* m, double gridUnit, double kd, const double* prop,
                    const double* incpol, double b1, double b2, double b3, double s_factor, double* cc_out,
                    double* cc_sqrt_out) {
  typedef std::complex<double> C;
  double vol = gridUnit * gridUnit * gridUnit;
  double temp = vol / (4 * (double)kPi);
  for (uint32_t i = 0; i < nmat; i++)
    for (int j = 0; j < 3; j++) {
      C mj(m[2 * (3 * i + j)], m[2 * (3 * i + j) + 1]);
      C m2 = mj * mj;
      C unc = 3.0 * temp * (m2 - C(1)) / (m2 + C(2));
      C cc = unc;
      if (kind != 0) {
        C t1(0, 2.0 / 3 * kd * kd * kd);
        double S;
        double p2[3] = {prop[0] * prop[0], prop[1] * prop[1], prop[2] * prop[2]};
        if (kind == 2) {
          S = 0.5 * (1 - (p2[0] * p2[0] + p2[1] * p2[1] + p2[2] * p2[2]));
        } else {
          S = p2[0] * incpol[0] * incpol[0] + p2[1] * incpol[1] * incpol[1] + p2[2] * incpol[2] * incpol[2];
        }
        S *= s_factor;
        t1 += (b1 + (b2 + b3 * S) * m2) * kd * kd;
        cc = unc / (C(1) - (unc / vol) * t1);
      }
      C cs = std::sqrt(cc);
      cc_out[2 * (3 * i + j)] = cc.real(); cc_out[2 * (3 * i + j) + 1] = cc.imag();
      cc_sqrt_out[2 * (3 * i + j)] = cs.real(); cc_sqrt_out[2 * (3 * i + j) + 1] = cs.imag();
    }
}

void orc_cc(int kind, uint32_t nmat, const double* m, double gridUnit, double kd, const double* prop,
            const double* incpol, double* cc_out, double* cc_sqrt_out) {
  // ldr_b1..b3, FPConst.hpp:48-50 (doi:10.1086/172396)
  orc_cc_planted(kind, nmat, m, gridUnit, kd, prop, incpol, 1.8915316, -0.1648469, 1.7700004, 1.0, cc_out,
                 cc_sqrt_out);
}

// Plane wave incident field, Beam.cpp:110-118: E_i = exp(i k r_i . prop) pol,
// r_i = origin + pos_i * gridUnit, origin from DipoleGeometry.cpp:75-80.
void orc_einc_plane(uint32_t nv, uint32_t vec_stride, const uint32_t* pos, const uint32_t* box,
                    double gridUnit, double waveNum, const double* prop, const double* pol, double* einc) {
  typedef std::complex<double> C;
  double origin[3];
  for (int a = 0; a < 3; a++) origin[a] = -((double)box[a] - 1.0) / 2 * gridUnit;
  std::memset(einc, 0, sizeof(double) * 2 * 3 * (size_t)vec_stride);
  for (uint32_t i = 0; i < nv; i++) {
    double r[3];
    for (int a = 0; a < 3; a++) r[a] = origin[a] + (double)pos[3 * i + a] * gridUnit;
    C e = std::exp(C(0, waveNum * (r[0] * prop[0] + r[1] * prop[1] + r[2] * prop[2])));
    for (int c = 0; c < 3; c++) {
      C v = e * pol[c];
      einc[2 * (i + c * (size_t)vec_stride)] = v.real();
      einc[2 * (i + c * (size_t)vec_stride) + 1] = v.imag();
    }
  }
}

// D matrix in the reference layout, DMatrixCpu.cpp:133-205:
// out[((j + Gy*(k + Gz*i))*6 + comp)] = FFT3(G clipped to |offset| < box)/(-Gx*Gy*Gz).
// Computed in long double, returned as double.
}  // extern "C"

template <typename TT>
static int dmatrix_impl(const uint32_t* box, const uint32_t* grid, double gridUnit, double waveNum, double* out) {
  typedef TT ld;
  typedef std::complex<ld> C;
  const int Bx = box[0], By = box[1], Bz = box[2];
  const int Gx = grid[0], Gy = grid[1], Gz = grid[2];
  if (Gx < 2 * Bx - 1 || Gy < 2 * By - 1 || Gz < 2 * Bz - 1) return ST_INVALID;
  const ld N = (ld)Gx * Gy * Gz;
  // the six components are independent: one thread each, with its own buffers
#pragma omp parallel for schedule(dynamic, 1)
  for (int comp = 0; comp < 6; comp++) {
    std::vector<C> d2((size_t)Gx * Gy * Gz, C(0));  // [gx][gy][gz] Fortran: x fastest
    auto D2 = [&](int x, int y, int z) -> C& { return d2[x + (size_t)Gx * (y + (size_t)Gy * z)]; };
    std::vector<C> slice((size_t)Gz * Gy), slice_tr((size_t)Gy * Gz);
    Fft<ld> fx(Gx), fy(Gy), fz(Gz);
    for (int k = 0; k < Gz; k++) {
      int kcor = k > Gz / 2 ? k - Gz : k;
      for (int j = 1 - By; j < By; j++)
        for (int i = 1 - Bx; i < Bx; i++) {
          C v = 0;
          if (i || j || kcor)
            v = green<ld>((ld)i * gridUnit, (ld)j * gridUnit, (ld)kcor * gridUnit, (ld)waveNum, kMu[comp], kNu[comp]);
          D2(clip(i, Gx), clip(j, Gy), k) = v;
        }
    }
    for (size_t l = 0; l < (size_t)Gy * Gz; l++) fx.run(&d2[l * Gx], 1, false);
    for (int i = 0; i < Gx; i++) {
      std::fill(slice.begin(), slice.end(), C(0));
      for (int j = 1 - By; j < By; j++)
        for (int k = 1 - Bz; k < Bz; k++) {
          int y = clip(j, Gy), z = clip(k, Gz);
          slice[z + (size_t)Gz * y] = D2(i, y, z);
        }
      for (int y = 0; y < Gy; y++) fz.run(&slice[(size_t)Gz * y], 1, false);
      for (int y = 0; y < Gy; y++)
        for (int z = 0; z < Gz; z++) slice_tr[y + (size_t)Gy * z] = slice[z + (size_t)Gz * y];
      for (int z = 0; z < Gz; z++) fy.run(&slice_tr[(size_t)Gy * z], 1, false);
      for (int k = 0; k < Gz; k++)
        for (int j = 0; j < Gy; j++) {
          C v = slice_tr[j + (size_t)Gy * k] / -N;
          size_t o = ((j + (size_t)Gy * (k + (size_t)Gz * i)) * 6 + comp) * 2;
          out[o] = (double)v.real();
          out[o + 1] = (double)v.imag();
        }
    }
  }
  return ST_OK;
}

extern "C" {

int orc_dmatrix(const uint32_t* box, const uint32_t* grid, double gridUnit, double waveNum, double* out) {
  return dmatrix_impl<long double>(box, grid, gridUnit, waveNum, out);
}

// The same in double (for large grids in tests).
int orc_dmatrix64(const uint32_t* box, const uint32_t* grid, double gridUnit, double waveNum, double* out) {
  return dmatrix_impl<double>(box, grid, gridUnit, waveNum, out);
}

// Green tensor component for one offset (metres), for closed-form checks.
void orc_green(double rx, double ry, double rz, double waveNum, int mu, int nu, double* out) {
  std::complex<ld> v = green<ld>(rx, ry, rz, waveNum, mu, nu);
  out[0] = (double)v.real();
  out[1] = (double)v.imag();
}

// Threads for the FFT matvec (OpenMP; 1 = the reference's serial CPU path).
void orc_set_threads(int n) {
#ifdef _OPENMP
  omp_set_num_threads(n > 0 ? n : 1);
#else
  (void)n;
#endif
}

// FFT matvec (MatVecCpu.cpp:67-150) in double (precision 64) or long double (80).
int orc_matvec(const orc_prob* p, const double* x, double* y, int conj, int precision) {
  Prob g = toProb(p);
  size_t n = 3 * (size_t)g.vec_stride;
  if (precision == 80) {
    auto a = loadVec<ld>(x, n);
    std::vector<std::complex<ld>> b(n);
    matvec<ld>(g, a.data(), b.data(), conj != 0);
    for (size_t i = 0; i < n; i++) { y[2 * i] = (double)b[i].real(); y[2 * i + 1] = (double)b[i].imag(); }
  } else {
    auto a = loadVec<double>(x, n);
    std::vector<std::complex<double>> b(n);
    matvec<double>(g, a.data(), b.data(), conj != 0);
    for (size_t i = 0; i < n; i++) { y[2 * i] = b[i].real(); y[2 * i + 1] = b[i].imag(); }
  }
  return ST_OK;
}

// Dense direct sum in long double, independent of the FFT and the circulant
// embedding: y_i = x_i - sqrt(a_i) sum_{j != i} G(r_i - r_j) sqrt(a_j) x_j,
// with y = conj(A conj(x)) when conj is set.
int orc_matvec_direct(const orc_prob* p, double gridUnit, double waveNum, const double* x, double* y, int conj) {
  typedef std::complex<ld> C;
  const size_t S = p->vec_stride;
  std::vector<C> u(3 * (size_t)p->nv);
  auto ccs = [&](uint32_t i, int c) { return C(p->cc_sqrt[2 * (3 * p->mat[i] + c)], p->cc_sqrt[2 * (3 * p->mat[i] + c) + 1]); };
  for (uint32_t i = 0; i < p->nv; i++)
    for (int c = 0; c < 3; c++) {
      C xv(x[2 * (i + c * S)], x[2 * (i + c * S) + 1]);
      if (conj) xv = std::conj(xv);
      u[3 * i + c] = ccs(i, c) * xv;
    }
  std::memset(y, 0, sizeof(double) * 2 * 3 * S);
  for (uint32_t i = 0; i < p->nv; i++) {
    C acc[3] = {0, 0, 0};
    for (uint32_t j = 0; j < p->nv; j++) {
      if (j == i) continue;
      ld r[3];
      for (int a = 0; a < 3; a++) r[a] = ((ld)p->pos[3 * i + a] - (ld)p->pos[3 * j + a]) * gridUnit;
      C G[6];
      for (int e = 0; e < 6; e++) G[e] = green<ld>(r[0], r[1], r[2], waveNum, kMu[e], kNu[e]);
      acc[0] += G[0] * u[3 * j] + G[1] * u[3 * j + 1] + G[2] * u[3 * j + 2];
      acc[1] += G[1] * u[3 * j] + G[3] * u[3 * j + 1] + G[4] * u[3 * j + 2];
      acc[2] += G[2] * u[3 * j] + G[4] * u[3 * j + 1] + G[5] * u[3 * j + 2];
    }
    for (int c = 0; c < 3; c++) {
      C xv(x[2 * (i + c * S)], x[2 * (i + c * S) + 1]);
      if (conj) xv = std::conj(xv);
      C v = xv - ccs(i, c) * acc[c];
      if (conj) v = std::conj(v);
      y[2 * (i + c * S)] = (double)v.real();
      y[2 * (i + c * S) + 1] = (double)v.imag();
    }
  }
  return ST_OK;
}

// IterativeSolverBase::getPolVec (IterativeSolverBase.cpp:40-103) with the CPU
// solvers, in double. Returns a status code; pvec_out = sqrt(cc) x.
// resid_out (nullable, length max_iter + 1) receives residScale * |r|^2 after
// each iteration (index 0 = initial). max_iter = 0 selects the reference default
// 3 * Bx * By * Bz (DDAMain.cpp:640-644).
// max_res_increase = 0 selects the solver's default (QmrCs.cpp:39,
// BicgStab.cpp:51, Cgnr.cpp:34, BicgCs.cpp:48).
int orc_solve_ex(const orc_prob* p, int kind, const double* einc, double eps, uint64_t max_iter,
                 const double* start, double* pvec_out, uint64_t* iters_out, double* resid_out,
                 uint64_t max_iter_sample, uint64_t max_res_increase) {
  Prob g = toProb(p);
  typedef double T;
  typedef std::complex<T> C;
  size_t n = 3 * (size_t)g.vec_stride;
  if (max_iter == 0) max_iter = 3ull * g.box[0] * g.box[1] * g.box[2];
  Solver<T> s(g, kind);
  Vec<T> e(n);
  e.v = loadVec<T>(einc, n);
  Vec<T> st(n);
  if (start) st.v = loadVec<T>(start, n);
  T ip = s.initGeneral(e, start ? &st : nullptr);
  s.inprodR = ip;
  s.epsB = (T)(eps * eps / s.residScale);
  size_t count = 0, counter = 0;
  size_t maxResIncrease = max_res_increase ? max_res_increase : maxResIncreaseFor(kind);
  if (resid_out) resid_out[0] = s.residScale * ip;
  s.init();
  int status = ST_OK;
  while (s.inprodR >= s.epsB) {
    if (count >= max_iter) { status = ST_MAXITER; break; }
    if (counter > maxResIncrease) { status = ST_STAGNATION; break; }
    if (max_iter_sample && count >= max_iter_sample) break;  // bounded timing sample
    T ip1 = s.iteration(count);
    if (s.status != ST_OK) { status = s.status; break; }
    count++;
    if (ip1 <= s.inprodR) { s.inprodR = ip1; counter = 0; } else { counter++; }
    if (resid_out && count <= max_iter) resid_out[count] = s.residScale * ip1;
  }
  if (iters_out) *iters_out = count;
  // CpuIterativeSolver::getResult, CpuIterativeSolver.cpp:92-98
  std::memset(pvec_out, 0, sizeof(double) * 2 * n);
  for (uint32_t i = 0; i < g.nv; i++)
    for (int c = 0; c < 3; c++) {
      C v = cplx<T>(g.cc_sqrt, 3 * g.mat[i] + c) * s.x[i + c * g.vec_stride];
      pvec_out[2 * (i + c * g.vec_stride)] = v.real();
      pvec_out[2 * (i + c * g.vec_stride) + 1] = v.imag();
    }
  return status;
}

int orc_solve(const orc_prob* p, int kind, const double* einc, double eps, uint64_t max_iter,
              const double* start, double* pvec_out, uint64_t* iters_out, double* resid_out,
              uint64_t max_iter_sample) {
  return orc_solve_ex(p, kind, einc, eps, max_iter, start, pvec_out, iters_out, resid_out, max_iter_sample, 0);
}

// Cext = 4 pi k sum Im(p_i . conj(E_i)), Beam.cpp:58-68.
double orc_ext_cross_sum(uint32_t nv, uint32_t vec_stride, const double* pvec, const double* einc, double waveNum) {
  typedef std::complex<ld> C;
  ld sum = 0;
  for (uint32_t i = 0; i < nv; i++) {
    C acc = 0;
    for (int c = 0; c < 3; c++) {
      size_t o = 2 * (i + c * (size_t)vec_stride);
      acc += C(pvec[o], pvec[o + 1]) * std::conj(C(einc[o], einc[o + 1]));
    }
    sum += acc.imag();
  }
  return (double)(sum * 4 * kPi * waveNum);
}

// Cabs, AbsCross.cpp:27-48.
double orc_abs_cross(uint32_t nv, uint32_t vec_stride, const double* pvec, const uint8_t* mat, uint32_t nmat,
                     const double* cc, double waveNum) {
  typedef std::complex<ld> C;
  ld temp1 = std::pow((ld)waveNum, (ld)3) * 2 / 3;
  std::vector<ld> multDr(3 * nmat);
  for (uint32_t i = 0; i < nmat; i++)
    for (int j = 0; j < 3; j++) multDr[3 * i + j] = -(C(1) / C(cc[2 * (3 * i + j)], cc[2 * (3 * i + j) + 1])).imag() - temp1;
  ld sum = 0;
  for (uint32_t i = 0; i < nv; i++)
    for (int j = 0; j < 3; j++) {
      size_t o = 2 * (i + j * (size_t)vec_stride);
      sum += multDr[3 * mat[i] + j] * ((ld)pvec[o] * pvec[o] + (ld)pvec[o + 1] * pvec[o + 1]);
    }
  return (double)(4 * kPi * waveNum * sum);
}

// Scattered far field in direction n, CpuFieldCalculator.cpp:48-98.
void orc_calc_field(uint32_t nv, uint32_t vec_stride, const double* pvec, const uint32_t* pos, const uint32_t* box,
                    double gridUnit, double waveNum, const double* n, double* out) {
  typedef std::complex<ld> C;
  ld kd = (ld)waveNum * gridUnit;
  C sum[3] = {0, 0, 0};
  for (uint32_t j = 0; j < nv; j++) {
    ld arg = -kd * ((ld)n[0] * pos[3 * j] + (ld)n[1] * pos[3 * j + 1] + (ld)n[2] * pos[3 * j + 2]);
    C a(std::cos(arg), std::sin(arg));
    for (int c = 0; c < 3; c++) {
      size_t o = 2 * (j + c * (size_t)vec_stride);
      sum[c] += C(pvec[o], pvec[o + 1]) * a;
    }
  }
  C ns = (ld)n[0] * sum[0] + (ld)n[1] * sum[1] + (ld)n[2] * sum[2];
  ld origin[3];
  for (int a = 0; a < 3; a++) origin[a] = -((ld)box[a] - 1) / 2 * gridUnit;
  ld ph = -(ld)waveNum * (origin[0] * n[0] + origin[1] * n[1] + origin[2] * n[2]);
  C f = C(0, (ld)waveNum * waveNum) * C(std::cos(ph), std::sin(ph));
  for (int c = 0; c < 3; c++) {
    C v = f * (sum[c] - (ld)n[c] * ns);
    out[2 * c] = (double)v.real();
    out[2 * c + 1] = (double)v.imag();
  }
}

// Cext from the forward amplitude, Beam.cpp:125-133 (plane wave override).
double orc_ext_cross_forward(uint32_t nv, uint32_t vec_stride, const double* pvec, const uint32_t* pos,
                             const uint32_t* box, double gridUnit, double waveNum, const double* prop,
                             const double* pol) {
  double e[6];
  orc_calc_field(nv, vec_stride, pvec, pos, box, gridUnit, waveNum, prop, e);
  double sum = -(e[0] * pol[0] + e[2] * pol[1] + e[4] * pol[2]);
  return sum * 4 * (double)kPi / waveNum;
}

// Jones matrices on a (theta, phi) grid, FarFieldCalc.cpp:119-174 and
// JonesCalculus.cpp:74-86. p1 solves polarisation 1, p2 polarisation 2 (for a
// symmetric target the reference passes p1 with rotated axes; callers do that
// by passing the same vector and `symmetric` = 1; symmetric = 2 and 3 are
// mistakes that tests plant to show their checks reject them). out:
// [count][2][2] complex.
void orc_jones(uint32_t nv, uint32_t vec_stride, const double* p1, const double* p2, int symmetric,
               const uint32_t* pos, const uint32_t* box, double gridUnit, double waveNum, const double* prop,
               const double* pol1, const double* pol2, uint32_t ntheta, double theta0, double dtheta,
               uint32_t nphi, double phi0, double dphi, double* out) {
  typedef std::complex<double> C;
  size_t count = (size_t)ntheta * nphi;
  auto field = [&](const double* pv, const double* px, const double* py, bool neg_py, size_t i, C& perp, C& par) {
    double th = theta0 + (double)(i / nphi) * dtheta, ph = phi0 + (double)(i % nphi) * dphi;
    double st = std::sin(th), ct = std::cos(th), sp = std::sin(ph), cp = std::cos(ph);
    double X[3], Y[3];
    for (int a = 0; a < 3; a++) { X[a] = px[a]; Y[a] = neg_py ? -py[a] : py[a]; }
    double n[3], pe[3], pa[3];
    for (int a = 0; a < 3; a++) {
      n[a] = ct * prop[a] + st * (cp * X[a] + sp * Y[a]);
      pe[a] = sp * X[a] - cp * Y[a];
      pa[a] = -st * prop[a] + ct * (cp * X[a] + sp * Y[a]);
    }
    double e[6];
    orc_calc_field(nv, vec_stride, pv, pos, box, gridUnit, waveNum, n, e);
    perp = C(0), par = C(0);
    for (int a = 0; a < 3; a++) { perp += C(e[2 * a], e[2 * a + 1]) * pe[a]; par += C(e[2 * a], e[2 * a + 1]) * pa[a]; }
  };
  for (size_t i = 0; i < count; i++) {
    C perp1, par1, perp2, par2;
    // eField1: incPolX = pol2, incPolY = pol1 (FarFieldCalc.cpp:165)
    field(p1, pol2, pol1, false, i, perp1, par1);
    if (symmetric == 1)  // FarFieldCalc.cpp:167-168: X = pol1, Y = -pol2
      field(p1, pol1, pol2, true, i, perp2, par2);
    else if (symmetric == 2)  // a mistake planted by tests: the minus on pol2 dropped
      field(p1, pol1, pol2, false, i, perp2, par2);
    else if (symmetric == 3)  // a mistake planted by tests: the rotation's axes swapped
      field(p1, pol2, pol1, true, i, perp2, par2);
    else
      field(p2, pol2, pol1, false, i, perp2, par2);
    double ph = phi0 + (double)(i % nphi) * dphi;
    double sp = std::sin(ph), cp = std::cos(ph);
    C J[4] = {cp * par2 + sp * par1, sp * par2 - cp * par1, cp * perp2 + sp * perp1, sp * perp2 - cp * perp1};
    for (int e = 0; e < 4; e++) { out[8 * i + 2 * e] = J[e].real(); out[8 * i + 2 * e + 1] = J[e].imag(); }
  }
}

// Bohren-Huffman Mie series, EMSim/Mie.cpp:40-174 (restated from the textbook
// algorithm). Returns Qext, Qsca, Qback, g.
void orc_mie(double x, double m_re, double m_im, double* qext, double* qsca, double* qback, double* gsca) {
  typedef std::complex<ld> C;
  C refrel(m_re, m_im);
  C cxy = (ld)x * refrel;
  ld xstop = x + 4 * std::pow((ld)x, (ld)1 / 3) + 2;
  size_t nstop = (size_t)xstop;
  ld ymod = std::abs(cxy);
  size_t nmx = (size_t)(std::max(xstop, ymod) + 15);
  std::vector<C> cxd(nmx);
  cxd[nmx - 1] = 0;
  for (size_t i = nmx - 1; i > 0; i--) cxd[i - 1] = (ld)(i + 1) / cxy - C(1) / (cxd[i] + (ld)(i + 1) / cxy);
  ld amu[2] = {1, -1};
  C s1[2] = {0, 0};
  ld pi0[2] = {0, 0}, pi1[2] = {1, 1}, pi[2], tau[2];
  ld psi0 = std::cos((ld)x), psi1 = std::sin((ld)x), chi0 = -std::sin((ld)x), chi1 = std::cos((ld)x);
  ld apsi1 = psi1;
  C cxxi1(apsi1, -chi1);
  ld qs = 0, gs = 0;
  C an, bn, an1, bn1;
  for (size_t n = 0; n < nstop; n++) {
    ld rn = n + 1;
    ld fn = (2 * rn + 1) / (rn * (rn + 1));
    ld psi = (2 * rn - 1) * psi1 / x - psi0;
    ld chi = (2 * rn - 1) * chi1 / x - chi0;
    C cxxi(psi, -chi);
    if (n > 0) { an1 = an; bn1 = bn; }
    an = ((cxd[n] / refrel + rn / x) * psi - apsi1) / ((cxd[n] / refrel + rn / x) * cxxi - cxxi1);
    bn = ((refrel * cxd[n] + rn / x) * psi - apsi1) / ((refrel * cxd[n] + rn / x) * cxxi - cxxi1);
    qs += (2 * rn + 1) * (std::norm(an) + std::norm(bn));
    gs += ((2 * rn + 1) / (rn * (rn + 1))) * std::real(an * std::conj(bn));
    if (n > 0) gs += ((rn - 1) * (rn + 1) / rn) * std::real(an1 * std::conj(an) + bn1 * std::conj(bn));
    for (int i = 0; i < 2; i++) {
      pi[i] = pi1[i];
      tau[i] = rn * amu[i] * pi[i] - (rn + 1) * pi0[i];
      s1[i] += fn * (an * pi[i] + bn * tau[i]);
    }
    psi0 = psi1; psi1 = psi; apsi1 = psi1; chi0 = chi1; chi1 = chi;
    cxxi1 = C(apsi1, -chi1);
    for (int i = 0; i < 2; i++) {
      pi1[i] = ((2 * rn + 1) * amu[i] * pi[i] - (rn + 1) * pi0[i]) / rn;
      pi0[i] = pi[i];
    }
  }
  *gsca = (double)(2 * gs / qs);
  *qsca = (double)((2 / ((ld)x * x)) * qs);
  *qext = (double)((4 / ((ld)x * x)) * s1[0].real());
  *qback = (double)((4 / ((ld)x * x)) * std::norm(s1[1]));
}

}  // extern "C"
