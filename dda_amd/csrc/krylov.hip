// Device-resident Krylov solvers: QMR-CS, BiCGStab, CGNR and BiCG-CS, for one
// or a batch of right-hand sides.
//
// Replaces IterativeSolverBase::getPolVec (DDA/IterativeSolverBase.cpp:40-103),
// GpuIterativeSolver::initGeneral/getResult (DDA/GpuIterativeSolver.cpp:78-144),
// the update loops of DDA/GpuQmrCs.cpp:77-198, DDA/GpuBicgStab.cpp:44-125,
// DDA/GpuCgnr.cpp:41-81, DDA/GpuBicgCs.cpp:41-95 and the linear combinations and
// reductions of LinAlg/GpuLinComb.cpp. The arithmetic of each iteration follows
// the reference's CPU solvers line by line (DDA/QmrCs.cpp:78-168,
// DDA/BicgStab.cpp:66-112, DDA/Cgnr.cpp:43-73, DDA/BicgCs.cpp:58-84).
//
// Every scalar lives in device memory (one SolverState per right-hand side,
// FP64). Reductions are two-stage with a fixed number of blocks per
// right-hand side that depends only on the vector length, so a right-hand
// side's result is bitwise the same alone and in any batch. Scalar algebra runs
// in a one-block-per-RHS kernel. The loop test, the maximum-iteration and
// stagnation aborts (IterativeSolverBase.cpp:56-70) and the breakdown tests are
// evaluated on the device: once a right-hand side's `done` is set every later
// kernel skips it (its matvec included), so the host enqueues iterations ahead
// and reads the states once per chunk of iterations.
//
// Vector kernels run on a grid (nb, nrhs): blockIdx.y is the right-hand side,
// vectors are [nrhs][n] back to back.

#include <chrono>
#include <thread>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "internal.hpp"

namespace dda {

constexpr int kRedThreads = 256;
constexpr int kMaxRedBlocks = 1024;
constexpr int kChunk = 16;  // most iterations enqueued between state reads

// Reduction blocks per right-hand side for vectors of n complex values.
inline int red_blocks(long long n) {
  long long b = (n + 8 * kRedThreads - 1) / (8 * kRedThreads);
  return (int)(b < 1 ? 1 : (b > kMaxRedBlocks ? kMaxRedBlocks : b));
}

struct SolverState {
  int done;     // 1 once the solve has finished for any reason
  int status;   // dda_status at exit
  int skip;     // done, or BiCGStab's early exit inside the current iteration
  int sel;      // initGeneral: 1 = start from x = 0
  unsigned long long count, counter, max_iter, max_res_increase, hist_cap;
  double eps, residScale, epsB, inprodR, inprodR_init, ip1, temp;
  // QMR-CS (DDA/QmrCs.hpp VarType)
  double omega_old, omega_new, c_old, c_new;
  double2 beta, mBeta, tautilda, s_old, s_new;
  double2 alpha, mAlpha, zetaInv, mEtaZeta, mThetaZeta, betaInv, tau, cnot, normSNew;
  // BiCGStab / CGNR / BiCG-CS
  double2 ro_new, ro_old, omega, bbeta, mbo, mOmega, one, mOne;
  double denumOmega, abs_ro_new, cg_ro_new, cg_ro_old;
  // plateau stop of the mixed mode's inner solves (0 = off): STAGNATION once
  // the smallest |r|^2 seen has not fallen below plateau_ratio times its value
  // plateau_window iterations earlier
  unsigned long long plateau_window, ckpt_count;
  double plateau_ratio, ckpt_r;
};
constexpr int kStateInts = sizeof(SolverState) / sizeof(int);
static_assert(sizeof(SolverState) % sizeof(int) == 0, "state stride");

__device__ inline double2 dmul(double2 a, double2 b) { return cmul(a, b); }
__device__ inline double2 ddiv(double2 a, double2 b) {
  // a / b with scaling against overflow
  const double s = fabs(b.x) + fabs(b.y);
  const double br = b.x / s, bi = b.y / s;
  const double den = br * br + bi * bi;
  const double ar = a.x / s, ai = a.y / s;
  return make_double2((ar * br + ai * bi) / den, (ai * br - ar * bi) / den);
}
__device__ inline double dabs(double2 a) { return hypot(a.x, a.y); }
__device__ inline double dnorm(double2 a) { return a.x * a.x + a.y * a.y; }
// principal square root (std::sqrt of std::complex)
__device__ inline double2 dsqrt(double2 z) {
  const double r = dabs(z);
  if (r == 0.0) return make_double2(0.0, 0.0);
  if (z.x >= 0.0) {
    const double t = sqrt(0.5 * (r + z.x));
    return make_double2(t, z.y / (2.0 * t));
  }
  const double t = sqrt(0.5 * (r - z.x));
  return make_double2(fabs(z.y) / (2.0 * t), z.y >= 0.0 ? t : -t);
}
__device__ inline double2 dre(double v) { return make_double2(v, 0.0); }
__device__ inline double2 dneg(double2 a) { return make_double2(-a.x, -a.y); }

// wave64 + LDS block sum; every thread returns the block total.
__device__ inline double block_sum(double v) {
  __shared__ double red[kRedThreads / 64];
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  const int w = threadIdx.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[w] = v;
  __syncthreads();
  double s = 0;
  for (int i = 0; i < (int)(blockDim.x >> 6); i++) s += red[i];
  return s;
}

// Per-block partial sums of up to three quantities; quantity q of block b of
// right-hand side r sits at part[(r * 4 + q) * nb + b].
__device__ inline void write_partials(double* part, int nq, double q0, double q1, double q2) {
  const int nb = gridDim.x;
  double* p = part + (long long)blockIdx.y * 4 * nb;
  const double t0 = block_sum(q0);
  const double t1 = nq > 1 ? block_sum(q1) : 0.0;
  const double t2 = nq > 2 ? block_sum(q2) : 0.0;
  if (threadIdx.x == 0) {
    p[blockIdx.x] = t0;
    if (nq > 1) p[nb + blockIdx.x] = t1;
    if (nq > 2) p[2 * nb + blockIdx.x] = t2;
  }
}

// Second stage, run by a k_scalar block (kScalarThreads) for one right-hand
// side: quantities 0 .. nq-1 of its partials, each in a fixed order (thread t
// adds entries t + j B, B = kScalarThreads, into chain j mod 4; then the
// chains, the lanes of a wave and the waves in a fixed order), with up to
// 4 nq loads in flight per thread.
constexpr int kScalarThreads = 1024;
__device__ inline double3 sum_partials(const double* __restrict__ part_rhs, int nb, int nq) {
  constexpr int B = kScalarThreads;
  double a0[4] = {0, 0, 0, 0}, a1[4] = {0, 0, 0, 0}, a2[4] = {0, 0, 0, 0};
  const double* p0 = part_rhs;
  const double* p1 = part_rhs + nb;
  const double* p2 = part_rhs + 2 * (long long)nb;
  int i = threadIdx.x;
  for (; i + 3 * B < nb; i += 4 * B) {
#pragma unroll
    for (int u = 0; u < 4; u++) {
      a0[u] += p0[i + u * B];
      if (nq > 1) a1[u] += p1[i + u * B];
      if (nq > 2) a2[u] += p2[i + u * B];
    }
  }
  for (; i < nb; i += B) {
    a0[0] += p0[i];
    if (nq > 1) a1[0] += p1[i];
    if (nq > 2) a2[0] += p2[i];
  }
  double s0 = (a0[0] + a0[1]) + (a0[2] + a0[3]);
  double s1 = (a1[0] + a1[1]) + (a1[2] + a1[3]);
  double s2 = (a2[0] + a2[1]) + (a2[2] + a2[3]);
  for (int off = 32; off > 0; off >>= 1) {
    s0 += __shfl_xor(s0, off, 64);
    s1 += __shfl_xor(s1, off, 64);
    s2 += __shfl_xor(s2, off, 64);
  }
  __shared__ double red[3][B / 64];
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    red[0][w] = s0;
    red[1][w] = s1;
    red[2][w] = s2;
  }
  __syncthreads();
  double3 r = make_double3(0.0, 0.0, 0.0);
  for (int k = 0; k < B / 64; k++) {
    r.x += red[0][k];
    r.y += red[1][k];
    r.z += red[2][k];
  }
  return r;
}

template <typename C> __device__ inline C coef(const double2* p) { return cmake<C>(p->x, p->y); }

#define DDA_GRID_LOOP(i) \
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)

// Per-RHS set-up of a vector kernel: the state and the vectors' offsets.
#define DDA_RHS_PROLOGUE(stp) \
  const int rhs_ = blockIdx.y; \
  const long long off_ = (long long)rhs_ * n; \
  (void)off_; \
  const SolverState* s_ = (stp) + rhs_; \
  (void)s_;

// ------------------------------------------------------------------------------------
// Vector kernels
// ------------------------------------------------------------------------------------

enum RedOp { RED_DOTU = 0, RED_DOTC = 1, RED_NORM = 2 };

// Partial sums of a.b (DOTU), a.conj(b) (DOTC) or |a|^2 (NORM). flag_off selects
// the state field that masks the kernel (offset of done or skip in ints).
template <typename R, int OP>
__global__ __launch_bounds__(kRedThreads) void k_reduce(const typename CxT<R>::T* __restrict__ a,
                                                        const typename CxT<R>::T* __restrict__ b, long long n,
                                                        double* part, const SolverState* st, int flag_off) {
  DDA_RHS_PROLOGUE(st)
  if (reinterpret_cast<const int*>(s_)[flag_off]) return;
  a += off_;
  b += off_;
  double s0 = 0, s1 = 0;
  DDA_GRID_LOOP(i) {
    const double ar = a[i].x, ai = a[i].y;
    if (OP == RED_NORM) {
      s0 += ar * ar + ai * ai;
    } else {
      const double br = b[i].x, bi = OP == RED_DOTC ? -(double)b[i].y : (double)b[i].y;
      s0 += ar * br - ai * bi;
      s1 += ar * bi + ai * br;
    }
  }
  write_partials(part, OP == RED_NORM ? 1 : 2, s0, s1, 0);
}

// out = c1 v1 + c2 v2 + v3, coefficients are state fields (byte offsets, -1 =
// term absent; v3 may be null), optional |out|^2 partial.
template <typename R>
__global__ __launch_bounds__(kRedThreads) void k_lincomb(typename CxT<R>::T* out, const typename CxT<R>::T* v1, int c1,
                                                         const typename CxT<R>::T* v2, int c2,
                                                         const typename CxT<R>::T* v3, long long n, double* part,
                                                         const SolverState* st, int flag_off) {
  typedef typename CxT<R>::T C;
  DDA_RHS_PROLOGUE(st)
  if (reinterpret_cast<const int*>(s_)[flag_off]) return;
  const char* sb = reinterpret_cast<const char*>(s_);
  const C k1 = c1 >= 0 ? coef<C>(reinterpret_cast<const double2*>(sb + c1)) : czero<C>();
  const C k2 = c2 >= 0 ? coef<C>(reinterpret_cast<const double2*>(sb + c2)) : czero<C>();
  out += off_;
  if (v1) v1 += off_;
  if (v2) v2 += off_;
  if (v3) v3 += off_;
  double s = 0;
  DDA_GRID_LOOP(i) {
    C o = czero<C>();
    if (c1 >= 0) o = cmul(v1[i], k1);
    if (c2 >= 0) o = cadd(o, cmul(v2[i], k2));
    if (v3) o = cadd(o, v3[i]);
    out[i] = o;
    s += (double)o.x * o.x + (double)o.y * o.y;
  }
  if (part) write_partials(part, 1, s, 0, 0);
}

// QMR-CS: vtilda = [vtilda * mBeta] + v * mAlpha + Av, with partials of
// vtilda.vtilda (quantities 0, 1) and |vtilda|^2 (2). QmrCs.cpp:96-104.
template <typename R>
__global__ __launch_bounds__(kRedThreads) void k_qmr_vt(typename CxT<R>::T* vt, const typename CxT<R>::T* __restrict__ v,
                                                        const typename CxT<R>::T* __restrict__ av, int first,
                                                        const SolverState* st, long long n, double* part) {
  typedef typename CxT<R>::T C;
  DDA_RHS_PROLOGUE(st)
  if (s_->done) return;
  vt += off_;
  v += off_;
  av += off_;
  const C mb = coef<C>(&s_->mBeta), ma = coef<C>(&s_->mAlpha);
  double s0 = 0, s1 = 0, s2 = 0;
  DDA_GRID_LOOP(i) {
    C o = cadd(cmul(v[i], ma), av[i]);
    if (!first) o = cadd(cmul(vt[i], mb), o);
    vt[i] = o;
    const double r = o.x, im = o.y;
    s0 += r * r - im * im;
    s1 += 2 * r * im;
    s2 += r * r + im * im;
  }
  write_partials(part, 3, s0, s1, s2);
}

// QMR-CS update, QmrCs.cpp:150-166: p_dst = [p_old mThetaZeta] + [p_new mEtaZeta]
// + v zetaInv; x += p_dst tau; vt *= betaInv; r = r normSNew + vt cnot.
// mode 0: nr == 0, 1: nr == 1, 2: later. Partial |r|^2.
template <typename R>
__global__ __launch_bounds__(kRedThreads) void k_qmr_upd(typename CxT<R>::T* p_dst, const typename CxT<R>::T* p_new,
                                                         const typename CxT<R>::T* __restrict__ v,
                                                         typename CxT<R>::T* vt, typename CxT<R>::T* x,
                                                         typename CxT<R>::T* r, int mode, const SolverState* st,
                                                         long long n, double* part) {
  typedef typename CxT<R>::T C;
  DDA_RHS_PROLOGUE(st)
  if (s_->done) return;
  p_dst += off_;
  p_new += off_;
  v += off_;
  vt += off_;
  x += off_;
  r += off_;
  const C zi = coef<C>(&s_->zetaInv), mez = coef<C>(&s_->mEtaZeta), mtz = coef<C>(&s_->mThetaZeta);
  const C tau = coef<C>(&s_->tau), bi = coef<C>(&s_->betaInv), cn = coef<C>(&s_->cnot), ns = coef<C>(&s_->normSNew);
  double s = 0;
  DDA_GRID_LOOP(i) {
    C p = cmul(v[i], zi);
    if (mode >= 1) p = cadd(cmul(p_new[i], mez), p);
    if (mode >= 2) p = cadd(cmul(p_dst[i], mtz), p);
    p_dst[i] = p;
    x[i] = cadd(cmul(p, tau), x[i]);
    const C vs = cmul(vt[i], bi);
    vt[i] = vs;
    const C rn = cadd(cmul(r[i], ns), cmul(vs, cn));
    r[i] = rn;
    s += (double)rn.x * rn.x + (double)rn.y * rn.y;
  }
  write_partials(part, 1, s, 0, 0);
}

// b = sqrt(a) einc (GpuIterativeSolverInst.cl setupPvec) with each RHS's couple
// constants; einc is FP64 [nrhs][3][S].
template <typename R>
__global__ __launch_bounds__(kRedThreads) void k_setup_b(typename CxT<R>::T* b, const double2* __restrict__ einc,
                                                         const typename CxT<R>::T* __restrict__ ccs,
                                                         const uint8_t* __restrict__ mat, int nv, int S, int nmat,
                                                         const int* member) {
  typedef typename CxT<R>::T C;
  const long long n = 3ll * S;
  const int rhs = blockIdx.y;
  b += rhs * n;
  einc += rhs * n;
  ccs += rhs * 3 * nmat;
  if (member) mat += (size_t)member[rhs] * nv;
  DDA_GRID_LOOP(t) {
    const int c = (int)(t / S), i = (int)(t - (long long)c * S);
    C o = czero<C>();
    if (i < nv) o = cmul(ccs[3 * mat[i] + c], cmake<C>(einc[t].x, einc[t].y));
    b[t] = o;
  }
}

// x = start / sqrt(a) (setupXvecStartValue), start given as p in d^3 units.
template <typename R>
__global__ void k_setup_start(typename CxT<R>::T* x, const double2* __restrict__ start,
                              const typename CxT<R>::T* __restrict__ ccs, const uint8_t* __restrict__ mat, int nv,
                              int S, int nmat, double scale, const int* member) {
  typedef typename CxT<R>::T C;
  const long long n = 3ll * S;
  const int rhs = blockIdx.y;
  x += rhs * n;
  start += rhs * n;
  ccs += rhs * 3 * nmat;
  if (member) mat += (size_t)member[rhs] * nv;
  DDA_GRID_LOOP(t) {
    const int c = (int)(t / S), i = (int)(t - (long long)c * S);
    double2 o = make_double2(0, 0);
    if (i < nv) {
      const C a = ccs[3 * mat[i] + c];
      o = ddiv(make_double2(start[t].x * scale, start[t].y * scale), make_double2(a.x, a.y));
    }
    x[t] = cmake<C>(o.x, o.y);
  }
}

// p = sqrt(a) x * scale into FP64 output (getResult).
template <typename R>
__global__ void k_get_result(double2* out, const typename CxT<R>::T* __restrict__ x,
                             const typename CxT<R>::T* __restrict__ ccs, const uint8_t* __restrict__ mat, int nv, int S,
                             int nmat, double scale, const int* member) {
  typedef typename CxT<R>::T C;
  const long long n = 3ll * S;
  const int rhs = blockIdx.y;
  out += rhs * n;
  x += rhs * n;
  ccs += rhs * 3 * nmat;
  if (member) mat += (size_t)member[rhs] * nv;
  DDA_GRID_LOOP(t) {
    const int c = (int)(t / S), i = (int)(t - (long long)c * S);
    double2 o = make_double2(0, 0);
    if (i < nv) {
      const C a = ccs[3 * mat[i] + c];
      const double2 p = dmul(make_double2(a.x, a.y), make_double2(x[t].x, x[t].y));
      o = make_double2(p.x * scale, p.y * scale);
    }
    out[t] = o;
  }
}

// initGeneral selection, CpuIterativeSolver.cpp:73-84: sel -> x = 0, r = b;
// else x = b (r already holds b - A b).
template <typename R>
__global__ void k_select_x0(typename CxT<R>::T* x, typename CxT<R>::T* r, const typename CxT<R>::T* __restrict__ b,
                            long long n, const SolverState* st) {
  typedef typename CxT<R>::T C;
  DDA_RHS_PROLOGUE(st)
  if (s_->done) return;
  x += off_;
  r += off_;
  b += off_;
  const int sel = s_->sel;
  DDA_GRID_LOOP(i) {
    if (sel) {
      x[i] = czero<C>();
      r[i] = b[i];
    } else {
      x[i] = b[i];
    }
  }
}

template <typename R>
__global__ void k_copy_flag(typename CxT<R>::T* dst, const typename CxT<R>::T* __restrict__ src, long long n,
                            const SolverState* st) {
  DDA_RHS_PROLOGUE(st)
  if (s_->skip) return;
  dst += off_;
  src += off_;
  DDA_GRID_LOOP(i) dst[i] = src[i];
}

// ------------------------------------------------------------------------------------
// Scalar kernel: one block per right-hand side
// ------------------------------------------------------------------------------------

enum Step {
  S_INIT_TEMP = 0,   // temp = |b|^2
  S_INIT_R = 1,      // inprodR from |b - A b|^2 and the x0 choice
  S_INIT_START = 2,  // inprodR from |b - A x0|^2 (start vector given)
  S_LOOP_CHECK = 3,  // the loop head of getPolVec before the first iteration
  S_QMR_INIT = 10,
  S_QMR_ALPHA = 11,
  S_QMR_GIVENS = 12,
  S_END_ITER = 13,   // LoopUpdate + loop head (all solvers)
  S_BS_RO = 20,
  S_BS_ALPHA = 21,
  S_BS_S = 22,
  S_BS_OMEGA = 23,
  S_BS_END = 24,
  S_BS_RO1 = 25,     // S_BS_RO with r.conj(rtilda) from the previous x/r update (partials 1, 2)
  S_CG_RO = 30,
  S_CG_ALPHA = 31,
  S_BI_RO = 40,
  S_BI_MU = 41,
};

__device__ inline void finish(SolverState* st, int status) {
  st->status = status;
  st->done = 1;
  st->skip = 1;
}

// Loop head of IterativeSolverBase::getPolVec (:56-61) plus the breakdown test
// at the start of the next QMR-CS iteration (QmrCs.cpp:89-93).
__device__ inline void loop_head(SolverState* st, int kind) {
  if (!(st->inprodR >= st->epsB)) {
    finish(st, DDA_OK);
    return;
  }
  if (st->count >= st->max_iter) {
    finish(st, DDA_MAX_ITER);
    return;
  }
  if (st->counter > st->max_res_increase) {
    finish(st, DDA_STAGNATION);
    return;
  }
  if (kind == DDA_SOLVER_QMR_CS) {
    const double rtmp1 = dnorm(st->beta) * st->residScale;
    if (st->count == 0) {
      if (rtmp1 > 1e+38f) finish(st, DDA_BREAKDOWN);
    } else if (rtmp1 < 1e-10f || rtmp1 > 1e+38f) {
      finish(st, DDA_BREAKDOWN);
    }
  }
  st->skip = st->done;
}

__global__ __launch_bounds__(kScalarThreads) void k_scalar(SolverState* st_all, const double* part_all, int nb, int step,
                                                        int kind, double* hist_all) {
  const int rhs = blockIdx.x;
  SolverState* st = st_all + rhs;
  const double* part = part_all + (long long)rhs * 4 * nb;
  double* hist = hist_all ? hist_all + (long long)rhs * st->hist_cap : nullptr;
  if (st->done) return;  // finished, or masked out of this solve
  // All threads take part in the reductions; thread 0 does the algebra.
  int nq = 0;
  switch (step) {
    case S_INIT_TEMP:
    case S_INIT_R:
    case S_INIT_START:
    case S_END_ITER:
    case S_CG_RO:
    case S_CG_ALPHA:
    case S_BS_S:
      nq = 1;
      break;
    case S_QMR_INIT:
    case S_QMR_ALPHA:
    case S_BS_RO:
    case S_BS_ALPHA:
    case S_BI_RO:
    case S_BI_MU:
      nq = 2;
      break;
    case S_QMR_GIVENS:
    case S_BS_OMEGA:
    case S_BS_RO1:
      nq = 3;
      break;
    default:
      break;
  }
  double q0 = 0, q1 = 0, q2 = 0;
  if (nq > 0) {
    const double3 t = sum_partials(part, nb, nq);
    q0 = t.x;
    q1 = t.y;
    q2 = t.z;
  }
  if (threadIdx.x != 0) return;
  switch (step) {
    case S_INIT_TEMP:
      st->temp = q0;
      if (!(q0 > 0)) {  // a zero right-hand side (residScale would be infinite)
        finish(st, DDA_INVALID_ARG);
        break;
      }
      st->residScale = 1.0 / q0;
      st->epsB = st->eps * st->eps / st->residScale;
      st->one = dre(1.0);
      st->mOne = dre(-1.0);
      break;
    case S_INIT_R:
      if (st->temp < q0) {
        st->sel = 1;
        st->inprodR = st->temp;
      } else {
        st->sel = 0;
        st->inprodR = q0;
      }
      st->inprodR_init = st->inprodR;
      if (hist && st->hist_cap > 0) hist[0] = st->residScale * st->inprodR;
      break;
    case S_INIT_START:
      st->sel = 0;
      st->inprodR = q0;
      st->inprodR_init = q0;
      if (hist && st->hist_cap > 0) hist[0] = st->residScale * st->inprodR;
      break;
    case S_QMR_INIT: {  // QmrCs.cpp:58-76
      const double2 rvec2 = make_double2(q0, q1);
      st->omega_old = 0;
      st->beta = dsqrt(rvec2);
      st->mBeta = dneg(st->beta);
      st->omega_new = sqrt(st->inprodR) / dabs(st->beta);
      st->tautilda = make_double2(st->omega_new * st->beta.x, st->omega_new * st->beta.y);
      st->c_old = 1;
      st->c_new = 1;
      st->s_old = dre(0);
      st->s_new = dre(0);
      st->betaInv = ddiv(dre(1.0), st->beta);
      break;
    }
    case S_LOOP_CHECK:
      loop_head(st, kind);
      break;
    case S_QMR_ALPHA:
      st->alpha = make_double2(q0, q1);
      st->mAlpha = dneg(st->alpha);
      break;
    case S_QMR_GIVENS: {  // QmrCs.cpp:106-148
      const double2 ctmp3 = make_double2(q0, q1);
      const double rtmp2 = q2;
      const double2 ctmp1 = make_double2(st->omega_old * st->beta.x, st->omega_old * st->beta.y);
      const double2 ctmp2 = make_double2(st->omega_new * st->alpha.x, st->omega_new * st->alpha.y);
      const double2 theta = dmul(make_double2(st->s_old.x, -st->s_old.y), ctmp1);
      const double2 eta = cadd(make_double2(st->c_old * st->c_new * ctmp1.x, st->c_old * st->c_new * ctmp1.y),
                               dmul(make_double2(st->s_new.x, -st->s_new.y), ctmp2));
      const double2 zetatilda = csub(make_double2(st->c_new * ctmp2.x, st->c_new * ctmp2.y),
                                     dmul(make_double2(st->c_old * st->s_new.x, st->c_old * st->s_new.y), ctmp1));
      st->beta = dsqrt(ctmp3);
      st->mBeta = dneg(st->beta);
      st->omega_old = st->omega_new;
      st->omega_new = sqrt(rtmp2) / dabs(st->beta);
      const double zetaabs = sqrt(dnorm(zetatilda) + rtmp2);
      const double rtmp3 = sqrt(dnorm(zetatilda));
      double2 zeta;
      if (rtmp3 < 1e-40f)
        zeta = dre(zetaabs);
      else
        zeta = make_double2(zetaabs / rtmp3 * zetatilda.x, zetaabs / rtmp3 * zetatilda.y);
      st->c_old = st->c_new;
      st->c_new = rtmp3 / zetaabs;
      st->s_old = st->s_new;
      const double2 bz = ddiv(st->beta, zeta);
      st->s_new = make_double2(st->omega_new * bz.x, st->omega_new * bz.y);
      st->tau = make_double2(st->c_new * st->tautilda.x, st->c_new * st->tautilda.y);
      st->tautilda = dneg(dmul(st->s_new, st->tautilda));
      st->zetaInv = ddiv(dre(1.0), zeta);
      st->mEtaZeta = dneg(ddiv(eta, zeta));
      st->mThetaZeta = dneg(ddiv(theta, zeta));
      st->betaInv = ddiv(dre(1.0), st->beta);
      st->normSNew = dre(dnorm(st->s_new));
      const double cw = st->c_new / st->omega_new;
      st->cnot = make_double2(cw * st->tautilda.x, cw * st->tautilda.y);
      break;
    }
    case S_END_ITER: {  // IterativeSolverBase.cpp:62-70 then the loop head
      if (kind == DDA_SOLVER_BICGSTAB && !st->skip) {  // |r|^2 of the x/r update
        st->ip1 = q0;
        st->ro_old = st->ro_new;
      }
      const double ip1 = (kind == DDA_SOLVER_BICGSTAB) ? st->ip1 : q0;
      st->count++;
      if (ip1 <= st->inprodR) {
        st->inprodR = ip1;
        st->counter = 0;
      } else {
        st->counter++;
      }
      if (hist && st->count < st->hist_cap) hist[st->count] = st->residScale * ip1;
      loop_head(st, kind);
      if (!st->done && st->plateau_window) {
        if (st->ckpt_count == 0 && st->ckpt_r == 0.0) {
          st->ckpt_r = st->inprodR;
          st->ckpt_count = st->count;
        } else if (st->count - st->ckpt_count >= st->plateau_window) {
          if (st->inprodR > st->plateau_ratio * st->ckpt_r) {
            finish(st, DDA_STAGNATION);
          } else {
            st->ckpt_r = st->inprodR;
            st->ckpt_count = st->count;
          }
        }
      }
      break;
    }
    // ---- BiCGStab, BicgStab.cpp:66-112 ----
    case S_BS_RO:
    case S_BS_RO1: {
      st->ro_new = step == S_BS_RO1 ? make_double2(q1, q2) : make_double2(q0, q1);
      const double dtmp = dabs(st->ro_new) / st->inprodR;
      if (!(dtmp >= 1e-16)) {
        finish(st, DDA_BREAKDOWN);
        break;
      }
      if (st->count > 0) {
        const double2 ro_new_alpha = dmul(st->ro_new, st->alpha);
        const double2 ro_old_omega = dmul(st->ro_old, st->omega);
        if (!(dabs(ro_old_omega) / dabs(ro_new_alpha) >= 10e-10)) {
          finish(st, DDA_BREAKDOWN);
          break;
        }
        st->bbeta = ddiv(ro_new_alpha, ro_old_omega);
        st->mbo = dneg(dmul(st->bbeta, st->omega));
      }
      break;
    }
    case S_BS_ALPHA:
      st->alpha = ddiv(st->ro_new, make_double2(q0, q1));
      st->mAlpha = dneg(st->alpha);
      break;
    case S_BS_S:
      st->ip1 = q0;
      if (q0 < st->epsB) st->skip = 1;  // early exit: x += alpha p only
      break;
    case S_BS_OMEGA:
      st->denumOmega = q0;
      st->omega = make_double2(q1 / q0, q2 / q0);
      st->mOmega = dneg(st->omega);
      break;
    case S_BS_END:
      break;
    // ---- CGNR, Cgnr.cpp:43-73 ----
    case S_CG_RO:
      st->cg_ro_new = q0;
      if (st->count > 0) st->bbeta = dre(st->cg_ro_new / st->cg_ro_old);
      break;
    case S_CG_ALPHA:
      st->alpha = dre(st->cg_ro_new / q0);
      st->mAlpha = dneg(st->alpha);
      st->cg_ro_old = st->cg_ro_new;
      break;
    // ---- BiCG-CS, BicgCs.cpp:58-84 ----
    case S_BI_RO: {
      st->ro_new = make_double2(q0, q1);
      st->abs_ro_new = dabs(st->ro_new);
      const double dtmp = st->abs_ro_new / st->inprodR;
      if (dtmp < 1e-10 || dtmp > 1e+10) {
        finish(st, DDA_BREAKDOWN);
        break;
      }
      if (st->count > 0) st->bbeta = ddiv(st->ro_new, st->ro_old);
      break;
    }
    case S_BI_MU: {
      const double2 mu = make_double2(q0, q1);
      const double dtmp2 = dabs(mu) / st->abs_ro_new;
      if (dtmp2 < 10e-10) {
        finish(st, DDA_BREAKDOWN);
        break;
      }
      st->alpha = ddiv(st->ro_new, mu);
      st->mAlpha = dneg(st->alpha);
      st->ro_old = st->ro_new;
      break;
    }
    default:
      break;
  }
}

// BiCGStab: p = p beta + v (-beta omega) + r (nr > 0) or p = r.
template <typename R>
__global__ __launch_bounds__(kRedThreads) void k_bs_p(typename CxT<R>::T* p, const typename CxT<R>::T* __restrict__ v,
                                                      const typename CxT<R>::T* __restrict__ r, int first,
                                                      const SolverState* st, long long n) {
  typedef typename CxT<R>::T C;
  DDA_RHS_PROLOGUE(st)
  if (s_->done) return;
  p += off_;
  v += off_;
  r += off_;
  const C b = coef<C>(&s_->bbeta), mbo = coef<C>(&s_->mbo);
  DDA_GRID_LOOP(i) {
    if (first)
      p[i] = r[i];
    else
      p[i] = cadd(cadd(cmul(p[i], b), cmul(v[i], mbo)), r[i]);
  }
}

// BiCGStab: partials |Av|^2 (0) and s.conj(Av) (1, 2).
template <typename R>
__global__ __launch_bounds__(kRedThreads) void k_bs_om(const typename CxT<R>::T* __restrict__ av,
                                                       const typename CxT<R>::T* __restrict__ s, long long n,
                                                       double* part, const SolverState* st) {
  DDA_RHS_PROLOGUE(st)
  if (s_->skip) return;
  av += off_;
  s += off_;
  double n2 = 0, dr = 0, di = 0;
  DDA_GRID_LOOP(i) {
    const double ar = av[i].x, ai = av[i].y, sr = s[i].x, si = s[i].y;
    n2 += ar * ar + ai * ai;
    dr += sr * ar + si * ai;
    di += si * ar - sr * ai;
  }
  write_partials(part, 3, n2, dr, di);
}

// BiCGStab: x = p alpha + s omega + x; r = Av (-omega) + s; partials |r|^2
// and r.conj(rtilda) (the next iteration's rho). On
// the early exit of an iteration (|s| small, BicgStab.cpp:109-110) only
// x = p alpha + x.
template <typename R>
__global__ __launch_bounds__(kRedThreads) void k_bs_upd(typename CxT<R>::T* x, typename CxT<R>::T* r,
                                                        const typename CxT<R>::T* __restrict__ p,
                                                        const typename CxT<R>::T* __restrict__ s,
                                                        const typename CxT<R>::T* __restrict__ av,
                                                        const typename CxT<R>::T* __restrict__ rt, long long n,
                                                        double* part, const SolverState* st) {
  typedef typename CxT<R>::T C;
  DDA_RHS_PROLOGUE(st)
  if (s_->done) return;
  x += off_;
  r += off_;
  p += off_;
  s += off_;
  av += off_;
  const C a = coef<C>(&s_->alpha), om = coef<C>(&s_->omega), mom = coef<C>(&s_->mOmega);
  if (s_->skip) {
    DDA_GRID_LOOP(i) x[i] = cadd(cmul(p[i], a), x[i]);
    return;
  }
  rt += off_;
  double sum = 0, d0 = 0, d1 = 0;
  DDA_GRID_LOOP(i) {
    x[i] = cadd(cadd(cmul(p[i], a), cmul(s[i], om)), x[i]);
    const C rn = cadd(cmul(av[i], mom), s[i]);
    r[i] = rn;
    const double ar = rn.x, ai = rn.y;
    sum += ar * ar + ai * ai;
    // the next iteration's r.conj(rtilda), summed as k_reduce<RED_DOTC> does
    const double br = rt[i].x, bi = -(double)rt[i].y;
    d0 += ar * br - ai * bi;
    d1 += ar * bi + ai * br;
  }
  write_partials(part, 3, sum, d0, d1);
}

// CGNR / BiCG: p = first ? a : p beta + a.
template <typename R>
__global__ void k_p_update(typename CxT<R>::T* p, const typename CxT<R>::T* __restrict__ a, int first,
                           const SolverState* st, long long n) {
  typedef typename CxT<R>::T C;
  DDA_RHS_PROLOGUE(st)
  if (s_->done) return;
  p += off_;
  a += off_;
  const C b = coef<C>(&s_->bbeta);
  DDA_GRID_LOOP(i) p[i] = first ? a[i] : cadd(cmul(p[i], b), a[i]);
}

// x = p alpha + x; r = Av mAlpha + r; partial |r|^2 (CGNR, BiCG-CS).
template <typename R>
__global__ __launch_bounds__(kRedThreads) void k_xr_update(typename CxT<R>::T* x, typename CxT<R>::T* r,
                                                           const typename CxT<R>::T* __restrict__ p,
                                                           const typename CxT<R>::T* __restrict__ av, long long n,
                                                           double* part, const SolverState* st) {
  typedef typename CxT<R>::T C;
  DDA_RHS_PROLOGUE(st)
  if (s_->done) return;
  x += off_;
  r += off_;
  p += off_;
  av += off_;
  const C a = coef<C>(&s_->alpha), ma = coef<C>(&s_->mAlpha);
  double sum = 0;
  DDA_GRID_LOOP(i) {
    x[i] = cadd(cmul(p[i], a), x[i]);
    const C rn = cadd(cmul(av[i], ma), r[i]);
    r[i] = rn;
    sum += (double)rn.x * rn.x + (double)rn.y * rn.y;
  }
  write_partials(part, 1, sum, 0, 0);
}

// Mixed precision helpers, per right-hand side (blockIdx.y): out = r64 * scale[r]
// in the inner solver's precision; x64 += d * scale[r]; r64 = b64 - t64 with
// partials of |r|^2 and r^T r (quantities 0; 1, 2). mask[r] != 0 skips a
// right-hand side.
template <typename C>
__global__ void k_load_rhs(C* out, const double2* __restrict__ in, long long n, const double* scale, const int* mask) {
  if (mask[blockIdx.y]) return;
  const double sc = scale[blockIdx.y];
  out += blockIdx.y * n;
  in += blockIdx.y * n;
  DDA_GRID_LOOP(i) out[i] = cmake<C>(in[i].x * sc, in[i].y * sc);
}
template <typename C>
__global__ void k_add_corr(double2* x, const C* __restrict__ d, long long n, const double* scale, const int* mask) {
  if (mask[blockIdx.y]) return;
  const double sc = scale[blockIdx.y];
  x += blockIdx.y * n;
  d += blockIdx.y * n;
  DDA_GRID_LOOP(i) {
    x[i].x += (double)d[i].x * sc;
    x[i].y += (double)d[i].y * sc;
  }
}
__global__ __launch_bounds__(kRedThreads) void k_resid64(double2* r, const double2* __restrict__ b,
                                                         const double2* __restrict__ t, long long n, double* part,
                                                         const int* mask) {
  if (mask[blockIdx.y]) return;
  r += blockIdx.y * n;
  b += blockIdx.y * n;
  t += blockIdx.y * n;
  double s = 0, tr = 0, ti = 0;
  DDA_GRID_LOOP(i) {
    const double2 v = make_double2(b[i].x - t[i].x, b[i].y - t[i].y);
    r[i] = v;
    s += v.x * v.x + v.y * v.y;
    tr += v.x * v.x - v.y * v.y;
    ti += 2 * v.x * v.y;
  }
  write_partials(part, 3, s, tr, ti);
}
// |b - t|^2 per right-hand side of vectors of any precision, summed in FP64.
template <typename C>
__global__ __launch_bounds__(kRedThreads) void k_resid_norm(const C* __restrict__ b, const C* __restrict__ t,
                                                            long long n, double* part, const int* mask) {
  if (mask[blockIdx.y]) return;
  b += blockIdx.y * n;
  t += blockIdx.y * n;
  double s = 0;
  DDA_GRID_LOOP(i) {
    const double x = (double)b[i].x - (double)t[i].x, y = (double)b[i].y - (double)t[i].y;
    s += x * x + y * y;
  }
  write_partials(part, 1, s, 0, 0);
}
__global__ __launch_bounds__(kRedThreads) void k_norm64(const double2* __restrict__ a, long long n, double* part) {
  a += blockIdx.y * n;
  double s = 0, tr = 0, ti = 0;
  DDA_GRID_LOOP(i) {
    s += a[i].x * a[i].x + a[i].y * a[i].y;
    tr += a[i].x * a[i].x - a[i].y * a[i].y;
    ti += 2 * a[i].x * a[i].y;
  }
  write_partials(part, 3, s, tr, ti);
}

// ------------------------------------------------------------------------------------
// Host side: one batched Krylov core per precision
// ------------------------------------------------------------------------------------

// Whether the process runs under rocprofv3, which hands its settings to the
// profiled program in ROCPROF_* environment variables.
static bool profiler_attached() {
  static const bool attached = std::getenv("ROCPROF_OUTPUT_PATH") != nullptr ||
                               std::getenv("ROCPROF_KERNEL_TRACE") != nullptr ||
                               std::getenv("ROCPROF_COUNTER_COLLECTION") != nullptr;
  return attached;
}

#define DDA_OFF(field) ((int)(offsetof(SolverState, field)))
#define DDA_IOFF(field) ((int)(offsetof(SolverState, field) / sizeof(int)))

// One Krylov solve A x = b for up to `cap` right-hand sides on the device for an
// operator of precision R. v[0] = b (set by the caller), v[2] = x (x0 when
// have_x0), v[1] = r, v[3] = Av, v[4..7] solver work vectors; each [cap][n].
template <typename R>
struct Core {
  typedef typename CxT<R>::T C;
  dda_op_s* op = nullptr;
  int kind = 0;
  hipStream_t st = nullptr;
  long long n = 0;
  int nb = 1;                       // reduction blocks per right-hand side
  int cap = 0;                      // right-hand sides allocated
  int nrhs = 1;                     // right-hand sides of the current solve
  SolverState* dst = nullptr;       // device [cap]
  SolverState* host2 = nullptr;     // pinned [2][cap]
  double* part = nullptr;           // [cap][4][nb]
  // [cap][4][at most 3 * n_nonempty]: partial sums formed by pass 5 of a
  // matvec (compile-time planned X passes), nb = op->p5_blocks
  double* part5 = nullptr;
  bool fused = false;               // this solve takes those sums from pass 5
  double* hist = nullptr;           // [cap][hist_cap]
  uint64_t hist_cap = 0;
  uint64_t max_iter = 0, max_res_increase = 0;
  C* v[8] = {};
  // Two steady-state iterations (nr >= 2, even) captured as a HIP graph and
  // replayed: the kernel sequence and its arguments repeat with period two
  // (QMR-CS swaps its vector pairs every iteration). Recaptured whenever
  // anything its launches reference changes.
  hipGraphExec_t gexec = nullptr;
  std::vector<const void*> gsig;
  // plateau stop for the next solves (SolverState::plateau_window; 0 = off)
  uint64_t plateau_window = 0;
  double plateau_ratio = 0.81;
  // GPU share of a long solve (dda_solver_set_gpu_share): once a solve has run
  // share_after_s seconds, each chunk is waited for and the stream then left
  // idle for (1 / share - 1) times as long, so that other streams' work runs
  // in between; share 1 = off
  double share_after_s = 0;
  double share = 1.0;

  dim3 g() const { return dim3(nb, nrhs); }
  dim3 b() const { return dim3(kRedThreads); }
  const int* flag(int ioff) const { return reinterpret_cast<const int*>(dst) + ioff; }
  dda_status matvec(const void* x, void* y, int conj, int ioff) {
    return launch_matvec<R>(op, x, y, conj, st, flag(ioff), nrhs, kStateInts);
  }
  void scalar(int step) {
    hipLaunchKernelGGL(k_scalar, dim3(nrhs), dim3(kScalarThreads), 0, st, dst, part, nb, step, kind, hist);
  }
  // y = A x, then scalar step `step` over the partial sums `dop` of x and y:
  // formed by the matvec's pass 5 when fused, else by a reduction kernel
  // (DOT_U: k_reduce DOTU; DOT_NORM: k_reduce NORM of y; DOT_OM: k_bs_om).
  dda_status matvec_sums(const C* x, C* y, int conj, int ioff, int dop, int step) {
    if (fused) {
      DDA_TRY(launch_matvec<R>(op, x, y, conj, st, flag(ioff), nrhs, kStateInts, part5, dop));
      hipLaunchKernelGGL(k_scalar, dim3(nrhs), dim3(kScalarThreads), 0, st, dst, part5, op->p5_blocks, step, kind,
                         hist);
      return DDA_OK;
    }
    DDA_TRY(matvec(x, y, conj, ioff));
    if (dop == DOT_U)
      reduce(RED_DOTU, x, y, ioff);
    else if (dop == DOT_NORM)
      reduce(RED_NORM, y, y, ioff);
    else
      hipLaunchKernelGGL(k_bs_om<R>, g(), b(), 0, st, y, x, n, part, dst);
    scalar(step);
    return DDA_OK;
  }
  void reduce(int opk, const C* a, const C* bb, int ioff) {
    if (opk == RED_NORM)
      hipLaunchKernelGGL((k_reduce<R, RED_NORM>), g(), b(), 0, st, a, bb, n, part, dst, ioff);
    else if (opk == RED_DOTU)
      hipLaunchKernelGGL((k_reduce<R, RED_DOTU>), g(), b(), 0, st, a, bb, n, part, dst, ioff);
    else
      hipLaunchKernelGGL((k_reduce<R, RED_DOTC>), g(), b(), 0, st, a, bb, n, part, dst, ioff);
  }
  void lincomb(C* out, const C* v1, int c1, const C* v2, int c2, const C* v3, bool withpart, int ioff) {
    hipLaunchKernelGGL(k_lincomb<R>, g(), b(), 0, st, out, v1, c1, v2, c2, v3, n, withpart ? part : nullptr, dst, ioff);
  }

  dda_status alloc(dda_op_s* o, int k, uint64_t mi, uint64_t mri, int capacity) {
    release();
    op = o;
    kind = k;
    st = o->ctx->stream;
    n = 3ll * o->dims.S;
    nb = red_blocks(n);
    cap = capacity;
    max_iter = mi;
    max_res_increase = mri;
    hist_cap = std::min<uint64_t>(mi + 1, 1u << 16);
    DDA_TRY_HIP(hipMalloc(&dst, cap * sizeof(SolverState)));
    DDA_TRY_HIP(hipHostMalloc((void**)&host2, 2 * cap * sizeof(SolverState), hipHostMallocDefault));
    DDA_TRY_HIP(hipMalloc(&part, (size_t)cap * 4 * nb * sizeof(double)));
    if (o->static_x)
      DDA_TRY_HIP(hipMalloc(&part5, (size_t)cap * 4 * std::max(1, 3 * o->n_nonempty) * sizeof(double)));
    DDA_TRY_HIP(hipMalloc(&hist, (size_t)cap * hist_cap * sizeof(double)));
    for (int i = 0; i < 8; i++) DDA_TRY_HIP(hipMalloc(&v[i], (size_t)cap * n * sizeof(C)));
    return DDA_OK;
  }
  void drop_graph() {
    if (gexec) (void)hipGraphExecDestroy(gexec);
    gexec = nullptr;
    gsig.clear();
  }
  std::vector<const void*> graph_signature() const {
    // every buffer a captured launch references that can be reallocated
    // (a larger block of right-hand sides, a pack's member table)
    std::vector<const void*> g = {dst, part, part5, hist, op->a1, op->a2, op->ccs, op->dhat, op->dims.member,
                                  op->dims.dhat_tab, reinterpret_cast<const void*>((intptr_t)nrhs),
                                  reinterpret_cast<const void*>((intptr_t)fused)};
    for (const dda_op_s::Tune* tn : {&op->tune, &op->tune_blk})
      for (int t : {tn->col_wide[0], tn->col_wide[1], tn->p3_wide, tn->x_wide[0], tn->x_wide[1], tn->split})
        g.push_back(reinterpret_cast<const void*>((intptr_t)t));
    for (int t : {(int)op->split_rhs, (int)op->tuned_blk}) g.push_back(reinterpret_cast<const void*>((intptr_t)t));
    for (C* p : v) g.push_back(p);
    return g;
  }
  void release() {
    drop_graph();
    if (dst) (void)hipFree(dst);
    if (host2) (void)hipHostFree(host2);
    if (part) (void)hipFree(part);
    if (part5) (void)hipFree(part5);
    if (hist) (void)hipFree(hist);
    for (C* p : v)
      if (p) (void)hipFree(p);
    dst = nullptr;
    host2 = nullptr;
    part = nullptr;
    part5 = nullptr;
    hist = nullptr;
    for (auto& p : v) p = nullptr;
    cap = 0;
  }

  // initGeneral + the getPolVec loop (GpuIterativeSolver.cpp:78-134,
  // IterativeSolverBase.cpp:40-103) for right-hand sides 0..k-1 with eps[r];
  // active[r] == 0 leaves r untouched (done from the start). Leaves x in v[2]
  // and the final states in fin[0..k-1].
  // profile_us (nullable): instead of iterating to convergence, run one
  // untimed iteration and then profile_n iterations, each enqueued eagerly and
  // waited for, and write their wall times (microseconds) there
  // (IterativeSolverBase::profilingRun, IterativeSolverBase.cpp:105-136).
  dda_status solve(int k, const double* eps, const int* active, bool have_x0, SolverState* fin,
                   const uint64_t* max_iters = nullptr, double* profile_us = nullptr, int profile_n = 0) {
    nrhs = k;
    // the sums after a matvec come from its pass 5 where the X passes have
    // a compile-time plan (DDA_NO_FUSED_SUMS=1: from reduction kernels)
    fused = part5 && !std::getenv("DDA_NO_FUSED_SUMS");
    C* bv = v[0];
    C* r = v[1];
    C* x = v[2];
    C* av = v[3];
    const int DONE = DDA_IOFF(done), SKIP = DDA_IOFF(skip);
    std::vector<SolverState> h(k);
    for (int i = 0; i < k; i++) {
      std::memset(&h[i], 0, sizeof(SolverState));
      h[i].max_iter = max_iters ? max_iters[i] : max_iter;
      h[i].max_res_increase = max_res_increase;
      h[i].plateau_window = plateau_window;
      h[i].plateau_ratio = plateau_ratio;
      h[i].hist_cap = hist_cap;
      h[i].eps = eps[i];
      if (active && !active[i]) {
        h[i].done = 1;
        h[i].skip = 1;
        h[i].status = DDA_OK;
      }
    }
    DDA_TRY_HIP(hipMemcpyAsync(dst, h.data(), k * sizeof(SolverState), hipMemcpyHostToDevice, st));
    reduce(RED_NORM, bv, bv, SKIP);
    scalar(S_INIT_TEMP);
    if (have_x0) {
      DDA_TRY(matvec(x, av, 0, DONE));
      lincomb(r, av, DDA_OFF(mOne), nullptr, -1, bv, true, SKIP);  // r = b - A x0
      scalar(S_INIT_START);
    } else {
      DDA_TRY(matvec(bv, av, 0, DONE));
      lincomb(r, av, DDA_OFF(mOne), nullptr, -1, bv, true, SKIP);  // r = b - A b
      scalar(S_INIT_R);
      hipLaunchKernelGGL(k_select_x0<R>, g(), b(), 0, st, x, r, bv, n, dst);
    }
    if (kind == DDA_SOLVER_QMR_CS) {
      reduce(RED_DOTU, r, r, SKIP);
      scalar(S_QMR_INIT);
      lincomb(v[4], r, DDA_OFF(betaInv), nullptr, -1, nullptr, false, SKIP);  // v = r / beta
    } else if (kind == DDA_SOLVER_BICGSTAB) {
      hipLaunchKernelGGL(k_copy_flag<R>, g(), b(), 0, st, v[7], r, n, dst);  // rtilda = r
    }
    scalar(S_LOOP_CHECK);
    DDA_TRY_HIP(hipGetLastError());

    int iv = 4, ivt = 5, ipo = 6, ipn = 7;
    uint64_t nr = 0;
    auto enqueue_iteration = [&]() -> dda_status {
      switch (kind) {
        case DDA_SOLVER_QMR_CS: {
          C* vv = v[iv];
          C* vt = v[ivt];
          DDA_TRY(matvec_sums(vv, av, 0, DONE, DOT_U, S_QMR_ALPHA));  // alpha = v.Av
          hipLaunchKernelGGL(k_qmr_vt<R>, g(), b(), 0, st, vt, vv, av, nr == 0 ? 1 : 0, dst, n, part);
          scalar(S_QMR_GIVENS);
          const int mode = nr == 0 ? 0 : (nr == 1 ? 1 : 2);
          hipLaunchKernelGGL(k_qmr_upd<R>, g(), b(), 0, st, v[ipo], v[ipn], vv, vt, x, r, mode, dst, n, part);
          scalar(S_END_ITER);
          std::swap(ipo, ipn);
          std::swap(iv, ivt);
          break;
        }
        case DDA_SOLVER_BICGSTAB: {
          C* p = v[4];
          C* vv = v[5];
          C* sv = v[6];
          C* rt = v[7];
          if (nr == 0) {
            reduce(RED_DOTC, r, rt, DONE);
            scalar(S_BS_RO);
          } else {
            scalar(S_BS_RO1);  // rho from the previous x/r update's partials
          }
          hipLaunchKernelGGL(k_bs_p<R>, g(), b(), 0, st, p, vv, r, nr == 0 ? 1 : 0, dst, n);
          DDA_TRY(matvec(p, vv, 0, DONE));
          reduce(RED_DOTC, vv, rt, DONE);
          scalar(S_BS_ALPHA);
          lincomb(sv, vv, DDA_OFF(mAlpha), nullptr, -1, r, true, DONE);  // s = v (-alpha) + r
          scalar(S_BS_S);
          DDA_TRY(matvec_sums(sv, av, 0, SKIP, DOT_OM, S_BS_OMEGA));  // t = A s; |t|^2, s.conj(t)
          hipLaunchKernelGGL(k_bs_upd<R>, g(), b(), 0, st, x, r, p, sv, av, rt, n, part, dst);
          scalar(S_END_ITER);
          break;
        }
        case DDA_SOLVER_CGNR: {
          C* p = v[4];
          DDA_TRY(matvec_sums(r, av, 1, DONE, DOT_NORM, S_CG_RO));  // A^H r, |A^H r|^2
          hipLaunchKernelGGL(k_p_update<R>, g(), b(), 0, st, p, av, nr == 0 ? 1 : 0, dst, n);
          DDA_TRY(matvec_sums(p, av, 0, DONE, DOT_NORM, S_CG_ALPHA));  // |A p|^2
          hipLaunchKernelGGL(k_xr_update<R>, g(), b(), 0, st, x, r, p, av, n, part, dst);
          scalar(S_END_ITER);
          break;
        }
        case DDA_SOLVER_BICG_CS: {
          C* p = v[4];
          reduce(RED_DOTU, r, r, DONE);
          scalar(S_BI_RO);
          hipLaunchKernelGGL(k_p_update<R>, g(), b(), 0, st, p, r, nr == 0 ? 1 : 0, dst, n);
          DDA_TRY(matvec_sums(p, av, 0, DONE, DOT_U, S_BI_MU));  // mu = p.Ap
          hipLaunchKernelGGL(k_xr_update<R>, g(), b(), 0, st, x, r, p, av, n, part, dst);
          scalar(S_END_ITER);
          break;
        }
      }
      DDA_TRY_HIP(hipGetLastError());
      nr++;
      return DDA_OK;
    };

    // The graph of two iterations: captured from the eager sequence (which
    // leaves the vector roles as they were), then replayed. Not under a
    // profiler: a kernel-trace run of a 128-RHS block solve ended in a host
    // segmentation fault inside the HIP runtime's graph launch (a read of a
    // device address) while the profiler intercepted the graph's dispatches;
    // the same graphs replay cleanly without it.
    const bool use_graph = !op->prof && !std::getenv("DDA_NO_GRAPHS") && !profiler_attached();
    auto launch_pair = [&]() -> dda_status {
      std::vector<const void*> sig = graph_signature();
      if (!gexec || sig != gsig) {
        drop_graph();
        hipGraph_t g = nullptr;
        DDA_TRY_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        dda_status rc1 = enqueue_iteration();
        if (rc1 == DDA_OK) rc1 = enqueue_iteration();
        hipError_t e = hipStreamEndCapture(st, &g);
        nr -= 2;
        if (rc1 != DDA_OK) {
          if (g) (void)hipGraphDestroy(g);
          return rc1;
        }
        DDA_TRY_HIP(e);
        e = hipGraphInstantiate(&gexec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        DDA_TRY_HIP(e);
        gsig = sig;
      }
      DDA_TRY_HIP(hipGraphLaunch(gexec, st));
      nr += 2;
      return DDA_OK;
    };

    if (profile_us) {
      for (int i = 0; i <= profile_n; i++) {
        const auto t0 = std::chrono::steady_clock::now();
        DDA_TRY(enqueue_iteration());
        DDA_TRY_HIP(hipStreamSynchronize(st));
        const double us = std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - t0).count();
        if (i > 0) profile_us[i - 1] = us;
      }
      DDA_TRY_HIP(hipMemcpyAsync(fin, dst, k * sizeof(SolverState), hipMemcpyDeviceToHost, st));
      DDA_TRY_HIP(hipStreamSynchronize(st));
      return DDA_OK;
    }

    // Chunks of iterations are enqueued ahead of the state reads: chunk c+1 is
    // queued before the states copied after chunk c are examined, so the GPU
    // never waits on the host; the chunk length follows the slowest observed
    // convergence rate so that few iterations are queued past the one that
    // meets the tolerance (those run as no-ops).
    hipEvent_t ev[2];
    DDA_TRY_HIP(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
    DDA_TRY_HIP(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
    DDA_TRY_HIP(hipMemcpyAsync(host2, dst, k * sizeof(SolverState), hipMemcpyDeviceToHost, st));
    DDA_TRY_HIP(hipEventRecord(ev[0], st));
    int next_len = 2, chunk = 0;
    // DDA_PROGRESS=<seconds>: a progress line on stderr at most that often, as
    // the reference logs the residual during getPolVec
    const char* pe = std::getenv("DDA_PROGRESS");
    const double progress_s = pe ? std::atof(pe) : 0.0;
    const auto t_start = std::chrono::steady_clock::now();
    double last_print = 0;
    std::vector<double> prev_r(k, -1.0);
    std::vector<uint64_t> prev_count(k, 0);
    bool finished = false;
    dda_status rc = DDA_OK;
    while (!finished && rc == DDA_OK) {
      for (int it = 0; it < next_len && rc == DDA_OK;) {
        if (use_graph && nr >= 2 && nr % 2 == 0 && next_len - it >= 2) {
          rc = launch_pair();
          it += 2;
        } else {
          rc = enqueue_iteration();
          it++;
        }
      }
      if (rc != DDA_OK) break;
      const int slot = (chunk + 1) & 1;
      DDA_TRY_HIP(hipMemcpyAsync(host2 + slot * cap, dst, k * sizeof(SolverState), hipMemcpyDeviceToHost, st));
      DDA_TRY_HIP(hipEventRecord(ev[slot], st));
      const int prev = chunk & 1;
      DDA_TRY_HIP(hipEventSynchronize(ev[prev]));
      const SolverState* hs = host2 + prev * cap;
      finished = true;
      int len = 1;
      for (int i = 0; i < k; i++) {
        if (hs[i].done) continue;
        finished = false;
        // half the iterations the recent convergence rate still asks for
        // (the residuals of these methods are erratic, and a longer guess
        // runs past convergence as short no-op launches), at least one graph
        // pair; measured 1% faster than the full estimate on a 2-RHS solve at 256^3
        int li = kChunk / 2;
        if (prev_r[i] > 0 && hs[i].inprodR > 0 && hs[i].count > prev_count[i] && hs[i].inprodR < prev_r[i] &&
            hs[i].epsB > 0) {
          const double rate = std::log(prev_r[i] / hs[i].inprodR) / (double)(hs[i].count - prev_count[i]);
          const double need = 0.5 * (std::log(hs[i].inprodR / hs[i].epsB) / rate - (double)next_len);
          li = need < 2 ? 2 : (need > kChunk ? kChunk : (int)std::ceil(need));
        }
        prev_r[i] = hs[i].inprodR;
        prev_count[i] = hs[i].count;
        len = std::max(len, li);
      }
      next_len = len;
      if (!finished && share < 1.0) {
        const auto t_now = std::chrono::steady_clock::now();
        if (std::chrono::duration<double>(t_now - t_start).count() >= share_after_s) {
          DDA_TRY_HIP(hipStreamSynchronize(st));  // the chunk queued behind the one just read
          const double busy = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_now).count();
          std::this_thread::sleep_for(std::chrono::duration<double>(busy * (1.0 / share - 1.0)));
        }
      }
      if (progress_s > 0) {
        const double now = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
        if (now - last_print >= progress_s) {
          last_print = now;
          int open = 0;
          for (int i = 0; i < k; i++) open += hs[i].done ? 0 : 1;
          std::fprintf(stderr, "dda: %.0f s, iteration %llu, |r|/|b| %.3e (rhs 0), %d of %d right-hand sides open\n",
                       now, (unsigned long long)hs[0].count, std::sqrt(hs[0].residScale * hs[0].inprodR), open, k);
        }
      }
      chunk++;
    }
    (void)hipStreamSynchronize(st);
    (void)hipEventDestroy(ev[0]);
    (void)hipEventDestroy(ev[1]);
    if (rc != DDA_OK) return rc;
    DDA_TRY_HIP(hipMemcpyAsync(fin, dst, k * sizeof(SolverState), hipMemcpyDeviceToHost, st));
    DDA_TRY_HIP(hipStreamSynchronize(st));
    return DDA_OK;
  }
};

template struct Core<float>;
template struct Core<double>;

}  // namespace dda

using namespace dda;

struct dda_solver_s {
  dda_op_s* op = nullptr;        // caller's operator (FP64; FP32 mode also takes an FP32 one)
  dda_op_s* op32 = nullptr;      // FP32 twin of an FP64 operator (mixed and FP32 modes), owned
  int kind = 0;
  int mode = DDA_MODE_FP64;
  uint64_t max_iter = 0, max_res_increase = 0;
  long long n = 0;
  int cap = 0;                   // right-hand sides allocated
  Core<double> c64;              // FP64 mode
  Core<float> c32;               // mixed mode inner solver; FP32 mode
  double2* x64 = nullptr;        // [cap][n] FP64 vectors
  double2* b64 = nullptr;
  double2* r64 = nullptr;
  double2* t64 = nullptr;
  double* part64 = nullptr;      // [cap][4][nb]
  double* scale = nullptr;       // [cap] (mixed)
  int* mask = nullptr;           // [cap] (mixed)
  double2* io = nullptr;         // [2][cap][n] FP64 staging for einc / start / result
  uint64_t last_iters = 0;
  std::vector<double> outer_hist;
  double inner_floor = 1e-4;
  uint64_t inner_plateau = 0;     // set at creation: kInnerPlateau
  double inner_plateau_ratio = 0.81;
};

namespace {

// The FP32 operator the iteration of the mixed and FP32 modes runs on.
dda_op_s* op_fp32(const dda_solver_s* s) { return s->op32 ? s->op32 : s->op; }

// Host sums of the per-RHS partials: quantity 0 into nrm; with rho, also
// |sum q1 + i q2| / q0 (|r^T r| / |r|^2 after k_resid64 / k_norm64).
dda_status read_sums(hipStream_t st, const double* part, int nb, int k, std::vector<double>* nrm,
                     std::vector<double>* rho = nullptr) {
  std::vector<double> h((size_t)k * 4 * nb);
  DDA_TRY_HIP(hipMemcpyAsync(h.data(), part, h.size() * sizeof(double), hipMemcpyDeviceToHost, st));
  DDA_TRY_HIP(hipStreamSynchronize(st));
  nrm->assign(k, 0.0);
  if (rho) rho->assign(k, 0.0);
  for (int r = 0; r < k; r++) {
    double q[3] = {0, 0, 0};
    for (int j = 0; j < (rho ? 3 : 1); j++)
      for (int i = 0; i < nb; i++) q[j] += h[((size_t)r * 4 + j) * nb + i];
    (*nrm)[r] = q[0];
    if (rho) (*rho)[r] = q[0] > 0 ? std::hypot(q[1], q[2]) / q[0] : 0.0;
  }
  return DDA_OK;
}

dda_status alloc_buffers(dda_solver_s* s, int k) {
  if (k <= s->cap) return DDA_OK;
  dda_op_s* op = s->op;
  DDA_TRY_HIP(hipStreamSynchronize(op->ctx->stream));
  void* bufs[] = {s->x64, s->b64, s->r64, s->t64, s->part64, s->scale, s->mask, s->io};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
  s->x64 = s->b64 = s->r64 = s->t64 = s->io = nullptr;
  s->part64 = s->scale = nullptr;
  s->mask = nullptr;
  const int nb = red_blocks(s->n);
  const size_t vb = (size_t)k * s->n * sizeof(double2);
  DDA_TRY_HIP(hipMalloc(&s->io, 2 * vb));
  if (s->mode != DDA_MODE_FP32) {
    DDA_TRY_HIP(hipMalloc(&s->t64, vb));
    DDA_TRY_HIP(hipMalloc(&s->r64, vb));
  }
  DDA_TRY_HIP(hipMalloc(&s->part64, (size_t)k * 4 * nb * sizeof(double)));
  DDA_TRY_HIP(hipMalloc(&s->scale, k * sizeof(double)));
  DDA_TRY_HIP(hipMalloc(&s->mask, k * sizeof(int)));
  if (s->mode == DDA_MODE_MIXED) {
    DDA_TRY_HIP(hipMalloc(&s->x64, vb));
    DDA_TRY_HIP(hipMalloc(&s->b64, vb));
    DDA_TRY(s->c32.alloc(s->op32, s->kind, s->max_iter, s->max_res_increase, k));
  } else if (s->mode == DDA_MODE_FP32) {
    DDA_TRY(s->c32.alloc(op_fp32(s), s->kind, s->max_iter, s->max_res_increase, k));
  } else {
    DDA_TRY(s->c64.alloc(op, s->kind, s->max_iter, s->max_res_increase, k));
  }
  s->cap = k;
  return DDA_OK;
}

// |b - A x|^2 per right-hand side (FP64 block matvec into t64, r64 = b - t64);
// right-hand sides with skip[r] != 0 are left out (their sums are undefined).
dda_status residual64(dda_solver_s* s, int k, double2* x, double2* b, const std::vector<int>& skip,
                      std::vector<double>* rr, std::vector<double>* rho = nullptr) {
  hipStream_t st = s->op->ctx->stream;
  const int nb = red_blocks(s->n);
  DDA_TRY_HIP(hipMemcpyAsync(s->mask, skip.data(), k * sizeof(int), hipMemcpyHostToDevice, st));
  DDA_TRY(launch_matvec<double>(s->op, x, s->t64, 0, st, s->mask, k, 1));
  hipLaunchKernelGGL(k_resid64, dim3(nb, k), dim3(kRedThreads), 0, st, s->r64, b, s->t64, s->n, s->part64, s->mask);
  DDA_TRY_HIP(hipGetLastError());
  return read_sums(st, s->part64, nb, k, rr, rho);
}

void fill_info(dda_solve_info* info, int r, uint64_t iters, uint64_t mvs, uint64_t refine, double rel, double trel,
               int status) {
  if (!info) return;
  info[r].iterations = iters;
  info[r].matvecs = mvs;
  info[r].refinements = refine;
  info[r].rel_residual = rel;
  info[r].true_rel_residual = trel;
  info[r].status = status;
}

dda_status run_fp64(dda_solver_s* s, int k, double eps, bool have_start, dda_solve_info* info) {
  dda_op_s* op = s->op;
  hipStream_t st = op->ctx->stream;
  const MvDims& d = op->dims;
  Core<double>& K = s->c64;
  const double2* ccs = (const double2*)op->ccs;
  const double gu = op->grid_unit;
  const dim3 G(red_blocks(s->n), k), B(kRedThreads);
  hipLaunchKernelGGL(k_setup_b<double>, G, B, 0, st, K.v[0], s->io, ccs, op->mat_dev, d.nv, d.S, d.nmat, d.member);
  if (have_start)
    hipLaunchKernelGGL(k_setup_start<double>, G, B, 0, st, K.v[2], s->io + (size_t)s->cap * s->n, ccs, op->mat_dev,
                       d.nv, d.S, d.nmat, 1.0 / (gu * gu * gu), d.member);
  DDA_TRY_HIP(hipGetLastError());
  std::vector<SolverState> fin(k);
  std::vector<double> epsv(k, eps);
  DDA_TRY(K.solve(k, epsv.data(), nullptr, have_start, fin.data()));
  hipLaunchKernelGGL(k_get_result<double>, G, B, 0, st, s->io, K.v[2], ccs, op->mat_dev, d.nv, d.S, d.nmat,
                     gu * gu * gu, d.member);
  DDA_TRY_HIP(hipGetLastError());
  std::vector<double> rr;
  std::vector<int> none(k, 0);
  for (int r = 0; r < k; r++) none[r] = fin[r].temp > 0 ? 0 : 1;
  DDA_TRY(residual64(s, k, K.v[2], K.v[0], none, &rr));
  dda_status worst = DDA_OK;
  s->last_iters = fin[0].count;
  const uint64_t per = s->kind == DDA_SOLVER_QMR_CS || s->kind == DDA_SOLVER_BICG_CS ? 1 : 2;
  for (int r = 0; r < k; r++) {
    fill_info(info, r, fin[r].count, fin[r].count * per + 2, 0, std::sqrt(fin[r].residScale * fin[r].inprodR),
              none[r] ? 0.0 : std::sqrt(rr[r] / fin[r].temp), fin[r].status);
    if (fin[r].status != DDA_OK && worst == DDA_OK) worst = (dda_status)fin[r].status;
  }
  return worst;
}

// The reference's float path (ddaCl<float>, DDA/DDAMain.cpp:723-727): the
// whole Krylov solve on FP32 vectors with the FP32 operator, no refinement
// (this build keeps the scalars and reductions in FP64). The true residual is
// recomputed with the FP32 operator, summed in FP64.
dda_status run_fp32(dda_solver_s* s, int k, double eps, bool have_start, dda_solve_info* info) {
  dda_op_s* op = op_fp32(s);
  hipStream_t st = op->ctx->stream;
  const MvDims& d = op->dims;
  Core<float>& K = s->c32;
  const float2* ccs = (const float2*)op->ccs;
  const double gu = op->grid_unit;
  const int nb = red_blocks(s->n);
  const dim3 G(nb, k), B(kRedThreads);
  hipLaunchKernelGGL(k_setup_b<float>, G, B, 0, st, K.v[0], s->io, ccs, op->mat_dev, d.nv, d.S, d.nmat, d.member);
  if (have_start)
    hipLaunchKernelGGL(k_setup_start<float>, G, B, 0, st, K.v[2], s->io + (size_t)s->cap * s->n, ccs, op->mat_dev,
                       d.nv, d.S, d.nmat, 1.0 / (gu * gu * gu), d.member);
  DDA_TRY_HIP(hipGetLastError());
  std::vector<SolverState> fin(k);
  std::vector<double> epsv(k, eps);
  DDA_TRY(K.solve(k, epsv.data(), nullptr, have_start, fin.data()));
  hipLaunchKernelGGL(k_get_result<float>, G, B, 0, st, s->io, K.v[2], ccs, op->mat_dev, d.nv, d.S, d.nmat,
                     gu * gu * gu, d.member);
  DDA_TRY_HIP(hipGetLastError());
  std::vector<int> none(k, 0);
  for (int r = 0; r < k; r++) none[r] = fin[r].temp > 0 ? 0 : 1;
  DDA_TRY_HIP(hipMemcpyAsync(s->mask, none.data(), k * sizeof(int), hipMemcpyHostToDevice, st));
  DDA_TRY(launch_matvec<float>(op, K.v[2], K.v[3], 0, st, s->mask, k, 1));
  hipLaunchKernelGGL(k_resid_norm<float2>, G, B, 0, st, (const float2*)K.v[0], (const float2*)K.v[3], s->n,
                     s->part64, s->mask);
  DDA_TRY_HIP(hipGetLastError());
  std::vector<double> rr;
  DDA_TRY(read_sums(st, s->part64, nb, k, &rr));
  dda_status worst = DDA_OK;
  s->last_iters = fin[0].count;
  const uint64_t per = s->kind == DDA_SOLVER_QMR_CS || s->kind == DDA_SOLVER_BICG_CS ? 1 : 2;
  for (int r = 0; r < k; r++) {
    fill_info(info, r, fin[r].count, fin[r].count * per + 2, 0, std::sqrt(fin[r].residScale * fin[r].inprodR),
              none[r] ? 0.0 : std::sqrt(rr[r] / fin[r].temp), fin[r].status);
    if (fin[r].status != DDA_OK && worst == DDA_OK) worst = (dda_status)fin[r].status;
  }
  return worst;
}

// FP64 iterative refinement around FP32 Krylov solves, per right-hand side:
//   r = b - A64 x; stop when |r|^2 < eps^2 |b|^2; solve A d = r / |r| to
//   max(floor, eps |b| / (2 |r|)); x += |r| d.
// The inner solve runs in FP32 unless it cannot make progress there, when it
// runs on the FP64 operator for that right-hand side and step:
//   - QMR-CS and BiCG-CS build their Lanczos basis with the unconjugated
//     product; when |r^T r| / |r|^2 < kRhoMin the first basis vector is
//     r / sqrt(r^T r) and FP32 rounding, amplified by |r|^2 / |r^T r|, stalls
//     the iteration (an incident wave whose phase pattern nearly cancels in
//     sum_i E_i^2 over the particle does this);
//   - after an FP32 step that did not halve the FP64 residual, that
//     right-hand side stays on FP64 for the rest of the solve.
// An FP32 inner solve also stops early (its status STAGNATION, its correction
// applied) once the smallest residual it has seen fell by less than 10% over
// kInnerPlateau iterations: the outer step then restarts it from the FP64
// residual. (FP32 QMR-CS on the 512-across sphere stalled at |r| / |b| =
// 1.3e-2 for 3000 iterations without it.)
// Right-hand sides that have converged are masked out of the later steps.
constexpr double kRhoMin = 1e-4;
constexpr uint64_t kInnerPlateau = 200;

dda_status run_mixed(dda_solver_s* s, int k, double eps, bool have_start, dda_solve_info* info) {
  dda_op_s* op = s->op;
  hipStream_t st = op->ctx->stream;
  const MvDims& d = op->dims;
  const double2* ccs = (const double2*)op->ccs;
  const double gu = op->grid_unit;
  const int nb = red_blocks(s->n);
  const dim3 G(nb, k), B(kRedThreads);
  const bool lanczos = s->kind == DDA_SOLVER_QMR_CS || s->kind == DDA_SOLVER_BICG_CS;
  hipLaunchKernelGGL(k_setup_b<double>, G, B, 0, st, s->b64, s->io, ccs, op->mat_dev, d.nv, d.S, d.nmat, d.member);
  hipLaunchKernelGGL(k_norm64, G, B, 0, st, s->b64, s->n, s->part64);
  DDA_TRY_HIP(hipGetLastError());
  std::vector<double> temp, rho;
  DDA_TRY(read_sums(st, s->part64, nb, k, &temp, &rho));
  std::vector<int> status(k, DDA_OK), conv(k, 0), sticky64(k, 0);
  for (int r = 0; r < k; r++)
    if (!(temp[r] > 0)) {
      status[r] = DDA_INVALID_ARG;  // a zero right-hand side
      conv[r] = 1;
    }
  if (have_start)
    hipLaunchKernelGGL(k_setup_start<double>, G, B, 0, st, s->x64, s->io + (size_t)s->cap * s->n, ccs, op->mat_dev,
                       d.nv, d.S, d.nmat, 1.0 / (gu * gu * gu), d.member);
  else
    DDA_TRY_HIP(hipMemsetAsync(s->x64, 0, (size_t)k * s->n * sizeof(double2), st));
  s->outer_hist.clear();
  std::vector<uint64_t> iters(k, 0), mvs(k, 0), outer(k, 0);
  std::vector<double> rel2(k, 1.0), rr;
  const uint64_t per = lanczos ? 1 : 2;
  const uint64_t max_outer = 40;
  for (uint64_t step = 0;; step++) {
    std::vector<double> prev = rel2;
    if (step == 0 && !have_start) {
      DDA_TRY_HIP(hipMemcpyAsync(s->r64, s->b64, (size_t)k * s->n * sizeof(double2), hipMemcpyDeviceToDevice, st));
      for (int r = 0; r < k; r++) rel2[r] = 1.0;
    } else {
      DDA_TRY(residual64(s, k, s->x64, s->b64, conv, &rr, &rho));
      for (int r = 0; r < k; r++) {
        if (conv[r]) continue;
        rel2[r] = rr[r] / temp[r];
        mvs[r]++;
      }
    }
    s->outer_hist.push_back(rel2[0]);
    bool any = false;
    std::vector<double> eps_in(k, 1.0), sc(k, 0.0);
    std::vector<int> act32(k, 0), act64(k, 0);
    for (int r = 0; r < k; r++) {
      if (conv[r]) continue;
      if (rel2[r] < eps * eps) {
        conv[r] = 1;
        continue;
      }
      if (outer[r] >= max_outer || iters[r] >= s->max_iter) {
        status[r] = DDA_MAX_ITER;
        conv[r] = 1;
        continue;
      }
      if (outer[r] > 0 && !sticky64[r] && rel2[r] > 0.25 * prev[r]) sticky64[r] = 1;
      any = true;
      (sticky64[r] || (lanczos && rho[r] < kRhoMin) ? act64 : act32)[r] = 1;
      sc[r] = 1.0 / std::sqrt(rel2[r] * temp[r]);
      eps_in[r] = std::max(s->inner_floor, 0.5 * eps / std::sqrt(rel2[r]));
    }
    if (!any) break;
    std::vector<double> inv(k, 0.0);
    for (int r = 0; r < k; r++) inv[r] = sc[r] > 0 ? 1.0 / sc[r] : 0.0;
    for (int pass = 0; pass < 2; pass++) {
      const std::vector<int>& act = pass == 0 ? act32 : act64;
      bool some = false;
      for (int r = 0; r < k; r++) some = some || act[r];
      if (!some) continue;
      std::vector<int> msk(k);
      std::vector<uint64_t> budget(k, 0);
      for (int r = 0; r < k; r++) {
        msk[r] = act[r] ? 0 : 1;
        if (act[r]) budget[r] = s->max_iter - iters[r];
      }
      DDA_TRY_HIP(hipMemcpyAsync(s->scale, sc.data(), k * sizeof(double), hipMemcpyHostToDevice, st));
      DDA_TRY_HIP(hipMemcpyAsync(s->mask, msk.data(), k * sizeof(int), hipMemcpyHostToDevice, st));
      std::vector<SolverState> fin(k);
      if (pass == 0) {
        hipLaunchKernelGGL(k_load_rhs<float2>, G, B, 0, st, (float2*)s->c32.v[0], s->r64, s->n, s->scale, s->mask);
        DDA_TRY_HIP(hipGetLastError());
        s->c32.plateau_window = s->inner_plateau;
        s->c32.plateau_ratio = s->inner_plateau_ratio;
        DDA_TRY(s->c32.solve(k, eps_in.data(), act.data(), false, fin.data(), budget.data()));
      } else {
        if (s->c64.cap < s->cap) DDA_TRY(s->c64.alloc(op, s->kind, s->max_iter, s->max_res_increase, s->cap));
        hipLaunchKernelGGL(k_load_rhs<double2>, G, B, 0, st, s->c64.v[0], s->r64, s->n, s->scale, s->mask);
        DDA_TRY_HIP(hipGetLastError());
        DDA_TRY(s->c64.solve(k, eps_in.data(), act.data(), false, fin.data(), budget.data()));
      }
      for (int r = 0; r < k; r++) {
        if (!act[r]) continue;
        iters[r] += fin[r].count;
        mvs[r] += fin[r].count * per + 1;
        if (fin[r].status == DDA_BREAKDOWN && fin[r].count == 0) {
          status[r] = DDA_BREAKDOWN;
          conv[r] = 1;
          msk[r] = 1;
        }
      }
      DDA_TRY_HIP(hipMemcpyAsync(s->scale, inv.data(), k * sizeof(double), hipMemcpyHostToDevice, st));
      DDA_TRY_HIP(hipMemcpyAsync(s->mask, msk.data(), k * sizeof(int), hipMemcpyHostToDevice, st));
      if (pass == 0)
        hipLaunchKernelGGL(k_add_corr<float2>, G, B, 0, st, s->x64, (const float2*)s->c32.v[2], s->n, s->scale,
                           s->mask);
      else
        hipLaunchKernelGGL(k_add_corr<double2>, G, B, 0, st, s->x64, (const double2*)s->c64.v[2], s->n, s->scale,
                           s->mask);
      DDA_TRY_HIP(hipGetLastError());
    }
    for (int r = 0; r < k; r++)
      if (act32[r] || act64[r]) outer[r]++;
  }
  s->last_iters = iters[0];
  hipLaunchKernelGGL(k_get_result<double>, G, B, 0, st, s->io, s->x64, ccs, op->mat_dev, d.nv, d.S, d.nmat,
                     gu * gu * gu, d.member);
  DDA_TRY_HIP(hipGetLastError());
  dda_status worst = DDA_OK;
  for (int r = 0; r < k; r++) {
    fill_info(info, r, iters[r], mvs[r], outer[r], std::sqrt(rel2[r]), std::sqrt(rel2[r]), status[r]);
    if (status[r] != DDA_OK && worst == DDA_OK) worst = (dda_status)status[r];
  }
  return worst;
}

}  // namespace

static void free_solver(dda_solver_s* s) {
  s->c64.release();
  s->c32.release();
  void* bufs[] = {s->x64, s->b64, s->r64, s->t64, s->part64, s->scale, s->mask, s->io};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
  if (s->op32) dda_op_destroy(s->op32);
  delete s;
}

extern "C" dda_status dda_solver_create_ex(dda_op_t op, int kind, uint64_t max_iter, uint64_t max_res_increase,
                                           int mode, dda_solver_t* out) {
  if (!op || !out || kind < 0 || kind > 3 ||
      (mode != DDA_MODE_FP64 && mode != DDA_MODE_MIXED && mode != DDA_MODE_FP32)) {
    set_error("dda_solver_create: null argument, unknown solver kind or mode");
    return DDA_INVALID_ARG;
  }
  DDA_TRY(bind_device(op->ctx));
  if (op->prec != 64 && mode != DDA_MODE_FP32) {
    set_error("dda_solver_create: FP64 and mixed modes need an FP64 operator (mixed mode builds its own FP32 twin)");
    return DDA_INVALID_ARG;
  }
  *out = nullptr;
  dda_solver_s* s = new dda_solver_s();
  s->op = op;
  s->kind = kind;
  s->mode = mode;
  const MvDims& d = op->dims;
  s->max_iter = max_iter ? max_iter : 3ull * d.Bx * d.By * d.Bz;  // DDA/DDAMain.cpp:640-644
  static const uint64_t defaults[4] = {50000, 30000, 10, 50000};  // QmrCs.cpp:39, BicgStab.cpp:51, Cgnr.cpp:34
  s->max_res_increase = max_res_increase ? max_res_increase : defaults[kind];
  s->inner_plateau = kInnerPlateau;
  s->n = 3ll * d.S;
  dda_status rc = DDA_OK;
  if (mode == DDA_MODE_MIXED || (mode == DDA_MODE_FP32 && op->prec == 64)) {
    rc = create_twin32(op, &s->op32);
    if (rc == DDA_OK && op->tuned) rc = dda_op_tune(s->op32, 1, nullptr, 0);
  }
  if (rc == DDA_OK) rc = alloc_buffers(s, 1);
  if (rc != DDA_OK) {
    free_solver(s);
    return rc;
  }
  *out = s;
  return DDA_OK;
}

extern "C" dda_status dda_solver_create(dda_op_t op, int kind, uint64_t max_iter, uint64_t max_res_increase,
                                        dda_solver_t* out) {
  return dda_solver_create_ex(op, kind, max_iter, max_res_increase, DDA_MODE_FP64, out);
}

extern "C" dda_status dda_solver_destroy(dda_solver_t s) {
  if (!s) return DDA_OK;
  if (s->op) (void)hipStreamSynchronize(s->op->ctx->stream);
  free_solver(s);
  return DDA_OK;
}

// A zero right-hand side has no solution to report: its output is zero.
static dda_status zero_invalid(dda_solver_s* s, int k, const dda_solve_info* info) {
  for (int r = 0; r < k; r++)
    if (info[r].status == DDA_INVALID_ARG)
      DDA_TRY_HIP(hipMemsetAsync(s->io + (size_t)r * s->n, 0, s->n * sizeof(double2), s->op->ctx->stream));
  return DDA_OK;
}

// dev: einc, start and pol_out are device buffers (FP64), else host arrays.
static dda_status solve_batch(dda_solver_t s, const void* einc, int nrhs, double eps, const void* start,
                              void* pol_out, dda_solve_info* info, bool dev) {
  if (!s || !einc || !pol_out || !(eps > 0) || nrhs < 1) {
    set_error("dda_solver_solve: null argument, eps <= 0 or nrhs < 1");
    return DDA_INVALID_ARG;
  }
  DDA_TRY(bind_device(s->op->ctx));
  dda_op_s* op = s->op;
  if (!op->cc_set) {
    set_error("dda_solver_solve: couple constants have not been set");
    return DDA_INVALID_ARG;
  }
  if (op->cc_rhs < nrhs) {
    if (op->cc_rhs != 1) {
      set_error("dda_solver_solve: couple constants set for fewer right-hand sides than solved");
      return DDA_INVALID_ARG;
    }
    // one set of couple constants for every right-hand side
    std::vector<double> rep;
    for (int r = 0; r < nrhs; r++) rep.insert(rep.end(), op->cc_sqrt_host.begin(), op->cc_sqrt_host.end());
    DDA_TRY(dda_op_set_cc_sqrt_batch(op, rep.data(), nrhs));
  }
  hipStream_t st = op->ctx->stream;
  DDA_TRY(alloc_buffers(s, nrhs));
  if (s->op32) {  // the FP32 twin of the mixed and FP32 modes follows the caller's operator
    DDA_TRY(dda_op_set_cc_sqrt_batch(s->op32, op->cc_sqrt_host.data(), nrhs));
    if (op->pack.nmember) DDA_TRY(dda_op_set_rhs_members(s->op32, op->pack.member_host.data(), nrhs));
    s->op32->prof = op->prof;
    s->op32->tune.split = op->tune.split;  // and its block split choice
    s->op32->split_rhs = op->split_rhs;
  }
  // a block of kBlockTuneRhs or more over a tuned operator: the operator that
  // runs the iterations is tuned for blocks too, once per shape (cached)
  dda_op_s* iter_op = s->op32 ? s->op32 : op;
  if (op->tuned && nrhs >= kBlockTuneRhs && !iter_op->tuned_blk) DDA_TRY(dda_op_tune(iter_op, nrhs, nullptr, 0));
  auto t0 = std::chrono::steady_clock::now();
  const size_t bytes = (size_t)nrhs * s->n * sizeof(double2);
  const hipMemcpyKind in_kind = dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
  DDA_TRY_HIP(hipMemcpyAsync(s->io, einc, bytes, in_kind, st));
  if (start) DDA_TRY_HIP(hipMemcpyAsync(s->io + (size_t)s->cap * s->n, start, bytes, in_kind, st));
  std::vector<dda_solve_info> inf(nrhs);
  dda_status rc = s->mode == DDA_MODE_MIXED  ? run_mixed(s, nrhs, eps, start != nullptr, inf.data())
                  : s->mode == DDA_MODE_FP32 ? run_fp32(s, nrhs, eps, start != nullptr, inf.data())
                                             : run_fp64(s, nrhs, eps, start != nullptr, inf.data());
  if (rc == DDA_HIP_ERROR || rc == DDA_OUT_OF_MEMORY || rc == DDA_NOT_SUPPORTED) return rc;
  DDA_TRY(zero_invalid(s, nrhs, inf.data()));
  DDA_TRY_HIP(hipMemcpyAsync(pol_out, s->io, bytes, dev ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, st));
  DDA_TRY_HIP(hipStreamSynchronize(st));
  const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  for (int r = 0; r < nrhs; r++) {
    inf[r].seconds = secs;
    if (info) info[r] = inf[r];
  }
  if (rc == DDA_BREAKDOWN) set_error("solver breakdown");
  if (rc == DDA_MAX_ITER) set_error("Too many iterations");
  if (rc == DDA_STAGNATION) set_error("Too many iterations w/o increase");
  if (rc == DDA_INVALID_ARG) set_error("a right-hand side is zero");
  return rc;
}

extern "C" dda_status dda_solver_solve_batch(dda_solver_t s, const double* einc, int nrhs, double eps,
                                             const double* start, double* pol_out, dda_solve_info* info) {
  return solve_batch(s, einc, nrhs, eps, start, pol_out, info, false);
}

extern "C" dda_status dda_solver_solve_batch_dev(dda_solver_t s, const void* einc_dev, int nrhs, double eps,
                                                 const void* start_dev, void* pol_dev, dda_solve_info* info) {
  return solve_batch(s, einc_dev, nrhs, eps, start_dev, pol_dev, info, true);
}

extern "C" dda_status dda_solver_solve(dda_solver_t s, const double* einc, double eps, const double* start,
                                       double* pol_out, dda_solve_info* info) {
  return dda_solver_solve_batch(s, einc, 1, eps, start, pol_out, info);
}

extern "C" dda_status dda_solver_history(dda_solver_t s, double* out, uint64_t n, uint64_t* written) {
  if (!s || (!out && n)) {
    set_error("dda_solver_history: null argument");
    return DDA_INVALID_ARG;
  }
  uint64_t m = 0;
  if (s->mode == DDA_MODE_MIXED) {
    m = std::min<uint64_t>(n, s->outer_hist.size());
    for (uint64_t i = 0; i < m; i++) out[i] = s->outer_hist[i];
  } else {
    const bool f32 = s->mode == DDA_MODE_FP32;
    const uint64_t cap = f32 ? s->c32.hist_cap : s->c64.hist_cap;
    const double* hist = f32 ? s->c32.hist : s->c64.hist;
    m = std::min<uint64_t>(std::min<uint64_t>(n, s->last_iters + 1), cap);
    if (m) {
      DDA_TRY_HIP(hipMemcpyAsync(out, hist, m * sizeof(double), hipMemcpyDeviceToHost, s->op->ctx->stream));
      DDA_TRY_HIP(hipStreamSynchronize(s->op->ctx->stream));
    }
  }
  if (written) *written = m;
  return DDA_OK;
}

extern "C" dda_status dda_solver_set_inner_floor(dda_solver_t s, double floor_) {
  if (!s || !(floor_ > 0 && floor_ < 1)) {
    set_error("dda_solver_set_inner_floor: floor must be in (0, 1)");
    return DDA_INVALID_ARG;
  }
  s->inner_floor = floor_;
  return DDA_OK;
}

extern "C" dda_status dda_solver_set_inner_plateau(dda_solver_t s, uint64_t window, double ratio) {
  if (!s || s->mode != DDA_MODE_MIXED || !(ratio > 0 && ratio <= 1)) {
    set_error("dda_solver_set_inner_plateau: a mixed-mode solver and a ratio in (0, 1]");
    return DDA_INVALID_ARG;
  }
  s->inner_plateau = window;
  s->inner_plateau_ratio = ratio;
  return DDA_OK;
}

extern "C" dda_status dda_solver_set_gpu_share(dda_solver_t s, double after_seconds, double share) {
  if (!s || !(after_seconds >= 0) || !(share > 0 && share <= 1)) {
    set_error("dda_solver_set_gpu_share: after_seconds >= 0 and a share in (0, 1]");
    return DDA_INVALID_ARG;
  }
  for (double* p : {&s->c64.share_after_s, &s->c32.share_after_s}) *p = after_seconds;
  for (double* p : {&s->c64.share, &s->c32.share}) *p = share;
  return DDA_OK;
}

extern "C" dda_status dda_solver_profile_read(dda_solver_t s, double ms[5], uint64_t counts[5], uint64_t* skipped) {
  if (!s) {
    set_error("dda_solver_profile_read: null solver");
    return DDA_INVALID_ARG;
  }
  return dda_op_profile_read(s->mode == DDA_MODE_FP64 ? s->op : op_fp32(s), ms, counts, skipped);
}

extern "C" dda_status dda_solver_profiling_run(dda_solver_t s, double us[5], double* median_us, double* spread_us) {
  if (!s || !us) {
    set_error("dda_solver_profiling_run: null argument");
    return DDA_INVALID_ARG;
  }
  DDA_TRY(bind_device(s->op->ctx));
  dda_op_s* op = s->op;
  if (!op->cc_set) {
    set_error("dda_solver_profiling_run: couple constants have not been set");
    return DDA_INVALID_ARG;
  }
  DDA_TRY(alloc_buffers(s, 1));
  if (s->op32) DDA_TRY(dda_op_set_cc_sqrt_batch(s->op32, op->cc_sqrt_host.data(), 1));
  // a fixed pseudo-random field (re, im uniform in [-1, 1) on every dipole):
  // the reference times iterations on whatever its vectors hold
  const MvDims& d = op->dims;
  std::vector<double2> e((size_t)s->n, make_double2(0, 0));
  uint64_t x = 0x5EEDull;
  auto next = [&]() {
    x ^= x << 13;
    x ^= x >> 7;
    x ^= x << 17;
    return (double)(x >> 11) * (2.0 / 9007199254740992.0) - 1.0;
  };
  for (int c = 0; c < 3; c++)
    for (int i = 0; i < d.nv; i++) e[(size_t)c * d.S + i] = make_double2(next(), next());
  hipStream_t st = op->ctx->stream;
  DDA_TRY_HIP(hipMemcpyAsync(s->io, e.data(), e.size() * sizeof(double2), hipMemcpyHostToDevice, st));
  const dim3 G(red_blocks(s->n), 1), B(kRedThreads);
  double t[5];
  SolverState fin;
  const double zero = 0.0;  // eps 0: the timed iterations never meet the tolerance
  const uint64_t big = ~0ull >> 1;
  if (s->mode == DDA_MODE_FP64) {
    hipLaunchKernelGGL(k_setup_b<double>, G, B, 0, st, s->c64.v[0], s->io, (const double2*)op->ccs, op->mat_dev,
                       d.nv, d.S, d.nmat, d.member);
    DDA_TRY_HIP(hipGetLastError());
    DDA_TRY(s->c64.solve(1, &zero, nullptr, false, &fin, &big, t, 5));
  } else {
    dda_op_s* o = op_fp32(s);
    hipLaunchKernelGGL(k_setup_b<float>, G, B, 0, st, s->c32.v[0], s->io, (const float2*)o->ccs, o->mat_dev, d.nv,
                       d.S, d.nmat, o->dims.member);
    DDA_TRY_HIP(hipGetLastError());
    DDA_TRY(s->c32.solve(1, &zero, nullptr, false, &fin, &big, t, 5));
  }
  if (fin.done && fin.status != DDA_OK) {
    set_error("dda_solver_profiling_run: the iteration stopped early (breakdown)");
    return (dda_status)fin.status;
  }
  for (int i = 0; i < 5; i++) us[i] = t[i];
  // median and the larger distance from it to the extremes, as the reference
  // reports them (IterativeSolverBase.cpp:126-134)
  double srt[5];
  std::copy(t, t + 5, srt);
  std::sort(srt, srt + 5);
  if (median_us) *median_us = srt[2];
  if (spread_us) *spread_us = std::max(srt[2] - srt[0], srt[4] - srt[2]);
  return DDA_OK;
}
