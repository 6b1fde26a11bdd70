// Device-resident Krylov solvers: QMR-CS, BiCGStab, CGNR and BiCG-CS.
//
// Replaces IterativeSolverBase::getPolVec (DDA/IterativeSolverBase.cpp:40-103),
// GpuIterativeSolver::initGeneral/getResult (DDA/GpuIterativeSolver.cpp:78-144),
// the update loops of DDA/GpuQmrCs.cpp:77-198, DDA/GpuBicgStab.cpp:44-125,
// DDA/GpuCgnr.cpp:41-81, DDA/GpuBicgCs.cpp:41-95 and the linear combinations and
// reductions of LinAlg/GpuLinComb.cpp. The arithmetic of each iteration follows
// the reference's CPU solvers line by line (DDA/QmrCs.cpp:78-168,
// DDA/BicgStab.cpp:66-112, DDA/Cgnr.cpp:43-73, DDA/BicgCs.cpp:58-84).
//
// Every scalar lives in device memory (SolverState, FP64). Reductions are
// two-stage with a fixed grid, so their summation order, and hence the result,
// depends only on the vector length. Scalar algebra runs in a one-block kernel.
// The loop test, the maximum-iteration and stagnation aborts
// (IterativeSolverBase.cpp:56-70) and the breakdown tests are evaluated on the
// device: once `done` is set every later kernel of the solve returns at entry,
// so the host can enqueue iterations ahead and only reads the flag once per
// chunk of iterations.

#include <chrono>
#include <cmath>
#include <cstring>

#include "internal.hpp"

namespace dda {

constexpr int kRedBlocks = 1024;
constexpr int kRedThreads = 256;
constexpr int kChunk = 16;  // most iterations enqueued between flag reads

struct SolverState {
  int done;     // 1 once the solve has finished for any reason
  int status;   // dda_status at exit
  int skip;     // done, or BiCGStab's early exit inside the current iteration
  int sel;      // initGeneral: 1 = start from x = 0
  unsigned long long count, counter, max_iter, max_res_increase, hist_cap;
  double residScale, epsB, inprodR, inprodR_init, ip1, temp;
  // QMR-CS (DDA/QmrCs.hpp VarType)
  double omega_old, omega_new, c_old, c_new;
  double2 beta, mBeta, tautilda, s_old, s_new;
  double2 alpha, mAlpha, zetaInv, mEtaZeta, mThetaZeta, betaInv, tau, cnot, normSNew;
  // BiCGStab / CGNR / BiCG-CS
  double2 ro_new, ro_old, omega, bbeta, mbo, mOmega, one, mOne;
  double denumOmega, abs_ro_new, cg_ro_new, cg_ro_old;
};

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

// Writes up to four per-block partial sums.
__device__ inline void write_partials(double* part, int nq, double q0, double q1, double q2, double q3) {
  double t0 = block_sum(q0);
  double t1 = nq > 1 ? block_sum(q1) : 0.0;
  double t2 = nq > 2 ? block_sum(q2) : 0.0;
  double t3 = nq > 3 ? block_sum(q3) : 0.0;
  if (threadIdx.x == 0) {
    part[0 * kRedBlocks + blockIdx.x] = t0;
    if (nq > 1) part[1 * kRedBlocks + blockIdx.x] = t1;
    if (nq > 2) part[2 * kRedBlocks + blockIdx.x] = t2;
    if (nq > 3) part[3 * kRedBlocks + blockIdx.x] = t3;
  }
}

// Second stage, run by a whole single block: fixed order.
__device__ inline double sum_partials(const double* part, int q) {
  double s = 0;
  for (int i = threadIdx.x; i < kRedBlocks; i += blockDim.x) s += part[q * kRedBlocks + i];
  return block_sum(s);
}

template <typename C> __device__ inline C coef(const double2* p) {
  return cmake<C>(p->x, p->y);
}

// ------------------------------------------------------------------------------------
// Vector kernels (grid kRedBlocks x kRedThreads, grid-stride)
// ------------------------------------------------------------------------------------

enum RedOp { RED_DOTU = 0, RED_DOTC = 1, RED_NORM = 2 };

// Partial sums of a.b (DOTU), a.conj(b) (DOTC) or |a|^2 (NORM).
template <typename R, int OP>
__global__ __launch_bounds__(kRedThreads) void k_reduce(const typename CxT<R>::T* __restrict__ a,
                                                        const typename CxT<R>::T* __restrict__ b, long long n,
                                                        double* part, const int* flag) {
  if (*flag) return;
  double s0 = 0, s1 = 0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const double ar = a[i].x, ai = a[i].y;
    if (OP == RED_NORM) {
      s0 += ar * ar + ai * ai;
    } else {
      const double br = b[i].x, bi = OP == RED_DOTC ? -(double)b[i].y : (double)b[i].y;
      s0 += ar * br - ai * bi;
      s1 += ar * bi + ai * br;
    }
  }
  write_partials(part, OP == RED_NORM ? 1 : 2, s0, s1, 0, 0);
}

// out = c1 v1 + c2 v2 + v3 (c = nullptr: the term is absent; v3 may be null),
// with an optional |out|^2 partial (quantity 0).
template <typename R>
__global__ __launch_bounds__(kRedThreads) void k_lincomb(typename CxT<R>::T* out, const typename CxT<R>::T* v1,
                                                         const double2* c1, const typename CxT<R>::T* v2,
                                                         const double2* c2, const typename CxT<R>::T* v3,
                                                         long long n, double* part, const int* flag) {
  typedef typename CxT<R>::T C;
  if (*flag) return;
  const C k1 = c1 ? coef<C>(c1) : czero<C>();
  const C k2 = c2 ? coef<C>(c2) : czero<C>();
  double s = 0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    C o = czero<C>();
    if (c1) o = cmul(v1[i], k1);
    if (c2) o = cadd(o, cmul(v2[i], k2));
    if (v3) o = cadd(o, v3[i]);
    out[i] = o;
    s += (double)o.x * o.x + (double)o.y * o.y;
  }
  if (part) write_partials(part, 1, s, 0, 0, 0);
}

// QMR-CS: vtilda = [vtilda * mBeta] + v * mAlpha + Av, with partials of
// vtilda.vtilda (quantities 0, 1) and |vtilda|^2 (2). QmrCs.cpp:96-104.
template <typename R>
__global__ __launch_bounds__(kRedThreads) void k_qmr_vt(typename CxT<R>::T* vt, const typename CxT<R>::T* __restrict__ v,
                                                        const typename CxT<R>::T* __restrict__ av, int first,
                                                        const SolverState* st, long long n, double* part) {
  typedef typename CxT<R>::T C;
  if (st->done) return;
  const C mb = coef<C>(&st->mBeta), ma = coef<C>(&st->mAlpha);
  double s0 = 0, s1 = 0, s2 = 0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    C o = cadd(cmul(v[i], ma), av[i]);
    if (!first) o = cadd(cmul(vt[i], mb), o);
    vt[i] = o;
    const double r = o.x, im = o.y;
    s0 += r * r - im * im;
    s1 += 2 * r * im;
    s2 += r * r + im * im;
  }
  write_partials(part, 3, s0, s1, s2, 0);
}

// QMR-CS update, QmrCs.cpp:150-166: p_dst = [p_old mThetaZeta] + [p_new mEtaZeta]
// + v zetaInv; x += p_dst tau; vt *= betaInv; r = r normSNew + vt cnot.
// mode 0: nr == 0, 1: nr == 1, 2: later. Partial |r|^2 (quantity 0).
template <typename R>
__global__ __launch_bounds__(kRedThreads) void k_qmr_upd(typename CxT<R>::T* p_dst, const typename CxT<R>::T* p_new,
                                                         const typename CxT<R>::T* __restrict__ v,
                                                         typename CxT<R>::T* vt, typename CxT<R>::T* x,
                                                         typename CxT<R>::T* r, int mode, const SolverState* st,
                                                         long long n, double* part) {
  typedef typename CxT<R>::T C;
  if (st->done) return;
  const C zi = coef<C>(&st->zetaInv), mez = coef<C>(&st->mEtaZeta), mtz = coef<C>(&st->mThetaZeta);
  const C tau = coef<C>(&st->tau), bi = coef<C>(&st->betaInv), cn = coef<C>(&st->cnot), ns = coef<C>(&st->normSNew);
  double s = 0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
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
  write_partials(part, 1, s, 0, 0, 0);
}

// b = sqrt(a) einc (GpuIterativeSolverInst.cl setupPvec), partial |b|^2.
template <typename R>
__global__ __launch_bounds__(kRedThreads) void k_setup_b(typename CxT<R>::T* b, const double2* __restrict__ einc,
                                                         const typename CxT<R>::T* __restrict__ ccs,
                                                         const uint8_t* __restrict__ mat, int nv, int S, double* part) {
  typedef typename CxT<R>::T C;
  double s = 0;
  const long long n = 3ll * S;
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x) {
    const int c = (int)(t / S), i = (int)(t - (long long)c * S);
    C o = czero<C>();
    if (i < nv) o = cmul(ccs[3 * mat[i] + c], cmake<C>(einc[t].x, einc[t].y));
    b[t] = o;
    s += (double)o.x * o.x + (double)o.y * o.y;
  }
  write_partials(part, 1, s, 0, 0, 0);
}

// x = start / sqrt(a) (setupXvecStartValue), start given as p in d^3 units.
template <typename R>
__global__ void k_setup_start(typename CxT<R>::T* x, const double2* __restrict__ start,
                              const typename CxT<R>::T* __restrict__ ccs, const uint8_t* __restrict__ mat, int nv,
                              int S, double scale) {
  typedef typename CxT<R>::T C;
  const long long n = 3ll * S;
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x) {
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
                             double scale) {
  typedef typename CxT<R>::T C;
  const long long n = 3ll * S;
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x) {
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
  const int sel = st->sel;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
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
                            const int* flag) {
  if (*flag) return;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

// ------------------------------------------------------------------------------------
// Scalar kernels (one block of kRedThreads)
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

__global__ __launch_bounds__(kRedThreads) void k_scalar(SolverState* st, const double* part, int step, int kind,
                                                        double eps, double* hist) {
  if (st->done && step >= S_LOOP_CHECK) return;
  // All threads take part in the reductions; thread 0 does the algebra.
  double q0 = 0, q1 = 0, q2 = 0;
  switch (step) {
    case S_INIT_TEMP:
    case S_INIT_R:
    case S_INIT_START:
    case S_END_ITER:
    case S_CG_RO:
    case S_CG_ALPHA:
      q0 = sum_partials(part, 0);
      break;
    case S_QMR_INIT:
    case S_QMR_ALPHA:
    case S_BS_RO:
    case S_BS_ALPHA:
    case S_BI_RO:
    case S_BI_MU:
      q0 = sum_partials(part, 0);
      q1 = sum_partials(part, 1);
      break;
    case S_QMR_GIVENS:
      q0 = sum_partials(part, 0);
      q1 = sum_partials(part, 1);
      q2 = sum_partials(part, 2);
      break;
    case S_BS_S:
      q0 = sum_partials(part, 0);
      break;
    case S_BS_OMEGA:
      q0 = sum_partials(part, 0);  // |Av|^2
      q1 = sum_partials(part, 1);  // s.conj(Av)
      q2 = sum_partials(part, 2);
      break;
    default:
      break;
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
      st->epsB = eps * eps / st->residScale;
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
      break;
    }
    // ---- BiCGStab, BicgStab.cpp:66-112 ----
    case S_BS_RO: {
      st->ro_new = make_double2(q0, q1);
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

// BiCGStab: x = p alpha + x, only on the early exit of an iteration.
template <typename R>
__global__ void k_bs_early(typename CxT<R>::T* x, const typename CxT<R>::T* __restrict__ p, long long n,
                           const SolverState* st) {
  typedef typename CxT<R>::T C;
  if (st->done || !st->skip) return;
  const C a = coef<C>(&st->alpha);
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    x[i] = cadd(cmul(p[i], a), x[i]);
}

// BiCGStab: p = p beta + v (-beta omega) + r (nr > 0) or p = r.
template <typename R>
__global__ __launch_bounds__(kRedThreads) void k_bs_p(typename CxT<R>::T* p, const typename CxT<R>::T* __restrict__ v,
                                                      const typename CxT<R>::T* __restrict__ r, int first,
                                                      const SolverState* st, long long n) {
  typedef typename CxT<R>::T C;
  if (st->done) return;
  const C b = coef<C>(&st->bbeta), mbo = coef<C>(&st->mbo);
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
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
  if (st->skip) return;
  double n2 = 0, dr = 0, di = 0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const double ar = av[i].x, ai = av[i].y, sr = s[i].x, si = s[i].y;
    n2 += ar * ar + ai * ai;
    dr += sr * ar + si * ai;
    di += si * ar - sr * ai;
  }
  write_partials(part, 3, n2, dr, di, 0);
}

// BiCGStab: x = p alpha + s omega + x; r = Av (-omega) + s; partial |r|^2.
template <typename R>
__global__ __launch_bounds__(kRedThreads) void k_bs_upd(typename CxT<R>::T* x, typename CxT<R>::T* r,
                                                        const typename CxT<R>::T* __restrict__ p,
                                                        const typename CxT<R>::T* __restrict__ s,
                                                        const typename CxT<R>::T* __restrict__ av, long long n,
                                                        double* part, SolverState* st) {
  typedef typename CxT<R>::T C;
  if (st->skip) return;
  const C a = coef<C>(&st->alpha), om = coef<C>(&st->omega), mom = coef<C>(&st->mOmega);
  double sum = 0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    x[i] = cadd(cadd(cmul(p[i], a), cmul(s[i], om)), x[i]);
    const C rn = cadd(cmul(av[i], mom), s[i]);
    r[i] = rn;
    sum += (double)rn.x * rn.x + (double)rn.y * rn.y;
  }
  write_partials(part, 1, sum, 0, 0, 0);
}

// BiCGStab end of iteration: ip1 from |r|^2 unless the early exit was taken.
__global__ __launch_bounds__(kRedThreads) void k_bs_fin(SolverState* st, const double* part) {
  if (st->done) return;
  const double q0 = sum_partials(part, 0);
  if (threadIdx.x == 0 && !st->skip) {
    st->ip1 = q0;
    st->ro_old = st->ro_new;
  }
}

// CGNR / BiCG: p = first ? a : p beta + a.
template <typename R>
__global__ void k_p_update(typename CxT<R>::T* p, const typename CxT<R>::T* __restrict__ a, int first,
                           const SolverState* st, long long n) {
  typedef typename CxT<R>::T C;
  if (st->done) return;
  const C b = coef<C>(&st->bbeta);
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    p[i] = first ? a[i] : cadd(cmul(p[i], b), a[i]);
}

// x = p alpha + x; r = Av mAlpha + r; partial |r|^2 (CGNR, BiCG-CS).
template <typename R>
__global__ __launch_bounds__(kRedThreads) void k_xr_update(typename CxT<R>::T* x, typename CxT<R>::T* r,
                                                           const typename CxT<R>::T* __restrict__ p,
                                                           const typename CxT<R>::T* __restrict__ av, long long n,
                                                           double* part, const SolverState* st) {
  typedef typename CxT<R>::T C;
  if (st->done) return;
  const C a = coef<C>(&st->alpha), ma = coef<C>(&st->mAlpha);
  double sum = 0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    x[i] = cadd(cmul(p[i], a), x[i]);
    const C rn = cadd(cmul(av[i], ma), r[i]);
    r[i] = rn;
    sum += (double)rn.x * rn.x + (double)rn.y * rn.y;
  }
  write_partials(part, 1, sum, 0, 0, 0);
}

// Mixed precision helpers: r32 = (float)(r64 * scale); x64 += (double)d32 * scale;
// r64 = b64 - t64 with partial |r|^2.
__global__ void k_to_f32(float2* out, const double2* __restrict__ in, long long n, double scale) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    out[i] = make_float2((float)(in[i].x * scale), (float)(in[i].y * scale));
}
__global__ void k_add_f32(double2* x, const float2* __restrict__ d, long long n, double scale) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    x[i].x += (double)d[i].x * scale;
    x[i].y += (double)d[i].y * scale;
  }
}
__global__ __launch_bounds__(kRedThreads) void k_resid64(double2* r, const double2* __restrict__ b,
                                                         const double2* __restrict__ t, long long n, double* part) {
  double s = 0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const double2 v = make_double2(b[i].x - t[i].x, b[i].y - t[i].y);
    r[i] = v;
    s += v.x * v.x + v.y * v.y;
  }
  write_partials(part, 1, s, 0, 0, 0);
}

// One Krylov solve A x = b on the device for an operator of precision R.
// v[0] = b (set by the caller), v[2] = x (x0 when have_x0), v[1] = r, v[3] = Av,
// v[4..7] solver work vectors.
template <typename R>
struct Core {
  typedef typename CxT<R>::T C;
  dda_op_s* op = nullptr;
  int kind = 0;
  hipStream_t st = nullptr;
  long long n = 0;
  SolverState* dst = nullptr;       // device
  SolverState* host2 = nullptr;     // pinned, two slots
  double* part = nullptr;
  double* hist = nullptr;
  uint64_t hist_cap = 0;
  uint64_t max_iter = 0, max_res_increase = 0;
  C* v[8] = {};

  dim3 g() const { return dim3(kRedBlocks); }
  dim3 b() const { return dim3(kRedThreads); }
  dda_status matvec(const void* x, void* y, int conj, const int* flag) {
    return launch_matvec<R>(op, x, y, conj, st, flag);
  }
  void scalar(int step, double eps = 0) {
    hipLaunchKernelGGL(k_scalar, dim3(1), dim3(kRedThreads), 0, st, dst, part, step, kind, eps, hist);
  }
  void reduce(int opk, const C* a, const C* bb, const int* flag) {
    if (opk == RED_NORM)
      hipLaunchKernelGGL((k_reduce<R, RED_NORM>), g(), b(), 0, st, a, bb, n, part, flag);
    else if (opk == RED_DOTU)
      hipLaunchKernelGGL((k_reduce<R, RED_DOTU>), g(), b(), 0, st, a, bb, n, part, flag);
    else
      hipLaunchKernelGGL((k_reduce<R, RED_DOTC>), g(), b(), 0, st, a, bb, n, part, flag);
  }
  void lincomb(C* out, const C* v1, const double2* c1, const C* v2, const double2* c2, const C* v3, bool withpart,
               const int* flag) {
    hipLaunchKernelGGL(k_lincomb<R>, g(), b(), 0, st, out, v1, c1, v2, c2, v3, n, withpart ? part : nullptr, flag);
  }

  dda_status alloc(dda_op_s* o, int k, uint64_t mi, uint64_t mri) {
    op = o;
    kind = k;
    st = o->ctx->stream;
    n = 3ll * o->dims.S;
    max_iter = mi;
    max_res_increase = mri;
    hist_cap = std::min<uint64_t>(mi + 1, 1u << 20);
    DDA_TRY_HIP(hipMalloc(&dst, sizeof(SolverState)));
    DDA_TRY_HIP(hipHostMalloc((void**)&host2, 2 * sizeof(SolverState), hipHostMallocDefault));
    DDA_TRY_HIP(hipMalloc(&part, 4 * kRedBlocks * sizeof(double)));
    DDA_TRY_HIP(hipMalloc(&hist, hist_cap * sizeof(double)));
    for (int i = 0; i < 8; i++) DDA_TRY_HIP(hipMalloc(&v[i], n * sizeof(C)));
    return DDA_OK;
  }
  void release() {
    if (dst) (void)hipFree(dst);
    if (host2) (void)hipHostFree(host2);
    if (part) (void)hipFree(part);
    if (hist) (void)hipFree(hist);
    for (C* p : v)
      if (p) (void)hipFree(p);
    dst = nullptr;
    host2 = nullptr;
    part = nullptr;
    hist = nullptr;
    for (auto& p : v) p = nullptr;
  }

  // Sum of partial quantity 0 on the host (synchronises).
  dda_status read_sum(double* out) {
    std::vector<double> h(kRedBlocks);
    DDA_TRY_HIP(hipMemcpyAsync(h.data(), part, kRedBlocks * sizeof(double), hipMemcpyDeviceToHost, st));
    DDA_TRY_HIP(hipStreamSynchronize(st));
    double t = 0;
    for (double x : h) t += x;
    *out = t;
    return DDA_OK;
  }

  // initGeneral + getPolVec loop (GpuIterativeSolver.cpp:78-134,
  // IterativeSolverBase.cpp:40-103); leaves x in v[2] and the final state in *fin.
  dda_status solve(double eps, bool have_x0, SolverState* fin) {
    const int* done = &dst->done;
    const int* skip = &dst->skip;
    C* bv = v[0];
    C* r = v[1];
    C* x = v[2];
    C* av = v[3];
    SolverState h;
    std::memset(&h, 0, sizeof h);
    h.max_iter = max_iter;
    h.max_res_increase = max_res_increase;
    h.hist_cap = hist_cap;
    DDA_TRY_HIP(hipMemcpyAsync(dst, &h, sizeof h, hipMemcpyHostToDevice, st));
    reduce(RED_NORM, bv, bv, skip);
    scalar(S_INIT_TEMP, eps);
    if (have_x0) {
      DDA_TRY(matvec(x, av, 0, nullptr));
      lincomb(r, av, &dst->mOne, nullptr, nullptr, bv, true, skip);  // r = b - A x0
      scalar(S_INIT_START);
    } else {
      DDA_TRY(matvec(bv, av, 0, nullptr));
      lincomb(r, av, &dst->mOne, nullptr, nullptr, bv, true, skip);  // r = b - A b
      scalar(S_INIT_R);
      hipLaunchKernelGGL(k_select_x0<R>, g(), b(), 0, st, x, r, bv, n, dst);
    }
    if (kind == DDA_SOLVER_QMR_CS) {
      reduce(RED_DOTU, r, r, skip);
      scalar(S_QMR_INIT);
      lincomb(v[4], r, &dst->betaInv, nullptr, nullptr, nullptr, false, skip);  // v = r / beta
    } else if (kind == DDA_SOLVER_BICGSTAB) {
      hipLaunchKernelGGL(k_copy_flag<R>, g(), b(), 0, st, v[7], r, n, skip);  // rtilda = r
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
          DDA_TRY(matvec(vv, av, 0, done));
          reduce(RED_DOTU, vv, av, done);
          scalar(S_QMR_ALPHA);
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
          reduce(RED_DOTC, r, rt, done);
          scalar(S_BS_RO);
          hipLaunchKernelGGL(k_bs_p<R>, g(), b(), 0, st, p, vv, r, nr == 0 ? 1 : 0, dst, n);
          DDA_TRY(matvec(p, vv, 0, done));
          reduce(RED_DOTC, vv, rt, done);
          scalar(S_BS_ALPHA);
          lincomb(sv, vv, &dst->mAlpha, nullptr, nullptr, r, true, done);  // s = v (-alpha) + r
          scalar(S_BS_S);
          hipLaunchKernelGGL(k_bs_early<R>, g(), b(), 0, st, x, p, n, dst);
          DDA_TRY(matvec(sv, av, 0, skip));
          hipLaunchKernelGGL(k_bs_om<R>, g(), b(), 0, st, av, sv, n, part, dst);
          scalar(S_BS_OMEGA);
          hipLaunchKernelGGL(k_bs_upd<R>, g(), b(), 0, st, x, r, p, sv, av, n, part, dst);
          hipLaunchKernelGGL(k_bs_fin, dim3(1), b(), 0, st, dst, part);
          scalar(S_END_ITER);
          break;
        }
        case DDA_SOLVER_CGNR: {
          C* p = v[4];
          DDA_TRY(matvec(r, av, 1, done));  // A^H r
          reduce(RED_NORM, av, av, done);
          scalar(S_CG_RO);
          hipLaunchKernelGGL(k_p_update<R>, g(), b(), 0, st, p, av, nr == 0 ? 1 : 0, dst, n);
          DDA_TRY(matvec(p, av, 0, done));
          reduce(RED_NORM, av, av, done);
          scalar(S_CG_ALPHA);
          hipLaunchKernelGGL(k_xr_update<R>, g(), b(), 0, st, x, r, p, av, n, part, dst);
          scalar(S_END_ITER);
          break;
        }
        case DDA_SOLVER_BICG_CS: {
          C* p = v[4];
          reduce(RED_DOTU, r, r, done);
          scalar(S_BI_RO);
          hipLaunchKernelGGL(k_p_update<R>, g(), b(), 0, st, p, r, nr == 0 ? 1 : 0, dst, n);
          DDA_TRY(matvec(p, av, 0, done));
          reduce(RED_DOTU, p, av, done);
          scalar(S_BI_MU);
          hipLaunchKernelGGL(k_xr_update<R>, g(), b(), 0, st, x, r, p, av, n, part, dst);
          scalar(S_END_ITER);
          break;
        }
      }
      DDA_TRY_HIP(hipGetLastError());
      nr++;
      return DDA_OK;
    };

    // Chunks of iterations are enqueued ahead of the flag reads: chunk c+1 is
    // queued before the state copied after chunk c is examined, so the GPU
    // never waits on the host, and the chunk length follows the observed
    // convergence rate so that few iterations are queued past the one that
    // meets the tolerance (those run as no-ops).
    hipEvent_t ev[2];
    DDA_TRY_HIP(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
    DDA_TRY_HIP(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
    DDA_TRY_HIP(hipMemcpyAsync(&host2[0], dst, sizeof(SolverState), hipMemcpyDeviceToHost, st));
    DDA_TRY_HIP(hipEventRecord(ev[0], st));
    int next_len = 2, chunk = 0;
    double prev_r = -1;
    uint64_t prev_count = 0;
    bool finished = false;
    dda_status rc = DDA_OK;
    while (!finished && rc == DDA_OK) {
      for (int k = 0; k < next_len && rc == DDA_OK; k++) rc = enqueue_iteration();
      if (rc != DDA_OK) break;
      const int slot = (chunk + 1) & 1;
      DDA_TRY_HIP(hipMemcpyAsync(&host2[slot], dst, sizeof(SolverState), hipMemcpyDeviceToHost, st));
      DDA_TRY_HIP(hipEventRecord(ev[slot], st));
      const int prev = chunk & 1;
      DDA_TRY_HIP(hipEventSynchronize(ev[prev]));
      const SolverState& hs = host2[prev];
      if (hs.done) {
        finished = true;
      } else {
        int len = kChunk;
        if (prev_r > 0 && hs.inprodR > 0 && hs.count > prev_count && hs.inprodR < prev_r && hs.epsB > 0) {
          const double rate = std::log(prev_r / hs.inprodR) / (double)(hs.count - prev_count);
          const double need = std::log(hs.inprodR / hs.epsB) / rate - (double)next_len;
          len = need < 1 ? 1 : (need > kChunk ? kChunk : (int)std::ceil(need));
        }
        prev_r = hs.inprodR;
        prev_count = hs.count;
        next_len = len;
      }
      chunk++;
    }
    (void)hipStreamSynchronize(st);
    (void)hipEventDestroy(ev[0]);
    (void)hipEventDestroy(ev[1]);
    if (rc != DDA_OK) return rc;
    DDA_TRY_HIP(hipMemcpy(fin, dst, sizeof(SolverState), hipMemcpyDeviceToHost));
    return DDA_OK;
  }
};

template struct Core<float>;
template struct Core<double>;

}  // namespace dda

using namespace dda;

struct dda_solver_s {
  dda_op_s* op = nullptr;        // FP64 operator (caller's)
  dda_op_s* op32 = nullptr;      // FP32 twin (mixed mode), owned
  int kind = 0;
  int mode = DDA_MODE_FP64;
  uint64_t max_iter = 0, max_res_increase = 0;
  long long n = 0;
  Core<double> c64;              // FP64 mode
  Core<float> c32;               // mixed mode inner solver
  double2* x64 = nullptr;        // mixed mode: FP64 b, x, r, t
  double2* b64 = nullptr;
  double2* r64 = nullptr;
  double2* t64 = nullptr;
  double* part64 = nullptr;
  double2* io = nullptr;         // FP64 staging for einc / start / result
  uint64_t last_iters = 0;
  std::vector<double> outer_hist;
  double inner_floor = 1e-4;
};

namespace {

dda_status true_residual64(dda_solver_s* s, double2* x, double2* b, double* rel2, double temp) {
  Core<double>& K = s->c64;
  (void)K;
  hipStream_t st = s->op->ctx->stream;
  DDA_TRY(launch_matvec<double>(s->op, x, s->t64, 0, st, nullptr));
  hipLaunchKernelGGL(k_resid64, dim3(kRedBlocks), dim3(kRedThreads), 0, st, s->r64, b, s->t64, s->n, s->part64);
  DDA_TRY_HIP(hipGetLastError());
  std::vector<double> h(kRedBlocks);
  DDA_TRY_HIP(hipMemcpyAsync(h.data(), s->part64, kRedBlocks * sizeof(double), hipMemcpyDeviceToHost, st));
  DDA_TRY_HIP(hipStreamSynchronize(st));
  double t = 0;
  for (double v : h) t += v;
  *rel2 = t / temp;
  return DDA_OK;
}

dda_status run_fp64(dda_solver_s* s, double eps, bool have_start, dda_solve_info* info) {
  dda_op_s* op = s->op;
  hipStream_t st = op->ctx->stream;
  const MvDims& d = op->dims;
  Core<double>& K = s->c64;
  const double2* ccs = (const double2*)op->ccs;
  const double gu = op->grid_unit;
  hipLaunchKernelGGL(k_setup_b<double>, dim3(kRedBlocks), dim3(kRedThreads), 0, st, K.v[0], s->io, ccs, op->mat_dev,
                     d.nv, d.S, K.part);
  if (have_start)
    hipLaunchKernelGGL(k_setup_start<double>, dim3(kRedBlocks), dim3(kRedThreads), 0, st, K.v[2], s->io + s->n, ccs,
                       op->mat_dev, d.nv, d.S, 1.0 / (gu * gu * gu));
  DDA_TRY_HIP(hipGetLastError());
  SolverState fin;
  DDA_TRY(K.solve(eps, have_start, &fin));
  s->last_iters = fin.count;
  hipLaunchKernelGGL(k_get_result<double>, dim3(kRedBlocks), dim3(kRedThreads), 0, st, s->io, K.v[2], ccs,
                     op->mat_dev, d.nv, d.S, gu * gu * gu);
  DDA_TRY_HIP(hipGetLastError());
  double rel2 = 0;
  DDA_TRY(true_residual64(s, K.v[2], K.v[0], &rel2, fin.temp));
  if (info) {
    info->iterations = fin.count;
    info->rel_residual = std::sqrt(fin.residScale * fin.inprodR);
    info->true_rel_residual = std::sqrt(rel2);
    info->refinements = 0;
    const uint64_t per = s->kind == DDA_SOLVER_QMR_CS || s->kind == DDA_SOLVER_BICG_CS ? 1 : 2;
    info->matvecs = fin.count * per + 2;
  }
  return (dda_status)fin.status;
}

// FP64 iterative refinement around an FP32 Krylov solve:
//   r = b - A64 x; stop when |r|^2 < eps^2 |b|^2; solve A32 d = r / |r| to
//   max(floor, eps |b| / (2 |r|)); x += |r| d.
dda_status run_mixed(dda_solver_s* s, double eps, bool have_start, dda_solve_info* info) {
  dda_op_s* op = s->op;
  hipStream_t st = op->ctx->stream;
  const MvDims& d = op->dims;
  Core<float>& K = s->c32;
  const double2* ccs = (const double2*)op->ccs;
  const double gu = op->grid_unit;
  const dim3 G(kRedBlocks), B(kRedThreads);
  hipLaunchKernelGGL(k_setup_b<double>, G, B, 0, st, s->b64, s->io, ccs, op->mat_dev, d.nv, d.S, s->part64);
  DDA_TRY_HIP(hipGetLastError());
  std::vector<double> h(kRedBlocks);
  DDA_TRY_HIP(hipMemcpyAsync(h.data(), s->part64, kRedBlocks * sizeof(double), hipMemcpyDeviceToHost, st));
  DDA_TRY_HIP(hipStreamSynchronize(st));
  double temp = 0;
  for (double v : h) temp += v;
  if (!(temp > 0)) {
    set_error("zero right-hand side");
    return DDA_INVALID_ARG;
  }
  if (have_start)
    hipLaunchKernelGGL(k_setup_start<double>, G, B, 0, st, s->x64, s->io + s->n, ccs, op->mat_dev, d.nv, d.S,
                       1.0 / (gu * gu * gu));
  else
    DDA_TRY_HIP(hipMemsetAsync(s->x64, 0, s->n * sizeof(double2), st));
  s->outer_hist.clear();
  uint64_t iters = 0, outer = 0, matvecs = 0;
  dda_status status = DDA_OK;
  double rel2 = 1.0;
  const uint64_t max_outer = 40;
  for (;;) {
    if (outer == 0 && !have_start) {
      DDA_TRY_HIP(hipMemcpyAsync(s->r64, s->b64, s->n * sizeof(double2), hipMemcpyDeviceToDevice, st));
      rel2 = 1.0;
    } else {
      DDA_TRY(true_residual64(s, s->x64, s->b64, &rel2, temp));
      matvecs++;
    }
    s->outer_hist.push_back(rel2);
    if (rel2 < eps * eps) break;
    if (outer >= max_outer || iters >= s->max_iter) {
      status = DDA_MAX_ITER;
      break;
    }
    const double rn = std::sqrt(rel2 * temp);
    hipLaunchKernelGGL(k_to_f32, G, B, 0, st, (float2*)K.v[0], s->r64, s->n, 1.0 / rn);
    DDA_TRY_HIP(hipGetLastError());
    const double eps_in = std::max(s->inner_floor, 0.5 * eps / std::sqrt(rel2));
    K.max_iter = s->max_iter - iters;
    SolverState fin;
    DDA_TRY(K.solve(eps_in, false, &fin));
    iters += fin.count;
    const uint64_t per = s->kind == DDA_SOLVER_QMR_CS || s->kind == DDA_SOLVER_BICG_CS ? 1 : 2;
    matvecs += fin.count * per + 1;
    outer++;
    if (fin.status == DDA_BREAKDOWN && fin.count == 0) {
      status = DDA_BREAKDOWN;
      break;
    }
    hipLaunchKernelGGL(k_add_f32, G, B, 0, st, s->x64, (const float2*)K.v[2], s->n, rn);
    DDA_TRY_HIP(hipGetLastError());
  }
  s->last_iters = iters;
  hipLaunchKernelGGL(k_get_result<double>, G, B, 0, st, s->io, s->x64, ccs, op->mat_dev, d.nv, d.S, gu * gu * gu);
  DDA_TRY_HIP(hipGetLastError());
  if (info) {
    info->iterations = iters;
    info->rel_residual = std::sqrt(rel2);
    info->true_rel_residual = std::sqrt(rel2);
    info->refinements = outer;
    info->matvecs = matvecs;
  }
  return status;
}

}  // namespace

static void free_solver(dda_solver_s* s) {
  s->c64.release();
  s->c32.release();
  void* bufs[] = {s->x64, s->b64, s->r64, s->t64, s->part64, s->io};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
  if (s->op32) dda_op_destroy(s->op32);
  delete s;
}

extern "C" dda_status dda_solver_create_ex(dda_op_t op, int kind, uint64_t max_iter, uint64_t max_res_increase,
                                           int mode, dda_solver_t* out) {
  if (!op || !out || kind < 0 || kind > 3 || (mode != DDA_MODE_FP64 && mode != DDA_MODE_MIXED)) {
    set_error("dda_solver_create: null argument, unknown solver kind or mode");
    return DDA_INVALID_ARG;
  }
  if (op->prec != 64) {
    set_error("dda_solver_create: the operator must be FP64 (mixed mode builds its own FP32 twin)");
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
  s->n = 3ll * d.S;
  dda_status rc = DDA_OK;
  auto dm = [&](void** p, size_t bytes) {
    if (rc != DDA_OK) return;
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
      set_error(std::string("dda_solver_create: ") + hipGetErrorString(e));
      rc = e == hipErrorOutOfMemory ? DDA_OUT_OF_MEMORY : DDA_HIP_ERROR;
    }
  };
  dm((void**)&s->io, 2 * s->n * sizeof(double2));
  dm((void**)&s->t64, s->n * sizeof(double2));
  dm((void**)&s->r64, s->n * sizeof(double2));
  dm((void**)&s->part64, 4 * kRedBlocks * sizeof(double));
  if (rc == DDA_OK && mode == DDA_MODE_FP64) rc = s->c64.alloc(op, kind, s->max_iter, s->max_res_increase);
  if (rc == DDA_OK && mode == DDA_MODE_MIXED) {
    dm((void**)&s->x64, s->n * sizeof(double2));
    dm((void**)&s->b64, s->n * sizeof(double2));
    if (rc == DDA_OK) {
      dda_op_desc desc;
      std::memset(&desc, 0, sizeof desc);
      desc.box[0] = d.Bx;
      desc.box[1] = d.By;
      desc.box[2] = d.Bz;
      desc.grid[0] = d.Gx;
      desc.grid[1] = d.Gy;
      desc.grid[2] = d.Gz;
      desc.nv = d.nv;
      desc.vec_stride = d.S;
      desc.pos = op->pos.data();
      desc.mat = op->mat.data();
      desc.nmat = op->nmat;
      desc.grid_unit = op->grid_unit;
      desc.wave_num = op->wave_num;
      rc = dda_op_create(op->ctx, &desc, 32, &s->op32);
    }
    if (rc == DDA_OK) rc = s->c32.alloc(s->op32, kind, s->max_iter, s->max_res_increase);
  }
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

extern "C" dda_status dda_solver_solve(dda_solver_t s, const double* einc, double eps, const double* start,
                                       double* pol_out, dda_solve_info* info) {
  if (!s || !einc || !pol_out || !(eps > 0)) {
    set_error("dda_solver_solve: null argument or eps <= 0");
    return DDA_INVALID_ARG;
  }
  dda_op_s* op = s->op;
  if (!op->cc_set) {
    set_error("dda_solver_solve: dda_op_set_cc_sqrt has not been called");
    return DDA_INVALID_ARG;
  }
  auto t0 = std::chrono::steady_clock::now();
  hipStream_t st = op->ctx->stream;
  if (s->mode == DDA_MODE_MIXED) {
    DDA_TRY(dda_op_set_cc_sqrt(s->op32, op->cc_sqrt_host.data()));
    s->op32->prof = op->prof;
  }
  DDA_TRY_HIP(hipMemcpyAsync(s->io, einc, s->n * sizeof(double2), hipMemcpyHostToDevice, st));
  if (start) DDA_TRY_HIP(hipMemcpyAsync(s->io + s->n, start, s->n * sizeof(double2), hipMemcpyHostToDevice, st));
  dda_status rc = s->mode == DDA_MODE_MIXED ? run_mixed(s, eps, start != nullptr, info)
                                            : run_fp64(s, eps, start != nullptr, info);
  if (rc == DDA_HIP_ERROR || rc == DDA_OUT_OF_MEMORY || rc == DDA_INVALID_ARG || rc == DDA_NOT_SUPPORTED) return rc;
  DDA_TRY_HIP(hipMemcpyAsync(pol_out, s->io, s->n * sizeof(double2), hipMemcpyDeviceToHost, st));
  DDA_TRY_HIP(hipStreamSynchronize(st));
  if (info) info->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (rc == DDA_BREAKDOWN) set_error("solver breakdown");
  if (rc == DDA_MAX_ITER) set_error("Too many iterations");
  if (rc == DDA_STAGNATION) set_error("Too many iterations w/o increase");
  return rc;
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
    m = std::min<uint64_t>(std::min<uint64_t>(n, s->last_iters + 1), s->c64.hist_cap);
    if (m) DDA_TRY_HIP(hipMemcpy(out, s->c64.hist, m * sizeof(double), hipMemcpyDeviceToHost));
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

extern "C" dda_status dda_solver_profile_read(dda_solver_t s, double ms[5], uint64_t counts[5], uint64_t* skipped) {
  if (!s) {
    set_error("dda_solver_profile_read: null solver");
    return DDA_INVALID_ARG;
  }
  return dda_op_profile_read(s->mode == DDA_MODE_MIXED ? s->op32 : s->op, ms, counts, skipped);
}
