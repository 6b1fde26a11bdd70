// The steps either side of the solve, on the device: the incident plane wave
// per right-hand side (DDA/Beam.cpp:110-118), LDR/CM polarisabilities
// (DDA/PolarizabilityDescription.cpp:65-110), Cext and Cabs
// (DDA/Beam.cpp:58-68, DDA/AbsCross.cpp:27-48), scattered far fields
// (DDA/CpuFieldCalculator.cpp:48-98) and Jones matrices on an angle grid
// (DDA/FarFieldCalc.cpp:119-174, EMSim/JonesCalculus.cpp:74-86). The reference
// evaluates the far field one direction per launch with a host read-back
// (DDA/GpuFieldCalculator.cpp:64-72); here all directions and right-hand sides
// go through one two-stage reduction.
#include <cmath>
#include <complex>
#include <cstring>
#include <mutex>
#include <vector>

#include "internal.hpp"

// The fixed device scratch of dda_cross_sections_noalloc: one copy of this
// array per GPU the library runs on, held until the process ends. Host code
// never passes the array's name to a kernel or a copy (on the host it is a
// stand-in, not a device address); it uses the per-device address that
// noalloc_scratch_setup takes with hipGetSymbolAddress.
static __device__ __align__(256) double scratch[DDA_NOALLOC_SCRATCH_DOUBLES];

namespace dda {
namespace {

constexpr int kObsThreads = 256;
constexpr double kPi = 3.14159265358979323846264338327950288;

// Grid point (0,0,0) sits at -(box - 1) / 2 grid units (DipoleGeometry::moveToCenter).
struct Frame {
  double ox, oy, oz, d;
};

// einc[r][c][i] = exp(i k r_i . prop_r) pol_r[c] for i < nv, 0 up to S.
__global__ void k_incident(double2* einc, const uint32_t* __restrict__ pos, const double* __restrict__ dirs, int nv,
                           int S, double k, Frame f) {
  const int r = blockIdx.y;
  const double* pr = dirs + 6 * r;  // prop[3], pol[3]
  double2* e = einc + (long long)r * 3 * S;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < S; i += gridDim.x * blockDim.x) {
    double2 ph = make_double2(0, 0);
    if (i < nv) {
      const double x = f.ox + pos[3 * i] * f.d, y = f.oy + pos[3 * i + 1] * f.d, z = f.oz + pos[3 * i + 2] * f.d;
      double s, c;
      sincos(k * (x * pr[0] + y * pr[1] + z * pr[2]), &s, &c);
      ph = make_double2(c, s);
    }
    for (int c = 0; c < 3; c++) e[(long long)c * S + i] = make_double2(ph.x * pr[3 + c], ph.y * pr[3 + c]);
  }
}

__device__ inline double block_sum(double v) {
  __shared__ double sh[kObsThreads / 64];
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  double t = 0;
  if (threadIdx.x == 0)
    for (int w = 0; w < kObsThreads / 64; w++) t += sh[w];
  __syncthreads();
  return t;
}

// Per-RHS partials of sum Im(p . conj(E)) and sum multDr |p|^2, where
// multDr = -Im(1/alpha) - 2/3 k^3 comes from mdr [nrhs][nmat][3] or, when mdr
// is null, from alpha itself (ccd [nrhs][nmat][3] complex, nullable: then the
// second sum is 0).
__global__ __launch_bounds__(kObsThreads) void k_cross(const double2* __restrict__ pol,
                                                       const double2* __restrict__ einc,
                                                       const uint8_t* __restrict__ mat,
                                                       const double* __restrict__ mdr,
                                                       const double2* __restrict__ ccd, double k3, int nv, int S,
                                                       int nmat, double* part, const int* member) {
  const int r = blockIdx.y;
  if (member) mat += (size_t)member[r] * nv;
  const double2* p = pol + (long long)r * 3 * S;
  const double2* e = einc + (long long)r * 3 * S;
  const double* md = mdr ? mdr + (long long)r * 3 * nmat : nullptr;
  const double2* ca = ccd ? ccd + (long long)r * 3 * nmat : nullptr;
  double se = 0, sa = 0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nv; i += gridDim.x * blockDim.x) {
    const int m = mat[i];
    for (int c = 0; c < 3; c++) {
      const double2 pv = p[(long long)c * S + i], ev = e[(long long)c * S + i];
      se += pv.y * ev.x - pv.x * ev.y;
      double w = 0;
      if (md) {
        w = md[3 * m + c];
      } else if (ca) {
        const double2 a = ca[3 * m + c];
        w = a.y / (a.x * a.x + a.y * a.y) - k3;  // -Im(1/a) - 2/3 k^3
      }
      sa += w * (pv.x * pv.x + pv.y * pv.y);
    }
  }
  const double te = block_sum(se), ta = block_sum(sa);
  if (threadIdx.x == 0) {
    part[((long long)r * gridDim.x + blockIdx.x) * 2] = te;
    part[((long long)r * gridDim.x + blockIdx.x) * 2 + 1] = ta;
  }
}

// Second stage of k_cross, one thread per right-hand side, blocks summed in
// order: out[r] = scale * sum_b part[r][b][0], out[nrhs + r] = scale * sum_b part[r][b][1].
__global__ void k_cross_final(const double* __restrict__ part, int nb, int nrhs, double scale, double* out) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrhs) return;
  double se = 0, sa = 0;
  for (int b = 0; b < nb; b++) {
    se += part[((long long)r * nb + b) * 2];
    sa += part[((long long)r * nb + b) * 2 + 1];
  }
  out[r] = scale * se;
  out[nrhs + r] = scale * sa;
}

// Far-field sums S_q = sum_j p_{rhs_q, j} exp(-i kd n_q . pos_j), first stage:
// block (chunk, query tile) -> partials[chunk][q][6].
constexpr int kQT = 4;
__global__ __launch_bounds__(kObsThreads) void k_far_partial(const double2* __restrict__ pol,
                                                             const uint32_t* __restrict__ pos,
                                                             const double* __restrict__ q, int nq, int nv, int S,
                                                             double kd, int chunk, double* part) {
  const int q0 = blockIdx.y * kQT;
  const int j0 = blockIdx.x * chunk, j1 = min(nv, j0 + chunk);
  double acc[kQT][6];
  double n[kQT][3];
  int rhs[kQT];
#pragma unroll
  for (int t = 0; t < kQT; t++) {
    const int qq = min(q0 + t, nq - 1);
    rhs[t] = (int)q[4 * qq];
    for (int a = 0; a < 3; a++) n[t][a] = -kd * q[4 * qq + 1 + a];
    for (int a = 0; a < 6; a++) acc[t][a] = 0;
  }
  for (int j = j0 + threadIdx.x; j < j1; j += blockDim.x) {
    const double x = pos[3 * j], y = pos[3 * j + 1], z = pos[3 * j + 2];
#pragma unroll
    for (int t = 0; t < kQT; t++) {
      double s, c;
      sincos(n[t][0] * x + n[t][1] * y + n[t][2] * z, &s, &c);
      const double2* p = pol + (long long)rhs[t] * 3 * S + j;
#pragma unroll
      for (int a = 0; a < 3; a++) {
        const double2 v = p[(long long)a * S];
        acc[t][2 * a] += v.x * c - v.y * s;
        acc[t][2 * a + 1] += v.x * s + v.y * c;
      }
    }
  }
#pragma unroll
  for (int t = 0; t < kQT; t++)
#pragma unroll
    for (int a = 0; a < 6; a++) {
      const double v = block_sum(acc[t][a]);
      if (threadIdx.x == 0 && q0 + t < nq) part[((long long)blockIdx.x * nq + q0 + t) * 6 + a] = v;
    }
}

// Second stage, one thread per query, chunks summed in order; then
// E = i k^2 exp(-i k origin . n) (S - n (n . S)).
__global__ void k_far_final(const double* __restrict__ part, const double* __restrict__ q, int nq, int nchunk,
                            double k, Frame f, double* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nq) return;
  double s[6] = {0, 0, 0, 0, 0, 0};
  for (int b = 0; b < nchunk; b++)
    for (int a = 0; a < 6; a++) s[a] += part[((long long)b * nq + i) * 6 + a];
  const double n0 = q[4 * i + 1], n1 = q[4 * i + 2], n2 = q[4 * i + 3];
  const double nsr = n0 * s[0] + n1 * s[2] + n2 * s[4], nsi = n0 * s[1] + n1 * s[3] + n2 * s[5];
  double ps, pc;
  sincos(-k * (f.ox * n0 + f.oy * n1 + f.oz * n2), &ps, &pc);
  // f = i k^2 (pc + i ps) = k^2 (-ps + i pc)
  const double fr = -k * k * ps, fi = k * k * pc;
  const double nn[3] = {n0, n1, n2};
  for (int a = 0; a < 3; a++) {
    const double vr = s[2 * a] - nn[a] * nsr, vi = s[2 * a + 1] - nn[a] * nsi;
    out[6 * i + 2 * a] = fr * vr - fi * vi;
    out[6 * i + 2 * a + 1] = fr * vi + fi * vr;
  }
}

Frame frame_of(const dda_op_s* op) {
  const MvDims& d = op->dims;
  Frame f;
  f.d = op->grid_unit;
  f.ox = -(d.Bx - 1) / 2.0 * f.d;
  f.oy = -(d.By - 1) / 2.0 * f.d;
  f.oz = -(d.Bz - 1) / 2.0 * f.d;
  return f;
}

// Device copy of the dipole positions, made on first use.
dda_status positions_on_device(dda_op_s* op) {
  if (op->pos_dev) return DDA_OK;
  const size_t b = op->pos.size() * sizeof(uint32_t);
  DDA_TRY_HIP(hipMalloc(&op->pos_dev, b));
  op->device_bytes += b;
  DDA_TRY_HIP(hipMemcpyAsync(op->pos_dev, op->pos.data(), b, hipMemcpyHostToDevice, op->ctx->stream));
  DDA_TRY_HIP(hipStreamSynchronize(op->ctx->stream));
  return DDA_OK;
}

// Device scratch of at least `bytes`, kept by the operator between calls.
dda_status obs_scratch(dda_op_s* op, size_t bytes, void** out) {
  if (op->obs_bytes < bytes) {
    if (op->obs_scratch) {
      DDA_TRY_HIP(hipStreamSynchronize(op->ctx->stream));
      (void)hipFree(op->obs_scratch);
      op->device_bytes -= op->obs_bytes;
      op->obs_scratch = nullptr;
      op->obs_bytes = 0;
    }
    DDA_TRY_HIP(hipMalloc(&op->obs_scratch, bytes));
    op->obs_bytes = bytes;
    op->device_bytes += bytes;
  }
  *out = op->obs_scratch;
  return DDA_OK;
}

}  // namespace

// Device addresses of `scratch`, one per device, taken once per device with
// that device current (dda_ctx_create calls this after hipSetDevice).
static std::mutex g_noalloc_mu;
static double* g_noalloc_ptr[kMaxNoAllocDevices] = {};

dda_status noalloc_scratch_setup(int device) {
  if (device < 0 || device >= kMaxNoAllocDevices) {
    set_error("noalloc scratch: device index out of range");
    return DDA_INVALID_ARG;
  }
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  if (g_noalloc_ptr[device]) return DDA_OK;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    set_error(std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
    return DDA_HIP_ERROR;
  }
  g_noalloc_ptr[device] = static_cast<double*>(p);
  return DDA_OK;
}

static double* noalloc_scratch(int device) {
  if (device < 0 || device >= kMaxNoAllocDevices) return nullptr;
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  return g_noalloc_ptr[device];
}

}  // namespace dda

using namespace dda;

// Alpha and sqrt(alpha) of material i, axis c, for the refractive index
// m[(i * mstride + c * cstride)] (complex): CM or LDR (with S from prop and
// pol, or the polarisation average).
static dda_status couple_constants_impl(const char* fn, int kind, uint32_t nmat, const double* m, int mstride,
                                        int cstride, double grid_unit, double kd, const double prop[3],
                                        const double pol[3], double* cc, double* cc_sqrt) {
  if (!m || !prop || !pol || !cc || nmat == 0 || kind < 0 || kind > 2 || !(grid_unit > 0) || !(kd > 0)) {
    set_error(std::string(fn) + ": null argument, nmat 0, unknown kind, or grid_unit / kd <= 0");
    return DDA_INVALID_ARG;
  }
  typedef std::complex<double> Cd;
  const double b1 = 1.8915316, b2 = -0.1648469, b3 = 1.7700004;  // FPConst.hpp:48-50 (doi:10.1086/172396)
  const double vol = grid_unit * grid_unit * grid_unit;
  double S = 0;
  if (kind == DDA_POL_LDR)
    for (int a = 0; a < 3; a++) S += prop[a] * prop[a] * pol[a] * pol[a];
  else if (kind == DDA_POL_LDR_AVGPOL) {
    double p4 = 0;
    for (int a = 0; a < 3; a++) p4 += prop[a] * prop[a] * prop[a] * prop[a];
    S = 0.5 * (1 - p4);
  }
  for (uint32_t i = 0; i < nmat; i++)
    for (int c = 0; c < 3; c++) {
      const double* mi = m + 2 * ((size_t)i * mstride + (size_t)c * cstride);
      const Cd mm(mi[0], mi[1]), m2 = mm * mm;
      const Cd cm = 3 * vol / (4 * kPi) * (m2 - 1.0) / (m2 + 2.0);
      Cd a = cm;
      if (kind != DDA_POL_CM) {
        const Cd t1 = (b1 + (b2 + b3 * S) * m2) * kd * kd + Cd(0, 2.0 / 3.0 * kd * kd * kd);
        a = cm / (1.0 - (cm / vol) * t1);
      }
      cc[2 * (3 * i + c)] = a.real();
      cc[2 * (3 * i + c) + 1] = a.imag();
      if (cc_sqrt) {
        const Cd sq = std::sqrt(a);
        cc_sqrt[2 * (3 * i + c)] = sq.real();
        cc_sqrt[2 * (3 * i + c) + 1] = sq.imag();
      }
    }
  return DDA_OK;
}

extern "C" dda_status dda_couple_constants(int kind, uint32_t nmat, const double* m, double grid_unit, double kd,
                                           const double prop[3], const double pol[3], double* cc, double* cc_sqrt) {
  return couple_constants_impl("dda_couple_constants", kind, nmat, m, 1, 0, grid_unit, kd, prop, pol, cc,
                               cc_sqrt);
}

extern "C" dda_status dda_couple_constants_diag(int kind, uint32_t nmat, const double* m, double grid_unit,
                                                double kd, const double prop[3], const double pol[3], double* cc,
                                                double* cc_sqrt) {
  return couple_constants_impl("dda_couple_constants_diag", kind, nmat, m, 3, 1, grid_unit, kd, prop, pol, cc,
                               cc_sqrt);
}

// Observables scale by the operator's wave number: a pack qualifies only when
// its members share one.
static dda_status one_wave_number(const dda_op_s* op, const char* fn) {
  for (double w : op->pack.wave_num)
    if (w != op->wave_num) {
      set_error(std::string(fn) + ": a pack of members with different wave numbers");
      return DDA_NOT_SUPPORTED;
    }
  return DDA_OK;
}

extern "C" dda_status dda_incident_plane(dda_op_t op, int nrhs, const double* prop, const double* pol,
                                         void* einc_dev) {
  if (!op || nrhs < 1 || !prop || !pol || !einc_dev) {
    set_error("dda_incident_plane: null argument or nrhs < 1");
    return DDA_INVALID_ARG;
  }
  if (op->pack.nmember > 1) {
    set_error("dda_incident_plane: a pack (the field of member m is its own operator's, scattered by the pack map)");
    return DDA_NOT_SUPPORTED;
  }
  DDA_TRY(bind_device(op->ctx));
  DDA_TRY(positions_on_device(op));
  void* buf = nullptr;
  DDA_TRY(obs_scratch(op, 6ull * nrhs * sizeof(double), &buf));
  std::vector<double> h(6ull * nrhs);
  for (int r = 0; r < nrhs; r++)
    for (int a = 0; a < 3; a++) {
      h[6 * r + a] = prop[3 * r + a];
      h[6 * r + 3 + a] = pol[3 * r + a];
    }
  hipStream_t st = op->ctx->stream;
  DDA_TRY_HIP(hipMemcpyAsync(buf, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, st));
  const MvDims& d = op->dims;
  hipLaunchKernelGGL(k_incident, dim3(std::min(1024, (d.S + 255) / 256), nrhs), dim3(256), 0, st, (double2*)einc_dev,
                     (const uint32_t*)op->pos_dev, (const double*)buf, d.nv, d.S, op->wave_num, frame_of(op));
  DDA_TRY_HIP(hipGetLastError());
  DDA_TRY_HIP(hipStreamSynchronize(st));
  return DDA_OK;
}

extern "C" dda_status dda_cross_sections(dda_op_t op, const void* pol_dev, const void* einc_dev, int nrhs,
                                         const double* cc, double* cext, double* cabs) {
  if (!op || !pol_dev || nrhs < 1 || (cext && !einc_dev) || (cabs && !cc)) {
    set_error("dda_cross_sections: null argument or nrhs < 1");
    return DDA_INVALID_ARG;
  }
  DDA_TRY(one_wave_number(op, "dda_cross_sections"));
  DDA_TRY(bind_device(op->ctx));
  const MvDims& d = op->dims;
  const int nb = std::max(1, std::min(256, (d.nv + 4095) / 4096));
  const double k = op->wave_num;
  std::vector<double> mdr(3ull * op->nmat * nrhs, 0.0);
  if (cc)
    for (size_t i = 0; i < mdr.size(); i++) {
      const std::complex<double> a(cc[2 * i], cc[2 * i + 1]);
      mdr[i] = -(1.0 / a).imag() - 2.0 / 3.0 * k * k * k;
    }
  void* buf = nullptr;
  const size_t pbytes = 2ull * nb * nrhs * sizeof(double);
  DDA_TRY(obs_scratch(op, pbytes + mdr.size() * sizeof(double), &buf));
  double* part = (double*)buf;
  double* mdr_dev = part + 2ull * nb * nrhs;
  hipStream_t st = op->ctx->stream;
  DDA_TRY_HIP(hipMemcpyAsync(mdr_dev, mdr.data(), mdr.size() * sizeof(double), hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_cross, dim3(nb, nrhs), dim3(kObsThreads), 0, st, (const double2*)pol_dev,
                     (const double2*)(einc_dev ? einc_dev : pol_dev), (const uint8_t*)op->mat_dev, mdr_dev,
                     (const double2*)nullptr, 0.0, d.nv, d.S, (int)op->nmat, part, d.member);
  DDA_TRY_HIP(hipGetLastError());
  std::vector<double> h(2ull * nb * nrhs);
  DDA_TRY_HIP(hipMemcpyAsync(h.data(), part, h.size() * sizeof(double), hipMemcpyDeviceToHost, st));
  DDA_TRY_HIP(hipStreamSynchronize(st));
  for (int r = 0; r < nrhs; r++) {
    double se = 0, sa = 0;
    for (int b = 0; b < nb; b++) {
      se += h[(r * (size_t)nb + b) * 2];
      sa += h[(r * (size_t)nb + b) * 2 + 1];
    }
    if (cext) cext[r] = 4 * kPi * k * se;
    if (cabs) cabs[r] = 4 * kPi * k * sa;
  }
  return DDA_OK;
}

// dda_cross_sections without allocating: alpha, the partial sums and the
// finished sums go through the device's fixed `scratch` array (its address
// from noalloc_scratch_setup), the per-right-hand-side sums are finished on the
// device and copied straight into cext / cabs. No hipMalloc, hipFree or host
// heap allocation happens here. Every k_cross block writes its own partials,
// so nothing needs clearing between calls. Calls that overlap on one GPU share
// the array and race.
extern "C" dda_status dda_cross_sections_noalloc(dda_op_t op, const void* pol_dev, const void* einc_dev, int nrhs,
                                                 const double* cc, double* cext, double* cabs) {
  if (!op || !pol_dev || nrhs < 1 || (cext && !einc_dev) || (cabs && !cc)) {
    set_error("dda_cross_sections_noalloc: null argument or nrhs < 1");
    return DDA_INVALID_ARG;
  }
  DDA_TRY(one_wave_number(op, "dda_cross_sections_noalloc"));
  const MvDims& d = op->dims;
  const int nb = std::max(1, std::min(256, (d.nv + 4095) / 4096));
  const size_t ncc = cc ? 6ull * op->nmat * nrhs : 0;  // alpha, [nrhs][nmat][3] complex
  const size_t npart = 2ull * nb * nrhs, nout = 2ull * nrhs;
  if (ncc + npart + nout > (size_t)DDA_NOALLOC_SCRATCH_DOUBLES) {
    set_error("dda_cross_sections_noalloc: needs " + std::to_string(ncc + npart + nout) +
              " doubles of scratch (6 nmat nrhs + 2 nrhs (blocks + 1)), above the fixed " +
              std::to_string(DDA_NOALLOC_SCRATCH_DOUBLES));
    return DDA_INVALID_ARG;
  }
  DDA_TRY(bind_device(op->ctx));
  double* base = noalloc_scratch(op->ctx->device);
  if (!base) {
    set_error("dda_cross_sections_noalloc: the device scratch of this context's device was not set up");
    return DDA_HIP_ERROR;
  }
  double* ccd = base;
  double* part = base + ncc;
  double* fin = part + npart;
  const double k = op->wave_num;
  hipStream_t st = op->ctx->stream;
  if (cc) DDA_TRY_HIP(hipMemcpyAsync(ccd, cc, ncc * sizeof(double), hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_cross, dim3(nb, nrhs), dim3(kObsThreads), 0, st, (const double2*)pol_dev,
                     (const double2*)(einc_dev ? einc_dev : pol_dev), (const uint8_t*)op->mat_dev,
                     (const double*)nullptr, cc ? (const double2*)ccd : (const double2*)nullptr,
                     2.0 / 3.0 * k * k * k, d.nv, d.S, (int)op->nmat, part, d.member);
  hipLaunchKernelGGL(k_cross_final, dim3((nrhs + 255) / 256), dim3(256), 0, st, part, nb, nrhs, 4 * kPi * k, fin);
  DDA_TRY_HIP(hipGetLastError());
  if (cext) DDA_TRY_HIP(hipMemcpyAsync(cext, fin, nrhs * sizeof(double), hipMemcpyDeviceToHost, st));
  if (cabs) DDA_TRY_HIP(hipMemcpyAsync(cabs, fin + nrhs, nrhs * sizeof(double), hipMemcpyDeviceToHost, st));
  DDA_TRY_HIP(hipStreamSynchronize(st));
  return DDA_OK;
}

extern "C" dda_status dda_far_field(dda_op_t op, const void* pol_dev, int nrhs, const int* rhs, const double* dirs,
                                    int nq, double* out) {
  if (!op || !pol_dev || nrhs < 1 || !dirs || nq < 1 || !out) {
    set_error("dda_far_field: null argument, nrhs < 1 or no directions");
    return DDA_INVALID_ARG;
  }
  DDA_TRY(one_wave_number(op, "dda_far_field"));
  DDA_TRY(bind_device(op->ctx));
  for (int i = 0; i < nq; i++)
    if (rhs && (rhs[i] < 0 || rhs[i] >= nrhs)) {
      set_error("dda_far_field: right-hand side index out of range");
      return DDA_INVALID_ARG;
    }
  DDA_TRY(positions_on_device(op));
  const MvDims& d = op->dims;
  const int chunk = 4096;
  const int nchunk = (d.nv + chunk - 1) / chunk;
  const size_t qbytes = 4ull * nq * sizeof(double), pbytes = 6ull * nchunk * nq * sizeof(double);
  const size_t obytes = 6ull * nq * sizeof(double);
  void* buf = nullptr;
  DDA_TRY(obs_scratch(op, qbytes + pbytes + obytes, &buf));
  double* qd = (double*)buf;
  double* part = qd + 4ull * nq;
  double* od = part + 6ull * nchunk * nq;
  std::vector<double> hq(4ull * nq);
  for (int i = 0; i < nq; i++) {
    hq[4 * i] = rhs ? rhs[i] : 0;
    for (int a = 0; a < 3; a++) hq[4 * i + 1 + a] = dirs[3 * i + a];
  }
  hipStream_t st = op->ctx->stream;
  DDA_TRY_HIP(hipMemcpyAsync(qd, hq.data(), qbytes, hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_far_partial, dim3(nchunk, (nq + kQT - 1) / kQT), dim3(kObsThreads), 0, st,
                     (const double2*)pol_dev, (const uint32_t*)op->pos_dev, qd, nq, d.nv, d.S,
                     op->wave_num * op->grid_unit, chunk, part);
  hipLaunchKernelGGL(k_far_final, dim3((nq + 255) / 256), dim3(256), 0, st, part, qd, nq, nchunk, op->wave_num,
                     frame_of(op), od);
  DDA_TRY_HIP(hipGetLastError());
  DDA_TRY_HIP(hipMemcpyAsync(out, od, obytes, hipMemcpyDeviceToHost, st));
  DDA_TRY_HIP(hipStreamSynchronize(st));
  return DDA_OK;
}

extern "C" dda_status dda_jones(dda_op_t op, const void* pol_dev, int symmetric, const double prop[3],
                                const double pol1[3], const double pol2[3], uint32_t ntheta, double theta0,
                                double dtheta, uint32_t nphi, double phi0, double dphi, double* out) {
  if (!op || !pol_dev || !prop || !pol1 || !pol2 || !out || ntheta == 0 || nphi == 0) {
    set_error("dda_jones: null argument or empty angle grid");
    return DDA_INVALID_ARG;
  }
  DDA_TRY(one_wave_number(op, "dda_jones"));
  const size_t count = (size_t)ntheta * nphi;
  // two field evaluations per angle: (p1; X = pol2, Y = pol1) and
  // (p2; X = pol2, Y = pol1) or, for a symmetric target, (p1; X = pol1, Y = -pol2)
  std::vector<double> dirs(6 * count), pe(12 * count), pa(12 * count);
  std::vector<int> rhs(2 * count);
  for (size_t i = 0; i < count; i++) {
    const double th = theta0 + (double)(i / nphi) * dtheta, ph = phi0 + (double)(i % nphi) * dphi;
    const double st = std::sin(th), ct = std::cos(th), sp = std::sin(ph), cp = std::cos(ph);
    for (int f = 0; f < 2; f++) {
      double X[3], Y[3];
      const bool rot = f == 1 && symmetric;
      for (int a = 0; a < 3; a++) {
        X[a] = rot ? pol1[a] : pol2[a];
        Y[a] = rot ? -pol2[a] : pol1[a];
      }
      const size_t qi = 2 * i + f;
      rhs[qi] = f == 1 && !symmetric ? 1 : 0;
      for (int a = 0; a < 3; a++) {
        dirs[3 * qi + a] = ct * prop[a] + st * (cp * X[a] + sp * Y[a]);
        pe[3 * qi + a] = sp * X[a] - cp * Y[a];
        pa[3 * qi + a] = -st * prop[a] + ct * (cp * X[a] + sp * Y[a]);
      }
    }
  }
  std::vector<double> e(12 * count);
  DDA_TRY(dda_far_field(op, pol_dev, symmetric ? 1 : 2, rhs.data(), dirs.data(), (int)(2 * count), e.data()));
  typedef std::complex<double> Cd;
  for (size_t i = 0; i < count; i++) {
    Cd perp[2], par[2];
    for (int f = 0; f < 2; f++) {
      const size_t qi = 2 * i + f;
      perp[f] = par[f] = 0;
      for (int a = 0; a < 3; a++) {
        const Cd v(e[6 * qi + 2 * a], e[6 * qi + 2 * a + 1]);
        perp[f] += v * pe[3 * qi + a];
        par[f] += v * pa[3 * qi + a];
      }
    }
    const double ph = phi0 + (double)(i % nphi) * dphi, sp = std::sin(ph), cp = std::cos(ph);
    const Cd J[4] = {cp * par[1] + sp * par[0], sp * par[1] - cp * par[0], cp * perp[1] + sp * perp[0],
                     sp * perp[1] - cp * perp[0]};
    for (int j = 0; j < 4; j++) {
      out[8 * i + 2 * j] = J[j].real();
      out[8 * i + 2 * j + 1] = J[j].imag();
    }
  }
  return DDA_OK;
}
