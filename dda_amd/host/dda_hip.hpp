// C++ adapters over the C ABI (include/dda_hip.h) with the shapes of the
// reference's classes, so that the reference's callers (DDA/DDAMain.cpp ddaCl,
// createResOutput) can swap the OpenCL path for this one:
//
//   Hip::GpuMatVec<T>         GpuMatVec<T>             DDA/GpuMatVec.hpp:85-99
//   Hip::MatVecHip<T>         MatVec<T> / MatVecGpu<T> DDA/MatVec.hpp:33-55, DDA/MatVecGpu.cpp:51-68
//   Hip::IterativeSolverBase<T>  IterativeSolverBase<T> DDA/IterativeSolverBase.hpp:59-62
//   Hip::GpuQmrCs<T>, GpuBicgStab<T>, GpuCgnr<T>, GpuBicgCs<T>  DDA/GpuQmrCs.hpp:81 etc.
//
// Status codes become exceptions (Hip::Error), as the reference's ASSERT /
// ABORT_MSG become Core::AssertionFailure (Core/Assert.hpp:42-64). Header-only,
// C++17, no Boost; link with libdda_hip.so.
#pragma once

#include <algorithm>
#include <array>
#include <complex>
#include <cstdint>
#include <memory>
#include <ostream>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "../../include/dda_hip.h"

namespace DDA {
namespace Hip {

class Error : public std::runtime_error {
 public:
  Error(dda_status s, const std::string& what) : std::runtime_error(what), status(s) {}
  dda_status status;
};

inline void check(dda_status s, const char* where) {
  if (s != DDA_OK) throw Error(s, std::string(where) + ": " + dda_last_error());
}

// One GPU and its in-order stream (the reference's cl::CommandQueue list).
class Context {
 public:
  explicit Context(int device = 0) { check(dda_ctx_create(device, &h_), "dda_ctx_create"); }
  ~Context() { dda_ctx_destroy(h_); }
  Context(const Context&) = delete;
  Context& operator=(const Context&) = delete;
  dda_ctx_t handle() const { return h_; }
  void finish() { check(dda_ctx_synchronize(h_), "dda_ctx_synchronize"); }

 private:
  dda_ctx_t h_ = nullptr;
};

// The parts of DDAParams / DipoleGeometry the matvec reads.
struct Problem {
  std::array<uint32_t, 3> box{};
  std::array<uint32_t, 3> grid{};           // 0 = fftFit smooth rule
  std::vector<std::array<uint32_t, 3>> positions;  // DipoleGeometry::positions()
  std::vector<uint8_t> materials;           // material index per dipole
  uint32_t matCount = 1;
  double gridUnit = 0;                      // DDAParams::gridUnit
  double waveNum = 0;                       // DDAParams::waveNum
  uint32_t vecStride() const { return (uint32_t)((positions.size() + 15) / 16 * 16); }
  size_t vecSize() const { return 3 * (size_t)vecStride(); }
};

// Device buffer of a dipole vector (DipVector<T>).
template <typename T>
class DeviceVector {
 public:
  DeviceVector(Context& ctx, size_t n) : ctx_(ctx), n_(n) {
    check(dda_malloc(ctx.handle(), n * sizeof(std::complex<T>), &p_), "dda_malloc");
  }
  ~DeviceVector() { dda_free(ctx_.handle(), p_); }
  DeviceVector(const DeviceVector&) = delete;
  DeviceVector& operator=(const DeviceVector&) = delete;
  void write(const std::vector<std::complex<T>>& h) {
    if (h.size() != n_) throw Error(DDA_INVALID_ARG, "DeviceVector::write: size mismatch");
    check(dda_memcpy_h2d(ctx_.handle(), p_, h.data(), n_ * sizeof(std::complex<T>)), "dda_memcpy_h2d");
  }
  std::vector<std::complex<T>> read() const {
    std::vector<std::complex<T>> h(n_);
    check(dda_memcpy_d2h(ctx_.handle(), h.data(), p_, n_ * sizeof(std::complex<T>)), "dda_memcpy_d2h");
    return h;
  }
  void* data() { return p_; }
  const void* data() const { return p_; }
  size_t size() const { return n_; }

 private:
  Context& ctx_;
  size_t n_;
  void* p_ = nullptr;
};

template <typename T> struct Precision;
template <> struct Precision<double> { static constexpr int value = 64; };
template <> struct Precision<float> { static constexpr int value = 32; };

// GpuMatVec<T>::create + setCoupleConstants + apply.
template <typename T>
class GpuMatVec {
 public:
  // dmatrix: optional host D matrix in the reference layout
  // (SymMatrix3<complex<double>>[Gy][Gz][Gx], Fortran order); null = generate
  // on the device.
  GpuMatVec(Context& ctx, const Problem& p, const std::complex<double>* dmatrix = nullptr) : ctx_(ctx), p_(p) {
    dda_op_desc d{};
    for (int a = 0; a < 3; a++) {
      d.box[a] = p.box[a];
      d.grid[a] = p.grid[a];
    }
    d.nv = (uint32_t)p.positions.size();
    d.vec_stride = p.vecStride();
    d.pos = p.positions.empty() ? nullptr : p.positions[0].data();
    d.mat = p.materials.empty() ? nullptr : p.materials.data();
    d.nmat = p.matCount;
    d.grid_unit = p.gridUnit;
    d.wave_num = p.waveNum;
    d.dmatrix = reinterpret_cast<const double*>(dmatrix);
    check(dda_op_create(ctx.handle(), &d, Precision<T>::value, &h_), "GpuMatVec::create");
    uint32_t g[3], vs;
    size_t bytes;
    check(dda_op_info(h_, g, &vs, &bytes), "dda_op_info");
    grid_ = {g[0], g[1], g[2]};
    deviceBytes_ = bytes;
  }
  ~GpuMatVec() { dda_op_destroy(h_); }
  GpuMatVec(const GpuMatVec&) = delete;
  GpuMatVec& operator=(const GpuMatVec&) = delete;

  // sqrt of the couple constants per material and axis (CoupleConstants::cc_sqrt()).
  void setCoupleConstants(const std::vector<std::array<std::complex<double>, 3>>& ccSqrt) {
    if (ccSqrt.size() != p_.matCount) throw Error(DDA_INVALID_ARG, "setCoupleConstants: one entry per material");
    check(dda_op_set_cc_sqrt(h_, reinterpret_cast<const double*>(ccSqrt.data())), "setCoupleConstants");
    cc_ = ccSqrt;
  }
  // One set per right-hand side of a block matvec / batched solve (no
  // counterpart in the reference, which handles one incident field at a time).
  void setCoupleConstantsBatch(const std::vector<std::vector<std::array<std::complex<double>, 3>>>& ccSqrt) {
    std::vector<std::array<std::complex<double>, 3>> flat;
    for (const auto& c : ccSqrt) {
      if (c.size() != p_.matCount) throw Error(DDA_INVALID_ARG, "setCoupleConstantsBatch: one entry per material");
      flat.insert(flat.end(), c.begin(), c.end());
    }
    check(dda_op_set_cc_sqrt_batch(h_, reinterpret_cast<const double*>(flat.data()), (int)ccSqrt.size()),
          "setCoupleConstantsBatch");
  }
  void apply(const DeviceVector<T>& arg, DeviceVector<T>& result, bool conj, void* stream = nullptr) {
    check(dda_op_apply(h_, arg.data(), result.data(), conj ? 1 : 0, stream), "GpuMatVec::apply");
  }
  // arg/result hold nrhs vectors of vecSize() back to back.
  void applyBatch(const DeviceVector<T>& arg, DeviceVector<T>& result, int nrhs, bool conj, void* stream = nullptr) {
    if (arg.size() < (size_t)nrhs * p_.vecSize() || result.size() < (size_t)nrhs * p_.vecSize())
      throw Error(DDA_INVALID_ARG, "GpuMatVec::applyBatch: vector size");
    check(dda_op_apply_batch(h_, arg.data(), result.data(), nrhs, conj ? 1 : 0, stream), "GpuMatVec::applyBatch");
  }
  const Problem& ddaParams() const { return p_; }
  const std::array<uint32_t, 3>& grid() const { return grid_; }
  size_t deviceBytes() const { return deviceBytes_; }
  dda_op_t handle() const { return h_; }
  Context& context() { return ctx_; }

 private:
  Context& ctx_;
  Problem p_;
  dda_op_t h_ = nullptr;
  std::array<uint32_t, 3> grid_{};
  size_t deviceBytes_ = 0;
  std::vector<std::array<std::complex<double>, 3>> cc_;
};

// MatVec<T> on host vectors: apply(arg, result, conj), as MatVecGpu<T>
// (vectors of std::complex<T>; the C ABI stages them as complex double).
template <typename T>
class MatVecHip {
 public:
  explicit MatVecHip(GpuMatVec<T>& op) : op_(op) {}
  void apply(const std::vector<std::complex<T>>& arg, std::vector<std::complex<T>>& result, bool conj) {
    const size_t n = op_.ddaParams().vecSize();
    if (arg.size() != n) throw Error(DDA_INVALID_ARG, "MatVecHip::apply: argument size");
    result.resize(n);
    if constexpr (std::is_same<T, double>::value) {
      check(dda_op_apply_host(op_.handle(), reinterpret_cast<const double*>(arg.data()),
                              reinterpret_cast<double*>(result.data()), conj ? 1 : 0),
            "MatVecHip::apply");
    } else {
      std::vector<std::complex<double>> a(arg.begin(), arg.end()), r(n);
      check(dda_op_apply_host(op_.handle(), reinterpret_cast<const double*>(a.data()),
                              reinterpret_cast<double*>(r.data()), conj ? 1 : 0),
            "MatVecHip::apply");
      for (size_t i = 0; i < n; i++) result[i] = std::complex<T>(r[i]);
    }
  }
  void setCoupleConstants(const std::vector<std::array<std::complex<double>, 3>>& ccSqrt) {
    op_.setCoupleConstants(ccSqrt);
  }

 private:
  GpuMatVec<T>& op_;
};

// IterativeSolverBase<T>::getPolVec over a GpuMatVec<T>: T = double runs the
// reference's double path (or, with mixed, FP32 Krylov inside FP64
// refinement), T = float its float path (ddaCl<float>, DDA/DDAMain.cpp:723-727).
template <typename T>
class IterativeSolverBase {
 public:
  typedef std::complex<T> ctype;
  virtual ~IterativeSolverBase() { dda_solver_destroy(h_); }
  IterativeSolverBase(const IterativeSolverBase&) = delete;
  IterativeSolverBase& operator=(const IterativeSolverBase&) = delete;

  void setCoupleConstants(const std::vector<std::array<std::complex<double>, 3>>& ccSqrt) {
    op_.setCoupleConstants(ccSqrt);
  }
  // einc: incident field SoA [3][vecStride]; returns p = sqrt(a) x. Stops at
  // |r| < eps |b|; throws Error with DDA_MAX_ITER / DDA_STAGNATION /
  // DDA_BREAKDOWN where the reference aborts. A residual line is written to
  // `log` as the reference does at the end of each solve.
  std::shared_ptr<std::vector<ctype>> getPolVec(const std::vector<ctype>& einc, T eps, std::ostream& log,
                                                const std::vector<ctype>& start = {}) {
    const size_t n = op_.ddaParams().vecSize();
    if (einc.size() != n || (!start.empty() && start.size() != n))
      throw Error(DDA_INVALID_ARG, "getPolVec: vector size");
    const std::vector<std::complex<double>> e = widen(einc), st = widen(start);
    std::vector<std::complex<double>> out(n);
    dda_status s = dda_solver_solve(h_, reinterpret_cast<const double*>(e.data()), eps,
                                    st.empty() ? nullptr : reinterpret_cast<const double*>(st.data()),
                                    reinterpret_cast<double*>(out.data()), &info_);
    log << "iterations = " << info_.iterations << ", |r|/|b| = " << info_.rel_residual << std::endl;
    check(s, "getPolVec");
    return std::make_shared<std::vector<ctype>>(out.begin(), out.end());
  }
  // Batched getPolVec: every field of `eincs` solved in lockstep, each result
  // bitwise the one getPolVec gives it alone. Per-field outcomes are in
  // lastInfos() (status per field); throws only for errors of the call itself.
  std::vector<std::shared_ptr<std::vector<ctype>>> getPolVecs(const std::vector<std::vector<ctype>>& eincs, T eps,
                                                              std::ostream& log) {
    const size_t n = op_.ddaParams().vecSize();
    std::vector<std::complex<double>> flat;
    flat.reserve(n * eincs.size());
    for (const auto& e : eincs) {
      if (e.size() != n) throw Error(DDA_INVALID_ARG, "getPolVecs: vector size");
      flat.insert(flat.end(), e.begin(), e.end());
    }
    std::vector<std::complex<double>> out(flat.size());
    infos_.assign(eincs.size(), dda_solve_info{});
    dda_status s = dda_solver_solve_batch(h_, reinterpret_cast<const double*>(flat.data()), (int)eincs.size(), eps,
                                          nullptr, reinterpret_cast<double*>(out.data()), infos_.data());
    if (s == DDA_HIP_ERROR || s == DDA_OUT_OF_MEMORY || s == DDA_NOT_SUPPORTED) check(s, "getPolVecs");
    std::vector<std::shared_ptr<std::vector<ctype>>> res;
    for (size_t r = 0; r < eincs.size(); r++) {
      res.push_back(std::make_shared<std::vector<ctype>>(out.begin() + r * n, out.begin() + (r + 1) * n));
      log << "rhs " << r << ": iterations = " << infos_[r].iterations << ", |r|/|b| = " << infos_[r].rel_residual
          << ", status = " << infos_[r].status << std::endl;
    }
    if (!infos_.empty()) info_ = infos_[0];
    return res;
  }
  // IterativeSolverBase::profilingRun (DDA/IterativeSolverBase.cpp:105-136):
  // one untimed iteration, five timed ones; prints "[i]time " sorted by time,
  // then "median +- spread", as the reference does.
  void profilingRun(std::ostream& out, std::ostream& log) {
    double us[5], med = 0, spread = 0;
    check(dda_solver_profiling_run(h_, us, &med, &spread), "profilingRun");
    std::array<std::pair<double, int>, 5> r;
    for (int i = 0; i < 5; i++) r[i] = {us[i], i};
    std::sort(r.begin(), r.end());
    for (const auto& e : r) out << "[" << e.second << "]" << e.first << "us ";
    out << std::endl << std::endl << med << "us +- " << spread << "us" << std::endl;
    log << "profiling run: " << med << "us +- " << spread << "us per iteration" << std::endl;
  }
  const dda_solve_info& lastInfo() const { return info_; }
  const std::vector<dda_solve_info>& lastInfos() const { return infos_; }
  std::vector<double> history() const {
    std::vector<double> h(info_.iterations + 1);
    uint64_t w = 0;
    check(dda_solver_history(h_, h.data(), h.size(), &w), "history");
    h.resize(w);
    return h;
  }

 protected:
  IterativeSolverBase(GpuMatVec<T>& op, int kind, uint64_t maxIter, bool mixed) : op_(op) {
    if (mixed && !std::is_same<T, double>::value)
      throw Error(DDA_INVALID_ARG, "mixed precision refines in FP64: it needs a GpuMatVec<double>");
    const int mode = std::is_same<T, float>::value ? DDA_MODE_FP32 : mixed ? DDA_MODE_MIXED : DDA_MODE_FP64;
    check(dda_solver_create_ex(op.handle(), kind, maxIter, 0, mode, &h_), "solver create");
  }
  static std::vector<std::complex<double>> widen(const std::vector<ctype>& v) {
    return std::vector<std::complex<double>>(v.begin(), v.end());
  }
  GpuMatVec<T>& op_;
  dda_solver_t h_ = nullptr;
  dda_solve_info info_{};
  std::vector<dda_solve_info> infos_;
};

// new GpuQmrCs<T>(pool, queues, ddaParams, matVec, maxIter, accounting) etc.
template <typename T>
class GpuQmrCs : public IterativeSolverBase<T> {
 public:
  GpuQmrCs(GpuMatVec<T>& op, uint64_t maxIter, bool mixed = false)
      : IterativeSolverBase<T>(op, DDA_SOLVER_QMR_CS, maxIter, mixed) {}
};
template <typename T>
class GpuBicgStab : public IterativeSolverBase<T> {
 public:
  GpuBicgStab(GpuMatVec<T>& op, uint64_t maxIter, bool mixed = false)
      : IterativeSolverBase<T>(op, DDA_SOLVER_BICGSTAB, maxIter, mixed) {}
};
template <typename T>
class GpuCgnr : public IterativeSolverBase<T> {
 public:
  GpuCgnr(GpuMatVec<T>& op, uint64_t maxIter, bool mixed = false)
      : IterativeSolverBase<T>(op, DDA_SOLVER_CGNR, maxIter, mixed) {}
};
template <typename T>
class GpuBicgCs : public IterativeSolverBase<T> {
 public:
  GpuBicgCs(GpuMatVec<T>& op, uint64_t maxIter, bool mixed = false)
      : IterativeSolverBase<T>(op, DDA_SOLVER_BICG_CS, maxIter, mixed) {}
};

}  // namespace Hip
}  // namespace DDA
