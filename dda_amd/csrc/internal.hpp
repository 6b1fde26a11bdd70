// Host-side structures shared by the translation units of libdda_hip.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/dda_hip.h"
#include "common.hpp"
#include "fft.hpp"

namespace dda {

void set_error(const std::string& msg);

#define DDA_TRY_HIP(expr)                                                                   \
  do {                                                                                      \
    hipError_t e_ = (expr);                                                                 \
    if (e_ != hipSuccess) {                                                                 \
      ::dda::set_error(std::string(#expr) + " failed: " + hipGetErrorString(e_));         \
      return (e_ == hipErrorOutOfMemory) ? DDA_OUT_OF_MEMORY : DDA_HIP_ERROR;             \
    }                                                                                       \
  } while (0)

#define DDA_TRY(expr)                    \
  do {                                   \
    dda_status s_ = (expr);              \
    if (s_ != DDA_OK) return s_;         \
  } while (0)

// Extents the matvec kernels need; passed by value.
struct MvDims {
  int Bx, By, Bz;
  int Gx, Gy, Gz;
  int nv;
  int S;       // vector stride
  int nrows;   // By * Bz
  int nmat;
  int sorted;  // dipoles already in row order (row_dip is the identity)
  // Packs of several particles on one grid (nullptr otherwise): member[rhs]
  // selects the material map (row_meta / mat at member * nv) and the tensor
  // dhat_tab[member] of each right-hand side.
  const int* member;
  const void* const* dhat_tab;
};

// Per-right-hand-side stop flags of a device-side solver: the flag of RHS r is
// done[r * stride]; null = never stop. Kernels whose blockIdx.y runs over
// batches map batch b to RHS b / per.
struct RhsFlags {
  const int* done = nullptr;
  int stride = 0;
  int per = 1;
};

// Build a plan: radix-8 stages first, then 4, 2, 3, 5, 7. Returns false if n
// has a prime factor above 7.
bool make_plan(int n, FftPlan* p);

// Device twiddle table exp(-2 pi i t / n), t < n, computed in long double.
dda_status make_twiddles(int n, int prec, void** out, hipStream_t st);

// The FP32 twin of an FP64 operator: same geometry, tensor converted from the
// FP64 one (or taken from the same cache).
dda_status create_twin32(const struct ::dda_op_s* op, struct ::dda_op_s** out);
// Whether this thread's last tensor-cache lookup found its entry, and the
// seconds this thread has spent in lookups so far (the tensor's generation on
// a miss, the wait for another caller's generation of it on a hit).
bool last_tensor_cache_hit();
double tensor_cache_seconds();

// Sums that pass 5 can form, per work-group, over the vector x it gathers and
// the y it writes (both as stored), for a solver's next scalar step: x.y
// (DOT_U: 2 quantities), |y|^2 (DOT_NORM: 1), |y|^2 and x.conj(y) (DOT_OM: 3).
enum PassDot { DOT_NONE = 0, DOT_U = 1, DOT_NORM = 2, DOT_OM = 3 };
struct PassDotReq {
  double* part = nullptr;  // quantity q of work-group b of right-hand side r at part[(r * 4 + q) * nb + b]
  int op = DOT_NONE;
  int r0 = 0;              // right-hand side of the launch's first block row
};

// Launch helpers implemented in matvec.hip.
// nrhs right-hand sides stored back to back ([nrhs][3][S]); done (nullable)
// points at the first RHS's stop flag, done_stride ints apart. dot_part /
// dot_op (compile-time planned X passes only): pass 5 also writes the partial
// sums dot_op of x and y to dot_part, nb = op->p5_blocks work-groups per
// right-hand side.
template <typename R>
dda_status launch_matvec(struct ::dda_op_s* op, const void* x, void* y, int conj, hipStream_t st,
                         const int* done, int nrhs = 1, int done_stride = 0, double* dot_part = nullptr,
                         int dot_op = 0);

// Takes the device address of the fixed scratch array of the allocation-free
// entry point (observe.hip) for `device`, which must be current; once per
// device, later calls return at once. dda_ctx_create calls it.
constexpr int kMaxNoAllocDevices = 64;
// Blocks of at least this many right-hand sides have a tuning of their own.
constexpr int kBlockTuneRhs = 2;
dda_status noalloc_scratch_setup(int device);

}  // namespace dda

struct dda_ctx_s {
  int device = 0;
  hipStream_t stream = nullptr;
  hipDeviceProp_t prop;
};

namespace dda {
// Makes the context's device current on the calling thread (handles may be
// used from any thread, one at a time).
inline dda_status bind_device(const dda_ctx_s* c) {
  int cur = -1;
  if (hipGetDevice(&cur) != hipSuccess || cur != c->device) {
    hipError_t e = hipSetDevice(c->device);
    if (e != hipSuccess) {
      set_error(std::string("hipSetDevice failed: ") + hipGetErrorString(e));
      return DDA_HIP_ERROR;
    }
  }
  return DDA_OK;
}
}  // namespace dda

struct dda_op_s {
  dda_ctx_s* ctx = nullptr;
  int prec = 64;
  dda::MvDims dims;
  uint32_t nmat = 0;
  double grid_unit = 0, wave_num = 0;
  bool cc_set = false;
  std::vector<uint32_t> pos;  // host copy [nv][3]
  std::vector<uint8_t> mat;   // host copy [nv]
  std::vector<double> cc_sqrt_host;  // [nmat][3] complex, reference units

  // device
  uint32_t* row_start = nullptr;  // [nrows + 1]
  uint32_t* row_dip = nullptr;    // [nv] dipole index, grouped by row
  uint32_t* row_meta = nullptr;   // [nv] x | mat << 16
  uint32_t* row_list = nullptr;   // rows holding at least one dipole
  uint8_t* mat_dev = nullptr;     // [nv] material index per dipole
  int n_nonempty = 0;
  void* ccs = nullptr;            // [nmat][3] complex R, d^(-3/2) scaled
  void* dhat = nullptr;           // octant [6][Gz/2+1][Gy/2+1][Gx/2+1] complex R, d^3 scaled
  void* a1 = nullptr;             // [3][Bz][By][Gx]
  void* a2 = nullptr;             // [3][Bz][Gy][Gx]
  void* tw[3] = {nullptr, nullptr, nullptr};
  dda::FftPlan plan[3];
  void* stage_x = nullptr;        // host-apply staging vectors
  void* stage_y = nullptr;
  size_t device_bytes = 0;

  // per-pass event timing (dda_op_profile)
  bool prof = false;
  struct ProfRec { int pass; hipEvent_t b, e; };
  std::vector<ProfRec> prof_recs;
  std::vector<hipEvent_t> prof_pool;

  // launch shapes
  int p1_rows = 1;   // rows per work-group in the X passes
  int p5_blocks = 0; // work-groups per right-hand side of the last static pass 5 launched
  int col_t = 8;     // columns per work-group in the Y/Z passes
  // Kernel variants (plan width of the static column passes: -1 default,
  // 0 narrow, 1 wide; the Z pass at 1024 points also 2, 3: the 32 x 32 plan
  // over 16 or 8 columns), chosen by dda_op_tune.
  struct Tune {
    int col_wide[2] = {-1, -1};  // forward / inverse Y pass
    int p3_wide = -1;            // Z pass with the tensor product
    int x_wide[2] = {-1, -1};    // X passes (scatter + forward, inverse + gather)
    int split = -1;              // a block as two half-blocks on two streams: -1 default (off), 0 off, 1 on
  } tune;
  bool tuned = false;
  // the tuning for blocks of kBlockTuneRhs or more right-hand sides (used when
  // tuned_blk; else `tune` serves every block size)
  Tune tune_blk;
  bool tuned_blk = false;
  struct dda_dhat_cache_s* dhat_cache = nullptr;  // set when dhat is a cache entry (not owned)
  // A pack of several particles on one grid (dda_op_create_pack): the union
  // of their positions, each member's materials on it (void = nmat - 1,
  // polarisability 0) and each member's tensor (its kd).
  struct Pack {
    int nmember = 0;
    std::vector<uint8_t> mats;                 // [nmember][nv]
    std::vector<double> wave_num;              // [nmember]
    std::vector<std::vector<uint32_t>> index;  // member dipole -> union index
    std::vector<void*> dhat;                   // [nmember], [0] = dhat
    void** dhat_tab = nullptr;                 // device copy of dhat
    int* member_dev = nullptr;                 // [cap_rhs] member of each right-hand side
    std::vector<int> member_host;
  } pack;
  void* pos_dev = nullptr;      // [nv][3] positions, uploaded on first use by the observables
  void* obs_scratch = nullptr;  // observables' work space
  size_t obs_bytes = 0;  // dda_op_tune ran (an FP32 twin of a tuned operator is tuned too)
  bool static_x = false;  // X passes with a compile-time plan (Gx has one)
  int cap_rhs = 1;   // right-hand sides the grid buffers hold
  int cc_rhs = 0;    // right-hand sides with couple constants set
  bool static_z = false;  // pass 3 uses a compile-time planned kernel
  // A block of right-hand sides runs as two halves on two streams (the Z pass
  // of one beside the X and Y passes of the other); the second stream and
  // the fork / join events are made on first use.
  bool split_rhs = false;  // the default of tune.split (DDA_SPLIT=1 turns it on)
  hipStream_t split_stream = nullptr;
  hipEvent_t split_ev[2] = {nullptr, nullptr};
};
