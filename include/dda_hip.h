/*
 * dda_hip.h - C ABI of the MI355X (gfx950) DDA FFT-convolution matvec and the
 * Krylov solvers around it.
 *
 * This boundary replaces, for one GPU, the reference's OpenCL path:
 *   GpuMatVec<T>              DDA/GpuMatVec.hpp:85-99, DDA/GpuMatVec.cpp:88-431
 *   MatVecGpu<T> (host data)  DDA/MatVecGpu.cpp:51-68
 *   MatVec<T>::apply          DDA/MatVec.hpp:54
 *   DMatrixGpu::createDMatrix DDA/DMatrixGpu.cpp:32-41 (here generated on the device)
 *   IterativeSolverBase<T>::getPolVec  DDA/IterativeSolverBase.cpp:40-103
 *   GpuQmrCs / GpuBicgStab / GpuCgnr / GpuBicgCs  DDA/GpuQmrCs.hpp:81,
 *                             DDA/GpuBicgStab.hpp:65, DDA/GpuCgnr.cpp:41, DDA/GpuBicgCs.cpp:41
 *
 * Conventions (as in the reference):
 *   - complex numbers are interleaved (re, im); device vectors are complex
 *     float (precision 32) or complex double (precision 64), host vectors are
 *     always complex double;
 *   - dipole vectors are SoA [3][vec_stride], vec_stride = roundUp(nv, 16)
 *     (DDA/DDAParams.cpp:113), entries [nv, vec_stride) are zero on output;
 *   - the operator is A x = x - sqrt(a) G sqrt(a) x (G(0) = 0), with the FFTs
 *     unnormalised and the -1/N folded into the FFT'd interaction tensor
 *     (DDA/DMatrixCpu.cpp:200); conj != 0 computes conj(A conj(x)) = A^H x;
 *   - a solve takes the incident field and returns the dipole polarisation
 *     p = sqrt(a) x (DDA/GpuIterativeSolver.cpp:136-144).
 *
 * Errors are status codes; dda_last_error() returns a thread-local message
 * for the last failing call on that thread. The library owns all device
 * memory it allocates; the caller owns host arrays, which are copied before a
 * call returns. A handle is used by one host thread at a time; distinct
 * handles may be used concurrently.
 */
#ifndef DDA_HIP_H
#define DDA_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum dda_status {
  DDA_OK = 0,
  DDA_INVALID_ARG = 1,
  DDA_OUT_OF_MEMORY = 2,
  DDA_HIP_ERROR = 3,
  DDA_BREAKDOWN = 4,   /* solver breakdown: the reference's ASSERTs, e.g. DDA/QmrCs.cpp:89-93 */
  DDA_MAX_ITER = 5,    /* "Too many iterations", DDA/IterativeSolverBase.cpp:57-58 */
  DDA_STAGNATION = 6,  /* "Too many iterations w/o increase", DDA/IterativeSolverBase.cpp:59-60 */
  DDA_NOT_SUPPORTED = 7
} dda_status;

typedef struct dda_ctx_s* dda_ctx_t;
typedef struct dda_op_s* dda_op_t;
typedef struct dda_solver_s* dda_solver_t;
typedef void* dda_stream_t; /* a hipStream_t; NULL selects the context's stream */
typedef struct dda_dhat_cache_s* dda_dhat_cache_t;

/* Message of the last failed call on this thread ("" if none). */
const char* dda_last_error(void);

/* Context: one GPU and one in-order stream (DDA/DDAMain.cpp:556-575). */
dda_status dda_ctx_create(int device, dda_ctx_t* out);
dda_status dda_ctx_destroy(dda_ctx_t ctx);
dda_status dda_ctx_synchronize(dda_ctx_t ctx);
/* Device name and HBM capacity, for reports. */
dda_status dda_ctx_info(dda_ctx_t ctx, char* name, size_t name_len, size_t* hbm_bytes, int* cu_count);

/* Plain device memory for callers without an allocator of their own. */
dda_status dda_malloc(dda_ctx_t ctx, size_t bytes, void** out);
dda_status dda_free(dda_ctx_t ctx, void* ptr);
dda_status dda_memcpy_h2d(dda_ctx_t ctx, void* dst, const void* src, size_t bytes);
dda_status dda_memcpy_d2h(dda_ctx_t ctx, void* dst, const void* src, size_t bytes);
dda_status dda_memcpy_d2d(dda_ctx_t ctx, void* dst, const void* src, size_t bytes);

/* FFT grid size for a box extent of x dipoles: fftFit, DDA/DDAParams.cpp:45-90.
 * x is 2*box; override != 0 returns override (0 if it is smaller than x). */
uint32_t dda_fft_fit(uint32_t x, uint32_t override_, int support_non_pot);
/* The smallest length >= x that has compile-time planned kernels (32, 48, 64,
 * 96, 128, 192, 256, 512, 1024), else dda_fft_fit's smooth rule: the faster
 * grid for a box extent of x/2 (any grid of at least twice the box gives the
 * same product, to rounding). The job server uses it for jobs without a grid;
 * the reference's GPU path likewise rounds to powers of two. */
uint32_t dda_fft_fit_static(uint32_t x);

/* Which passes of an operator on `grid` run compile-time planned kernels (no
 * counterpart in the reference, whose FFT plans are chosen at run time,
 * DDA/GpuFFTPlans.cpp:34-61): static_x for the X passes (Gx has a static
 * plan), static_z for the Z pass with the tensor product (Gz has one and every
 * byte offset of that pass fits 32 bits: the tensor octant's components and
 * Gz/8 planes of the grid each below 4 GiB). Other grids use the run-time
 * planned kernels. Host-only. */
dda_status dda_grid_plan(const uint32_t grid[3], int precision, int* static_x, int* static_z);

/* Operator description: the parts of DDAParams/DipoleGeometry the matvec reads
 * (DDA/GpuMatVec.cpp:88-136). */
typedef struct dda_op_desc {
  uint32_t box[3];       /* dipole box (Bx, By, Bz) */
  uint32_t grid[3];      /* FFT grid, even and >= 2*box; 0 = fftFit smooth rule */
  uint32_t nv;           /* non-void dipoles */
  uint32_t vec_stride;   /* 0 = roundUp(nv, 16) */
  const uint32_t* pos;   /* host [nv][3] grid coordinates (DipoleGeometry::positions()) */
  const uint8_t* mat;    /* host [nv] material index (DipoleGeometry::materialIndices) */
  uint32_t nmat;         /* number of materials */
  double grid_unit;      /* d, metres (DDAParams::gridUnit) */
  double wave_num;       /* k, 1/metres (DDAParams::waveNum) */
  /* Optional host D matrix in the reference layout (DDA/DMatrixCpu.cpp:133-205,
   * SymMatrix3<complex<double>> [Gy][Gz][Gx] Fortran order, 1/m^3). NULL =
   * generate it on the device in FP64 (DDA/DMatrixGpu.cpp:32-41 does it on the
   * CPU with FFTW and uploads). */
  const double* dmatrix;
} dda_op_desc;

/* GpuMatVec::create, DDA/GpuMatVec.hpp:85-86. precision is 32 or 64. */
dda_status dda_op_create(dda_ctx_t ctx, const dda_op_desc* desc, int precision, dda_op_t* out);
dda_status dda_op_destroy(dda_op_t op);
/* Device cache of FFT'd interaction tensors (no counterpart in the reference,
 * which builds the tensor once per run on the CPU, DDA/DMatrixGpu.cpp:32-41).
 * The d^3-scaled tensor depends only on the grid and kd = k d
 * (DDA/DMatrixCpu.cpp:49-63), so operators of any geometry on the same grid
 * and kd share one entry. Cached tensors clip the Green's function at G/2 per
 * axis instead of at the box: the product is the same for every dipole inside
 * the box (the embedding never pairs dipoles further apart than the box), to
 * rounding. Entries are built on first use (FP32 entries from a cached FP64
 * one when present) and evicted least-recently-used once unreferenced and
 * over max_bytes. Thread-safe; one cache per device. */
dda_status dda_dhat_cache_create(dda_ctx_t ctx, size_t max_bytes, dda_dhat_cache_t* out);
/* Fails (INVALID_ARG) while operators still reference entries. */
dda_status dda_dhat_cache_destroy(dda_dhat_cache_t cache);
/* counts = {hits, misses, evictions}; bytes and entries held now. */
dda_status dda_dhat_cache_stats(dda_dhat_cache_t cache, uint64_t counts[3], size_t* bytes, int* entries);
/* dda_op_create with the tensor taken from (or added to) `cache`; desc->dmatrix
 * must be NULL. The operator holds a reference until it is destroyed; the FP32
 * twin of a mixed-mode solver over it uses the same cache. */
dda_status dda_op_create_cached(dda_ctx_t ctx, const dda_op_desc* desc, int precision, dda_dhat_cache_t cache,
                                dda_op_t* out);
/* Packs (no counterpart in the reference, which runs one particle per
 * operator): one operator for nmember particles on the same grid and grid unit
 * (each descs[m] with its own box <= the union box, positions, materials and
 * wave number; grid[] equal in all, 0 = fitted to the largest box). It works on
 * the union of the members' positions in (z, y, x) order; its materials are
 * 0 .. max(nmat_m) - 1 plus a last, void material (polarisability forced to 0)
 * that marks the union positions where a member has no dipole. Every
 * right-hand side belongs to one member (dda_op_set_rhs_members, default 0) and
 * sees that member's materials and tensor; a vector that is zero off the
 * member's dipoles stays so, so a block solve over the pack solves each
 * member's own system. cache may be NULL (tensors generated, shared between
 * members of equal wave number). Observables (dda_cross_sections ...) reject
 * packs whose members differ in wave number. */
dda_status dda_op_create_pack(dda_ctx_t ctx, const dda_op_desc* descs, int nmember, int precision,
                              dda_dhat_cache_t cache, dda_op_t* out);
/* Members, union dipole count and union material count (with the void one) of
 * a pack; nmember = 0 for an operator that is not one. */
dda_status dda_op_pack_info(dda_op_t op, int* nmember, uint32_t* nv_union, uint32_t* nmat_union);
/* index[i] = union position of member `member`'s dipole i (descs[member].nv
 * entries): a member's vector entry i lives at union entry index[i]. */
dda_status dda_op_pack_map(dda_op_t op, int member, uint32_t* index);
/* Member of each of the first nrhs right-hand sides of a pack. */
dda_status dda_op_set_rhs_members(dda_op_t op, const int* member, int nrhs);
/* Grid, vector stride and device bytes held by the operator. */
dda_status dda_op_info(dda_op_t op, uint32_t grid[3], uint32_t* vec_stride, size_t* device_bytes);
/* GpuMatVec::setCoupleConstants, DDA/GpuMatVec.cpp:139-142: sqrt of the
 * polarisabilities, [nmat][3] complex double in m^(3/2). */
dda_status dda_op_set_cc_sqrt(dda_op_t op, const double* cc_sqrt);
/* GpuMatVec::apply, DDA/GpuMatVec.cpp:144-431: y = A x (or A^H x) on device
 * vectors of the operator's precision, enqueued on `stream`. x and y must not
 * alias. */
dda_status dda_op_apply(dda_op_t op, const void* x_dev, void* y_dev, int conj, dda_stream_t stream);
/* Block matvec (no counterpart in the reference, which solves one incident
 * field at a time): nrhs right-hand sides that share the geometry and the
 * interaction tensor, each with its own couple constants (LDR makes them
 * depend on the propagation direction and polarisation,
 * DDA/PolarizabilityDescription.cpp:96-104). Vectors are [nrhs][3][vec_stride]
 * back to back. dda_op_set_cc_sqrt_batch takes [nrhs][nmat][3] complex double
 * and sizes the operator's buffers for nrhs; each right-hand side's result is
 * bitwise the one dda_op_apply gives it alone. */
dda_status dda_op_set_cc_sqrt_batch(dda_op_t op, const double* cc_sqrt, int nrhs);
dda_status dda_op_apply_batch(dda_op_t op, const void* x_dev, void* y_dev, int nrhs, int conj, dda_stream_t stream);
/* Autotuner (no counterpart in the reference, whose FFT plans are fixed per
 * length): times each variant of the column passes (plan width: 8 or 16
 * points per thread; in FP32 every pass also with another tile, i.e.
 * another number of columns or lines per work-group, and 1024-point Z columns
 * with the 32 x 32 plan over 16 or 8 columns) on this operator's own buffers for nrhs
 * right-hand sides, each twice (the faster timing counts), and keeps per pass
 * the fastest where it beats the precision's default by more than 2% (the
 * default otherwise). nrhs >= 2 tunes a separate set for blocks of right-hand
 * sides (matvecs of blocks use it; single right-hand sides, the set of
 * nrhs = 1). The choices are cached for the process by (grid, box,
 * precision, block or not), so later operators of that shape start tuned. A
 * solver over a tuned operator tunes the operator its iterations run on (the
 * FP32 twin in the mixed and FP32 modes) for one right-hand side at creation
 * and for blocks at its first block solve. Results differ
 * between variants only by rounding, so with a block tuning a block's results
 * equal its right-hand sides solved alone to rounding rather than bit for bit.
 * `report` (nullable) receives a one-line summary of the timings. */
dda_status dda_op_tune(dda_op_t op, int nrhs, char* report, size_t report_len);
/* The kernel variant of each pass, in the tuner's order: Y forward, Z with
 * the tensor product, Y inverse, X forward, X inverse; -1 = the precision's
 * default, 0 = narrow plan, 1 = wide plan; in FP32 every pass also takes 2
 * and 3 (narrow and wide with the other tile), except the Z pass of a
 * 1024-point grid, where 2 and 3 are the 32 x 32 plan over 16 or 8 columns;
 * the Y passes also 4 and 5 (a third tile, four times the default columns
 * where they fit).
 * Entry 5:
 * a block of right-hand sides as two half-blocks on two streams (-1 default
 * = off, 0 off, 1 on). Getting reads the set used for single right-hand sides;
 * setting pins the variants for every block size (e.g. a stored tuning, or a
 * parity test of one variant); values a pass does not have run as its
 * default. No counterpart in the reference. */
dda_status dda_op_get_variants(dda_op_t op, int variants[6]);
dda_status dda_op_set_variants(dda_op_t op, const int variants[6]);
/* MatVecGpu::apply, DDA/MatVecGpu.cpp:51-68: host complex double in and out;
 * synchronous. */
dda_status dda_op_apply_host(dda_op_t op, const double* x, double* y, int conj);
/* Copies the operator's FFT'd interaction tensor to the host in the reference
 * layout (1/m^3), for parity checks of the device generation. */
dda_status dda_op_get_dmatrix(dda_op_t op, double* out);
/* Kernel timing, in place of the reference's --sync per-stage profiling
 * (DDA/GpuMatVec.cpp:149-164): while enabled, every matvec of this operator
 * records HIP events around each of its five kernels on the stream it runs
 * on. dda_op_profile_read synchronises, returns the summed durations (ms) and
 * launch counts per pass since the last read, and clears the record. Launches
 * that a finished solver turned into no-ops (shorter than 2 us) are left out
 * of both and counted in *skipped. */
dda_status dda_op_profile(dda_op_t op, int enable);
dda_status dda_op_profile_read(dda_op_t op, double ms[5], uint64_t counts[5], uint64_t* skipped);

typedef enum dda_solver_kind {
  DDA_SOLVER_QMR_CS = 0,   /* DDA/GpuQmrCs.cpp */
  DDA_SOLVER_BICGSTAB = 1, /* DDA/GpuBicgStab.cpp */
  DDA_SOLVER_CGNR = 2,     /* DDA/GpuCgnr.cpp */
  DDA_SOLVER_BICG_CS = 3   /* DDA/GpuBicgCs.cpp */
} dda_solver_kind;

typedef struct dda_solve_info {
  uint64_t iterations;   /* Krylov iterations (summed over refinement steps) */
  uint64_t matvecs;      /* operator applications */
  uint64_t refinements;  /* outer FP64 residual evaluations (mixed mode), else 0 */
  double rel_residual;   /* sqrt(residScale * inprodR) at exit, as the reference logs it */
  double true_rel_residual; /* |b - A x| / |b| recomputed in FP64 after the solve */
  double seconds;        /* wall time of the solve (of the whole batch) */
  int32_t status;        /* dda_status of this right-hand side */
  int32_t reserved;
} dda_solve_info;

typedef enum dda_solver_mode {
  DDA_MODE_FP64 = 0,  /* every operation in FP64, as the reference's double path */
  DDA_MODE_MIXED = 1, /* FP32 Krylov solves (FP64 scalars) inside FP64 iterative refinement */
  DDA_MODE_FP32 = 2   /* the reference's float path (ddaCl<float>, DDA/DDAMain.cpp:723-727): FP32
                         operator and vectors, no refinement; scalars and reductions in FP64 */
} dda_solver_mode;

/* new GpuQmrCs(...) etc. (DDA/DDAMain.cpp:647-656) in FP64 mode; op must be an
 * FP64 operator (DDA_MODE_FP32 through dda_solver_create_ex also takes an FP32
 * one, as GpuQmrCs<float> takes GpuMatVec<float>). max_iter = 0 selects 3*Bx*By*Bz (DDA/DDAMain.cpp:640-644);
 * max_res_increase = 0 selects the solver's reference default (QMR 50000,
 * BiCGStab 30000, CGNR 10, BiCG 50000). */
dda_status dda_solver_create(dda_op_t op, int kind, uint64_t max_iter, uint64_t max_res_increase,
                             dda_solver_t* out);
/* As dda_solver_create with a mode. DDA_MODE_MIXED has no counterpart in the
 * reference: it builds an FP32 twin of op (tensor generated in FP64), runs the
 * Krylov iteration in FP32 with FP64 scalars and reductions, and refines in
 * FP64: r = b - A x, solve A d = r / |r| in FP32 to max(floor, eps |b| / 2|r|),
 * x += |r| d, until the FP64 residual meets eps. max_iter bounds the summed
 * inner iterations; the history then holds the FP64 residuals of the outer
 * steps. */
dda_status dda_solver_create_ex(dda_op_t op, int kind, uint64_t max_iter, uint64_t max_res_increase, int mode,
                                dda_solver_t* out);
/* DDA_MODE_FP32 runs the Krylov iteration on the FP32 operator: op itself if it
 * is precision 32, else an FP32 twin built like the mixed mode's; the true
 * residual in dda_solve_info is then recomputed with that FP32 operator. */
/* Tolerance floor of the FP32 inner solves in mixed mode (default 1e-4). */
dda_status dda_solver_set_inner_floor(dda_solver_t s, double floor_);
/* Plateau stop of the FP32 inner solves in mixed mode: an inner solve ends
 * (its correction applied, the outer step restarting it from the FP64
 * residual) once the smallest |r|^2 it has seen did not fall below ratio times
 * its value `window` iterations earlier. Default window 200, ratio 0.81 (|r|
 * down by less than 10%); window 0 turns it off. */
dda_status dda_solver_set_inner_plateau(dda_solver_t s, uint64_t window, double ratio);
/* GPU share of long solves (no counterpart in the reference; for servers that
 * run several solves on one GPU): once a solve (in mixed mode: an inner or a
 * refinement solve) has run after_seconds, the host waits for each chunk of
 * iterations and leaves the stream idle for (1 / share - 1) times as long
 * before the next, so other streams' work runs in between. share 1 (the
 * default) turns it off; results do not depend on it. */
dda_status dda_solver_set_gpu_share(dda_solver_t s, double after_seconds, double share);
/* dda_op_profile_read for the operator the Krylov iteration runs on: the FP32
 * twin in mixed mode (it follows the FP64 operator's dda_op_profile switch),
 * else the FP64 operator itself. */
dda_status dda_solver_profile_read(dda_solver_t s, double ms[5], uint64_t counts[5], uint64_t* skipped);
dda_status dda_solver_destroy(dda_solver_t s);
/* IterativeSolverBase::profilingRun, DDA/IterativeSolverBase.cpp:105-136: with
 * the couple constants already set, one untimed iteration, then five
 * iterations timed one by one (each waited for) on a fixed pseudo-random
 * right-hand side; us[5] receives their wall times in microseconds in the
 * order run, median_us / spread_us (nullable) the median and the larger
 * distance from it to the fastest or slowest, as the reference prints them.
 * The iteration timed is the one the solves run: FP64, or FP32 in the mixed
 * and FP32 modes. */
dda_status dda_solver_profiling_run(dda_solver_t s, double us[5], double* median_us, double* spread_us);
/* IterativeSolverBase::getPolVec, DDA/IterativeSolverBase.cpp:40-103.
 * einc: host [3*vec_stride] complex double; start (nullable) is a previous
 * polarisation (--load-start-dip-pol, x0 = start / sqrt(a)); pol_out receives
 * p = sqrt(a) x. Stops when |r|^2 < eps^2 |b|^2. */
dda_status dda_solver_solve(dda_solver_t s, const double* einc, double eps, const double* start,
                            double* pol_out, dda_solve_info* info);
/* A batch of nrhs incident fields ([nrhs][3][vec_stride] each for einc, start,
 * pol_out), solved in lockstep with one block matvec per iteration and
 * per-right-hand-side convergence masking (no counterpart in the reference,
 * which solves one field at a time). info has nrhs entries; the return value
 * is the first non-OK status among them. Couple constants come from
 * dda_op_set_cc_sqrt_batch, or dda_op_set_cc_sqrt for all alike. A right-hand
 * side's result is bitwise the one it gets when solved alone. */
dda_status dda_solver_solve_batch(dda_solver_t s, const double* einc, int nrhs, double eps, const double* start,
                                  double* pol_out, dda_solve_info* info);
/* dda_solver_solve_batch with device-resident vectors: einc_dev, start_dev
 * (nullable) and pol_dev are FP64 complex [nrhs][3][vec_stride] device
 * buffers; nothing crosses PCIe. */
dda_status dda_solver_solve_batch_dev(dda_solver_t s, const void* einc_dev, int nrhs, double eps, const void* start_dev,
                                      void* pol_dev, dda_solve_info* info);
/* residScale * |r|^2 after each iteration of the last solve (index 0 = start;
 * of right-hand side 0 after a batch; in mixed mode the FP64 |r|^2/|b|^2 after
 * each refinement step).
 * Writes min(n, iterations + 1) values and returns that count in *written. */
dda_status dda_solver_history(dda_solver_t s, double* out, uint64_t n, uint64_t* written);

/* ---------------------------------------------------------------------------
 * Around the solve
 * ------------------------------------------------------------------------- */
typedef enum dda_polarizability {
  DDA_POL_CM = 0,         /* Clausius-Mossotti, DDA/PolarizabilityDescription.cpp:65-77 */
  DDA_POL_LDR = 1,        /* lattice dispersion relation, :85-110 */
  DDA_POL_LDR_AVGPOL = 2  /* LDR with the polarisation-averaged S */
} dda_polarizability;
/* CoupleConstants (DDA/DDAParams.cpp:315-331): alpha and sqrt(alpha) per
 * material and axis for refractive indices m ([nmat] complex), grid unit d,
 * kd, propagation direction and incident polarisation; cc and cc_sqrt (the
 * latter nullable) are [nmat][3] complex (m^3, m^(3/2)). Host-only. */
dda_status dda_couple_constants(int kind, uint32_t nmat, const double* m, double grid_unit, double kd,
                                const double prop[3], const double pol[3], double* cc, double* cc_sqrt);
/* The same for anisotropic materials: m is [nmat][3] complex, the diagonal of
 * each material's refractive-index tensor (the reference's
 * DiagMatrix3<cldouble> materials, PolarizabilityDescription.cpp:69-75, 90-96:
 * every axis uses its own m). */
dda_status dda_couple_constants_diag(int kind, uint32_t nmat, const double* m, double grid_unit, double kd,
                                     const double prop[3], const double pol[3], double* cc, double* cc_sqrt);
/* Plane-wave incident fields on the device, DDA/Beam.cpp:110-118:
 * einc_dev[r][c][i] = exp(i k r_i . prop_r) pol_r[c] with r_i measured from the
 * box centre (DipoleGeometry::moveToCenter); prop, pol host [nrhs][3];
 * einc_dev FP64 complex [nrhs][3][vec_stride]. */
dda_status dda_incident_plane(dda_op_t op, int nrhs, const double* prop, const double* pol, void* einc_dev);
/* Cext = 4 pi k sum Im(p . conj(E)) (DDA/Beam.cpp:58-68) and
 * Cabs = 4 pi k sum (-Im(1/alpha) - 2/3 k^3) |p|^2 (DDA/AbsCross.cpp:27-48)
 * per right-hand side from device vectors (FP64 [nrhs][3][vec_stride]); cc is
 * host alpha [nrhs][nmat][3] complex (needed for cabs). cext / cabs: host
 * [nrhs], either may be NULL. */
dda_status dda_cross_sections(dda_op_t op, const void* pol_dev, const void* einc_dev, int nrhs, const double* cc,
                              double* cext, double* cabs);
/* dda_cross_sections without any allocation. No counterpart in the reference:
 * for callers that cannot allocate device (or host heap) memory — serving
 * loops and harnesses that forbid hipMalloc. Its partial sums live in a fixed
 * device array of DDA_NOALLOC_SCRATCH_DOUBLES doubles per GPU (8 MiB, held
 * until the process ends), which needs 6 nmat nrhs (with cc) + 2 nrhs (blocks
 * + 1) doubles, blocks = clamp(ceil(nv / 4096), 1, 256); larger calls fail
 * with DDA_INVALID_ARG. Same arguments and results as dda_cross_sections (sums
 * in the same order; Cabs's -Im(1/alpha) is formed on the device, so it may
 * differ in the last bits). Calls that overlap on one GPU share the array and
 * race: serialise them. */
#define DDA_NOALLOC_SCRATCH_DOUBLES (1 << 20)
dda_status dda_cross_sections_noalloc(dda_op_t op, const void* pol_dev, const void* einc_dev, int nrhs,
                                      const double* cc, double* cext, double* cabs);
/* Scattered far fields, DDA/CpuFieldCalculator.cpp:48-98: for query q, the
 * field of right-hand side rhs[q] (NULL = all 0) in unit direction dirs[q]
 * (host [nq][3]); out host [nq][3] complex. All queries run in one two-stage
 * device reduction. */
dda_status dda_far_field(dda_op_t op, const void* pol_dev, int nrhs, const int* rhs, const double* dirs, int nq,
                         double* out);
/* Jones matrices on theta0 + i dtheta (i < ntheta) x phi0 + j dphi (j < nphi)
 * from the solutions of both polarisations (pol_dev [2][3][vec_stride]; with
 * symmetric != 0 only polarisation 1, rotated as DDA/FarFieldCalc.cpp:167-168
 * does), FarFieldCalc.cpp:119-174 and EMSim/JonesCalculus.cpp:74-86; out host
 * [ntheta * nphi][2][2] complex, the matrix of getJonesMatrix
 * (JonesCalculus.cpp:74-86; its text output lists s1..s4 = J[1][1], J[0][0],
 * J[0][1], J[1][0], :88-116). */
dda_status dda_jones(dda_op_t op, const void* pol_dev, int symmetric, const double prop[3], const double pol1[3],
                     const double pol2[3], uint32_t ntheta, double theta0, double dtheta, uint32_t nphi, double phi0,
                     double dphi, double* out);

/* ---------------------------------------------------------------------------
 * Job server (no counterpart in the reference, which runs one geometry per
 * process): scattering jobs are queued, admitted while their estimated device
 * footprint fits the HBM budget, and run by `slots` workers, each on its own
 * context and stream, so admitted jobs overlap on the GPU and a freed slot
 * takes the next job at once (continuous batching). Queued jobs of the same
 * particle (geometry, grid and wave number) are packed into one block solve of
 * at most max_pack right-hand sides. Jobs without a grid get
 * dda_fft_fit_static(2 box) per axis. Interaction tensors come from a device
 * cache keyed by (grid, kd) (dda_dhat_cache_*). A job's result does not
 * depend on what else runs: block solves give each right-hand side bitwise
 * its single-solve result, and the cached tensor is the same for every job of
 * that grid and kd.
 * ------------------------------------------------------------------------- */
typedef struct dda_server_s* dda_server_t;
typedef struct dda_server_config {
  int device;          /* GPU */
  int slots;           /* concurrent block solves, 1..16, each on its own stream */
  size_t hbm_budget;   /* bytes for admitted jobs plus the tensor cache (0 = 90% of HBM) */
  size_t cache_bytes;  /* tensor cache limit, part of hbm_budget (0 = a quarter of it) */
  int solver_kind;     /* dda_solver_kind */
  int mode;            /* dda_solver_mode */
  double eps;          /* relative residual */
  uint64_t max_iter;   /* 0 = 3 * box volume */
  int max_pack;        /* right-hand sides per block solve, >= 1 */
  int pack_grids;      /* 1: also pack different particles on one grid and grid unit
                          (dda_op_create_pack; a job's result then equals its solve
                          alone to the solver's tolerance, not bit for bit) */
  double long_solve_s;     /* with long_solve_share in (0, 1): dda_solver_set_gpu_share */
  double long_solve_share; /* of every job's solver (0 or 1 = off) */
} dda_server_config;
typedef struct dda_job_desc {
  dda_op_desc geometry;  /* as for dda_op_create; dmatrix must be NULL */
  int nrhs;              /* incident fields of this job */
  const double* cc_sqrt; /* [nrhs][nmat][3] complex double, one set per field */
  const double* einc;    /* [nrhs][3][vec_stride] complex double */
} dda_job_desc;
typedef struct dda_job_result {
  double t_submit, t_start, t_end; /* seconds since the server was created */
  int32_t status;                  /* first non-OK dda_status of the job's fields */
  int32_t packed_rhs;              /* right-hand sides of the block solve it ran in */
  size_t footprint;                /* bytes admitted for that block solve */
  int32_t cache_hit;               /* tensor found in the cache */
  int32_t iterations;              /* most Krylov iterations among the job's fields */
  double setup_s;                  /* operator, tensor and solver set-up of the block solve */
  double tensor_s;                 /* of setup_s: interaction tensors from the cache (their generation
                                      on a miss, the wait for another slot's generation on a hit) */
} dda_job_result;
dda_status dda_server_create(const dda_server_config* cfg, dda_server_t* out);
/* Copies the job's inputs and returns its id at once. */
dda_status dda_server_submit(dda_server_t s, const dda_job_desc* job, uint64_t* id);
/* Blocks until job `id` is done; fills pol_out ([nrhs][3][vec_stride]), info
 * ([nrhs]) and res (all nullable) and forgets the job. Returns its status.
 * One thread waits per job: a second wait on a job already being waited for
 * returns DDA_INVALID_ARG. dda_server_destroy must not overlap a wait. */
dda_status dda_server_wait(dda_server_t s, uint64_t id, double* pol_out, dda_solve_info* info, dda_job_result* res);
/* counts = {submitted, completed, cache hits, cache misses, cache evictions};
 * peak_bytes = the largest admitted footprint held at once. */
dda_status dda_server_stats(dda_server_t s, uint64_t counts[5], size_t* peak_bytes);
/* Runs the queued jobs to completion, then stops the workers. */
dda_status dda_server_destroy(dda_server_t s);

#ifdef __cplusplus
}
#endif

#endif /* DDA_HIP_H */
