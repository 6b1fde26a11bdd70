// The FFT-convolution matvec y = x - sqrt(a) G sqrt(a) x on one MI355X.
//
// Replaces GpuMatVec::apply (DDA/GpuMatVec.cpp:144-431) and its kernels
// initXMatrix / innerLoop / createResVec (DDA/GpuMatVecInst.cl:31-121), the LDS
// transposes (DDA/GpuTransposePlanInst.cl) and the Apple FFT plans
// (DDA/GpuFFTPlanCl.cpp) with five passes over HBM and no transposes: the
// contiguous x axis is transformed first, after which every later axis is a
// strided column FFT over a tile of adjacent x columns, so each row of the tile
// is one coalesced segment.
//
//   pass 1  X-FFT of each (z, y) row, fused with the scatter of sqrt(a) x
//           (initXMatrix):         x -> A1[3][Bz][By][Gx]
//   pass 2  Y-FFT, y zero-padded:   A1 -> A2[3][Bz][Gy][Gx]
//   pass 3  Z-FFT (z zero-padded), 3x3 complex-symmetric D multiply (innerLoop),
//           inverse Z-FFT cropped to z < Bz, in place on A2
//   pass 4  inverse Y-FFT cropped to y < By: A2 -> A1
//   pass 5  inverse X-FFT fused with the gather y = sqrt(a) X + x (createResVec)
//
// The FFT'd interaction tensor is stored as one octant: G_xx, G_yy, G_zz are
// even in every offset component and G_xy (G_xz, G_yz) odd in x and y (x and z,
// y and z), so D(G - k) = +-D(k) per axis (DDA/DMatrixCpu.cpp:49-63).

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <chrono>
#include <condition_variable>
#include <map>
#include <mutex>
#include <tuple>

#include "fft_static.hpp"
#include "internal.hpp"

namespace dda {

// ------------------------------------------------------------------------------------
// Kernels
// ------------------------------------------------------------------------------------

__device__ inline bool stopped(const RhsFlags& f, int rhs) { return f.done != nullptr && f.done[rhs * f.stride] != 0; }

// Pass 1: rows r0 .. r0+Lr-1 of the (z, y) plane, all three components. Tile
// line l = c * Lr + lr.
template <typename R, int PPT>
__global__ __launch_bounds__(256) void k_pass1(const typename CxT<R>::T* __restrict__ x,
                                               typename CxT<R>::T* __restrict__ a1,
                                               const uint32_t* __restrict__ row_start,
                                               const uint32_t* __restrict__ row_dip,
                                               const uint32_t* __restrict__ row_meta,
                                               const typename CxT<R>::T* __restrict__ ccs, MvDims d, int lr_n,
                                               FftPlan px, FastDiv ldiv, FastDiv ndiv,
                                               const typename CxT<R>::T* __restrict__ tw, int conj, RhsFlags fl) {
  typedef typename CxT<R>::T C;
  const int rhs = blockIdx.y;
  if (stopped(fl, rhs)) return;
  x += (long long)rhs * 3 * d.S;
  a1 += (long long)rhs * 3 * d.nrows * d.Gx;
  ccs += rhs * 3 * d.nmat;
  if (d.member) row_meta += (size_t)d.member[rhs] * d.nv;
  extern __shared__ char smem[];
  C* tile = reinterpret_cast<C*>(smem);
  const int L = 3 * lr_n;
  const int lp = lds_stride(L);
  const int N = d.Gx;
  const int pts = N * L;
  const int r0 = blockIdx.x * lr_n;
  for (int i = threadIdx.x; i < N * lp; i += blockDim.x) tile[i] = czero<C>();
  __syncthreads();
  for (int lr = 0; lr < lr_n; lr++) {
    const int row = r0 + lr;
    if (row >= d.nrows) break;
    const uint32_t s0 = row_start[row];
    const int cnt = (int)(row_start[row + 1] - s0);
    for (int q = threadIdx.x; q < cnt; q += blockDim.x) {
      const uint32_t i = row_dip[s0 + q];
      const uint32_t meta = row_meta[s0 + q];
      const int xp = meta & 0xffff;
      const int m = meta >> 16;
#pragma unroll
      for (int c = 0; c < 3; c++) {
        C v = x[(size_t)c * d.S + i];
        if (conj) v = cconj(v);
        tile[xp * lp + c * lr_n + lr] = cmul(ccs[3 * m + c], v);
      }
    }
  }
  __syncthreads();
  fft_tile<PPT, false>(tile, px, ldiv, lp, tw);
  for (int i = threadIdx.x; i < pts; i += blockDim.x) {
    const int l = ndiv.div(i);
    const int p = i - l * N;
    const int c = l / lr_n;
    const int lr = l - c * lr_n;
    const int row = r0 + lr;
    if (row < d.nrows) a1[((size_t)c * d.nrows + row) * N + p] = tile[p * lp + l];
  }
}

// Pass 5: rows taken from the list of non-empty rows.
template <typename R, int PPT>
__global__ __launch_bounds__(256) void k_pass5(const typename CxT<R>::T* __restrict__ a1,
                                               const typename CxT<R>::T* __restrict__ x,
                                               typename CxT<R>::T* __restrict__ y,
                                               const uint32_t* __restrict__ row_list, int n_list,
                                               const uint32_t* __restrict__ row_start,
                                               const uint32_t* __restrict__ row_dip,
                                               const uint32_t* __restrict__ row_meta,
                                               const typename CxT<R>::T* __restrict__ ccs, MvDims d, int lr_n,
                                               FftPlan px, FastDiv ldiv, FastDiv ndiv,
                                               const typename CxT<R>::T* __restrict__ tw, int conj, RhsFlags fl) {
  typedef typename CxT<R>::T C;
  const int rhs = blockIdx.y;
  if (stopped(fl, rhs)) return;
  a1 += (long long)rhs * 3 * d.nrows * d.Gx;
  x += (long long)rhs * 3 * d.S;
  y += (long long)rhs * 3 * d.S;
  ccs += rhs * 3 * d.nmat;
  if (d.member) row_meta += (size_t)d.member[rhs] * d.nv;
  extern __shared__ char smem[];
  C* tile = reinterpret_cast<C*>(smem);
  const int L = 3 * lr_n;
  const int lp = lds_stride(L);
  const int N = d.Gx;
  const int pts = N * L;
  const int s_base = blockIdx.x * lr_n;
  for (int i = threadIdx.x; i < pts; i += blockDim.x) {
    const int l = ndiv.div(i);
    const int p = i - l * N;
    const int c = l / lr_n;
    const int lr = l - c * lr_n;
    const int slot = s_base + lr;
    C v = czero<C>();
    if (slot < n_list) v = a1[((size_t)c * d.nrows + row_list[slot]) * N + p];
    tile[p * lp + l] = v;
  }
  __syncthreads();
  fft_tile<PPT, true>(tile, px, ldiv, lp, tw);
  for (int lr = 0; lr < lr_n; lr++) {
    const int slot = s_base + lr;
    if (slot >= n_list) break;
    const int row = row_list[slot];
    const uint32_t s0 = row_start[row];
    const int cnt = (int)(row_start[row + 1] - s0);
    for (int q = threadIdx.x; q < cnt; q += blockDim.x) {
      const uint32_t i = row_dip[s0 + q];
      const uint32_t meta = row_meta[s0 + q];
      const int xp = meta & 0xffff;
      const int m = meta >> 16;
#pragma unroll
      for (int c = 0; c < 3; c++) {
        C a = x[(size_t)c * d.S + i];
        if (conj) a = cconj(a);
        C r = cadd(cmul(ccs[3 * m + c], tile[xp * lp + c * lr_n + lr]), a);
        if (conj) r = cconj(r);
        y[(size_t)c * d.S + i] = r;
      }
    }
  }
  if (blockIdx.x == 0) {
    const int pad = d.S - d.nv;
    for (int t = threadIdx.x; t < 3 * pad; t += blockDim.x) {
      const int c = t / pad;
      y[(size_t)c * d.S + d.nv + (t - c * pad)] = czero<C>();
    }
  }
}

// Strided column FFT over T adjacent columns. Column `col` of batch b, point p
// lives at base[b * bstride + p * pstride + col]; inputs beyond in_len are zero,
// only the first out_len points are stored.
struct ColParams {
  int ncols;
  int in_len, out_len;
  long long pstride_in, pstride_out;
  long long bstride_in, bstride_out;
};

template <typename R, int PPT, bool INV>
__global__ __launch_bounds__(256) void k_colfft(const typename CxT<R>::T* in, typename CxT<R>::T* out, ColParams cp,
                                                FftPlan p, FastDiv tdiv, const typename CxT<R>::T* __restrict__ tw,
                                                RhsFlags fl) {
  typedef typename CxT<R>::T C;
  if (stopped(fl, blockIdx.y / fl.per)) return;
  extern __shared__ char smem[];
  C* tile = reinterpret_cast<C*>(smem);
  const int T = (int)tdiv.d;
  const int c0 = blockIdx.x * T;
  const long long b = blockIdx.y;
  const C* src = in + b * cp.bstride_in;
  C* dst = out + b * cp.bstride_out;
  const int lp = lds_stride(T);
  const int pts = p.n * T;
  for (int i = threadIdx.x; i < pts; i += blockDim.x) {
    const int pp = tdiv.div(i);
    const int l = i - pp * T;
    const int col = c0 + l;
    C v = czero<C>();
    if (pp < cp.in_len && col < cp.ncols) v = src[pp * cp.pstride_in + col];
    tile[pp * lp + l] = v;
  }
  __syncthreads();
  fft_tile<PPT, INV>(tile, p, tdiv, lp, tw);
  const int opts = cp.out_len * T;
  for (int i = threadIdx.x; i < opts; i += blockDim.x) {
    const int pp = tdiv.div(i);
    const int l = i - pp * T;
    const int col = c0 + l;
    if (col < cp.ncols) dst[pp * cp.pstride_out + col] = tile[pp * lp + l];
  }
}

// Pass 3 over the (ky, kx tile) pencil: for each component load z < Bz, FFT,
// keep the spectrum in registers; multiply by D(kx, ky, kz); inverse FFT each
// component and store z < Bz back in place.
//
// Work-groups are mapped XCD-aware: the four pencils that read the same octant
// of the tensor, (kx tile, ky), its x mirror, its y mirror and both, are
// consecutive work items, and consecutive work items go to one XCD (work-group
// b runs on XCD b mod 8), so three of the four read the tensor from that XCD's
// L2. Mirror duplicates (ky = 0, Gy/2, self-mirrored tiles) return at once.
template <typename R, int PPT>
__global__ __launch_bounds__(256) void k_pass3(typename CxT<R>::T* a2, const typename CxT<R>::T* __restrict__ dhat,
                                               MvDims d, FftPlan pz, FastDiv tdiv,
                                               const typename CxT<R>::T* __restrict__ tw, int nwork, int nper,
                                               int nrhs, RhsFlags fl) {
  typedef typename CxT<R>::T C;
  const int T = (int)tdiv.d;
  const int ntx = (d.Gx + T - 1) / T;
  const int ftn = (ntx + 1) / 2;
  const int hx = d.Gx / 2, hy = d.Gy / 2, hz = d.Gz / 2;
  const int w = (blockIdx.x & 7) * nper + (blockIdx.x >> 3);
  if (w >= nwork) return;
  const int sub = w & 3, q = w >> 2;
  const int rhs = q % nrhs, grp = q / nrhs;
  if (stopped(fl, rhs)) return;
  a2 += (long long)rhs * 3 * d.Bz * d.Gy * d.Gx;
  if (d.dhat_tab) dhat = (const typename CxT<R>::T*)d.dhat_tab[d.member[rhs]];
  const int ft = grp % ftn, fyp = grp / ftn;
  int ky = fyp, tile_x = ft;
  if (sub & 2) {
    if (fyp == 0 || fyp == hy) return;
    ky = d.Gy - fyp;
  }
  if (sub & 1) {
    tile_x = ntx - 1 - ft;
    if (tile_x == ft) return;
  }
  extern __shared__ char smem[];
  C* tile = reinterpret_cast<C*>(smem);
  const int c0 = tile_x * T;
  const int lp = lds_stride(T);
  const int N = d.Gz;
  const int pts = N * T;
  const long long cstride = (long long)d.Bz * d.Gy * d.Gx;
  const long long zstride = (long long)d.Gy * d.Gx;
  const long long oct = (long long)(hx + 1) * (hy + 1) * (hz + 1);
  C* base = a2 + (long long)ky * d.Gx;
  C reg[3][PPT];
#pragma unroll
  for (int c = 0; c < 3; c++) {
    for (int i = threadIdx.x; i < pts; i += blockDim.x) {
      const int z = tdiv.div(i);
      const int l = i - z * T;
      const int col = c0 + l;
      C v = czero<C>();
      if (z < d.Bz && col < d.Gx) v = base[c * cstride + z * zstride + col];
      tile[z * lp + l] = v;
    }
    __syncthreads();
    fft_tile<PPT, false>(tile, pz, tdiv, lp, tw);
#pragma unroll
    for (int k = 0; k < PPT; k++) {
      const int i = threadIdx.x + k * blockDim.x;
      if (i < pts) {
        const int z = tdiv.div(i);
        reg[c][k] = tile[z * lp + i - z * T];
      }
    }
    __syncthreads();
  }
  const int fy = ky <= hy ? ky : d.Gy - ky;
  const R sy = ky <= hy ? (R)1 : (R)-1;
#pragma unroll
  for (int k = 0; k < PPT; k++) {
    const int i = threadIdx.x + k * blockDim.x;
    if (i < pts) {
      const int kz = tdiv.div(i);
      const int kx = c0 + i - kz * T;
      if (kx < d.Gx) {
        const int fx = kx <= hx ? kx : d.Gx - kx;
        const int fz = kz <= hz ? kz : d.Gz - kz;
        const R sx = kx <= hx ? (R)1 : (R)-1;
        const R sz = kz <= hz ? (R)1 : (R)-1;
        const C* D = dhat + ((long long)fz * (hy + 1) + fy) * (hx + 1) + fx;
        const C d0 = D[0], d3 = D[3 * oct], d5 = D[5 * oct];
        const C d1 = cscale(D[oct], sx * sy), d2 = cscale(D[2 * oct], sx * sz), d4 = cscale(D[4 * oct], sy * sz);
        const C x0 = reg[0][k], x1 = reg[1][k], x2 = reg[2][k];
        reg[0][k] = cdot3(d0, x0, d1, x1, d2, x2);
        reg[1][k] = cdot3(d1, x0, d3, x1, d4, x2);
        reg[2][k] = cdot3(d2, x0, d4, x1, d5, x2);
      }
    }
  }
#pragma unroll
  for (int c = 0; c < 3; c++) {
#pragma unroll
    for (int k = 0; k < PPT; k++) {
      const int i = threadIdx.x + k * blockDim.x;
      if (i < pts) {
        const int z = tdiv.div(i);
        tile[z * lp + i - z * T] = reg[c][k];
      }
    }
    __syncthreads();
    fft_tile<PPT, true>(tile, pz, tdiv, lp, tw);
    const int opts = d.Bz * T;
    for (int i = threadIdx.x; i < opts; i += blockDim.x) {
      const int z = tdiv.div(i);
      const int l = i - z * T;
      const int col = c0 + l;
      if (col < d.Gx) base[c * cstride + z * zstride + col] = tile[z * lp + l];
    }
    __syncthreads();
  }
}


// ------------------------------------------------------------------------------------
// Compile-time planned variants of the column passes (fft_static.hpp). Thread
// (l, t) = (tid mod T, tid / T) owns column c0 + l and the plan's points t + ...
// ------------------------------------------------------------------------------------

// Buffer resources (V#) for the Z pass: raw (stride 0) buffers whose range
// check turns loads beyond `bytes` into zeros and drops such stores, so that
// the zero padding and the cropping need neither branches nor 64-bit address
// arithmetic per access (the offset is a 32-bit byte offset from a uniform
// base). Word 3 selects 32-bit data format, as gfx9-family buffers require.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t zrsrc(const void* p, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ float2 zload(__amdgpu_buffer_rsrc_t r, uint32_t off, float2*) {
  return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
}
__device__ __forceinline__ double2 zload(__amdgpu_buffer_rsrc_t r, uint32_t off, double2*) {
  return __builtin_bit_cast(double2, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0));
}
typedef int zi2 __attribute__((ext_vector_type(2)));
typedef int zi4 __attribute__((ext_vector_type(4)));
constexpr int kStoreNT = 2;  // buffer cache policy: nt
// The Z pass's stores and those of the X-forward and Y-inverse passes are
// non-temporal (nt): their outputs are read by the next pass only after this
// kernel's end, and streaming them leaves little dirty data in the L2s to be
// written back when the kernel ends (the gap before the next kernel).
// Measured in a solve at 256^3: 591 -> 560 us per 2-RHS matvec; the Y-forward
// pass keeps ordinary stores (non-temporal ones made it 5% slower).
__device__ __forceinline__ void zstore(__amdgpu_buffer_rsrc_t r, uint32_t off, float2 v) {
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(zi2, v), r, off, 0, kStoreNT);
}
__device__ __forceinline__ void zstore(__amdgpu_buffer_rsrc_t r, uint32_t off, double2 v) {
  __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(zi4, v), r, off, 0, kStoreNT);
}
template <bool NT> __device__ __forceinline__ void cstore(float2* p, float2 v) {
  if constexpr (NT) {
    typedef float f2 __attribute__((ext_vector_type(2)));
    __builtin_nontemporal_store(__builtin_bit_cast(f2, v), reinterpret_cast<f2*>(p));
  } else {
    *p = v;
  }
}
template <bool NT> __device__ __forceinline__ void cstore(double2* p, double2 v) {
  if constexpr (NT) {
    typedef double d2 __attribute__((ext_vector_type(2)));
    __builtin_nontemporal_store(__builtin_bit_cast(d2, v), reinterpret_cast<d2*>(p));
  } else {
    *p = v;
  }
}

// Waves per SIMD the Z pass is compiled for, a register budget: FP32
// 16-point plans at 3 (168 VGPRs: the tensor loads of several points stay in
// flight), other FP32 plans at 4, FP64 left to the compiler. At 1024 points
// the 16-point plan runs 1024-thread work-groups, which fit a CU only at 4
// waves per SIMD (128 VGPRs, a few spilled).
// 32-point plans (FP32) hold 192 VGPRs of data: 2 waves per SIMD (256).
template <typename R, int N, int PPT> struct P3Waves {
  static constexpr int value = sizeof(R) == 4 ? (PPT >= 32 ? 2 : PPT >= 16 && N < 1024 ? 3 : 4) : 1;
};

// Pass 3 with a static plan of length N = Gz (see k_pass3 for the mapping).
// Thread (l, t) holds, of every component, the points z = t + K_e of the plan's
// first stage (K_e a compile-time constant per register e), so one lane offset
// t * Gy * Gx + l addresses all of them from per-(component, e) uniform bases;
// each such buffer's range ends at z = Bz (points at and beyond it read as zero
// and are not stored) and spans at most N / PPT planes (it stays below 4 GiB up
// to 1024^3). The tensor is read through one buffer per component with a lane
// offset per point.
// PL is the plan (StaticPlan<N, W> or a PlanOf variant).
template <typename R, int N, int T, typename PL>
__global__ __launch_bounds__(T * (N / PL::PPT))
__attribute__((amdgpu_waves_per_eu(P3Waves<R, N, PL::PPT>::value))) void k_pass3s(typename CxT<R>::T* a2,
                                                                              const typename CxT<R>::T* __restrict__ dhat,
                                                MvDims d, const typename CxT<R>::T* __restrict__ tw, int nwork,
                                                int nper, int nrhs, RhsFlags fl) {
  typedef typename CxT<R>::T C;
  typedef PL P;
  constexpr int PPT = P::PPT;
  constexpr int TPC = N / PPT;
  typedef typename P::Fwd F;
  typedef typename Reverse<F>::type Inv;
  const int ntx = (d.Gx + T - 1) / T;
  const int ftn = (ntx + 1) / 2;
  const int hx = d.Gx / 2, hy = d.Gy / 2, hz = N / 2;
  const int w = (blockIdx.x & 7) * nper + (blockIdx.x >> 3);
  if (w >= nwork) return;
  const int sub = w & 3, q = w >> 2;
  const int rhs = q % nrhs, grp = q / nrhs;
  if (stopped(fl, rhs)) return;
  a2 += (long long)rhs * 3 * d.Bz * d.Gy * d.Gx;
  if (d.dhat_tab) dhat = (const typename CxT<R>::T*)d.dhat_tab[d.member[rhs]];
  const int ft = grp % ftn, fyp = grp / ftn;
  int ky = fyp, tile_x = ft;
  if (sub & 2) {
    if (fyp == 0 || fyp == hy) return;
    ky = d.Gy - fyp;
  }
  if (sub & 1) {
    tile_x = ntx - 1 - ft;
    if (tile_x == ft) return;
  }
  extern __shared__ char smem[];
  C* tile = reinterpret_cast<C*>(smem);
  constexpr int lp = (T & 1) ? T : T + 1;
  const int l = threadIdx.x % T, t = threadIdx.x / T;
  const int kx = tile_x * T + l;
  const bool colok = kx < d.Gx;
  const long long cstride = (long long)d.Bz * d.Gy * d.Gx;
  const uint32_t zstride = (uint32_t)d.Gy * d.Gx;
  C* base = a2 + (long long)ky * d.Gx + (long long)tile_x * T;
  // lanes of columns beyond Gx address past every range
  const uint32_t voff = colok ? ((uint32_t)t * zstride + l) * (uint32_t)sizeof(C) : 0xfffffff0u;
  auto range = [&](const C* b, int c, int e) {
    const int z0 = first_point<N, PPT, F>(0, e);
    const int nz = d.Bz - z0 < TPC ? d.Bz - z0 : TPC;
    return zrsrc(b + c * cstride + (long long)z0 * zstride, nz > 0 ? (uint32_t)nz * zstride * sizeof(C) : 0u);
  };
  C reg[3][PPT];
#pragma unroll
  for (int c = 0; c < 3; c++)
#pragma unroll
    for (int e = 0; e < PPT; e++)
      reg[c][e] = upper_half<F>(e) ? czero<C>() : zload(range(base, c, e), voff, (C*)nullptr);
  // Bz <= N/2 (the grid holds at least twice the box): the forward transforms
  // skip the zero upper half and the inverse ones never form it; two-stage
  // plans form the second-stage twiddle powers once for all six transforms
  // (FP64 and 32-point plans: no registers to spare)
  constexpr bool PRE = TwoStage<N, PPT, F>::value && sizeof(R) == 4 && PPT <= 16;
  C wp[PRE ? PPT - 1 : 1];
  if constexpr (PRE) twiddle_powers<PPT>(wp, t, tw);
  const C* twp = PRE ? wp : tw;
#pragma unroll
  for (int c = 0; c < 3; c++) RunStages<N, PPT, false, 1, C, F, true, PRE>::run(reg[c], t, tile, lp, l, twp);
  const uint32_t plane = (uint32_t)(hy + 1) * (hx + 1);
  const uint32_t oct = plane * (uint32_t)(hz + 1);
  const int fy = ky <= hy ? ky : d.Gy - ky;
  const R sy = ky <= hy ? (R)1 : (R)-1;
  const int fx = kx <= hx ? kx : d.Gx - kx;
  const R sx = kx <= hx ? (R)1 : (R)-1;
  const uint32_t fyx = colok ? (uint32_t)fy * (hx + 1) + fx : 0u;
  const R sxy = sx * sy;
  // (issuing point e + 1's tensor loads before point e is formed was measured
  // slower: 275 -> 310 us per 2-RHS FP32 launch at 256^3, 300 -> 347 us FP64)
#pragma unroll
  for (int e = 0; e < PPT; e++) {
    const int kz = first_point<N, PPT, Inv>(t, e);
    const int fz = kz <= hz ? kz : N - kz;
    const R sz = kz <= hz ? (R)1 : (R)-1;
    const uint32_t o = ((uint32_t)fz * plane + fyx) * (uint32_t)sizeof(C);
    C dc[6];
#pragma unroll
    for (int c = 0; c < 6; c++) dc[c] = zload(zrsrc(dhat + (size_t)c * oct, oct * sizeof(C)), o, (C*)nullptr);
    const C d0 = dc[0], d1 = cscale(dc[1], sxy), d2 = cscale(dc[2], sx * sz), d3 = dc[3], d4 = cscale(dc[4], sy * sz),
            d5 = dc[5];
    const C x0 = reg[0][e], x1 = reg[1][e], x2 = reg[2][e];
    reg[0][e] = cdot3(d0, x0, d1, x1, d2, x2);
    reg[1][e] = cdot3(d1, x0, d3, x1, d4, x2);
    reg[2][e] = cdot3(d2, x0, d4, x1, d5, x2);
  }
  // the store resources are formed anew (an opaque copy of the base keeps the
  // compiler from holding the load ones in scalar registers across the
  // transforms)
  C* sbase = base;
  asm volatile("" : "+s"(sbase));
#pragma unroll
  for (int c = 0; c < 3; c++) {
    RunStages<N, PPT, true, 1, C, Inv, false, PRE>::run(reg[c], t, tile, lp, l, twp);
#pragma unroll
    for (int e = 0; e < PPT; e++)
      if (!upper_half<F>(e)) zstore(range(sbase, c, e), voff, reg[c][e]);
  }
}

// Pass 1 with a static plan of length N = Gx. A work-group takes LN consecutive
// lines g = c * nrows + row of the A1 layout (contiguous in A1); thread (l, t),
// t = tid mod TPC along the line, scatters line l's dipoles into the LDS tile,
// then runs the transform from its first-stage points (only x < N/2 can hold
// dipoles: the upper half is never read) and stores its outputs, TPC
// consecutive points per line, coalesced.
template <typename R, int N, bool W, int BS = 256>
__global__ __launch_bounds__(BS) void k_pass1s(const typename CxT<R>::T* __restrict__ x,
                                                typename CxT<R>::T* __restrict__ a1,
                                                const uint32_t* __restrict__ row_start,
                                                const uint32_t* __restrict__ row_dip,
                                                const uint32_t* __restrict__ row_meta,
                                                const typename CxT<R>::T* __restrict__ ccs, MvDims d,
                                                const typename CxT<R>::T* __restrict__ tw, int conj, RhsFlags fl) {
  typedef typename CxT<R>::T C;
  typedef StaticPlan<N, W> P;
  constexpr int PPT = P::PPT;
  constexpr int TPC = N / PPT;
  constexpr int LN = BS / TPC;
  constexpr int lp = (LN & 1) ? LN : LN + 1;
  typedef typename P::Fwd F;
  typedef typename Reverse<F>::type Out;
  const int rhs = blockIdx.y;
  if (stopped(fl, rhs)) return;
  x += (long long)rhs * 3 * d.S;
  a1 += (long long)rhs * 3 * d.nrows * d.Gx;
  ccs += rhs * 3 * d.nmat;
  if (d.member) row_meta += (size_t)d.member[rhs] * d.nv;
  extern __shared__ char smem[];
  C* tile = reinterpret_cast<C*>(smem);
  const int t = threadIdx.x % TPC, l = threadIdx.x / TPC;
  const int nlines = 3 * d.nrows;
  const int g = blockIdx.x * LN + l;
  for (int i = threadIdx.x; i < (N / 2) * lp; i += blockDim.x) tile[i] = czero<C>();
  __syncthreads();
  if (g < nlines) {
    const int c = g / d.nrows, row = g - c * d.nrows;
    const uint32_t s0 = row_start[row], s1 = row_start[row + 1];
    for (uint32_t q = s0 + t; q < s1; q += TPC) {
      const uint32_t i = d.sorted ? q : row_dip[q], meta = row_meta[q];
      C v = x[(size_t)c * d.S + i];
      if (conj) v = cconj(v);
      tile[(meta & 0xffff) * lp + l] = cmul(ccs[3 * (meta >> 16) + c], v);
    }
  }
  __syncthreads();
  C v[PPT];
#pragma unroll
  for (int e = 0; e < PPT; e++) v[e] = upper_half<F>(e) ? czero<C>() : tile[first_point<N, PPT, F>(t, e) * lp + l];
  __syncthreads();
  RunStages<N, PPT, false, 1, C, F, true>::run(v, t, tile, lp, l, tw);
  if (g < nlines) {
    C* dst = a1 + (size_t)g * N;
#pragma unroll
    for (int e = 0; e < PPT; e++) cstore<true>(dst + first_point<N, PPT, Out>(t, e), v[e]);
  }
}

// The sums a pass 5 work-group forms over the entries it writes, in double,
// with x and y as stored: x.y (DOT_U: q0, q1), |y|^2 (DOT_NORM: q0) or |y|^2
// and x.conj(y) (DOT_OM: q0, q1, q2). Block sums of them, in a fixed order,
// are stored as a solver's partial sums:
// quantity q of work-group b of the launch's right-hand side r at
// dot[(r * 4 + q) * gridDim.x + b]. Every thread of the block calls it.
template <int BS>
__device__ inline void dot_partials(double* dot, int op, double q0, double q1, double q2) {
  __shared__ double red[3][BS / 64];
  for (int off = 32; off > 0; off >>= 1) {
    q0 += __shfl_xor(q0, off, 64);
    q1 += __shfl_xor(q1, off, 64);
    q2 += __shfl_xor(q2, off, 64);
  }
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    red[0][w] = q0;
    red[1][w] = q1;
    red[2][w] = q2;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const int nq = op == DOT_U ? 2 : op == DOT_NORM ? 1 : 3;
    double* p = dot + (long long)blockIdx.y * 4 * gridDim.x + blockIdx.x;
    for (int q = 0; q < nq; q++) {
      double t = 0;
      for (int i = 0; i < BS / 64; i++) t += red[q][i];
      p[(long long)q * gridDim.x] = t;
    }
  }
}

// Pass 5 with a static plan: lines g = c * n_list + slot over the non-empty
// rows; inverse transform from A1 (loads coalesced), the kept half (x < N/2)
// through the tile, then line l's dipoles gathered by its TPC threads.
// dot_op != DOT_NONE: also the partial sums of x and y (dot_partials) into dot.
template <typename R, int N, bool W, int BS = 256>
__global__ __launch_bounds__(BS) void k_pass5s(const typename CxT<R>::T* __restrict__ a1,
                                                const typename CxT<R>::T* __restrict__ x,
                                                typename CxT<R>::T* __restrict__ y,
                                                const uint32_t* __restrict__ row_list, int n_list,
                                                const uint32_t* __restrict__ row_start,
                                                const uint32_t* __restrict__ row_dip,
                                                const uint32_t* __restrict__ row_meta,
                                                const typename CxT<R>::T* __restrict__ ccs, MvDims d,
                                                const typename CxT<R>::T* __restrict__ tw, int conj, RhsFlags fl,
                                                double* __restrict__ dot, int dot_op) {
  typedef typename CxT<R>::T C;
  typedef StaticPlan<N, W> P;
  constexpr int PPT = P::PPT;
  constexpr int TPC = N / PPT;
  constexpr int LN = BS / TPC;
  constexpr int lp = (LN & 1) ? LN : LN + 1;
  typedef typename P::Fwd F;
  typedef typename Reverse<F>::type Out;
  const int rhs = blockIdx.y;
  if (stopped(fl, rhs)) return;
  a1 += (long long)rhs * 3 * d.nrows * d.Gx;
  x += (long long)rhs * 3 * d.S;
  y += (long long)rhs * 3 * d.S;
  ccs += rhs * 3 * d.nmat;
  if (d.member) row_meta += (size_t)d.member[rhs] * d.nv;
  extern __shared__ char smem[];
  C* tile = reinterpret_cast<C*>(smem);
  const int t = threadIdx.x % TPC, l = threadIdx.x / TPC;
  const int g = blockIdx.x * LN + l;
  const bool ok = g < 3 * n_list;
  const int c = ok ? g / n_list : 0;
  const int row = ok ? (int)row_list[g - c * n_list] : 0;
  const C* src = a1 + ((size_t)c * d.nrows + row) * N;
  C v[PPT];
#pragma unroll
  for (int e = 0; e < PPT; e++) v[e] = ok ? src[first_point<N, PPT, F>(t, e)] : czero<C>();
  // the dipoles this thread gathers (a row holds at most N/2: the grid spans
  // at least twice the box), their x entries and positions loaded before the
  // transform so that those loads wait while it runs
  constexpr int MAXD = (N / 2 + TPC - 1) / TPC;
  const uint32_t s0 = ok ? row_start[row] : 0u, s1 = ok ? row_start[row + 1] : 0u;
  uint32_t gi[MAXD], gm[MAXD];
  C ga[MAXD];
#pragma unroll
  for (int k = 0; k < MAXD; k++) {
    const uint32_t q = s0 + t + k * TPC;
    gi[k] = q < s1 ? (d.sorted ? q : row_dip[q]) : 0u;
    gm[k] = q < s1 ? row_meta[q] : 0u;
    ga[k] = q < s1 ? x[(size_t)c * d.S + gi[k]] : czero<C>();
  }
  RunStages<N, PPT, true, 1, C, F>::run(v, t, tile, lp, l, tw);
  __syncthreads();
#pragma unroll
  for (int e = 0; e < PPT; e++)
    if (!upper_half<Out>(e)) tile[first_point<N, PPT, Out>(t, e) * lp + l] = v[e];
  __syncthreads();
  double q0 = 0, q1 = 0, q2 = 0;
#pragma unroll
  for (int k = 0; k < MAXD; k++) {
    if (s0 + t + k * TPC < s1) {
      C a = ga[k];
      if (conj) a = cconj(a);
      C r = cadd(cmul(ccs[3 * (gm[k] >> 16) + c], tile[(gm[k] & 0xffff) * lp + l]), a);
      if (conj) r = cconj(r);
      y[(size_t)c * d.S + gi[k]] = r;
      if (dot_op != DOT_NONE) {
        const double xr = ga[k].x, xi = ga[k].y, yr = r.x, yi = r.y;
        const bool du = dot_op == DOT_U;
        q0 += du ? xr * yr - xi * yi : yr * yr + yi * yi;
        q1 += du ? xr * yi + xi * yr : xr * yr + xi * yi;
        q2 += xi * yr - xr * yi;
      }
    }
  }
  if (dot_op != DOT_NONE) dot_partials<BS>(dot, dot_op, q0, q1, q2);
  if (blockIdx.x == 0) {
    const int pad = d.S - d.nv;
    for (int q = threadIdx.x; q < 3 * pad; q += blockDim.x) {
      const int cc = q / pad;
      y[(size_t)cc * d.S + d.nv + (q - cc * pad)] = czero<C>();
    }
  }
}

// Column FFT with a static plan (pass 2 / pass 4 and the tensor generation).
// HIN: in_len <= N/2, the zero upper half of the input is skipped; HOUT:
// out_len <= N/2, the upper half of the output is never formed.
template <typename R, int N, int T, bool W, bool INV, bool HIN, bool HOUT>
__global__ __launch_bounds__(T * (N / StaticPlan<N, W>::PPT)) void k_colfft_s(const typename CxT<R>::T* in,
                                                                                typename CxT<R>::T* out, ColParams cp,
                                                  const typename CxT<R>::T* __restrict__ tw, RhsFlags fl) {
  typedef typename CxT<R>::T C;
  typedef StaticPlan<N, W> P;
  constexpr int PPT = P::PPT;
  typedef typename P::Fwd F;
  typedef typename Reverse<F>::type Out;
  if (stopped(fl, blockIdx.y / fl.per)) return;
  extern __shared__ char smem[];
  C* tile = reinterpret_cast<C*>(smem);
  constexpr int lp = (T & 1) ? T : T + 1;
  const int l = threadIdx.x % T, t = threadIdx.x / T;
  const int col = blockIdx.x * T + l;
  const bool colok = col < cp.ncols;
  const C* src = in + (long long)blockIdx.y * cp.bstride_in + col;
  C* dst = out + (long long)blockIdx.y * cp.bstride_out + col;
  C v[PPT];
#pragma unroll
  for (int e = 0; e < PPT; e++) {
    const int p = first_point<N, PPT, F>(t, e);
    v[e] = (!(HIN && upper_half<F>(e)) && colok && p < cp.in_len) ? src[p * cp.pstride_in] : czero<C>();
  }
  RunStages<N, PPT, INV, 1, C, F, HIN>::run(v, t, tile, lp, l, tw);
#pragma unroll
  for (int e = 0; e < PPT; e++) {
    const int p = first_point<N, PPT, Out>(t, e);
    if (!(HOUT && upper_half<Out>(e)) && colok && p < cp.out_len)
      cstore<INV>(dst + p * cp.pstride_out, v[e]);
  }
}

// Row FFT over contiguous lines (used for the X axis of the tensor generation).
template <typename R, int PPT, bool INV>
__global__ __launch_bounds__(256) void k_rowfft(typename CxT<R>::T* data, long long nlines, FftPlan p, FastDiv ldiv,
                                                FastDiv ndiv, const typename CxT<R>::T* __restrict__ tw) {
  typedef typename CxT<R>::T C;
  extern __shared__ char smem[];
  C* tile = reinterpret_cast<C*>(smem);
  const int L = (int)ldiv.d;
  const int lp = lds_stride(L);
  const int N = p.n;
  const int pts = N * L;
  const long long l0 = (long long)blockIdx.x * L;
  for (int i = threadIdx.x; i < pts; i += blockDim.x) {
    const int l = ndiv.div(i);
    const int pp = i - l * N;
    C v = czero<C>();
    if (l0 + l < nlines) v = data[(l0 + l) * N + pp];
    tile[pp * lp + l] = v;
  }
  __syncthreads();
  fft_tile<PPT, INV>(tile, p, ldiv, lp, tw);
  for (int i = threadIdx.x; i < pts; i += blockDim.x) {
    const int l = ndiv.div(i);
    const int pp = i - l * N;
    if (l0 + l < nlines) data[(l0 + l) * N + pp] = tile[pp * lp + l];
  }
}

// d^3-scaled Green tensor component on the full grid, clipped to |offset| < clip
// per axis (DDA/DMatrixCpu.cpp:49-63,124-131,169-193). Offsets and kd are in
// units of the grid spacing.
__global__ void k_green_fill(double2* w, int comp, int Gx, int Gy, int Gz, int cx, int cy, int cz, double kd) {
  const long long total = (long long)Gx * Gy * Gz;
  const int mu = comp < 3 ? 0 : (comp < 5 ? 1 : 2);
  const int nu = comp == 0 ? 0 : (comp == 1 || comp == 3) ? 1 : 2;
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < total; t += (long long)gridDim.x * blockDim.x) {
    const int x = (int)(t % Gx);
    const int y = (int)((t / Gx) % Gy);
    const int z = (int)(t / ((long long)Gx * Gy));
    const int i = x <= Gx / 2 ? x : x - Gx;
    const int j = y <= Gy / 2 ? y : y - Gy;
    const int k = z <= Gz / 2 ? z : z - Gz;
    double2 v = make_double2(0.0, 0.0);
    if (abs(i) < cx && abs(j) < cy && abs(k) < cz && (i | j | k) != 0) {
      const double r[3] = {(double)i, (double)j, (double)k};
      const double rr = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
      const double kr = kd * rr;
      const double kr2 = kr * kr;
      const double qmn = (r[mu] / rr) * (r[nu] / rr);
      double sn, cs;
      sincos(kr, &sn, &cs);
      const double inv3 = 1.0 / (rr * rr * rr);
      // exp(i kr) / r^3 * [(3 - kr^2 - 3 i kr) q_mu q_nu + delta (kr^2 - 1 + i kr)]
      double br = (3.0 - kr2) * qmn, bi = -3.0 * kr * qmn;
      if (mu == nu) {
        br += kr2 - 1.0;
        bi += kr;
      }
      v.x = (cs * br - sn * bi) * inv3;
      v.y = (cs * bi + sn * br) * inv3;
    }
    w[t] = v;
  }
}

// Octant extraction with the -1/N of DDA/DMatrixCpu.cpp:200.
template <typename R>
__global__ void k_dhat_extract(const double2* __restrict__ w, typename CxT<R>::T* dhat, int comp, int Gx, int Gy,
                               int Gz, double scale) {
  typedef typename CxT<R>::T C;
  const int hx = Gx / 2 + 1, hy = Gy / 2 + 1, hz = Gz / 2 + 1;
  const long long total = (long long)hx * hy * hz;
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < total; t += (long long)gridDim.x * blockDim.x) {
    const int fx = (int)(t % hx);
    const int fy = (int)((t / hx) % hy);
    const int fz = (int)(t / ((long long)hx * hy));
    const double2 v = w[((long long)fz * Gy + fy) * Gx + fx];
    dhat[comp * total + t] = cmake<C>(v.x * scale, v.y * scale);
  }
}

// ------------------------------------------------------------------------------------
// Host side
// ------------------------------------------------------------------------------------

bool make_plan(int n, FftPlan* p) {
  std::memset(p, 0, sizeof(*p));
  p->n = n;
  int m = n, s = 0;
  auto push = [&](int r) {
    p->radix[s++] = r;
    m /= r;
  };
  while (m % 8 == 0 && s < kMaxStages) push(8);
  while (m % 4 == 0 && s < kMaxStages) push(4);
  while (m % 2 == 0 && s < kMaxStages) push(2);
  for (int r : {3, 5, 7})
    while (m % r == 0 && s < kMaxStages) push(r);
  if (m != 1) return false;
  p->nst = s;
  int ns = 1;
  for (int i = 0; i < s; i++) {
    p->ns[i] = ns;
    p->nsdiv[i] = make_fastdiv(ns);
    p->twmul[i] = n / (ns * p->radix[i]);
    ns *= p->radix[i];
  }
  return true;
}

// Host <-> device copies ordered on a stream (never on the null stream, so
// that graph capture on another thread's stream is undisturbed), synchronous.
static hipError_t copy_sync(void* dst, const void* src, size_t bytes, hipMemcpyKind kind, hipStream_t st) {
  hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, st);
  return e == hipSuccess ? hipStreamSynchronize(st) : e;
}

dda_status make_twiddles(int n, int prec, void** out, hipStream_t st) {
  const long double pi = 3.1415926535897932384626433832795028L;
  size_t esz = prec == 64 ? sizeof(double2) : sizeof(float2);
  std::vector<char> h(esz * n);
  for (int t = 0; t < n; t++) {
    long double a = -2 * pi * (long double)t / (long double)n;
    double c = (double)cosl(a), s = (double)sinl(a);
    if (prec == 64)
      reinterpret_cast<double2*>(h.data())[t] = make_double2(c, s);
    else
      reinterpret_cast<float2*>(h.data())[t] = make_float2((float)c, (float)s);
  }
  DDA_TRY_HIP(hipMalloc(out, h.size()));
  DDA_TRY_HIP(copy_sync(*out, h.data(), h.size(), hipMemcpyHostToDevice, st));
  return DDA_OK;
}

namespace {

inline int round_up(int v, int m) { return (v + m - 1) / m * m; }

// Threads per block for a tile of `pts` points with PPT points per thread.
inline int block_for(int pts, int ppt) {
  int b = round_up((pts + ppt - 1) / ppt, 64);
  return b < 64 ? 64 : b;
}

// Columns per work-group: `pref`, halved until the tile fits 16 points per thread.
inline int col_tile(int n, int pref) {
  int t = pref;
  while (t > 1 && n * t > 16 * 256) t /= 2;
  return t;
}

template <typename K>
dda_status set_lds(K kernel, size_t bytes) {
  if (bytes > 64 * 1024) DDA_TRY_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                                                         hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
  return DDA_OK;
}

template <typename R, bool INV>
dda_status launch_col(const void* in, void* out, const ColParams& cp, const FftPlan& p, int T, const void* tw,
                      int nbatch, hipStream_t st, RhsFlags done) {
  typedef typename CxT<R>::T C;
  const int pts = p.n * T;
  const size_t lds = (size_t)p.n * lds_stride(T) * sizeof(C);
  dim3 grid((cp.ncols + T - 1) / T, nbatch);
  FastDiv tdiv = make_fastdiv(T);
  if (pts <= 8 * 256) {
    auto k = k_colfft<R, 8, INV>;
    DDA_TRY(set_lds(k, lds));
    hipLaunchKernelGGL(k, grid, dim3(block_for(pts, 8)), lds, st, (const C*)in, (C*)out, cp, p, tdiv, (const C*)tw,
                       done);
  } else if (pts <= 16 * 256) {
    auto k = k_colfft<R, 16, INV>;
    DDA_TRY(set_lds(k, lds));
    hipLaunchKernelGGL(k, grid, dim3(block_for(pts, 16)), lds, st, (const C*)in, (C*)out, cp, p, tdiv,
                       (const C*)tw, done);
  } else {
    set_error("column FFT tile too large");
    return DDA_NOT_SUPPORTED;
  }
  DDA_TRY_HIP(hipGetLastError());
  return DDA_OK;
}

template <typename R>
dda_status launch_row(void* data, long long nlines, const FftPlan& p, const void* tw, hipStream_t st) {
  typedef typename CxT<R>::T C;
  int L = (8 * 256) / p.n;
  if (L < 1) L = 1;
  int ppt = 8;
  if (p.n * L > 8 * 256) ppt = 16;
  const int pts = p.n * L;
  if (pts > 16 * 256) {
    set_error("row FFT too long");
    return DDA_NOT_SUPPORTED;
  }
  const size_t lds = (size_t)p.n * lds_stride(L) * sizeof(C);
  dim3 grid((unsigned)((nlines + L - 1) / L));
  FastDiv ldiv = make_fastdiv(L), ndiv = make_fastdiv(p.n);
  if (ppt == 8) {
    auto k = k_rowfft<R, 8, false>;
    DDA_TRY(set_lds(k, lds));
    hipLaunchKernelGGL(k, grid, dim3(block_for(pts, 8)), lds, st, (C*)data, nlines, p, ldiv, ndiv, (const C*)tw);
  } else {
    auto k = k_rowfft<R, 16, false>;
    DDA_TRY(set_lds(k, lds));
    hipLaunchKernelGGL(k, grid, dim3(block_for(pts, 16)), lds, st, (C*)data, nlines, p, ldiv, ndiv, (const C*)tw);
  }
  DDA_TRY_HIP(hipGetLastError());
  return DDA_OK;
}


// Columns per work-group for a static plan: 128-byte row segments, at most 256
// threads.
// Long columns (512, 1024 points) keep 128-byte segments with larger
// work-groups: up to 1024 threads for a single-component transform, 512 for
// the Z pass, whose threads hold three components. PREF (if nonzero) asks for
// that many columns instead.
template <typename R, int N, bool W, int MAXT = 256, int PREF = 0>
constexpr int static_cols() {
  constexpr int pref = PREF ? PREF : sizeof(R) == 8 ? 8 : 16;
  constexpr int tpc = N / StaticPlan<N, W>::PPT;
  constexpr int cap = N >= 512 ? MAXT : 256;
  // at least one full wavefront per work-group
  constexpr int t = pref * tpc < 64 ? 64 / tpc : pref;
  return t * tpc <= cap ? t : cap / tpc;
}

// Plan width of a static kernel: 1 = wide (16 points per thread where the plan
// has them), 0 = narrow, -1 = the precision's default (wide for complex float).
template <typename R>
inline bool plan_wide(int w) {
  return w < 0 ? sizeof(R) == 4 : w != 0;
}

// The other tile the tuner tries for the FP32 Y passes: twice the default
// columns per work-group where that stays within 1024 threads, else half
// (keeping at least one wavefront).
template <int T, int TPC>
constexpr int other_cols() {
  return 2 * T * TPC <= 1024 ? 2 * T : (T / 2) * TPC >= 64 ? T / 2 : T;
}
// and the third: four times the default columns where that stays within 1024
// threads and the LDS (FP32 tiles of N x (4T + 1) points), else the second
template <int T, int TPC, int N>
constexpr int third_cols() {
  return 4 * T * TPC <= 1024 && (size_t)N * (4 * T + 1) * 8 <= 160 * 1024 ? 4 * T : other_cols<T, TPC>();
}

// ALT: 0 the default tile, 1 other_cols, 2 third_cols
template <typename R, int N, bool W, bool INV, int ALT = 0>
dda_status launch_col_static_w(const void* in, void* out, const ColParams& cp, const void* tw, int nbatch,
                               hipStream_t st, RhsFlags done) {
  typedef typename CxT<R>::T C;
  constexpr int T0 = static_cols<R, N, W, 1024>();
  constexpr int TPC0 = N / StaticPlan<N, W>::PPT;
  constexpr int T = ALT == 0 ? T0 : ALT == 1 ? other_cols<T0, TPC0>() : third_cols<T0, TPC0, N>();
  constexpr int threads = T * (N / StaticPlan<N, W>::PPT);
  constexpr int lp = (T & 1) ? T : T + 1;
  const size_t lds = (size_t)N * lp * sizeof(C);
  // the matvec's forward pass reads at most N/2 points and its inverse pass
  // keeps at most N/2; the tensor generation transforms full lengths
  const bool hin = !INV && 2 * cp.in_len <= N, hout = INV && 2 * cp.out_len <= N;
  void (*k)(const C*, C*, ColParams, const C*, RhsFlags);
  if constexpr (ALT != 0) {
    // the other tiles are matvec variants: full-length transforms keep the default
    if (!hin && !hout) return launch_col_static_w<R, N, W, INV, 0>(in, out, cp, tw, nbatch, st, done);
    k = INV ? k_colfft_s<R, N, T, W, INV, false, INV> : k_colfft_s<R, N, T, W, INV, !INV, false>;
  } else {
    k = hin ? k_colfft_s<R, N, T, W, INV, !INV, false>
            : hout ? k_colfft_s<R, N, T, W, INV, false, INV> : k_colfft_s<R, N, T, W, INV, false, false>;
  }
  DDA_TRY(set_lds(k, lds));
  hipLaunchKernelGGL(k, dim3((cp.ncols + T - 1) / T, nbatch), dim3(threads), lds, st, (const C*)in, (C*)out, cp,
                     (const C*)tw, done);
  DDA_TRY_HIP(hipGetLastError());
  return DDA_OK;
}

// wide: -1 the precision's default plan, 0 narrow, 1 wide; FP32 also 2 and 3,
// narrow and wide with the second tile (other_cols), and 4 and 5 with the
// third (third_cols).
template <typename R, int N, bool INV>
dda_status launch_col_static(const void* in, void* out, const ColParams& cp, const void* tw, int nbatch, hipStream_t st,
                             RhsFlags done, int wide) {
  if constexpr (sizeof(R) == 4) {
    if (wide >= 4)
      return (wide & 1) ? launch_col_static_w<R, N, true, INV, 2>(in, out, cp, tw, nbatch, st, done)
                        : launch_col_static_w<R, N, false, INV, 2>(in, out, cp, tw, nbatch, st, done);
    if (wide >= 2)
      return (wide & 1) ? launch_col_static_w<R, N, true, INV, 1>(in, out, cp, tw, nbatch, st, done)
                        : launch_col_static_w<R, N, false, INV, 1>(in, out, cp, tw, nbatch, st, done);
  }
  return plan_wide<R>(wide >= 2 ? (wide & 1) : wide) ? launch_col_static_w<R, N, true, INV>(in, out, cp, tw, nbatch, st, done)
                                                    : launch_col_static_w<R, N, false, INV>(in, out, cp, tw, nbatch, st, done);
}

// The buffers of right-hand sides r0 .. r0 + nrhs - 1 of an operator: what one
// sequence of the five passes works on (the whole block, or one half of it
// when the halves run on two streams).
struct MvView {
  void* a1;
  void* a2;
  const void* ccs;
  MvDims d;  // member table shifted to r0
};

inline MvView view_of(const dda_op_s* op, int r0) {
  const MvDims& d = op->dims;
  const size_t csz = op->prec == 64 ? sizeof(double2) : sizeof(float2);
  MvView v;
  v.a1 = (char*)op->a1 + (size_t)r0 * 3 * d.Bz * d.By * d.Gx * csz;
  v.a2 = (char*)op->a2 + (size_t)r0 * 3 * d.Bz * d.Gy * d.Gx * csz;
  v.ccs = (const char*)op->ccs + (size_t)r0 * 3 * d.nmat * csz;
  v.d = d;
  if (d.member) v.d.member = d.member + r0;
  return v;
}

template <typename R, int N, typename PL, int T>
dda_status launch_pass3_plan(dda_op_s* op, const MvView& v, hipStream_t st, RhsFlags done, int nrhs) {
  typedef typename CxT<R>::T C;
  constexpr int threads = T * (N / PL::PPT);
  constexpr int lp = (T & 1) ? T : T + 1;
  const MvDims& d = v.d;
  const size_t lds = (size_t)N * lp * sizeof(C);
  const int ntx = (d.Gx + T - 1) / T;
  const int nwork = (d.Gy / 2 + 1) * ((ntx + 1) / 2) * 4 * nrhs;
  const int nper = (nwork + 7) / 8;
  auto k = k_pass3s<R, N, T, PL>;
  DDA_TRY(set_lds(k, lds));
  hipLaunchKernelGGL(k, dim3(8 * nper), dim3(threads), lds, st, (C*)v.a2, (const C*)op->dhat, d,
                     (const C*)op->tw[2], nwork, nper, nrhs, done);
  DDA_TRY_HIP(hipGetLastError());
  return DDA_OK;
}

// The other tile the tuner tries for the FP32 Z pass: 32 columns where the
// default is 16 and 32 fit the work-group size the kernel's register budget
// allows (MAXTH threads), else half the default.
template <int T, int TPC, int MAXTH>
constexpr int other_cols_z() {
  return T != 32 && 32 * TPC <= MAXTH ? 32 : (T / 2) * TPC >= 64 ? T / 2 : T;
}

template <typename R, int N, bool W, bool ALT = false>
dda_status launch_pass3_static_w(dda_op_s* op, const MvView& v, hipStream_t st, RhsFlags done, int nrhs) {
  // FP32 1024-point columns: 16 columns (128-byte segments) in a 1024-thread
  // work-group at 128 VGPRs, the CU's whole register file (17.5 ms at 1024^3,
  // against 22.7 ms for two 512-thread work-groups of 8 columns)
  // short FP32 columns (at most 64 points): 32 columns (256-byte segments)
  // per work-group; measured on a 128-RHS block at a 64-point Z grid: 935 ->
  // 781 us per launch
  constexpr int T0 = static_cols<R, N, W, (N == 1024 && sizeof(R) == 4 && W) ? 1024 : 512,
                                 (sizeof(R) == 4 && N <= 64) ? 32 : 0>();
  constexpr int TPC = N / StaticPlan<N, W>::PPT;
  constexpr int T = ALT ? other_cols_z<T0, TPC, 256 * P3Waves<R, N, StaticPlan<N, W>::PPT>::value>() : T0;
  return launch_pass3_plan<R, N, StaticPlan<N, W>, T>(op, v, st, done, nrhs);
}

// Z-pass variants (Tune::p3_wide): -1 the precision's default, 0 narrow,
// 1 wide; at 1024 points also 2 and 3, the 32 x 32 plan (32 points per
// thread, one LDS exchange per transform, 2 waves per SIMD, FP32) over 16
// columns (512 threads, 128-byte segments, one work-group per CU by LDS) or
// 8 columns (256 threads, 64-byte segments, two work-groups per CU).
// FP32 only: in FP64 the 32-point plan's data alone needs 384 VGPRs. The FP32
// default at 1024 points is 2: measured at 1024^3 (sphere 512 across) 14.8 ms
// against 17.8 ms (wide), 19.2 ms (variant 3) and 33.0 ms (narrow).
// (Two further forms were measured and dropped. The wide plan without held
// twiddle powers at 4 waves per SIMD, 128 VGPRs: 308.8 against 293.2 us per
// 2-RHS FP32 launch at 256^3, and it spills at 32, 128 and 512 points. A
// persistent kernel, resident work-groups looping over the tiles and loading
// each component of the next tile into the registers its store frees:
// 303.5 against 289.3 us (2-RHS FP32, 256^3), 315.0 against 291.4 us (FP64),
// 15.1 against 14.6 ms (FP32, 1024^3); its loop costs 13% more VALU at 1024
// points than the hidden loads save.)
typedef PlanOf<1024, 32, RadixList<32, 32>> Plan1024x32;
// FP32 below 1024 points: 2 and 3 are the narrow and wide plans with the
// other tile (other_cols_z).
inline int p3_variants(int n, int prec) { return prec == 32 ? 4 : 2; }

template <typename R, int N>
dda_status launch_pass3_static(dda_op_s* op, const MvView& v, hipStream_t st, RhsFlags done, int nrhs,
                               const dda_op_s::Tune& tn) {
  const int var = tn.p3_wide;
  if constexpr (N == 1024 && sizeof(R) == 4) {
    if (var == 2 || var < 0) return launch_pass3_plan<R, N, Plan1024x32, 16>(op, v, st, done, nrhs);
    if (var == 3) return launch_pass3_plan<R, N, Plan1024x32, 8>(op, v, st, done, nrhs);
  } else if constexpr (sizeof(R) == 4) {
    if (var >= 2)
      return (var & 1) ? launch_pass3_static_w<R, N, true, true>(op, v, st, done, nrhs)
                       : launch_pass3_static_w<R, N, false, true>(op, v, st, done, nrhs);
  }
  return plan_wide<R>(var >= 2 ? (var & 1) : var) ? launch_pass3_static_w<R, N, true>(op, v, st, done, nrhs)
                           : launch_pass3_static_w<R, N, false>(op, v, st, done, nrhs);
}

template <typename R, int N, bool W, int BS = 256>
dda_status launch_x_static_w(dda_op_s* op, const MvView& v, bool inverse, const void* x, void* y, int conj,
                             hipStream_t st, RhsFlags done, int nrhs, const PassDotReq& dr) {
  typedef typename CxT<R>::T C;
  constexpr int TPC = N / StaticPlan<N, W>::PPT;
  constexpr int LN = BS / TPC;
  constexpr int lp = (LN & 1) ? LN : LN + 1;
  const MvDims& d = v.d;
  const size_t lds = (size_t)N * lp * sizeof(C);
  if (!inverse) {
    auto k = k_pass1s<R, N, W, BS>;
    DDA_TRY(set_lds(k, lds));
    hipLaunchKernelGGL(k, dim3((3 * d.nrows + LN - 1) / LN, nrhs), dim3(LN * TPC), lds, st, (const C*)x, (C*)v.a1,
                       op->row_start, op->row_dip, op->row_meta, (const C*)v.ccs, d, (const C*)op->tw[0], conj,
                       done);
  } else {
    auto k = k_pass5s<R, N, W, BS>;
    DDA_TRY(set_lds(k, lds));
    const int nl = 3 * op->n_nonempty;
    const int nbx = nl > 0 ? (nl + LN - 1) / LN : 1;
    op->p5_blocks = nbx;
    double* dot = dr.op != DOT_NONE ? dr.part + (size_t)dr.r0 * 4 * nbx : nullptr;
    hipLaunchKernelGGL(k, dim3(nbx, nrhs), dim3(LN * TPC), lds, st, (const C*)v.a1, (const C*)x, (C*)y, op->row_list,
                       op->n_nonempty, op->row_start, op->row_dip, op->row_meta, (const C*)v.ccs, d,
                       (const C*)op->tw[0], conj, done, dot, dr.op);
  }
  DDA_TRY_HIP(hipGetLastError());
  return DDA_OK;
}

template <typename R, int N>
dda_status launch_x_static(dda_op_s* op, const MvView& v, bool inverse, const void* x, void* y, int conj,
                           hipStream_t st, RhsFlags done, int nrhs, const dda_op_s::Tune& tn,
                           const PassDotReq& dr = PassDotReq()) {
  // default narrow for both precisions: half the LDS per wave, twice the waves
  // per CU (measured 52 -> 39 us at 256^3 FP32); FP32 also 2 and 3, narrow
  // and wide over 512-thread work-groups (twice the lines each)
  const int var = tn.x_wide[inverse];
  if constexpr (sizeof(R) == 4) {
    if (var >= 2)
      return (var & 1) ? launch_x_static_w<R, N, true, 512>(op, v, inverse, x, y, conj, st, done, nrhs, dr)
                       : launch_x_static_w<R, N, false, 512>(op, v, inverse, x, y, conj, st, done, nrhs, dr);
  }
  return var >= 0 && (var & 1) ? launch_x_static_w<R, N, true>(op, v, inverse, x, y, conj, st, done, nrhs, dr)
                               : launch_x_static_w<R, N, false>(op, v, inverse, x, y, conj, st, done, nrhs, dr);
}

// Lengths with a static plan; others use the run-time planned kernels.
#define DDA_STATIC_SIZES(X) X(32) X(48) X(64) X(96) X(128) X(192) X(256) X(512) X(1024)

inline bool has_static_plan(int n) {
  switch (n) {
#define X(N) case N:
    DDA_STATIC_SIZES(X)
#undef X
    return true;
    default:
      return false;
  }
}

// The static Z pass (k_pass3s) addresses A2 and each component of the tensor
// octant with 32-bit byte offsets from uniform bases: a tensor component spans
// (Gx/2+1)(Gy/2+1)(Gz/2+1) elements, and one buffer range of A2 at most
// Gz/8 planes of Gy*Gx elements (its plans have at least 8 points per
// thread). Grids beyond that run the run-time planned Z pass, whose addressing
// is 64-bit.
inline bool static_z_fits(int Gx, int Gy, int Gz, size_t csz) {
  const uint64_t lim = 1ull << 32;
  const uint64_t oct = (uint64_t)(Gx / 2 + 1) * (Gy / 2 + 1) * (Gz / 2 + 1);
  const uint64_t span = (uint64_t)(Gz / 8 + 1) * Gy * Gx;
  return has_static_plan(Gz) && oct * csz < lim && span * csz < lim;
}

template <typename R, bool INV>
dda_status launch_col_any(const void* in, void* out, const ColParams& cp, const FftPlan& p, int T, const void* tw,
                          int nbatch, hipStream_t st, RhsFlags done, int wide = -1) {
  switch (p.n) {
#define X(N) \
  case N:    \
    return launch_col_static<R, N, INV>(in, out, cp, tw, nbatch, st, done, wide);
    DDA_STATIC_SIZES(X)
#undef X
    default:
      return launch_col<R, INV>(in, out, cp, p, T, tw, nbatch, st, done);
  }
}

}  // namespace

// Timing events of dda_op_profile, recycled between reads.
static hipEvent_t pool_event(dda_op_s* op) {
  if (!op->prof_pool.empty()) {
    hipEvent_t e = op->prof_pool.back();
    op->prof_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

// Brackets pass `pass` with events when profiling is on.
struct PassTimer {
  dda_op_s* op;
  hipStream_t st;
  int pass;
  hipEvent_t b = nullptr, e = nullptr;
  PassTimer(dda_op_s* o, hipStream_t s, int p) : op(o), st(s), pass(p) {
    if (op->prof) {
      b = pool_event(op);
      e = pool_event(op);
      if (b) (void)hipEventRecord(b, st);
    }
  }
  ~PassTimer() {
    if (b && e) {
      (void)hipEventRecord(e, st);
      op->prof_recs.push_back({pass, b, e});
    }
  }
};

// Passes p0 .. p1 (0-based: X forward, Y forward, Z + tensor, Y inverse, X
// inverse) of the matvec for right-hand sides r0 .. r0 + nrhs - 1 on `st`.
template <typename R>
static dda_status launch_passes(dda_op_s* op, const void* xv, void* yv, int conj, hipStream_t st, RhsFlags done,
                                int r0, int nrhs, int p0, int p1, const dda_op_s::Tune& tn,
                                const PassDotReq& dot = PassDotReq()) {
  typedef typename CxT<R>::T C;
  const MvView v = view_of(op, r0);
  const MvDims& d = v.d;
  const C* x = (const C*)xv + (size_t)r0 * 3 * d.S;
  C* y = (C*)yv + (size_t)r0 * 3 * d.S;
  if (done.done) done.done += (long long)r0 * done.stride;
  C* a1 = (C*)v.a1;
  C* a2 = (C*)v.a2;
  const C* ccs = (const C*)v.ccs;
  RhsFlags col_done = done;
  col_done.per = 3 * d.Bz;
  // The Y passes launch one grid row per (RHS, component, z) column batch; a
  // launch holds at most 65535 rows, so larger blocks run as several launches
  // of whole right-hand sides.
  const int rhs_per_col_launch = 65535 / (3 * d.Bz);

  // pass 1 / pass 5 shapes
  const int lr = op->p1_rows;
  const int L = 3 * lr;
  const int pts_x = d.Gx * L;
  const int ppt_x = pts_x <= 8 * 256 ? 8 : 16;
  const size_t lds_x = (size_t)d.Gx * lds_stride(L) * sizeof(C);
  const FastDiv ldiv = make_fastdiv(L), ndiv = make_fastdiv(d.Gx);
  if (p0 <= 0 && 0 <= p1) {
    PassTimer pt(op, st, 0);
    if (op->static_x) {
      switch (d.Gx) {
#define X(N)                                                                         \
  case N:                                                                            \
    DDA_TRY((launch_x_static<R, N>(op, v, false, x, nullptr, conj, st, done, nrhs, tn))); \
    break;
        DDA_STATIC_SIZES(X)
#undef X
      }
    } else {
      dim3 grid((d.nrows + lr - 1) / lr, nrhs);
      if (ppt_x == 8) {
        auto k = k_pass1<R, 8>;
        DDA_TRY(set_lds(k, lds_x));
        hipLaunchKernelGGL(k, grid, dim3(block_for(pts_x, 8)), lds_x, st, x, a1, op->row_start, op->row_dip,
                           op->row_meta, ccs, d, lr, op->plan[0], ldiv, ndiv, (const C*)op->tw[0], conj, done);
      } else {
        auto k = k_pass1<R, 16>;
        DDA_TRY(set_lds(k, lds_x));
        hipLaunchKernelGGL(k, grid, dim3(block_for(pts_x, 16)), lds_x, st, x, a1, op->row_start, op->row_dip,
                           op->row_meta, ccs, d, lr, op->plan[0], ldiv, ndiv, (const C*)op->tw[0], conj, done);
      }
      DDA_TRY_HIP(hipGetLastError());
    }
  }
  // pass 2: Y forward, A1 -> A2, batch = (c, z)
  if (p0 <= 1 && 1 <= p1) {
    PassTimer pt(op, st, 1);
    ColParams cp;
    cp.ncols = d.Gx;
    cp.in_len = d.By;
    cp.out_len = d.Gy;
    cp.pstride_in = d.Gx;
    cp.pstride_out = d.Gx;
    cp.bstride_in = (long long)d.By * d.Gx;
    cp.bstride_out = (long long)d.Gy * d.Gx;
    for (int q0 = 0; q0 < nrhs; q0 += rhs_per_col_launch) {
      const int k = std::min(rhs_per_col_launch, nrhs - q0);
      const long long nb0 = 3ll * d.Bz * q0;
      RhsFlags f = col_done;
      if (f.done) f.done += (long long)q0 * f.stride;
      DDA_TRY((launch_col_any<R, false>(a1 + nb0 * cp.bstride_in, a2 + nb0 * cp.bstride_out, cp, op->plan[1],
                                        op->col_t, op->tw[1], 3 * d.Bz * k, st, f, tn.col_wide[0])));
    }
  }
  // pass 3: Z forward, D multiply, Z inverse, in place on A2
  if (p0 <= 2 && 2 <= p1) {
    PassTimer pt(op, st, 2);
    if (op->static_z) {
      switch (d.Gz) {
#define X(N)                                                         \
  case N:                                                            \
    DDA_TRY((launch_pass3_static<R, N>(op, v, st, done, nrhs, tn))); \
    break;
        DDA_STATIC_SIZES(X)
#undef X
        default:
          break;
      }
    } else {
      const int T = op->col_t;
      const int pts = d.Gz * T;
      const size_t lds = (size_t)d.Gz * lds_stride(T) * sizeof(C);
      const int ntx = (d.Gx + T - 1) / T;
      const int nwork = (d.Gy / 2 + 1) * ((ntx + 1) / 2) * 4 * nrhs;
      const int nper = (nwork + 7) / 8;
      dim3 grid(8 * nper);
      FastDiv tdiv = make_fastdiv(T);
      if (pts <= 8 * 256) {
        auto k = k_pass3<R, 8>;
        DDA_TRY(set_lds(k, lds));
        hipLaunchKernelGGL(k, grid, dim3(block_for(pts, 8)), lds, st, a2, (const C*)op->dhat, d, op->plan[2], tdiv,
                           (const C*)op->tw[2], nwork, nper, nrhs, done);
      } else if (pts <= 16 * 256) {
        auto k = k_pass3<R, 16>;
        DDA_TRY(set_lds(k, lds));
        hipLaunchKernelGGL(k, grid, dim3(block_for(pts, 16)), lds, st, a2, (const C*)op->dhat, d, op->plan[2], tdiv,
                           (const C*)op->tw[2], nwork, nper, nrhs, done);
      } else {
        set_error("Z pass tile too large");
        return DDA_NOT_SUPPORTED;
      }
      DDA_TRY_HIP(hipGetLastError());
    }
  }
  // pass 4: Y inverse, A2 -> A1 cropped to y < By
  if (p0 <= 3 && 3 <= p1) {
    PassTimer pt(op, st, 3);
    ColParams cp;
    cp.ncols = d.Gx;
    cp.in_len = d.Gy;
    cp.out_len = d.By;
    cp.pstride_in = d.Gx;
    cp.pstride_out = d.Gx;
    cp.bstride_in = (long long)d.Gy * d.Gx;
    cp.bstride_out = (long long)d.By * d.Gx;
    for (int q0 = 0; q0 < nrhs; q0 += rhs_per_col_launch) {
      const int k = std::min(rhs_per_col_launch, nrhs - q0);
      const long long nb0 = 3ll * d.Bz * q0;
      RhsFlags f = col_done;
      if (f.done) f.done += (long long)q0 * f.stride;
      DDA_TRY((launch_col_any<R, true>(a2 + nb0 * cp.bstride_in, a1 + nb0 * cp.bstride_out, cp, op->plan[1],
                                       op->col_t, op->tw[1], 3 * d.Bz * k, st, f, tn.col_wide[1])));
    }
  }
  // pass 5
  if (p0 <= 4 && 4 <= p1) {
    PassTimer pt(op, st, 4);
    PassDotReq dr = dot;
    dr.r0 += r0;
    if (dot.op != DOT_NONE && !op->static_x) {
      set_error("matvec: partial sums in pass 5 need a compile-time planned X pass");
      return DDA_NOT_SUPPORTED;
    }
    if (op->static_x) {
      switch (d.Gx) {
#define X(N)                                                                    \
  case N:                                                                       \
    DDA_TRY((launch_x_static<R, N>(op, v, true, x, y, conj, st, done, nrhs, tn, dr))); \
    break;
        DDA_STATIC_SIZES(X)
#undef X
      }
    } else {
      dim3 grid((op->n_nonempty + lr - 1) / lr > 0 ? (op->n_nonempty + lr - 1) / lr : 1, nrhs);
      if (ppt_x == 8) {
        auto k = k_pass5<R, 8>;
        DDA_TRY(set_lds(k, lds_x));
        hipLaunchKernelGGL(k, grid, dim3(block_for(pts_x, 8)), lds_x, st, a1, x, y, op->row_list, op->n_nonempty,
                           op->row_start, op->row_dip, op->row_meta, ccs, d, lr, op->plan[0], ldiv, ndiv,
                           (const C*)op->tw[0], conj, done);
      } else {
        auto k = k_pass5<R, 16>;
        DDA_TRY(set_lds(k, lds_x));
        hipLaunchKernelGGL(k, grid, dim3(block_for(pts_x, 16)), lds_x, st, a1, x, y, op->row_list, op->n_nonempty,
                           op->row_start, op->row_dip, op->row_meta, ccs, d, lr, op->plan[0], ldiv, ndiv,
                           (const C*)op->tw[0], conj, done);
      }
      DDA_TRY_HIP(hipGetLastError());
    }
  }
  return DDA_OK;
}

// The second stream and the two events of a split block matvec, made on first use.
static dda_status split_resources(dda_op_s* op) {
  if (op->split_stream) return DDA_OK;
  DDA_TRY_HIP(hipStreamCreateWithFlags(&op->split_stream, hipStreamNonBlocking));
  for (auto& e : op->split_ev) DDA_TRY_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  return DDA_OK;
}

// The five passes of one matvec on `st`. done (nullable) lets a device-side
// solver turn the launches into no-ops once it has finished.
//
// A block of two or more right-hand sides with the split enabled runs as two
// halves on two streams, staggered by two passes: the second half starts once
// the first has finished its Y pass, so the Z pass of each half (bound by its
// arithmetic) can run beside the memory-bound X and Y passes of the other. The
// fork and the join are event waits, which a stream capture records as graph
// edges. Per-pass profiling keeps the block on one stream. Off unless the
// tuner (or dda_op_set_variants) picks it: measured on the 2-RHS FP32 block
// at 256^3, 707 us split against 642 us on one stream.
template <typename R>
dda_status launch_matvec(dda_op_s* op, const void* xv, void* yv, int conj, hipStream_t st, const int* done_ptr,
                         int nrhs, int done_stride, double* dot_part, int dot_op) {
  if (nrhs < 1 || nrhs > op->cap_rhs) {
    set_error("matvec: number of right-hand sides exceeds the operator's capacity");
    return DDA_INVALID_ARG;
  }
  if (nrhs > 65535 || 3 * op->dims.Bz > 65535) {
    set_error("matvec: more than 65535 right-hand sides, or 3 * box z above 65535");
    return DDA_NOT_SUPPORTED;
  }
  RhsFlags done;
  done.done = done_ptr;
  done.stride = done_stride;
  done.per = 1;
  // blocks of kBlockTuneRhs or more run with their own tuning, when there is one
  const dda_op_s::Tune& tn = nrhs >= kBlockTuneRhs && op->tuned_blk ? op->tune_blk : op->tune;
  const bool split = (tn.split < 0 ? op->split_rhs : tn.split != 0) && nrhs >= 2 && !op->prof;
  PassDotReq dot;
  dot.part = dot_part;
  dot.op = dot_part ? dot_op : DOT_NONE;
  if (!split) return launch_passes<R>(op, xv, yv, conj, st, done, 0, nrhs, 0, 4, tn, dot);
  DDA_TRY(split_resources(op));
  const int h = (nrhs + 1) / 2;
  hipStream_t aux = op->split_stream;
  DDA_TRY((launch_passes<R>(op, xv, yv, conj, st, done, 0, h, 0, 1, tn)));
  DDA_TRY_HIP(hipEventRecord(op->split_ev[0], st));
  DDA_TRY_HIP(hipStreamWaitEvent(aux, op->split_ev[0], 0));
  DDA_TRY((launch_passes<R>(op, xv, yv, conj, st, done, 0, h, 2, 4, tn, dot)));
  DDA_TRY((launch_passes<R>(op, xv, yv, conj, aux, done, h, nrhs - h, 0, 4, tn, dot)));
  DDA_TRY_HIP(hipEventRecord(op->split_ev[1], aux));
  DDA_TRY_HIP(hipStreamWaitEvent(st, op->split_ev[1], 0));
  return DDA_OK;
}

template dda_status launch_matvec<float>(dda_op_s*, const void*, void*, int, hipStream_t, const int*, int, int,
                                         double*, int);
template dda_status launch_matvec<double>(dda_op_s*, const void*, void*, int, hipStream_t, const int*, int, int,
                                          double*, int);

// Generate the octant of the FFT'd, d^3-scaled tensor on the device in FP64:
// fill one component on the full grid, 3-D FFT (X rows, Y and Z columns), keep
// the octant times -1/N, converted to the operator's precision.
// clip: offsets with |offset| >= clip per axis are zero (the reference clips at
// the box; a tensor clipped anywhere between the box and G/2 gives the same
// products for every dipole inside the box). Writes the octant into dst in the
// operator's precision.
static dda_status generate_dhat(dda_op_s* op, int cx, int cy, int cz, void* dst) {
  const MvDims& d = op->dims;
  hipStream_t st = op->ctx->stream;
  const long long total = (long long)d.Gx * d.Gy * d.Gz;
  double2* w = nullptr;
  DDA_TRY_HIP(hipMalloc(&w, total * sizeof(double2)));
  void* tw64[3] = {nullptr, nullptr, nullptr};
  dda_status rc = DDA_OK;
  for (int a = 0; a < 3 && rc == DDA_OK; a++) rc = make_twiddles(op->plan[a].n, 64, &tw64[a], st);
  const double kd = op->wave_num * op->grid_unit;
  const double scale = -1.0 / (double)total;
  for (int comp = 0; comp < 6 && rc == DDA_OK; comp++) {
    hipLaunchKernelGGL(k_green_fill, dim3(2048), dim3(256), 0, st, w, comp, d.Gx, d.Gy, d.Gz, cx, cy, cz, kd);
    rc = launch_row<double>(w, (long long)d.Gy * d.Gz, op->plan[0], tw64[0], st);
    if (rc != DDA_OK) break;
    ColParams cy;
    cy.ncols = d.Gx;
    cy.in_len = cy.out_len = d.Gy;
    cy.pstride_in = cy.pstride_out = d.Gx;
    cy.bstride_in = cy.bstride_out = (long long)d.Gx * d.Gy;
    rc = launch_col_any<double, false>(w, w, cy, op->plan[1], col_tile(d.Gy, 8), tw64[1], d.Gz, st, RhsFlags());
    if (rc != DDA_OK) break;
    ColParams cz;
    cz.ncols = d.Gx * d.Gy;
    cz.in_len = cz.out_len = d.Gz;
    cz.pstride_in = cz.pstride_out = (long long)d.Gx * d.Gy;
    cz.bstride_in = cz.bstride_out = 0;
    rc = launch_col_any<double, false>(w, w, cz, op->plan[2], col_tile(d.Gz, 8), tw64[2], 1, st, RhsFlags());
    if (rc != DDA_OK) break;
    if (op->prec == 64)
      hipLaunchKernelGGL(k_dhat_extract<double>, dim3(1024), dim3(256), 0, st, w, (double2*)dst, comp, d.Gx, d.Gy,
                         d.Gz, scale);
    else
      hipLaunchKernelGGL(k_dhat_extract<float>, dim3(1024), dim3(256), 0, st, w, (float2*)dst, comp, d.Gx, d.Gy,
                         d.Gz, scale);
    if (hipGetLastError() != hipSuccess) {
      set_error("tensor generation launch failed");
      rc = DDA_HIP_ERROR;
    }
  }
  hipError_t e = hipStreamSynchronize(st);
  for (int a = 0; a < 3; a++)
    if (tw64[a]) (void)hipFree(tw64[a]);
  (void)hipFree(w);
  if (rc != DDA_OK) return rc;
  DDA_TRY_HIP(e);
  return DDA_OK;
}

__global__ void k_narrow(const double2* __restrict__ in, float2* out, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    out[i] = make_float2((float)in[i].x, (float)in[i].y);
}

}  // namespace dda

using namespace dda;

// ------------------------------------------------------------------------------------
// Device cache of FFT'd interaction tensors keyed by (grid, kd, precision)
// ------------------------------------------------------------------------------------

struct dda_dhat_cache_s {
  struct Entry {
    void* ptr = nullptr;
    size_t bytes = 0;
    int refs = 0;
    bool ready = false;
    uint64_t last_use = 0;
  };
  typedef std::tuple<int, int, int, uint64_t, int> Key;  // grid, bits of kd, precision
  int device = 0;
  size_t max_bytes = 0, bytes = 0;
  uint64_t clock = 0, hits = 0, misses = 0, evictions = 0;
  std::mutex mu;
  std::condition_variable cv;
  std::map<Key, Entry> map;
};

static dda_dhat_cache_s::Key cache_key(const MvDims& d, double kd, int prec) {
  uint64_t bits;
  std::memcpy(&bits, &kd, sizeof bits);
  return dda_dhat_cache_s::Key(d.Gx, d.Gy, d.Gz, bits, prec);
}

// Frees unreferenced entries, least recently used first, until `need` more
// bytes fit under the limit (or nothing is left to free). Caller holds mu.
static void cache_evict(dda_dhat_cache_s* c, size_t need) {
  while (c->bytes + need > c->max_bytes) {
    auto victim = c->map.end();
    for (auto it = c->map.begin(); it != c->map.end(); ++it)
      if (it->second.ready && it->second.refs == 0 &&
          (victim == c->map.end() || it->second.last_use < victim->second.last_use))
        victim = it;
    if (victim == c->map.end()) return;
    (void)hipFree(victim->second.ptr);
    c->bytes -= victim->second.bytes;
    c->map.erase(victim);
    c->evictions++;
  }
}

// The operator's tensor from the cache: a hit takes a reference; a miss builds
// it on the operator's stream (an FP32 entry from the FP64 one when that is
// cached, else from the Green's function clipped at G/2) while other callers of
// the same key wait.
static thread_local bool t_cache_hit = false;
static thread_local double t_cache_seconds = 0;
bool dda::last_tensor_cache_hit() { return t_cache_hit; }
double dda::tensor_cache_seconds() { return t_cache_seconds; }

static dda_status cache_get(dda_dhat_cache_s* c, dda_op_s* op, void** out) {
  const MvDims& d = op->dims;
  const double kd = op->wave_num * op->grid_unit;
  const auto key = cache_key(d, kd, op->prec);
  const size_t bytes = 6ull * (d.Gx / 2 + 1) * (d.Gy / 2 + 1) * (d.Gz / 2 + 1) *
                       (op->prec == 64 ? sizeof(double2) : sizeof(float2));
  const auto t0 = std::chrono::steady_clock::now();
  auto elapsed = [&]() { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); };
  std::unique_lock<std::mutex> lk(c->mu);
  auto it = c->map.find(key);
  while (it != c->map.end() && !it->second.ready) {
    c->cv.wait(lk);
    it = c->map.find(key);
  }
  t_cache_hit = it != c->map.end();
  if (it != c->map.end()) {
    it->second.refs++;
    it->second.last_use = ++c->clock;
    c->hits++;
    *out = it->second.ptr;
    t_cache_seconds += elapsed();  // waiting for another caller's build counts
    return DDA_OK;
  }
  c->misses++;
  cache_evict(c, bytes);
  // the entry's bytes are reserved now, so that concurrent misses see them
  // when they evict; the pointer is published under the lock once built
  c->bytes += bytes;
  c->map[key].bytes = bytes;
  void* src64 = nullptr;
  if (op->prec == 32) {
    auto f = c->map.find(cache_key(d, kd, 64));
    if (f != c->map.end() && f->second.ready) {
      src64 = f->second.ptr;
      f->second.refs++;  // held while converting
    }
  }
  lk.unlock();
  dda_status rc = DDA_OK;
  void* ptr = nullptr;
  hipError_t e = hipMalloc(&ptr, bytes);
  if (e != hipSuccess) {
    set_error(std::string("tensor cache: hipMalloc failed: ") + hipGetErrorString(e));
    rc = e == hipErrorOutOfMemory ? DDA_OUT_OF_MEMORY : DDA_HIP_ERROR;
    ptr = nullptr;
  } else if (src64) {
    hipLaunchKernelGGL(k_narrow, dim3(1024), dim3(256), 0, op->ctx->stream, (const double2*)src64, (float2*)ptr,
                       (long long)(bytes / sizeof(float2)));
    e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(op->ctx->stream);
    if (e != hipSuccess) {
      set_error(std::string("tensor cache: conversion failed: ") + hipGetErrorString(e));
      rc = DDA_HIP_ERROR;
    }
  } else {
    rc = generate_dhat(op, d.Gx / 2, d.Gy / 2, d.Gz / 2, ptr);
  }
  lk.lock();
  if (src64) c->map.find(cache_key(d, kd, 64))->second.refs--;
  if (rc != DDA_OK) {
    if (ptr) (void)hipFree(ptr);
    c->bytes -= bytes;
    c->map.erase(key);
    c->cv.notify_all();
    return rc;
  }
  dda_dhat_cache_s::Entry& en = c->map[key];
  en.ptr = ptr;
  en.ready = true;
  en.refs = 1;
  en.last_use = ++c->clock;
  c->cv.notify_all();
  *out = ptr;
  t_cache_seconds += elapsed();
  return DDA_OK;
}

static dda_status cache_acquire(dda_dhat_cache_s* c, dda_op_s* op) {
  DDA_TRY(cache_get(c, op, &op->dhat));
  op->dhat_cache = c;
  return DDA_OK;
}

static void cache_release(dda_dhat_cache_s* c, void* ptr) {
  std::lock_guard<std::mutex> g(c->mu);
  for (auto& kv : c->map)
    if (kv.second.ptr == ptr) {
      kv.second.refs--;
      break;
    }
  cache_evict(c, 0);
}

extern "C" dda_status dda_dhat_cache_create(dda_ctx_t ctx, size_t max_bytes, dda_dhat_cache_t* out) {
  if (!ctx || !out) {
    set_error("dda_dhat_cache_create: null argument");
    return DDA_INVALID_ARG;
  }
  dda_dhat_cache_s* c = new dda_dhat_cache_s();
  c->device = ctx->device;
  c->max_bytes = max_bytes;
  *out = c;
  return DDA_OK;
}

extern "C" dda_status dda_dhat_cache_destroy(dda_dhat_cache_t c) {
  if (!c) return DDA_OK;
  {
    std::lock_guard<std::mutex> g(c->mu);
    for (auto& kv : c->map)
      if (kv.second.refs > 0) {
        set_error("dda_dhat_cache_destroy: operators still use cached tensors");
        return DDA_INVALID_ARG;
      }
    for (auto& kv : c->map) (void)hipFree(kv.second.ptr);
  }
  delete c;
  return DDA_OK;
}

extern "C" dda_status dda_dhat_cache_stats(dda_dhat_cache_t c, uint64_t counts[3], size_t* bytes, int* entries) {
  if (!c) {
    set_error("dda_dhat_cache_stats: null cache");
    return DDA_INVALID_ARG;
  }
  std::lock_guard<std::mutex> g(c->mu);
  if (counts) {
    counts[0] = c->hits;
    counts[1] = c->misses;
    counts[2] = c->evictions;
  }
  if (bytes) *bytes = c->bytes;
  if (entries) *entries = (int)c->map.size();
  return DDA_OK;
}

extern "C" dda_status dda_grid_plan(const uint32_t grid[3], int precision, int* static_x, int* static_z) {
  if (!grid || (precision != 32 && precision != 64)) {
    set_error("dda_grid_plan: null grid or precision not 32/64");
    return DDA_INVALID_ARG;
  }
  const size_t csz = precision == 64 ? sizeof(double2) : sizeof(float2);
  if (static_x) *static_x = has_static_plan((int)grid[0]) ? 1 : 0;
  if (static_z) *static_z = static_z_fits((int)grid[0], (int)grid[1], (int)grid[2], csz) ? 1 : 0;
  return DDA_OK;
}

extern "C" uint32_t dda_fft_fit_static(uint32_t x) {
  static const uint32_t sizes[] = {
#define X(N) N,
      DDA_STATIC_SIZES(X)
#undef X
  };
  for (uint32_t n : sizes)
    if (n >= x) return n;
  return dda_fft_fit(x, 0, 1);
}

extern "C" uint32_t dda_fft_fit(uint32_t x, uint32_t override_, int support_non_pot) {
  // DDA/DDAParams.cpp:45-90
  if (override_ != 0) return x <= override_ ? override_ : 0;
  if (x == 0) return 0;
  if (!support_non_pot) {
    uint32_t ret = 1;
    while (x > 1) {
      ret *= 2;
      x = (x + 1) / 2;
    }
    return ret;
  }
  uint32_t ret = x + (x & 1);
  for (;;) {
    uint32_t y = ret;
    for (uint32_t p : {2u, 3u, 5u, 7u})
      while (y % p == 0) y /= p;
    if (y == 1) return ret;
    ret += 2;
  }
}

static void free_op(dda_op_s* op) {
  if (op->split_stream) {
    (void)hipStreamSynchronize(op->split_stream);
    (void)hipStreamDestroy(op->split_stream);
  }
  for (auto e : op->split_ev)
    if (e) (void)hipEventDestroy(e);
  for (auto& r : op->prof_recs) {
    (void)hipEventDestroy(r.b);
    (void)hipEventDestroy(r.e);
  }
  for (auto e : op->prof_pool) (void)hipEventDestroy(e);
  for (size_t m = 1; m < op->pack.dhat.size(); m++) {
    void* p = op->pack.dhat[m];
    if (!p) continue;
    if (op->dhat_cache) {
      cache_release(op->dhat_cache, p);
    } else if (p != op->dhat &&
               std::find(op->pack.dhat.begin() + 1, op->pack.dhat.begin() + m, p) == op->pack.dhat.begin() + m) {
      (void)hipFree(p);  // owned; members of equal wave number share one
    }
  }
  if (op->dhat_cache) {
    cache_release(op->dhat_cache, op->dhat);
    op->dhat = nullptr;
  }
  if (op->pack.dhat_tab) (void)hipFree(op->pack.dhat_tab);
  if (op->pack.member_dev) (void)hipFree(op->pack.member_dev);
  void* bufs[] = {op->row_start, op->row_dip, op->row_meta, op->row_list, op->mat_dev, op->ccs, op->dhat, op->a1,
                  op->a2, op->tw[0], op->tw[1], op->tw[2], op->stage_x, op->stage_y, op->pos_dev, op->obs_scratch};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
  delete op;
}

void apply_cached_tuning(dda_op_s* op);

// The tensor comes from desc->dmatrix, else from `cache`, else from the FP64
// operator `narrow_from` of the same shape (converted), else it is generated.
static dda_status create_op(dda_ctx_t ctx, const dda_op_desc* desc, int precision, dda_dhat_cache_s* cache,
                            const dda_op_s* narrow_from, dda_op_t* out) {
  if (!ctx || !desc || !out || (precision != 32 && precision != 64)) {
    set_error("dda_op_create: null argument or precision not 32/64");
    return DDA_INVALID_ARG;
  }
  DDA_TRY(bind_device(ctx));
  *out = nullptr;
  const dda_op_desc& ds = *desc;
  if (ds.nv == 0 || !ds.pos || ds.nmat == 0 || ds.nmat > 256 || !(ds.grid_unit > 0) || !(ds.wave_num > 0)) {
    set_error("dda_op_create: need nv > 0, positions, 1..256 materials, grid_unit > 0, wave_num > 0");
    return DDA_INVALID_ARG;
  }
  for (int a = 0; a < 3; a++)
    if (ds.box[a] == 0 || ds.box[a] > 65535) {
      set_error("dda_op_create: box extents must be in 1..65535");
      return DDA_INVALID_ARG;
    }
  uint32_t grid[3];
  for (int a = 0; a < 3; a++) {
    grid[a] = ds.grid[a] ? ds.grid[a] : dda_fft_fit(2 * ds.box[a], 0, 1);
    if (grid[a] < 2 * ds.box[a] || grid[a] % 2) {
      set_error("dda_op_create: grid must be even and at least twice the box (DDA/DDAParams.cpp:45-50,131-133)");
      return DDA_INVALID_ARG;
    }
  }
  const uint32_t S = ds.vec_stride ? ds.vec_stride : (ds.nv + 15) / 16 * 16;
  if (S < ds.nv) {
    set_error("dda_op_create: vec_stride < nv");
    return DDA_INVALID_ARG;
  }
  dda_op_s* op = new dda_op_s();
  op->ctx = ctx;
  op->prec = precision;
  op->nmat = ds.nmat;
  op->grid_unit = ds.grid_unit;
  op->wave_num = ds.wave_num;
  MvDims& d = op->dims;
  d.Bx = ds.box[0]; d.By = ds.box[1]; d.Bz = ds.box[2];
  d.Gx = grid[0]; d.Gy = grid[1]; d.Gz = grid[2];
  d.nv = ds.nv;
  d.S = S;
  d.nrows = d.By * d.Bz;
  d.nmat = ds.nmat;
  for (int a = 0; a < 3; a++)
    if (!make_plan(grid[a], &op->plan[a])) {
      set_error("dda_op_create: grid sizes must factor into 2, 3, 5, 7");
      free_op(op);
      return DDA_NOT_SUPPORTED;
    }
  op->pos.assign(ds.pos, ds.pos + 3 * (size_t)ds.nv);
  op->mat.assign(ds.nv, 0);
  if (ds.mat) op->mat.assign(ds.mat, ds.mat + ds.nv);

  // Rows of the (z, y) plane: counting sort of the dipoles by row, stable.
  std::vector<uint32_t> row_start(d.nrows + 1, 0), row_dip(ds.nv), row_meta(ds.nv), row_list;
  for (uint32_t i = 0; i < ds.nv; i++) {
    const uint32_t* p = &op->pos[3 * i];
    if ((int)p[0] >= d.Bx || (int)p[1] >= d.By || (int)p[2] >= d.Bz || op->mat[i] >= ds.nmat) {
      set_error("dda_op_create: dipole position outside the box or material index >= nmat");
      free_op(op);
      return DDA_INVALID_ARG;
    }
    row_start[p[2] * d.By + p[1] + 1]++;
  }
  for (int r = 0; r < d.nrows; r++) {
    if (row_start[r + 1]) row_list.push_back(r);
    row_start[r + 1] += row_start[r];
  }
  {
    std::vector<uint32_t> fill(row_start.begin(), row_start.end() - 1);
    for (uint32_t i = 0; i < ds.nv; i++) {
      const uint32_t* p = &op->pos[3 * i];
      const uint32_t r = p[2] * d.By + p[1];
      const uint32_t k = fill[r]++;
      row_dip[k] = i;
      row_meta[k] = p[0] | ((uint32_t)op->mat[i] << 16);
    }
  }
  op->n_nonempty = (int)row_list.size();
  d.sorted = 1;  // the reference's (z, y, x) dipole order makes the row index the identity
  for (uint32_t k = 0; k < ds.nv && d.sorted; k++) d.sorted = row_dip[k] == k;
  const size_t csz = precision == 64 ? sizeof(double2) : sizeof(float2);
  op->static_z = static_z_fits(d.Gx, d.Gy, d.Gz, csz) && !std::getenv("DDA_GENERIC_FFT");
  op->static_x = has_static_plan(d.Gx) && !std::getenv("DDA_GENERIC_FFT");
  op->split_rhs = std::getenv("DDA_SPLIT") && std::atoi(std::getenv("DDA_SPLIT")) != 0;

  // Launch shapes: X passes hold ~2048 points per work-group; column passes
  // take 128-byte row segments (8 complex doubles, 16 complex floats).
  op->p1_rows = std::max(1, 2048 / (3 * d.Gx));
  op->col_t = precision == 64 ? 8 : 16;
  while (op->col_t > 1 && std::max(d.Gy, d.Gz) * op->col_t > 16 * 256) op->col_t /= 2;
  if (3 * d.Gx > 16 * 256) {
    set_error("dda_op_create: grid x extent too large for the single-tile X pass");
    free_op(op);
    return DDA_NOT_SUPPORTED;
  }

  auto dmalloc = [&](void** p, size_t bytes) -> dda_status {
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
      set_error(std::string("dda_op_create: hipMalloc failed: ") + hipGetErrorString(e));
      return e == hipErrorOutOfMemory ? DDA_OUT_OF_MEMORY : DDA_HIP_ERROR;
    }
    op->device_bytes += bytes;
    return DDA_OK;
  };
  const size_t a1n = 3ull * d.Bz * d.By * d.Gx;
  const size_t a2n = 3ull * d.Bz * d.Gy * d.Gx;
  const size_t dn = 6ull * (d.Gx / 2 + 1) * (d.Gy / 2 + 1) * (d.Gz / 2 + 1);
  dda_status rc = DDA_OK;
  if (rc == DDA_OK) rc = dmalloc((void**)&op->row_start, row_start.size() * 4);
  if (rc == DDA_OK) rc = dmalloc((void**)&op->row_dip, ds.nv * 4ull);
  if (rc == DDA_OK) rc = dmalloc((void**)&op->row_meta, ds.nv * 4ull);
  if (rc == DDA_OK) rc = dmalloc((void**)&op->row_list, std::max<size_t>(1, row_list.size()) * 4);
  if (rc == DDA_OK) rc = dmalloc((void**)&op->mat_dev, ds.nv);
  if (rc == DDA_OK) rc = dmalloc(&op->ccs, 3ull * ds.nmat * csz);
  if (rc == DDA_OK && (ds.dmatrix || !cache)) rc = dmalloc(&op->dhat, dn * csz);
  if (rc == DDA_OK) rc = dmalloc(&op->a1, a1n * csz);
  if (rc == DDA_OK) rc = dmalloc(&op->a2, a2n * csz);
  for (int a = 0; a < 3 && rc == DDA_OK; a++) rc = make_twiddles(grid[a], precision, &op->tw[a], ctx->stream);
  if (rc != DDA_OK) {
    free_op(op);
    return rc;
  }
  hipError_t e = hipSuccess;
  hipStream_t cst = ctx->stream;
  if (e == hipSuccess) e = copy_sync(op->row_start, row_start.data(), row_start.size() * 4, hipMemcpyHostToDevice, cst);
  if (e == hipSuccess) e = copy_sync(op->row_dip, row_dip.data(), ds.nv * 4ull, hipMemcpyHostToDevice, cst);
  if (e == hipSuccess) e = copy_sync(op->row_meta, row_meta.data(), ds.nv * 4ull, hipMemcpyHostToDevice, cst);
  if (e == hipSuccess && !row_list.empty())
    e = copy_sync(op->row_list, row_list.data(), row_list.size() * 4, hipMemcpyHostToDevice, cst);
  if (e == hipSuccess) e = copy_sync(op->mat_dev, op->mat.data(), ds.nv, hipMemcpyHostToDevice, cst);
  if (e == hipSuccess) e = hipMemsetAsync(op->ccs, 0, 3ull * ds.nmat * csz, cst);
  if (e == hipSuccess) e = hipStreamSynchronize(cst);
  if (e != hipSuccess) {
    set_error(std::string("dda_op_create: upload failed: ") + hipGetErrorString(e));
    free_op(op);
    return DDA_HIP_ERROR;
  }

  if (ds.dmatrix) {
    // Reference layout ((j + Gy*(k + Gz*i))*6 + comp), 1/m^3 -> octant, d^3 scaled.
    const double d3 = ds.grid_unit * ds.grid_unit * ds.grid_unit;
    const int hx = d.Gx / 2 + 1, hy = d.Gy / 2 + 1, hz = d.Gz / 2 + 1;
    std::vector<char> h(dn * csz);
    for (int fz = 0; fz < hz; fz++)
      for (int fy = 0; fy < hy; fy++)
        for (int fx = 0; fx < hx; fx++)
          for (int c = 0; c < 6; c++) {
            const size_t src = (((size_t)fy + (size_t)d.Gy * (fz + (size_t)d.Gz * fx)) * 6 + c) * 2;
            const size_t dst = (size_t)c * hx * hy * hz + (((size_t)fz * hy + fy) * hx + fx);
            const double re = ds.dmatrix[src] * d3, im = ds.dmatrix[src + 1] * d3;
            if (precision == 64)
              reinterpret_cast<double2*>(h.data())[dst] = make_double2(re, im);
            else
              reinterpret_cast<float2*>(h.data())[dst] = make_float2((float)re, (float)im);
          }
    e = copy_sync(op->dhat, h.data(), h.size(), hipMemcpyHostToDevice, ctx->stream);
    if (e != hipSuccess) {
      set_error(std::string("dda_op_create: tensor upload failed: ") + hipGetErrorString(e));
      free_op(op);
      return DDA_HIP_ERROR;
    }
  } else if (cache) {
    rc = cache_acquire(cache, op);
    if (rc != DDA_OK) {
      free_op(op);
      return rc;
    }
  } else if (narrow_from && precision == 32 && narrow_from->prec == 64) {
    hipLaunchKernelGGL(k_narrow, dim3(1024), dim3(256), 0, ctx->stream, (const double2*)narrow_from->dhat,
                       (float2*)op->dhat, (long long)dn);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
      set_error(std::string("dda_op_create: tensor conversion failed: ") + hipGetErrorString(e));
      free_op(op);
      return DDA_HIP_ERROR;
    }
  } else {
    rc = generate_dhat(op, d.Bx, d.By, d.Bz, op->dhat);
    if (rc != DDA_OK) {
      free_op(op);
      return rc;
    }
  }
  apply_cached_tuning(op);
  *out = op;
  return DDA_OK;
}

extern "C" dda_status dda_op_create(dda_ctx_t ctx, const dda_op_desc* desc, int precision, dda_op_t* out) {
  return create_op(ctx, desc, precision, nullptr, nullptr, out);
}

extern "C" dda_status dda_op_create_cached(dda_ctx_t ctx, const dda_op_desc* desc, int precision,
                                           dda_dhat_cache_t cache, dda_op_t* out) {
  if (!cache || (desc && desc->dmatrix)) {
    set_error("dda_op_create_cached: null cache, or a host tensor was given");
    return DDA_INVALID_ARG;
  }
  if (ctx && ctx->device != cache->device) {
    set_error("dda_op_create_cached: the cache belongs to another device");
    return DDA_INVALID_ARG;
  }
  return create_op(ctx, desc, precision, cache, nullptr, out);
}


// ------------------------------------------------------------------------------------
// Packs: several particles on one grid in one operator. The operator runs on the
// union of the members' dipole positions; right-hand side r belongs to member
// member[r], whose materials (the void material, polarisability 0, where the
// member has no dipole) and tensor (its wave number) the kernels select. A
// vector of member m is zero off m's dipoles and stays so under the operator
// (void rows are the identity), so a block solve over the pack solves each
// member's system.
// ------------------------------------------------------------------------------------

static dda_status ensure_rhs_capacity(dda_op_s* op, int nrhs);

// Gives `op` (created on the union with member 0's materials and tensor) the
// other members' material maps and tensors. From `src` (an FP64 pack of the
// same union) the tensors are converted, else taken from the cache or generated.
static dda_status pack_extend(dda_op_s* op, const std::vector<uint8_t>& mats, const std::vector<double>& wave_num,
                              const std::vector<std::vector<uint32_t>>& index, const dda_op_s* src) {
  MvDims& d = op->dims;
  const int nm = (int)wave_num.size();
  hipStream_t st = op->ctx->stream;
  op->pack.nmember = nm;
  op->pack.mats = mats;
  op->pack.wave_num = wave_num;
  op->pack.index = index;
  op->pack.dhat.assign(nm, nullptr);
  op->pack.dhat[0] = op->dhat;
  const size_t csz = op->prec == 64 ? sizeof(double2) : sizeof(float2);
  const size_t dn = 6ull * (d.Gx / 2 + 1) * (d.Gy / 2 + 1) * (d.Gz / 2 + 1);
  const double w0 = op->wave_num;
  for (int m = 1; m < nm; m++) {
    if (!op->dhat_cache) {
      int same = -1;
      for (int k = 0; k < m && same < 0; k++)
        if (wave_num[k] == wave_num[m]) same = k;
      if (same >= 0) {
        op->pack.dhat[m] = op->pack.dhat[same];
        continue;
      }
      DDA_TRY_HIP(hipMalloc(&op->pack.dhat[m], dn * csz));
      op->device_bytes += dn * csz;
      if (src) {
        hipLaunchKernelGGL(k_narrow, dim3(1024), dim3(256), 0, st, (const double2*)src->pack.dhat[m],
                           (float2*)op->pack.dhat[m], (long long)dn);
        DDA_TRY_HIP(hipGetLastError());
        DDA_TRY_HIP(hipStreamSynchronize(st));
        continue;
      }
    }
    op->wave_num = wave_num[m];
    const dda_status rc = op->dhat_cache ? cache_get(op->dhat_cache, op, &op->pack.dhat[m])
                                         : generate_dhat(op, d.Bx, d.By, d.Bz, op->pack.dhat[m]);
    op->wave_num = w0;
    DDA_TRY(rc);
  }
  // materials and row metadata of every member ([nm][nv]); the union is in
  // (z, y, x) order, so row k holds dipole k
  std::vector<uint32_t> meta((size_t)nm * d.nv);
  for (int m = 0; m < nm; m++)
    for (int k = 0; k < d.nv; k++)
      meta[(size_t)m * d.nv + k] = op->pos[3 * (size_t)k] | ((uint32_t)mats[(size_t)m * d.nv + k] << 16);
  uint32_t* row_meta = nullptr;
  uint8_t* mat_dev = nullptr;
  DDA_TRY_HIP(hipMalloc(&row_meta, meta.size() * 4));
  hipError_t e = hipMalloc(&mat_dev, mats.size());
  if (e != hipSuccess) (void)hipFree(row_meta);
  DDA_TRY_HIP(e);
  (void)hipFree(op->row_meta);
  (void)hipFree(op->mat_dev);
  op->row_meta = row_meta;
  op->mat_dev = mat_dev;
  op->device_bytes += (size_t)(nm - 1) * d.nv * 5;
  DDA_TRY_HIP(copy_sync(row_meta, meta.data(), meta.size() * 4, hipMemcpyHostToDevice, st));
  DDA_TRY_HIP(copy_sync(mat_dev, mats.data(), mats.size(), hipMemcpyHostToDevice, st));
  DDA_TRY_HIP(hipMalloc(&op->pack.dhat_tab, nm * sizeof(void*)));
  DDA_TRY_HIP(copy_sync(op->pack.dhat_tab, op->pack.dhat.data(), nm * sizeof(void*), hipMemcpyHostToDevice, st));
  op->pack.member_host.assign(op->cap_rhs, 0);
  DDA_TRY_HIP(hipMalloc(&op->pack.member_dev, op->cap_rhs * sizeof(int)));
  DDA_TRY_HIP(copy_sync(op->pack.member_dev, op->pack.member_host.data(), op->cap_rhs * sizeof(int),
                        hipMemcpyHostToDevice, st));
  d.member = op->pack.member_dev;
  d.dhat_tab = op->pack.dhat_tab;
  return DDA_OK;
}

extern "C" dda_status dda_op_create_pack(dda_ctx_t ctx, const dda_op_desc* descs, int nmember, int precision,
                                         dda_dhat_cache_t cache, dda_op_t* out) {
  if (!ctx || !descs || !out || nmember < 1 || nmember > 4096) {
    set_error("dda_op_create_pack: null argument or nmember not in 1..4096");
    return DDA_INVALID_ARG;
  }
  *out = nullptr;
  uint32_t box[3] = {0, 0, 0};
  uint32_t nmat = 0;
  for (int m = 0; m < nmember; m++) {
    const dda_op_desc& ds = descs[m];
    if (ds.dmatrix || !ds.pos || ds.nv == 0 || ds.nmat == 0 || ds.nmat > 255 || ds.grid_unit != descs[0].grid_unit ||
        std::memcmp(ds.grid, descs[0].grid, sizeof ds.grid) != 0 || !(ds.wave_num > 0)) {
      set_error("dda_op_create_pack: members need positions, 1..255 materials, no host tensor, and one grid and "
                "grid unit");
      return DDA_INVALID_ARG;
    }
    for (int a = 0; a < 3; a++) box[a] = std::max(box[a], ds.box[a]);
    nmat = std::max(nmat, ds.nmat + 1);
  }
  for (int a = 0; a < 3; a++)
    if (box[a] == 0 || box[a] > 65535) {
      set_error("dda_op_create_pack: box extents must be in 1..65535");
      return DDA_INVALID_ARG;
    }
  // the union of the positions in (z, y, x) order
  auto key = [&](const uint32_t* p) {
    return ((uint64_t)p[2] * box[1] + p[1]) * box[0] + p[0];
  };
  std::vector<uint64_t> keys;
  for (int m = 0; m < nmember; m++) {
    const dda_op_desc& ds = descs[m];
    for (uint32_t i = 0; i < ds.nv; i++) {
      const uint32_t* p = ds.pos + 3 * (size_t)i;
      if (p[0] >= ds.box[0] || p[1] >= ds.box[1] || p[2] >= ds.box[2] || (ds.mat && ds.mat[i] >= ds.nmat)) {
        set_error("dda_op_create_pack: a dipole outside its box or a material index >= nmat");
        return DDA_INVALID_ARG;
      }
      keys.push_back(key(p));
    }
  }
  std::sort(keys.begin(), keys.end());
  keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
  const size_t nu = keys.size();
  if (nu > 0xffffffffu) {
    set_error("dda_op_create_pack: union too large");
    return DDA_INVALID_ARG;
  }
  std::vector<uint32_t> pos(3 * nu);
  for (size_t k = 0; k < nu; k++) {
    pos[3 * k] = (uint32_t)(keys[k] % box[0]);
    pos[3 * k + 1] = (uint32_t)(keys[k] / box[0] % box[1]);
    pos[3 * k + 2] = (uint32_t)(keys[k] / box[0] / box[1]);
  }
  const uint8_t vd = (uint8_t)(nmat - 1);
  std::vector<uint8_t> mats((size_t)nmember * nu, vd);
  std::vector<std::vector<uint32_t>> index(nmember);
  std::vector<double> wave_num(nmember);
  for (int m = 0; m < nmember; m++) {
    const dda_op_desc& ds = descs[m];
    wave_num[m] = ds.wave_num;
    index[m].resize(ds.nv);
    uint8_t* mm = &mats[(size_t)m * nu];
    for (uint32_t i = 0; i < ds.nv; i++) {
      const uint64_t k = key(ds.pos + 3 * (size_t)i);
      const uint32_t u = (uint32_t)(std::lower_bound(keys.begin(), keys.end(), k) - keys.begin());
      if (mm[u] != vd) {
        set_error("dda_op_create_pack: two dipoles of one member at the same position");
        return DDA_INVALID_ARG;
      }
      index[m][i] = u;
      mm[u] = ds.mat ? ds.mat[i] : 0;
    }
  }
  dda_op_desc ud;
  std::memset(&ud, 0, sizeof ud);
  for (int a = 0; a < 3; a++) {
    ud.box[a] = box[a];
    ud.grid[a] = descs[0].grid[a];
  }
  ud.nv = (uint32_t)nu;
  ud.pos = pos.data();
  ud.mat = mats.data();
  ud.nmat = nmat;
  ud.grid_unit = descs[0].grid_unit;
  ud.wave_num = descs[0].wave_num;
  dda_op_t op = nullptr;
  DDA_TRY(create_op(ctx, &ud, precision, cache, nullptr, &op));
  const dda_status rc = pack_extend(op, mats, wave_num, index, nullptr);
  if (rc != DDA_OK) {
    free_op(op);
    return rc;
  }
  *out = op;
  return DDA_OK;
}

extern "C" dda_status dda_op_pack_info(dda_op_t op, int* nmember, uint32_t* nv_union, uint32_t* nmat_union) {
  if (!op) {
    set_error("dda_op_pack_info: null operator");
    return DDA_INVALID_ARG;
  }
  if (nmember) *nmember = op->pack.nmember;
  if (nv_union) *nv_union = (uint32_t)op->dims.nv;
  if (nmat_union) *nmat_union = op->nmat;
  return DDA_OK;
}

extern "C" dda_status dda_op_pack_map(dda_op_t op, int member, uint32_t* index) {
  if (!op || !index || member < 0 || member >= op->pack.nmember) {
    set_error("dda_op_pack_map: null argument, not a pack, or member out of range");
    return DDA_INVALID_ARG;
  }
  std::copy(op->pack.index[member].begin(), op->pack.index[member].end(), index);
  return DDA_OK;
}

extern "C" dda_status dda_op_set_rhs_members(dda_op_t op, const int* member, int nrhs) {
  if (!op || !member || nrhs < 1 || op->pack.nmember == 0) {
    set_error("dda_op_set_rhs_members: null argument, nrhs < 1, or not a pack");
    return DDA_INVALID_ARG;
  }
  for (int r = 0; r < nrhs; r++)
    if (member[r] < 0 || member[r] >= op->pack.nmember) {
      set_error("dda_op_set_rhs_members: member index out of range");
      return DDA_INVALID_ARG;
    }
  DDA_TRY(bind_device(op->ctx));
  DDA_TRY(ensure_rhs_capacity(op, nrhs));
  std::copy(member, member + nrhs, op->pack.member_host.begin());
  DDA_TRY_HIP(copy_sync(op->pack.member_dev, op->pack.member_host.data(), nrhs * sizeof(int), hipMemcpyHostToDevice,
                        op->ctx->stream));
  return DDA_OK;
}

// The FP32 twin of an FP64 operator (same geometry; tensor converted from the
// FP64 one, or from the same cache).
dda_status dda::create_twin32(const dda_op_s* op, dda_op_t* out) {
  dda_op_desc desc;
  std::memset(&desc, 0, sizeof desc);
  const MvDims& d = op->dims;
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
  DDA_TRY(create_op(op->ctx, &desc, 32, op->dhat_cache, op, out));
  if (op->pack.nmember) {
    dda_status rc = pack_extend(*out, op->pack.mats, op->pack.wave_num, op->pack.index, op);
    if (rc != DDA_OK) {
      free_op(*out);
      *out = nullptr;
      return rc;
    }
  }
  return DDA_OK;
}

extern "C" dda_status dda_op_destroy(dda_op_t op) {
  if (!op) return DDA_OK;
  (void)hipStreamSynchronize(op->ctx->stream);
  free_op(op);
  return DDA_OK;
}

extern "C" dda_status dda_op_info(dda_op_t op, uint32_t grid[3], uint32_t* vec_stride, size_t* device_bytes) {
  if (!op) {
    set_error("dda_op_info: null operator");
    return DDA_INVALID_ARG;
  }
  if (grid) {
    grid[0] = op->dims.Gx;
    grid[1] = op->dims.Gy;
    grid[2] = op->dims.Gz;
  }
  if (vec_stride) *vec_stride = op->dims.S;
  if (device_bytes) *device_bytes = op->device_bytes;
  return DDA_OK;
}

// Grid buffers and couple constants for `nrhs` right-hand sides.
static dda_status ensure_rhs_capacity(dda_op_s* op, int nrhs) {
  if (nrhs <= op->cap_rhs) return DDA_OK;
  DDA_TRY_HIP(hipStreamSynchronize(op->ctx->stream));
  const MvDims& d = op->dims;
  const size_t csz = op->prec == 64 ? sizeof(double2) : sizeof(float2);
  const size_t a1n = 3ull * d.Bz * d.By * d.Gx, a2n = 3ull * d.Bz * d.Gy * d.Gx;
  void *a1 = nullptr, *a2 = nullptr, *ccs = nullptr;
  hipError_t e = hipMalloc(&a1, a1n * csz * nrhs);
  if (e == hipSuccess) e = hipMalloc(&a2, a2n * csz * nrhs);
  if (e == hipSuccess) e = hipMalloc(&ccs, 3ull * d.nmat * csz * nrhs);
  if (e == hipSuccess)
    e = copy_sync(ccs, op->ccs, 3ull * d.nmat * csz * op->cap_rhs, hipMemcpyDeviceToDevice, op->ctx->stream);
  if (e != hipSuccess) {
    if (a1) (void)hipFree(a1);
    if (a2) (void)hipFree(a2);
    if (ccs) (void)hipFree(ccs);
    set_error(std::string("grid buffers for a batch: ") + hipGetErrorString(e));
    return e == hipErrorOutOfMemory ? DDA_OUT_OF_MEMORY : DDA_HIP_ERROR;
  }
  op->device_bytes += (a1n + a2n + 3ull * d.nmat) * csz * (nrhs - op->cap_rhs);
  (void)hipFree(op->a1);
  (void)hipFree(op->a2);
  (void)hipFree(op->ccs);
  op->a1 = a1;
  op->a2 = a2;
  op->ccs = ccs;
  if (op->pack.nmember) {
    int* mem = nullptr;
    DDA_TRY_HIP(hipMalloc(&mem, nrhs * sizeof(int)));
    op->pack.member_host.resize(nrhs, 0);
    DDA_TRY_HIP(copy_sync(mem, op->pack.member_host.data(), nrhs * sizeof(int), hipMemcpyHostToDevice,
                          op->ctx->stream));
    (void)hipFree(op->pack.member_dev);
    op->pack.member_dev = mem;
    op->dims.member = mem;
  }
  op->cap_rhs = nrhs;
  return DDA_OK;
}

// Uploads sqrt(cc) for RHS first .. first+nrhs-1 ([nrhs][nmat][3] complex double,
// m^(3/2)), d^(-3/2) scaled into the operator's precision.
static dda_status upload_ccs(dda_op_s* op, const double* cc_sqrt, int first, int nrhs) {
  const double f = std::pow(op->grid_unit, -1.5);
  const size_t n = 3ull * op->nmat * nrhs;
  const size_t off = 3ull * op->nmat * first;
  // the last material of a pack is the void one: no polarisability
  const size_t nm3 = 3ull * op->nmat;
  auto keep = [&](size_t i) { return op->pack.nmember == 0 || i % nm3 < nm3 - 3 ? 1.0 : 0.0; };
  hipError_t e;
  if (op->prec == 64) {
    std::vector<double2> h(n);
    for (size_t i = 0; i < n; i++) h[i] = make_double2(cc_sqrt[2 * i] * f * keep(i), cc_sqrt[2 * i + 1] * f * keep(i));
    e = hipMemcpyAsync((double2*)op->ccs + off, h.data(), n * sizeof(double2), hipMemcpyHostToDevice, op->ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(op->ctx->stream);
  } else {
    std::vector<float2> h(n);
    for (size_t i = 0; i < n; i++)
      h[i] = make_float2((float)(cc_sqrt[2 * i] * f * keep(i)), (float)(cc_sqrt[2 * i + 1] * f * keep(i)));
    e = hipMemcpyAsync((float2*)op->ccs + off, h.data(), n * sizeof(float2), hipMemcpyHostToDevice, op->ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(op->ctx->stream);
  }
  DDA_TRY_HIP(e);
  return DDA_OK;
}

extern "C" dda_status dda_op_set_cc_sqrt(dda_op_t op, const double* cc_sqrt) {
  if (!op || !cc_sqrt) {
    set_error("dda_op_set_cc_sqrt: null argument");
    return DDA_INVALID_ARG;
  }
  DDA_TRY(bind_device(op->ctx));
  op->cc_sqrt_host.assign(cc_sqrt, cc_sqrt + 6ull * op->nmat);
  DDA_TRY(upload_ccs(op, cc_sqrt, 0, 1));
  op->cc_set = true;
  op->cc_rhs = 1;  // a batch call must set the others again
  return DDA_OK;
}

extern "C" dda_status dda_op_set_cc_sqrt_batch(dda_op_t op, const double* cc_sqrt, int nrhs) {
  if (!op || !cc_sqrt || nrhs < 1) {
    set_error("dda_op_set_cc_sqrt_batch: null argument or nrhs < 1");
    return DDA_INVALID_ARG;
  }
  DDA_TRY(bind_device(op->ctx));
  DDA_TRY(ensure_rhs_capacity(op, nrhs));
  DDA_TRY(upload_ccs(op, cc_sqrt, 0, nrhs));
  op->cc_sqrt_host.assign(cc_sqrt, cc_sqrt + 6ull * op->nmat * nrhs);
  op->cc_set = true;
  op->cc_rhs = nrhs;
  return DDA_OK;
}

extern "C" dda_status dda_op_apply(dda_op_t op, const void* x_dev, void* y_dev, int conj, dda_stream_t stream) {
  return dda_op_apply_batch(op, x_dev, y_dev, 1, conj, stream);
}

extern "C" dda_status dda_op_apply_batch(dda_op_t op, const void* x_dev, void* y_dev, int nrhs, int conj,
                                         dda_stream_t stream) {
  if (!op || !x_dev || !y_dev || x_dev == y_dev || nrhs < 1) {
    set_error("dda_op_apply: null or aliasing vectors, or nrhs < 1");
    return DDA_INVALID_ARG;
  }
  DDA_TRY(bind_device(op->ctx));
  if (!op->cc_set || nrhs > op->cc_rhs) {
    set_error("dda_op_apply: couple constants not set for every right-hand side");
    return DDA_INVALID_ARG;
  }
  hipStream_t st = stream ? (hipStream_t)stream : op->ctx->stream;
  if (op->prec == 64) return launch_matvec<double>(op, x_dev, y_dev, conj, st, nullptr, nrhs, 0);
  return launch_matvec<float>(op, x_dev, y_dev, conj, st, nullptr, nrhs, 0);
}

extern "C" dda_status dda_op_apply_host(dda_op_t op, const double* x, double* y, int conj) {
  if (!op || !x || !y) {
    set_error("dda_op_apply_host: null argument");
    return DDA_INVALID_ARG;
  }
  DDA_TRY(bind_device(op->ctx));
  const size_t n = 3ull * op->dims.S;
  const size_t csz = op->prec == 64 ? sizeof(double2) : sizeof(float2);
  if (!op->stage_x) {
    DDA_TRY_HIP(hipMalloc(&op->stage_x, n * csz));
    DDA_TRY_HIP(hipMalloc(&op->stage_y, n * csz));
    op->device_bytes += 2 * n * csz;
  }
  hipStream_t st = op->ctx->stream;
  if (op->prec == 64) {
    DDA_TRY_HIP(hipMemcpyAsync(op->stage_x, x, n * csz, hipMemcpyHostToDevice, st));
  } else {
    std::vector<float2> h(n);
    for (size_t i = 0; i < n; i++) h[i] = make_float2((float)x[2 * i], (float)x[2 * i + 1]);
    DDA_TRY_HIP(hipMemcpyAsync(op->stage_x, h.data(), n * csz, hipMemcpyHostToDevice, st));
    DDA_TRY_HIP(hipStreamSynchronize(st));
  }
  DDA_TRY(dda_op_apply(op, op->stage_x, op->stage_y, conj, st));
  if (op->prec == 64) {
    DDA_TRY_HIP(hipMemcpyAsync(y, op->stage_y, n * csz, hipMemcpyDeviceToHost, st));
    DDA_TRY_HIP(hipStreamSynchronize(st));
  } else {
    std::vector<float2> h(n);
    DDA_TRY_HIP(hipMemcpyAsync(h.data(), op->stage_y, n * csz, hipMemcpyDeviceToHost, st));
    DDA_TRY_HIP(hipStreamSynchronize(st));
    for (size_t i = 0; i < n; i++) {
      y[2 * i] = h[i].x;
      y[2 * i + 1] = h[i].y;
    }
  }
  return DDA_OK;
}

extern "C" dda_status dda_op_get_dmatrix(dda_op_t op, double* out) {
  if (!op || !out) {
    set_error("dda_op_get_dmatrix: null argument");
    return DDA_INVALID_ARG;
  }
  const MvDims& d = op->dims;
  const int hx = d.Gx / 2 + 1, hy = d.Gy / 2 + 1, hz = d.Gz / 2 + 1;
  const size_t dn = 6ull * hx * hy * hz;
  const size_t csz = op->prec == 64 ? sizeof(double2) : sizeof(float2);
  std::vector<char> h(dn * csz);
  DDA_TRY_HIP(hipStreamSynchronize(op->ctx->stream));
  DDA_TRY_HIP(copy_sync(h.data(), op->dhat, dn * csz, hipMemcpyDeviceToHost, op->ctx->stream));
  const double inv_d3 = 1.0 / (op->grid_unit * op->grid_unit * op->grid_unit);
  for (int i = 0; i < d.Gx; i++)
    for (int k = 0; k < d.Gz; k++)
      for (int j = 0; j < d.Gy; j++) {
        const int fx = i <= d.Gx / 2 ? i : d.Gx - i;
        const int fy = j <= d.Gy / 2 ? j : d.Gy - j;
        const int fz = k <= d.Gz / 2 ? k : d.Gz - k;
        const double sx = i <= d.Gx / 2 ? 1 : -1, sy = j <= d.Gy / 2 ? 1 : -1, sz = k <= d.Gz / 2 ? 1 : -1;
        const double sgn[6] = {1, sx * sy, sx * sz, 1, sy * sz, 1};
        for (int c = 0; c < 6; c++) {
          const size_t src = (size_t)c * hx * hy * hz + (((size_t)fz * hy + fy) * hx + fx);
          double re, im;
          if (op->prec == 64) {
            double2 v = reinterpret_cast<double2*>(h.data())[src];
            re = v.x;
            im = v.y;
          } else {
            float2 v = reinterpret_cast<float2*>(h.data())[src];
            re = v.x;
            im = v.y;
          }
          const size_t dst = (((size_t)j + (size_t)d.Gy * (k + (size_t)d.Gz * i)) * 6 + c) * 2;
          out[dst] = re * sgn[c] * inv_d3;
          out[dst + 1] = im * sgn[c] * inv_d3;
        }
      }
  return DDA_OK;
}

// ------------------------------------------------------------------------------------
// Autotuning of the kernel variants
// ------------------------------------------------------------------------------------

namespace {

typedef std::tuple<int, int, int, int, int, int, int, int> TuneKey;  // grid, box, precision, block
std::mutex g_tune_mu;
std::map<TuneKey, dda_op_s::Tune> g_tune_cache;

TuneKey tune_key(const dda_op_s* op, bool blk) {
  const MvDims& d = op->dims;
  return TuneKey(d.Gx, d.Gy, d.Gz, d.Bx, d.By, d.Bz, op->prec, blk ? 1 : 0);
}

// Mean time of `reps` launches of one pass (after one untimed launch).
template <typename F>
dda_status time_pass(hipStream_t st, int reps, F&& launch, double* us) {
  hipEvent_t e0, e1;
  DDA_TRY_HIP(hipEventCreate(&e0));
  DDA_TRY_HIP(hipEventCreate(&e1));
  dda_status rc = launch();
  if (rc == DDA_OK && hipEventRecord(e0, st) != hipSuccess) rc = DDA_HIP_ERROR;
  for (int i = 0; i < reps && rc == DDA_OK; i++) rc = launch();
  float ms = 0;
  if (rc == DDA_OK && (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                       hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
    rc = DDA_HIP_ERROR;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *us = 1e3 * ms / reps;
  return rc;
}

template <typename R>
dda_status tune_impl(dda_op_s* op, int nrhs, std::string* report) {
  const MvDims& d = op->dims;
  hipStream_t st = op->ctx->stream;
  const size_t csz = sizeof(typename CxT<R>::T);
  // finite data in the grid buffers (their contents are scratch between matvecs)
  DDA_TRY_HIP(hipMemsetAsync(op->a1, 0, 3ull * d.Bz * d.By * d.Gx * csz * nrhs, st));
  DDA_TRY_HIP(hipMemsetAsync(op->a2, 0, 3ull * d.Bz * d.Gy * d.Gx * csz * nrhs, st));
  const int reps = 5;
  char buf[160];
  dda_op_s::Tune best = op->tune;
  void* xv = nullptr;  // dipole vectors for the X passes
  if (op->static_x) {
    DDA_TRY_HIP(hipMalloc(&xv, 2 * 3ull * d.S * csz * nrhs));
    const hipError_t e = hipMemsetAsync(xv, 0, 2 * 3ull * d.S * csz * nrhs, st);
    if (e != hipSuccess) {
      (void)hipFree(xv);
      DDA_TRY_HIP(e);
    }
  }
  for (int pass = 0; pass < 5; pass++) {
    const bool is_static = pass == 1 ? op->static_z : pass >= 3 ? op->static_x : has_static_plan(d.Gy);
    if (!is_static) continue;
    // Z: p3_variants; FP32 X: narrow and wide, each with the other tile too;
    // FP32 Y: with the second and the third tile
    const int nvar = pass == 1 ? p3_variants(d.Gz, op->prec) : op->prec != 32 ? 2 : pass >= 3 ? 4 : 6;
    // two rounds over the variants, the faster of each variant's two timings
    double t[6] = {0, 0, 0, 0, 0, 0};
    for (int it = 0; it < 2 * nvar; it++) {
      const int w = it % nvar;
      double tt = 0;
      dda_status rc;
      if (pass >= 3) {
        op->tune.x_wide[pass - 3] = w;
        char* xb = (char*)xv;
        rc = time_pass(st, reps, [&]() {
          switch (d.Gx) {
#define X(N) \
  case N:    \
    return launch_x_static<R, N>(op, view_of(op, 0), pass == 4, xb, xb + 3ull * d.S * csz * nrhs, 0, st, RhsFlags(), \
                                 nrhs, op->tune);
            DDA_STATIC_SIZES(X)
#undef X
            default:
              return DDA_NOT_SUPPORTED;
          }
        }, &tt);
      } else if (pass == 1) {
        op->tune.p3_wide = w;
        rc = time_pass(st, reps, [&]() {
          switch (d.Gz) {
#define X(N) \
  case N:    \
    return launch_pass3_static<R, N>(op, view_of(op, 0), st, RhsFlags(), nrhs, op->tune);
            DDA_STATIC_SIZES(X)
#undef X
            default:
              return DDA_NOT_SUPPORTED;
          }
        }, &tt);
      } else {
        const bool inv = pass == 2;
        ColParams cp;
        cp.ncols = d.Gx;
        cp.in_len = inv ? d.Gy : d.By;
        cp.out_len = inv ? d.By : d.Gy;
        cp.pstride_in = cp.pstride_out = d.Gx;
        cp.bstride_in = (long long)(inv ? d.Gy : d.By) * d.Gx;
        cp.bstride_out = (long long)(inv ? d.By : d.Gy) * d.Gx;
        const void* in = inv ? op->a2 : op->a1;
        void* out = inv ? op->a1 : op->a2;
        // one launch's worth of column batches (a matvec splits larger blocks)
        const int nb = 3 * d.Bz * std::min(nrhs, 65535 / (3 * d.Bz));
        rc = time_pass(st, reps, [&]() {
          return inv ? launch_col_any<R, true>(in, out, cp, op->plan[1], op->col_t, op->tw[1], nb, st, RhsFlags(), w)
                     : launch_col_any<R, false>(in, out, cp, op->plan[1], op->col_t, op->tw[1], nb, st, RhsFlags(), w);
        }, &tt);
      }
      if (rc != DDA_OK) {
        op->tune = best;
        if (xv) (void)hipFree(xv);
        return rc;
      }
      t[w] = it < nvar ? tt : std::min(t[w], tt);
    }
    // the precision's default unless another variant is faster by more than
    // 2% (closer timings change places between runs)
    // (the X passes default to the narrow plan in both precisions)
    const int def = pass == 1 && d.Gz == 1024 && sizeof(R) == 4 ? 2 : pass < 3 && sizeof(R) == 4 ? 1 : 0;
    int fastest = 0;
    for (int w = 1; w < nvar; w++)
      if (t[w] < t[fastest]) fastest = w;
    const int pick = t[fastest] < 0.98 * t[def] ? fastest : def;
    if (pass == 1)
      best.p3_wide = pick;
    else if (pass >= 3)
      best.x_wide[pass - 3] = pick;
    else
      best.col_wide[pass == 2] = pick;
    static const char* names[5] = {"Y forward", "Z + tensor", "Y inverse", "X forward", "X inverse"};
    static const char* vn1024[4] = {"narrow", "wide", "32x32/16 cols", "32x32/8 cols"};
    static const char* vntile[6] = {"narrow", "wide", "narrow other tile", "wide other tile", "narrow third tile",
                                    "wide third tile"};
    const char* const* vnames = pass == 1 && d.Gz == 1024 ? vn1024 : vntile;
    std::snprintf(buf, sizeof buf, "%s%s: narrow %.1f us, wide %.1f us", report->empty() ? "" : "; ", names[pass],
                  t[0], t[1]);
    *report += buf;
    for (int w = 2; w < nvar; w++) {
      std::snprintf(buf, sizeof buf, ", %s %.1f us", vnames[w], t[w]);
      *report += buf;
    }
    std::snprintf(buf, sizeof buf, " -> %s", vnames[pick]);
    *report += buf;
    op->tune = best;
  }
  if (xv) {
    (void)hipStreamSynchronize(st);
    (void)hipFree(xv);
  }
  if (report->empty()) *report = "no static-plan passes to tune at this grid";
  // a block of right-hand sides: the whole matvec on one stream or split in
  // two halves on two streams (on the grid buffers as input and output)
  if (nrhs >= 2 && !op->prof) {
    void* xy = nullptr;
    DDA_TRY_HIP(hipMalloc(&xy, 2 * 3ull * d.S * csz * nrhs));
    hipError_t e = hipMemsetAsync(xy, 0, 2 * 3ull * d.S * csz * nrhs, st);
    double ts[2] = {0, 0};
    dda_status rc = e == hipSuccess ? DDA_OK : DDA_HIP_ERROR;
    for (int sp = 0; sp < 2 && rc == DDA_OK; sp++) {
      op->tune.split = sp;
      char* b = (char*)xy;
      rc = time_pass(st, reps, [&]() {
        return launch_matvec<R>(op, b, b + 3ull * d.S * csz * nrhs, 0, st, nullptr, nrhs, 0);
      }, &ts[sp]);
    }
    (void)hipStreamSynchronize(st);
    (void)hipFree(xy);
    if (rc != DDA_OK) {
      op->tune = best;
      return rc;
    }
    best.split = ts[1] < ts[0] ? 1 : 0;
    std::snprintf(buf, sizeof buf, "; block of %d: one stream %.1f us, two streams %.1f us -> %s", nrhs, ts[0],
                  ts[1], best.split ? "two" : "one");
    *report += buf;
    op->tune = best;
  }
  return DDA_OK;
}

}  // namespace

// Applies the cached tunings for this operator's shape (one RHS and blocks),
// where they exist.
void apply_cached_tuning(dda_op_s* op) {
  std::lock_guard<std::mutex> g(g_tune_mu);
  auto it = g_tune_cache.find(tune_key(op, false));
  if (it != g_tune_cache.end()) op->tune = it->second;
  it = g_tune_cache.find(tune_key(op, true));
  if (it != g_tune_cache.end()) {
    op->tune_blk = it->second;
    op->tuned_blk = true;
  }
}

extern "C" dda_status dda_op_tune(dda_op_t op, int nrhs, char* report, size_t report_len) {
  if (!op || nrhs < 1 || nrhs > op->cap_rhs) {
    set_error("dda_op_tune: null operator or nrhs outside 1 .. the operator's capacity");
    return DDA_INVALID_ARG;
  }
  DDA_TRY(bind_device(op->ctx));
  std::string rep;
  DDA_TRY_HIP(hipStreamSynchronize(op->ctx->stream));
  // a block of kBlockTuneRhs or more is tuned into tune_blk; while it is timed
  // every launch runs the candidate in op->tune
  const bool blk = nrhs >= kBlockTuneRhs;
  const dda_op_s::Tune keep = op->tune;
  const bool keep_blk = op->tuned_blk;
  if (blk) {
    if (op->tuned_blk) op->tune = op->tune_blk;
    op->tuned_blk = false;
  }
  dda_status rc = op->prec == 64 ? tune_impl<double>(op, nrhs, &rep) : tune_impl<float>(op, nrhs, &rep);
  if (blk) {
    if (rc == DDA_OK) op->tune_blk = op->tune;
    op->tuned_blk = rc == DDA_OK || keep_blk;
    op->tune = keep;
  }
  if (rc != DDA_OK) return rc;
  op->tuned = true;
  {
    std::lock_guard<std::mutex> g(g_tune_mu);
    g_tune_cache[tune_key(op, blk)] = blk ? op->tune_blk : op->tune;
  }
  if (report && report_len) {
    std::strncpy(report, rep.c_str(), report_len - 1);
    report[report_len - 1] = 0;
  }
  return DDA_OK;
}

extern "C" dda_status dda_op_get_variants(dda_op_t op, int variants[6]) {
  if (!op || !variants) {
    set_error("dda_op_get_variants: null argument");
    return DDA_INVALID_ARG;
  }
  const dda_op_s::Tune& t = op->tune;
  const int v[6] = {t.col_wide[0], t.p3_wide, t.col_wide[1], t.x_wide[0], t.x_wide[1], t.split};
  for (int i = 0; i < 6; i++) variants[i] = v[i];
  return DDA_OK;
}

extern "C" dda_status dda_op_set_variants(dda_op_t op, const int variants[6]) {
  if (!op || !variants) {
    set_error("dda_op_set_variants: null argument");
    return DDA_INVALID_ARG;
  }
  for (int i = 0; i < 6; i++)
    if (variants[i] < -1 || variants[i] > (i == 0 || i == 2 ? 5 : i <= 4 ? 3 : 1)) {
      set_error("dda_op_set_variants: variant out of range (-1 .. 5 for the Y passes, -1 .. 3 for the others, "
                "-1 .. 1 for the split)");
      return DDA_INVALID_ARG;
    }
  DDA_TRY(bind_device(op->ctx));
  // launches already queued keep the variants they were made with
  DDA_TRY_HIP(hipStreamSynchronize(op->ctx->stream));
  dda_op_s::Tune& t = op->tune;
  t.col_wide[0] = variants[0];
  t.p3_wide = variants[1];
  t.col_wide[1] = variants[2];
  t.x_wide[0] = variants[3];
  t.x_wide[1] = variants[4];
  t.split = variants[5];
  op->tuned_blk = false;  // pinned variants serve every block size
  return DDA_OK;
}

extern "C" dda_status dda_op_profile(dda_op_t op, int enable) {
  if (!op) {
    set_error("dda_op_profile: null operator");
    return DDA_INVALID_ARG;
  }
  op->prof = enable != 0;
  return DDA_OK;
}

extern "C" dda_status dda_op_profile_read(dda_op_t op, double ms[5], uint64_t counts[5], uint64_t* skipped) {
  if (!op || !ms || !counts) {
    set_error("dda_op_profile_read: null argument");
    return DDA_INVALID_ARG;
  }
  for (int i = 0; i < 5; i++) {
    ms[i] = 0;
    counts[i] = 0;
  }
  // Launches of a device-side solve after some or all of its right-hand sides
  // have finished do less work or none (kernels return at entry for finished
  // ones): launches shorter than 60% of their pass's median are counted in
  // `skipped`, not in the times, so that ms / counts is the full launch.
  uint64_t sk = 0;
  std::vector<float> t(op->prof_recs.size());
  std::vector<float> by_pass[5];
  for (size_t i = 0; i < op->prof_recs.size(); i++) {
    auto& r = op->prof_recs[i];
    DDA_TRY_HIP(hipEventSynchronize(r.e));
    DDA_TRY_HIP(hipEventElapsedTime(&t[i], r.b, r.e));
    by_pass[r.pass].push_back(t[i]);
  }
  float med[5] = {0, 0, 0, 0, 0};
  for (int p = 0; p < 5; p++)
    if (!by_pass[p].empty()) {
      std::nth_element(by_pass[p].begin(), by_pass[p].begin() + by_pass[p].size() / 2, by_pass[p].end());
      med[p] = by_pass[p][by_pass[p].size() / 2];
    }
  for (size_t i = 0; i < op->prof_recs.size(); i++) {
    auto& r = op->prof_recs[i];
    if (t[i] < 0.6f * med[r.pass] || t[i] < 0.002f) {
      sk++;
    } else {
      ms[r.pass] += t[i];
      counts[r.pass]++;
    }
    op->prof_pool.push_back(r.b);
    op->prof_pool.push_back(r.e);
  }
  op->prof_recs.clear();
  if (skipped) *skipped = sk;
  return DDA_OK;
}
