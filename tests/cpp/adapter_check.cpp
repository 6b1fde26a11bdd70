// Drives the C++ adapters (dda_amd/host/dda_hip.hpp) the way the reference's
// ddaCl<double> and ddaCl<float> / createResOutput drive GpuMatVec and GpuQmrCs
// (DDA/DDAMain.cpp:550-663, 306-410, 723-727): build the operator, set the
// couple constants, apply the matvec on host vectors in FP64 and FP32, then
// getPolVec with QMR-CS on the double path and on the float path, and the
// profiling run (--profiling-run, DDA/DDAMain.cpp:313-317).
//
// usage: adapter_check <case.bin> <out.bin>
//   case.bin: u32 nv, nmat, box[3], grid[3]; f64 gridUnit, waveNum, eps;
//             u32 pos[nv][3]; u8 mat[nv]; c128 ccSqrt[nmat][3];
//             c128 x[3 S]; c128 einc[3 S]      (S = roundUp(nv, 16))
//   out.bin:  c128 y64[3 S]; c128 y32[3 S]; c128 p[3 S]; c128 p32[3 S];
//             u64 iterations, iterations32
#include <cstdio>
#include <fstream>
#include <iostream>

#include "../../dda_amd/host/dda_hip.hpp"

using namespace DDA::Hip;
typedef std::complex<double> cd;

template <typename T> static void rd(std::ifstream& f, T* p, size_t n) {
  f.read(reinterpret_cast<char*>(p), sizeof(T) * n);
  if (!f) throw std::runtime_error("short case file");
}

int main(int argc, char** argv) {
  if (argc != 3) {
    std::cerr << "usage: adapter_check case.bin out.bin\n";
    return 2;
  }
  try {
    std::ifstream f(argv[1], std::ios::binary);
    uint32_t hdr[8];
    double fd[3];
    rd(f, hdr, 8);
    rd(f, fd, 3);
    Problem p;
    const uint32_t nv = hdr[0];
    p.matCount = hdr[1];
    p.box = {hdr[2], hdr[3], hdr[4]};
    p.grid = {hdr[5], hdr[6], hdr[7]};
    p.gridUnit = fd[0];
    p.waveNum = fd[1];
    p.positions.resize(nv);
    rd(f, p.positions[0].data(), 3 * (size_t)nv);
    p.materials.resize(nv);
    rd(f, p.materials.data(), nv);
    std::vector<std::array<cd, 3>> cc(p.matCount);
    rd(f, cc[0].data(), 3 * (size_t)p.matCount);
    std::vector<cd> x(p.vecSize()), einc(p.vecSize());
    rd(f, x.data(), x.size());
    rd(f, einc.data(), einc.size());

    Context ctx(0);
    GpuMatVec<double> op64(ctx, p);
    GpuMatVec<float> op32(ctx, p);
    MatVecHip<double> mv64(op64);
    MatVecHip<float> mv32(op32);
    mv64.setCoupleConstants(cc);
    mv32.setCoupleConstants(cc);
    std::vector<cd> y64;
    std::vector<std::complex<float>> y32f, xf(x.begin(), x.end());
    mv64.apply(x, y64, false);
    mv32.apply(xf, y32f, false);
    std::vector<cd> y32(y32f.begin(), y32f.end());
    // the float path: GpuQmrCs<float> over GpuMatVec<float> (ddaCl<float>)
    std::vector<cd> p32;
    uint64_t it32 = 0;
    {
      GpuQmrCs<float> s32(op32, 10000);
      s32.setCoupleConstants(cc);
      std::vector<std::complex<float>> ef(einc.begin(), einc.end());
      auto pf = s32.getPolVec(ef, 1e-5f, std::cout);
      p32.assign(pf->begin(), pf->end());
      it32 = s32.lastInfo().iterations;
      s32.profilingRun(std::cout, std::cout);
    }
    GpuQmrCs<double> solver(op64, 10000);
    solver.setCoupleConstants(cc);
    auto pol = solver.getPolVec(einc, fd[2], std::cout);
    uint64_t it = solver.lastInfo().iterations;

    // batch of two fields sharing the couple constants: each bitwise the single solve
    {
      std::vector<cd> e2(einc.size());
      for (size_t i = 0; i < einc.size(); i++) e2[i] = einc[i] * cd(0.5, -0.25);
      auto pols = solver.getPolVecs({e2, einc}, fd[2], std::cout);
      if (pols.size() != 2 || *pols[1] != *pol || solver.lastInfos()[1].iterations != it ||
          solver.lastInfos()[1].status != DDA_OK) {
        std::cerr << "batched solve differs from the single solve\n";
        return 1;
      }
      std::cout << "batch ok" << std::endl;
    }
    std::ofstream o(argv[2], std::ios::binary);
    o.write(reinterpret_cast<const char*>(y64.data()), sizeof(cd) * y64.size());
    o.write(reinterpret_cast<const char*>(y32.data()), sizeof(cd) * y32.size());
    o.write(reinterpret_cast<const char*>(pol->data()), sizeof(cd) * pol->size());
    o.write(reinterpret_cast<const char*>(p32.data()), sizeof(cd) * p32.size());
    o.write(reinterpret_cast<const char*>(&it), sizeof it);
    o.write(reinterpret_cast<const char*>(&it32), sizeof it32);
    solver.profilingRun(std::cout, std::cout);
    // error path: a zero-iteration limit aborts like the reference ("Too many iterations")
    GpuQmrCs<double> limited(op64, 1);
    limited.setCoupleConstants(cc);
    try {
      limited.getPolVec(einc, 1e-14, std::cout);
      std::cerr << "expected an exception\n";
      return 1;
    } catch (const Error& e) {
      if (e.status != DDA_MAX_ITER) return 1;
    }
    std::cout << "adapter_check ok" << std::endl;
  } catch (const std::exception& e) {
    std::cerr << "adapter_check: " << e.what() << std::endl;
    return 1;
  }
  return 0;
}
