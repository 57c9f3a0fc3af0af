// A program shaped like the reference's timing demos
// galaxySimulationPMTiming and galaxySimulationP3MTiming (source/demos.cpp:
// 779-836, 953-1015), written against the drop-in C++ API: a disk galaxy in
// a 60 x 60 x 30 kpc box on a 128 x 128 x 64 grid, H = 60 / 64, DT = 1,
// G = 4.5e-3, and the bulge as an external field passed as a lambda that
// wraps sphRadDecrField, exactly as the demos pass it.  The disk comes from a
// binary file (the reference samples it with DiskSamplerLinear):
//   int32 n, simLength, p3m;  float32 pos[3n], vel[3n], mass[n]  (original units)
// Outputs in <outdir>: the StateRecorder files (run with diagnostics, as the
// demos do) and final.bin, the particles' positions, velocities and
// accelerations (float32 [3n] each) as getParticles() returns them after run().
#include <chrono>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <stdexcept>
#include <vector>

#include "pmsim/externalFields.h"
#include "pmsim/grid.h"
#include "pmsim/p3mMethod.h"
#include "pmsim/pmMethod.h"
#include "pmsim/stateRecorder.h"

namespace {

class NoHostFFT : public FFTAdapter<float> {
 public:
  std::vector<std::complex<float>>& fft(std::vector<std::complex<float>>&,
                                        std::vector<std::complex<float>>& out) override {
    throw std::logic_error("the GPU force step must not call the host FFT");
    return out;
  }
  std::vector<std::complex<float>>& ifft(std::vector<std::complex<float>>&,
                                         std::vector<std::complex<float>>& out) override {
    throw std::logic_error("the GPU force step must not call the host FFT");
    return out;
  }
};

template <typename T>
void readn(std::ifstream& f, T* p, size_t n) {
  f.read(reinterpret_cast<char*>(p), (std::streamsize)(n * sizeof(T)));
  if (!f) throw std::runtime_error("short input file");
}

}  // namespace

int main(int argc, char** argv) {
  if (argc != 3) {
    std::fprintf(stderr, "usage: %s <input.bin> <outdir>\n", argv[0]);
    return 2;
  }
  try {
    std::ifstream in(argv[1], std::ios::binary);
    int hdr[3];
    readn(in, hdr, 3);
    const int n = hdr[0], simLength = hdr[1], p3mRun = hdr[2];
    std::vector<float> pos(3 * (size_t)n), vel(3 * (size_t)n), masses(n);
    readn(in, pos.data(), pos.size());
    readn(in, vel.data(), vel.size());
    readn(in, masses.data(), masses.size());
    std::vector<Vec3> state(2 * (size_t)n);
    for (int i = 0; i < n; ++i) {
      state[i] = Vec3(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
      state[n + i] = Vec3(vel[3 * i], vel[3 * i + 1], vel[3 * i + 2]);
    }

    Vec3 galaxyCenter = Vec3(30, 30, 15);
    float rb = 3.0f;
    float mb = 60.0f;
    float G = 4.5e-3f;
    auto externalField = [galaxyCenter, rb, mb, G](Vec3 pos) -> Vec3 {
      return sphRadDecrField(pos, galaxyCenter, rb, mb, G);
    };
    auto externalPotential = [galaxyCenter, rb, mb, G](Vec3 pos) -> float {
      return sphRadDecrFieldPotential(pos, galaxyCenter, rb, mb, G);
    };
    auto gridPoints = std::make_tuple(128, 128, 64);
    NoHostFFT fftAdapter;
    Grid grid(gridPoints, fftAdapter);
    auto effectiveBoxSize = std::make_tuple(60.0f, 60.0f, 30.0f);
    float H = std::get<0>(effectiveBoxSize) / (std::get<0>(gridPoints) / 2);
    float DT = 1;
    StateRecorder stateRecorder(n, simLength + 1, argv[2]);
    auto start = std::chrono::high_resolution_clock::now();
    if (!p3mRun) {
      PMMethod pm(state, masses, effectiveBoxSize, externalField, externalPotential, H, DT, G,
                  InterpolationScheme::TSC, FiniteDiffScheme::TWO_POINT, GreensFunction::DISCRETE_LAPLACIAN, 0,
                  grid);
      pm.run(stateRecorder, simLength, true /*diagnostics*/);
      std::ofstream out(std::string(argv[2]) + "/final.bin", std::ios::binary);
      for (int which = 0; which < 3; ++which)
        for (const auto& p : pm.getParticles()) {
          const Vec3& v = which == 0 ? p.position : which == 1 ? p.velocity : p.acceleration;
          out.write(reinterpret_cast<const char*>(&v), sizeof(Vec3));
        }
    } else {
      float a = 3 * H;
      PMMethod pm(state, masses, effectiveBoxSize, externalField, externalPotential, H, DT, G,
                  InterpolationScheme::TSC, FiniteDiffScheme::TWO_POINT, GreensFunction::S1_OPTIMAL, a, grid);
      float re = 0.7f * a;
      float softeningLength = 1.5f;
      P3MMethod p3m(pm, effectiveBoxSize, re, a, H, softeningLength, CloudShape::S1);
      p3m.run(stateRecorder, simLength, true /*diagnostics*/);
      std::ofstream out(std::string(argv[2]) + "/final.bin", std::ios::binary);
      for (int which = 0; which < 3; ++which)
        for (const auto& p : pm.getParticles()) {
          const Vec3& v = which == 0 ? p.position : which == 1 ? p.velocity : p.acceleration;
          out.write(reinterpret_cast<const char*>(&v), sizeof(Vec3));
        }
    }
    auto end = std::chrono::high_resolution_clock::now();
    std::chrono::duration<double> elapsed = end - start;
    std::cout << '\n' << n << ' ' << elapsed.count() << std::endl;
  } catch (const std::exception& e) {
    std::fprintf(stderr, "galaxy_timing: %s\n", e.what());
    return 1;
  }
  return 0;
}
