// A program written against the drop-in C++ API the way the reference's
// demos use PMMethod / P3MMethod (source/demos.cpp): build the state vector
// (positions then velocities), a Grid, a PMMethod, optionally a P3MMethod,
// and call run() with a StateRecorder.  The inputs come from a binary file so
// the Python test can rerun the same workload through the C ABI:
//   int32 n, steps, p3m, grid; float32 H, DT, G, diameter, cutoff;
//   float32 pos[3n], vel[3n], mass[n]            (original units)
// Outputs in <outdir>: the StateRecorder files and final.bin with the
// particles' positions, velocities, accelerations and total short-range
// forces (shortRangeForce plus the shortRangeFromNeighbor slots, as
// correctAccelerations sums them) (float32 [3n] each) as getParticles()
// returns them after run().
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <stdexcept>
#include <vector>

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
    int hdr[4];
    float prm[5];
    readn(in, hdr, 4);
    readn(in, prm, 5);
    const int n = hdr[0], steps = hdr[1], p3m = hdr[2], g = hdr[3];
    const float H = prm[0], DT = prm[1], G = prm[2], diameter = prm[3], cutoff = prm[4];
    std::vector<float> pos(3 * (size_t)n), vel(3 * (size_t)n), mass(n);
    readn(in, pos.data(), pos.size());
    readn(in, vel.data(), vel.size());
    readn(in, mass.data(), mass.size());

    std::vector<Vec3> state(2 * (size_t)n);
    for (int i = 0; i < n; ++i) {
      state[i] = Vec3(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
      state[n + i] = Vec3(vel[3 * i], vel[3 * i + 1], vel[3 * i + 2]);
    }
    const float L = g * H;
    NoHostFFT fft;
    Grid grid(std::make_tuple(g, g, g), fft);
    PMMethod pm(state, mass, std::make_tuple(L, L, L), [](Vec3) { return Vec3::zero(); },
                [](Vec3) { return 0.0f; }, H, DT, G, InterpolationScheme::TSC,
                FiniteDiffScheme::TWO_POINT,
                p3m ? GreensFunction::S2_OPTIMAL : GreensFunction::DISCRETE_LAPLACIAN, diameter, grid);
    StateRecorder rec(n, steps + 1, argv[2]);
    if (p3m) {
      P3MMethod sr(pm, std::make_tuple(L, L, L), cutoff, diameter, H, 0.0f, CloudShape::S2, true);
      sr.run(rec, steps, true);
    } else {
      pm.run(rec, steps, true);
    }
    const auto& ps = pm.getParticles();
    std::ofstream out(std::string(argv[2]) + "/final.bin", std::ios::binary);
    for (int which = 0; which < 4; ++which)
      for (const auto& p : ps) {
        Vec3 v = which == 0 ? p.position : which == 1 ? p.velocity : p.acceleration;
        if (which == 3) {
          v = p.shortRangeForce;
          for (const Vec3& s : p.shortRangeFromNeighbor) v += s;
        }
        out.write(reinterpret_cast<const char*>(&v), sizeof(Vec3));
      }
    std::printf("\ndropin_check: %d particles, %d steps, p3m=%d, code units after run: %d\n", n, steps, p3m,
                (int)pm.codeUnits());
  } catch (const std::exception& e) {
    std::fprintf(stderr, "dropin_check: %s\n", e.what());
    return 1;
  }
  return 0;
}
