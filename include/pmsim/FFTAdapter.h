// The FFT seam of the reference (include/FFTAdapter.h:6-15): complex-to-
// complex 3-D transforms of length Nx*Ny*Nz, x fastest, the inverse divided
// by the length.  The GPU force step does not go through it (its FFT is part
// of libpmgpu); Grid still accepts and uses an adapter for the host-side
// Grid::fftDensity / invFftPotential calls.
#pragma once

#include <complex>
#include <vector>

template <typename T>
class FFTAdapter {
 public:
  virtual ~FFTAdapter() {}

  virtual std::vector<std::complex<T>>& fft(std::vector<std::complex<T>>& in,
                                            std::vector<std::complex<T>>& out) = 0;
  virtual std::vector<std::complex<T>>& ifft(std::vector<std::complex<T>>& in,
                                             std::vector<std::complex<T>>& out) = 0;
};
