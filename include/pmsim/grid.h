// Grid of the drop-in host API (reference include/grid.h:10-70): flat index
// x + y*Nx + z*Nx*Ny.  Once a PMMethod is attached the density, potential and
// field live on the GPU; the accessors read host mirrors refreshed from the
// device after each force step.  The Green's function accessor returns the
// inversion-symmetrised influence function the GPU multiplies with
// (G_sym(k) = (G(k) + G(-k))/2; it yields the same real potential as the
// reference's unsymmetrised G, source/grid.cpp:66-68).
#pragma once

#include <complex>
#include <tuple>
#include <vector>

#include "FFTAdapter.h"
#include "vec3.h"

struct pmgpu_ctx;

class Grid {
 public:
  Grid(std::tuple<int, int, int> gridPoints, FFTAdapter<float>& fftAdapter);

  std::tuple<int, int, int> indexTripleFromFlat(int flatIndex) const;
  void assignDensity(int x, int y, int z, float density);
  void clearDensity();
  float getDensity(int x, int y, int z) const;
  void assignField(int x, int y, int z, Vec3 fieldVal);
  Vec3 getField(int x, int y, int z) const;
  int getLength() const { return length; }
  std::tuple<int, int, int> getGridPoints() const {
    return std::make_tuple(gridPointsX, gridPointsY, gridPointsZ);
  }
  const std::vector<std::complex<float>>& fftDensity();
  const std::vector<std::complex<float>>& invFftPotential();
  void setPotentialFourier(int i, int j, int k, std::complex<float> value);
  std::complex<float> getDensityFourier(int i, int j, int k) const;
  float getPotential(int i, int j, int k) const;
  std::complex<float> getGreensFunction(int i, int j, int k) const;
  void setGreensFunction(int i, int j, int k, std::complex<float> value);

  // drop-in plumbing (not in the reference): the device context whose grids
  // back this object, and invalidation after a device step
  void attach(pmgpu_ctx* ctx);
  void markDeviceNewer();

 private:
  int getIndx(int i, int j, int k) const { return i + j * gridPointsX + k * gridPointsX * gridPointsY; }
  int getWrappedIndx(int i, int j, int k) const;
  void syncDensity() const;
  void syncPotential() const;
  void syncField() const;
  void syncGreens() const;

  int gridPointsX, gridPointsY, gridPointsZ, length;
  FFTAdapter<float>& fftAdapter;
  pmgpu_ctx* ctx = nullptr;
  mutable bool densityStale = false, potentialStale = false, fieldStale = false, greensStale = false;
  mutable std::vector<float> density, potential, field, greensHalf;
  std::vector<std::complex<float>> densityComplex, densityFourier, potentialComplex, potentialFourier;
  std::vector<std::complex<float>> greensOverride;
};
