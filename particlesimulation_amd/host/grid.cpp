// Grid (reference source/grid.cpp) backed by the device grids of libpmgpu.
#include <stdexcept>
#include <string>

#include "pmgpu.h"
#include "pmsim/grid.h"

static void check(int rc, pmgpu_ctx* ctx, const char* what) {
  if (rc != PMGPU_OK) throw std::runtime_error(std::string(what) + ": " + pmgpu_last_error(ctx));
}

Grid::Grid(std::tuple<int, int, int> gridPoints, FFTAdapter<float>& fftAdapter_)
    : gridPointsX(std::get<0>(gridPoints)),
      gridPointsY(std::get<1>(gridPoints)),
      gridPointsZ(std::get<2>(gridPoints)),
      length(gridPointsX * gridPointsY * gridPointsZ),
      fftAdapter(fftAdapter_),
      density(length, 0.0f),
      potential(length, 0.0f),
      field(3 * (size_t)length, 0.0f) {}

std::tuple<int, int, int> Grid::indexTripleFromFlat(int flatIndex) const {
  int x = flatIndex % gridPointsX;
  int y = (flatIndex / gridPointsX) % gridPointsY;
  int z = flatIndex / (gridPointsX * gridPointsY);
  return std::make_tuple(x, y, z);
}

int Grid::getWrappedIndx(int i, int j, int k) const {
  auto mod = [](int a, int b) { return (a % b + b) % b; };
  return mod(i, gridPointsX) + mod(j, gridPointsY) * gridPointsX + mod(k, gridPointsZ) * gridPointsX * gridPointsY;
}

void Grid::attach(pmgpu_ctx* c) {
  ctx = c;
  markDeviceNewer();
}

void Grid::markDeviceNewer() {
  if (!ctx) return;
  densityStale = potentialStale = fieldStale = greensStale = true;
}

void Grid::syncDensity() const {
  if (!densityStale) return;
  check(pmgpu_get_grid(ctx, PMGPU_GRID_DENSITY, density.data()), ctx, "Grid density");
  densityStale = false;
}

void Grid::syncPotential() const {
  if (!potentialStale) return;
  check(pmgpu_get_grid(ctx, PMGPU_GRID_POTENTIAL, potential.data()), ctx, "Grid potential");
  potentialStale = false;
}

void Grid::syncField() const {
  if (!fieldStale) return;
  check(pmgpu_get_grid(ctx, PMGPU_GRID_FIELD, field.data()), ctx, "Grid field");
  fieldStale = false;
}

void Grid::syncGreens() const {
  if (!greensStale) return;
  greensHalf.assign((size_t)(gridPointsX / 2 + 1) * gridPointsY * gridPointsZ, 0.0f);
  check(pmgpu_get_grid(ctx, PMGPU_GRID_GREENS, greensHalf.data()), ctx, "Grid greens");
  greensStale = false;
}

// Host-side writes change the mirrors only; the device force step recomputes
// its grids from the particles.
void Grid::assignDensity(int x, int y, int z, float d) {
  syncDensity();
  density[getIndx(x, y, z)] += d;
}

void Grid::clearDensity() {
  densityStale = false;
  std::fill(density.begin(), density.end(), 0.0f);
}

float Grid::getDensity(int x, int y, int z) const {
  syncDensity();
  return density[getIndx(x, y, z)];
}

void Grid::assignField(int x, int y, int z, Vec3 v) {
  syncField();
  const int i = getIndx(x, y, z);
  field[3 * (size_t)i] = v.x;
  field[3 * (size_t)i + 1] = v.y;
  field[3 * (size_t)i + 2] = v.z;
}

Vec3 Grid::getField(int x, int y, int z) const {
  syncField();
  const size_t i = (size_t)getIndx(x, y, z);
  return Vec3(field[3 * i], field[3 * i + 1], field[3 * i + 2]);
}

const std::vector<std::complex<float>>& Grid::fftDensity() {
  syncDensity();
  densityComplex.assign(density.begin(), density.end());
  densityFourier.assign(length, std::complex<float>(0.0f, 0.0f));
  return fftAdapter.fft(densityComplex, densityFourier);
}

const std::vector<std::complex<float>>& Grid::invFftPotential() {
  potentialFourier.resize(length);
  potentialComplex.assign(length, std::complex<float>(0.0f, 0.0f));
  fftAdapter.ifft(potentialFourier, potentialComplex);
  potentialStale = false;
  for (int i = 0; i < length; ++i) potential[i] = potentialComplex[i].real();
  return potentialComplex;
}

void Grid::setPotentialFourier(int i, int j, int k, std::complex<float> value) {
  potentialFourier.resize(length);
  potentialFourier[getIndx(i, j, k)] = value;
}

std::complex<float> Grid::getDensityFourier(int i, int j, int k) const {
  if (densityFourier.empty()) throw std::logic_error("call fftDensity first");
  return densityFourier[getIndx(i, j, k)];
}

float Grid::getPotential(int i, int j, int k) const {
  syncPotential();
  return potential[getWrappedIndx(i, j, k)];
}

std::complex<float> Grid::getGreensFunction(int i, int j, int k) const {
  if (!greensOverride.empty()) return greensOverride[getIndx(i, j, k)];
  if (!ctx) return std::complex<float>(0.0f, 0.0f);
  syncGreens();
  const int nxh = gridPointsX / 2 + 1;
  if (i >= nxh) {  // G_sym(k) = G_sym(-k)
    i = (gridPointsX - i) % gridPointsX;
    j = (gridPointsY - j) % gridPointsY;
    k = (gridPointsZ - k) % gridPointsZ;
  }
  return std::complex<float>(greensHalf[(size_t)i + (size_t)nxh * (j + (size_t)gridPointsY * k)], 0.0f);
}

void Grid::setGreensFunction(int i, int j, int k, std::complex<float> value) {
  if (greensOverride.empty()) greensOverride.assign(length, std::complex<float>(0.0f, 0.0f));
  greensOverride[getIndx(i, j, k)] = value;
}
