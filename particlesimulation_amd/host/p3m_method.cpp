// P3MMethod on the GPU (reference source/p3mMethod.cpp).
#include <stdexcept>
#include <string>

#include "pmgpu.h"
#include "pmsim/p3mMethod.h"
#include "pmsim/unitConversions.h"

P3MMethod::P3MMethod(PMMethod& pm, std::tuple<float, float, float> compBoxSize_, float cutoffRadius_,
                     float particleDiameter_, float H_, float softeningLength_, CloudShape cloudShape_,
                     bool useSRForceTable, bool enableYSorting)
    : pmMethod(pm),
      compBoxSize(compBoxSize_),
      // the reference converts the cutoff with the PMMethod's H and the other
      // lengths with its own argument (source/p3mMethod.cpp:34-37)
      cutoffRadius(lengthToCodeUnits(cutoffRadius_, pm.getH())),
      particleDiameter(lengthToCodeUnits(particleDiameter_, H_)),
      softeningLength(lengthToCodeUnits(softeningLength_, H_)),
      cloudShape(cloudShape_),
      mode(useSRForceTable ? 1 : 0),
      H(H_) {
  (void)enableYSorting;
  configure();
}

void P3MMethod::configure() {
  pmgpu_p3m_params q{};
  q.cutoff = cutoffRadius;
  q.particle_diameter = particleDiameter;
  q.softening = softeningLength;
  q.cloud_shape = cloudShape == CloudShape::S1 ? PMGPU_S1 : PMGPU_S2;
  q.use_table = mode;
  q.box[0] = std::get<0>(compBoxSize) / H;
  q.box[1] = std::get<1>(compBoxSize) / H;
  q.box[2] = std::get<2>(compBoxSize) / H;
  const int rc = pmgpu_p3m_configure(pmMethod.context(), &q);
  if (rc == PMGPU_ERR_INVALID_ARGUMENT) throw std::invalid_argument(pmgpu_last_error(pmMethod.context()));
  if (rc != PMGPU_OK) throw std::runtime_error(pmgpu_last_error(pmMethod.context()));
  pmMethod.recordShortRange();
}

void P3MMethod::setShortRangeMode(int m) {
  if (m < 0 || m > 2) throw std::invalid_argument("short-range mode must be 0, 1 or 2");
  mode = m;
  configure();
}

// calculateShortRangeForces + correctAccelerations (p3mMethod.cpp:50-57,
// 168-192) on the current device state: adds the short-range correction to
// the accelerations of the last PM step.
void P3MMethod::calculateShortRangeForces() {
  pmMethod.uploadIfModified();
  const int rc = pmgpu_p3m_short_range(pmMethod.context());
  if (rc != PMGPU_OK) throw std::runtime_error(pmgpu_last_error(pmMethod.context()));
  pmMethod.markDeviceNewer();
}

void P3MMethod::run(StateRecorder& stateRecorder, const int simLength, bool collectDiagnostics,
                    bool recordField) {
  pmMethod.runLoop(stateRecorder, simLength, collectDiagnostics, recordField, true);
}
