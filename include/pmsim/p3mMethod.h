// P3MMethod of the drop-in host API (reference include/p3mMethod.h:11-63):
// the short-range pair pass runs on the GPU (pmgpu_p3m_short_range).  The
// reference's enableYSorting only changes its CPU summation order and is
// accepted and ignored.  useSRForceTable selects the reference's 500-entry
// table; the accurate table of pmgpu.h is available through
// setShortRangeMode.
#pragma once

#include <tuple>

#include "greensFunctions.h"
#include "pmMethod.h"
#include "simInfo.h"

class P3MMethod {
 public:
  P3MMethod(PMMethod& pmMethod, std::tuple<float, float, float> compBoxSize, float cutoffRadius,
            float particleDiameter, float H, float softeningLength, CloudShape cloudShape,
            bool useSRForceTable = true, bool enableYSorting = true);

  void run(StateRecorder& stateRecorder, const int simLength, bool collectDiagnostics = false,
           bool recordField = false);

  // drop-in extension: 0 direct formula, 1 reference table, 2 accurate table
  void setShortRangeMode(int mode);
  void calculateShortRangeForces();

 private:
  void configure();

  PMMethod& pmMethod;
  SimInfo simInfo;
  std::tuple<float, float, float> compBoxSize;
  float cutoffRadius;      // code units (converted with pmMethod.getH())
  float particleDiameter;  // code units (converted with H)
  float softeningLength;   // code units (converted with H)
  CloudShape cloudShape;
  int mode;
  float H;
};
