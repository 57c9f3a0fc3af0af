// PMMethod of the drop-in host API: the reference's constructor and members
// (include/pmMethod.h:12-76) running the force step on the GPU through the
// C ABI of include/pmgpu.h.  The particles live on the device in code units;
// getParticles() returns a host mirror refreshed from the device, in the
// units the reference would hold at that point (original units before run(),
// code units after it).  Changes made through that reference are detected and
// uploaded before the next device operation.
#pragma once

#include <cstdint>
#include <functional>
#include <string>
#include <tuple>
#include <vector>

#include "grid.h"
#include "particle.h"
#include "pmConfig.h"
#include "stateRecorder.h"
#include "vec3.h"

struct pmgpu_ctx;

class PMMethod {
 public:
  PMMethod(const std::vector<Vec3>& state, const std::vector<float>& masses,
           const std::tuple<float, float, float> effectiveBoxSize,
           const std::function<Vec3(Vec3)> externalField,
           const std::function<float(Vec3)> externalPotential, const float H, const float DT,
           const float G, const InterpolationScheme is, const FiniteDiffScheme fds,
           const GreensFunction gFunc, const float particleDiameter, Grid& grid);
  ~PMMethod();
  PMMethod(const PMMethod&) = delete;
  PMMethod& operator=(const PMMethod&) = delete;

  std::string run(StateRecorder& stateRecorder, const int simLength, bool collectDiagnostics = false,
                  bool recordField = false);

  std::vector<Particle>& getParticles();
  float getH() const { return H; }
  float getDT() const { return DT; }
  float getG() const { return G; }
  const Grid& getGrid() const { return grid; }
  std::function<float(Vec3)> getExternalPotential() const { return externalPotential; }

  void pmMethodStep();
  bool escapedComputationalBox();
  Vec3 totalExternalForceOrigUnits();
  void initGreensFunction();

  // drop-in plumbing (not in the reference)
  // the run loop shared with P3MMethod::run; p3m adds the short-range pass
  std::string runLoop(StateRecorder& stateRecorder, int simLength, bool collectDiagnostics,
                      bool recordField, bool p3m);
  pmgpu_ctx* context() const { return ctx; }
  bool codeUnits() const { return inCodeUnits; }
  void setCodeUnits(bool code);           // the unit system the mirror is in
  void uploadIfModified();                // push host-side edits of the mirror
  void downloadMirror();                  // refresh the mirror from the device
  void markDeviceNewer() { mirrorStale = true; }
  // P3MMethod asks for the short-range part of each force, which the mirror
  // then shows as Particle::shortRangeForce
  void recordShortRange();
  Vec3 externalFieldAt(Vec3 posOriginal) const { return externalField(posOriginal); }

 private:
  void upload();
  uint64_t mirrorHash() const;
  void applyHostField();  // adds a host-evaluated external field to the last force

  Grid& grid;
  float effectiveBoxSizeX, effectiveBoxSizeY, effectiveBoxSizeZ;
  std::vector<Particle> particles;
  int N;
  std::function<Vec3(Vec3)> externalField;
  std::function<float(Vec3)> externalPotential;
  float H, DT, G;
  InterpolationScheme is;
  FiniteDiffScheme fds;
  GreensFunction gFunc;
  float particleDiameter;  // code units, as the reference stores it

  pmgpu_ctx* ctx = nullptr;
  bool inCodeUnits = false;   // units of the host mirror (the device is always in code units)
  bool mirrorStale = false;   // device newer than the mirror
  uint64_t mirrorSeen = 0;    // hash of the mirror when it was last synchronised
  bool greensReady = false;
  bool hostField = false;     // the external field is a host callable (not SphRadDecr / ZeroField)
  bool srRecorded = false;    // the device records the short-range accelerations
};
