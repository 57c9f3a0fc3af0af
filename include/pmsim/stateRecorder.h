// StateRecorder with the reference's constructor, methods and file formats
// (include/stateRecorder.h, source/stateRecorder.cpp): binary positions /
// field files (int32 particle count, int32 frame count, then float32 x y z
// per particle per frame, flushed every maxRecords records) and text energy
// and momentum files.
#pragma once

#include <filesystem>
#include <fstream>
#include <string>
#include <vector>

#include "particle.h"
#include "vec3.h"

class StateRecorder {
 public:
  StateRecorder(int particlesCnt, int framesCnt, const std::filesystem::path& outputDirPath,
                const char* positionsFile = "positions.dat", const char* energyFile = "energy.txt",
                const char* momentumFile = "momentum.txt",
                const char* expectedMomentumFile = "expected_momentum.txt",
                const char* angularMomentumFile = "angular_momentum.txt",
                const char* fieldFile = "field.dat", int maxRecords = 500);
  ~StateRecorder();

  void recordPositions(std::vector<Vec3>::iterator begin, std::vector<Vec3>::iterator end);
  void recordPositions(const std::vector<Particle>& particles);
  void recordEnergy(float pe, float ke);
  void recordTotalMomentum(Vec3 momentum);
  void recordExpectedMomentum(Vec3 expectedMomentum);
  void recordTotalAngularMomentum(Vec3 angularMomentum);
  void recordField(const std::vector<Particle>& particles, float H, float DT);
  std::string flush();

 private:
  void textRecord(std::ofstream& of, std::string& buf, int& count, const std::string& line);
  void binRecord(std::ofstream& of, std::vector<Vec3>& buf, const Vec3& v);

  int particlesCnt_;
  int framesCnt_;
  int maxRecords_;
  std::filesystem::path dir_;
  std::ofstream positions_, field_, energy_, momentum_, expectedMomentum_, angularMomentum_;
  std::vector<Vec3> positionsBuf_, fieldBuf_;
  std::string energyStr_, momentumStr_, expectedMomentumStr_, angularMomentumStr_;
  int energyCnt_ = 0, momentumCnt_ = 0, expectedMomentumCnt_ = 0, angularMomentumCnt_ = 0;
};
