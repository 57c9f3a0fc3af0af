// StateRecorder (reference source/stateRecorder.cpp:5-137): same files and
// binary layout; records are buffered and written every maxRecords records.
#include <string>

#include "pmsim/stateRecorder.h"
#include "pmsim/unitConversions.h"

StateRecorder::StateRecorder(int particlesCnt, int framesCnt, const std::filesystem::path& outputDirPath,
                             const char* positionsFile, const char* energyFile, const char* momentumFile,
                             const char* expectedMomentumFile, const char* angularMomentumFile,
                             const char* fieldFile, int maxRecords)
    : particlesCnt_(particlesCnt), framesCnt_(framesCnt), maxRecords_(maxRecords), dir_(outputDirPath) {
  auto open_bin = [&](std::ofstream& f, const char* name) {
    f.open(dir_ / name, std::ios::binary | std::ios::trunc);
    f.write(reinterpret_cast<const char*>(&particlesCnt_), sizeof(int));
    f.write(reinterpret_cast<const char*>(&framesCnt_), sizeof(int));
  };
  open_bin(positions_, positionsFile);
  open_bin(field_, fieldFile);
  energy_.open(dir_ / energyFile, std::ios::trunc);
  momentum_.open(dir_ / momentumFile, std::ios::trunc);
  expectedMomentum_.open(dir_ / expectedMomentumFile, std::ios::trunc);
  angularMomentum_.open(dir_ / angularMomentumFile, std::ios::trunc);
}

StateRecorder::~StateRecorder() {
  positions_.close();
  field_.close();
  energy_.close();
  momentum_.close();
  expectedMomentum_.close();
  angularMomentum_.close();
}

void StateRecorder::binRecord(std::ofstream& of, std::vector<Vec3>& buf, const Vec3& v) {
  buf.push_back(v);
  if ((int)buf.size() >= maxRecords_) {
    of.write(reinterpret_cast<const char*>(buf.data()), (std::streamsize)(buf.size() * sizeof(Vec3)));
    buf.clear();
  }
}

void StateRecorder::textRecord(std::ofstream& of, std::string& buf, int& count, const std::string& line) {
  buf += line;
  buf += '\n';
  if (++count >= maxRecords_) {
    of << buf;
    buf.clear();
    count = 0;
  }
}

void StateRecorder::recordPositions(std::vector<Vec3>::iterator begin, std::vector<Vec3>::iterator end) {
  for (auto it = begin; it != end; ++it) binRecord(positions_, positionsBuf_, *it);
}

void StateRecorder::recordPositions(const std::vector<Particle>& particles) {
  for (const auto& p : particles) binRecord(positions_, positionsBuf_, p.position);
}

void StateRecorder::recordEnergy(float pe, float ke) {
  textRecord(energy_, energyStr_, energyCnt_, std::to_string(pe) + ' ' + std::to_string(ke));
}

static std::string vecText(const Vec3& v) {
  char single[64], buf[128];
  return v.toString(single, sizeof(single), buf, sizeof(buf));
}

void StateRecorder::recordTotalMomentum(Vec3 m) { textRecord(momentum_, momentumStr_, momentumCnt_, vecText(m)); }

void StateRecorder::recordExpectedMomentum(Vec3 m) {
  textRecord(expectedMomentum_, expectedMomentumStr_, expectedMomentumCnt_, vecText(m));
}

void StateRecorder::recordTotalAngularMomentum(Vec3 m) {
  textRecord(angularMomentum_, angularMomentumStr_, angularMomentumCnt_, vecText(m));
}

void StateRecorder::recordField(const std::vector<Particle>& particles, float H, float DT) {
  for (const auto& p : particles) binRecord(field_, fieldBuf_, accelerationToOriginalUnits(p.acceleration, H, DT));
}

std::string StateRecorder::flush() {
  positions_.write(reinterpret_cast<const char*>(positionsBuf_.data()),
                   (std::streamsize)(positionsBuf_.size() * sizeof(Vec3)));
  positionsBuf_.clear();
  field_.write(reinterpret_cast<const char*>(fieldBuf_.data()), (std::streamsize)(fieldBuf_.size() * sizeof(Vec3)));
  fieldBuf_.clear();
  energy_ << energyStr_;
  momentum_ << momentumStr_;
  expectedMomentum_ << expectedMomentumStr_;
  angularMomentum_ << angularMomentumStr_;
  energyStr_.clear();
  momentumStr_.clear();
  expectedMomentumStr_.clear();
  angularMomentumStr_.clear();
  positions_.flush();
  field_.flush();
  energy_.flush();
  momentum_.flush();
  expectedMomentum_.flush();
  angularMomentum_.flush();
  return dir_.string();
}
