// Vec3 of the drop-in host API: the same members and operators as the
// reference's include/vec3.h:5-40 (x, y, z floats; 12 bytes).
#pragma once

#include <cmath>
#include <cstddef>
#include <cstdio>
#include <cstring>

struct Vec3 {
  float x{};
  float y{};
  float z{};

  Vec3() = default;
  Vec3(float x_, float y_, float z_) : x(x_), y(y_), z(z_) {}

  static Vec3 create(float x, float y, float z) { return Vec3(x, y, z); }
  static Vec3 zero() { return Vec3(0.0f, 0.0f, 0.0f); }

  Vec3& operator+=(const Vec3 o) {
    x += o.x;
    y += o.y;
    z += o.z;
    return *this;
  }
  Vec3& operator/=(float s) {
    x /= s;
    y /= s;
    z /= s;
    return *this;
  }
  Vec3 cross(const Vec3 o) const { return Vec3(y * o.z - z * o.y, z * o.x - x * o.z, x * o.y - y * o.x); }
  float getMagnitude() const { return std::sqrt(x * x + y * y + z * z); }
  float getMagnitudeSquared() const { return x * x + y * y + z * z; }

  // "%.6e %.6e %.6e" into vecBuf (the reference's text format of the momentum files)
  char* toString(char* singleBuf, size_t singleBufSize, char* vecBuf, size_t vecBufSize) const {
    (void)singleBuf;
    (void)singleBufSize;
    std::snprintf(vecBuf, vecBufSize, "%.6e %.6e %.6e", x, y, z);
    return vecBuf;
  }
};

inline Vec3 operator+(const Vec3& a, const Vec3& b) { return Vec3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline Vec3 operator-(const Vec3& a, const Vec3& b) { return Vec3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline Vec3 operator*(float s, const Vec3& a) { return Vec3(s * a.x, s * a.y, s * a.z); }
inline Vec3 operator/(const Vec3& a, float s) { return Vec3(a.x / s, a.y / s, a.z / s); }

static_assert(sizeof(Vec3) == 12, "Vec3 must stay three packed floats");
