// Standalone subset of Rmagine's sensor models (same fields, layout and formulas as
// src/rmagine_core/include/rmagine/types/sensor_models.h:55-503) plus the example models
// of src/rmagine_core/src/types/sensors.cpp:6-137.
#pragma once

#include <cmath>
#include <cstdint>

#include <rmagine/math/types.h>
#include <rmagine/types/Memory.hpp>

namespace rmagine {

constexpr float DEG_TO_RAD_F = static_cast<float>(M_PI / 180.0);

struct Interval {
  float min;
  float max;
  float invalidValue() const { return max + 1.0; }
  bool inside(const float& v) const { return v >= min && v <= max; }
};

struct DiscreteInterval {
  float min;
  float inc;
  uint32_t size;
  float max() const { return min + static_cast<float>(size - 1) * inc; }
  float getValue(uint32_t id) const { return min + static_cast<float>(id) * inc; }
  static float IncFromMinMaxSize(float min, float max, uint32_t size) {
    return (max - min) / static_cast<float>(size - 1);
  }
};

struct SphericalModel {
  static constexpr char name[] = "Sphere";
  DiscreteInterval phi;
  DiscreteInterval theta;
  Interval range;
  uint32_t getWidth() const { return theta.size; }
  uint32_t getHeight() const { return phi.size; }
  uint32_t size() const { return getWidth() * getHeight(); }
  uint32_t getBufferId(uint32_t vid, uint32_t hid) const { return vid * theta.size + hid; }
  Vector getDirection(uint32_t vid, uint32_t hid) const {
    const float p = phi.getValue(vid), t = theta.getValue(hid);
    return {std::cos(p) * std::cos(t), std::cos(p) * std::sin(t), std::sin(p)};
  }
  Vector getOrigin(uint32_t, uint32_t) const { return {0, 0, 0}; }
};
using LiDARModel = SphericalModel;

struct PinholeModel {
  static constexpr char name[] = "Pinhole";
  uint32_t width;
  uint32_t height;
  Interval range;
  float f[2];
  float c[2];
  uint32_t getWidth() const { return width; }
  uint32_t getHeight() const { return height; }
  uint32_t size() const { return width * height; }
  uint32_t getBufferId(uint32_t vid, uint32_t hid) const { return vid * width + hid; }
  Vector getDirectionOptical(uint32_t vid, uint32_t hid) const {
    const float pX = (static_cast<float>(hid) - c[0]) / f[0];
    const float pY = (static_cast<float>(vid) - c[1]) / f[1];
    const Vector d{pX, pY, 1.0f};
    return d.normalize();
  }
  Vector getDirection(uint32_t vid, uint32_t hid) const {
    const Vector d = getDirectionOptical(vid, hid);
    return {d.z, -d.x, -d.y};
  }
  Vector getOrigin(uint32_t, uint32_t) const { return {0, 0, 0}; }
};
using CameraModel = PinholeModel;
using DepthCameraModel = PinholeModel;

template <typename MemT>
struct O1DnModel_ {
  static constexpr char name[] = "O1Dn";
  uint32_t width;
  uint32_t height;
  Interval range;
  Vector orig;
  Memory<Vector, MemT> dirs;
  uint32_t getWidth() const { return width; }
  uint32_t getHeight() const { return height; }
  uint32_t size() const { return width * height; }
  uint32_t getBufferId(uint32_t vid, uint32_t hid) const { return vid * width + hid; }
};
using O1DnModel = O1DnModel_<RAM>;

template <typename MemT>
struct OnDnModel_ {
  static constexpr char name[] = "OnDn";
  uint32_t width;
  uint32_t height;
  Interval range;
  Memory<Vector, MemT> origs;
  Memory<Vector, MemT> dirs;
  uint32_t getWidth() const { return width; }
  uint32_t getHeight() const { return height; }
  uint32_t size() const { return width * height; }
  uint32_t getBufferId(uint32_t vid, uint32_t hid) const { return vid * width + hid; }
};
using OnDnModel = OnDnModel_<RAM>;

static_assert(sizeof(SphericalModel) == 32, "SphericalModel must be 32 bytes");
static_assert(sizeof(PinholeModel) == 32, "PinholeModel must be 32 bytes");

inline SphericalModel vlp16_900() {
  SphericalModel m;
  m.theta.min = -M_PI;
  m.theta.inc = (360.0 / 900.0) * DEG_TO_RAD_F;
  m.theta.size = 900;
  m.phi.min = -15.0 * DEG_TO_RAD_F;
  m.phi.inc = 2.0 * DEG_TO_RAD_F;
  m.phi.size = 16;
  m.range.min = 0.5;
  m.range.max = 130.0;
  return m;
}

inline SphericalModel example_spherical() {
  SphericalModel m;
  m.theta.min = -M_PI;
  m.theta.inc = 0.4 * M_PI / 180.0;
  m.theta.size = 900;
  m.phi.min = -15.0 * M_PI / 180.0;
  m.phi.inc = 2.0 * M_PI / 180.0;
  m.phi.size = 16;
  m.range.min = 0.0;
  m.range.max = 100.0;
  return m;
}

inline PinholeModel example_pinhole() {
  PinholeModel m;
  m.width = 200;
  m.height = 150;
  m.c[0] = 100.0;
  m.c[1] = 75.0;
  m.f[0] = 100.0;
  m.f[1] = 100.0;
  m.range.min = 0.0;
  m.range.max = 100.0;
  return m;
}

inline O1DnModel example_o1dn() {
  O1DnModel m;
  m.orig = {0, 0, 0};
  m.width = 200;
  m.height = 1;
  m.dirs.resize(200);
  const float step = 0.05f;
  for (int i = 0; i < 200; i++) {
    const float y = -static_cast<float>(i - 100) * step;
    const float x = std::cos(y) * 2.0 + 2.0;
    m.dirs[i] = Vector{x, y, -1.0f}.normalize();
  }
  m.range.min = 0.0;
  m.range.max = 100.0;
  return m;
}

inline OnDnModel example_ondn() {
  OnDnModel m;
  m.width = 200;
  m.height = 1;
  m.dirs.resize(200);
  m.origs.resize(200);
  const float step = 0.05f;
  for (int i = 0; i < 200; i++) {
    const float percent = static_cast<float>(i) / 200.0f;
    const float s = -static_cast<float>(i - 100) * step;
    m.origs[i] = {0.0f, std::sin(s) * percent, std::cos(s) * percent};
    m.dirs[i] = {1.0f, 0.0f, 0.0f};
  }
  m.range.min = 0.0;
  m.range.max = 100.0;
  return m;
}

}  // namespace rmagine
