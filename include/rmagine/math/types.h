// Standalone subset of Rmagine's math types (FP32 PODs with identical layout and the
// reference's operation order), for builds without rmagine_core:
//   Vector3_   src/rmagine_core/include/rmagine/math/types/Vector3.hpp:54-56, Vector3.tcc:80-218
//   Quaternion_ Quaternion.hpp:61-64, Quaternion.tcc:19-75
//   Transform_ Transform.hpp:63-68 (32 bytes with the stamp), Transform.tcc:24-62
//   EulerAngles_ -> Quaternion  EulerAngles.tcc:40-61
//   Quaternion -> EulerAngles_  Quaternion.tcc:206-250
//   AABB_ (min, max)            AABB.hpp:50-53
//   Matrix3x3 (column-major) Matrix.hpp:184
//   CrossStatistics_  CrossStatistics.hpp:50-60, CrossStatistics.tcc:9-58
#pragma once

#include <cmath>
#include <cstdint>

namespace rmagine {

template <typename DataT>
struct EulerAngles_;

template <typename DataT>
struct Vector3_ {
  DataT x, y, z;
  Vector3_ add(const Vector3_& b) const { return {x + b.x, y + b.y, z + b.z}; }
  Vector3_ sub(const Vector3_& b) const { return {x - b.x, y - b.y, z - b.z}; }
  Vector3_ mult(DataT s) const { return {x * s, y * s, z * s}; }
  Vector3_ div(DataT s) const { return {x / s, y / s, z / s}; }
  DataT dot(const Vector3_& b) const { return x * b.x + y * b.y + z * b.z; }
  Vector3_ cross(const Vector3_& b) const { return {y * b.z - z * b.y, z * b.x - x * b.z, x * b.y - y * b.x}; }
  DataT l2normSquared() const { return x * x + y * y + z * z; }
  DataT l2norm() const { return std::sqrt(l2normSquared()); }
  Vector3_ normalize() const { return div(l2norm()); }
  void normalizeInplace() { *this = normalize(); }
  Vector3_ operator+(const Vector3_& b) const { return add(b); }
  Vector3_ operator-(const Vector3_& b) const { return sub(b); }
  Vector3_ operator-() const { return {-x, -y, -z}; }
  Vector3_ operator*(DataT s) const { return mult(s); }
  DataT operator*(const Vector3_& b) const { return dot(b); }
  Vector3_ operator/(DataT s) const { return div(s); }
  Vector3_& operator+=(const Vector3_& b) { return *this = add(b); }
  Vector3_& operator*=(DataT s) { return *this = mult(s); }
};

template <typename DataT>
struct Quaternion_ {
  DataT x, y, z, w;
  static Quaternion_ Identity() { return {0, 0, 0, 1}; }
  void setIdentity() { *this = Identity(); }
  Quaternion_ inv() const { return {-x, -y, -z, w}; }
  Quaternion_ mult(const Quaternion_& q2) const {
    return {w * q2.x + x * q2.w + y * q2.z - z * q2.y,
            w * q2.y - x * q2.z + y * q2.w + z * q2.x,
            w * q2.z + x * q2.y - y * q2.x + z * q2.w,
            w * q2.w - x * q2.x - y * q2.y - z * q2.z};
  }
  Vector3_<DataT> mult(const Vector3_<DataT>& p) const {
    const Quaternion_ P{p.x, p.y, p.z, 0};
    const Quaternion_ PT = mult(P).mult(inv());
    return {PT.x, PT.y, PT.z};
  }
  DataT l2norm() const { return std::sqrt(x * x + y * y + z * z + w * w); }
  void normalizeInplace() {
    const DataT n = l2norm();
    x /= n; y /= n; z /= n; w /= n;
  }
  Quaternion_ operator*(const Quaternion_& q) const { return mult(q); }
  Vector3_<DataT> operator*(const Vector3_<DataT>& v) const { return mult(v); }
  Quaternion_ operator~() const { return inv(); }
  operator EulerAngles_<DataT>() const;
};

template <typename DataT>
struct EulerAngles_ {
  DataT roll, pitch, yaw;
  operator Quaternion_<DataT>() const {
    const DataT cr = std::cos(roll / 2.0f), sr = std::sin(roll / 2.0f);
    const DataT cp = std::cos(pitch / 2.0f), sp = std::sin(pitch / 2.0f);
    const DataT cy = std::cos(yaw / 2.0f), sy = std::sin(yaw / 2.0f);
    return {sr * cp * cy - cr * sp * sy, cr * sp * cy + sr * cp * sy, cr * cp * sy - sr * sp * cy,
            cr * cp * cy + sr * sp * sy};
  }
};

// Quaternion -> roll, pitch, yaw (Quaternion.tcc:206-250); pitch is +-pi/2 where |sin| >= 1
template <typename DataT>
Quaternion_<DataT>::operator EulerAngles_<DataT>() const {
  const DataT sinr_cosp = 2.0 * (w * x + y * z);
  const DataT cosr_cosp = 1.0 - 2.0 * (x * x + y * y);
  const DataT sinp = 2.0 * (w * y - z * x);
  const DataT siny_cosp = 2.0 * (w * z + x * y);
  const DataT cosy_cosp = 1.0 - 2.0 * (y * y + z * z);
  EulerAngles_<DataT> e;
  e.roll = std::atan2(sinr_cosp, cosr_cosp);
  e.pitch = std::fabs(sinp) >= DataT(1) ? std::copysign(DataT(M_PI_2), sinp) : std::asin(sinp);
  e.yaw = std::atan2(siny_cosp, cosy_cosp);
  return e;
}

template <typename DataT>
struct Transform_ {
  Quaternion_<DataT> R;
  Vector3_<DataT> t;
  uint32_t stamp;
  static Transform_ Identity() {
    Transform_ T;
    T.setIdentity();
    return T;
  }
  void setIdentity() {
    R.setIdentity();
    t = {0, 0, 0};
    stamp = 0;
  }
  Transform_ inv() const {
    Transform_ Ti;
    Ti.R = R.inv();
    Ti.t = -(Ti.R * t);
    Ti.stamp = stamp;
    return Ti;
  }
  Transform_ mult(const Transform_& T2) const {
    Transform_ T3;
    T3.t = R * T2.t;
    T3.R = R * T2.R;
    T3.t += t;
    T3.stamp = stamp;
    return T3;
  }
  Vector3_<DataT> mult(const Vector3_<DataT>& v) const { return R * v + t; }
  Transform_ operator*(const Transform_& T2) const { return mult(T2); }
  Vector3_<DataT> operator*(const Vector3_<DataT>& v) const { return mult(v); }
  Transform_ operator~() const { return inv(); }
  Transform_& operator*=(const Transform_& T2) { return *this = mult(T2); }
};

// column-major 3x3 (Matrix.hpp:184, Matrix.tcc:29)
template <typename DataT>
struct Matrix3x3_ {
  DataT data[9];
  DataT& operator()(unsigned r, unsigned c) { return data[c * 3 + r]; }
  const DataT& operator()(unsigned r, unsigned c) const { return data[c * 3 + r]; }
  void setZeros() {
    for (auto& v : data) v = 0;
  }
};

// FP64 statistics layout used by librmh's reductions: dataset mean, model mean,
// covariance (column-major), count; the reference's CrossStatistics_ with DataT=double.
template <typename DataT>
struct CrossStatistics_ {
  Vector3_<DataT> dataset_mean;
  Vector3_<DataT> model_mean;
  Matrix3x3_<DataT> covariance;
  unsigned int n_meas;
};

template <typename DataT>
struct AABB_ {
  Vector3_<DataT> min;
  Vector3_<DataT> max;
};

using Vector3 = Vector3_<float>;
using Vector = Vector3;
using Point = Vector3;
using Quaternion = Quaternion_<float>;
using EulerAngles = EulerAngles_<float>;
using Transform = Transform_<float>;
using Matrix3x3 = Matrix3x3_<float>;
using CrossStatistics = CrossStatistics_<float>;
using CrossStatisticsD = CrossStatistics_<double>;
using AABB = AABB_<float>;

static_assert(sizeof(Transform) == 32, "Transform must be 32 bytes");
static_assert(sizeof(CrossStatistics) == 64, "CrossStatistics_<float> is 64 bytes");
static_assert(sizeof(CrossStatisticsD) == 128, "CrossStatistics_<double> has the layout of rmh_cross_stats");
static_assert(sizeof(AABB) == 24, "AABB is 24 bytes");

struct Face {
  uint32_t v0, v1, v2;
};

}  // namespace rmagine
