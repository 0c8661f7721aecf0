// Batched pose algebra and Umeyama on device arrays (VRAM_HIP), the HIP counterpart of
// src/rmagine_cuda/include/rmagine/math/memory_math.cuh (multNxN / multNx1 / mult1xN of
// Transforms and of Transform x Vector :54-250, invert :357-361, pack :460-470, setIdentity
// :502, umeyama_transform :580-618) over librmh (rmh_transform_*, rmh_umeyama), and the host
// umeyama_transform of src/rmagine_core/include/rmagine/math/optimization.h:56-75 over
// rmh_umeyama_host.  Launches go to the legacy default stream; the functions returning Memory
// allocate their result, the others write into the caller's views (sizes as in the reference:
// N x N equal sizes, N x 1 / 1 x N one operand of size 1).
#pragma once

#include <rmh.h>

#include <stdexcept>

#include <rmagine/math/types.h>
#include <rmagine/types/MemoryHip.hpp>
#include <rmagine/util/hip/HipException.hpp>

namespace rmagine {

namespace detail {

inline const rmh_transform* tf(const MemoryView<Transform, VRAM_HIP>& T) {
  return reinterpret_cast<const rmh_transform*>(T.raw());
}
inline rmh_transform* tf(MemoryView<Transform, VRAM_HIP>& T) { return reinterpret_cast<rmh_transform*>(T.raw()); }
inline const rmh_vec3* v3(const MemoryView<Vector, VRAM_HIP>& v) { return reinterpret_cast<const rmh_vec3*>(v.raw()); }
inline rmh_vec3* v3(MemoryView<Vector, VRAM_HIP>& v) { return reinterpret_cast<rmh_vec3*>(v.raw()); }

inline void tf_mul(const MemoryView<Transform, VRAM_HIP>& A, const MemoryView<Transform, VRAM_HIP>& B,
                   MemoryView<Transform, VRAM_HIP>& C) {
  const size_t n = C.size();
  if ((A.size() != 1 && A.size() != n) || (B.size() != 1 && B.size() != n))
    throw std::runtime_error("[memory_math] operand sizes do not match the result");
  RM_RMH_CHECK(rmh_transform_mul(tf(A), static_cast<uint32_t>(A.size()), tf(B), static_cast<uint32_t>(B.size()), tf(C),
                                 static_cast<uint32_t>(n), nullptr));
}

inline void tf_points(const MemoryView<Transform, VRAM_HIP>& T, const MemoryView<Vector, VRAM_HIP>& x,
                      MemoryView<Vector, VRAM_HIP>& c) {
  if (x.size() != c.size() || (T.size() != 1 && T.size() != x.size()))
    throw std::runtime_error("[memory_math] operand sizes do not match the result");
  RM_RMH_CHECK(rmh_transform_points(tf(T), static_cast<uint32_t>(T.size()), v3(x), v3(c),
                                    static_cast<uint32_t>(x.size()), nullptr));
}

}  // namespace detail

// ---- Transform x Transform ----------------------------------------------------------------

inline void multNxN(const MemoryView<Transform, VRAM_HIP>& T1, const MemoryView<Transform, VRAM_HIP>& T2,
                    MemoryView<Transform, VRAM_HIP>& Tr) {
  if (T1.size() != Tr.size() || T2.size() != Tr.size()) throw std::runtime_error("[multNxN] sizes differ");
  detail::tf_mul(T1, T2, Tr);
}
inline Memory<Transform, VRAM_HIP> multNxN(const MemoryView<Transform, VRAM_HIP>& T1,
                                           const MemoryView<Transform, VRAM_HIP>& T2) {
  Memory<Transform, VRAM_HIP> Tr(T1.size());
  multNxN(T1, T2, Tr);
  return Tr;
}

inline void multNx1(const MemoryView<Transform, VRAM_HIP>& T1, const MemoryView<Transform, VRAM_HIP>& t2,
                    MemoryView<Transform, VRAM_HIP>& Tr) {
  if (t2.size() != 1 || T1.size() != Tr.size()) throw std::runtime_error("[multNx1] sizes differ");
  detail::tf_mul(T1, t2, Tr);
}
inline Memory<Transform, VRAM_HIP> multNx1(const MemoryView<Transform, VRAM_HIP>& T1,
                                           const MemoryView<Transform, VRAM_HIP>& t2) {
  Memory<Transform, VRAM_HIP> Tr(T1.size());
  multNx1(T1, t2, Tr);
  return Tr;
}

inline void mult1xN(const MemoryView<Transform, VRAM_HIP>& t1, const MemoryView<Transform, VRAM_HIP>& T2,
                    MemoryView<Transform, VRAM_HIP>& Tr) {
  if (t1.size() != 1 || T2.size() != Tr.size()) throw std::runtime_error("[mult1xN] sizes differ");
  detail::tf_mul(t1, T2, Tr);
}
inline Memory<Transform, VRAM_HIP> mult1xN(const MemoryView<Transform, VRAM_HIP>& t1,
                                           const MemoryView<Transform, VRAM_HIP>& T2) {
  Memory<Transform, VRAM_HIP> Tr(T2.size());
  mult1xN(t1, T2, Tr);
  return Tr;
}

// ---- Transform x Vector -------------------------------------------------------------------

inline void multNxN(const MemoryView<Transform, VRAM_HIP>& T, const MemoryView<Vector, VRAM_HIP>& x,
                    MemoryView<Vector, VRAM_HIP>& c) {
  if (T.size() != x.size()) throw std::runtime_error("[multNxN] sizes differ");
  detail::tf_points(T, x, c);
}
inline Memory<Vector, VRAM_HIP> multNxN(const MemoryView<Transform, VRAM_HIP>& T, const MemoryView<Vector, VRAM_HIP>& x) {
  Memory<Vector, VRAM_HIP> c(x.size());
  multNxN(T, x, c);
  return c;
}

inline void mult1xN(const MemoryView<Transform, VRAM_HIP>& t, const MemoryView<Vector, VRAM_HIP>& x,
                    MemoryView<Vector, VRAM_HIP>& c) {
  if (t.size() != 1) throw std::runtime_error("[mult1xN] the transform view must hold one transform");
  detail::tf_points(t, x, c);
}
inline Memory<Vector, VRAM_HIP> mult1xN(const MemoryView<Transform, VRAM_HIP>& t, const MemoryView<Vector, VRAM_HIP>& x) {
  Memory<Vector, VRAM_HIP> c(x.size());
  mult1xN(t, x, c);
  return c;
}

// ---- invert, setIdentity, pack --------------------------------------------------------------

inline void invert(const MemoryView<Transform, VRAM_HIP>& T, MemoryView<Transform, VRAM_HIP>& Ti) {
  if (T.size() != Ti.size()) throw std::runtime_error("[invert] sizes differ");
  RM_RMH_CHECK(rmh_transform_inv(detail::tf(T), detail::tf(Ti), static_cast<uint32_t>(T.size()), nullptr));
}
inline Memory<Transform, VRAM_HIP> invert(const MemoryView<Transform, VRAM_HIP>& T) {
  Memory<Transform, VRAM_HIP> Ti(T.size());
  invert(T, Ti);
  return Ti;
}

inline void setIdentity(MemoryView<Transform, VRAM_HIP>& Ts) {
  RM_RMH_CHECK(rmh_transform_set_identity(detail::tf(Ts), static_cast<uint32_t>(Ts.size()), nullptr));
}

// (R, t) -> Transform; R as rotation matrices (the reference's FP32 matrix -> quaternion
// steps) or quaternions; the stamps of T are kept
inline void pack(const MemoryView<Matrix3x3, VRAM_HIP>& R, const MemoryView<Vector, VRAM_HIP>& t,
                 MemoryView<Transform, VRAM_HIP>& T) {
  if (R.size() != t.size() || R.size() != T.size()) throw std::runtime_error("[pack] sizes differ");
  RM_RMH_CHECK(rmh_transform_pack_matrix(reinterpret_cast<const float*>(R.raw()), detail::v3(t), detail::tf(T),
                                         static_cast<uint32_t>(T.size()), nullptr));
}
inline void pack(const MemoryView<Quaternion, VRAM_HIP>& R, const MemoryView<Vector, VRAM_HIP>& t,
                 MemoryView<Transform, VRAM_HIP>& T) {
  if (R.size() != t.size() || R.size() != T.size()) throw std::runtime_error("[pack] sizes differ");
  RM_RMH_CHECK(rmh_transform_pack_quat(reinterpret_cast<const rmh_quat*>(R.raw()), detail::v3(t), detail::tf(T),
                                       static_cast<uint32_t>(T.size()), nullptr));
}

// ---- Umeyama ------------------------------------------------------------------------------
// FP64 solve per entry (rmh_umeyama); n_meas == 0 gives the identity.

inline void umeyama_transform(MemoryView<Transform, VRAM_HIP>& Ts, const MemoryView<CrossStatisticsD, VRAM_HIP>& stats) {
  if (Ts.size() != stats.size()) throw std::runtime_error("[umeyama_transform] sizes differ");
  RM_RMH_CHECK(rmh_umeyama(reinterpret_cast<const rmh_cross_stats*>(stats.raw()), static_cast<uint32_t>(Ts.size()),
                           detail::tf(Ts), nullptr, nullptr));
}

inline void umeyama_transform(MemoryView<Transform, VRAM_HIP>& Ts, const MemoryView<Vector3, VRAM_HIP>& ds,
                              const MemoryView<Vector3, VRAM_HIP>& ms, const MemoryView<Matrix3x3, VRAM_HIP>& Cs,
                              const MemoryView<unsigned int, VRAM_HIP>& n_meas) {
  const size_t n = Ts.size();
  if (ds.size() != n || ms.size() != n || Cs.size() != n || (n_meas.size() != 0 && n_meas.size() != n))
    throw std::runtime_error("[umeyama_transform] sizes differ");
  Memory<CrossStatisticsD, VRAM_HIP> st(n);
  RM_RMH_CHECK(rmh_cross_stats_pack(detail::v3(ds), detail::v3(ms), reinterpret_cast<const float*>(Cs.raw()),
                                    n_meas.size() ? n_meas.raw() : nullptr, reinterpret_cast<rmh_cross_stats*>(st.raw()),
                                    static_cast<uint32_t>(n), nullptr));
  umeyama_transform(Ts, st);
}

inline void umeyama_transform(MemoryView<Transform, VRAM_HIP>& Ts, const MemoryView<Vector3, VRAM_HIP>& ds,
                              const MemoryView<Vector3, VRAM_HIP>& ms, const MemoryView<Matrix3x3, VRAM_HIP>& Cs) {
  umeyama_transform(Ts, ds, ms, Cs, MemoryView<unsigned int, VRAM_HIP>::Empty());
}

inline Memory<Transform, VRAM_HIP> umeyama_transform(const MemoryView<Vector3, VRAM_HIP>& ds,
                                                     const MemoryView<Vector3, VRAM_HIP>& ms,
                                                     const MemoryView<Matrix3x3, VRAM_HIP>& Cs,
                                                     const MemoryView<unsigned int, VRAM_HIP>& n_meas) {
  Memory<Transform, VRAM_HIP> Ts(ds.size());
  umeyama_transform(Ts, ds, ms, Cs, n_meas);
  return Ts;
}

inline Memory<Transform, VRAM_HIP> umeyama_transform(const MemoryView<Vector3, VRAM_HIP>& ds,
                                                     const MemoryView<Vector3, VRAM_HIP>& ms,
                                                     const MemoryView<Matrix3x3, VRAM_HIP>& Cs) {
  Memory<Transform, VRAM_HIP> Ts(ds.size());
  umeyama_transform(Ts, ds, ms, Cs);
  return Ts;
}

// host: umeyama_transform(const CrossStatistics&) and (d, m, C, n_meas)
// (optimization.h:56-75), the same FP64 solve on the CPU
inline Transform umeyama_transform(const CrossStatisticsD& stats) {
  rmh_transform r;
  RM_RMH_CHECK(rmh_umeyama_host(reinterpret_cast<const rmh_cross_stats*>(&stats), 1, &r, nullptr));
  Transform T;
  T.R = {r.R.x, r.R.y, r.R.z, r.R.w};
  T.t = {r.t.x, r.t.y, r.t.z};
  T.stamp = r.stamp;
  return T;
}

inline Transform umeyama_transform(const Vector3& d, const Vector3& m, const Matrix3x3& C, const unsigned int n_meas = 1) {
  CrossStatisticsD s;
  s.dataset_mean = {d.x, d.y, d.z};
  s.model_mean = {m.x, m.y, m.z};
  for (int k = 0; k < 9; k++) s.covariance.data[k] = C.data[k];
  s.n_meas = n_meas;
  return umeyama_transform(s);
}

inline Transform umeyama_transform(const CrossStatistics& stats) {
  return umeyama_transform(stats.dataset_mean, stats.model_mean, stats.covariance, stats.n_meas);
}

}  // namespace rmagine
