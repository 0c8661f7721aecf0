// Correspondence statistics on device point clouds (VRAM_HIP), the HIP counterpart of
// src/rmagine_cuda/include/rmagine/math/statistics.cuh:76-158 (statistics_p2p,
// statistics_p2l, statistics_objectwise_p2l) over librmh's FP64 reductions (rmh_statistics,
// rmh_statistics_objectwise_p2l).
//
// As in the reference, a statistics view of size k > 1 receives k partial results: output b
// covers the elements [b * chunk, min(N, (b + 1) * chunk)) with chunk = ceil(N / (512 k)) *
// 512 (the block partition of statistics.cu:197-213), so that merging the k partials gives
// the total.  The sums are FP64 on the device; the CrossStatistics (FP32) overloads round the
// FP64 results once, the CrossStatisticsD overloads return them as they are.  Launches go to
// the legacy default stream, which orders them after the simulators' blocking map streams.
#pragma once

#include <rmh.h>

#include <rmagine/math/types.h>
#include <rmagine/types/MemoryHip.hpp>
#include <rmagine/types/PointCloud.hpp>
#include <rmagine/types/UmeyamaReductionConstraints.hpp>
#include <rmagine/util/hip/HipException.hpp>

namespace rmagine {

namespace detail {

inline rmh_cloud to_rmh_cloud(const PointCloudView_<VRAM_HIP>& c, size_t offset, uint64_t stride) {
  rmh_cloud r{};
  r.points = c.points.size() ? reinterpret_cast<const rmh_vec3*>(c.points.raw()) + offset : nullptr;
  r.mask = c.mask.size() ? c.mask.raw() + offset : nullptr;
  r.normals = c.normals.size() ? reinterpret_cast<const rmh_vec3*>(c.normals.raw()) + offset : nullptr;
  r.ids = c.ids.size() ? c.ids.raw() + offset : nullptr;
  r.segment_stride = stride;
  return r;
}

inline rmh_umeyama_constraints to_rmh(const UmeyamaReductionConstraints& p) {
  return {p.max_dist, p.dataset_id, p.model_id};
}

inline CrossStatistics to_f32(const CrossStatisticsD& s) {
  CrossStatistics r;
  r.dataset_mean = {static_cast<float>(s.dataset_mean.x), static_cast<float>(s.dataset_mean.y),
                    static_cast<float>(s.dataset_mean.z)};
  r.model_mean = {static_cast<float>(s.model_mean.x), static_cast<float>(s.model_mean.y),
                  static_cast<float>(s.model_mean.z)};
  for (int k = 0; k < 9; k++) r.covariance.data[k] = static_cast<float>(s.covariance.data[k]);
  r.n_meas = s.n_meas;
  return r;
}

// k = stats.size() partial statistics over the reference's block partition
inline void statistics_partitioned(int kind, const Transform& pre_transform, const PointCloudView_<VRAM_HIP>& dataset,
                                   const PointCloudView_<VRAM_HIP>& model, const UmeyamaReductionConstraints& params,
                                   MemoryView<CrossStatisticsD, VRAM_HIP>& stats) {
  const size_t k = stats.size();
  if (k == 0) return;
  const size_t n = dataset.points.size();
  if (model.points.size() < n) throw std::runtime_error("[statistics] model holds fewer points than the dataset");
  Memory<Transform, VRAM_HIP> pre(1);
  {
    Transform t = pre_transform;
    pre = make_view(t);
  }
  const size_t chunk = (n + 512 * k - 1) / (512 * k) * 512;
  const size_t full = chunk ? n / chunk : 0;
  const size_t rest = n - full * chunk;
  const rmh_umeyama_constraints prm = to_rmh(params);
  rmh_cross_stats* out = reinterpret_cast<rmh_cross_stats*>(stats.raw());
  if (full > 0) {
    const rmh_cloud d = to_rmh_cloud(dataset, 0, chunk), m = to_rmh_cloud(model, 0, chunk);
    RM_RMH_CHECK(rmh_statistics(kind, reinterpret_cast<const rmh_transform*>(pre.raw()), 1, &d, &m, chunk,
                                static_cast<uint32_t>(full), &prm, out, nullptr, 0, nullptr));
  }
  size_t done = full;
  if (rest > 0) {
    const rmh_cloud d = to_rmh_cloud(dataset, full * chunk, rest), m = to_rmh_cloud(model, full * chunk, rest);
    RM_RMH_CHECK(rmh_statistics(kind, reinterpret_cast<const rmh_transform*>(pre.raw()), 1, &d, &m, rest, 1, &prm,
                                out + done, nullptr, 0, nullptr));
    done++;
  }
  if (done < k) {  // outputs past the end of the cloud: empty statistics
    const rmh_cloud d = to_rmh_cloud(dataset, 0, 0), m = to_rmh_cloud(model, 0, 0);
    RM_RMH_CHECK(rmh_statistics(kind, reinterpret_cast<const rmh_transform*>(pre.raw()), 1, &d, &m, 0,
                                static_cast<uint32_t>(k - done), &prm, out + done, nullptr, 0, nullptr));
  }
}

inline void statistics_partitioned_f32(int kind, const Transform& pre, const PointCloudView_<VRAM_HIP>& dataset,
                                       const PointCloudView_<VRAM_HIP>& model, const UmeyamaReductionConstraints& params,
                                       MemoryView<CrossStatistics, VRAM_HIP>& stats) {
  Memory<CrossStatisticsD, VRAM_HIP> s64(stats.size());
  statistics_partitioned(kind, pre, dataset, model, params, s64);
  RM_RMH_CHECK(rmh_cross_stats_to_f32(reinterpret_cast<const rmh_cross_stats*>(s64.raw()), stats.raw(),
                                      static_cast<uint32_t>(stats.size()), nullptr));
}

inline CrossStatisticsD statistics_total(int kind, const Transform& pre, const PointCloudView_<VRAM_HIP>& dataset,
                                         const PointCloudView_<VRAM_HIP>& model,
                                         const UmeyamaReductionConstraints& params) {
  Memory<CrossStatisticsD, VRAM_HIP> s64(1);
  statistics_partitioned(kind, pre, dataset, model, params, s64);
  Memory<CrossStatisticsD, RAM> h;
  h = s64;
  return h[0];
}

}  // namespace detail

// ---- point to point (statistics.cuh:76-104) -----------------------------------------------

inline void statistics_p2p(const Transform& pre_transform, const PointCloudView_<VRAM_HIP>& dataset,
                           const PointCloudView_<VRAM_HIP>& model, const UmeyamaReductionConstraints params,
                           MemoryView<CrossStatistics, VRAM_HIP>& stats) {
  detail::statistics_partitioned_f32(RMH_STATS_P2P, pre_transform, dataset, model, params, stats);
}

inline void statistics_p2p(const Transform& pre_transform, const PointCloudView_<VRAM_HIP>& dataset,
                           const PointCloudView_<VRAM_HIP>& model, const UmeyamaReductionConstraints params,
                           MemoryView<CrossStatisticsD, VRAM_HIP>& stats) {
  detail::statistics_partitioned(RMH_STATS_P2P, pre_transform, dataset, model, params, stats);
}

inline void statistics_p2p(const Transform& pre_transform, const PointCloudView_<VRAM_HIP>& dataset,
                           const PointCloudView_<VRAM_HIP>& model, const UmeyamaReductionConstraints params,
                           CrossStatistics& stats) {
  stats = detail::to_f32(detail::statistics_total(RMH_STATS_P2P, pre_transform, dataset, model, params));
}

inline CrossStatistics statistics_p2p(const Transform& pre_transform, const PointCloudView_<VRAM_HIP>& dataset,
                                      const PointCloudView_<VRAM_HIP>& model, const UmeyamaReductionConstraints params) {
  return detail::to_f32(detail::statistics_total(RMH_STATS_P2P, pre_transform, dataset, model, params));
}

// ---- point to plane (statistics.cuh:106-137) ----------------------------------------------

inline void statistics_p2l(const Transform& pre_transform, const PointCloudView_<VRAM_HIP>& dataset,
                           const PointCloudView_<VRAM_HIP>& model, const UmeyamaReductionConstraints params,
                           MemoryView<CrossStatistics, VRAM_HIP>& stats) {
  detail::statistics_partitioned_f32(RMH_STATS_P2L, pre_transform, dataset, model, params, stats);
}

inline void statistics_p2l(const Transform& pre_transform, const PointCloudView_<VRAM_HIP>& dataset,
                           const PointCloudView_<VRAM_HIP>& model, const UmeyamaReductionConstraints params,
                           MemoryView<CrossStatisticsD, VRAM_HIP>& stats) {
  detail::statistics_partitioned(RMH_STATS_P2L, pre_transform, dataset, model, params, stats);
}

inline void statistics_p2l(const Transform& pre_transform, const PointCloudView_<VRAM_HIP>& dataset,
                           const PointCloudView_<VRAM_HIP>& model, const UmeyamaReductionConstraints params,
                           CrossStatistics& stats) {
  stats = detail::to_f32(detail::statistics_total(RMH_STATS_P2L, pre_transform, dataset, model, params));
}

inline CrossStatistics statistics_p2l(const Transform& pre_transform, const PointCloudView_<VRAM_HIP>& dataset,
                                      const PointCloudView_<VRAM_HIP>& model, const UmeyamaReductionConstraints params) {
  return detail::to_f32(detail::statistics_total(RMH_STATS_P2L, pre_transform, dataset, model, params));
}

// ---- objectwise point to plane (statistics.cuh:139-158) -----------------------------------
// One statistics entry per object: its pixel window bboxes[o] (x columns, y rows) of
// image-shaped scans with row stride `width`, pre-transform pre_transforms[o] and
// constraints params[o]; the element set is the reference kernel's (rmh.h,
// rmh_statistics_objectwise_p2l).

inline void statistics_objectwise_p2l(const MemoryView<Transform, VRAM_HIP>& pre_transforms,
                                      const PointCloudView_<VRAM_HIP>& dataset, const PointCloudView_<VRAM_HIP>& model,
                                      const unsigned int& width, const unsigned int& height,
                                      const MemoryView<UmeyamaReductionConstraints, VRAM_HIP>& params,
                                      const MemoryView<AABB, VRAM_HIP>& bboxes,
                                      MemoryView<CrossStatisticsD, VRAM_HIP>& stats) {
  const rmh_cloud d = detail::to_rmh_cloud(dataset, 0, 0), m = detail::to_rmh_cloud(model, 0, 0);
  RM_RMH_CHECK(rmh_statistics_objectwise_p2l(
      reinterpret_cast<const rmh_transform*>(pre_transforms.raw()), &d, &m, width, height,
      static_cast<uint32_t>(dataset.points.size()), reinterpret_cast<const rmh_umeyama_constraints*>(params.raw()),
      reinterpret_cast<const rmh_aabb*>(bboxes.raw()), static_cast<uint32_t>(stats.size()), 0u,
      reinterpret_cast<rmh_cross_stats*>(stats.raw()), nullptr));
}

inline void statistics_objectwise_p2l(const MemoryView<Transform, VRAM_HIP>& pre_transforms,
                                      const PointCloudView_<VRAM_HIP>& dataset, const PointCloudView_<VRAM_HIP>& model,
                                      const unsigned int& width, const unsigned int& height,
                                      const MemoryView<UmeyamaReductionConstraints, VRAM_HIP>& params,
                                      const MemoryView<AABB, VRAM_HIP>& bboxes,
                                      MemoryView<CrossStatistics, VRAM_HIP>& stats) {
  Memory<CrossStatisticsD, VRAM_HIP> s64(stats.size());
  statistics_objectwise_p2l(pre_transforms, dataset, model, width, height, params, bboxes, s64);
  RM_RMH_CHECK(rmh_cross_stats_to_f32(reinterpret_cast<const rmh_cross_stats*>(s64.raw()), stats.raw(),
                                      static_cast<uint32_t>(stats.size()), nullptr));
}

// host-side per-object inputs and results (statistics.cuh:150-158)
inline void statistics_objectwise_p2l(const MemoryView<Transform, RAM>& pre_transforms,
                                      const PointCloudView_<VRAM_HIP>& dataset, const PointCloudView_<VRAM_HIP>& model,
                                      const unsigned int& width, const unsigned int& height,
                                      const MemoryView<UmeyamaReductionConstraints, RAM>& params,
                                      const MemoryView<AABB, RAM>& bboxes, MemoryView<CrossStatistics, RAM>& stats) {
  Memory<Transform, VRAM_HIP> pre_d(pre_transforms);
  Memory<UmeyamaReductionConstraints, VRAM_HIP> prm_d(params);
  Memory<AABB, VRAM_HIP> bb_d(bboxes);
  Memory<CrossStatistics, VRAM_HIP> st_d(stats.size());
  statistics_objectwise_p2l(pre_d, dataset, model, width, height, prm_d, bb_d, st_d);
  copy(static_cast<const MemoryView<CrossStatistics, VRAM_HIP>&>(st_d), stats);
}

static_assert(sizeof(UmeyamaReductionConstraints) == sizeof(rmh_umeyama_constraints), "constraints layout");

}  // namespace rmagine
