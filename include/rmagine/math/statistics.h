// Correspondence statistics on host point clouds (RAM), for builds without rmagine_core:
// the CPU statistics_p2p / statistics_p2l of src/rmagine_core/src/math/statistics.cpp:20-196.
// Each element passes the mask and id filters, D = pre_transform * dataset point and the
// distance test are evaluated in FP32 as the reference does (Transform::mult, Vector3 ops,
// statistics.cpp:40-50,105-119); the kept correspondences are accumulated in FP64 (means,
// then the centred cross covariance (1/n) sum (m - m_mean)(d - d_mean)^T) and rounded once
// into the FP32 CrossStatistics, where the reference merges FP32 partial statistics.
#pragma once

#include <cmath>
#include <vector>

#include <rmagine/math/types.h>
#include <rmagine/types/Memory.hpp>
#include <rmagine/types/PointCloud.hpp>
#include <rmagine/types/UmeyamaReductionConstraints.hpp>

namespace rmagine {

namespace detail {

template <bool P2L>
inline CrossStatistics statistics_host(const Transform& pre, const PointCloudView_<RAM>& dataset,
                                       const PointCloudView_<RAM>& model, const UmeyamaReductionConstraints& params) {
  std::vector<Vector3_<double>> D, M;
  for (size_t i = 0; i < dataset.points.size(); i++) {
    if (!dataset.mask.empty() && dataset.mask[i] == 0) continue;
    if (!model.mask.empty() && model.mask[i] == 0) continue;
    if (!dataset.ids.empty() && dataset.ids[i] != params.dataset_id) continue;
    if (!model.ids.empty() && model.ids[i] != params.model_id) continue;
    const Vector Di = pre * dataset.points[i];
    Vector Mi = model.points[i];
    if (P2L) {
      const Vector Ni = model.normals[i];
      const float spd = (Mi - Di).dot(Ni);
      if (!(std::fabs(spd) < params.max_dist)) continue;
      Mi = Di + Ni * spd;
    } else {
      if (!(std::fabs((Mi - Di).l2norm()) < params.max_dist)) continue;
    }
    D.push_back({Di.x, Di.y, Di.z});
    M.push_back({Mi.x, Mi.y, Mi.z});
  }
  CrossStatistics r;
  r.dataset_mean = {0.f, 0.f, 0.f};
  r.model_mean = {0.f, 0.f, 0.f};
  r.covariance.setZeros();
  r.n_meas = static_cast<unsigned int>(D.size());
  if (D.empty()) return r;
  Vector3_<double> dm{0, 0, 0}, mm{0, 0, 0};
  for (size_t k = 0; k < D.size(); k++) {
    dm += D[k];
    mm += M[k];
  }
  const double n = static_cast<double>(D.size());
  dm = dm / n;
  mm = mm / n;
  double C[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};  // column-major: C[c * 3 + r] = sum (m_r - mm_r)(d_c - dm_c)
  for (size_t k = 0; k < D.size(); k++) {
    const double dv[3] = {D[k].x - dm.x, D[k].y - dm.y, D[k].z - dm.z};
    const double mv[3] = {M[k].x - mm.x, M[k].y - mm.y, M[k].z - mm.z};
    for (int c = 0; c < 3; c++)
      for (int row = 0; row < 3; row++) C[c * 3 + row] += mv[row] * dv[c];
  }
  r.dataset_mean = {static_cast<float>(dm.x), static_cast<float>(dm.y), static_cast<float>(dm.z)};
  r.model_mean = {static_cast<float>(mm.x), static_cast<float>(mm.y), static_cast<float>(mm.z)};
  for (int k = 0; k < 9; k++) r.covariance.data[k] = static_cast<float>(C[k] / n);
  return r;
}

}  // namespace detail

inline void statistics_p2p(const Transform& pre_transform, const PointCloudView_<RAM>& dataset,
                           const PointCloudView_<RAM>& model, const UmeyamaReductionConstraints params,
                           MemoryView<CrossStatistics, RAM>& stats) {
  stats[0] = detail::statistics_host<false>(pre_transform, dataset, model, params);
}
inline void statistics_p2p(const Transform& pre_transform, const PointCloudView_<RAM>& dataset,
                           const PointCloudView_<RAM>& model, const UmeyamaReductionConstraints params,
                           CrossStatistics& stats) {
  stats = detail::statistics_host<false>(pre_transform, dataset, model, params);
}
inline CrossStatistics statistics_p2p(const Transform& pre_transform, const PointCloudView_<RAM>& dataset,
                                      const PointCloudView_<RAM>& model, const UmeyamaReductionConstraints params) {
  return detail::statistics_host<false>(pre_transform, dataset, model, params);
}

inline void statistics_p2l(const Transform& pre_transform, const PointCloudView_<RAM>& dataset,
                           const PointCloudView_<RAM>& model, const UmeyamaReductionConstraints params,
                           MemoryView<CrossStatistics, RAM>& stats) {
  stats[0] = detail::statistics_host<true>(pre_transform, dataset, model, params);
}
inline void statistics_p2l(const Transform& pre_transform, const PointCloudView_<RAM>& dataset,
                           const PointCloudView_<RAM>& model, const UmeyamaReductionConstraints params,
                           CrossStatistics& stats) {
  stats = detail::statistics_host<true>(pre_transform, dataset, model, params);
}
inline CrossStatistics statistics_p2l(const Transform& pre_transform, const PointCloudView_<RAM>& dataset,
                                      const PointCloudView_<RAM>& model, const UmeyamaReductionConstraints params) {
  return detail::statistics_host<true>(pre_transform, dataset, model, params);
}

}  // namespace rmagine
