// Ouster sensor metadata -> O1Dn model: o1dn_from_ouster_meta_file, the counterpart of
// src/rmagine_ouster/include/rmagine/types/ouster_sensors.h:48 (implementation
// src/rmagine_ouster/src/types/ouster_sensors.cpp:8-80), over librmh's rmh_ouster_o1dn
// (no jsoncpp needed).
#pragma once

#include <rmh.h>

#include <string>

#include <rmagine/types/sensor_models.h>
#include <rmagine/util/hip/HipException.hpp>

namespace rmagine {

inline O1DnModel o1dn_from_ouster_meta_file(const std::string& filename) {
  uint32_t W = 0, H = 0;
  RM_RMH_CHECK(rmh_ouster_o1dn(filename.c_str(), &W, &H, nullptr, 0));
  O1DnModel model;
  model.width = W;
  model.height = H;
  model.orig = {0.0f, 0.0f, 0.0f};
  model.range.min = 0.1f;
  model.range.max = 80.0f;
  model.dirs.resize(static_cast<size_t>(W) * H);
  RM_RMH_CHECK(rmh_ouster_o1dn(filename.c_str(), &W, &H, reinterpret_cast<float*>(model.dirs.raw()),
                               3 * static_cast<size_t>(W) * H));
  return model;
}

}  // namespace rmagine
