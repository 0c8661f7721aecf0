// Point clouds for builds without rmagine_core: same shape as
// src/rmagine_core/include/rmagine/types/PointCloud.hpp:46-98.  PointCloud_ owns its buffers,
// PointCloudView_ points at existing ones (points required; mask, normals, ids and time
// offsets empty unless set).
#pragma once

#include <cstdint>

#include <rmagine/math/types.h>
#include <rmagine/types/Memory.hpp>

namespace rmagine {

template <typename MemT>
struct PointCloud_ {
  Memory<Vector, MemT> points;
  Memory<uint8_t, MemT> mask;
  Memory<Vector, MemT> normals;
  Memory<unsigned int, MemT> ids;
  Memory<float, MemT> time_offsets;
};

using PointCloud = PointCloud_<RAM>;

template <typename MemT>
struct PointCloudView_ {
  MemoryView<Vector, MemT> points;
  MemoryView<uint8_t, MemT> mask = MemoryView<uint8_t, MemT>::Empty();
  MemoryView<Vector, MemT> normals = MemoryView<Vector, MemT>::Empty();
  MemoryView<unsigned int, MemT> ids = MemoryView<unsigned int, MemT>::Empty();
  MemoryView<float, MemT> time_offsets = MemoryView<float, MemT>::Empty();
};

using PointCloudView = PointCloudView_<RAM>;

// deep copy of the viewed buffers into another memory space
template <typename MemTto, typename MemTfrom>
PointCloud_<MemTto> transfer(const PointCloudView_<MemTfrom>& from) {
  PointCloud_<MemTto> r;
  r.points = from.points;
  r.mask = from.mask;
  r.normals = from.normals;
  r.ids = from.ids;
  r.time_offsets = from.time_offsets;
  return r;
}

template <typename MemT>
const PointCloudView_<MemT> watch(const PointCloud_<MemT>& from) {
  PointCloudView_<MemT> v{MemoryView<Vector, MemT>(const_cast<Vector*>(from.points.raw()), from.points.size())};
  v.mask = MemoryView<uint8_t, MemT>(const_cast<uint8_t*>(from.mask.raw()), from.mask.size());
  v.normals = MemoryView<Vector, MemT>(const_cast<Vector*>(from.normals.raw()), from.normals.size());
  v.ids = MemoryView<unsigned int, MemT>(const_cast<unsigned int*>(from.ids.raw()), from.ids.size());
  v.time_offsets = MemoryView<float, MemT>(const_cast<float*>(from.time_offsets.raw()), from.time_offsets.size());
  return v;
}

}  // namespace rmagine
