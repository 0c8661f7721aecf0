// HipMap: a triangle mesh and its BVH resident in HBM (librmh), the role of OptixMap /
// EmbreeMap (src/rmagine_optix/include/rmagine/map/OptixMap.hpp:88-110,
// src/rmagine_embree/include/rmagine/map/EmbreeMap.hpp:70-117).  The map owns the mesh copy,
// the BVH and a blocking HIP stream on which simulators enqueue their launches
// (SimulatorOptix.cpp:38-42 uses the map's stream the same way).
#pragma once

#include <rmh.h>

#include <memory>
#include <vector>

#include <rmagine/math/types.h>
#include <rmagine/types/MemoryHip.hpp>
#include <rmagine/util/hip/HipException.hpp>

namespace rmagine {

class HipMap {
 public:
  // Single mesh from host buffers.
  HipMap(const MemoryView<Vector, RAM>& vertices, const MemoryView<Face, RAM>& faces,
         const rmh_bvh_params* params = nullptr)
      : HipMap(Memory<Vector, VRAM_HIP>(vertices), Memory<Face, VRAM_HIP>(faces), {}, params) {}

  // Meshes from device buffers; mesh_face_offsets (n_meshes+1 entries) assigns geometry ids.
  HipMap(const MemoryView<Vector, VRAM_HIP>& vertices, const MemoryView<Face, VRAM_HIP>& faces,
         const std::vector<uint32_t>& mesh_face_offsets = {}, const rmh_bvh_params* params = nullptr) {
    RM_RMH_CHECK(rmh_stream_create(&m_stream));
    rmh_map* m = nullptr;
    const uint32_t n_meshes = mesh_face_offsets.empty() ? 0u : static_cast<uint32_t>(mesh_face_offsets.size() - 1);
    const int rc = rmh_map_build(reinterpret_cast<const float*>(vertices.raw()), static_cast<uint32_t>(vertices.size()),
                                 reinterpret_cast<const uint32_t*>(faces.raw()), static_cast<uint32_t>(faces.size()),
                                 mesh_face_offsets.empty() ? nullptr : mesh_face_offsets.data(), n_meshes, params,
                                 m_stream, &m);
    if (rc != RMH_OK) {
      (void)rmh_stream_destroy(m_stream);
      RM_RMH_CHECK(rc);
    }
    m_map = m;
  }

  HipMap(const HipMap&) = delete;
  HipMap& operator=(const HipMap&) = delete;

  ~HipMap() {
    if (m_map) (void)rmh_map_destroy(m_map);
    if (m_stream) (void)rmh_stream_destroy(m_stream);
  }

  const rmh_map* handle() const { return m_map; }
  void* stream() const { return m_stream; }

  rmh_map_info info() const {
    rmh_map_info i{};
    RM_RMH_CHECK(rmh_map_get_info(m_map, &i));
    return i;
  }

 private:
  rmh_map* m_map = nullptr;
  void* m_stream = nullptr;
};

using HipMapPtr = std::shared_ptr<HipMap>;

}  // namespace rmagine
