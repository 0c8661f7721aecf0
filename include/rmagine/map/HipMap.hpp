// HipMap: a triangle mesh and its BVH resident in HBM (librmh), the role of OptixMap /
// EmbreeMap (src/rmagine_optix/include/rmagine/map/OptixMap.hpp:88-110,
// src/rmagine_embree/include/rmagine/map/EmbreeMap.hpp:70-117).  The map owns the mesh copy,
// the BVH and a blocking HIP stream on which simulators enqueue their launches
// (SimulatorOptix.cpp:38-42 uses the map's stream the same way).
#pragma once

#include <rmh.h>

#include <limits>
#include <memory>
#include <stdexcept>
#include <string>
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

  // Nearest map point strictly within max_distance (EmbreeMap::closestPoint,
  // src/rmagine_embree/src/map/EmbreeMap.cpp:33-37; result {d, p, n, primID, geomID},
  // d = FLT_MAX and ids 0xFFFFFFFF when nothing is that close).  Blocking, one query; see
  // rmh_closest_points for batches.
  struct ClosestPointResult {
    float d;
    Vector p;
    Vector n;
    unsigned int primID;
    unsigned int geomID;
  };
  ClosestPointResult closestPoint(const Vector& qp, float max_distance = std::numeric_limits<float>::max()) const {
    Memory<Vector, RAM> q(1);
    q[0] = qp;
    Memory<Vector, VRAM_HIP> qd;
    qd = q;
    Memory<float, VRAM_HIP> d(1);
    Memory<Vector, VRAM_HIP> p(1), n(1);
    Memory<unsigned int, VRAM_HIP> f(1), g(1);
    rmh_closest_outputs out{};
    out.dist = d.raw();
    out.points = reinterpret_cast<rmh_vec3*>(p.raw());
    out.normals = reinterpret_cast<rmh_vec3*>(n.raw());
    out.face_ids = f.raw();
    out.geom_ids = g.raw();
    RM_RMH_CHECK(rmh_closest_points(m_map, reinterpret_cast<const rmh_vec3*>(qd.raw()), 1, max_distance, &out,
                                    m_stream));
    Memory<float, RAM> dh = d;
    Memory<Vector, RAM> ph = p, nh = n;
    Memory<unsigned int, RAM> fh = f, gh = g;
    return ClosestPointResult{dh[0], ph[0], nh[0], fh[0], gh[0]};
  }

  // New vertex positions (same count and faces): updates triangles, normals and BVH boxes
  // in place (no rebuild).  The role of OptiX's GAS update / EmbreeMesh::apply + commit.
  void refit(const MemoryView<Vector, VRAM_HIP>& vertices) {
    RM_RMH_CHECK(rmh_map_refit(m_map, reinterpret_cast<const float*>(vertices.raw()), m_stream));
  }

  // Node layout of the ray kernels: the 4-wide quantised tree (true) or the binary tree
  // (false, the default).  Returns whether the 4-wide tree is in use.
  bool setLayout(bool wide) {
    RM_RMH_CHECK(rmh_map_set_layout(m_map, wide ? 1 : 0, m_stream));
    return info().n_wide_nodes > 0;
  }

  // Geometry / object id reported for each mesh (empty vector keeps the mesh index).
  void setMeshIds(const std::vector<uint32_t>& geom_ids, const std::vector<uint32_t>& object_ids) {
    const uint32_t n = info().n_meshes;
    if ((!geom_ids.empty() && geom_ids.size() != n) || (!object_ids.empty() && object_ids.size() != n))
      throw std::runtime_error("HipMap::setMeshIds: one id per mesh expected");
    RM_RMH_CHECK(rmh_map_set_mesh_ids(m_map, geom_ids.empty() ? nullptr : geom_ids.data(),
                                      object_ids.empty() ? nullptr : object_ids.data(), n));
  }

 private:
  rmh_map* m_map = nullptr;
  void* m_stream = nullptr;
};

using HipMapPtr = std::shared_ptr<HipMap>;

// Mesh file (PLY, OBJ, STL, COLLADA) as a map, with the scene graph resolved and the ids
// assigned as Rmagine's import_embree_map / import_optix_map do
// (src/rmagine_embree/include/rmagine/map/EmbreeMap.hpp:94-117,
//  src/rmagine_optix/include/rmagine/map/OptixMap.hpp:88-109).  flags:
// RMH_SCENE_POLYGONS_FAN triangulates polygons (default: their first three indices).
inline HipMapPtr import_hip_map(const std::string& meshfile, uint32_t flags = 0,
                                const rmh_bvh_params* params = nullptr) {
  rmh_scene* sc = nullptr;
  RM_RMH_CHECK(rmh_scene_load(meshfile.c_str(), flags, &sc));
  uint32_t nv = 0, nf = 0, nm = 0;
  (void)rmh_scene_counts(sc, &nv, &nf, &nm);
  Memory<Vector, RAM> V(nv);
  Memory<Face, RAM> F(nf);
  std::vector<uint32_t> off(nm + 1), geom(nm), obj(nm);
  const int rc = rmh_scene_copy(sc, reinterpret_cast<float*>(V.raw()), reinterpret_cast<uint32_t*>(F.raw()),
                                off.data(), geom.data(), obj.data());
  (void)rmh_scene_destroy(sc);
  RM_RMH_CHECK(rc);
  auto map = std::make_shared<HipMap>(Memory<Vector, VRAM_HIP>(V), Memory<Face, VRAM_HIP>(F), off, params);
  map->setMeshIds(geom, obj);
  return map;
}

}  // namespace rmagine
