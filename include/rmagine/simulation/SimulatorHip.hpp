// SimulatorHip: the base of Rmagine's GPU Simulator API on librmh (map, Tsb, simulate).  The
// model classes derive from it in SphereSimulatorHip.hpp, PinholeSimulatorHip.hpp,
// O1DnSimulatorHip.hpp and OnDnSimulatorHip.hpp.
//
// API shape of SimulatorOptix / SphereSimulatorOptix / PinholeSimulatorOptix /
// O1DnSimulatorOptix / OnDnSimulatorOptix
// (src/rmagine_optix/include/rmagine/simulation/SimulatorOptix.hpp:22-79,
//  SphereSimulatorOptix.hpp:122-191, SphereSimulatorOptix.tcc:11-100,
//  O1DnSimulatorOptix.cpp:41-102): setMap / setTsb / setModel / simulate<BundleT>(Tbm[, res])
// with poses and results in device memory (VRAM_HIP), asynchronous on the map's stream,
// and the results of the Embree simulators
// (src/rmagine_embree/include/rmagine/simulation/SphereSimulatorEmbree.tcc:50-254).
// An attribute is produced iff its buffer in the bundle has size > 0
// (embree_common.h:79-139, optix/common.h:155-232).
#pragma once

#include <rmh.h>

#include <memory>
#include <stdexcept>

#include <rmagine/map/HipMap.hpp>
#include <rmagine/math/types.h>
#include <rmagine/simulation/SimulationResults.hpp>
#include <rmagine/types/MemoryHip.hpp>
#include <rmagine/types/sensor_models.h>

namespace rmagine {

namespace detail {

template <typename BundleT>
rmh_outputs outputs_of(BundleT& res, size_t n) {
  rmh_outputs o{};
  auto take = [n](auto& mem) -> void* {
    if (mem.size() == 0) return nullptr;
    if (mem.size() != n) throw std::runtime_error("[SimulatorHip] result buffer size does not match W*H*N");
    return mem.raw();
  };
  if constexpr (BundleT::template has<Hits<VRAM_HIP>>()) o.hits = static_cast<uint8_t*>(take(res.Hits<VRAM_HIP>::hits));
  if constexpr (BundleT::template has<Ranges<VRAM_HIP>>()) o.ranges = static_cast<float*>(take(res.Ranges<VRAM_HIP>::ranges));
  if constexpr (BundleT::template has<Points<VRAM_HIP>>()) o.points = static_cast<rmh_vec3*>(take(res.Points<VRAM_HIP>::points));
  if constexpr (BundleT::template has<Normals<VRAM_HIP>>()) o.normals = static_cast<rmh_vec3*>(take(res.Normals<VRAM_HIP>::normals));
  if constexpr (BundleT::template has<FaceIds<VRAM_HIP>>()) o.face_ids = static_cast<uint32_t*>(take(res.FaceIds<VRAM_HIP>::face_ids));
  if constexpr (BundleT::template has<GeomIds<VRAM_HIP>>()) o.geom_ids = static_cast<uint32_t*>(take(res.GeomIds<VRAM_HIP>::geom_ids));
  if constexpr (BundleT::template has<ObjectIds<VRAM_HIP>>()) o.object_ids = static_cast<uint32_t*>(take(res.ObjectIds<VRAM_HIP>::object_ids));
  return o;
}

struct SensorDeleter {
  void operator()(rmh_sensor* s) const { (void)rmh_sensor_destroy(s); }
};

inline rmh_transform to_rmh(const Transform& T) {
  rmh_transform r;
  r.R = {T.R.x, T.R.y, T.R.z, T.R.w};
  r.t = {T.t.x, T.t.y, T.t.z};
  r.stamp = T.stamp;
  return r;
}

}  // namespace detail

class SimulatorHip {
 public:
  SimulatorHip() { m_Tsb.setIdentity(); }
  explicit SimulatorHip(HipMapPtr map) : m_map(std::move(map)) { m_Tsb.setIdentity(); }
  virtual ~SimulatorHip() = default;

  void setMap(HipMapPtr map) { m_map = std::move(map); }
  HipMapPtr map() const { return m_map; }

  void setTsb(const MemoryView<Transform, RAM>& Tsb) { m_Tsb = Tsb[0]; }
  void setTsb(const Transform& Tsb) { m_Tsb = Tsb; }

  uint32_t getWidth() const { return m_width; }
  uint32_t getHeight() const { return m_height; }

  // SimulatorOptix::simulate<BundleT>(Tbm, res)
  template <typename BundleT>
  void simulate(const MemoryView<Transform, VRAM_HIP>& Tbm, BundleT& res) const {
    if (!m_map) throw std::runtime_error("[SimulatorHip] simulate(): no map set");
    if (!m_sensor) throw std::runtime_error("[SimulatorHip] simulate(): no sensor model set");
    const size_t n = static_cast<size_t>(m_width) * m_height * Tbm.size();
    const rmh_outputs out = detail::outputs_of(res, n);
    const rmh_transform Tsb = detail::to_rmh(m_Tsb);
    RM_RMH_CHECK(rmh_simulate(m_map->handle(), m_sensor.get(), &Tsb, reinterpret_cast<const rmh_transform*>(Tbm.raw()),
                              static_cast<uint32_t>(Tbm.size()), &out, m_map->stream()));
  }

  template <typename BundleT>
  BundleT simulate(const MemoryView<Transform, VRAM_HIP>& Tbm) const {
    BundleT res;
    resize_memory_bundle<VRAM_HIP>(res, m_width, m_height, static_cast<unsigned int>(Tbm.size()));
    simulate(Tbm, res);
    return res;
  }

  // poses in host memory (a Memory<Transform, RAM> or a make_view of one pose): uploaded
  // first, as the reference's simulate(const Memory<Transform, VRAM_CUDA>&) converts them
  template <typename BundleT>
  void simulate(const MemoryView<Transform, RAM>& Tbm, BundleT& res) const {
    Memory<Transform, VRAM_HIP> T(Tbm);
    simulate(T, res);
  }

  template <typename BundleT>
  BundleT simulate(const MemoryView<Transform, RAM>& Tbm) const {
    BundleT res;
    resize_memory_bundle<VRAM_HIP>(res, m_width, m_height, static_cast<unsigned int>(Tbm.size()));
    simulate(Tbm, res);
    return res;
  }

  // single pose from the host (SimulatorOptix.tcc:6-24 uploads it)
  template <typename BundleT>
  void simulate(const Transform& Tbm, BundleT& res) const {
    Memory<Transform, VRAM_HIP> T(1);
    T = make_view(const_cast<Transform&>(Tbm));
    simulate(T, res);
  }

  template <typename BundleT>
  BundleT simulate(const Transform& Tbm) const {
    BundleT res;
    resize_memory_bundle<VRAM_HIP>(res, m_width, m_height, 1);
    simulate(Tbm, res);
    return res;
  }

  // OptiX builds a pipeline per attribute set here; the HIP kernels are specialised at
  // build time, so this only checks that a map and a model are set.
  template <typename BundleT>
  void preBuildProgram() {
    if (!m_sensor) throw std::runtime_error("[SimulatorHip] preBuildProgram(): no sensor model set");
  }

  // wait for all launches of this simulator's map stream
  void synchronize() const {
    if (m_map) RM_RMH_CHECK(rmh_stream_synchronize(m_map->stream()));
  }

  // O1Dn/OnDn: cast the rays in direction-sorted order (the default) or in list order; the
  // results are identical, only the speed differs (image models ignore it)
  void setRayOrder(bool sorted) { setRayOrder(sorted ? RayOrder::Sorted : RayOrder::List); }

  // O1Dn/OnDn traversal order: list, direction-sorted (the default) or the model's W x H
  // grid in 4 x 16 tiles (as image sensors are cast); the results are identical
  enum class RayOrder { List = RMH_RAY_ORDER_LIST, Sorted = RMH_RAY_ORDER_SORTED, Tiles = RMH_RAY_ORDER_TILES };
  void setRayOrder(RayOrder order) {
    if (!m_sensor) throw std::runtime_error("[SimulatorHip] setRayOrder(): no sensor model set");
    RM_RMH_CHECK(rmh_sensor_set_ray_order(m_sensor.get(), static_cast<int>(order)));
  }

  // O1Dn/OnDn: walk the BVH speculatively (the default, fastest for scattered rays) or the
  // classic way (fastest for a scanner's coherent ray grid); the results are identical, only
  // the speed differs (image models ignore it)
  void setWalk(bool speculative) {
    if (!m_sensor) throw std::runtime_error("[SimulatorHip] setWalk(): no sensor model set");
    RM_RMH_CHECK(rmh_sensor_set_walk(m_sensor.get(), speculative ? 1 : 0));
  }

  // the librmh sensor of the current model (shared: a batch keeps it alive) and Tsb, for
  // BatchSimulatorHip and poseGradHip
  std::shared_ptr<rmh_sensor> sensorHandle() const { return m_sensor; }
  const Transform& getTsb() const { return m_Tsb; }

 protected:
  void setSensor(int kind, const void* model) {
    rmh_sensor* s = nullptr;
    RM_RMH_CHECK(rmh_sensor_create(kind, model, &s));
    m_sensor.reset(s, detail::SensorDeleter());
    RM_RMH_CHECK(rmh_sensor_dims(s, &m_width, &m_height));
  }

  HipMapPtr m_map;
  Transform m_Tsb;
  std::shared_ptr<rmh_sensor> m_sensor;
  uint32_t m_width = 0, m_height = 0;
};

using SimulatorHipPtr = std::shared_ptr<SimulatorHip>;

}  // namespace rmagine
