// Batched, graph-captured and differentiable forms of the *SimulatorHip classes.
//
// No counterpart in the reference: it launches one sensor model per optixLaunch
// (src/rmagine_optix/src/simulation/SphereSimulatorOptix.cpp:64-81; the closest piece is
// the model union of src/rmagine_optix/include/rmagine/simulation/optix/sim_program_data.h:32-44)
// and has no differentiable path (closest: jacobian_and_residual_p2l,
// src/rmagine_core/include/rmagine/math/optimization.h:113-137).
//
//   BatchSimulatorHip  many simulators' sensor models, poses and result bundles, cast in one
//                      persistent launch (rmh_batch_*); each bundle receives exactly what
//                      the simulator's own simulate() would write
//   HipGraph           a recorded sequence of non-allocating librmh calls on one stream,
//                      replayed with one call (rmh_graph_*)
//   poseGradHip        Huber range loss and its gradient with respect to 6-DoF poses
//                      (rmh_pose_grad)
#pragma once

#include <rmh.h>

#include <memory>
#include <stdexcept>
#include <utility>
#include <vector>

#include <rmagine/simulation/SimulatorHip.hpp>

namespace rmagine {

// rmh_batch_params; 0 keeps the current value (defaults: 256 rays per grab, 32 waves per
// compute unit, 4 image rows per tile, 10 LDS stack rows, one persistent launch)
struct BatchParamsHip {
  uint32_t chunk_rays = 0;
  uint32_t waves_per_cu = 0;
  uint32_t tile_rows = 0;
  uint32_t lds_stack = 0;
  uint32_t launches = 0;
};

class BatchSimulatorHip {
 public:
  explicit BatchSimulatorHip(HipMapPtr map) : m_map(std::move(map)) {
    if (!m_map) throw std::runtime_error("[BatchSimulatorHip] no map set");
  }
  BatchSimulatorHip(const BatchSimulatorHip&) = delete;
  BatchSimulatorHip& operator=(const BatchSimulatorHip&) = delete;
  ~BatchSimulatorHip() {
    if (m_batch) (void)rmh_batch_destroy(m_batch);
  }

  // Adds the current model and Tsb of `sim`, the poses Tbm and the bundle `res` (sized as
  // for sim.simulate(Tbm, res); buffers of size 0 are not written).  The pose and result
  // buffers must stay valid while the batch is used.  Takes effect at commit().
  template <typename BundleT>
  void add(const SimulatorHip& sim, const MemoryView<Transform, VRAM_HIP>& Tbm, BundleT& res) {
    auto sensor = sim.sensorHandle();
    if (!sensor) throw std::runtime_error("[BatchSimulatorHip] add(): simulator without a sensor model");
    rmh_batch_entry e{};
    e.sensor = sensor.get();
    e.Tsb = detail::to_rmh(sim.getTsb());
    e.Tbm_dev = reinterpret_cast<const rmh_transform*>(Tbm.raw());
    e.n_poses = static_cast<uint32_t>(Tbm.size());
    e.out = detail::outputs_of(res, static_cast<size_t>(sim.getWidth()) * sim.getHeight() * Tbm.size());
    m_entries.push_back(e);
    m_sensors.push_back(std::move(sensor));
  }

  // Builds the device table of every entry added so far.
  void commit(const BatchParamsHip& p = BatchParamsHip()) {
    if (m_batch) {
      (void)rmh_batch_destroy(m_batch);
      m_batch = nullptr;
    }
    const rmh_batch_params prm = to_rmh(p);
    RM_RMH_CHECK(rmh_batch_create(m_map->handle(), m_entries.data(), static_cast<uint32_t>(m_entries.size()), &prm,
                                  &m_batch));
  }

  void setParams(const BatchParamsHip& p) {
    const rmh_batch_params prm = to_rmh(p);
    RM_RMH_CHECK(rmh_batch_set_params(batch(), &prm));
  }

  BatchParamsHip params() const {
    rmh_batch_params prm{};
    RM_RMH_CHECK(rmh_batch_get_info(batch(), nullptr, nullptr, &prm));
    return BatchParamsHip{prm.chunk_rays, prm.waves_per_cu, prm.tile_rows, prm.lds_stack, prm.launches};
  }

  uint64_t totalRays() const {
    uint64_t rays = 0;
    RM_RMH_CHECK(rmh_batch_get_info(batch(), &rays, nullptr, nullptr));
    return rays;
  }

  // asynchronous on the map's stream (or `stream`); launches of one batch must not overlap
  void simulate() const { simulate(m_map->stream()); }
  void simulate(void* stream) const { RM_RMH_CHECK(rmh_batch_simulate(batch(), stream)); }

 private:
  static rmh_batch_params to_rmh(const BatchParamsHip& p) {
    return rmh_batch_params{p.chunk_rays, p.waves_per_cu, p.tile_rows, p.lds_stack, p.launches};
  }
  rmh_batch* batch() const {
    if (!m_batch) throw std::runtime_error("[BatchSimulatorHip] commit() has not been called");
    return m_batch;
  }

  HipMapPtr m_map;
  std::vector<rmh_batch_entry> m_entries;
  std::vector<std::shared_ptr<rmh_sensor>> m_sensors;
  rmh_batch* m_batch = nullptr;
};

class HipGraph {
 public:
  // Records fn() (non-allocating librmh calls enqueued on `stream`, which must not be the
  // default stream) into a graph.
  template <typename Fn>
  static HipGraph capture(void* stream, Fn&& fn) {
    if (!stream) throw std::runtime_error("[HipGraph] capture needs a non-default stream");
    RM_RMH_CHECK(rmh_graph_begin_capture(stream));
    rmh_graph* g = nullptr;
    try {
      fn();
    } catch (...) {
      if (rmh_graph_end_capture(stream, &g) == RMH_OK && g) (void)rmh_graph_destroy(g);
      throw;
    }
    RM_RMH_CHECK(rmh_graph_end_capture(stream, &g));
    return HipGraph(g);
  }
  HipGraph(HipGraph&& o) noexcept : m_graph(o.m_graph) { o.m_graph = nullptr; }
  HipGraph& operator=(HipGraph&& o) noexcept {
    std::swap(m_graph, o.m_graph);
    return *this;
  }
  HipGraph(const HipGraph&) = delete;
  HipGraph& operator=(const HipGraph&) = delete;
  ~HipGraph() {
    if (m_graph) (void)rmh_graph_destroy(m_graph);
  }
  void launch(void* stream) const { RM_RMH_CHECK(rmh_graph_launch(m_graph, stream)); }

 private:
  explicit HipGraph(rmh_graph* g) : m_graph(g) {}
  rmh_graph* m_graph = nullptr;
};

// Per pose of Tbm: {Huber loss, dL/dtau (3), dL/domega (3), rays used} into out (8 doubles
// per pose, device memory) for the simulator's map, model and Tsb.  target: W*H ranges shared
// by all poses, or W*H per pose when per_pose_targets; target_mask may be empty (then a
// target counts when it lies in the model's range).  Asynchronous on the map's stream;
// workspace may be empty (the call then allocates and waits).
inline void poseGradHip(const SimulatorHip& sim, const MemoryView<Transform, VRAM_HIP>& Tbm,
                        const MemoryView<float, VRAM_HIP>& target, const MemoryView<uint8_t, VRAM_HIP>& target_mask,
                        bool per_pose_targets, float huber_delta, MemoryView<double, VRAM_HIP>& out,
                        MemoryView<uint8_t, VRAM_HIP> workspace = MemoryView<uint8_t, VRAM_HIP>(nullptr, 0)) {
  const HipMapPtr map = sim.map();
  auto sensor = sim.sensorHandle();
  if (!map) throw std::runtime_error("[poseGradHip] no map set");
  if (!sensor) throw std::runtime_error("[poseGradHip] no sensor model set");
  const size_t rays = static_cast<size_t>(sim.getWidth()) * sim.getHeight();
  if (target.size() != (per_pose_targets ? rays * Tbm.size() : rays))
    throw std::runtime_error("[poseGradHip] target size does not match the model (and poses)");
  if (target_mask.size() != 0 && target_mask.size() != target.size())
    throw std::runtime_error("[poseGradHip] target_mask size does not match target");
  if (out.size() != 8 * Tbm.size()) throw std::runtime_error("[poseGradHip] out must hold 8 doubles per pose");
  const rmh_transform Tsb = detail::to_rmh(sim.getTsb());
  RM_RMH_CHECK(rmh_pose_grad(map->handle(), sensor.get(), &Tsb, reinterpret_cast<const rmh_transform*>(Tbm.raw()),
                             static_cast<uint32_t>(Tbm.size()), target.raw(),
                             target_mask.size() ? target_mask.raw() : nullptr, per_pose_targets ? 1 : 0, huber_delta,
                             out.raw(), workspace.size() ? workspace.raw() : nullptr, workspace.size(), map->stream()));
}

}  // namespace rmagine
