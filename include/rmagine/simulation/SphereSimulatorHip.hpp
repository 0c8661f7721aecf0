// SphereSimulatorHip: spherical (LiDAR) sensor model on SimulatorHip.
//
// Counterpart of SphereSimulatorOptix (src/rmagine_optix/include/rmagine/simulation/
// SphereSimulatorOptix.hpp:122-191) and SphereSimulatorEmbree: setModel takes the model by value
// or as a one-element RAM view; the rays are generated on the device from it.
#pragma once

#include <memory>

#include <rmagine/simulation/SimulatorHip.hpp>

namespace rmagine {

class SphereSimulatorHip : public SimulatorHip {
 public:
  using SimulatorHip::SimulatorHip;
  void setModel(const SphericalModel& model) {
    m_model = model;
    setSensor(RMH_MODEL_SPHERICAL, &model);
  }
  void setModel(const MemoryView<SphericalModel, RAM>& model) { setModel(model[0]); }
  const SphericalModel& model() const { return m_model; }

 private:
  SphericalModel m_model{};
};

using SphereSimulatorHipPtr = std::shared_ptr<SphereSimulatorHip>;

}  // namespace rmagine
