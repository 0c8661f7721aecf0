// PinholeSimulatorHip: pinhole (depth camera) sensor model on SimulatorHip.
//
// Counterpart of PinholeSimulatorOptix (src/rmagine_optix/include/rmagine/simulation/
// PinholeSimulatorOptix.hpp) and PinholeSimulatorEmbree: setModel takes the model by value or as a
// one-element RAM view; the rays are generated on the device from it.
#pragma once

#include <memory>

#include <rmagine/simulation/SimulatorHip.hpp>

namespace rmagine {

class PinholeSimulatorHip : public SimulatorHip {
 public:
  using SimulatorHip::SimulatorHip;
  void setModel(const PinholeModel& model) {
    m_model = model;
    setSensor(RMH_MODEL_PINHOLE, &model);
  }
  void setModel(const MemoryView<PinholeModel, RAM>& model) { setModel(model[0]); }
  const PinholeModel& model() const { return m_model; }

 private:
  PinholeModel m_model{};
};

using PinholeSimulatorHipPtr = std::shared_ptr<PinholeSimulatorHip>;

}  // namespace rmagine
