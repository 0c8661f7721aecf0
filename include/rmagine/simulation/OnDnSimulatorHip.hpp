// OnDnSimulatorHip: n origins, n directions, on SimulatorHip.
//
// Counterpart of OnDnSimulatorOptix (src/rmagine_optix/include/rmagine/simulation/
// OnDnSimulatorOptix.hpp) and OnDnSimulatorEmbree: a RAM model's origins and directions are
// uploaded once; a VRAM_HIP model's buffers are used in place and must outlive the simulator.
#pragma once

#include <memory>

#include <rmagine/simulation/SimulatorHip.hpp>

namespace rmagine {

class OnDnSimulatorHip : public SimulatorHip {
 public:
  using SimulatorHip::SimulatorHip;
  void setModel(const OnDnModel_<RAM>& model) {
    Memory<Vector, VRAM_HIP> dirs(model.dirs), origs(model.origs);
    setDev(model, origs, dirs);
  }
  void setModel(const OnDnModel_<VRAM_HIP>& model) { setDev(model, model.origs, model.dirs); }

 private:
  template <typename M>
  void setDev(const M& model, const MemoryView<Vector, VRAM_HIP>& origs, const MemoryView<Vector, VRAM_HIP>& dirs) {
    rmh_ondn_model m{};
    m.width = model.width;
    m.height = model.height;
    m.range = {model.range.min, model.range.max};
    m.origs_dev = reinterpret_cast<const rmh_vec3*>(origs.raw());
    m.dirs_dev = reinterpret_cast<const rmh_vec3*>(dirs.raw());
    setSensor(RMH_MODEL_ONDN, &m);
  }
};

using OnDnSimulatorHipPtr = std::shared_ptr<OnDnSimulatorHip>;

}  // namespace rmagine
