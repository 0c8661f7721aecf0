// O1DnSimulatorHip: one origin, n directions, on SimulatorHip.
//
// Counterpart of O1DnSimulatorOptix (src/rmagine_optix/include/rmagine/simulation/
// O1DnSimulatorOptix.hpp, O1DnSimulatorOptix.cpp:41-83): a RAM model's directions are uploaded
// once; a VRAM_HIP model's directions are used in place and must outlive the simulator.
#pragma once

#include <memory>

#include <rmagine/simulation/SimulatorHip.hpp>

namespace rmagine {

class O1DnSimulatorHip : public SimulatorHip {
 public:
  using SimulatorHip::SimulatorHip;
  // uploads the directions, like O1DnSimulatorOptix::setModel (O1DnSimulatorOptix.cpp:41-83)
  void setModel(const O1DnModel_<RAM>& model) {
    Memory<Vector, VRAM_HIP> dirs(model.dirs);
    setDev(model, dirs);
  }
  void setModel(const O1DnModel_<VRAM_HIP>& model) { setDev(model, model.dirs); }

 private:
  template <typename M>
  void setDev(const M& model, const MemoryView<Vector, VRAM_HIP>& dirs) {
    rmh_o1dn_model m{};
    m.width = model.width;
    m.height = model.height;
    m.range = {model.range.min, model.range.max};
    m.orig = {model.orig.x, model.orig.y, model.orig.z};
    m.dirs_dev = reinterpret_cast<const rmh_vec3*>(dirs.raw());
    setSensor(RMH_MODEL_O1DN, &m);
  }
};

using O1DnSimulatorHipPtr = std::shared_ptr<O1DnSimulatorHip>;

}  // namespace rmagine
