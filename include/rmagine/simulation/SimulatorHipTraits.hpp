// The Hip computing backend tag and its simulator traits, the counterpart of
// src/rmagine_optix/include/rmagine/simulation/SimulatorOptixTraits.hpp:50-90 (struct Optix)
// and SimulatorEmbreeTraits.hpp:50-90 (struct Embree): make_simulator<SphericalModel, Hip>()
// returns a SphereSimulatorHipPtr, and so on for the pinhole, O1Dn and OnDn models.
#pragma once

#include <rmagine/simulation/Simulator.hpp>
#include <rmagine/simulation/O1DnSimulatorHip.hpp>
#include <rmagine/simulation/OnDnSimulatorHip.hpp>
#include <rmagine/simulation/PinholeSimulatorHip.hpp>
#include <rmagine/simulation/SphereSimulatorHip.hpp>

namespace rmagine {

struct Hip {};

template <>
class SimulatorType<SphericalModel, Hip> {
 public:
  using Class = SphereSimulatorHip;
  using Ptr = SphereSimulatorHipPtr;
};

template <>
class SimulatorType<PinholeModel, Hip> {
 public:
  using Class = PinholeSimulatorHip;
  using Ptr = PinholeSimulatorHipPtr;
};

template <>
class SimulatorType<O1DnModel, Hip> {
 public:
  using Class = O1DnSimulatorHip;
  using Ptr = O1DnSimulatorHipPtr;
};

template <>
class SimulatorType<OnDnModel, Hip> {
 public:
  using Class = OnDnSimulatorHip;
  using Ptr = OnDnSimulatorHipPtr;
};

}  // namespace rmagine
