// SphereSimulatorHip (counterpart of SphereSimulatorOptix / SphereSimulatorEmbree); defined in SimulatorHip.hpp.
#pragma once
#include <rmagine/simulation/SimulatorHip.hpp>
