// PinholeSimulatorHip (counterpart of PinholeSimulatorOptix / PinholeSimulatorEmbree); defined in SimulatorHip.hpp.
#pragma once
#include <rmagine/simulation/SimulatorHip.hpp>
