// O1DnSimulatorHip (counterpart of O1DnSimulatorOptix / O1DnSimulatorEmbree); defined in SimulatorHip.hpp.
#pragma once
#include <rmagine/simulation/SimulatorHip.hpp>
