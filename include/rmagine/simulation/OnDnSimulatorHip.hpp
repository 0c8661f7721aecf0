// OnDnSimulatorHip (counterpart of OnDnSimulatorOptix / OnDnSimulatorEmbree); defined in SimulatorHip.hpp.
#pragma once
#include <rmagine/simulation/SimulatorHip.hpp>
