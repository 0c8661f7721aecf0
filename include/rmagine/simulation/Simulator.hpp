// Backend-generic simulator traits, for builds without rmagine_core: same names and meaning
// as src/rmagine_core/include/rmagine/simulation/Simulator.hpp:45-78.  SimulatorType<ModelT,
// CompT> names the simulator class of a sensor model on a computing backend (a tag type);
// backends specialise it (SimulatorHipTraits.hpp for the Hip tag).  Inside an Rmagine tree
// the reference's own Simulator.hpp takes this file's place.
#pragma once

#include <type_traits>

namespace rmagine {

struct NoneType {};

template <typename ModelT, typename CompT>
class SimulatorType {
 public:
  using Class = NoneType;
  using Ptr = NoneType;
};

template <typename ModelT, typename CompT>
using Simulator = typename SimulatorType<ModelT, CompT>::Class;

template <typename ModelT, typename CompT>
using SimulatorPtr = typename SimulatorType<ModelT, CompT>::Ptr;

// A new simulator of ModelT on backend CompT, or false when the backend has none for the model.
template <typename ModelT, typename CompT>
auto make_simulator() {
  using SimT = SimulatorType<ModelT, CompT>;
  if constexpr (std::is_same<typename SimT::Class, NoneType>::value) {
    return false;
  } else {
    typename SimT::Ptr sim(new typename SimT::Class);
    return sim;
  }
}

}  // namespace rmagine
