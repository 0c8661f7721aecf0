// rmagine::UmeyamaReductionConstraints for builds without rmagine_core
// (src/rmagine_core/include/rmagine/types/UmeyamaReductionConstraints.hpp:46-59): the
// correspondence gate and the dataset / model ids the statistics keep.  12 bytes, the layout
// of rmh_umeyama_constraints.
#pragma once

namespace rmagine {

struct UmeyamaReductionConstraints {
  float max_dist;           // ignore correspondences farther apart
  unsigned int dataset_id;  // ignore dataset ids other than this one (when ids are given)
  unsigned int model_id;    // ignore model ids other than this one (when ids are given)
};

}  // namespace rmagine
