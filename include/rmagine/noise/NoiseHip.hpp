// Range noise on the device (librmh rmh_noise_apply): GaussianNoiseHip, RelGaussianNoiseHip
// and UniformDustNoiseHip, the counterparts of Rmagine's NoiseCuda family
// (src/rmagine_cuda/include/rmagine/noise/NoiseCuda.hpp, GaussianNoiseCuda.hpp,
//  RelGaussianNoiseCuda.hpp, UniformDustNoiseCuda.hpp).  apply() works in place on a
// device range buffer, asynchronously on `stream` (default: the null stream).  The random
// numbers are counter-based (no per-element state): Options keeps seed and max_range; the
// reference's memory options have nothing to manage.
#pragma once

#include <rmh.h>

#include <cstdint>
#include <memory>

#include <rmagine/types/MemoryHip.hpp>
#include <rmagine/util/hip/HipException.hpp>

namespace rmagine {

class NoiseHip {
 public:
  struct Options {
    unsigned int seed = 42;
    float max_range = 10000.0f;
  };

  virtual ~NoiseHip() = default;

  void apply(MemoryView<float, VRAM_HIP>& ranges, void* stream = nullptr) {
    RM_RMH_CHECK(rmh_noise_apply(&m_params, ranges.raw(), ranges.size(), m_calls++, stream));
  }

 protected:
  NoiseHip(uint32_t kind, float p0, float p1, float p2, Options options) {
    m_params.kind = kind;
    m_params.p0 = p0;
    m_params.p1 = p1;
    m_params.p2 = p2;
    m_params.max_range = options.max_range;
    m_params.seed = options.seed;
  }

  rmh_noise_params m_params{};
  uint32_t m_calls = 0;
};

class GaussianNoiseHip : public NoiseHip {
 public:
  GaussianNoiseHip(float mean, float stddev, Options options = {})
      : NoiseHip(RMH_NOISE_GAUSSIAN, mean, stddev, 0.0f, options) {}
};

class RelGaussianNoiseHip : public NoiseHip {
 public:
  RelGaussianNoiseHip(float mean, float stddev, float range_exp, Options options = {})
      : NoiseHip(RMH_NOISE_REL_GAUSSIAN, mean, stddev, range_exp, options) {}
};

class UniformDustNoiseHip : public NoiseHip {
 public:
  UniformDustNoiseHip(float hit_probability, float return_probability, Options options = {})
      : NoiseHip(RMH_NOISE_UNIFORM_DUST, hit_probability, return_probability, 0.0f, options) {}
};

using NoiseHipPtr = std::shared_ptr<NoiseHip>;

}  // namespace rmagine
