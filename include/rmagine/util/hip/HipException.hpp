// Error mapping of the C ABI to exceptions, in the style of RM_CUDA_CHECK / CudaException
// (src/rmagine_cuda/include/rmagine/util/cuda/CudaDebug.hpp:43-53,
// src/rmagine_core/include/rmagine/util/exceptions.h:78-85).
#pragma once

#include <rmh.h>

#include <stdexcept>
#include <string>

namespace rmagine {

class HipException : public std::runtime_error {
 public:
  HipException(int code, const std::string& msg, const char* file, int line)
      : std::runtime_error(std::string("[HipException] ") + msg + " (code " + std::to_string(code) + ", " + file +
                           ":" + std::to_string(line) + ")"),
        m_code(code) {}
  int code() const { return m_code; }

 private:
  int m_code;
};

inline void rmh_throw_on_error(int code, const char* file, int line) {
  if (code == RMH_OK) return;
  const char* msg = rmh_last_error();
  if (code == RMH_ERR_NOMAP) throw std::runtime_error(std::string("[SimulatorHip] ") + (msg ? msg : "no map"));
  throw HipException(code, msg ? msg : "", file, line);
}

}  // namespace rmagine

#define RM_RMH_CHECK(expr) ::rmagine::rmh_throw_on_error((expr), __FILE__, __LINE__)
