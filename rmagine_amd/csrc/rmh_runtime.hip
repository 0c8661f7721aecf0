// Runtime part of librmh: thread-local errors, device memory, streams and sensors.
//
// Memory entry points play the role of Rmagine's VRAM_CUDA tag
// (src/rmagine_cuda/src/types/MemoryCuda.cu:13-89): blocking copies, plain device
// allocations.  Sensors are the device-side copy of a sensor model that the reference's
// setModel() uploads (src/rmagine_optix/src/simulation/SphereSimulatorOptix.cpp:20-60,
// O1DnSimulatorOptix.cpp:41-83).
#include "rmh_internal.hpp"

#include <cmath>
#include <cstring>
#include <new>
#include <vector>

namespace rmh {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

int hip_fail(hipError_t e, const char* what) {
  g_last_error = std::string("HIP error '") + hipGetErrorString(e) + "' in " + what;
  return RMH_ERR_HIP;
}

}  // namespace rmh

using namespace rmh;

extern "C" {

const char* rmh_last_error(void) { return g_last_error.c_str(); }

int rmh_version(void) { return RMH_VERSION_MAJOR * 1000 + RMH_VERSION_MINOR; }

int rmh_get_device_count(int* count) {
  if (!count) return fail(RMH_ERR_INVALID, "rmh_get_device_count: count is NULL");
  RMH_HIP_TRY(hipGetDeviceCount(count));
  return RMH_OK;
}

int rmh_set_device(int device) {
  RMH_HIP_TRY(hipSetDevice(device));
  return RMH_OK;
}

int rmh_malloc(void** ptr, size_t bytes) {
  if (!ptr) return fail(RMH_ERR_INVALID, "rmh_malloc: ptr is NULL");
  *ptr = nullptr;
  if (bytes == 0) return RMH_OK;
  hipError_t e = hipMalloc(ptr, bytes);
  if (e != hipSuccess) {
    *ptr = nullptr;
    return fail(RMH_ERR_ALLOC, std::string("rmh_malloc: ") + hipGetErrorString(e));
  }
  return RMH_OK;
}

int rmh_free(void* ptr) {
  if (!ptr) return RMH_OK;
  RMH_HIP_TRY(hipFree(ptr));
  return RMH_OK;
}

int rmh_memcpy(void* dst, const void* src, size_t bytes, int kind) {
  if (bytes == 0) return RMH_OK;
  if (!dst || !src) return fail(RMH_ERR_INVALID, "rmh_memcpy: NULL pointer");
  hipMemcpyKind k;
  switch (kind) {
    case 0: k = hipMemcpyHostToDevice; break;
    case 1: k = hipMemcpyDeviceToHost; break;
    case 2: k = hipMemcpyDeviceToDevice; break;
    case 3: k = hipMemcpyHostToHost; break;
    default: return fail(RMH_ERR_INVALID, "rmh_memcpy: bad kind");
  }
  RMH_HIP_TRY(hipMemcpy(dst, src, bytes, k));
  return RMH_OK;
}

int rmh_memset(void* ptr, int value, size_t bytes) {
  if (bytes == 0) return RMH_OK;
  if (!ptr) return fail(RMH_ERR_INVALID, "rmh_memset: NULL pointer");
  RMH_HIP_TRY(hipMemset(ptr, value, bytes));
  return RMH_OK;
}

int rmh_stream_create(void** stream) {
  if (!stream) return fail(RMH_ERR_INVALID, "rmh_stream_create: NULL");
  hipStream_t s;
  // blocking stream at default priority, like CudaStream (src/rmagine_cuda/src/util/cuda/CudaStream.cpp:9-20)
  RMH_HIP_TRY(hipStreamCreate(&s));
  *stream = s;
  return RMH_OK;
}

int rmh_stream_destroy(void* stream) {
  if (!stream) return RMH_OK;
  RMH_HIP_TRY(hipStreamDestroy(static_cast<hipStream_t>(stream)));
  return RMH_OK;
}

int rmh_stream_synchronize(void* stream) {
  RMH_HIP_TRY(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
  return RMH_OK;
}

int rmh_device_synchronize(void) {
  RMH_HIP_TRY(hipDeviceSynchronize());
  return RMH_OK;
}

// ---- sensors ---------------------------------------------------------------------------

// Direction table of a spherical model.  The reference evaluates
// phi = min + id*inc in FP32 (sensor_models.h:89-92) and then cosf/sinf
// (sensor_models.h:170-175).  The table keeps the FP32 angle but takes the cosine and
// sine in FP64 and rounds once, i.e. the correctly rounded FP32 value of the reference's
// cosf/sinf, so that every device computes identical direction bits.
static void spherical_lut(const rmh_spherical_model& m, std::vector<float2>& lut) {
  const uint32_t H = m.phi.size, W = m.theta.size;
  lut.resize(static_cast<size_t>(H) + W);
  for (uint32_t i = 0; i < H; i++) {
    const float a = m.phi.min + static_cast<float>(i) * m.phi.inc;
    lut[i] = make_float2(static_cast<float>(std::cos(static_cast<double>(a))),
                         static_cast<float>(std::sin(static_cast<double>(a))));
  }
  for (uint32_t j = 0; j < W; j++) {
    const float a = m.theta.min + static_cast<float>(j) * m.theta.inc;
    lut[H + j] = make_float2(static_cast<float>(std::cos(static_cast<double>(a))),
                             static_cast<float>(std::sin(static_cast<double>(a))));
  }
}

static void sensor_free(rmh_sensor* s) {
  if (!s) return;
  if (s->lut) (void)hipFree(s->lut);
  if (s->dirs) (void)hipFree(s->dirs);
  if (s->origs) (void)hipFree(s->origs);
  delete s;
}

int rmh_sensor_create(int kind, const void* model, rmh_sensor** out) {
  if (!model || !out) return fail(RMH_ERR_INVALID, "rmh_sensor_create: NULL argument");
  *out = nullptr;
  rmh_sensor* s = new (std::nothrow) rmh_sensor();
  if (!s) return fail(RMH_ERR_ALLOC, "rmh_sensor_create: out of host memory");
  s->kind = kind;
  (void)hipGetDevice(&s->device);
  hipError_t e = hipSuccess;
  switch (kind) {
    case RMH_MODEL_SPHERICAL: {
      const auto* m = static_cast<const rmh_spherical_model*>(model);
      if (m->phi.size == 0 || m->theta.size == 0) {
        sensor_free(s);
        return fail(RMH_ERR_INVALID, "rmh_sensor_create: spherical model with zero size");
      }
      s->sphere = *m;
      s->width = m->theta.size;
      s->height = m->phi.size;
      s->range = m->range;
      std::vector<float2> lut;
      spherical_lut(*m, lut);
      e = hipMalloc(&s->lut, lut.size() * sizeof(float2));
      if (e == hipSuccess)
        e = hipMemcpy(s->lut, lut.data(), lut.size() * sizeof(float2), hipMemcpyHostToDevice);
      break;
    }
    case RMH_MODEL_PINHOLE: {
      const auto* m = static_cast<const rmh_pinhole_model*>(model);
      if (m->width == 0 || m->height == 0) {
        sensor_free(s);
        return fail(RMH_ERR_INVALID, "rmh_sensor_create: pinhole model with zero size");
      }
      s->pinhole = *m;
      s->width = m->width;
      s->height = m->height;
      s->range = m->range;
      break;
    }
    case RMH_MODEL_O1DN: {
      const auto* m = static_cast<const rmh_o1dn_model*>(model);
      const size_t n = static_cast<size_t>(m->width) * m->height;
      if (n == 0 || !m->dirs_dev) {
        sensor_free(s);
        return fail(RMH_ERR_INVALID, "rmh_sensor_create: O1Dn model needs width*height > 0 and dirs");
      }
      s->width = m->width;
      s->height = m->height;
      s->range = m->range;
      s->orig = m->orig;
      e = hipMalloc(&s->dirs, n * sizeof(rmh_vec3));
      if (e == hipSuccess)
        e = hipMemcpy(s->dirs, m->dirs_dev, n * sizeof(rmh_vec3), hipMemcpyDeviceToDevice);
      break;
    }
    case RMH_MODEL_ONDN: {
      const auto* m = static_cast<const rmh_ondn_model*>(model);
      const size_t n = static_cast<size_t>(m->width) * m->height;
      if (n == 0 || !m->dirs_dev || !m->origs_dev) {
        sensor_free(s);
        return fail(RMH_ERR_INVALID, "rmh_sensor_create: OnDn model needs width*height > 0, dirs and origs");
      }
      s->width = m->width;
      s->height = m->height;
      s->range = m->range;
      e = hipMalloc(&s->dirs, n * sizeof(rmh_vec3));
      if (e == hipSuccess)
        e = hipMemcpy(s->dirs, m->dirs_dev, n * sizeof(rmh_vec3), hipMemcpyDeviceToDevice);
      if (e == hipSuccess) e = hipMalloc(&s->origs, n * sizeof(rmh_vec3));
      if (e == hipSuccess)
        e = hipMemcpy(s->origs, m->origs_dev, n * sizeof(rmh_vec3), hipMemcpyDeviceToDevice);
      break;
    }
    default:
      sensor_free(s);
      return fail(RMH_ERR_INVALID, "rmh_sensor_create: unknown model kind");
  }
  if (e != hipSuccess) {
    sensor_free(s);
    return hip_fail(e, "rmh_sensor_create");
  }
  *out = s;
  return RMH_OK;
}

int rmh_sensor_destroy(rmh_sensor* sensor) {
  sensor_free(sensor);
  return RMH_OK;
}

int rmh_sensor_dims(const rmh_sensor* sensor, uint32_t* width, uint32_t* height) {
  if (!sensor) return fail(RMH_ERR_INVALID, "rmh_sensor_dims: NULL sensor");
  if (width) *width = sensor->width;
  if (height) *height = sensor->height;
  return RMH_OK;
}

}  // extern "C"
