// Runtime part of librmh: thread-local errors, device memory, streams and sensors.
//
// Memory entry points play the role of Rmagine's VRAM_CUDA tag
// (src/rmagine_cuda/src/types/MemoryCuda.cu:13-89): blocking copies, plain device
// allocations.  Sensors are the device-side copy of a sensor model that the reference's
// setModel() uploads (src/rmagine_optix/src/simulation/SphereSimulatorOptix.cpp:20-60,
// O1DnSimulatorOptix.cpp:41-83).
#include "rmh_grid_order.hpp"
#include "rmh_internal.hpp"

#include <rocprim/device/device_radix_sort.hpp>

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
  if (s->order) (void)hipFree(s->order);
  if (s->dirs_list) (void)hipFree(s->dirs_list);
  if (s->origs_list) (void)hipFree(s->origs_list);
  if (s->perm) (void)hipFree(s->perm);
  delete s;
}

// ---- traversal order of O1Dn / OnDn rays -------------------------------------------------
// An O1Dn/OnDn model is an arbitrary list of rays; in list order the 64 rays of a wave can
// point anywhere, and their traversals share almost nothing.  The sensor therefore keeps
// its rays sorted by direction (2D Morton code of the octahedral map of the direction,
// 16 bits per axis) and the kernels trace in that order, writing each result at the
// ray's own buffer index.  Results do not depend on the order.

__device__ inline uint32_t spread16(uint32_t x) {
  x &= 0xFFFFu;
  x = (x | (x << 8)) & 0x00FF00FFu;
  x = (x | (x << 4)) & 0x0F0F0F0Fu;
  x = (x | (x << 2)) & 0x33333333u;
  x = (x | (x << 1)) & 0x55555555u;
  return x;
}

__global__ void k_dir_keys(const rmh_vec3* __restrict__ d, uint32_t n, uint32_t* __restrict__ keys,
                           uint32_t* __restrict__ idx) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const rmh_vec3 v = d[i];
  const float l1 = fabsf(v.x) + fabsf(v.y) + fabsf(v.z);
  float u = 0.0f, w = 0.0f;
  if (l1 > 0.0f) {  // octahedral map: the unit sphere onto the square [-1, 1]^2
    u = v.x / l1;
    w = v.y / l1;
    if (v.z < 0.0f) {
      const float uu = (1.0f - fabsf(w)) * (u < 0.0f ? -1.0f : 1.0f);
      w = (1.0f - fabsf(u)) * (w < 0.0f ? -1.0f : 1.0f);
      u = uu;
    }
  }
  // NaN-safe clamp (fminf/fmaxf return the number) and 16-bit quantisation
  const float qu = fminf(fmaxf((u + 1.0f) * 32767.5f, 0.0f), 65535.0f);
  const float qw = fminf(fmaxf((w + 1.0f) * 32767.5f, 0.0f), 65535.0f);
  keys[i] = spread16(static_cast<uint32_t>(qu)) | (spread16(static_cast<uint32_t>(qw)) << 1);
  idx[i] = i;
}

__global__ void k_iota(uint32_t* __restrict__ p, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = i;
}

__global__ void k_gather_rays(const uint32_t* __restrict__ order, uint32_t n, const rmh_vec3* __restrict__ d_in,
                              const rmh_vec3* __restrict__ o_in, rmh_vec3* __restrict__ d_out,
                              rmh_vec3* __restrict__ o_out) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const uint32_t i = order[s];
  d_out[s] = d_in[i];
  if (o_in) o_out[s] = o_in[i];
}

// Uploads n rays (device pointers from the caller) into s->dirs / s->origs in direction
// order and fills s->order.  Blocking (sensor creation is not on a timed path).
static hipError_t upload_sorted_rays(rmh_sensor* s, const rmh_vec3* dirs_in, const rmh_vec3* origs_in, uint32_t n) {
  uint32_t *keys = nullptr, *keys2 = nullptr, *idx = nullptr;
  void* tmp = nullptr;
  size_t tmp_bytes = 0;
  const uint32_t blocks = (n + 255) / 256;
  hipError_t e = hipMalloc(&s->dirs, n * sizeof(rmh_vec3));
  if (e == hipSuccess && origs_in) e = hipMalloc(&s->origs, n * sizeof(rmh_vec3));
  if (e == hipSuccess) e = hipMalloc(&s->order, n * sizeof(uint32_t));
  if (e == hipSuccess) e = hipMalloc(&s->perm, n * sizeof(uint32_t));
  if (e == hipSuccess) e = hipMalloc(&s->dirs_list, n * sizeof(rmh_vec3));
  if (e == hipSuccess && origs_in) e = hipMalloc(&s->origs_list, n * sizeof(rmh_vec3));
  if (e == hipSuccess) e = hipMemcpy(s->dirs_list, dirs_in, n * sizeof(rmh_vec3), hipMemcpyDeviceToDevice);
  if (e == hipSuccess && origs_in)
    e = hipMemcpy(s->origs_list, origs_in, n * sizeof(rmh_vec3), hipMemcpyDeviceToDevice);
  if (e == hipSuccess) e = hipMalloc(&keys, n * sizeof(uint32_t));
  if (e == hipSuccess) e = hipMalloc(&keys2, n * sizeof(uint32_t));
  if (e == hipSuccess) e = hipMalloc(&idx, n * sizeof(uint32_t));
  if (e == hipSuccess) {
    k_dir_keys<<<blocks, 256>>>(dirs_in, n, keys, idx);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = rocprim::radix_sort_pairs(nullptr, tmp_bytes, keys, keys2, idx, s->order, n);
  if (e == hipSuccess) e = hipMalloc(&tmp, tmp_bytes > 0 ? tmp_bytes : 1);
  if (e == hipSuccess) e = rocprim::radix_sort_pairs(tmp, tmp_bytes, keys, keys2, idx, s->order, n);
  if (e == hipSuccess) e = hipMemcpy(s->perm, s->order, n * sizeof(uint32_t), hipMemcpyDeviceToDevice);
  if (e == hipSuccess) {
    k_gather_rays<<<blocks, 256>>>(s->order, n, dirs_in, origs_in, s->dirs, s->origs);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipDeviceSynchronize();
  s->sorted = 1;
  if (keys) (void)hipFree(keys);
  if (keys2) (void)hipFree(keys2);
  if (idx) (void)hipFree(idx);
  if (tmp) (void)hipFree(tmp);
  return e;
}

// Rewrites the traversal arrays in place (same buffers, so batch tables and captured
// graphs that point at them stay valid): list order, direction order, or the grid's tiles.
int rmh_sensor_set_ray_order(rmh_sensor* s, int order) {
  if (!s) return fail(RMH_ERR_INVALID, "rmh_sensor_set_ray_order: NULL sensor");
  if (order != RMH_RAY_ORDER_LIST && order != RMH_RAY_ORDER_SORTED && order != RMH_RAY_ORDER_TILES)
    return fail(RMH_ERR_INVALID, "rmh_sensor_set_ray_order: order must be 0 (list), 1 (sorted) or 2 (tiles)");
  if (s->kind != RMH_MODEL_O1DN && s->kind != RMH_MODEL_ONDN) return RMH_OK;  // images: no list
  if (order == s->sorted) return RMH_OK;
  const uint32_t n = s->width * s->height;
  const uint32_t blocks = (n + 255) / 256;
  RMH_HIP_TRY(hipDeviceSynchronize());  // no launch may read the arrays meanwhile
  if (order == RMH_RAY_ORDER_SORTED) {
    RMH_HIP_TRY(hipMemcpy(s->order, s->perm, n * sizeof(uint32_t), hipMemcpyDeviceToDevice));
  } else if (order == RMH_RAY_ORDER_LIST) {
    k_iota<<<blocks, 256>>>(s->order, n);
    RMH_HIP_TRY(hipGetLastError());
  } else {
    const std::vector<uint32_t> t = rmh::grid_tile_order(s->width, s->height);
    RMH_HIP_TRY(hipMemcpy(s->order, t.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice));
  }
  k_gather_rays<<<blocks, 256>>>(s->order, n, s->dirs_list, s->origs_list, s->dirs, s->origs);
  RMH_HIP_TRY(hipGetLastError());
  RMH_HIP_TRY(hipDeviceSynchronize());
  s->sorted = order;
  return RMH_OK;
}

int rmh_sensor_get_ray_order(const rmh_sensor* s, int* order) {
  if (!s || !order) return fail(RMH_ERR_INVALID, "rmh_sensor_get_ray_order: NULL argument");
  *order = (s->kind == RMH_MODEL_O1DN || s->kind == RMH_MODEL_ONDN) ? s->sorted : RMH_RAY_ORDER_LIST;
  return RMH_OK;
}

// A flag read when rmh_simulate launches: no device work, nothing to wait for.
int rmh_sensor_set_walk(rmh_sensor* s, int speculative) {
  if (!s) return fail(RMH_ERR_INVALID, "rmh_sensor_set_walk: NULL sensor");
  if (s->kind != RMH_MODEL_O1DN && s->kind != RMH_MODEL_ONDN) return RMH_OK;  // images: classic walk
  s->speculative = speculative ? 1 : 0;
  return RMH_OK;
}

int rmh_sensor_get_walk(const rmh_sensor* s, int* speculative) {
  if (!s || !speculative) return fail(RMH_ERR_INVALID, "rmh_sensor_get_walk: NULL argument");
  *speculative = (s->kind == RMH_MODEL_O1DN || s->kind == RMH_MODEL_ONDN) ? s->speculative : 0;
  return RMH_OK;
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
      if (n == 0 || n > 0xFFFFFFFFu || !m->dirs_dev) {
        sensor_free(s);
        return fail(RMH_ERR_INVALID, "rmh_sensor_create: O1Dn model needs 0 < width*height < 2^32 and dirs");
      }
      s->width = m->width;
      s->height = m->height;
      s->range = m->range;
      s->orig = m->orig;
      e = upload_sorted_rays(s, m->dirs_dev, nullptr, static_cast<uint32_t>(n));
      break;
    }
    case RMH_MODEL_ONDN: {
      const auto* m = static_cast<const rmh_ondn_model*>(model);
      const size_t n = static_cast<size_t>(m->width) * m->height;
      if (n == 0 || n > 0xFFFFFFFFu || !m->dirs_dev || !m->origs_dev) {
        sensor_free(s);
        return fail(RMH_ERR_INVALID, "rmh_sensor_create: OnDn model needs 0 < width*height < 2^32, dirs and origs");
      }
      s->width = m->width;
      s->height = m->height;
      s->range = m->range;
      e = upload_sorted_rays(s, m->dirs_dev, m->origs_dev, static_cast<uint32_t>(n));
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
  // A ray list shaped as a scanner's grid (at least one 4 x 16 tile) is cast like an image
  // by default: in the grid's tiles with the classic walk (an Ouster-style 128 x 1024 OnDn
  // list on the terrain: 9.3 ms instead of 15.4 sorted and speculative).  Lists of one or a
  // few rows keep the direction order and the speculative walk that scattered rays need.
  if ((s->kind == RMH_MODEL_O1DN || s->kind == RMH_MODEL_ONDN) && s->height >= rmh::kGridTileRows &&
      s->width >= rmh::kGridTileCols) {
    if (int rc = rmh_sensor_set_ray_order(s, RMH_RAY_ORDER_TILES)) {
      sensor_free(s);
      return rc;
    }
    s->speculative = 0;
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
