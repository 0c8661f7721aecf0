// Pose adjoint of librmh: gradients of a Huber range-residual loss with respect to 6-DoF
// sensor poses, through the hit triangle's plane.  The reference has no differentiable
// path; the closest pieces are its point-to-plane Jacobian
// (src/rmagine_core/include/rmagine/math/optimization.h:113-137) and the SE(3)
// exponential / SO(3) left Jacobian (src/rmagine_core/include/rmagine/math/lie.tcc:252-277).
//
// For pose p (base -> map transform Tbm = (R, t)) and ray k the forward pass is the same
// closest-hit query as simulate() (FP64 distance r to the hit triangle with unit world
// normal n, ray origin o and direction d in the map frame).  With a target range r*_k the
// loss is L_p = sum_k rho(r_k - r*_k), rho the Huber function with threshold delta, over
// rays that hit (r >= range.min) and whose target is valid.  The pose is perturbed on the
// left in the map frame, t -> t + tau, R -> Exp(omega) R; then
//   dr/dtau   = g,                      g = -n / (n . d)
//   dr/domega = (o + r d - t) x g       (lever arm from the base origin to the hit point)
// and dL/d(tau, omega) = sum_k rho'(e_k) dr_k/d(tau, omega).  Per ray in FP64, reduced per
// pose in FP64 (wave shuffles within each 64-ray tile, then the tiles of a pose in a fixed
// order).
//
// Output per pose: 8 doubles {L, dL/dtau (3), dL/domega (3), number of rays used}.
#include "rmh_trace.hpp"

#include <algorithm>

namespace {

using namespace rmh;

struct GradArgs {
  MapData map;
  RayGen g;
  Tf Tsb;
  const rmh_transform* Tbm;
  uint32_t pose_offset;
  const float* target;
  const uint8_t* target_mask;
  uint32_t target_stride;     // 0: one target scan shared by all poses; else ray_count
  double delta;
  Tiling tl;                  // one 64-ray tile per workgroup; chunks = tiles per pose
  double* partial;            // (pose, tile) -> 8 doubles
};

constexpr int kWave = 64;

template <int KIND, int STACK, bool WIDE>
__global__ __launch_bounds__(kWave, 8) void k_pose_grad(const GradArgs a) {
  __shared__ int2 stk[kLdsStack][kWave];
  const uint32_t pose = a.pose_offset + blockIdx.y;
  uint32_t vid, hid, lid;
  const bool active = tile_ray(a.tl, a.g.grid_w, a.g.grid_h, blockIdx.x, threadIdx.x, vid, hid);
  double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  if (active) {
    V3 ds, os;
    gen_ray<KIND>(a.g, vid, hid, lid, ds, os);
    const Tf Tbm = load_tf(a.Tbm[pose]);
    const Tf Tsm = tmul(Tbm, a.Tsb);
    const V3 dm = qrot(Tsm.R, ds);
    const V3 om = map_origin<KIND>(Tsm, os);
    Hit h;
    trace<STACK, kWave, WIDE, spec_mode<KIND>()>(a.map, om, dm, a.g.range_max, stk, h);
    const size_t ti = static_cast<size_t>(pose) * a.target_stride + lid;
    const float tr = a.target[ti];
    const bool tvalid = a.target_mask ? (a.target_mask[ti] != 0) : (tr >= a.g.range_min && tr <= a.g.range_max);
    if (h.slot != 0xFFFFFFFFu && static_cast<float>(h.t) >= a.g.range_min && tvalid) {
      const float4 nn = a.map.normals[h.slot];
      const double nx = nn.x, ny = nn.y, nz = nn.z;
      const double dx = dm.x, dy = dm.y, dz = dm.z;
      const double nd = nx * dx + ny * dy + nz * dz;
      if (nd != 0.0) {
        const double e = h.t - static_cast<double>(tr);
        const double ae = fabs(e);
        const double rho = ae <= a.delta ? 0.5 * e * e : a.delta * (ae - 0.5 * a.delta);
        const double drho = ae <= a.delta ? e : (e > 0 ? a.delta : -a.delta);
        const double gx = -nx / nd, gy = -ny / nd, gz = -nz / nd;
        const double lx = (double)om.x + h.t * dx - (double)Tbm.t.x;
        const double ly = (double)om.y + h.t * dy - (double)Tbm.t.y;
        const double lz = (double)om.z + h.t * dz - (double)Tbm.t.z;
        acc[0] = rho;
        acc[1] = drho * gx;
        acc[2] = drho * gy;
        acc[3] = drho * gz;
        acc[4] = drho * (ly * gz - lz * gy);
        acc[5] = drho * (lz * gx - lx * gz);
        acc[6] = drho * (lx * gy - ly * gx);
        acc[7] = 1.0;
      }
    }
  }
  for (int k = 0; k < 8; k++) {
    double x = acc[k];
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    acc[k] = x;
  }
  if (threadIdx.x == 0) {
    double* p = a.partial + (static_cast<size_t>(pose) * a.tl.tiles + blockIdx.x) * 8;
    for (int k = 0; k < 8; k++) p[k] = acc[k];
  }
}

__global__ void k_pose_grad_combine(const double* partial, uint32_t chunks, uint32_t n_poses, double* out) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n_poses * 8u) return;
  const uint32_t p = i >> 3, k = i & 7u;
  double s = 0.0;
  for (uint32_t c = 0; c < chunks; c++) s += partial[(static_cast<size_t>(p) * chunks + c) * 8 + k];
  out[i] = s;
}

template <int KIND, int STACK, bool WIDE>
hipError_t launch_grad(const GradArgs& a, uint32_t np, hipStream_t stream) {
  k_pose_grad<KIND, STACK, WIDE><<<dim3(a.tl.tiles, np), kWave, 0, stream>>>(a);
  return hipGetLastError();
}

template <int STACK, bool WIDE>
hipError_t launch_grad_stack(int kind, const GradArgs& a, uint32_t np, hipStream_t stream) {
  switch (kind) {
    case RMH_MODEL_SPHERICAL: return launch_grad<RMH_MODEL_SPHERICAL, STACK, WIDE>(a, np, stream);
    case RMH_MODEL_PINHOLE: return launch_grad<RMH_MODEL_PINHOLE, STACK, WIDE>(a, np, stream);
    case RMH_MODEL_O1DN: return launch_grad<RMH_MODEL_O1DN, STACK, WIDE>(a, np, stream);
    default: return launch_grad<RMH_MODEL_ONDN, STACK, WIDE>(a, np, stream);
  }
}

hipError_t launch_grad_sel(const TravSel& ts, int kind, const GradArgs& a, uint32_t np, hipStream_t stream) {
  if (ts.wide) {
    if (ts.stack == 32) return launch_grad_stack<32, true>(kind, a, np, stream);
    if (ts.stack == 64) return launch_grad_stack<64, true>(kind, a, np, stream);
    return launch_grad_stack<128, true>(kind, a, np, stream);
  }
  if (ts.stack == 32) return launch_grad_stack<32, false>(kind, a, np, stream);
  if (ts.stack == 64) return launch_grad_stack<64, false>(kind, a, np, stream);
  return launch_grad_stack<128, false>(kind, a, np, stream);
}

struct Scratch {
  void* p = nullptr;
  ~Scratch() { if (p) (void)hipFree(p); }
};

}  // namespace

extern "C" {

int rmh_pose_grad_workspace_size(const rmh_map* map, const rmh_sensor* sensor, uint32_t n_poses, size_t* bytes) {
  if (!map || !sensor || !bytes) return fail(RMH_ERR_INVALID, "rmh_pose_grad_workspace_size: NULL argument");
  const Tiling tl = sensor_tiling(sensor);
  *bytes = sizeof(double) * 8 * static_cast<uint64_t>(tl.tiles) * n_poses;
  return RMH_OK;
}

int rmh_pose_grad(const rmh_map* map, const rmh_sensor* sensor, const rmh_transform* Tsb, const rmh_transform* Tbm,
                  uint32_t n_poses, const float* target, const uint8_t* target_mask, int per_pose_targets,
                  float huber_delta, double* out, void* workspace, size_t workspace_bytes, void* stream_) {
  if (!map) return fail(RMH_ERR_NOMAP, "rmh_pose_grad: no map set");
  if (!sensor || !out || !target) return fail(RMH_ERR_INVALID, "rmh_pose_grad: NULL argument");
  if (n_poses == 0) return RMH_OK;
  if (n_poses > (1u << 28)) return fail(RMH_ERR_UNSUPPORTED, "rmh_pose_grad: too many poses");
  if (!Tbm) return fail(RMH_ERR_INVALID, "rmh_pose_grad: Tbm is NULL");
  if (!(huber_delta > 0.0f)) return fail(RMH_ERR_INVALID, "rmh_pose_grad: huber_delta must be > 0");
  const TravSel ts = trav_select(map);
  if (ts.stack < 0) return fail(RMH_ERR_UNSUPPORTED, "rmh_pose_grad: BVH deeper than 128 levels");
  const uint64_t rc = static_cast<uint64_t>(sensor->width) * sensor->height;
  if (rc == 0 || rc > (1ull << 30)) return fail(RMH_ERR_UNSUPPORTED, "rmh_pose_grad: bad ray count");
  GradArgs a;
  a.map = make_mapdata(map);
  a.g = make_raygen(sensor);
  rmh_transform I{};
  I.R.w = 1.0f;
  a.Tsb = load_tf(Tsb ? *Tsb : I);
  a.Tbm = Tbm;
  a.target = target;
  a.target_mask = target_mask;
  a.target_stride = per_pose_targets ? static_cast<uint32_t>(rc) : 0u;
  a.delta = huber_delta;
  a.tl = sensor_tiling(sensor);
  if (a.tl.tiles > 65535u) return fail(RMH_ERR_UNSUPPORTED, "rmh_pose_grad: more than 65535 ray tiles per pose");
  const size_t need = sizeof(double) * 8 * static_cast<size_t>(a.tl.tiles) * n_poses;
  Scratch scratch;
  if (workspace) {
    if (workspace_bytes < need) return fail(RMH_ERR_INVALID, "rmh_pose_grad: workspace too small");
    a.partial = static_cast<double*>(workspace);
  } else {
    RMH_HIP_TRY(hipMalloc(&scratch.p, need));
    a.partial = static_cast<double*>(scratch.p);
  }
  hipStream_t stream = static_cast<hipStream_t>(stream_);
  for (uint32_t p0 = 0; p0 < n_poses; p0 += 65535u) {
    const uint32_t np = std::min<uint32_t>(65535u, n_poses - p0);
    a.pose_offset = p0;
    const hipError_t e = launch_grad_sel(ts, sensor->kind, a, np, stream);
    if (e != hipSuccess) return hip_fail(e, "rmh_pose_grad launch");
  }
  k_pose_grad_combine<<<(n_poses * 8u + 255u) / 256u, 256, 0, stream>>>(a.partial, a.tl.tiles, n_poses, out);
  RMH_HIP_TRY(hipGetLastError());
  if (scratch.p) RMH_HIP_TRY(hipStreamSynchronize(stream));
  return RMH_OK;
}

}  // extern "C"
