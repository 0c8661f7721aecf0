// Ray-casting kernels of librmh: the hand-written replacement of the OptiX ray
// generation programs (src/rmagine_optix/src/simulation/optix/SphereProgramGen.cu:59-124,
// PinholeProgramGen.cu:13-77, O1DnProgramGen.cu:12-80, OnDnProgramGen.cu:12-81), of the
// hardware traversal behind optixTrace / rtcIntersect1, and of the hit/miss attribute
// writers (SensorProgramHit.cu:264-346), with the result contract of the Embree
// simulators (src/rmagine_embree/include/rmagine/simulation/SphereSimulatorEmbree.tcc:55-253,
// PinholeSimulatorEmbree.tcc, O1DnSimulatorEmbree.tcc:50-254, OnDnSimulatorEmbree.tcc).
//
// One lane per ray, wave64, one wave per workgroup.  A wave casts a 4 x 16 tile of one
// sensor image (1 x 64 for O1Dn/OnDn lists), so its rays stay close in direction.  Per lane:
//   * ray set-up in FP32 in the reference's operation order (contraction off);
//   * while-while traversal of the binary BVH with an LDS stack (one slot column per
//     lane, so the 64 lanes of a wave hit 64 different banks), near child first,
//     conservative FP32 slab tests against outward-padded boxes; speculative walks
//     (a lane holds its first leaf and walks on) for O1Dn/OnDn ray lists;
//   * watertight ray/triangle test (Woop, Benthin, Wald, JCGT 2013) evaluated in FP64:
//     the decision and the distance are exact to FP64 rounding, which makes hit masks,
//     face ids and ranges agree with an FP64 evaluation of the same predicate; ties at
//     equal distance go to the lower global face index;
//   * fused attribute epilogue, writing only the requested attributes.
//
// Two launch forms:
//   rmh_simulate        one sensor model, n poses, one workgroup per (pose, tile);
//   rmh_batch_simulate  many sensors (any mix of models, poses and ray counts) in one
//                       persistent launch (or, as a tuning choice, two: image sensors,
//                       then ray lists, each with its own walk): waves grab chunks of tiles
//                       of one sensor from an atomic counter until the work is exhausted.
//                       All pointers and sizes live in a device table built once, so the
//                       launch (counter reset kernel + persistent kernel) can be captured
//                       into a graph.
// The reference has no batched form; its closest piece is the per-launch model union
// (src/rmagine_optix/include/rmagine/simulation/optix/sim_program_data.h:32-44).
#include "rmh_trace.hpp"

#include <algorithm>
#include <cmath>
#include <new>
#include <vector>

namespace {

using namespace rmh;

// One wave per workgroup: a wave that finishes frees its LDS stack at once, so
// occupancy follows the live waves instead of the slowest wave of a larger workgroup.
constexpr int kWave = 64;

// The launch arguments of a kernel whose only argument is a T, read through the kernarg
// segment pointer (the argument starts the segment) hidden behind a barrier.  Reads made
// through it after a traversal are scalar loads again (served by the scalar cache), where
// the compiler would otherwise load the fields at the kernel's start and hold them in SGPRs
// throughout: k_simulate's arguments outgrow the SGPRs, and the excess was spilled into VGPR
// lanes, taking VGPRs from the traversal (terrain 8.80 -> 8.67 ms without them; the batched
// kernel, whose entries live in device memory, measured slower with it and does not use it).
// (The pointer is returned as a generic one; the cast from the constant address space stays
// visible to the compiler, which reads through it with scalar loads.)
template <class T>
__device__ __forceinline__ const T* kernel_args() {
  typedef const __attribute__((address_space(4))) T KArg;
  KArg* p = (KArg*)__builtin_amdgcn_kernarg_segment_ptr();
  asm volatile("" : "+s"(p));
  return (const T*)p;
}

struct SimArgs {
  MapData map;
  RayGen g;
  Tf Tsb;
  const rmh_transform* Tbm;
  uint32_t pose_offset;              // first pose of this launch
  Tiling tl;
  rmh_outputs out;
};


// Launch bounds ask for 8 waves per SIMD (64 VGPRs); with the writers' values recomputed or
// re-read after the traversal, nothing spills (62-64 VGPRs, 62 SGPRs).
template <int KIND, int STACK, bool WIDE, int SPEC>
__global__ __launch_bounds__(kWave, 8) void k_simulate(const SimArgs a) {
  __shared__ int2 stk[kLdsStack][kWave];
  // block, pose and tile indices are wave-uniform: divided on the scalar unit, they stay in
  // SGPRs through the traversal instead of taking VGPRs (the scarce registers at 8 waves
  // per SIMD; `/` would run on the vector unit and leave its quotient there)
  const uint32_t lb = xcd_contiguous(blockIdx.x, gridDim.x);
  const uint32_t pl = fast_div(lb, a.tl.div_tiles);
  const uint32_t tile = lb - pl * a.tl.tiles;
  uint32_t vid, hid;
  if (!tile_ray(a.tl, a.g.grid_w, a.g.grid_h, tile, threadIdx.x, vid, hid)) return;
  const uint32_t pose = a.pose_offset + pl;
  Hit h;
  {
    uint32_t lid;
    V3 ds, os;
    gen_ray<KIND>(a.g, vid, hid, lid, ds, os);
    const Tf Tsm = tmul(load_tf(a.Tbm[pose]), a.Tsb);  // Tsm = Tbm * Tsb (SphereSimulatorEmbree.tcc:69-70)
    const V3 dm = qrot(Tsm.R, ds);
    const V3 om = map_origin<KIND>(Tsm, os);
    trace<STACK, kWave, WIDE, SPEC>(a.map, om, dm, a.g.range_max, stk, h);
  }
  // The ray set-up is recomputed for the attribute writers instead of being held through
  // the traversal: at 64 VGPRs those values would otherwise be spilled to scratch.  The
  // launch arguments the writers use are read again (kernel_args).
  asm volatile("" : "+v"(vid), "+v"(hid));
  const SimArgs* ap = kernel_args<SimArgs>();
  uint32_t lid;
  V3 ds, os;
  gen_ray<KIND>(ap->g, vid, hid, lid, ds, os);
  const Tf Tsm = tmul(load_tf(ap->Tbm[pose]), ap->Tsb);
  const uint64_t glob = static_cast<uint64_t>(pose) * ap->g.ray_count + lid;
  write_outputs<KIND>(ap->out, glob, ap->g, ap->map, Tsm, ds, os, h);
}

// Image models walk the classic way; ray lists speculatively or the classic way, as the
// sensor is set (rmh_sensor_set_walk).
template <int KIND, int STACK, bool WIDE>
hipError_t launch_kind(const SimArgs& a, bool spec, uint32_t n_poses, hipStream_t stream) {
  if constexpr (speculative_walk(KIND)) {
    if (spec) {
      k_simulate<KIND, STACK, WIDE, kSpecOn><<<n_poses * a.tl.tiles, kWave, 0, stream>>>(a);
      return hipGetLastError();
    }
  }
  k_simulate<KIND, STACK, WIDE, kSpecOff><<<n_poses * a.tl.tiles, kWave, 0, stream>>>(a);
  return hipGetLastError();
}

template <int STACK, bool WIDE>
hipError_t launch_stack(int kind, bool spec, const SimArgs& a, uint32_t n_poses, hipStream_t stream) {
  switch (kind) {
    case RMH_MODEL_SPHERICAL: return launch_kind<RMH_MODEL_SPHERICAL, STACK, WIDE>(a, spec, n_poses, stream);
    case RMH_MODEL_PINHOLE: return launch_kind<RMH_MODEL_PINHOLE, STACK, WIDE>(a, spec, n_poses, stream);
    case RMH_MODEL_O1DN: return launch_kind<RMH_MODEL_O1DN, STACK, WIDE>(a, spec, n_poses, stream);
    default: return launch_kind<RMH_MODEL_ONDN, STACK, WIDE>(a, spec, n_poses, stream);
  }
}

hipError_t launch_sim(const TravSel& ts, int kind, bool spec, const SimArgs& a, uint32_t n_poses, hipStream_t stream) {
  if (ts.wide) {
    if (ts.stack == 32) return launch_stack<32, true>(kind, spec, a, n_poses, stream);
    if (ts.stack == 64) return launch_stack<64, true>(kind, spec, a, n_poses, stream);
    return launch_stack<128, true>(kind, spec, a, n_poses, stream);
  }
  if (ts.stack == 32) return launch_stack<32, false>(kind, spec, a, n_poses, stream);
  if (ts.stack == 64) return launch_stack<64, false>(kind, spec, a, n_poses, stream);
  return launch_stack<128, false>(kind, spec, a, n_poses, stream);
}

// ---- traversal counts (diagnostic) ------------------------------------------------------

struct VisitArgs {
  MapData map;
  RayGen g;
  Tf Tsb;
  const rmh_transform* Tbm;
  Tiling tl;
  unsigned long long* counts;  // 6 sums, see VisitCount
};

// k_simulate's rays and traversal (binary nodes, same speculation mode), counting instead
// of writing: per-lane counts are summed over the wave and added to the totals.
template <int KIND, int STACK, int SPEC>
__global__ __launch_bounds__(kWave, 8) void k_visit_stats(const VisitArgs a) {
  __shared__ int2 stk[kLdsStack][kWave];
  const uint32_t lb = xcd_contiguous(blockIdx.x, gridDim.x);
  const uint32_t pl = fast_div(lb, a.tl.div_tiles);
  const uint32_t tile = lb - pl * a.tl.tiles;
  uint32_t vid, hid;
  VisitCount vc = {0u, 0u, 0u, 0u, 0u, 0u};
  if (tile_ray(a.tl, a.g.grid_w, a.g.grid_h, tile, threadIdx.x, vid, hid)) {
    uint32_t lid;
    V3 ds, os;
    gen_ray<KIND>(a.g, vid, hid, lid, ds, os);
    const Tf Tsm = tmul(load_tf(a.Tbm[pl]), a.Tsb);
    Hit h;
    trace<STACK, kWave, false, SPEC, true>(a.map, map_origin<KIND>(Tsm, os), qrot(Tsm.R, ds),
                                                       a.g.range_max, stk, h, false, &vc);
  }
  uint32_t v[6] = {vc.vec_nodes, vc.uni_nodes, vc.leaves, vc.tris, vc.wave_iters, vc.wave_uni_iters};
  for (int k = 0; k < 6; k++)
    for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_xor(v[k], off, 64);
  if (threadIdx.x == 0)
    for (int k = 0; k < 6; k++) atomicAdd(&a.counts[k], static_cast<unsigned long long>(v[k]));
}

template <int STACK>
hipError_t launch_visit(int kind, bool spec, const VisitArgs& a, uint32_t n_poses, hipStream_t stream) {
  const uint32_t blocks = n_poses * a.tl.tiles;
  switch (kind) {
    case RMH_MODEL_SPHERICAL: k_visit_stats<RMH_MODEL_SPHERICAL, STACK, kSpecOff><<<blocks, kWave, 0, stream>>>(a); break;
    case RMH_MODEL_PINHOLE: k_visit_stats<RMH_MODEL_PINHOLE, STACK, kSpecOff><<<blocks, kWave, 0, stream>>>(a); break;
    case RMH_MODEL_O1DN:
      if (spec) k_visit_stats<RMH_MODEL_O1DN, STACK, kSpecOn><<<blocks, kWave, 0, stream>>>(a);
      else k_visit_stats<RMH_MODEL_O1DN, STACK, kSpecOff><<<blocks, kWave, 0, stream>>>(a);
      break;
    default:
      if (spec) k_visit_stats<RMH_MODEL_ONDN, STACK, kSpecOn><<<blocks, kWave, 0, stream>>>(a);
      else k_visit_stats<RMH_MODEL_ONDN, STACK, kSpecOff><<<blocks, kWave, 0, stream>>>(a);
      break;
  }
  return hipGetLastError();
}

// ---- batched, persistent ---------------------------------------------------------------

struct BatchEntryDev {
  RayGen g;
  Tf Tsb;
  const rmh_transform* Tbm;
  rmh_outputs out;
  Tiling tl;
  uint32_t n_poses;
  uint64_t n_slots;  // n_poses * tiles * 64 ray slots, enumerated tile by tile
};

// A persistent launch covers the entries [entry_lo, entry_hi), whose chunks are
// [chunk_prefix[entry_lo], chunk_prefix[entry_hi]).
struct BatchArgs {
  MapData map;
  const BatchEntryDev* entries;
  const uint32_t* chunk_prefix;  // n_entries + 1
  uint32_t entry_lo, entry_hi;
  uint32_t chunk_lo, chunk_hi;
  uint32_t chunk_rays;           // multiple of 64
  uint32_t* counter;
};

// Sensor-frame ray and map-frame ray of one slot of a batch entry.
template <int KIND>
__device__ __forceinline__ void batch_gen(const BatchEntryDev& e, uint32_t vid, uint32_t hid, uint32_t& lid,
                                          const Tf& Tsm, V3& ds, V3& os, V3& dm, V3& om) {
  gen_ray<KIND>(e.g, vid, hid, lid, ds, os);
  dm = qrot(Tsm.R, ds);
  om = map_origin<KIND>(Tsm, os);
}

template <int K>
struct KindTag {
  static constexpr int value = K;
};

// Calls fn(KindTag<kind>) for the sensor kinds a batched kernel of mode SPEC serves: the
// image models (kSpecOff), the ray-list models (kSpecOn) or all four (kSpecRuntime).
template <int SPEC, class Fn>
__device__ __forceinline__ void with_kind(int kind, Fn&& fn) {
  if constexpr (SPEC != kSpecOn) {
    if (kind == RMH_MODEL_SPHERICAL) { fn(KindTag<RMH_MODEL_SPHERICAL>{}); return; }
    if (SPEC == kSpecOff || kind == RMH_MODEL_PINHOLE) { fn(KindTag<RMH_MODEL_PINHOLE>{}); return; }
  }
  if (kind == RMH_MODEL_O1DN) { fn(KindTag<RMH_MODEL_O1DN>{}); return; }
  fn(KindTag<RMH_MODEL_ONDN>{});
}

// One batch slot.  Ray generation and attribute writing depend on the sensor kind, the
// traversal does not, so it is instantiated once per kernel (one copy of the
// octant-specialised walks keeps the kernel inside the instruction cache).  As in
// k_simulate, the ray set-up is recomputed for the writers rather than held (spilled)
// through the traversal.
template <int STACK, int BLOCK, bool WIDE, int LDS, int SPEC>
__device__ __forceinline__ void batch_ray(const BatchArgs& a, const BatchEntryDev* ep, int kind, uint32_t pose,
                                          uint32_t tile, uint32_t vid, uint32_t hid, int2 (*stk)[BLOCK]) {
  Hit h;
  {
    const BatchEntryDev& e = *ep;
    const Tf Tsm = tmul(load_tf(e.Tbm[pose]), e.Tsb);
    uint32_t lid;
    V3 ds, os, dm, om;
    with_kind<SPEC>(kind, [&](auto k) { batch_gen<decltype(k)::value>(e, vid, hid, lid, Tsm, ds, os, dm, om); });
    trace<STACK, BLOCK, WIDE, SPEC, false, LDS>(a.map, om, dm, e.g.range_max, stk, h, speculative_walk(kind));
  }
  // the entry's fields are read again after the traversal (the barrier hides that they are
  // the same), and the ray's tile coordinates recomputed from the lane: holding either
  // through the traversal spilled them to scratch once per tile
  asm volatile("" : "+s"(ep));
  uint32_t lane = threadIdx.x & 63u;
  asm volatile("" : "+v"(lane));
  const BatchEntryDev& e = *ep;
  tile_ray(e.tl, e.g.grid_w, e.g.grid_h, tile, lane, vid, hid);
  const Tf Tsm = tmul(load_tf(e.Tbm[pose]), e.Tsb);
  uint32_t lid;
  V3 ds, os;
  with_kind<SPEC>(kind, [&](auto k) { gen_ray<decltype(k)::value>(e.g, vid, hid, lid, ds, os); });
  const uint64_t r = static_cast<uint64_t>(pose) * e.g.ray_count + lid;
  with_kind<SPEC>(kind, [&](auto k) { write_outputs<decltype(k)::value>(e.out, r, e.g, a.map, Tsm, ds, os, h); });
}

__global__ void k_reset_counter(uint32_t* counter) {
  if (threadIdx.x == 0) *counter = 0u;  // a kernel store (not a memset node) keeps graph replays correct
}

// Persistent kernel over a range of entries.  SPEC = kSpecRuntime: every sensor kind in one
// launch, ray lists walking speculatively and image sensors with the leaf holding switched
// off; kSpecOff / kSpecOn: image sensors only (classic walk) / ray lists only, for a batch
// cast as two launches.
template <int STACK, int BLOCK, bool WIDE, int LDS, int SPEC>
__global__ __launch_bounds__(BLOCK, 8) void k_batch(const BatchArgs a) {
  __shared__ int2 stk[LDS][BLOCK];
  const int lane = threadIdx.x & 63;
  for (;;) {
    uint32_t c = 0;
    if (lane == 0) c = atomicAdd(a.counter, 1u);
    c = __shfl(c, 0, 64);
    c = __builtin_amdgcn_readfirstlane(c) + a.chunk_lo;
    if (c >= a.chunk_hi) break;  // every wave leaves once the counter passes the end
    uint32_t lo = a.entry_lo, hi = a.entry_hi;
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) >> 1;
      if (a.chunk_prefix[mid] <= c) lo = mid; else hi = mid;
    }
    const BatchEntryDev& e = a.entries[lo];
    const uint64_t tile0 = static_cast<uint64_t>(c - a.chunk_prefix[lo]) * (a.chunk_rays / 64);
    const uint64_t n_tiles = static_cast<uint64_t>(e.n_poses) * e.tl.tiles;
    const int kind = e.g.kind;
    for (uint32_t k = 0; k < a.chunk_rays / 64; k++) {
      if (tile0 + k >= n_tiles) break;
      const uint32_t tg = static_cast<uint32_t>(tile0 + k);  // wave-uniform; n_tiles < 2^32
      const uint32_t pose = fast_div(tg, e.tl.div_tiles);
      const uint32_t tile = tg - pose * e.tl.tiles;
      uint32_t vid, hid;
      if (!tile_ray(e.tl, e.g.grid_w, e.g.grid_h, tile, lane, vid, hid)) continue;
      batch_ray<STACK, BLOCK, WIDE, LDS, SPEC>(a, &e, kind, pose, tile, vid, hid, stk);
    }
  }
}

template <int STACK, bool WIDE, int LDS>
hipError_t launch_batch(const BatchArgs& a, int spec, uint32_t blocks, hipStream_t stream) {
  k_reset_counter<<<1, 64, 0, stream>>>(a.counter);
  if (spec == kSpecRuntime) k_batch<STACK, kWave, WIDE, LDS, kSpecRuntime><<<blocks, kWave, 0, stream>>>(a);
  else if (spec == kSpecOn) k_batch<STACK, kWave, WIDE, LDS, kSpecOn><<<blocks, kWave, 0, stream>>>(a);
  else k_batch<STACK, kWave, WIDE, LDS, kSpecOff><<<blocks, kWave, 0, stream>>>(a);
  return hipGetLastError();
}

// LDS stack depth of the batched kernel (a launch parameter the autotuner chooses): 8, 10
// or 16 entries of 8 B per lane (4, 5 or 8 KiB per wave; 16 halves the resident waves);
// deeper entries go to scratch.
template <int STACK, bool WIDE>
hipError_t launch_batch_lds(uint32_t lds, const BatchArgs& a, int spec, uint32_t blocks, hipStream_t stream) {
  if (lds == 8) return launch_batch<STACK, WIDE, 8>(a, spec, blocks, stream);
  if (lds == 16) return launch_batch<STACK, WIDE, 16>(a, spec, blocks, stream);
  return launch_batch<STACK, WIDE, 10>(a, spec, blocks, stream);
}

hipError_t launch_batch_sel(const TravSel& ts, uint32_t lds, const BatchArgs& a, int spec, uint32_t blocks,
                            hipStream_t stream) {
  if (ts.wide) {
    if (ts.stack == 32) return launch_batch_lds<32, true>(lds, a, spec, blocks, stream);
    if (ts.stack == 64) return launch_batch_lds<64, true>(lds, a, spec, blocks, stream);
    return launch_batch_lds<128, true>(lds, a, spec, blocks, stream);
  }
  if (ts.stack == 32) return launch_batch_lds<32, false>(lds, a, spec, blocks, stream);
  if (ts.stack == 64) return launch_batch_lds<64, false>(lds, a, spec, blocks, stream);
  return launch_batch_lds<128, false>(lds, a, spec, blocks, stream);
}

bool valid_lds_stack(uint32_t v) { return v == 8 || v == 10 || v == 16; }

}  // namespace

struct rmh_batch {
  const rmh_map* map = nullptr;
  BatchEntryDev* entries = nullptr;
  uint32_t* chunk_prefix = nullptr;
  uint32_t* counter = nullptr;        // one work counter per launch group
  uint32_t n_entries = 0;
  uint32_t n_image_entries = 0;       // the table holds the image entries first, then the ray lists
  uint32_t image_chunks = 0;
  uint32_t total_chunks = 0;
  uint64_t total_rays = 0;
  rmh_batch_params params{};
  uint32_t n_cu = 256;
  std::vector<uint64_t> entry_rays;
  std::vector<BatchEntryDev> host;     // host copy of the table (re-tiled by set_params)
  std::vector<const rmh_sensor*> sensors;
};

namespace {

void batch_free(rmh_batch* b) {
  if (!b) return;
  if (b->entries) (void)hipFree(b->entries);
  if (b->chunk_prefix) (void)hipFree(b->chunk_prefix);
  if (b->counter) (void)hipFree(b->counter);
  delete b;
}

bool valid_tile_rows(uint32_t r) { return r >= 1 && r <= 64 && (r & (r - 1)) == 0; }

// Tiling, slot and ray counts of every entry for the batch's tile_rows.
void batch_tile(rmh_batch* b) {
  b->total_rays = 0;
  for (uint32_t i = 0; i < b->n_entries; i++) {
    BatchEntryDev& h = b->host[i];
    h.tl = sensor_tiling(b->sensors[i], b->params.tile_rows);
    h.n_slots = static_cast<uint64_t>(h.n_poses) * h.tl.tiles * 64u;
    b->entry_rays[i] = h.n_slots;
    b->total_rays += static_cast<uint64_t>(h.n_poses) * h.g.ray_count;
  }
}

int batch_upload(rmh_batch* b, const std::vector<BatchEntryDev>& host) {
  std::vector<uint32_t> prefix(b->n_entries + 1, 0);
  uint64_t chunks = 0;
  for (uint32_t i = 0; i < b->n_entries; i++) {
    // the kernel enumerates an entry's tiles with 32-bit indices
    if (b->entry_rays[i] / 64u >= (1ull << 32)) return fail(RMH_ERR_UNSUPPORTED, "rmh_batch: an entry holds 2^32 or more ray tiles");
    prefix[i] = static_cast<uint32_t>(chunks);
    chunks += (b->entry_rays[i] + b->params.chunk_rays - 1) / b->params.chunk_rays;
    if (chunks >= (1ull << 31)) return fail(RMH_ERR_UNSUPPORTED, "rmh_batch: too many ray chunks");
  }
  prefix[b->n_entries] = static_cast<uint32_t>(chunks);
  b->total_chunks = static_cast<uint32_t>(chunks);
  b->image_chunks = prefix[b->n_image_entries];
  RMH_HIP_TRY(hipMemcpy(b->chunk_prefix, prefix.data(), sizeof(uint32_t) * prefix.size(), hipMemcpyHostToDevice));
  if (!host.empty())
    RMH_HIP_TRY(hipMemcpy(b->entries, host.data(), sizeof(BatchEntryDev) * host.size(), hipMemcpyHostToDevice));
  return RMH_OK;
}

int check_depth(const rmh_map* map) {
  if (trav_select(map).stack < 0) return fail(RMH_ERR_UNSUPPORTED, "BVH deeper than 128 levels");
  return RMH_OK;
}

}  // namespace

extern "C" {

int rmh_simulate(const rmh_map* map, const rmh_sensor* sensor, const rmh_transform* Tsb, const rmh_transform* Tbm,
                 uint32_t n_poses, const rmh_outputs* out, void* stream_) {
  // SphereSimulatorOptix.tcc:49-53 throws when no map is set
  if (!map) return fail(RMH_ERR_NOMAP, "rmh_simulate: no map set");
  if (!sensor) return fail(RMH_ERR_INVALID, "rmh_simulate: no sensor model set");
  if (!out) return fail(RMH_ERR_INVALID, "rmh_simulate: outputs is NULL");
  if (n_poses == 0) return RMH_OK;
  if (!Tbm) return fail(RMH_ERR_INVALID, "rmh_simulate: Tbm is NULL");
  const uint64_t ray_count = static_cast<uint64_t>(sensor->width) * sensor->height;
  if (ray_count == 0) return RMH_OK;
  if (ray_count > (1ull << 31)) return fail(RMH_ERR_UNSUPPORTED, "rmh_simulate: more than 2^31 rays per pose");
  if (int rc = check_depth(map)) return rc;
  const TravSel ts = trav_select(map);

  SimArgs a;
  a.map = make_mapdata(map);
  a.g = make_raygen(sensor);
  rmh_transform I{};
  I.R.w = 1.0f;
  a.Tsb = load_tf(Tsb ? *Tsb : I);
  a.Tbm = Tbm;
  a.out = *out;
  hipStream_t stream = static_cast<hipStream_t>(stream_);

  a.tl = sensor_tiling(sensor);
  const uint32_t poses_per_launch = static_cast<uint32_t>(std::max<uint64_t>(1, kMaxTileBlocks / a.tl.tiles));
  for (uint32_t p0 = 0; p0 < n_poses; p0 += poses_per_launch) {
    const uint32_t np = std::min<uint32_t>(poses_per_launch, n_poses - p0);
    a.pose_offset = p0;
    const hipError_t e = launch_sim(ts, sensor->kind, sensor->speculative != 0, a, np, stream);
    if (e != hipSuccess) return hip_fail(e, "rmh_simulate launch");
  }
  return RMH_OK;
}

int rmh_simulate_visit_stats(const rmh_map* map, const rmh_sensor* sensor, const rmh_transform* Tsb,
                             const rmh_transform* Tbm, uint32_t n_poses, uint64_t* counts, void* stream_) {
  if (!map) return fail(RMH_ERR_NOMAP, "rmh_simulate_visit_stats: no map set");
  if (!sensor || !counts) return fail(RMH_ERR_INVALID, "rmh_simulate_visit_stats: NULL argument");
  for (int k = 0; k < 6; k++) counts[k] = 0;
  if (n_poses == 0) return RMH_OK;
  if (!Tbm) return fail(RMH_ERR_INVALID, "rmh_simulate_visit_stats: Tbm is NULL");
  const uint64_t ray_count = static_cast<uint64_t>(sensor->width) * sensor->height;
  if (ray_count == 0) return RMH_OK;
  const uint32_t depth = map->info.max_depth;
  if (depth > 128) return fail(RMH_ERR_UNSUPPORTED, "rmh_simulate_visit_stats: BVH deeper than 128 levels");
  VisitArgs a;
  a.map = make_mapdata(map);
  a.g = make_raygen(sensor);
  rmh_transform I{};
  I.R.w = 1.0f;
  a.Tsb = load_tf(Tsb ? *Tsb : I);
  a.Tbm = Tbm;
  a.tl = sensor_tiling(sensor);
  if (static_cast<uint64_t>(n_poses) * a.tl.tiles > kMaxTileBlocks)
    return fail(RMH_ERR_UNSUPPORTED, "rmh_simulate_visit_stats: too many rays for one launch");
  unsigned long long* dev = nullptr;
  RMH_HIP_TRY(hipMalloc(&dev, sizeof(unsigned long long) * 6));
  hipStream_t stream = static_cast<hipStream_t>(stream_);
  hipError_t e = hipMemsetAsync(dev, 0, sizeof(unsigned long long) * 6, stream);
  a.counts = dev;
  if (e == hipSuccess) {
    if (depth <= 32) e = launch_visit<32>(sensor->kind, sensor->speculative != 0, a, n_poses, stream);
    else if (depth <= 64) e = launch_visit<64>(sensor->kind, sensor->speculative != 0, a, n_poses, stream);
    else e = launch_visit<128>(sensor->kind, sensor->speculative != 0, a, n_poses, stream);
  }
  unsigned long long h[6] = {0, 0, 0, 0, 0, 0};
  if (e == hipSuccess) e = hipMemcpyAsync(h, dev, sizeof(h), hipMemcpyDeviceToHost, stream);
  if (e == hipSuccess) e = hipStreamSynchronize(stream);
  (void)hipFree(dev);
  if (e != hipSuccess) return hip_fail(e, "rmh_simulate_visit_stats");
  for (int k = 0; k < 6; k++) counts[k] = h[k];
  return RMH_OK;
}

int rmh_batch_create(const rmh_map* map, const rmh_batch_entry* entries, uint32_t n_entries,
                     const rmh_batch_params* params, rmh_batch** out) {
  if (!out) return fail(RMH_ERR_INVALID, "rmh_batch_create: out is NULL");
  *out = nullptr;
  if (!map) return fail(RMH_ERR_NOMAP, "rmh_batch_create: no map set");
  if (n_entries > 0 && !entries) return fail(RMH_ERR_INVALID, "rmh_batch_create: entries is NULL");
  if (int rc = check_depth(map)) return rc;
  rmh_batch* b = new (std::nothrow) rmh_batch();
  if (!b) return fail(RMH_ERR_ALLOC, "rmh_batch_create: out of host memory");
  b->map = map;
  b->n_entries = n_entries;
  b->params.chunk_rays = 256;
  b->params.waves_per_cu = 32;
  b->params.tile_rows = kDefaultTileRows;
  b->params.lds_stack = kLdsStack;
  b->params.launches = 1;
  if (params) {
    if (params->chunk_rays) b->params.chunk_rays = params->chunk_rays;
    if (params->waves_per_cu) b->params.waves_per_cu = params->waves_per_cu;
    if (params->tile_rows) b->params.tile_rows = params->tile_rows;
    if (params->lds_stack) b->params.lds_stack = params->lds_stack;
    if (params->launches) b->params.launches = params->launches;
  }
  if (b->params.chunk_rays % 64 != 0 || b->params.chunk_rays > (1u << 20)) {
    batch_free(b);
    return fail(RMH_ERR_INVALID, "rmh_batch_create: chunk_rays must be a multiple of 64 (<= 2^20)");
  }
  if (b->params.waves_per_cu > 64 || !valid_tile_rows(b->params.tile_rows)) {
    batch_free(b);
    return fail(RMH_ERR_INVALID, "rmh_batch_create: waves_per_cu must be <= 64 and tile_rows a power of two <= 64");
  }
  if (!valid_lds_stack(b->params.lds_stack)) {
    batch_free(b);
    return fail(RMH_ERR_INVALID, "rmh_batch_create: lds_stack must be 8, 10 or 16");
  }
  if (b->params.launches != 1 && b->params.launches != 2) {
    batch_free(b);
    return fail(RMH_ERR_INVALID, "rmh_batch_create: launches must be 1 or 2");
  }
  int dev = 0, ncu = 0;
  (void)hipGetDevice(&dev);
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && ncu > 0)
    b->n_cu = static_cast<uint32_t>(ncu);
  b->host.resize(n_entries);
  b->sensors.resize(n_entries);
  b->entry_rays.resize(n_entries);
  // device table order: image sensors first, then ray lists (each entry keeps its own
  // outputs, so the order is not visible to the caller)
  std::vector<uint32_t> order;
  order.reserve(n_entries);
  for (int lists = 0; lists < 2; lists++)
    for (uint32_t i = 0; i < n_entries; i++) {
      if (!entries[i].sensor) { batch_free(b); return fail(RMH_ERR_INVALID, "rmh_batch_create: entry without sensor"); }
      if (speculative_walk(entries[i].sensor->kind) == (lists == 1)) order.push_back(i);
    }
  for (uint32_t i = 0; i < n_entries; i++)
    if (!speculative_walk(entries[i].sensor->kind)) b->n_image_entries++;
  for (uint32_t i = 0; i < n_entries; i++) {
    const rmh_batch_entry& e = entries[order[i]];
    if (e.n_poses > 0 && !e.Tbm_dev) { batch_free(b); return fail(RMH_ERR_INVALID, "rmh_batch_create: entry without poses"); }
    b->sensors[i] = e.sensor;
    b->host[i].g = make_raygen(e.sensor);
    b->host[i].Tsb = load_tf(e.Tsb);
    b->host[i].Tbm = e.Tbm_dev;
    b->host[i].out = e.out;
    b->host[i].n_poses = e.n_poses;
  }
  batch_tile(b);
  hipError_t err = hipMalloc(&b->entries, sizeof(BatchEntryDev) * std::max<uint32_t>(1, n_entries));
  if (err == hipSuccess) err = hipMalloc(&b->chunk_prefix, sizeof(uint32_t) * (n_entries + 1));
  if (err == hipSuccess) err = hipMalloc(&b->counter, 2 * sizeof(uint32_t));
  if (err != hipSuccess) { batch_free(b); return hip_fail(err, "rmh_batch_create"); }
  if (int rc = batch_upload(b, b->host)) { batch_free(b); return rc; }
  *out = b;
  return RMH_OK;
}

int rmh_batch_set_params(rmh_batch* b, const rmh_batch_params* params) {
  if (!b || !params) return fail(RMH_ERR_INVALID, "rmh_batch_set_params: NULL argument");
  rmh_batch_params p = b->params;
  if (params->chunk_rays) p.chunk_rays = params->chunk_rays;
  if (params->waves_per_cu) p.waves_per_cu = params->waves_per_cu;
  if (params->tile_rows) p.tile_rows = params->tile_rows;
  if (params->lds_stack) p.lds_stack = params->lds_stack;
  if (params->launches) p.launches = params->launches;
  if (!valid_lds_stack(p.lds_stack))
    return fail(RMH_ERR_INVALID, "rmh_batch_set_params: lds_stack must be 8, 10 or 16");
  if (p.launches != 1 && p.launches != 2)
    return fail(RMH_ERR_INVALID, "rmh_batch_set_params: launches must be 1 or 2");
  if (p.chunk_rays % 64 != 0 || p.chunk_rays > (1u << 20))
    return fail(RMH_ERR_INVALID, "rmh_batch_set_params: chunk_rays must be a multiple of 64 (<= 2^20)");
  if (p.waves_per_cu > 64) return fail(RMH_ERR_INVALID, "rmh_batch_set_params: waves_per_cu must be <= 64");
  if (!valid_tile_rows(p.tile_rows))
    return fail(RMH_ERR_INVALID, "rmh_batch_set_params: tile_rows must be a power of two <= 64");
  const bool retile = p.tile_rows != b->params.tile_rows;
  const bool rechunk = retile || p.chunk_rays != b->params.chunk_rays;
  b->params = p;
  if (rechunk) {
    RMH_HIP_TRY(hipDeviceSynchronize());  // the tables may be in use by a running launch
    if (retile) {
      batch_tile(b);
      return batch_upload(b, b->host);
    }
    return batch_upload(b, {});
  }
  return RMH_OK;
}

int rmh_batch_get_info(const rmh_batch* b, uint64_t* total_rays, uint32_t* total_chunks, rmh_batch_params* params) {
  if (!b) return fail(RMH_ERR_INVALID, "rmh_batch_get_info: NULL batch");
  if (total_rays) *total_rays = b->total_rays;
  if (total_chunks) *total_chunks = b->total_chunks;
  if (params) *params = b->params;
  return RMH_OK;
}

int rmh_batch_simulate(rmh_batch* b, void* stream_) {
  if (!b) return fail(RMH_ERR_INVALID, "rmh_batch_simulate: NULL batch");
  if (b->total_chunks == 0) return RMH_OK;
  const TravSel ts = trav_select(b->map);
  if (ts.stack < 0) return fail(RMH_ERR_UNSUPPORTED, "rmh_batch_simulate: BVH deeper than 128 levels");
  hipStream_t stream = static_cast<hipStream_t>(stream_);
  const uint32_t waves = b->n_cu * b->params.waves_per_cu;
  BatchArgs a;
  a.map = make_mapdata(b->map);
  a.entries = b->entries;
  a.chunk_prefix = b->chunk_prefix;
  a.chunk_rays = b->params.chunk_rays;
  if (b->params.launches == 1) {
    // one persistent launch over every entry (image entries first, then ray lists)
    a.entry_lo = 0;
    a.entry_hi = b->n_entries;
    a.chunk_lo = 0;
    a.chunk_hi = b->total_chunks;
    a.counter = b->counter;
    const uint32_t blocks = std::max<uint32_t>(1, std::min<uint32_t>(waves, b->total_chunks));
    const hipError_t e = launch_batch_sel(ts, b->params.lds_stack, a, kSpecRuntime, blocks, stream);
    if (e != hipSuccess) return hip_fail(e, "rmh_batch_simulate launch");
    return RMH_OK;
  }
  // two persistent launches in stream order: the image entries, then the ray lists
  for (int lists = 0; lists < 2; lists++) {
    a.entry_lo = lists ? b->n_image_entries : 0u;
    a.entry_hi = lists ? b->n_entries : b->n_image_entries;
    a.chunk_lo = lists ? b->image_chunks : 0u;
    a.chunk_hi = lists ? b->total_chunks : b->image_chunks;
    a.counter = b->counter + lists;
    if (a.chunk_hi == a.chunk_lo) continue;
    const uint32_t blocks = std::max<uint32_t>(1, std::min<uint32_t>(waves, a.chunk_hi - a.chunk_lo));
    const hipError_t e = launch_batch_sel(ts, b->params.lds_stack, a, lists ? kSpecOn : kSpecOff, blocks, stream);
    if (e != hipSuccess) return hip_fail(e, "rmh_batch_simulate launch");
  }
  return RMH_OK;
}

int rmh_batch_destroy(rmh_batch* b) {
  batch_free(b);
  return RMH_OK;
}

// ---- graphs ------------------------------------------------------------------------------

struct rmh_graph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

int rmh_graph_begin_capture(void* stream) {
  if (!stream) return fail(RMH_ERR_INVALID, "rmh_graph_begin_capture: the legacy default stream cannot be captured");
  RMH_HIP_TRY(hipStreamBeginCapture(static_cast<hipStream_t>(stream), hipStreamCaptureModeThreadLocal));
  return RMH_OK;
}

int rmh_graph_end_capture(void* stream, rmh_graph** out) {
  if (!stream || !out) return fail(RMH_ERR_INVALID, "rmh_graph_end_capture: NULL argument");
  *out = nullptr;
  rmh_graph* g = new (std::nothrow) rmh_graph();
  if (!g) return fail(RMH_ERR_ALLOC, "rmh_graph_end_capture: out of host memory");
  hipError_t e = hipStreamEndCapture(static_cast<hipStream_t>(stream), &g->graph);
  if (e == hipSuccess) e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    if (g->graph) (void)hipGraphDestroy(g->graph);
    delete g;
    return hip_fail(e, "rmh_graph_end_capture");
  }
  *out = g;
  return RMH_OK;
}

int rmh_graph_launch(rmh_graph* g, void* stream) {
  if (!g) return fail(RMH_ERR_INVALID, "rmh_graph_launch: NULL graph");
  RMH_HIP_TRY(hipGraphLaunch(g->exec, static_cast<hipStream_t>(stream)));
  return RMH_OK;
}

int rmh_graph_destroy(rmh_graph* g) {
  if (!g) return RMH_OK;
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  delete g;
  return RMH_OK;
}

}  // extern "C"
