// Closest-point queries of librmh: for a batch of query points, the nearest point of the
// map strictly within max_distance.  Replaces EmbreeMap::closestPoint / EmbreeScene::closestPoint
// (src/rmagine_embree/src/map/EmbreeMap.cpp:33-37, src/rmagine_embree/src/map/embree/EmbreeScene.cpp:292-315)
// and its point-query callback (EmbreeMesh.cpp:22-131: closest point of a triangle,
// result {d, p, n, primID, geomID}, updated while d < the query radius).
//
// One lane per query.  A bound first (one descent to the nearer child at every level and
// that leaf's triangles), then a depth-first BVH walk, nearer child first; a subtree is
// skipped when a lower bound of the squared distance from the query to its (outward padded)
// box, computed in FP32 (box_d2), exceeds the best squared distance so far.  The walk's
// stack is the ray kernels' (rmh_trace.hpp): the top rows in LDS, deeper entries in private
// memory; held in registers, the 32 entries and their distances took the kernel to 190
// VGPRs (two waves per SIMD).  An unbounded search pushes a subtree at almost every level of
// its first descent, so the stack is deeper than a ray's: 16 LDS rows (8 KB per wave; the
// kernel's registers allow five waves per SIMD, which the 160 KB of LDS of a CU then just
// holds).  The triangle's closest point (Ericson, Real-Time Collision Detection, 5.1.5:
// Voronoi regions of the vertices, edges and face) and its distance are evaluated in FP64,
// so the result is the FP64 nearest triangle; equal distances go to the lower global face
// index.  Outputs per query (each pointer may be NULL): distance (FLT_MAX when nothing lies
// within max_distance), closest point, unit face normal (not flipped), face id within its
// mesh, geom id, object id (0xFFFFFFFF for none; point and normal NaN).
//
// Lanes of a wave that walk unrelated parts of the tree fetch its nodes separately: random
// query sets fetch about 3x the node bytes of the same points in spatial order.  Batches of
// at least kSortQueries queries are therefore sorted by the Morton code of their positions in
// the map's bounding box (30-bit codes, on the device, in stream-ordered scratch), and walked
// in that order when a sample of their waves spans kSpreadRatio times less space in it than
// in the caller's order (the walk then reads queries and writes results scattered, which
// costs a walk in an already coherent order more than it gains).  Each lane reads and writes
// its own query's slots, so the results do not depend on the order.  Under stream capture,
// or when the scratch cannot be had, the batch is walked in the caller's order.  The scratch
// (17 B per query) comes from a pool of the library's own that keeps kPoolKeep bytes per
// device between calls.
#include "rmh_trace.hpp"

#include <rocprim/device/device_radix_sort.hpp>

#include <algorithm>
#include <cfloat>
#include <mutex>

namespace {

using namespace rmh;

struct D3 { double x, y, z; };

__device__ __forceinline__ D3 d3(double x, double y, double z) { D3 r; r.x = x; r.y = y; r.z = z; return r; }
__device__ __forceinline__ D3 sub(const D3& a, const D3& b) { return d3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ double ddot(const D3& a, const D3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ D3 madd(const D3& a, const D3& b, double s) { return d3(a.x + b.x * s, a.y + b.y * s, a.z + b.z * s); }

// Closest point of triangle (a, b, c) to p.
__device__ D3 closest_on_triangle(const D3& p, const D3& a, const D3& b, const D3& c) {
  const D3 ab = sub(b, a), ac = sub(c, a), ap = sub(p, a);
  const double d1 = ddot(ab, ap), d2 = ddot(ac, ap);
  if (d1 <= 0.0 && d2 <= 0.0) return a;
  const D3 bp = sub(p, b);
  const double d3v = ddot(ab, bp), d4 = ddot(ac, bp);
  if (d3v >= 0.0 && d4 <= d3v) return b;
  const double vc = d1 * d4 - d3v * d2;
  if (vc <= 0.0 && d1 >= 0.0 && d3v <= 0.0) return madd(a, ab, d1 / (d1 - d3v));
  const D3 cp = sub(p, c);
  const double d5 = ddot(ab, cp), d6 = ddot(ac, cp);
  if (d6 >= 0.0 && d5 <= d6) return c;
  const double vb = d5 * d2 - d1 * d6;
  if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0) return madd(a, ac, d2 / (d2 - d6));
  const double va = d3v * d6 - d5 * d4;
  if (va <= 0.0 && (d4 - d3v) >= 0.0 && (d5 - d6) >= 0.0)
    return madd(b, sub(c, b), (d4 - d3v) / ((d4 - d3v) + (d5 - d6)));
  const double den = 1.0 / (va + vb + vc);
  return madd(madd(a, ab, vb * den), ac, vc * den);
}

// A lower bound of the squared distance from the query (qx, qy, qz: FP32, as given) to the
// box [lo, hi], in FP32: each difference of two floats is rounded once (relative error at
// most 2^-24), the squares and the sum add three more roundings, so shrinking the result by
// 2^-20 bounds it from below; an overflow to infinity is clamped to FLT_MAX (still below any
// FP64 bound the walk can hold for such a box).  Pruning with a lower bound never drops a
// box that holds the nearest point; the triangle distances stay FP64.
__device__ __forceinline__ double box_d2(float qx, float qy, float qz, float lox, float hix, float loy, float hiy,
                                         float loz, float hiz) {
  const float dx = fmaxf(fmaxf(lox - qx, qx - hix), 0.0f);
  const float dy = fmaxf(fmaxf(loy - qy, qy - hiy), 0.0f);
  const float dz = fmaxf(fmaxf(loz - qz, qz - hiz), 0.0f);
  const float d2 = (dx * dx + dy * dy + dz * dz) * (1.0f - 0x1p-20f);
  return static_cast<double>(fminf(d2, FLT_MAX));
}

struct CPArgs {
  MapData map;
  const rmh_vec3* q;
  const uint32_t* order;  // lane i takes query order[i] when the spread test below holds
  const float* spread;    // summed wave extents of a sample: [0] caller's order, [1] sorted
  uint32_t n;
  double best0;  // max_distance^2: a result needs d^2 < best0
  rmh_closest_outputs out;
};

constexpr int kClosestLds = 16;
constexpr float kSpreadRatio = 4.0f;

template <int STACK>
__global__ __launch_bounds__(64, 4) void k_closest(const CPArgs a) {
  __shared__ int2 stk[kClosestLds][64];
  const uint32_t lane_q = blockIdx.x * 64 + threadIdx.x;
  if (lane_q >= a.n) return;
  const bool reorder = a.order && a.spread[1] * kSpreadRatio < a.spread[0];
  const uint32_t i = reorder ? a.order[lane_q] : lane_q;
  const rmh_vec3 qv = a.q[i];
  const D3 q = d3(qv.x, qv.y, qv.z);
  double best = a.best0;
  uint32_t best_slot = 0xFFFFFFFFu, best_face = 0xFFFFFFFFu;
  D3 best_p = d3(0, 0, 0);
  volatile int spill_mem[STACK > kClosestLds ? 2 * (STACK - kClosestLds) : 2];
  PrivInt* spill = (PrivInt*)spill_mem;
  int sp = 0;
  // the triangles of leaf reference `leaf` against the query
  auto leaf_test = [&](int leaf) {
    const uint32_t e = ~static_cast<uint32_t>(leaf);
    const uint32_t first = e >> 3, cnt = (e & 7u) + 1u;
    for (uint32_t k = 0; k < cnt; k++) {
      const size_t s = static_cast<size_t>(first + k) * 3;
      const float4 v0 = a.map.tris[s], v1 = a.map.tris[s + 1], v2 = a.map.tris[s + 2];
      const D3 p = closest_on_triangle(q, d3(v0.x, v0.y, v0.z), d3(v1.x, v1.y, v1.z), d3(v2.x, v2.y, v2.z));
      const D3 dv = sub(p, q);
      const double d2 = ddot(dv, dv);
      const uint32_t face = __float_as_uint(v2.w);
      if (d2 < best || (d2 == best && best_slot != 0xFFFFFFFFu && face < best_face)) {
        best = d2;
        best_slot = first + k;
        best_face = face;
        best_p = p;
      }
    }
  };
  int ref = kDone;
  if (a.map.n_faces) {
    // a bound before the walk: one descent to the nearer child at every level, then that
    // leaf's triangles.  Without it the walk starts unbounded and pushes the farther child
    // at every level of its first descent, deeper than the LDS rows hold.  The walk tests
    // that leaf again, which changes nothing (same distance, same face).
    int g = a.map.root;
    while (g >= 0) {
      const Node2 nd = a.map.nodes[g];
      const double b0 = box_d2(qv.x, qv.y, qv.z, nd.a.x, nd.a.y, nd.a.z, nd.a.w, nd.c.x, nd.c.y);
      const double b1 = box_d2(qv.x, qv.y, qv.z, nd.b.x, nd.b.y, nd.b.z, nd.b.w, nd.c.z, nd.c.w);
      g = b0 <= b1 ? nd.d.x : nd.d.y;
    }
    leaf_test(g);
    ref = a.map.root;
  }
  while (ref != kDone) {
    if (ref >= 0) {
      const Node2 nd = a.map.nodes[ref];
      const double b0 = box_d2(qv.x, qv.y, qv.z, nd.a.x, nd.a.y, nd.a.z, nd.a.w, nd.c.x, nd.c.y);
      const double b1 = box_d2(qv.x, qv.y, qv.z, nd.b.x, nd.b.y, nd.b.z, nd.b.w, nd.c.z, nd.c.w);
      const bool h0 = b0 <= best, h1 = b1 <= best;
      if (h0 && h1) {
        const bool first0 = b0 <= b1;
        // one entry per level: the host picks STACK >= tree depth; the entry's distance is
        // rounded down, so a pop never skips a box that could still hold the nearest point
        push_ref<STACK, 64, kClosestLds, true>(stk, spill, sp, first0 ? nd.d.y : nd.d.x,
                                               __double2float_rd(first0 ? b1 : b0));
        ref = first0 ? nd.d.x : nd.d.y;
        continue;
      }
      if (h0) { ref = nd.d.x; continue; }
      if (h1) { ref = nd.d.y; continue; }
    } else {
      leaf_test(ref);
    }
    // pop the next subtree that can still hold something nearer: the bound rounded up to
    // FP32 pops every entry with distance <= best (and at most a few more, whose children
    // are tested again)
    ref = pop_ref<STACK, 64, kClosestLds, true>(stk, spill, sp, __double2float_ru(best));
  }
  const rmh_closest_outputs& o = a.out;
  if (best_slot != 0xFFFFFFFFu) {
    const size_t s = static_cast<size_t>(best_slot) * 3;
    const float4 nn = a.map.normals[best_slot];
    if (o.dist) o.dist[i] = static_cast<float>(sqrt(best));
    if (o.points) { rmh_vec3 v; v.x = (float)best_p.x; v.y = (float)best_p.y; v.z = (float)best_p.z; o.points[i] = v; }
    if (o.normals) { rmh_vec3 v; v.x = nn.x; v.y = nn.y; v.z = nn.z; o.normals[i] = v; }
    if (o.face_ids) o.face_ids[i] = __float_as_uint(a.map.tris[s].w);
    if (o.geom_ids) o.geom_ids[i] = __float_as_uint(a.map.tris[s + 1].w);
    if (o.object_ids) o.object_ids[i] = __float_as_uint(nn.w);
  } else {
    const float qnan = __builtin_nanf("");
    rmh_vec3 v; v.x = qnan; v.y = qnan; v.z = qnan;
    if (o.dist) o.dist[i] = FLT_MAX;
    if (o.points) o.points[i] = v;
    if (o.normals) o.normals[i] = v;
    if (o.face_ids) o.face_ids[i] = 0xFFFFFFFFu;
    if (o.geom_ids) o.geom_ids[i] = 0xFFFFFFFFu;
    if (o.object_ids) o.object_ids[i] = 0xFFFFFFFFu;
  }
}

constexpr uint32_t kSortQueries = 1u << 16;

// Spreads the low 10 bits of v to every third bit.
__device__ __forceinline__ uint32_t spread3(uint32_t v) {
  v &= 0x3FFu;
  v = (v | (v << 16)) & 0x030000FFu;
  v = (v | (v << 8)) & 0x0300F00Fu;
  v = (v | (v << 4)) & 0x030C30C3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}

// 30-bit Morton code of each query's cell in a grid of cubic cells, 1024 along the longest
// side of the root's box (queries outside it clamp to the border cells; cubic cells keep a
// wave's queries in a compact region of a flat map, not a thin slab), and the identity
// permutation beside it.
__global__ __launch_bounds__(256) void k_query_keys(const Node2* nodes, int root, const rmh_vec3* q, uint32_t n,
                                                   uint32_t* keys, uint32_t* idx, float* spread) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i == 0) { spread[0] = 0.0f; spread[1] = 0.0f; }
  if (i >= n) return;
  const Node2 r = nodes[root];
  const float lx = fminf(r.a.x, r.b.x), hx = fmaxf(r.a.y, r.b.y);
  const float ly = fminf(r.a.z, r.b.z), hy = fmaxf(r.a.w, r.b.w);
  const float lz = fminf(r.c.x, r.c.z), hz = fmaxf(r.c.y, r.c.w);
  const float scale = 1024.0f / fmaxf(fmaxf(fmaxf(hx - lx, hy - ly), hz - lz), 1e-30f);
  const rmh_vec3 p = q[i];
  auto cell = [scale](float v, float lo) {
    return static_cast<uint32_t>(fminf(fmaxf((v - lo) * scale, 0.0f), 1023.0f));  // NaN -> 0
  };
  keys[i] = spread3(cell(p.x, lx)) | (spread3(cell(p.y, ly)) << 1) | (spread3(cell(p.z, lz)) << 2);
  idx[i] = i;
}

constexpr uint32_t kSpreadWaves = 4096;

// Extent (x + y + z) of the queries of sampled waves, in the caller's order and in the
// sorted one, summed into spread[0] and spread[1].  Wave k of the grid takes the 64 queries
// from (k * stride) * 64; lanes past n repeat query n - 1, which belongs to the same wave.
__global__ __launch_bounds__(256) void k_order_spread(const rmh_vec3* q, const uint32_t* order, uint32_t n,
                                                     uint32_t stride, float* spread) {
  __shared__ float part[2][4];
  const uint32_t wave = blockIdx.x * 4 + threadIdx.x / 64, lane = threadIdx.x & 63u;
  const uint64_t first = static_cast<uint64_t>(wave) * stride * 64;
  float ext[2] = {0.0f, 0.0f};
  if (first < n) {
    const uint32_t j = first + lane < n ? static_cast<uint32_t>(first + lane) : n - 1;
    const rmh_vec3 pts[2] = {q[j], q[order[j]]};
    for (int o = 0; o < 2; o++) {
      float lo[3] = {pts[o].x, pts[o].y, pts[o].z}, hi[3] = {pts[o].x, pts[o].y, pts[o].z};
      for (int m = 32; m >= 1; m >>= 1)
        for (int c = 0; c < 3; c++) {
          lo[c] = fminf(lo[c], __shfl_xor(lo[c], m, 64));
          hi[c] = fmaxf(hi[c], __shfl_xor(hi[c], m, 64));
        }
      ext[o] = (hi[0] - lo[0]) + (hi[1] - lo[1]) + (hi[2] - lo[2]);
      if (!(ext[o] <= FLT_MAX)) ext[o] = FLT_MAX;  // NaN or overflow
    }
  }
  if (lane == 0) { part[0][threadIdx.x / 64] = ext[0]; part[1][threadIdx.x / 64] = ext[1]; }
  __syncthreads();
  if (threadIdx.x < 2) {
    const float v = part[threadIdx.x][0] + part[threadIdx.x][1] + part[threadIdx.x][2] + part[threadIdx.x][3];
    atomicAdd(&spread[threadIdx.x], v);
  }
}

// Stream-ordered pool of the sorting scratch, one per device, created at first use.  It keeps
// up to kPoolKeep bytes after a synchronisation (the default pool would return them to the
// driver, and map them again at the next call: 0.3 ms per call for 4M queries).
constexpr uint64_t kPoolKeep = 256ull << 20;
constexpr int kMaxPoolDevices = 64;

hipMemPool_t scratch_pool(int device) {
  static std::mutex mu;
  static hipMemPool_t pools[kMaxPoolDevices] = {};
  if (device < 0 || device >= kMaxPoolDevices) return nullptr;
  std::lock_guard<std::mutex> lock(mu);
  if (!pools[device]) {
    hipMemPoolProps props = {};
    props.allocType = hipMemAllocationTypePinned;
    props.location.type = hipMemLocationTypeDevice;
    props.location.id = device;
    hipMemPool_t p = nullptr;
    if (hipMemPoolCreate(&p, &props) != hipSuccess) { (void)hipGetLastError(); return nullptr; }
    uint64_t keep = kPoolKeep;
    (void)hipMemPoolSetAttribute(p, hipMemPoolAttrReleaseThreshold, &keep);
    (void)hipGetLastError();
    pools[device] = p;
  }
  return pools[device];
}

// Morton order of the queries in `scratch` (stream-ordered; freed by the caller after the
// walk) and the spread sums that decide whether the walk takes it, or NULL when the batch is
// walked in the caller's order.
uint32_t* query_order(const rmh_map* map, const rmh_vec3* q, uint32_t n, hipStream_t s, void** scratch,
                      const float** spread_out) {
  *scratch = nullptr;
  if (n < kSortQueries || map->root < 0 || map->n_faces == 0) return nullptr;
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(s, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) {
    (void)hipGetLastError();
    return nullptr;
  }
  size_t tmp_bytes = 0;
  if (rocprim::radix_sort_pairs(nullptr, tmp_bytes, static_cast<uint32_t*>(nullptr), static_cast<uint32_t*>(nullptr),
                                static_cast<uint32_t*>(nullptr), static_cast<uint32_t*>(nullptr), n, 0u, 30u,
                                s) != hipSuccess) {
    (void)hipGetLastError();
    return nullptr;
  }
  const size_t arr = (static_cast<size_t>(n) * sizeof(uint32_t) + 255) & ~static_cast<size_t>(255);
  hipMemPool_t pool = scratch_pool(map->device);
  if (!pool || hipMallocFromPoolAsync(scratch, 4 * arr + 256 + tmp_bytes, pool, s) != hipSuccess) {
    (void)hipGetLastError();
    *scratch = nullptr;
    return nullptr;
  }
  char* base = static_cast<char*>(*scratch);
  uint32_t* keys = reinterpret_cast<uint32_t*>(base);
  uint32_t* keys2 = reinterpret_cast<uint32_t*>(base + arr);
  uint32_t* idx = reinterpret_cast<uint32_t*>(base + 2 * arr);
  uint32_t* order = reinterpret_cast<uint32_t*>(base + 3 * arr);
  float* spread = reinterpret_cast<float*>(base + 4 * arr);
  const uint32_t waves = (n + 63) / 64, sampled = std::min(waves, kSpreadWaves);
  const uint32_t stride = waves / sampled;  // >= 1; sampled * stride <= waves
  k_query_keys<<<(n + 255) / 256, 256, 0, s>>>(map->nodes, map->root, q, n, keys, idx, spread);
  bool ok = hipGetLastError() == hipSuccess &&
            rocprim::radix_sort_pairs(base + 4 * arr + 256, tmp_bytes, keys, keys2, idx, order, n, 0u, 30u, s) ==
                hipSuccess;
  if (ok) {
    k_order_spread<<<(sampled + 3) / 4, 256, 0, s>>>(q, order, n, stride, spread);
    ok = hipGetLastError() == hipSuccess;
  }
  if (!ok) {
    (void)hipGetLastError();
    (void)hipFreeAsync(*scratch, s);
    *scratch = nullptr;
    return nullptr;
  }
  *spread_out = spread;
  return order;
}

}  // namespace

extern "C" int rmh_closest_points(const rmh_map* map, const rmh_vec3* queries, uint32_t n, float max_distance,
                                  const rmh_closest_outputs* out, void* stream) {
  if (!map) return fail(RMH_ERR_NOMAP, "rmh_closest_points: no map set");
  if (!out) return fail(RMH_ERR_INVALID, "rmh_closest_points: NULL outputs");
  if (n == 0) return RMH_OK;
  if (!queries) return fail(RMH_ERR_INVALID, "rmh_closest_points: NULL queries");
  if (!(max_distance >= 0.0f)) return fail(RMH_ERR_INVALID, "rmh_closest_points: max_distance must be >= 0");
  if (map->info.max_depth > 128) return fail(RMH_ERR_UNSUPPORTED, "rmh_closest_points: BVH deeper than 128 levels");
  CPArgs a;
  a.map.tris = map->tris;
  a.map.normals = map->normals;
  a.map.nodes = map->nodes;
  a.map.root = map->root;
  a.map.n_faces = map->n_faces;
  a.q = queries;
  a.n = n;
  a.best0 = static_cast<double>(max_distance) * static_cast<double>(max_distance);
  a.out = *out;
  const uint32_t blocks = (n + 63) / 64;
  hipStream_t s = static_cast<hipStream_t>(stream);
  void* scratch = nullptr;
  a.spread = nullptr;
  a.order = query_order(map, queries, n, s, &scratch, &a.spread);
  const uint32_t depth = map->info.max_depth;
  if (depth <= 32) k_closest<32><<<blocks, 64, 0, s>>>(a);
  else if (depth <= 64) k_closest<64><<<blocks, 64, 0, s>>>(a);
  else k_closest<128><<<blocks, 64, 0, s>>>(a);
  const hipError_t e = hipGetLastError();
  if (scratch) RMH_HIP_TRY(hipFreeAsync(scratch, s));
  RMH_HIP_TRY(e);
  return RMH_OK;
}
