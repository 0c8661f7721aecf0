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
#include "rmh_trace.hpp"

#include <cfloat>

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
  uint32_t n;
  double best0;  // max_distance^2: a result needs d^2 < best0
  rmh_closest_outputs out;
};

constexpr int kClosestLds = 16;

template <int STACK>
__global__ __launch_bounds__(64, 4) void k_closest(const CPArgs a) {
  __shared__ int2 stk[kClosestLds][64];
  const uint32_t i = blockIdx.x * 64 + threadIdx.x;
  if (i >= a.n) return;
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
  const uint32_t depth = map->info.max_depth;
  if (depth <= 32) k_closest<32><<<blocks, 64, 0, s>>>(a);
  else if (depth <= 64) k_closest<64><<<blocks, 64, 0, s>>>(a);
  else k_closest<128><<<blocks, 64, 0, s>>>(a);
  RMH_HIP_TRY(hipGetLastError());
  return RMH_OK;
}
