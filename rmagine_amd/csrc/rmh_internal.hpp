// Internal declarations of librmh: error handling, map/sensor objects and the FP32
// pose/ray algebra shared by the kernels.
//
// The whole library is compiled with -ffp-contract=off.  Pose and ray set-up below
// repeat Rmagine's operation order exactly (no fused multiply-add), so a pose and a
// model direction produce the same FP32 ray bits as the reference's CPU code:
//   quaternion product        src/rmagine_core/include/rmagine/math/types/Quaternion.tcc:35-42
//   rotate vector (q p q^-1)  Quaternion.tcc:70-75, inverse = conjugate Quaternion.tcc:19-24
//   Transform mult / inv      src/rmagine_core/include/rmagine/math/types/Transform.tcc:24-62
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <string>
#include <vector>
#include <rmh.h>

#define RMH_HD __host__ __device__ __forceinline__

namespace rmh {

// ---- error handling --------------------------------------------------------------
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);
int hip_fail(hipError_t e, const char* what);

#define RMH_HIP_TRY(expr)                                   \
  do {                                                      \
    hipError_t _e = (expr);                                 \
    if (_e != hipSuccess) return ::rmh::hip_fail(_e, #expr); \
  } while (0)

// ---- small vector algebra ------------------------------------------------------------
struct V3 { float x, y, z; };
struct Q4 { float x, y, z, w; };
struct Tf { Q4 R; V3 t; };

RMH_HD V3 v3(float x, float y, float z) { V3 r; r.x = x; r.y = y; r.z = z; return r; }
RMH_HD V3 add(const V3& a, const V3& b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
RMH_HD V3 scale(const V3& a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
RMH_HD float dot(const V3& a, const V3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

// Vector3::l2norm / normalize (Vector3.tcc:180-211): sqrtf of (x*x + y*y + z*z), then
// three divisions.
RMH_HD V3 normalize(const V3& a) {
  const float n = sqrtf(a.x * a.x + a.y * a.y + a.z * a.z);
  return v3(a.x / n, a.y / n, a.z / n);
}

// Quaternion::mult(Quaternion) (Quaternion.tcc:35-42)
RMH_HD Q4 qmul(const Q4& a, const Q4& b) {
  Q4 r;
  r.x = a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y;
  r.y = a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x;
  r.z = a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w;
  r.w = a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z;
  return r;
}

RMH_HD Q4 qconj(const Q4& q) { Q4 r; r.x = -q.x; r.y = -q.y; r.z = -q.z; r.w = q.w; return r; }

// Quaternion::mult(Vector) = q * (p,0) * q^-1 (Quaternion.tcc:70-75), inverse is the
// conjugate (Quaternion.tcc:19-24).
RMH_HD V3 qrot(const Q4& q, const V3& p) {
  Q4 P; P.x = p.x; P.y = p.y; P.z = p.z; P.w = 0.0f;
  const Q4 r = qmul(qmul(q, P), qconj(q));
  return v3(r.x, r.y, r.z);
}

// Transform::mult(Transform) (Transform.tcc:34-42): t = R*T2.t; R = R*T2.R; t += t1
RMH_HD Tf tmul(const Tf& a, const Tf& b) {
  Tf r;
  r.t = qrot(a.R, b.t);
  r.R = qmul(a.R, b.R);
  r.t = add(r.t, a.t);
  return r;
}

// Transform::mult(Vector) (Transform.tcc:57-62): R*v + t
RMH_HD V3 tapply(const Tf& T, const V3& v) { return add(qrot(T.R, v), T.t); }

RMH_HD Tf load_tf(const rmh_transform& T) {
  Tf r;
  r.R.x = T.R.x; r.R.y = T.R.y; r.R.z = T.R.z; r.R.w = T.R.w;
  r.t = v3(T.t.x, T.t.y, T.t.z);
  return r;
}

// ---- library objects -------------------------------------------------------------------
// Node of the binary BVH, 64 B: the boxes of both children and their references.
//   a = (c0.lo.x, c0.hi.x, c0.lo.y, c0.hi.y)
//   b = (c1.lo.x, c1.hi.x, c1.lo.y, c1.hi.y)
//   c = (c0.lo.z, c0.hi.z, c1.lo.z, c1.hi.z)
//   d = (ref0, ref1, -, -)
// A reference r >= 0 is an interior node index; r < 0 is a leaf: ~r = first << 3 | (count-1)
// where [first, first+count) is a range of the Morton-ordered triangle array.
struct alignas(64) Node2 {
  float4 a, b, c;
  int4 d;
};

// Node of the 4-wide BVH, 64 B (16 words), collapsed from the binary tree.  Child boxes are
// quantised to 8 bits per plane on a per-node grid: plane = origin + q * 2^e per axis.
// The origin is a multiple of 2^e below 2^(24+e) in magnitude, so every grid plane is an
// FP32 number and fma(q, 2^e, origin) computes it exactly; planes are rounded outward, so
// each quantised box contains the padded box the binary node holds for that child.
//   w[0..2]  origin x, y, z (float bits)
//   w[3]     byte 0..2: e_x + 127, e_y + 127, e_z + 127; byte 3: number of children (2..4)
//   w[4..9]  lo x, hi x, lo y, hi y, lo z, hi z: byte k = the plane of child k
//   w[10..13] child references (as in Node2; interior references index Node4 nodes)
//   w[14..15] unused
struct alignas(64) Node4 {
  uint32_t w[16];
};

constexpr int kMaxLeaf = 8;
// deepest binary tree the ray kernels trace (one traversal-stack entry per level)
constexpr uint32_t kMaxBinaryDepth = 128;

RMH_HD int make_leaf_ref(uint32_t first, uint32_t count) {
  return ~static_cast<int>((first << 3) | (count - 1u));
}

struct MapData {
  // Morton-ordered triangle records, 3 x float4 each:
  //   [0] = v0.xyz, face id within its mesh (bits)   [1] = v1.xyz, geom id (bits)
  //   [2] = v2.xyz, original global face index (bits)
  const float4* tris;
  // Morton-ordered unit face normals (FP64 cross product, rounded once), w unused.
  const float4* normals;
  const Node2* nodes;
  int root;           // root reference (interior index or leaf)
  uint32_t n_faces;
  const Node4* nodes4;  // 4-wide tree (NULL when the map has none)
  int root4;
};

}  // namespace rmh

struct rmh_map {
  uint32_t n_vertices = 0, n_faces = 0, n_meshes = 0;
  float* vertices = nullptr;      // owned copy, 3 floats per vertex
  uint32_t* faces = nullptr;      // owned copy
  float4* tris = nullptr;
  float4* normals = nullptr;
  rmh::Node2* nodes = nullptr;
  int root = 0;
  uint32_t n_internal = 0;
  rmh_map_info info{};
  int device = 0;
  std::vector<uint32_t> mesh_offsets;  // host copy: faces of mesh g are [off[g], off[g+1])
  int* parent = nullptr;               // parent of each interior node (refit), -1 at the root
  uint8_t* reach = nullptr;            // 1 for the interior nodes of the final tree
  // 4-wide quantised tree collapsed from the binary one; the ray kernels traverse it when
  // present (nodes4 != NULL), closest-point queries and refit work on the binary tree
  rmh::Node4* nodes4 = nullptr;
  int root4 = 0;
  uint32_t wide_capacity = 0;          // allocated Node4 slots
  // refit schedule, made at the first refit: the interior nodes of the final tree ordered by
  // depth, deepest first, and where each depth starts (refit_levels[d] .. refit_levels[d+1]
  // in reverse depth order)
  uint32_t* refit_order = nullptr;
  std::vector<uint32_t> refit_levels;
};

struct rmh_sensor {
  int kind = 0;
  uint32_t width = 0, height = 0;
  rmh_interval range{};
  rmh_spherical_model sphere{};
  rmh_pinhole_model pinhole{};
  rmh_vec3 orig{};
  float2* lut = nullptr;          // spherical: H rows (cos,sin phi) then W cols (cos,sin theta)
  // O1Dn / OnDn rays in traversal order (sorted by direction); order[s] is the buffer
  // index (row * W + column) of traversal slot s
  rmh_vec3* dirs = nullptr;       // W*H
  rmh_vec3* origs = nullptr;      // OnDn only, W*H
  uint32_t* order = nullptr;      // W*H
  // the rays in list order and the direction-sorted permutation, kept so that the
  // traversal order can be switched (rmh_sensor_set_ray_order) without the caller's arrays
  rmh_vec3* dirs_list = nullptr;
  rmh_vec3* origs_list = nullptr;
  uint32_t* perm = nullptr;
  int sorted = 1;                 // traversal order: RMH_RAY_ORDER_LIST / _SORTED / _TILES
  int speculative = 1;            // O1Dn / OnDn walk in rmh_simulate: speculative or classic
  int device = 0;
};
