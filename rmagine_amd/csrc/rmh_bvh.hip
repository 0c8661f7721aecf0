// GPU BVH builder of librmh: the role of OptixScene's GAS build
// (src/rmagine_optix/src/map/optix/OptixScene.cpp:230-379) and EmbreeScene::commit
// (src/rmagine_embree/src/map/embree/EmbreeScene.cpp:187-191).
//
// Steps, all on the device:
//   1. per face: bounds, centroid, scene centroid bounds (wave64 shuffle + one atomic
//      per block), face-index validation;
//   2. 63-bit Morton codes of the centroids;
//   3. radix sort of (code, face) pairs (rocPRIM);
//   4. Karras 2012 radix tree over the sorted codes (index-augmented for duplicates);
//   4b. SAH refinement: three bottom-up rounds of treelet restructuring (Karras and Aila,
//      HPG 2013; rmh_treelet.hpp), then depth-first renumbering of the leaves so that
//      every subtree is one range of the triangle array again;
//   5. bottom-up pass: node boxes, primitive counts and the area-weighted SAH cost; a
//      subtree of at most max_leaf_size triangles becomes a leaf when that is cheaper
//      under SAH.  The bottom-up passes (4b, 5) run one launch per tree depth, deepest
//      first, over the interior nodes sorted by depth (make_levels);
//   6. layout: 64-byte binary nodes that hold both child boxes (padded outward), plus
//      depth statistics of the reachable tree;
//   7. gather of the triangles in Morton order (3 x float4 per triangle with the face,
//      geometry and global ids in the w lanes) and of FP64-derived unit face normals.
// The normals play the part of EmbreeMesh::computeFaceNormals
// (src/rmagine_embree/src/map/embree/EmbreeMesh.cpp:275-295) and of
// computeFaceNormals_kernel (src/rmagine_cuda/src/map/mesh_preprocessing.cu:6-20).
#include "rmh_internal.hpp"
#include "rmh_treelet.hpp"

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <new>
#include <utility>
#include <vector>

namespace {

using namespace rmh;

constexpr int kBlock = 256;
constexpr uint32_t kLeafBit = kTreeLeafBit;
constexpr int kTreeletRounds = 3;
// largest map: node references must stay below 2^26 (32-bit node byte offsets in the ray kernels)
constexpr uint32_t kMaxFaces = 1u << 26;

__device__ inline void atomic_min_float(float* addr, float v) {
  if (v >= 0.0f) atomicMin(reinterpret_cast<int*>(addr), __float_as_int(v));
  else atomicMax(reinterpret_cast<unsigned int*>(addr), __float_as_uint(v));
}

__device__ inline void atomic_max_float(float* addr, float v) {
  if (v >= 0.0f) atomicMax(reinterpret_cast<int*>(addr), __float_as_int(v));
  else atomicMin(reinterpret_cast<unsigned int*>(addr), __float_as_uint(v));
}

__device__ inline float wave_min(float v) {
  for (int off = 32; off > 0; off >>= 1) v = fminf(v, __shfl_xor(v, off, 64));
  return v;
}

__device__ inline float wave_max(float v) {
  for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
  return v;
}

__global__ void k_init(float* bounds, uint32_t* err) {
  if (threadIdx.x < 3) bounds[threadIdx.x] = FLT_MAX;
  else if (threadIdx.x < 6) bounds[threadIdx.x] = -FLT_MAX;
  else if (threadIdx.x == 6) *err = 0u;
}

// Step 1.
__global__ __launch_bounds__(kBlock) void k_prep(const float* __restrict__ V, uint32_t nv,
                                                 const uint32_t* __restrict__ F, uint32_t nf,
                                                 float4* __restrict__ lo, float4* __restrict__ hi,
                                                 float* bounds, uint32_t* err) {
  const uint32_t f = blockIdx.x * kBlock + threadIdx.x;
  float cmin[3] = {FLT_MAX, FLT_MAX, FLT_MAX};
  float cmax[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
  if (f < nf) {
    uint32_t idx[3] = {F[3 * f + 0], F[3 * f + 1], F[3 * f + 2]};
    bool bad = false;
    for (int k = 0; k < 3; k++) {
      if (idx[k] >= nv) { bad = true; idx[k] = 0; }
    }
    if (bad) atomicOr(err, 1u);
    float b0[3], b1[3];
    for (int a = 0; a < 3; a++) {
      const float x0 = V[3 * idx[0] + a], x1 = V[3 * idx[1] + a], x2 = V[3 * idx[2] + a];
      b0[a] = fminf(x0, fminf(x1, x2));
      b1[a] = fmaxf(x0, fmaxf(x1, x2));
      if (!(isfinite(x0) && isfinite(x1) && isfinite(x2))) atomicOr(err, 2u);
      const float c = 0.5f * b0[a] + 0.5f * b1[a];
      cmin[a] = c;
      cmax[a] = c;
    }
    lo[f] = make_float4(b0[0], b0[1], b0[2], 0.0f);
    hi[f] = make_float4(b1[0], b1[1], b1[2], 0.0f);
  }
  __shared__ float red[2][3][kBlock / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  for (int a = 0; a < 3; a++) {
    const float mn = wave_min(cmin[a]), mx = wave_max(cmax[a]);
    if (lane == 0) { red[0][a][wid] = mn; red[1][a][wid] = mx; }
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    const int a = threadIdx.x;
    float mn = red[0][a][0], mx = red[1][a][0];
    for (int w = 1; w < kBlock / 64; w++) { mn = fminf(mn, red[0][a][w]); mx = fmaxf(mx, red[1][a][w]); }
    atomic_min_float(&bounds[a], mn);
    atomic_max_float(&bounds[3 + a], mx);
  }
}

// Step 2.
__global__ __launch_bounds__(kBlock) void k_morton(const float4* __restrict__ lo, const float4* __restrict__ hi,
                                                   uint32_t nf, const float* __restrict__ bounds,
                                                   uint64_t* __restrict__ keys, uint32_t* __restrict__ vals) {
  const uint32_t f = blockIdx.x * kBlock + threadIdx.x;
  if (f >= nf) return;
  const float4 a = lo[f], b = hi[f];
  const float l[3] = {a.x, a.y, a.z}, h[3] = {b.x, b.y, b.z};
  const float bmin[3] = {bounds[0], bounds[1], bounds[2]};
  const float ext = fmaxf(fmaxf(bounds[3] - bounds[0], bounds[4] - bounds[1]), bounds[5] - bounds[2]);
  keys[f] = morton_key(l, h, bmin, ext);
  vals[f] = f;
}

// Step 4 (karras_node: Karras 2012, Fig. 4).
__global__ __launch_bounds__(kBlock) void k_karras(const uint64_t* __restrict__ keys, int n,
                                                   uint32_t* __restrict__ child, uint32_t* __restrict__ range,
                                                   int* __restrict__ parent_int, int* __restrict__ parent_leaf) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n - 1) return;
  karras_node(keys, n, i, child, range, parent_int, parent_leaf);
}

__device__ inline float half_area(float4 a, float4 b) {
  const float dx = b.x - a.x, dy = b.y - a.y, dz = b.z - a.z;
  return dx * dy + dy * dz + dz * dx;
}

// Step 5.  node_lo.w = SAH cost (area weighted), node_hi.w = primitive count.  One launch
// per depth of the radix tree, deepest first (make_levels): every node of the level reads
// its children, which earlier launches finished, so no arrival counters are needed (an
// agent-scope acquire-release atomic writes back and invalidates the L2 of an XCD each time,
// which made the counter-driven pass 7 ms on the 2.1M-triangle terrain).
__global__ __launch_bounds__(kBlock) void k_bottom_up_level(const uint32_t* __restrict__ order, uint32_t lo, uint32_t hi,
                                                            const uint32_t* __restrict__ vals,
                                                            const float4* __restrict__ flo, const float4* __restrict__ fhi,
                                                            const uint32_t* __restrict__ child,
                                                            float4* node_lo, float4* node_hi,
                                                            uint8_t* collapse, uint32_t* subtree,
                                                            uint32_t max_leaf, float c_node, float c_tri) {
  const uint32_t k = lo + blockIdx.x * kBlock + threadIdx.x;
  if (k >= hi) return;
  const uint32_t p = order[k];
  float4 lo2[2], hi2[2];
  float cost[2], cnt[2];
  uint32_t below = 0;  // interior nodes of the final tree in the children's subtrees
  for (int c = 0; c < 2; c++) {
    const uint32_t ch = child[2 * p + c];
    if (ch & kLeafBit) {
      const uint32_t f = vals[ch & ~kLeafBit];
      lo2[c] = flo[f];
      hi2[c] = fhi[f];
      cnt[c] = 1.0f;
      cost[c] = c_tri * half_area(lo2[c], hi2[c]);
    } else {
      lo2[c] = node_lo[ch];
      hi2[c] = node_hi[ch];
      cost[c] = lo2[c].w;
      cnt[c] = hi2[c].w;
      below += subtree[ch];
    }
  }
  float4 L = make_float4(fminf(lo2[0].x, lo2[1].x), fminf(lo2[0].y, lo2[1].y), fminf(lo2[0].z, lo2[1].z), 0.0f);
  float4 H = make_float4(fmaxf(hi2[0].x, hi2[1].x), fmaxf(hi2[0].y, hi2[1].y), fmaxf(hi2[0].z, hi2[1].z), 0.0f);
  const float count = cnt[0] + cnt[1];
  const float area = half_area(L, H);
  const float split_cost = c_node * area + cost[0] + cost[1];
  float node_cost = split_cost;
  uint8_t col = 0;
  if (count <= static_cast<float>(max_leaf)) {
    const float leaf_cost = c_tri * area * count;
    if (leaf_cost <= split_cost) { node_cost = leaf_cost; col = 1; }
  }
  L.w = node_cost;
  H.w = count;
  node_lo[p] = L;
  node_hi[p] = H;
  collapse[p] = col;
  subtree[p] = col ? 0u : 1u + below;  // a collapsed node becomes a leaf of the final tree
}

// Sort key of every interior node of the radix tree for the level-by-level passes: its
// depth (the root, node 0, has parent -1), deepest first, and within a depth the nodes
// that get a treelet (count >= gamma; none when count is NULL) before the others.  A parent
// chain that leaves the node range or runs longer than kMaxTreeWalk steps (a cycle) sets err
// bit 8.
__global__ __launch_bounds__(kBlock) void k_level_keys(int n, const int* __restrict__ parent_int,
                                                       const uint32_t* __restrict__ count, uint32_t gamma,
                                                       uint32_t* __restrict__ keys, uint32_t* __restrict__ idx,
                                                       uint32_t* err) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n - 1) return;
  uint32_t d = 0;
  for (int p = parent_int[i]; p >= 0; p = parent_int[p]) {
    if (++d > static_cast<uint32_t>(kMaxTreeWalk) || p >= n - 1) { atomicOr(err, 8u); d = 0; break; }
  }
  const bool need = count && count[i] >= gamma;
  keys[i] = 2u * (static_cast<uint32_t>(kMaxTreeWalk) - d) + (need ? 0u : 1u);
  idx[i] = static_cast<uint32_t>(i);
}

// start[key] = first position of `key` in the sorted keys (untouched for absent keys)
__global__ __launch_bounds__(kBlock) void k_level_starts(uint32_t ni, const uint32_t* __restrict__ skeys,
                                                         uint32_t* __restrict__ start) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= ni) return;
  if (i == 0 || skeys[i] != skeys[i - 1]) start[skeys[i]] = i;
}

// Triangles below every interior node of the radix tree: its range of sorted leaves.
__global__ __launch_bounds__(kBlock) void k_range_counts(uint32_t ni, const uint32_t* __restrict__ range,
                                                         uint32_t* __restrict__ count) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i < ni) count[i] = range[2 * i + 1] - range[2 * i] + 1u;
}

// Outward padding so that the FP32 slab test stays conservative for boxes of zero
// thickness and for values that are not exactly representable after subtraction.
__device__ inline float pad_lo(float x) { return x - fabsf(x) * 2.4e-7f - 1e-30f; }
__device__ inline float pad_hi(float x) { return x + fabsf(x) * 2.4e-7f + 1e-30f; }

// Step 6.
__global__ __launch_bounds__(kBlock) void k_layout(int n, const uint32_t* __restrict__ vals,
                                                   const float4* __restrict__ flo, const float4* __restrict__ fhi,
                                                   const uint32_t* __restrict__ child, const uint32_t* __restrict__ range,
                                                   const int* __restrict__ parent_int,
                                                   const float4* __restrict__ node_lo, const float4* __restrict__ node_hi,
                                                   const uint8_t* __restrict__ collapse, Node2* __restrict__ nodes,
                                                   uint32_t* stats /* [0]=reachable interior, [1]=max depth */,
                                                   uint8_t* __restrict__ reach) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n - 1) return;
  float4 lo[2], hi[2];
  int ref[2];
  for (int c = 0; c < 2; c++) {
    const uint32_t ch = child[2 * i + c];
    if (ch & kLeafBit) {
      const uint32_t k = ch & ~kLeafBit;
      const uint32_t f = vals[k];
      lo[c] = flo[f];
      hi[c] = fhi[f];
      ref[c] = make_leaf_ref(k, 1u);
    } else {
      lo[c] = node_lo[ch];
      hi[c] = node_hi[ch];
      if (collapse[ch]) {
        const uint32_t first = range[2 * ch], last = range[2 * ch + 1];
        ref[c] = make_leaf_ref(first, last - first + 1u);
      } else {
        ref[c] = static_cast<int>(ch);
      }
    }
  }
  Node2 nd;
  nd.a = make_float4(pad_lo(lo[0].x), pad_hi(hi[0].x), pad_lo(lo[0].y), pad_hi(hi[0].y));
  nd.b = make_float4(pad_lo(lo[1].x), pad_hi(hi[1].x), pad_lo(lo[1].y), pad_hi(hi[1].y));
  nd.c = make_float4(pad_lo(lo[0].z), pad_hi(hi[0].z), pad_lo(lo[1].z), pad_hi(hi[1].z));
  nd.d = make_int4(ref[0], ref[1], 0, 0);
  nodes[i] = nd;
  // reachable iff neither this node nor an ancestor became a leaf
  bool reachable = collapse[i] == 0;
  uint32_t depth = 0;
  int p = i;
  while (reachable && p != 0) {
    p = parent_int[p];
    depth++;
    if (collapse[p]) reachable = false;
  }
  reach[i] = reachable ? 1 : 0;
  if (reachable) {
    atomicAdd(&stats[0], 1u);
    atomicMax(&stats[1], depth + 1u);
  }
}

// Step 8: node order.  Interior indices of the radix tree follow split positions, and the
// subtrees collapsed into leaves leave about half of them unused, so the nodes a ray visits
// one after another lie far apart.  The final tree is rewritten in depth-first preorder
// (a node, its first child's subtree, its second child's subtree): a descent through first
// children reads consecutive 64-B nodes, every subtree is one contiguous range, and the
// array holds only reachable nodes.  The interior-node count of every subtree of the final
// tree (size) comes from the bottom-up pass of step 5.
// 8a: preorder position = sum over the path from the root of 1 per step, plus the size
// of the first child's subtree wherever the path takes the second child.
__global__ __launch_bounds__(kBlock) void k_preorder_rank(uint32_t ni, const uint8_t* __restrict__ reach,
                                                          const int* __restrict__ parent, const Node2* __restrict__ nodes,
                                                          const uint32_t* __restrict__ size, uint32_t* __restrict__ rank,
                                                          uint32_t* err) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= ni || !reach[i]) return;
  uint32_t r = 0;
  int j = static_cast<int>(i);
  for (int steps = 0; j != 0; steps++) {
    if (steps >= kMaxTreeWalk) { atomicOr(err, 1u); return; }
    const int p = parent[j];
    const int first = nodes[p].d.x;
    r += 1u + ((first != j && first >= 0) ? size[first] : 0u);
    j = p;
  }
  rank[i] = r;
}

// 8b: move every reachable node to its preorder slot, with its child references, parent
// link and reach flag renumbered.
__global__ __launch_bounds__(kBlock) void k_preorder_scatter(uint32_t ni, uint32_t n_out, const uint8_t* __restrict__ reach,
                                                             const int* __restrict__ parent, const Node2* __restrict__ nodes,
                                                             const uint32_t* __restrict__ rank, Node2* __restrict__ out,
                                                             int* __restrict__ out_parent, uint8_t* __restrict__ out_reach,
                                                             uint32_t* err) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= ni || !reach[i]) return;
  const uint32_t s = rank[i];
  if (s >= n_out) { atomicOr(err, 1u); return; }
  Node2 nd = nodes[i];
  if (nd.d.x >= 0) nd.d.x = static_cast<int>(rank[nd.d.x]);
  if (nd.d.y >= 0) nd.d.y = static_cast<int>(rank[nd.d.y]);
  out[s] = nd;
  out_parent[s] = i == 0 ? -1 : static_cast<int>(rank[parent[i]]);
  out_reach[s] = 1;
}

// Step 7.
__global__ __launch_bounds__(kBlock) void k_gather(uint32_t nf, const uint32_t* __restrict__ vals,
                                                   const float* __restrict__ V, const uint32_t* __restrict__ F,
                                                   const uint32_t* __restrict__ mesh_off, uint32_t n_meshes,
                                                   float4* __restrict__ tris, float4* __restrict__ normals) {
  const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
  if (k >= nf) return;
  const uint32_t f = vals[k];
  // geometry of face f: largest g with mesh_off[g] <= f
  uint32_t lo = 0, hi = n_meshes;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (mesh_off[mid] <= f) lo = mid; else hi = mid;
  }
  const uint32_t g = lo, local = f - mesh_off[lo];
  const uint32_t i0 = F[3 * f], i1 = F[3 * f + 1], i2 = F[3 * f + 2];  // validated in k_prep
  const float x0 = V[3 * i0], y0 = V[3 * i0 + 1], z0 = V[3 * i0 + 2];
  const float x1 = V[3 * i1], y1 = V[3 * i1 + 1], z1 = V[3 * i1 + 2];
  const float x2 = V[3 * i2], y2 = V[3 * i2 + 1], z2 = V[3 * i2 + 2];
  tris[3 * static_cast<size_t>(k) + 0] = make_float4(x0, y0, z0, __uint_as_float(local));
  tris[3 * static_cast<size_t>(k) + 1] = make_float4(x1, y1, z1, __uint_as_float(g));
  tris[3 * static_cast<size_t>(k) + 2] = make_float4(x2, y2, z2, __uint_as_float(f));
  const double ax = (double)x1 - (double)x0, ay = (double)y1 - (double)y0, az = (double)z1 - (double)z0;
  const double bx = (double)x2 - (double)x0, by = (double)y2 - (double)y0, bz = (double)z2 - (double)z0;
  const double nx = ay * bz - az * by, ny = az * bx - ax * bz, nz = ax * by - ay * bx;
  const double len = sqrt(nx * nx + ny * ny + nz * nz);
  // w: object id of the face (the mesh index until rmh_map_set_mesh_ids says otherwise)
  float4 nrm = make_float4(0.0f, 0.0f, 0.0f, __uint_as_float(g));
  if (len > 0.0) nrm = make_float4((float)(nx / len), (float)(ny / len), (float)(nz / len), __uint_as_float(g));
  normals[k] = nrm;
}

// Per-mesh geom / object ids: for every Morton slot, the mesh of its global face (v2.w)
// selects the ids (NULL table = keep).
__global__ __launch_bounds__(kBlock) void k_set_ids(uint32_t nf, const uint32_t* __restrict__ mesh_off, uint32_t n_meshes,
                                                    const uint32_t* __restrict__ geom, const uint32_t* __restrict__ obj,
                                                    float4* __restrict__ tris, float4* __restrict__ normals) {
  const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
  if (k >= nf) return;
  const uint32_t f = __float_as_uint(tris[3 * static_cast<size_t>(k) + 2].w);
  uint32_t lo = 0, hi = n_meshes;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (mesh_off[mid] <= f) lo = mid; else hi = mid;
  }
  if (geom) tris[3 * static_cast<size_t>(k) + 1].w = __uint_as_float(geom[lo]);
  if (obj) normals[k].w = __uint_as_float(obj[lo]);
}

// ---- SAH refinement: treelet restructuring (rmh_treelet.hpp) ------------------------------------

__global__ __launch_bounds__(kBlock) void k_leaf_boxes(uint32_t nf, const uint32_t* __restrict__ vals,
                                                       const float4* __restrict__ flo, const float4* __restrict__ fhi,
                                                       float* __restrict__ leaf_box) {
  const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
  if (k >= nf) return;
  const uint32_t f = vals[k];
  const float4 a = flo[f], b = fhi[f];
  float* o = leaf_box + 6 * static_cast<size_t>(k);
  o[0] = a.x; o[1] = a.y; o[2] = a.z; o[3] = b.x; o[4] = b.y; o[5] = b.z;
}

// Treelet restructuring of `root` by one whole wave (all 64 lanes call it with the same
// root): every lane forms the same treelet; the areas of the leaf subsets (two per lane)
// and then, subset size by subset size, their optimal costs and splits are computed in
// LDS (a subset's splits only use smaller subsets); lane 0 rebuilds the treelet.  The
// per-subset choices are those of the serial treelet_optimize().
__device__ void treelet_wave(const TreeArrays& t, int root, float c_node, float c_tri, float* s_area, float* s_copt,
                             uint8_t* s_part) {
  // the treelet (formed by lane 0) and its leaf boxes (one lane each) live in LDS: held per
  // lane, the arrays were private memory that all 64 lanes wrote and read
  __shared__ uint32_t s_leaf[kTreeletLeaves];
  __shared__ int s_inner[kTreeletLeaves - 1];
  __shared__ float s_lb[kTreeletLeaves][6];
  __shared__ int s_nl;
  const int lane = threadIdx.x;
  if (lane == 0) s_nl = treelet_form(t, root, s_leaf, s_inner);
  __syncthreads();
  const int nl = s_nl;
  if (nl < 3) return;  // the same for every lane
  if (lane < nl) ref_box(t, s_leaf[lane], s_lb[lane]);
  __syncthreads();
  const uint32_t full = (1u << nl) - 1u;
  for (uint32_t S = lane; S <= full; S += 64)
    if (S >= 1) s_area[S] = subset_area(s_lb, nl, S);
  if (lane < nl) {
    s_copt[1u << lane] = ref_cost(t, s_leaf[lane], c_tri);
    s_part[1u << lane] = 0;
  }
  __syncthreads();
  for (int size = 2; size <= nl; size++) {
    for (uint32_t S = lane; S <= full; S += 64) {
      if (S >= 1 && popcount7(S) == size) {
        uint32_t bp;
        s_copt[S] = best_split(s_copt, S, s_area[S], c_node, &bp);
        s_part[S] = static_cast<uint8_t>(bp);
      }
    }
    __syncthreads();
  }
  const bool better = s_copt[full] < t.cost[root] * (1.0f - 1e-6f);
  if (better && lane == 0) treelet_rebuild(t, root, s_leaf, s_inner, nl, s_part, c_node, c_tri);
  __syncthreads();
}

// One bottom-up treelet pass runs level by level over the current tree, deepest first
// (make_levels): at each depth, k_refresh_nodes sets the box, count and cost of the nodes
// with fewer than `gamma` triangles below from their children (thread per node), and
// k_treelet_nodes does the same for the others and restructures each with one wave
// (treelet_wave).  Children are final when their parent's level runs (earlier launches), a
// treelet only involves nodes of its subtree, and the nodes of one depth have disjoint
// subtrees, so no counters or fences are needed.  Which nodes get a treelet is known in
// advance: a node's triangle count depends only on its leaf set, which restructuring below
// it does not change.
__global__ __launch_bounds__(kBlock) void k_refresh_nodes(const uint32_t* __restrict__ order, uint32_t lo, uint32_t hi,
                                                          TreeArrays t, float c_node, float c_tri) {
  const uint32_t k = lo + blockIdx.x * kBlock + threadIdx.x;
  if (k < hi) refresh_node(t, static_cast<int>(order[k]), c_node, c_tri);
}

__global__ __launch_bounds__(64) void k_treelet_nodes(const uint32_t* __restrict__ order, uint32_t lo, TreeArrays t,
                                                      float c_node, float c_tri) {
  __shared__ float s_area[1 << kTreeletLeaves];
  __shared__ float s_copt[1 << kTreeletLeaves];
  __shared__ uint8_t s_part[1 << kTreeletLeaves];
  const int p = static_cast<int>(order[lo + blockIdx.x]);
  if (threadIdx.x == 0) refresh_node(t, p, c_node, c_tri);
  __syncthreads();  // lane 0's box and cost before the wave reads them
  treelet_wave(t, p, c_node, c_tri, s_area, s_copt, s_part);
}

// Structure check of the refined tree before anything walks it: every child slot points
// into range at a node whose parent link names the slot's owner, no node is referenced
// twice (refs[] counts references), and every node reaches the root by parent links within
// kMaxTreeWalk steps (no cycle).  Any failure sets *err.
__global__ __launch_bounds__(kBlock) void k_check_tree(int n, TreeArrays t, uint32_t* refs, uint32_t* err) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n - 1) {
    const uint32_t c0 = t.child[2 * i], c1 = t.child[2 * i + 1];
    if (c0 == c1) atomicOr(err, 4u);
    for (int c = 0; c < 2; c++) {
      const uint32_t r = c ? c1 : c0;
      if (r & kTreeLeafBit) {
        const uint32_t k = r & ~kTreeLeafBit;
        if (k >= static_cast<uint32_t>(n) || t.parent_leaf[k] != i) { atomicOr(err, 4u); continue; }
        atomicAdd(&refs[n - 1 + k], 1u);
      } else {
        if (r == 0u || r >= static_cast<uint32_t>(n - 1) || t.parent_int[r] != i) { atomicOr(err, 4u); continue; }
        atomicAdd(&refs[r], 1u);
      }
    }
  }
  if (i < n) {
    int p = t.parent_leaf[i];
    int steps = 0;
    while (p >= 0 && steps < kMaxTreeWalk) {
      if (p >= n - 1) { p = -2; break; }
      p = t.parent_int[p];
      steps++;
    }
    if (p != -1) atomicOr(err, 4u);
  }
}

__global__ __launch_bounds__(kBlock) void k_check_refs(int n, const uint32_t* refs, uint32_t* err) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= 2 * n - 1) return;
  const uint32_t want = i == 0 ? 0u : 1u;  // the root has no parent slot
  if (refs[i] != want) atomicOr(err, 4u);
}

// Depth-first position of every leaf and of the first leaf below every interior node.
__global__ __launch_bounds__(kBlock) void k_dfs_positions(int n, TreeArrays t, uint32_t* __restrict__ pos,
                                                          uint32_t* __restrict__ off) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) pos[i] = dfs_offset(t, static_cast<uint32_t>(i) | kTreeLeafBit);
  if (i < n - 1) off[i] = dfs_offset(t, static_cast<uint32_t>(i));
}

// Leaves renumbered depth-first: the sorted face list, the leaf parents and the leaf
// references follow, and every interior node covers the range [off, off + count - 1].
__global__ __launch_bounds__(kBlock) void k_apply_order(int n, TreeArrays t, const uint32_t* __restrict__ pos,
                                                        const uint32_t* __restrict__ off, const uint32_t* __restrict__ vals,
                                                        uint32_t* __restrict__ vals_out, int* __restrict__ pleaf_out,
                                                        uint32_t* __restrict__ range) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) {
    vals_out[pos[i]] = vals[i];
    pleaf_out[pos[i]] = t.parent_leaf[i];
  }
  if (i < n - 1) {
    for (int c = 0; c < 2; c++) {
      const uint32_t r = t.child[2 * i + c];
      if (r & kTreeLeafBit) t.child[2 * i + c] = pos[r & ~kTreeLeafBit] | kTreeLeafBit;
    }
    range[2 * i] = off[i];
    range[2 * i + 1] = off[i] + t.count[i] - 1u;
  }
}

// ---- 4-wide quantised tree ------------------------------------------------------------------
//
// Collapsed top-down from the binary tree, one launch pair per level: a wide node made
// from binary node b starts with b's two children and repeatedly opens the interior
// candidate of largest surface area until it has four children or only leaves.  A
// level's children are numbered by an exclusive scan of the per-node interior-child
// counts, so the build is deterministic and each level lies in memory in the order of its
// parents (Morton order of the binary tree).

constexpr int kEmptyRef = 0x7fffffff;  // unused child slot (masked by the child count)

struct Cand {
  int ref;
  float lo[3], hi[3];
};

__device__ inline Cand child_cand(const Node2& nd, int c) {
  Cand k;
  if (c == 0) {
    k.ref = nd.d.x;
    k.lo[0] = nd.a.x; k.hi[0] = nd.a.y; k.lo[1] = nd.a.z; k.hi[1] = nd.a.w; k.lo[2] = nd.c.x; k.hi[2] = nd.c.y;
  } else {
    k.ref = nd.d.y;
    k.lo[0] = nd.b.x; k.hi[0] = nd.b.y; k.lo[1] = nd.b.z; k.hi[1] = nd.b.w; k.lo[2] = nd.c.z; k.hi[2] = nd.c.w;
  }
  return k;
}

__device__ inline float cand_area(const Cand& c) {
  const float dx = c.hi[0] - c.lo[0], dy = c.hi[1] - c.lo[1], dz = c.hi[2] - c.lo[2];
  return dx * dy + dy * dz + dz * dx;
}

// Children of the wide node made from binary node b, in left-to-right order.
__device__ inline int select_children(const Node2* __restrict__ nodes, int b, Cand c[4]) {
  const Node2 nd = nodes[b];
  c[0] = child_cand(nd, 0);
  c[1] = child_cand(nd, 1);
  int n = 2;
  while (n < 4) {
    int best = -1;
    float best_area = -1.0f;
    for (int k = 0; k < n; k++)
      if (c[k].ref >= 0 && cand_area(c[k]) > best_area) { best_area = cand_area(c[k]); best = k; }
    if (best < 0) break;
    const Node2 cn = nodes[c[best].ref];
    for (int k = n; k > best + 1; k--) c[k] = c[k - 1];
    c[best] = child_cand(cn, 0);
    c[best + 1] = child_cand(cn, 1);
    n++;
  }
  return n;
}

// Grid of one axis for the planes in [lo, hi]: the smallest exponent e >= -126 with an
// origin o = floor(lo / 2^e) 2^e such that o + 255 2^e >= hi and every o + q 2^e
// (q = 0..255) is an FP32 number (|o| and |o + 255 2^e| below 2^(24+e) and FLT_MAX).
// All of it is exact in FP64.  False when no exponent up to 127 fits.
__device__ inline bool quant_frame(float lo, float hi, int& e_out, double& origin_out) {
  const double L = lo, H = hi;
  int e = 0;
  (void)frexp((H - L) / 255.0, &e);
  e = max(e - 1, -126);
  for (; e <= 127; e++) {
    const double s = ldexp(1.0, e);
    const double o = floor(L / s) * s;
    const double top = o + 255.0 * s;
    if (top < H) continue;
    const double lim = fmin(ldexp(1.0, 24 + e), static_cast<double>(FLT_MAX));
    if (fabs(o) >= lim || fabs(top) >= lim) continue;
    e_out = e;
    origin_out = o;
    return true;
  }
  return false;
}

// Largest grid plane <= x and smallest grid plane >= x (x inside [o, o + 255 s]); the
// checks are exact FP64 comparisons, so a rounded quotient cannot move a plane inward.
__device__ inline uint32_t quant_lo(float x, double o, double s) {
  double q = fmin(fmax(floor((static_cast<double>(x) - o) / s), 0.0), 255.0);
  while (q > 0.0 && o + q * s > static_cast<double>(x)) q -= 1.0;
  return static_cast<uint32_t>(q);
}

__device__ inline uint32_t quant_hi(float x, double o, double s) {
  double q = fmin(fmax(ceil((static_cast<double>(x) - o) / s), 0.0), 255.0);
  while (q < 255.0 && o + q * s < static_cast<double>(x)) q += 1.0;
  return static_cast<uint32_t>(q);
}

__global__ __launch_bounds__(kBlock) void k_wide_count(const Node2* __restrict__ nodes, const int* __restrict__ front,
                                                       uint32_t n, uint32_t* __restrict__ counts) {
  const uint32_t j = blockIdx.x * kBlock + threadIdx.x;
  if (j >= n) return;
  Cand c[4];
  const int k = select_children(nodes, front[j], c);
  uint32_t cnt = 0;
  for (int i = 0; i < k; i++) cnt += c[i].ref >= 0 ? 1u : 0u;
  counts[j] = cnt;
}

// Writes the wide nodes of one level (slots base .. base + n - 1) and the binary indices
// of the next level's nodes (slot next_base + offsets[j] + t for the t-th interior child).
__global__ __launch_bounds__(kBlock) void k_wide_emit(const Node2* __restrict__ nodes, const int* __restrict__ front,
                                                      uint32_t n, uint32_t base, uint32_t next_base, uint32_t cap,
                                                      const uint32_t* __restrict__ offsets, Node4* __restrict__ out,
                                                      int* __restrict__ next_front, uint32_t* err) {
  const uint32_t j = blockIdx.x * kBlock + threadIdx.x;
  if (j >= n) return;
  if (base + j >= cap) { atomicOr(err, 2u); return; }
  Cand c[4];
  const int k = select_children(nodes, front[j], c);
  Node4 nd;
  for (int i = 0; i < 16; i++) nd.w[i] = 0u;
  uint32_t meta = static_cast<uint32_t>(k) << 24;
  for (int a = 0; a < 3; a++) {
    float lo = c[0].lo[a], hi = c[0].hi[a];
    for (int i = 1; i < k; i++) { lo = fminf(lo, c[i].lo[a]); hi = fmaxf(hi, c[i].hi[a]); }
    int e;
    double o;
    if (!quant_frame(lo, hi, e, o)) { atomicOr(err, 1u); return; }
    const double s = ldexp(1.0, e);
    nd.w[a] = __float_as_uint(static_cast<float>(o));  // exact: o is an FP32 number
    meta |= static_cast<uint32_t>(e + 127) << (8 * a);
    for (int i = 0; i < k; i++) {
      nd.w[4 + 2 * a] |= quant_lo(c[i].lo[a], o, s) << (8 * i);
      nd.w[5 + 2 * a] |= quant_hi(c[i].hi[a], o, s) << (8 * i);
    }
  }
  nd.w[3] = meta;
  uint32_t slot = offsets[j];
  for (int i = 0; i < 4; i++) {
    int ref = kEmptyRef;
    if (i < k) {
      ref = c[i].ref;
      if (ref >= 0) {
        if (next_base + slot >= cap) { atomicOr(err, 2u); return; }
        next_front[slot] = ref;
        ref = static_cast<int>(next_base + slot);
        slot++;
      }
    }
    nd.w[10 + i] = static_cast<uint32_t>(ref);
  }
  out[base + j] = nd;
}

struct DevBuf {
  void* p = nullptr;
  ~DevBuf() { if (p) (void)hipFree(p); }
  template <class T> T* as() { return static_cast<T*>(p); }
};

inline uint32_t blocks_for(uint64_t n) { return static_cast<uint32_t>((n + kBlock - 1) / kBlock); }

// One launch range of the level-by-level passes: order[lo, hi) are nodes of one depth, with
// or without a treelet.
struct LevelClass {
  uint32_t lo, hi;
  bool treelet;
};

// Scratch of make_levels for a tree of ni interior nodes.
struct LevelBufs {
  DevBuf keys, keys2, idx, order, start, tmp;
  size_t tmp_bytes = 0;
};

constexpr uint32_t kLevelKeys = 2u * (static_cast<uint32_t>(kMaxTreeWalk) + 1u);

hipError_t level_bufs(LevelBufs& b, uint32_t ni, hipStream_t stream) {
  hipError_t e;
  if ((e = hipMalloc(&b.keys.p, sizeof(uint32_t) * ni)) != hipSuccess) return e;
  if ((e = hipMalloc(&b.keys2.p, sizeof(uint32_t) * ni)) != hipSuccess) return e;
  if ((e = hipMalloc(&b.idx.p, sizeof(uint32_t) * ni)) != hipSuccess) return e;
  if ((e = hipMalloc(&b.order.p, sizeof(uint32_t) * ni)) != hipSuccess) return e;
  if ((e = hipMalloc(&b.start.p, sizeof(uint32_t) * kLevelKeys)) != hipSuccess) return e;
  e = rocprim::radix_sort_pairs(nullptr, b.tmp_bytes, b.keys.as<uint32_t>(), b.keys2.as<uint32_t>(),
                                b.idx.as<uint32_t>(), b.order.as<uint32_t>(), ni, 0u, 14u, stream);
  if (e != hipSuccess) return e;
  return hipMalloc(&b.tmp.p, b.tmp_bytes);
}

// Interior nodes of the radix tree (n leaves, n - 1 interior nodes) grouped by depth,
// deepest first, into b.order; `classes` receives the launch ranges in that order (within a
// depth the treelet nodes first).  count = NULL: no treelets.  Synchronous: the range sizes
// size the launches.  A broken tree sets err bit 8.
hipError_t make_levels(int n, const int* parent_int, const uint32_t* count, uint32_t gamma, LevelBufs& b,
                       uint32_t* err, hipStream_t stream, std::vector<LevelClass>& classes) {
  const uint32_t ni = static_cast<uint32_t>(n - 1);
  hipError_t e;
  k_level_keys<<<blocks_for(ni), kBlock, 0, stream>>>(n, parent_int, count, gamma, b.keys.as<uint32_t>(),
                                                      b.idx.as<uint32_t>(), err);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  if ((e = rocprim::radix_sort_pairs(b.tmp.p, b.tmp_bytes, b.keys.as<uint32_t>(), b.keys2.as<uint32_t>(),
                                     b.idx.as<uint32_t>(), b.order.as<uint32_t>(), ni, 0u, 14u, stream)) != hipSuccess)
    return e;
  if ((e = hipMemsetAsync(b.start.p, 0xFF, sizeof(uint32_t) * kLevelKeys, stream)) != hipSuccess) return e;
  k_level_starts<<<blocks_for(ni), kBlock, 0, stream>>>(ni, b.keys2.as<uint32_t>(), b.start.as<uint32_t>());
  if ((e = hipGetLastError()) != hipSuccess) return e;
  std::vector<uint32_t> st(kLevelKeys);
  if ((e = hipMemcpyAsync(st.data(), b.start.p, sizeof(uint32_t) * kLevelKeys, hipMemcpyDeviceToHost, stream)) !=
      hipSuccess)
    return e;
  if ((e = hipStreamSynchronize(stream)) != hipSuccess) return e;
  classes.clear();
  for (uint32_t k = 0; k < kLevelKeys; k++) {
    if (st[k] == 0xFFFFFFFFu) continue;
    if (!classes.empty()) classes.back().hi = st[k];
    classes.push_back(LevelClass{st[k], ni, (k & 1u) == 0u});
  }
  return hipSuccess;
}

void map_free(rmh_map* m) {
  if (!m) return;
  if (m->vertices) (void)hipFree(m->vertices);
  if (m->faces) (void)hipFree(m->faces);
  if (m->tris) (void)hipFree(m->tris);
  if (m->normals) (void)hipFree(m->normals);
  if (m->nodes) (void)hipFree(m->nodes);
  if (m->parent) (void)hipFree(m->parent);
  if (m->reach) (void)hipFree(m->reach);
  if (m->nodes4) (void)hipFree(m->nodes4);
  if (m->refit_order) (void)hipFree(m->refit_order);
  delete m;
}

// Most levels of the 4-wide tree the ray kernels accept: a lane's stack holds at most 3
// entries per level, and the largest stack template has 128.
constexpr uint32_t kMaxWideLevels = 42;

void free_wide(rmh_map* m) {
  if (m->nodes4) (void)hipFree(m->nodes4);
  m->nodes4 = nullptr;
  m->wide_capacity = 0;
  m->root4 = m->root;
  m->info.n_wide_nodes = 0;
  m->info.wide_levels = 0;
}

// (Re)builds the 4-wide tree of m from its binary tree, in place when the existing buffer
// is large enough (a refit keeps the node count, so the buffer address stays the same).
// Frees it (the ray kernels then use the binary tree) for a single-leaf tree, for a tree
// of more than kMaxWideLevels wide levels, and for coordinates no 8-bit grid can hold.
hipError_t build_wide(rmh_map* m, hipStream_t stream) {
  const uint32_t cap = m->info.n_nodes;  // one wide node per binary node at most
  if (m->root < 0 || cap == 0 || m->n_internal == 0) { free_wide(m); return hipSuccess; }
  DevBuf fa, fb, counts, offsets, err, tmp;
  hipError_t e;
  if (!m->nodes4 || m->wide_capacity < cap) {
    free_wide(m);
    if ((e = hipMalloc(&m->nodes4, sizeof(Node4) * cap)) != hipSuccess) { m->nodes4 = nullptr; return e; }
    m->wide_capacity = cap;
  }
  if ((e = hipMalloc(&fa.p, sizeof(int) * cap)) != hipSuccess) return e;
  if ((e = hipMalloc(&fb.p, sizeof(int) * cap)) != hipSuccess) return e;
  if ((e = hipMalloc(&counts.p, sizeof(uint32_t) * cap)) != hipSuccess) return e;
  if ((e = hipMalloc(&offsets.p, sizeof(uint32_t) * cap)) != hipSuccess) return e;
  if ((e = hipMalloc(&err.p, sizeof(uint32_t))) != hipSuccess) return e;
  size_t tmp_bytes = 0;
  if ((e = rocprim::exclusive_scan(nullptr, tmp_bytes, counts.as<uint32_t>(), offsets.as<uint32_t>(), 0u, cap,
                                   rocprim::plus<uint32_t>(), stream)) != hipSuccess) return e;
  if ((e = hipMalloc(&tmp.p, tmp_bytes)) != hipSuccess) return e;
  if ((e = hipMemsetAsync(err.p, 0, sizeof(uint32_t), stream)) != hipSuccess) return e;
  const int root = 0;
  if ((e = hipMemcpyAsync(fa.p, &root, sizeof(int), hipMemcpyHostToDevice, stream)) != hipSuccess) return e;
  int* front = fa.as<int>();
  int* next = fb.as<int>();
  uint32_t n = 1, base = 0, levels = 0;
  bool ok = true;
  while (n > 0) {
    const uint32_t next_base = base + n;
    k_wide_count<<<blocks_for(n), kBlock, 0, stream>>>(m->nodes, front, n, counts.as<uint32_t>());
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if ((e = rocprim::exclusive_scan(tmp.p, tmp_bytes, counts.as<uint32_t>(), offsets.as<uint32_t>(), 0u, n,
                                     rocprim::plus<uint32_t>(), stream)) != hipSuccess) return e;
    k_wide_emit<<<blocks_for(n), kBlock, 0, stream>>>(m->nodes, front, n, base, next_base, cap, offsets.as<uint32_t>(),
                                                      m->nodes4, next, err.as<uint32_t>());
    if ((e = hipGetLastError()) != hipSuccess) return e;
    uint32_t h[3] = {0, 0, 0};
    if ((e = hipMemcpyAsync(&h[0], offsets.as<uint32_t>() + (n - 1), sizeof(uint32_t), hipMemcpyDeviceToHost,
                            stream)) != hipSuccess) return e;
    if ((e = hipMemcpyAsync(&h[1], counts.as<uint32_t>() + (n - 1), sizeof(uint32_t), hipMemcpyDeviceToHost,
                            stream)) != hipSuccess) return e;
    if ((e = hipMemcpyAsync(&h[2], err.p, sizeof(uint32_t), hipMemcpyDeviceToHost, stream)) != hipSuccess) return e;
    if ((e = hipStreamSynchronize(stream)) != hipSuccess) return e;
    levels++;
    const uint64_t children = static_cast<uint64_t>(h[0]) + h[1];
    if (h[2] != 0u || next_base + children > cap || levels > kMaxWideLevels) { ok = false; break; }
    base = next_base;
    n = static_cast<uint32_t>(children);
    std::swap(front, next);
  }
  if (!ok) {
    free_wide(m);
    return hipSuccess;
  }
  m->root4 = 0;
  m->info.n_wide_nodes = base;
  m->info.wide_levels = levels;
  return hipSuccess;
}

#define BUILD_TRY(expr)                                             \
  do {                                                              \
    hipError_t _e = (expr);                                         \
    if (_e != hipSuccess) { map_free(m); return hip_fail(_e, #expr); } \
  } while (0)

}  // namespace

// Rewrites m->nodes / m->parent / m->reach (ni radix-tree slots, n_reach of them reachable,
// root at slot 0, size = interior nodes of the final tree per subtree) into preorder, in
// buffers of exactly n_reach entries.
static hipError_t preorder_nodes(rmh_map* m, uint32_t ni, uint32_t n_reach, const uint32_t* size,
                                 hipStream_t stream) {
  DevBuf rank, err;
  Node2* nodes = nullptr;
  int* parent = nullptr;
  uint8_t* reach = nullptr;
  hipError_t e = hipMalloc(&rank.p, sizeof(uint32_t) * ni);
  if (e == hipSuccess) e = hipMalloc(&err.p, sizeof(uint32_t));
  if (e == hipSuccess) e = hipMalloc(&nodes, sizeof(Node2) * n_reach);
  if (e == hipSuccess) e = hipMalloc(&parent, sizeof(int) * n_reach);
  if (e == hipSuccess) e = hipMalloc(&reach, n_reach);
  if (e == hipSuccess) e = hipMemsetAsync(err.p, 0, sizeof(uint32_t), stream);
  if (e == hipSuccess) {
    k_preorder_rank<<<blocks_for(ni), kBlock, 0, stream>>>(ni, m->reach, m->parent, m->nodes, size,
                                                           rank.as<uint32_t>(), err.as<uint32_t>());
    e = hipGetLastError();
  }
  if (e == hipSuccess) {
    k_preorder_scatter<<<blocks_for(ni), kBlock, 0, stream>>>(ni, n_reach, m->reach, m->parent, m->nodes,
                                                              rank.as<uint32_t>(), nodes, parent, reach, err.as<uint32_t>());
    e = hipGetLastError();
  }
  uint32_t herr = 0, root_size = 0;
  if (e == hipSuccess) e = hipMemcpyAsync(&herr, err.p, sizeof(uint32_t), hipMemcpyDeviceToHost, stream);
  if (e == hipSuccess) e = hipMemcpyAsync(&root_size, size, sizeof(uint32_t), hipMemcpyDeviceToHost, stream);
  if (e == hipSuccess) e = hipStreamSynchronize(stream);
  // a walk that did not end at the root, or a count that differs from the reachable
  // nodes, means the links do not form one tree: the build fails
  if (e == hipSuccess && (herr != 0 || root_size != n_reach)) e = hipErrorInvalidValue;
  if (e != hipSuccess) {
    if (nodes) (void)hipFree(nodes);
    if (parent) (void)hipFree(parent);
    if (reach) (void)hipFree(reach);
    return e;
  }
  (void)hipFree(m->nodes);
  (void)hipFree(m->parent);
  (void)hipFree(m->reach);
  m->nodes = nodes;
  m->parent = parent;
  m->reach = reach;
  m->n_internal = n_reach;
  return hipSuccess;
}

static int map_build(const float* vertices, uint32_t nv, const uint32_t* faces, uint32_t nf,
                     const uint32_t* mesh_face_offsets, uint32_t n_meshes, const rmh_bvh_params* params,
                     void* stream_, rmh_map** out);

// The ray kernels hold one stack entry per tree level, at most 128.  The radix tree is at
// most 95 levels deep (63 key bits, then at most 32 levels of index bits among equal keys),
// but treelet restructuring may lengthen paths; a refined tree deeper than 128 levels is
// replaced by the unrefined one, so every map the build accepts can be traced.
extern "C" int rmh_map_build(const float* vertices, uint32_t nv, const uint32_t* faces, uint32_t nf,
                             const uint32_t* mesh_face_offsets, uint32_t n_meshes,
                             const rmh_bvh_params* params, void* stream_, rmh_map** out) {
  const int rc = map_build(vertices, nv, faces, nf, mesh_face_offsets, n_meshes, params, stream_, out);
  if (rc != RMH_OK || (*out)->info.max_depth <= kMaxBinaryDepth) return rc;
  if (params && (params->flags & RMH_BVH_NO_REFINE)) return rc;  // not reached, by the bound above
  map_free(*out);
  *out = nullptr;
  rmh_bvh_params prm{4u, 1.0f, 1.5f, 0u};
  if (params) prm = *params;
  prm.flags |= RMH_BVH_NO_REFINE;
  return map_build(vertices, nv, faces, nf, mesh_face_offsets, n_meshes, &prm, stream_, out);
}

static int map_build(const float* vertices, uint32_t nv, const uint32_t* faces, uint32_t nf,
                     const uint32_t* mesh_face_offsets, uint32_t n_meshes, const rmh_bvh_params* params,
                     void* stream_, rmh_map** out) {
  if (!out) return fail(RMH_ERR_INVALID, "rmh_map_build: out is NULL");
  *out = nullptr;
  if (nf > 0 && (!vertices || !faces || nv == 0))
    return fail(RMH_ERR_INVALID, "rmh_map_build: faces given without vertices");
  // The ray kernels address a node as a 32-bit byte offset (index << 6, 64-byte nodes), so
  // every tree must have fewer than 2^26 interior nodes; a binary tree over n_faces leaves
  // has at most n_faces - 1 of them (and the 4-wide tree fewer still).
  if (nf > kMaxFaces)
    return fail(RMH_ERR_UNSUPPORTED, "rmh_map_build: at most 2^26 faces (BVH node offsets are 32-bit)");
  rmh_bvh_params prm{4u, 1.0f, 1.5f, 0u};
  if (params) prm = *params;
  if (prm.max_leaf_size < 1 || prm.max_leaf_size > static_cast<uint32_t>(kMaxLeaf))
    return fail(RMH_ERR_INVALID, "rmh_map_build: max_leaf_size must be in [1, 8]");
  std::vector<uint32_t> off;
  if (mesh_face_offsets && n_meshes > 0) {
    off.assign(mesh_face_offsets, mesh_face_offsets + n_meshes + 1);
    if (off.front() != 0 || off.back() != nf)
      return fail(RMH_ERR_INVALID, "rmh_map_build: mesh_face_offsets must start at 0 and end at n_faces");
    for (uint32_t g = 0; g < n_meshes; g++)
      if (off[g] > off[g + 1]) return fail(RMH_ERR_INVALID, "rmh_map_build: mesh_face_offsets not sorted");
  } else {
    n_meshes = 1;
    off = {0u, nf};
  }
  hipStream_t stream = static_cast<hipStream_t>(stream_);
  rmh_map* m = new (std::nothrow) rmh_map();
  if (!m) return fail(RMH_ERR_ALLOC, "rmh_map_build: out of host memory");
  m->n_vertices = nv;
  m->n_faces = nf;
  m->n_meshes = n_meshes;
  (void)hipGetDevice(&m->device);
  m->info.n_vertices = nv;
  m->info.n_faces = nf;
  m->info.n_meshes = n_meshes;
  m->mesh_offsets = off;
  if (nf == 0) {  // empty map: every ray misses
    m->root = 0;
    *out = m;
    return RMH_OK;
  }

  hipEvent_t ev0, ev1;
  BUILD_TRY(hipEventCreate(&ev0));
  BUILD_TRY(hipEventCreate(&ev1));
  struct EvGuard { hipEvent_t a, b; ~EvGuard() { (void)hipEventDestroy(a); (void)hipEventDestroy(b); } } evg{ev0, ev1};

  BUILD_TRY(hipMalloc(&m->vertices, sizeof(float) * 3 * nv));
  BUILD_TRY(hipMalloc(&m->faces, sizeof(uint32_t) * 3 * nf));
  BUILD_TRY(hipMemcpyAsync(m->vertices, vertices, sizeof(float) * 3 * nv, hipMemcpyDeviceToDevice, stream));
  BUILD_TRY(hipMemcpyAsync(m->faces, faces, sizeof(uint32_t) * 3 * nf, hipMemcpyDeviceToDevice, stream));

  const int n = static_cast<int>(nf);
  const uint32_t ni = nf > 1 ? nf - 1 : 0;
  DevBuf flo, fhi, bounds, err, keys, keys2, vals, vals2, child, range, pint, pleaf, nlo, nhi, coll, stats, moff, tmp,
      subtree;
  LevelBufs lb;                    // scratch of the level-by-level passes (make_levels)
  std::vector<LevelClass> levels;  // their launch ranges
  BUILD_TRY(hipMalloc(&flo.p, sizeof(float4) * nf));
  BUILD_TRY(hipMalloc(&fhi.p, sizeof(float4) * nf));
  BUILD_TRY(hipMalloc(&bounds.p, sizeof(float) * 8));
  BUILD_TRY(hipMalloc(&err.p, sizeof(uint32_t) * 4));
  BUILD_TRY(hipMalloc(&keys.p, sizeof(uint64_t) * nf));
  BUILD_TRY(hipMalloc(&keys2.p, sizeof(uint64_t) * nf));
  BUILD_TRY(hipMalloc(&vals.p, sizeof(uint32_t) * nf));
  BUILD_TRY(hipMalloc(&vals2.p, sizeof(uint32_t) * nf));
  BUILD_TRY(hipMalloc(&child.p, sizeof(uint32_t) * 2 * (ni + 1)));
  BUILD_TRY(hipMalloc(&range.p, sizeof(uint32_t) * 2 * (ni + 1)));
  BUILD_TRY(hipMalloc(&pint.p, sizeof(int) * (ni + 1)));
  BUILD_TRY(hipMalloc(&pleaf.p, sizeof(int) * nf));
  BUILD_TRY(hipMalloc(&nlo.p, sizeof(float4) * (ni + 1)));
  BUILD_TRY(hipMalloc(&nhi.p, sizeof(float4) * (ni + 1)));
  BUILD_TRY(hipMalloc(&coll.p, sizeof(uint8_t) * (ni + 1)));
  BUILD_TRY(hipMalloc(&subtree.p, sizeof(uint32_t) * (ni + 1)));
  if (ni > 0) BUILD_TRY(level_bufs(lb, ni, stream));
  BUILD_TRY(hipMalloc(&stats.p, sizeof(uint32_t) * 4));
  BUILD_TRY(hipMalloc(&moff.p, sizeof(uint32_t) * (n_meshes + 1)));
  BUILD_TRY(hipMalloc(&m->tris, sizeof(float4) * 3 * static_cast<size_t>(nf)));
  BUILD_TRY(hipMalloc(&m->normals, sizeof(float4) * nf));
  BUILD_TRY(hipMalloc(&m->nodes, sizeof(Node2) * (ni + 1)));
  BUILD_TRY(hipMalloc(&m->parent, sizeof(int) * (ni + 1)));
  BUILD_TRY(hipMalloc(&m->reach, sizeof(uint8_t) * (ni + 1)));
  BUILD_TRY(hipMemcpy(moff.p, off.data(), sizeof(uint32_t) * (n_meshes + 1), hipMemcpyHostToDevice));

  BUILD_TRY(hipEventRecord(ev0, stream));
  k_init<<<1, 64, 0, stream>>>(bounds.as<float>(), err.as<uint32_t>());
  k_prep<<<blocks_for(nf), kBlock, 0, stream>>>(m->vertices, nv, m->faces, nf, flo.as<float4>(), fhi.as<float4>(),
                                                bounds.as<float>(), err.as<uint32_t>());
  BUILD_TRY(hipGetLastError());
  uint32_t herr = 0;
  BUILD_TRY(hipMemcpyAsync(&herr, err.p, sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
  BUILD_TRY(hipStreamSynchronize(stream));
  if (herr & 1u) { map_free(m); return fail(RMH_ERR_INVALID, "rmh_map_build: face index >= n_vertices"); }
  if (herr & 2u) { map_free(m); return fail(RMH_ERR_INVALID, "rmh_map_build: non-finite vertex coordinate"); }

  k_morton<<<blocks_for(nf), kBlock, 0, stream>>>(flo.as<float4>(), fhi.as<float4>(), nf, bounds.as<float>(),
                                                  keys.as<uint64_t>(), vals.as<uint32_t>());
  BUILD_TRY(hipGetLastError());
  size_t tmp_bytes = 0;
  BUILD_TRY(rocprim::radix_sort_pairs(nullptr, tmp_bytes, keys.as<uint64_t>(), keys2.as<uint64_t>(),
                                      vals.as<uint32_t>(), vals2.as<uint32_t>(), nf, 0u, 63u, stream));
  BUILD_TRY(hipMalloc(&tmp.p, tmp_bytes));
  BUILD_TRY(rocprim::radix_sort_pairs(tmp.p, tmp_bytes, keys.as<uint64_t>(), keys2.as<uint64_t>(),
                                      vals.as<uint32_t>(), vals2.as<uint32_t>(), nf, 0u, 63u, stream));
  uint64_t* skeys = keys2.as<uint64_t>();
  uint32_t* svals = vals2.as<uint32_t>();

  BUILD_TRY(hipMemsetAsync(stats.p, 0, sizeof(uint32_t) * 4, stream));
  DevBuf lbox, tbox, tcost, tcount, pos, offs, vals3, pleaf2;
  if (n >= 2) {
    k_karras<<<blocks_for(ni), kBlock, 0, stream>>>(skeys, n, child.as<uint32_t>(), range.as<uint32_t>(),
                                                    pint.as<int>(), pleaf.as<int>());
    BUILD_TRY(hipGetLastError());
    if (!(prm.flags & RMH_BVH_NO_REFINE) && n >= kTreeletLeaves) {
      // SAH refinement of the radix tree, then depth-first leaf order so that every
      // subtree is again one range of the triangle array
      BUILD_TRY(hipMalloc(&lbox.p, sizeof(float) * 6 * static_cast<size_t>(nf)));
      BUILD_TRY(hipMalloc(&tbox.p, sizeof(float) * 6 * static_cast<size_t>(ni)));
      BUILD_TRY(hipMalloc(&tcost.p, sizeof(float) * ni));
      BUILD_TRY(hipMalloc(&tcount.p, sizeof(uint32_t) * ni));
      BUILD_TRY(hipMalloc(&pos.p, sizeof(uint32_t) * nf));
      BUILD_TRY(hipMalloc(&offs.p, sizeof(uint32_t) * ni));
      BUILD_TRY(hipMalloc(&vals3.p, sizeof(uint32_t) * nf));
      BUILD_TRY(hipMalloc(&pleaf2.p, sizeof(int) * nf));
      TreeArrays ta;
      ta.child = child.as<uint32_t>();
      ta.parent_int = pint.as<int>();
      ta.parent_leaf = pleaf.as<int>();
      ta.box = tbox.as<float>();
      ta.cost = tcost.as<float>();
      ta.count = tcount.as<uint32_t>();
      ta.leaf_box = lbox.as<float>();
      k_leaf_boxes<<<blocks_for(nf), kBlock, 0, stream>>>(nf, svals, flo.as<float4>(), fhi.as<float4>(), lbox.as<float>());
      BUILD_TRY(hipGetLastError());
      BUILD_TRY(hipMemsetAsync(err.p, 0, sizeof(uint32_t), stream));
      // triangle counts of the radix tree decide the first round's treelets; each round
      // leaves every count current for the next
      k_range_counts<<<blocks_for(ni), kBlock, 0, stream>>>(ni, range.as<uint32_t>(), ta.count);
      BUILD_TRY(hipGetLastError());
      for (int r = 0; r < kTreeletRounds; r++) {
        // the depths change with every round's restructuring
        BUILD_TRY(make_levels(n, pint.as<int>(), ta.count, static_cast<uint32_t>(kTreeletLeaves), lb, err.as<uint32_t>(),
                              stream, levels));
        for (const LevelClass& c : levels) {
          if (c.treelet)
            k_treelet_nodes<<<c.hi - c.lo, 64, 0, stream>>>(lb.order.as<uint32_t>(), c.lo, ta, prm.sah_cost_node,
                                                            prm.sah_cost_tri);
          else
            k_refresh_nodes<<<blocks_for(c.hi - c.lo), kBlock, 0, stream>>>(lb.order.as<uint32_t>(), c.lo, c.hi, ta,
                                                                             prm.sah_cost_node, prm.sah_cost_tri);
          BUILD_TRY(hipGetLastError());
        }
      }
      {
        DevBuf refs;
        BUILD_TRY(hipMalloc(&refs.p, sizeof(uint32_t) * (2 * static_cast<size_t>(nf) - 1)));
        BUILD_TRY(hipMemsetAsync(refs.p, 0, sizeof(uint32_t) * (2 * static_cast<size_t>(nf) - 1), stream));
        k_check_tree<<<blocks_for(nf), kBlock, 0, stream>>>(n, ta, refs.as<uint32_t>(), err.as<uint32_t>());
        BUILD_TRY(hipGetLastError());
        k_check_refs<<<blocks_for(2 * static_cast<uint64_t>(nf)), kBlock, 0, stream>>>(n, refs.as<uint32_t>(),
                                                                                       err.as<uint32_t>());
        BUILD_TRY(hipGetLastError());
        uint32_t terr = 0;
        BUILD_TRY(hipMemcpyAsync(&terr, err.p, sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
        BUILD_TRY(hipStreamSynchronize(stream));
        if (terr) { map_free(m); return fail(RMH_ERR_UNSUPPORTED, "rmh_map_build: BVH refinement produced an invalid tree"); }
      }
      k_dfs_positions<<<blocks_for(nf), kBlock, 0, stream>>>(n, ta, pos.as<uint32_t>(), offs.as<uint32_t>());
      BUILD_TRY(hipGetLastError());
      k_apply_order<<<blocks_for(nf), kBlock, 0, stream>>>(n, ta, pos.as<uint32_t>(), offs.as<uint32_t>(), svals,
                                                           vals3.as<uint32_t>(), pleaf2.as<int>(), range.as<uint32_t>());
      BUILD_TRY(hipGetLastError());
      svals = vals3.as<uint32_t>();
      std::swap(pleaf.p, pleaf2.p);
    }
    BUILD_TRY(make_levels(n, pint.as<int>(), nullptr, 0u, lb, err.as<uint32_t>(), stream, levels));
    for (const LevelClass& c : levels) {
      const uint32_t cnt = c.hi - c.lo;
      k_bottom_up_level<<<blocks_for(cnt), kBlock, 0, stream>>>(lb.order.as<uint32_t>(), c.lo, c.hi, svals,
                                                                flo.as<float4>(), fhi.as<float4>(), child.as<uint32_t>(),
                                                                nlo.as<float4>(), nhi.as<float4>(), coll.as<uint8_t>(),
                                                                subtree.as<uint32_t>(), prm.max_leaf_size,
                                                                prm.sah_cost_node, prm.sah_cost_tri);
      BUILD_TRY(hipGetLastError());
    }
    k_layout<<<blocks_for(ni), kBlock, 0, stream>>>(n, svals, flo.as<float4>(), fhi.as<float4>(), child.as<uint32_t>(),
                                                    range.as<uint32_t>(), pint.as<int>(), nlo.as<float4>(),
                                                    nhi.as<float4>(), coll.as<uint8_t>(), m->nodes, stats.as<uint32_t>(),
                                                    m->reach);
    BUILD_TRY(hipGetLastError());
    BUILD_TRY(hipMemcpyAsync(m->parent, pint.p, sizeof(int) * ni, hipMemcpyDeviceToDevice, stream));
  }
  k_gather<<<blocks_for(nf), kBlock, 0, stream>>>(nf, svals, m->vertices, m->faces, moff.as<uint32_t>(), n_meshes,
                                                  m->tris, m->normals);
  BUILD_TRY(hipGetLastError());
  BUILD_TRY(hipEventRecord(ev1, stream));
  BUILD_TRY(hipStreamSynchronize(stream));
  float ms = 0.0f;
  BUILD_TRY(hipEventElapsedTime(&ms, ev0, ev1));

  uint32_t hstats[4] = {0, 0, 0, 0};
  uint8_t root_collapsed = 0;
  float4 root_lo = make_float4(0, 0, 0, 0), root_hi = make_float4(0, 0, 0, 0);
  if (n >= 2) {
    BUILD_TRY(hipMemcpy(hstats, stats.p, sizeof(hstats), hipMemcpyDeviceToHost));
    BUILD_TRY(hipMemcpy(&root_collapsed, coll.p, 1, hipMemcpyDeviceToHost));
    BUILD_TRY(hipMemcpy(&root_lo, nlo.p, sizeof(float4), hipMemcpyDeviceToHost));
    BUILD_TRY(hipMemcpy(&root_hi, nhi.p, sizeof(float4), hipMemcpyDeviceToHost));
  }
  m->n_internal = ni;
  if (n == 1 || root_collapsed) {
    m->root = make_leaf_ref(0u, nf);
    m->info.n_nodes = 0;
    m->info.n_leaves = 1;
    m->info.max_depth = 0;
  } else {
    m->root = 0;
    m->info.n_nodes = hstats[0];
    m->info.n_leaves = hstats[0] + 1;
    m->info.max_depth = hstats[1];
  }
  {
    const float dx = root_hi.x - root_lo.x, dy = root_hi.y - root_lo.y, dz = root_hi.z - root_lo.z;
    const float area = dx * dy + dy * dz + dz * dx;
    m->info.sah_cost = (n >= 2 && area > 0.0f) ? root_lo.w / area : 0.0f;
  }
  if (m->root == 0 && hstats[0] > 0) {
    BUILD_TRY(hipEventRecord(ev0, stream));
    BUILD_TRY(preorder_nodes(m, ni, hstats[0], subtree.as<uint32_t>(), stream));
    BUILD_TRY(hipEventRecord(ev1, stream));
    BUILD_TRY(hipStreamSynchronize(stream));
    float order_ms = 0.0f;
    BUILD_TRY(hipEventElapsedTime(&order_ms, ev0, ev1));
    ms += order_ms;
  } else {
    m->n_internal = 0;
  }
  m->root4 = m->root;
  if (prm.flags & RMH_BVH_WIDE) {
    BUILD_TRY(hipEventRecord(ev0, stream));
    BUILD_TRY(build_wide(m, stream));
    BUILD_TRY(hipEventRecord(ev1, stream));
    BUILD_TRY(hipStreamSynchronize(stream));
    float wide_ms = 0.0f;
    BUILD_TRY(hipEventElapsedTime(&wide_ms, ev0, ev1));
    ms += wide_ms;
  }
  m->info.build_ms = ms;
  m->info.device_bytes = sizeof(float4) * 4 * static_cast<uint64_t>(nf) + sizeof(Node2) * static_cast<uint64_t>(m->n_internal) +
                         sizeof(Node4) * static_cast<uint64_t>(m->info.n_wide_nodes);
  *out = m;
  return RMH_OK;
}

extern "C" int rmh_map_destroy(rmh_map* map) {
  map_free(map);
  return RMH_OK;
}

extern "C" int rmh_map_set_mesh_ids(rmh_map* map, const uint32_t* geom_ids, const uint32_t* object_ids,
                                    uint32_t n_meshes) {
  if (!map) return fail(RMH_ERR_INVALID, "rmh_map_set_mesh_ids: NULL map");
  if (n_meshes != map->n_meshes)
    return fail(RMH_ERR_INVALID, "rmh_map_set_mesh_ids: n_meshes differs from the map's mesh count");
  if (map->n_faces == 0 || (!geom_ids && !object_ids)) return RMH_OK;
  const uint32_t n = map->n_meshes;
  std::vector<uint32_t> host;
  host.reserve(3 * static_cast<size_t>(n) + 1);
  host.insert(host.end(), map->mesh_offsets.begin(), map->mesh_offsets.end());      // n + 1
  if (geom_ids) host.insert(host.end(), geom_ids, geom_ids + n);
  if (object_ids) host.insert(host.end(), object_ids, object_ids + n);
  uint32_t* dev = nullptr;
  RMH_HIP_TRY(hipMalloc(&dev, sizeof(uint32_t) * host.size()));
  hipError_t e = hipMemcpy(dev, host.data(), sizeof(uint32_t) * host.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    const uint32_t* g = geom_ids ? dev + n + 1 : nullptr;
    const uint32_t* o = object_ids ? dev + n + 1 + (geom_ids ? n : 0) : nullptr;
    k_set_ids<<<blocks_for(map->n_faces), kBlock>>>(map->n_faces, dev, n, g, o, map->tris, map->normals);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipDeviceSynchronize();
  (void)hipFree(dev);
  if (e != hipSuccess) return hip_fail(e, "rmh_map_set_mesh_ids");
  return RMH_OK;
}

// ---- refit: new vertex positions, same topology -------------------------------------------------

__global__ __launch_bounds__(kBlock) void k_refit_tris(uint32_t nf, const float* __restrict__ V, const uint32_t* __restrict__ F,
                                                       float4* __restrict__ tris, float4* __restrict__ normals) {
  const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
  if (k >= nf) return;
  float4 t0 = tris[3 * static_cast<size_t>(k)], t1 = tris[3 * static_cast<size_t>(k) + 1], t2 = tris[3 * static_cast<size_t>(k) + 2];
  const uint32_t f = __float_as_uint(t2.w);
  const uint32_t i0 = F[3 * f], i1 = F[3 * f + 1], i2 = F[3 * f + 2];
  t0.x = V[3 * i0]; t0.y = V[3 * i0 + 1]; t0.z = V[3 * i0 + 2];
  t1.x = V[3 * i1]; t1.y = V[3 * i1 + 1]; t1.z = V[3 * i1 + 2];
  t2.x = V[3 * i2]; t2.y = V[3 * i2 + 1]; t2.z = V[3 * i2 + 2];
  tris[3 * static_cast<size_t>(k)] = t0;
  tris[3 * static_cast<size_t>(k) + 1] = t1;
  tris[3 * static_cast<size_t>(k) + 2] = t2;
  const double ax = (double)t1.x - (double)t0.x, ay = (double)t1.y - (double)t0.y, az = (double)t1.z - (double)t0.z;
  const double bx = (double)t2.x - (double)t0.x, by = (double)t2.y - (double)t0.y, bz = (double)t2.z - (double)t0.z;
  const double nx = ay * bz - az * by, ny = az * bx - ax * bz, nz = ax * by - ay * bx;
  const double len = sqrt(nx * nx + ny * ny + nz * nz);
  float4 nrm = normals[k];
  nrm.x = len > 0.0 ? (float)(nx / len) : 0.0f;
  nrm.y = len > 0.0 ? (float)(ny / len) : 0.0f;
  nrm.z = len > 0.0 ? (float)(nz / len) : 0.0f;
  normals[k] = nrm;
}

__device__ inline void set_child_box(Node2& nd, int c, float4 lo, float4 hi) {
  if (c == 0) {
    nd.a.x = pad_lo(lo.x); nd.a.y = pad_hi(hi.x); nd.a.z = pad_lo(lo.y); nd.a.w = pad_hi(hi.y);
    nd.c.x = pad_lo(lo.z); nd.c.y = pad_hi(hi.z);
  } else {
    nd.b.x = pad_lo(lo.x); nd.b.y = pad_hi(hi.x); nd.b.z = pad_lo(lo.y); nd.b.w = pad_hi(hi.y);
    nd.c.z = pad_lo(lo.z); nd.c.w = pad_hi(hi.z);
  }
}

// Depth of every interior node of the final tree (0 at the root; unreachable slots get
// UINT32_MAX).  The parent chain of a valid tree ends at the root.
__global__ __launch_bounds__(kBlock) void k_node_depth(uint32_t ni, const uint8_t* __restrict__ reach,
                                                       const int* __restrict__ parent, uint32_t* __restrict__ depth) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= ni) return;
  if (!reach[i]) { depth[i] = 0xFFFFFFFFu; return; }
  uint32_t d = 0;
  for (int p = parent[i]; p >= 0; p = parent[p]) d++;
  depth[i] = d;
}

// One depth of the level-by-level refit: every node of the level boxes its two children,
// leaf slots from their triangles, interior slots as the union of the child's own two
// (already padded) boxes, refitted by the previous launch.  Each node is written by one
// thread and read only by later launches, so no atomics or fences are needed (agent-scope
// release and acquire write back and invalidate the L2 of an XCD on every such atomic,
// which made the bottom-up pass with arrival counters 4 ms on the terrain).
__global__ __launch_bounds__(kBlock) void k_refit_level(const uint32_t* __restrict__ order, uint32_t lo, uint32_t hi,
                                                        const float4* __restrict__ tris, Node2* nodes) {
  const uint32_t k = lo + blockIdx.x * kBlock + threadIdx.x;
  if (k >= hi) return;
  const uint32_t i = order[k];
  Node2 nd = nodes[i];
  for (int c = 0; c < 2; c++) {
    const int ref = c ? nd.d.y : nd.d.x;
    float4 blo, bhi;
    if (ref < 0) {
      const uint32_t e = ~static_cast<uint32_t>(ref);
      const uint32_t first = e >> 3, cnt = (e & 7u) + 1u;
      blo = make_float4(INFINITY, INFINITY, INFINITY, 0.0f);
      bhi = make_float4(-INFINITY, -INFINITY, -INFINITY, 0.0f);
      for (uint32_t t = first; t < first + cnt; t++)
        for (int v = 0; v < 3; v++) {
          const float4 p = tris[3 * static_cast<size_t>(t) + v];
          blo.x = fminf(blo.x, p.x); blo.y = fminf(blo.y, p.y); blo.z = fminf(blo.z, p.z);
          bhi.x = fmaxf(bhi.x, p.x); bhi.y = fmaxf(bhi.y, p.y); bhi.z = fmaxf(bhi.z, p.z);
        }
    } else {
      const Node2 ch = nodes[ref];
      blo = make_float4(fminf(ch.a.x, ch.b.x), fminf(ch.a.z, ch.b.z), fminf(ch.c.x, ch.c.z), 0.0f);
      bhi = make_float4(fmaxf(ch.a.y, ch.b.y), fmaxf(ch.a.w, ch.b.w), fmaxf(ch.c.y, ch.c.w), 0.0f);
    }
    set_child_box(nd, c, blo, bhi);
  }
  nodes[i].a = nd.a;
  nodes[i].b = nd.b;
  nodes[i].c = nd.c;
}

// The refit schedule of a map (once: the tree's topology never changes after the build).
hipError_t make_refit_schedule(rmh_map* m, hipStream_t stream) {
  const uint32_t ni = m->n_internal;
  uint32_t* depth = nullptr;
  hipError_t e = hipMalloc(&depth, sizeof(uint32_t) * ni);
  if (e != hipSuccess) return e;
  k_node_depth<<<blocks_for(ni), kBlock, 0, stream>>>(ni, m->reach, m->parent, depth);
  std::vector<uint32_t> hd(ni);
  e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(hd.data(), depth, sizeof(uint32_t) * ni, hipMemcpyDeviceToHost, stream);
  if (e == hipSuccess) e = hipStreamSynchronize(stream);
  (void)hipFree(depth);
  if (e != hipSuccess) return e;
  uint32_t max_d = 0, n = 0;
  for (uint32_t d : hd)
    if (d != 0xFFFFFFFFu) { max_d = std::max(max_d, d); n++; }
  // counting sort, deepest level first
  std::vector<uint32_t> start(max_d + 2, 0);
  for (uint32_t d : hd)
    if (d != 0xFFFFFFFFu) start[max_d - d + 1]++;
  for (uint32_t l = 1; l < start.size(); l++) start[l] += start[l - 1];
  std::vector<uint32_t> order(std::max<uint32_t>(1, n));
  std::vector<uint32_t> fill(start.begin(), start.end() - 1);
  for (uint32_t i = 0; i < ni; i++)
    if (hd[i] != 0xFFFFFFFFu) order[fill[max_d - hd[i]]++] = i;
  uint32_t* dev = nullptr;
  e = hipMalloc(&dev, sizeof(uint32_t) * order.size());
  if (e == hipSuccess) e = hipMemcpy(dev, order.data(), sizeof(uint32_t) * order.size(), hipMemcpyHostToDevice);
  if (e != hipSuccess) { if (dev) (void)hipFree(dev); return e; }
  m->refit_order = dev;
  m->refit_levels = std::move(start);
  return hipSuccess;
}

extern "C" int rmh_map_refit(rmh_map* map, const float* vertices, void* stream_) {
  if (!map) return fail(RMH_ERR_INVALID, "rmh_map_refit: NULL map");
  if (map->n_faces == 0) return RMH_OK;
  if (!vertices) return fail(RMH_ERR_INVALID, "rmh_map_refit: NULL vertices");
  hipStream_t stream = static_cast<hipStream_t>(stream_);
  RMH_HIP_TRY(hipMemcpyAsync(map->vertices, vertices, sizeof(float) * 3 * static_cast<size_t>(map->n_vertices),
                             hipMemcpyDeviceToDevice, stream));
  k_refit_tris<<<blocks_for(map->n_faces), kBlock, 0, stream>>>(map->n_faces, map->vertices, map->faces, map->tris,
                                                                 map->normals);
  RMH_HIP_TRY(hipGetLastError());
  if (map->root >= 0 && map->n_internal > 0) {
    if (!map->refit_order) {
      const hipError_t e = make_refit_schedule(map, stream);
      if (e != hipSuccess) return hip_fail(e, "rmh_map_refit: schedule");
    }
    // one launch per depth, deepest first: a level reads only what earlier launches wrote
    const std::vector<uint32_t>& lv = map->refit_levels;
    hipError_t e = hipSuccess;
    for (size_t l = 0; l + 1 < lv.size() && e == hipSuccess; l++) {
      if (lv[l + 1] == lv[l]) continue;
      k_refit_level<<<blocks_for(lv[l + 1] - lv[l]), kBlock, 0, stream>>>(map->refit_order, lv[l], lv[l + 1],
                                                                           map->tris, map->nodes);
      e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    if (e != hipSuccess) return hip_fail(e, "rmh_map_refit");
    // the 4-wide tree is collapsed again from the refitted binary boxes
    if (map->nodes4) {
      const uint64_t before = sizeof(Node4) * static_cast<uint64_t>(map->info.n_wide_nodes);
      e = build_wide(map, stream);
      if (e != hipSuccess) {
        free_wide(map);  // never leave a half-written tree for the ray kernels
        return hip_fail(e, "rmh_map_refit: 4-wide tree");
      }
      map->info.device_bytes = map->info.device_bytes - before + sizeof(Node4) * static_cast<uint64_t>(map->info.n_wide_nodes);
    }
  }
  return RMH_OK;
}

extern "C" int rmh_map_set_layout(rmh_map* map, int wide, void* stream_) {
  if (!map) return fail(RMH_ERR_INVALID, "rmh_map_set_layout: NULL map");
  const uint64_t before = sizeof(Node4) * static_cast<uint64_t>(map->info.n_wide_nodes);
  if (wide) {
    if (!map->nodes4) {
      const hipError_t e = build_wide(map, static_cast<hipStream_t>(stream_));
      if (e != hipSuccess) {
        free_wide(map);
        return hip_fail(e, "rmh_map_set_layout");
      }
    }
  } else if (map->nodes4) {
    RMH_HIP_TRY(hipDeviceSynchronize());
    free_wide(map);
  }
  map->info.device_bytes = map->info.device_bytes - before + sizeof(Node4) * static_cast<uint64_t>(map->info.n_wide_nodes);
  return RMH_OK;
}

extern "C" int rmh_map_get_info(const rmh_map* map, rmh_map_info* info) {
  if (!map || !info) return fail(RMH_ERR_INVALID, "rmh_map_get_info: NULL argument");
  *info = map->info;
  return RMH_OK;
}
