// Binary BVH topology code shared by the device builder (rmh_bvh.hip) and a host check
// program (tests/cpp/treelet_check.cpp): Morton keys, the Karras radix tree, treelet
// restructuring and the depth-first re-numbering of the leaves.  Plain C++ on plain
// arrays, compiled for both the host and the device.
//
// Tree arrays (Karras index space): n leaves = triangles in sorted order, n - 1 interior
// nodes, node 0 the root.  child[2i + c] is interior node j or, with bit 31 set, leaf k;
// parent_int[i] / parent_leaf[k] point back (-1 at the root).
#pragma once

#include <cstdint>

#if defined(__HIPCC__)
#define RMH_TREE_HD __host__ __device__ inline
#else
#define RMH_TREE_HD inline
#endif

namespace rmh {

constexpr uint32_t kTreeLeafBit = 0x80000000u;

RMH_TREE_HD uint64_t split3_21(uint32_t x) {
  uint64_t v = x & 0x1fffffu;
  v = (v | (v << 32)) & 0x001f00000000ffffull;
  v = (v | (v << 16)) & 0x001f0000ff0000ffull;
  v = (v | (v << 8)) & 0x100f00f00f00f00full;
  v = (v | (v << 4)) & 0x10c30c30c30c30c3ull;
  v = (v | (v << 2)) & 0x1249249249249249ull;
  return v;
}

// 63-bit Morton key of a triangle with bounds lo/hi inside the scene's centroid bounds
// (bmin, ext = the largest extent: one cubic scale for all three axes, so a flat scene
// such as a 100 m x 100 m x 16 m terrain does not get as many z splits as x splits).
RMH_TREE_HD uint64_t morton_key(const float lo[3], const float hi[3], const float bmin[3], float ext) {
  uint32_t q[3];
  for (int k = 0; k < 3; k++) {
    const float c = 0.5f * lo[k] + 0.5f * hi[k];
    float u = ext > 0.0f ? (c - bmin[k]) / ext : 0.5f;
    u = u < 0.0f ? 0.0f : (u > 1.0f ? 1.0f : u);
    q[k] = static_cast<uint32_t>(u * 2097151.0f);
  }
  return (split3_21(q[0]) << 2) | (split3_21(q[1]) << 1) | split3_21(q[2]);
}

RMH_TREE_HD int clz64(uint64_t x) {
  int n = 0;
  for (uint64_t bit = 1ull << 63; bit && !(x & bit); bit >>= 1) n++;
  return n;
}

RMH_TREE_HD int clz32(uint32_t x) {
  int n = 0;
  for (uint32_t bit = 1u << 31; bit && !(x & bit); bit >>= 1) n++;
  return n;
}

// Common prefix length of keys i and j (index-augmented for equal keys), -1 outside.
RMH_TREE_HD int karras_delta(const uint64_t* keys, int n, int i, int j) {
  if (j < 0 || j >= n) return -1;
  const uint64_t a = keys[i], b = keys[j];
#if defined(__HIP_DEVICE_COMPILE__)
  if (a == b) return 64 + __clz(static_cast<uint32_t>(i ^ j));
  return __clzll(a ^ b);
#else
  if (a == b) return 64 + clz32(static_cast<uint32_t>(i ^ j));
  return clz64(a ^ b);
#endif
}

// Interior node i of the radix tree over the sorted keys: Karras, "Maximizing Parallelism
// in the Construction of BVHs, Octrees, and k-d Trees", HPG 2012, Fig. 4.
RMH_TREE_HD void karras_node(const uint64_t* keys, int n, int i, uint32_t* child, uint32_t* range, int* parent_int,
                             int* parent_leaf) {
  const int d = (karras_delta(keys, n, i, i + 1) - karras_delta(keys, n, i, i - 1)) >= 0 ? 1 : -1;
  const int dmin = karras_delta(keys, n, i, i - d);
  int lmax = 2;
  while (karras_delta(keys, n, i, i + lmax * d) > dmin) lmax <<= 1;
  int l = 0;
  for (int t = lmax >> 1; t >= 1; t >>= 1)
    if (karras_delta(keys, n, i, i + (l + t) * d) > dmin) l += t;
  const int j = i + l * d;
  const int dnode = karras_delta(keys, n, i, j);
  int s = 0;
  int t = l;
  do {
    t = (t + 1) >> 1;
    if (karras_delta(keys, n, i, i + (s + t) * d) > dnode) s += t;
  } while (t > 1);
  const int gamma = i + s * d + (d < 0 ? -1 : 0);
  const int first = i < j ? i : j, last = i < j ? j : i;
  uint32_t left, right;
  if (first == gamma) { left = static_cast<uint32_t>(gamma) | kTreeLeafBit; parent_leaf[gamma] = i; }
  else { left = static_cast<uint32_t>(gamma); parent_int[gamma] = i; }
  if (last == gamma + 1) { right = static_cast<uint32_t>(gamma + 1) | kTreeLeafBit; parent_leaf[gamma + 1] = i; }
  else { right = static_cast<uint32_t>(gamma + 1); parent_int[gamma + 1] = i; }
  child[2 * i + 0] = left;
  child[2 * i + 1] = right;
  range[2 * i + 0] = static_cast<uint32_t>(first);
  range[2 * i + 1] = static_cast<uint32_t>(last);
  if (i == 0) parent_int[0] = -1;
}

// ---- treelet restructuring -------------------------------------------------------------
//
// Karras and Aila, "Fast Parallel Construction of High-Quality Bounding Volume
// Hierarchies", HPG 2013: a treelet of up to 7 leaves is grown below a node by repeatedly
// opening its largest-area interior leaf; dynamic programming over the 127 leaf subsets
// finds the topology of least SAH cost (c_node * area per interior node + the leaves'
// subtree costs); the treelet is rebuilt with that topology, reusing its interior nodes.

constexpr int kTreeletLeaves = 7;

struct TreeArrays {
  uint32_t* child;
  int* parent_int;
  int* parent_leaf;
  float* box;             // interior nodes: lo x, y, z, hi x, y, z
  float* cost;            // interior nodes: SAH cost of the subtree (area weighted)
  uint32_t* count;        // interior nodes: triangles below
  const float* leaf_box;  // leaves (sorted order): lo x, y, z, hi x, y, z
};

RMH_TREE_HD float half_area6(const float* b) {
  const float dx = b[3] - b[0], dy = b[4] - b[1], dz = b[5] - b[2];
  return dx * dy + dy * dz + dz * dx;
}

RMH_TREE_HD void ref_box(const TreeArrays& t, uint32_t ref, float* b) {
  const float* s = (ref & kTreeLeafBit) ? t.leaf_box + 6 * static_cast<uint64_t>(ref & ~kTreeLeafBit)
                                         : t.box + 6 * static_cast<uint64_t>(ref);
  for (int k = 0; k < 6; k++) b[k] = s[k];
}

RMH_TREE_HD float ref_cost(const TreeArrays& t, uint32_t ref, float c_tri) {
  if (ref & kTreeLeafBit) return c_tri * half_area6(t.leaf_box + 6 * static_cast<uint64_t>(ref & ~kTreeLeafBit));
  return t.cost[ref];
}

RMH_TREE_HD uint32_t ref_count(const TreeArrays& t, uint32_t ref) { return (ref & kTreeLeafBit) ? 1u : t.count[ref]; }

RMH_TREE_HD void set_parent(const TreeArrays& t, uint32_t ref, int p) {
  if (ref & kTreeLeafBit) t.parent_leaf[ref & ~kTreeLeafBit] = p;
  else t.parent_int[ref] = p;
}

// Box, count and cost of interior node i from its children.
RMH_TREE_HD void refresh_node(const TreeArrays& t, int i, float c_node, float c_tri) {
  float a[6], b[6];
  const uint32_t c0 = t.child[2 * i], c1 = t.child[2 * i + 1];
  ref_box(t, c0, a);
  ref_box(t, c1, b);
  float* o = t.box + 6 * static_cast<uint64_t>(i);
  for (int k = 0; k < 3; k++) {
    o[k] = a[k] < b[k] ? a[k] : b[k];
    o[3 + k] = a[3 + k] > b[3 + k] ? a[3 + k] : b[3 + k];
  }
  t.count[i] = ref_count(t, c0) + ref_count(t, c1);
  t.cost[i] = c_node * half_area6(o) + ref_cost(t, c0, c_tri) + ref_cost(t, c1, c_tri);
}

RMH_TREE_HD int popcount7(uint32_t x) {
  int n = 0;
  for (; x; x &= x - 1) n++;
  return n;
}

RMH_TREE_HD int lowest_bit(uint32_t x) {
  int i = 0;
  while (!(x & 1u)) { x >>= 1; i++; }
  return i;
}

// Treelet below `root`: its leaves (subtree references) and interior nodes (inner[0] =
// root).  Returns the number of leaves (2..7).
RMH_TREE_HD int treelet_form(const TreeArrays& t, int root, uint32_t* leaf, int* inner) {
  int nl = 2, ni = 1;
  inner[0] = root;
  leaf[0] = t.child[2 * root];
  leaf[1] = t.child[2 * root + 1];
  while (nl < kTreeletLeaves) {
    int best = -1;
    float best_area = -1.0f;
    for (int j = 0; j < nl; j++) {
      if (leaf[j] & kTreeLeafBit) continue;
      const float a = half_area6(t.box + 6 * static_cast<uint64_t>(leaf[j]));
      if (a > best_area) { best_area = a; best = j; }
    }
    if (best < 0) break;
    const uint32_t x = leaf[best];
    inner[ni++] = static_cast<int>(x);
    leaf[best] = t.child[2 * x];
    leaf[nl++] = t.child[2 * x + 1];
  }
  return nl;
}

// Surface area (half) of the union of the leaf boxes in subset S.
RMH_TREE_HD float subset_area(const float (*lb)[6], int nl, uint32_t S) {
  float b[6] = {3.0e38f, 3.0e38f, 3.0e38f, -3.0e38f, -3.0e38f, -3.0e38f};
  for (int j = 0; j < nl; j++) {
    if (!(S & (1u << j))) continue;
    for (int k = 0; k < 3; k++) {
      b[k] = lb[j][k] < b[k] ? lb[j][k] : b[k];
      b[3 + k] = lb[j][3 + k] > b[3 + k] ? lb[j][3 + k] : b[3 + k];
    }
  }
  return half_area6(b);
}

// Least cost split of subset S (at least two members) from the optimal costs of its proper
// subsets: every split is seen once, as the part P that holds S's lowest member; ties keep
// the first P in this order.  Returns the cost of S; *bp = P.
RMH_TREE_HD float best_split(const float* copt, uint32_t S, float area_S, float c_node, uint32_t* bp) {
  const uint32_t low = S & (0u - S);
  float best = 3.0e38f;
  uint32_t b = low;
  for (uint32_t P = (S - 1u) & S; P > 0; P = (P - 1u) & S) {
    if (!(P & low)) continue;
    const float c = copt[P] + copt[S ^ P];
    if (c < best) { best = c; b = P; }
  }
  *bp = b;
  return c_node * area_S + best;
}

// Rebuilds the treelet with the splits part[] (subset -> the part holding its lowest member),
// reusing its interior nodes, then refreshes their boxes, counts and costs bottom-up.
RMH_TREE_HD void treelet_rebuild(const TreeArrays& t, int root, const uint32_t* leaf, const int* inner, int nl,
                                 const uint8_t* part, float c_node, float c_tri) {
  uint32_t stack_set[kTreeletLeaves];
  int stack_node[kTreeletLeaves];
  int made[kTreeletLeaves - 1];
  int sp = 0, nmade = 0, next_inner = 1;
  stack_set[sp] = (1u << nl) - 1u;
  stack_node[sp] = root;
  sp++;
  while (sp > 0) {
    sp--;
    const uint32_t S = stack_set[sp];
    const int node = stack_node[sp];
    made[nmade++] = node;
    const uint32_t P = part[S], Q = S ^ P;
    for (int c = 0; c < 2; c++) {
      const uint32_t sub = c ? Q : P;
      uint32_t ref;
      if (popcount7(sub) == 1) {
        ref = leaf[lowest_bit(sub)];
      } else {
        const int j = inner[next_inner++];
        ref = static_cast<uint32_t>(j);
        stack_set[sp] = sub;
        stack_node[sp] = j;
        sp++;
      }
      t.child[2 * node + c] = ref;
      set_parent(t, ref, node);
    }
  }
  for (int m = nmade - 1; m >= 0; m--) refresh_node(t, made[m], c_node, c_tri);
}

// Restructures the treelet rooted at interior node `root` (whose children are final), with
// the dynamic program run serially.  Returns true when the topology changed.
RMH_TREE_HD bool treelet_optimize(const TreeArrays& t, int root, float c_node, float c_tri) {
  uint32_t leaf[kTreeletLeaves];
  int inner[kTreeletLeaves - 1];
  const int nl = treelet_form(t, root, leaf, inner);
  if (nl < 3) return false;  // two leaves have one topology
  float lb[kTreeletLeaves][6];
  float copt[1 << kTreeletLeaves];
  uint8_t part[1 << kTreeletLeaves];
  for (int j = 0; j < nl; j++) ref_box(t, leaf[j], lb[j]);
  const uint32_t full = (1u << nl) - 1u;
  for (uint32_t S = 1; S <= full; S++) {
    if (popcount7(S) == 1) {
      copt[S] = ref_cost(t, leaf[lowest_bit(S)], c_tri);
      part[S] = 0;
      continue;
    }
    uint32_t bp;
    copt[S] = best_split(copt, S, subset_area(lb, nl, S), c_node, &bp);
    part[S] = static_cast<uint8_t>(bp);
  }
  if (!(copt[full] < t.cost[root] * (1.0f - 1e-6f))) return false;
  treelet_rebuild(t, root, leaf, inner, nl, part, c_node, c_tri);
  return true;
}

// Depth-first position of the first leaf below a node: the leaves of every left sibling
// on the path to the root come before it.  ref is a leaf (bit 31) or an interior node.
// The walk stops after kMaxTreeWalk steps (a checked tree is far shallower).
constexpr int kMaxTreeWalk = 4096;

RMH_TREE_HD uint32_t dfs_offset(const TreeArrays& t, uint32_t ref) {
  uint32_t pos = 0;
  int p = (ref & kTreeLeafBit) ? t.parent_leaf[ref & ~kTreeLeafBit] : t.parent_int[ref];
  uint32_t cur = ref;
  for (int steps = 0; p >= 0 && steps < kMaxTreeWalk; steps++) {
    if (t.child[2 * p + 1] == cur) pos += ref_count(t, t.child[2 * p]);
    cur = static_cast<uint32_t>(p);
    p = t.parent_int[p];
  }
  return pos;
}

}  // namespace rmh
