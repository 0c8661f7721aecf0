// Host check of the BVH topology code that the device builder runs (rmh_treelet.hpp):
// builds the Karras tree of a mesh on the CPU, runs treelet restructuring rounds in
// post-order (the order the device's bottom-up pass respects), re-numbers the leaves
// depth-first, and verifies after every step that the arrays form one binary tree over
// all triangles (each node reached once from the root, parent links consistent, boxes and
// counts the unions and sums of the children), that the SAH cost never rises, and that
// every subtree covers one contiguous range of the new leaf order.
//
// usage: treelet_check <mesh.bin> [rounds]   (mesh.bin: uint32 nv, nf; float32 V[nv*3];
// uint32 F[nf*3]).  Prints the SAH cost before and after; exit status 0 when all holds.
#include "../../rmagine_amd/csrc/rmh_treelet.hpp"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <vector>

using namespace rmh;

static void require(bool ok, const std::string& what) {
  if (!ok) throw std::runtime_error("FAILED: " + what);
}

struct Tree {
  int n = 0;
  std::vector<uint32_t> child, range, count;
  std::vector<int> pint, pleaf;
  std::vector<float> box, cost, leaf_box;
  TreeArrays arrays() {
    TreeArrays t;
    t.child = child.data();
    t.parent_int = pint.data();
    t.parent_leaf = pleaf.data();
    t.box = box.data();
    t.cost = cost.data();
    t.count = count.data();
    t.leaf_box = leaf_box.data();
    return t;
  }
};

// Interior nodes in post-order (children before parents).
static std::vector<int> post_order(const Tree& T) {
  std::vector<int> out, st = {0};
  std::vector<int> order;
  while (!st.empty()) {
    const int i = st.back();
    st.pop_back();
    order.push_back(i);
    for (int c = 0; c < 2; c++) {
      const uint32_t r = T.child[2 * i + c];
      if (!(r & kTreeLeafBit)) st.push_back(static_cast<int>(r));
    }
  }
  out.assign(order.rbegin(), order.rend());
  return out;
}

static void validate(Tree& T, const char* stage) {
  const int n = T.n;
  std::vector<int> seen_int(n - 1, 0), seen_leaf(n, 0);
  std::vector<int> st = {0};
  require(T.pint[0] == -1, std::string(stage) + ": root parent");
  size_t steps = 0;
  while (!st.empty()) {
    const int i = st.back();
    st.pop_back();
    require(i >= 0 && i < n - 1, std::string(stage) + ": interior index in range");
    require(seen_int[i]++ == 0, std::string(stage) + ": interior node reached twice");
    require(++steps <= static_cast<size_t>(n), std::string(stage) + ": walk too long");
    for (int c = 0; c < 2; c++) {
      const uint32_t r = T.child[2 * i + c];
      if (r & kTreeLeafBit) {
        const uint32_t k = r & ~kTreeLeafBit;
        require(k < static_cast<uint32_t>(n), std::string(stage) + ": leaf index in range");
        require(seen_leaf[k]++ == 0, std::string(stage) + ": leaf reached twice");
        require(T.pleaf[k] == i, std::string(stage) + ": leaf parent link");
      } else {
        require(T.pint[r] == i, std::string(stage) + ": interior parent link");
        st.push_back(static_cast<int>(r));
      }
    }
  }
  for (int i = 0; i < n - 1; i++) require(seen_int[i] == 1, std::string(stage) + ": interior node unreached");
  for (int k = 0; k < n; k++) require(seen_leaf[k] == 1, std::string(stage) + ": leaf unreached");
  // boxes and counts
  TreeArrays t = T.arrays();
  for (int i = 0; i < n - 1; i++) {
    float a[6], b[6];
    ref_box(t, T.child[2 * i], a);
    ref_box(t, T.child[2 * i + 1], b);
    for (int k = 0; k < 3; k++) {
      require(T.box[6 * i + k] == std::min(a[k], b[k]), std::string(stage) + ": box lo is the union");
      require(T.box[6 * i + 3 + k] == std::max(a[3 + k], b[3 + k]), std::string(stage) + ": box hi is the union");
    }
    require(T.count[i] == ref_count(t, T.child[2 * i]) + ref_count(t, T.child[2 * i + 1]), std::string(stage) + ": count");
  }
}

int main(int argc, char** argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: treelet_check mesh.bin [rounds]\n");
    return 2;
  }
  const int rounds = argc > 2 ? std::atoi(argv[2]) : 3;
  FILE* f = std::fopen(argv[1], "rb");
  if (!f) { std::perror(argv[1]); return 2; }
  uint32_t nv = 0, nf = 0;
  if (std::fread(&nv, 4, 1, f) != 1 || std::fread(&nf, 4, 1, f) != 1) return 2;
  std::vector<float> V(3 * static_cast<size_t>(nv));
  std::vector<uint32_t> F(3 * static_cast<size_t>(nf));
  if (std::fread(V.data(), 4, V.size(), f) != V.size() || std::fread(F.data(), 4, F.size(), f) != F.size()) return 2;
  std::fclose(f);
  try {
    require(nf >= 2, "at least two faces");
    const int n = static_cast<int>(nf);
    // face boxes and the cubic Morton frame (as the device's k_prep / k_morton)
    std::vector<float> flo(3 * n), fhi(3 * n);
    float cmin[3] = {3e38f, 3e38f, 3e38f}, cmax[3] = {-3e38f, -3e38f, -3e38f};
    for (int i = 0; i < n; i++)
      for (int a = 0; a < 3; a++) {
        const float x0 = V[3 * F[3 * i] + a], x1 = V[3 * F[3 * i + 1] + a], x2 = V[3 * F[3 * i + 2] + a];
        flo[3 * i + a] = std::min(x0, std::min(x1, x2));
        fhi[3 * i + a] = std::max(x0, std::max(x1, x2));
        const float c = 0.5f * flo[3 * i + a] + 0.5f * fhi[3 * i + a];
        cmin[a] = std::min(cmin[a], c);
        cmax[a] = std::max(cmax[a], c);
      }
    const float ext = std::max(cmax[0] - cmin[0], std::max(cmax[1] - cmin[1], cmax[2] - cmin[2]));
    std::vector<std::pair<uint64_t, uint32_t>> kv(n);
    for (int i = 0; i < n; i++) kv[i] = {morton_key(&flo[3 * i], &fhi[3 * i], cmin, ext), static_cast<uint32_t>(i)};
    std::sort(kv.begin(), kv.end());
    std::vector<uint64_t> keys(n);
    std::vector<uint32_t> vals(n);
    for (int i = 0; i < n; i++) { keys[i] = kv[i].first; vals[i] = kv[i].second; }
    Tree T;
    T.n = n;
    T.child.assign(2 * (n - 1), 0);
    T.range.assign(2 * (n - 1), 0);
    T.count.assign(n - 1, 0);
    T.pint.assign(n - 1, -2);
    T.pleaf.assign(n, -2);
    T.box.assign(6 * static_cast<size_t>(n - 1), 0.0f);
    T.cost.assign(n - 1, 0.0f);
    T.leaf_box.resize(6 * static_cast<size_t>(n));
    for (int k = 0; k < n; k++)
      for (int a = 0; a < 3; a++) {
        T.leaf_box[6 * k + a] = flo[3 * vals[k] + a];
        T.leaf_box[6 * k + 3 + a] = fhi[3 * vals[k] + a];
      }
    for (int i = 0; i < n - 1; i++) karras_node(keys.data(), n, i, T.child.data(), T.range.data(), T.pint.data(), T.pleaf.data());
    const float c_node = 1.0f, c_tri = 1.5f;
    TreeArrays t = T.arrays();
    for (int i : post_order(T)) refresh_node(t, i, c_node, c_tri);
    validate(T, "karras");
    const float root_area = half_area6(&T.box[0]);
    const float sah0 = T.cost[0] / root_area;
    float prev = T.cost[0];
    long changed = 0;
    for (int r = 0; r < rounds; r++) {
      for (int i : post_order(T)) {
        refresh_node(t, i, c_node, c_tri);
        if (T.count[i] >= static_cast<uint32_t>(kTreeletLeaves)) changed += treelet_optimize(t, i, c_node, c_tri) ? 1 : 0;
      }
      validate(T, "treelet round");
      require(T.cost[0] <= prev * (1.0f + 1e-5f), "SAH cost does not rise");
      prev = T.cost[0];
    }
    // depth-first leaf numbering
    std::vector<uint32_t> pos(n), off(n - 1);
    for (int k = 0; k < n; k++) pos[k] = dfs_offset(t, static_cast<uint32_t>(k) | kTreeLeafBit);
    for (int i = 0; i < n - 1; i++) off[i] = dfs_offset(t, static_cast<uint32_t>(i));
    std::vector<int> hit(n, 0);
    for (int k = 0; k < n; k++) {
      require(pos[k] < static_cast<uint32_t>(n) && hit[pos[k]]++ == 0, "leaf positions form a permutation");
    }
    // the order of a depth-first walk (left child first) equals the positions
    {
      std::vector<uint32_t> st = {0};
      uint32_t next = 0;
      while (!st.empty()) {
        const uint32_t r = st.back();
        st.pop_back();
        if (r & kTreeLeafBit) {
          require(pos[r & ~kTreeLeafBit] == next++, "positions follow the depth-first order");
        } else {
          require(off[r] == next, "interior offset is the first leaf position");
          st.push_back(T.child[2 * r + 1]);
          st.push_back(T.child[2 * r]);
        }
      }
    }
    std::printf("{\"faces\": %d, \"sah_karras\": %.4f, \"sah_treelets\": %.4f, \"rounds\": %d, \"restructured\": %ld}\n",
                n, sah0, T.cost[0] / root_area, rounds, changed);
  } catch (const std::exception& e) {
    std::printf("%s\n", e.what());
    return 1;
  }
  return 0;
}
