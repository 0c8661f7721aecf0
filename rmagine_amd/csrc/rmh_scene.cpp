// Map files of librmh: mesh file loading without Assimp, flattened into the arrays
// rmh_map_build takes, with the geometry and object ids Rmagine's import gives.
//
// Replaces import_embree_map / import_optix_map
// (src/rmagine_embree/include/rmagine/map/EmbreeMap.hpp:94-117,
//  src/rmagine_optix/include/rmagine/map/OptixMap.hpp:88-109), which read any format through
// Assimp (aiImporter::ReadFile(path, 0): no post-processing) and build the scene with
// make_embree_scene (src/rmagine_embree/src/map/embree/EmbreeScene.cpp:317-406) followed by
// EmbreeScene::freeze (EmbreeScene.cpp:246-290):
//   * every Assimp mesh with triangles becomes a mesh; a face keeps its first three indices
//     (EmbreeMesh::init, src/rmagine_embree/src/map/embree/EmbreeMesh.cpp:195-225); a mesh
//     without any triangle is skipped (EmbreeScene.cpp:334-342);
//   * every scene node with meshes, in depth-first pre-order (get_nodes_with_meshes,
//     src/rmagine_core/src/util/assimp/helper.cpp:10-32), becomes an instance of its meshes
//     under the node's global transform (helper.cpp:49-62), decomposed into rotation +
//     translation and a scale vector (decompose, src/rmagine_core/src/math/linalg.cpp:45-88);
//     instance i gets top-level id i; meshes no node refers to follow, unchanged;
//   * freeze() replaces each single-mesh instance by its mesh, transformed
//     (v' = R (s * v) + t, EmbreeMesh::apply, EmbreeMesh.cpp:297-306); the mesh takes the
//     id the instance had.  Instances of several meshes stay instances: their meshes report
//     geom id = index inside the instance and object id = the instance id.
// Here instances are flattened too (their meshes transformed into the map), keeping those
// ids.  Formats: PLY (ascii, binary little/big endian), OBJ, STL (ascii, binary) and COLLADA
// (.dae: <triangles>/<polylist>/<polygons>, node <matrix>/<translate>/<rotate>/<scale>,
// <instance_geometry>/<instance_node>, <unit meter>; the up axis is ignored, as the
// reference configures Assimp, src/rmagine_core/src/map/AssimpIO.cpp:6-12).
#include <rmh.h>

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <new>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace rmh {
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);
}  // namespace rmh

namespace {

struct LoadError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

// One Assimp-like mesh: local vertices and polygons (index lists of any length).
struct Mesh {
  std::string name;
  std::vector<double> v;                 // 3 per vertex
  std::vector<uint32_t> poly;            // concatenated indices
  std::vector<uint32_t> poly_start;      // size = polygons + 1
  void add_poly(const std::vector<uint32_t>& idx) {
    if (poly_start.empty()) poly_start.push_back(0);
    poly.insert(poly.end(), idx.begin(), idx.end());
    poly_start.push_back(static_cast<uint32_t>(poly.size()));
  }
  size_t n_poly() const { return poly_start.empty() ? 0 : poly_start.size() - 1; }
};

struct Node {
  std::string name;
  double M[16];                          // row-major local transform
  std::vector<int> meshes;
  std::vector<int> children;
  int parent = -1;
};

struct SceneData {
  std::vector<Mesh> meshes;
  std::vector<Node> nodes;               // nodes[0] = root
};

void identity(double* M) {
  for (int i = 0; i < 16; i++) M[i] = (i % 5 == 0) ? 1.0 : 0.0;
}

void matmul(const double* A, const double* B, double* C) {  // C = A * B (row-major 4x4)
  double T[16];
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) {
      double s = 0.0;
      for (int k = 0; k < 4; k++) s += A[4 * i + k] * B[4 * k + j];
      T[4 * i + j] = s;
    }
  std::memcpy(C, T, sizeof(T));
}

std::string read_file(const std::string& path) {
  std::ifstream f(path, std::ios::binary);
  if (!f) throw LoadError("cannot open '" + path + "'");
  std::ostringstream ss;
  ss << f.rdbuf();
  return ss.str();
}

std::string lower(std::string s) {
  for (auto& c : s) c = static_cast<char>(std::tolower(static_cast<unsigned char>(c)));
  return s;
}

// ---- PLY ---------------------------------------------------------------------------------

enum class PType { I8, U8, I16, U16, I32, U32, F32, F64, NONE };

PType ply_type(const std::string& t) {
  if (t == "char" || t == "int8") return PType::I8;
  if (t == "uchar" || t == "uint8") return PType::U8;
  if (t == "short" || t == "int16") return PType::I16;
  if (t == "ushort" || t == "uint16") return PType::U16;
  if (t == "int" || t == "int32") return PType::I32;
  if (t == "uint" || t == "uint32") return PType::U32;
  if (t == "float" || t == "float32") return PType::F32;
  if (t == "double" || t == "float64") return PType::F64;
  throw LoadError("PLY: unknown property type '" + t + "'");
}

int ptype_size(PType t) {
  switch (t) {
    case PType::I8: case PType::U8: return 1;
    case PType::I16: case PType::U16: return 2;
    case PType::I32: case PType::U32: case PType::F32: return 4;
    case PType::F64: return 8;
    default: return 0;
  }
}

struct PProp {
  std::string name;
  PType type = PType::NONE;       // scalar type, or list item type
  PType count_type = PType::NONE; // != NONE: list property
};

struct PElem {
  std::string name;
  uint64_t count = 0;
  std::vector<PProp> props;
};

struct BinReader {
  const unsigned char* p;
  const unsigned char* end;
  bool big;
  double get(PType t) {
    const int n = ptype_size(t);
    if (p + n > end) throw LoadError("PLY: file ends inside the data");
    unsigned char b[8];
    for (int i = 0; i < n; i++) b[i] = p[big ? n - 1 - i : i];
    p += n;
    switch (t) {
      case PType::I8: { int8_t v; std::memcpy(&v, b, 1); return v; }
      case PType::U8: return b[0];
      case PType::I16: { int16_t v; std::memcpy(&v, b, 2); return v; }
      case PType::U16: { uint16_t v; std::memcpy(&v, b, 2); return v; }
      case PType::I32: { int32_t v; std::memcpy(&v, b, 4); return v; }
      case PType::U32: { uint32_t v; std::memcpy(&v, b, 4); return v; }
      case PType::F32: { float v; std::memcpy(&v, b, 4); return v; }
      case PType::F64: { double v; std::memcpy(&v, b, 8); return v; }
      default: throw LoadError("PLY: bad type");
    }
  }
};

struct AsciiReader {
  const char* p;
  const char* end;
  double get() {
    while (p < end && std::isspace(static_cast<unsigned char>(*p))) p++;
    if (p >= end) throw LoadError("PLY: file ends inside the data");
    char* q = nullptr;
    const double v = std::strtod(p, &q);
    if (q == p) throw LoadError("PLY: bad number");
    p = q;
    return v;
  }
};

void load_ply(const std::string& data, SceneData& sc) {
  const size_t hdr_end = data.find("end_header");
  if (data.compare(0, 3, "ply") != 0 || hdr_end == std::string::npos) throw LoadError("PLY: no header");
  std::istringstream hs(data.substr(0, hdr_end));
  std::string line, format;
  std::vector<PElem> elems;
  while (std::getline(hs, line)) {
    std::istringstream ls(line);
    std::string kw;
    ls >> kw;
    if (kw == "format") {
      ls >> format;
    } else if (kw == "element") {
      PElem e;
      ls >> e.name >> e.count;
      elems.push_back(e);
    } else if (kw == "property") {
      if (elems.empty()) throw LoadError("PLY: property before element");
      PProp pr;
      std::string t;
      ls >> t;
      if (t == "list") {
        std::string ct, it;
        ls >> ct >> it >> pr.name;
        pr.count_type = ply_type(ct);
        pr.type = ply_type(it);
      } else {
        pr.type = ply_type(t);
        ls >> pr.name;
      }
      elems.back().props.push_back(pr);
    }
  }
  size_t body = data.find('\n', hdr_end);
  body = body == std::string::npos ? data.size() : body + 1;
  const bool ascii = format == "ascii";
  if (!ascii && format != "binary_little_endian" && format != "binary_big_endian")
    throw LoadError("PLY: unknown format '" + format + "'");
  BinReader br{reinterpret_cast<const unsigned char*>(data.data()) + body,
               reinterpret_cast<const unsigned char*>(data.data()) + data.size(), format == "binary_big_endian"};
  AsciiReader ar{data.data() + body, data.data() + data.size()};
  auto get = [&](PType t) { return ascii ? ar.get() : br.get(t); };
  Mesh m;
  m.name = "ply";
  std::vector<uint32_t> idx;
  for (const PElem& e : elems) {
    const bool is_v = e.name == "vertex", is_f = e.name == "face";
    int px = -1, py = -1, pz = -1, pf = -1;
    for (size_t k = 0; k < e.props.size(); k++) {
      const std::string& n = e.props[k].name;
      if (is_v && n == "x") px = static_cast<int>(k);
      if (is_v && n == "y") py = static_cast<int>(k);
      if (is_v && n == "z") pz = static_cast<int>(k);
      if (is_f && e.props[k].count_type != PType::NONE && (n == "vertex_indices" || n == "vertex_index")) pf = static_cast<int>(k);
    }
    if (is_v && (px < 0 || py < 0 || pz < 0)) throw LoadError("PLY: vertex element without x, y, z");
    if (is_v) m.v.reserve(3 * e.count);
    for (uint64_t i = 0; i < e.count; i++) {
      double xyz[3] = {0, 0, 0};
      for (size_t k = 0; k < e.props.size(); k++) {
        const PProp& pr = e.props[k];
        if (pr.count_type != PType::NONE) {
          const double c = get(pr.count_type);
          if (!(c >= 0) || c > 1e7) throw LoadError("PLY: bad list length");
          const uint64_t cnt = static_cast<uint64_t>(c);
          if (static_cast<int>(k) == pf) idx.clear();
          for (uint64_t j = 0; j < cnt; j++) {
            const double x = get(pr.type);
            if (static_cast<int>(k) == pf) {
              if (!(x >= 0) || x > 4294967295.0) throw LoadError("PLY: bad vertex index");
              idx.push_back(static_cast<uint32_t>(x));
            }
          }
          if (static_cast<int>(k) == pf) m.add_poly(idx);
        } else {
          const double x = get(pr.type);
          if (static_cast<int>(k) == px) xyz[0] = x;
          if (static_cast<int>(k) == py) xyz[1] = x;
          if (static_cast<int>(k) == pz) xyz[2] = x;
        }
      }
      if (is_v) m.v.insert(m.v.end(), xyz, xyz + 3);
    }
  }
  sc.meshes.push_back(std::move(m));
  Node root;
  root.name = "root";
  identity(root.M);
  root.meshes.push_back(0);
  sc.nodes.push_back(root);
}

// ---- OBJ ---------------------------------------------------------------------------------
// Vertices are global; `o` starts an object (a node), `g` and `usemtl` start a new mesh of
// the current object.  Face entries "v", "v/vt", "v//vn", "v/vt/vn"; negative indices count
// back from the last vertex.
void load_obj(const std::string& data, SceneData& sc) {
  std::vector<double> V;
  struct Obj { std::string name; std::vector<Mesh> meshes; };
  std::vector<Obj> objs;
  auto cur_mesh = [&]() -> Mesh& {
    if (objs.empty()) objs.push_back(Obj{"default", {}});
    if (objs.back().meshes.empty()) objs.back().meshes.push_back(Mesh{});
    return objs.back().meshes.back();
  };
  auto new_mesh = [&](const std::string& name) {
    if (objs.empty()) objs.push_back(Obj{"default", {}});
    auto& ms = objs.back().meshes;
    if (ms.empty() || ms.back().n_poly() > 0) ms.push_back(Mesh{});
    ms.back().name = name;
  };
  std::istringstream in(data);
  std::string line;
  std::vector<uint32_t> idx;
  while (std::getline(in, line)) {
    if (!line.empty() && line.back() == '\r') line.pop_back();
    std::istringstream ls(line);
    std::string kw;
    ls >> kw;
    if (kw == "v") {
      double x = 0, y = 0, z = 0;
      ls >> x >> y >> z;
      V.insert(V.end(), {x, y, z});
    } else if (kw == "f") {
      idx.clear();
      std::string tok;
      while (ls >> tok) {
        const long long i = std::strtoll(tok.c_str(), nullptr, 10);
        const long long nv = static_cast<long long>(V.size() / 3);
        const long long a = i < 0 ? nv + i : i - 1;
        if (i == 0 || a < 0 || a >= nv) throw LoadError("OBJ: face index out of range");
        idx.push_back(static_cast<uint32_t>(a));
      }
      cur_mesh().add_poly(idx);
    } else if (kw == "o") {
      std::string name;
      std::getline(ls, name);
      if (objs.empty() || objs.back().meshes.empty() ||
          std::any_of(objs.back().meshes.begin(), objs.back().meshes.end(), [](const Mesh& m) { return m.n_poly() > 0; }))
        objs.push_back(Obj{name, {}});
      else
        objs.back().name = name;
    } else if (kw == "g" || kw == "usemtl") {
      std::string name;
      std::getline(ls, name);
      new_mesh(name);
    }
  }
  Node root;
  root.name = "root";
  identity(root.M);
  sc.nodes.push_back(root);
  for (auto& o : objs) {
    Node nd;
    nd.name = o.name;
    identity(nd.M);
    nd.parent = 0;
    for (auto& m : o.meshes) {
      if (m.n_poly() == 0) continue;
      m.v = V;  // global vertex list (unused vertices are harmless)
      nd.meshes.push_back(static_cast<int>(sc.meshes.size()));
      sc.meshes.push_back(std::move(m));
    }
    if (nd.meshes.empty()) continue;
    sc.nodes[0].children.push_back(static_cast<int>(sc.nodes.size()));
    sc.nodes.push_back(nd);
  }
}

// ---- STL ---------------------------------------------------------------------------------
void load_stl(const std::string& data, SceneData& sc) {
  Mesh m;
  m.name = "stl";
  const bool binary = data.size() >= 84 && [&] {
    uint32_t n;
    std::memcpy(&n, data.data() + 80, 4);
    return data.size() == 84 + 50ull * n;
  }();
  if (binary) {
    uint32_t n;
    std::memcpy(&n, data.data() + 80, 4);
    for (uint32_t i = 0; i < n; i++) {
      const char* r = data.data() + 84 + 50ull * i + 12;
      for (int k = 0; k < 3; k++) {
        float xyz[3];
        std::memcpy(xyz, r + 12 * k, 12);
        m.v.insert(m.v.end(), {xyz[0], xyz[1], xyz[2]});
      }
      m.add_poly({3 * i, 3 * i + 1, 3 * i + 2});
    }
  } else {
    std::istringstream in(data);
    std::string kw;
    std::vector<uint32_t> idx;
    while (in >> kw) {
      if (kw == "vertex") {
        double x, y, z;
        in >> x >> y >> z;
        idx.push_back(static_cast<uint32_t>(m.v.size() / 3));
        m.v.insert(m.v.end(), {x, y, z});
      } else if (kw == "endloop" || kw == "endfacet") {
        if (!idx.empty()) m.add_poly(idx);
        idx.clear();
      }
    }
  }
  sc.meshes.push_back(std::move(m));
  Node root;
  root.name = "root";
  identity(root.M);
  root.meshes.push_back(0);
  sc.nodes.push_back(root);
}

// ---- COLLADA (minimal XML) -----------------------------------------------------------------
struct XNode {
  std::string tag;
  std::map<std::string, std::string> attr;
  std::string text;
  std::vector<std::unique_ptr<XNode>> kids;
  const XNode* child(const std::string& t) const {
    for (auto& k : kids) if (k->tag == t) return k.get();
    return nullptr;
  }
  std::string get(const std::string& a) const {
    auto it = attr.find(a);
    return it == attr.end() ? std::string() : it->second;
  }
};

// Element tree of an XML document (comments, processing instructions and CDATA skipped;
// entities are not expanded, COLLADA data does not need them).
std::unique_ptr<XNode> parse_xml(const std::string& s) {
  auto root = std::make_unique<XNode>();
  root->tag = "#document";
  std::vector<XNode*> stack{root.get()};
  size_t i = 0;
  const size_t n = s.size();
  while (i < n) {
    if (s[i] != '<') {
      const size_t j = s.find('<', i);
      const size_t e = j == std::string::npos ? n : j;
      stack.back()->text.append(s, i, e - i);
      i = e;
      continue;
    }
    if (s.compare(i, 4, "<!--") == 0) { const size_t j = s.find("-->", i); i = j == std::string::npos ? n : j + 3; continue; }
    if (s.compare(i, 9, "<![CDATA[") == 0) { const size_t j = s.find("]]>", i); i = j == std::string::npos ? n : j + 3; continue; }
    if (s.compare(i, 2, "<?") == 0 || s.compare(i, 2, "<!") == 0) { const size_t j = s.find('>', i); i = j == std::string::npos ? n : j + 1; continue; }
    const size_t j = s.find('>', i);
    if (j == std::string::npos) throw LoadError("COLLADA: unterminated tag");
    std::string body = s.substr(i + 1, j - i - 1);
    i = j + 1;
    if (!body.empty() && body[0] == '/') {
      if (stack.size() > 1) stack.pop_back();
      continue;
    }
    bool self = !body.empty() && body.back() == '/';
    if (self) body.pop_back();
    auto nd = std::make_unique<XNode>();
    size_t p = 0;
    while (p < body.size() && !std::isspace(static_cast<unsigned char>(body[p]))) p++;
    nd->tag = body.substr(0, p);
    while (p < body.size()) {
      while (p < body.size() && std::isspace(static_cast<unsigned char>(body[p]))) p++;
      const size_t eq = body.find('=', p);
      if (eq == std::string::npos) break;
      std::string key = body.substr(p, eq - p);
      while (!key.empty() && std::isspace(static_cast<unsigned char>(key.back()))) key.pop_back();
      size_t q = eq + 1;
      while (q < body.size() && std::isspace(static_cast<unsigned char>(body[q]))) q++;
      if (q >= body.size()) break;
      const char quote = body[q];
      const size_t qe = body.find(quote, q + 1);
      if (qe == std::string::npos) break;
      nd->attr[key] = body.substr(q + 1, qe - q - 1);
      p = qe + 1;
    }
    XNode* raw = nd.get();
    stack.back()->kids.push_back(std::move(nd));
    if (!self) stack.push_back(raw);
  }
  return root;
}

void collect_ids(const XNode* n, std::map<std::string, const XNode*>& ids) {
  const std::string id = n->get("id");
  if (!id.empty()) ids[id] = n;
  for (auto& k : n->kids) collect_ids(k.get(), ids);
}

std::vector<double> numbers(const std::string& t) {
  std::vector<double> out;
  const char* p = t.c_str();
  while (*p) {
    char* q = nullptr;
    const double v = std::strtod(p, &q);
    if (q == p) { p++; continue; }
    out.push_back(v);
    p = q;
  }
  return out;
}

const XNode* ref(const std::map<std::string, const XNode*>& ids, const std::string& url) {
  const std::string key = (!url.empty() && url[0] == '#') ? url.substr(1) : url;
  auto it = ids.find(key);
  return it == ids.end() ? nullptr : it->second;
}

// Positions of a <source> (float_array with stride from its accessor).
std::vector<double> source_positions(const XNode* src) {
  const XNode* fa = src ? src->child("float_array") : nullptr;
  if (!fa) throw LoadError("COLLADA: position source without float_array");
  std::vector<double> v = numbers(fa->text);
  size_t stride = 3;
  if (const XNode* tc = src->child("technique_common"))
    if (const XNode* acc = tc->child("accessor")) {
      const std::string st = acc->get("stride");
      if (!st.empty()) stride = static_cast<size_t>(std::stoul(st));
    }
  if (stride < 3) throw LoadError("COLLADA: position stride below 3");
  std::vector<double> out;
  for (size_t i = 0; i + 3 <= v.size(); i += stride) out.insert(out.end(), {v[i], v[i + 1], v[i + 2]});
  return out;
}

// One Assimp mesh per <triangles>/<polylist>/<polygons> block of a <geometry>.
std::vector<int> collada_geometry(const XNode* geom, const std::map<std::string, const XNode*>& ids, SceneData& sc) {
  std::vector<int> out;
  const XNode* mesh = geom->child("mesh");
  if (!mesh) return out;
  for (auto& blk : mesh->kids) {
    const std::string t = blk->tag;
    if (t != "triangles" && t != "polylist" && t != "polygons") continue;
    int voff = -1, max_off = 0;
    std::vector<double> pos;
    for (auto& in : blk->kids) {
      if (in->tag != "input") continue;
      const int off = std::stoi(in->get("offset").empty() ? "0" : in->get("offset"));
      max_off = std::max(max_off, off);
      if (in->get("semantic") == "VERTEX") {
        voff = off;
        const XNode* verts = ref(ids, in->get("source"));
        const XNode* pin = nullptr;
        if (verts)
          for (auto& vi : verts->kids)
            if (vi->tag == "input" && vi->get("semantic") == "POSITION") pin = vi.get();
        if (!pin) throw LoadError("COLLADA: <vertices> without POSITION");
        pos = source_positions(ref(ids, pin->get("source")));
      }
    }
    if (voff < 0) continue;
    const size_t stride = static_cast<size_t>(max_off) + 1;
    Mesh m;
    m.name = geom->get("name");
    m.v = pos;
    const size_t nv = pos.size() / 3;
    std::vector<uint32_t> idx;
    auto take = [&](const std::vector<double>& p, size_t first, size_t count) {
      idx.clear();
      for (size_t c = 0; c < count; c++) {
        const size_t at = (first + c) * stride + static_cast<size_t>(voff);
        if (at >= p.size()) throw LoadError("COLLADA: index list too short");
        const double x = p[at];
        if (!(x >= 0) || static_cast<size_t>(x) >= nv) throw LoadError("COLLADA: vertex index out of range");
        idx.push_back(static_cast<uint32_t>(x));
      }
      m.add_poly(idx);
    };
    if (t == "triangles") {
      const XNode* p = blk->child("p");
      const std::vector<double> pv = p ? numbers(p->text) : std::vector<double>();
      for (size_t f = 0; f + 1 <= pv.size() / (3 * stride); f++) take(pv, 3 * f, 3);
    } else if (t == "polylist") {
      const XNode* p = blk->child("p");
      const XNode* vc = blk->child("vcount");
      const std::vector<double> pv = p ? numbers(p->text) : std::vector<double>();
      const std::vector<double> cv = vc ? numbers(vc->text) : std::vector<double>();
      size_t first = 0;
      for (double c : cv) { take(pv, first, static_cast<size_t>(c)); first += static_cast<size_t>(c); }
    } else {
      for (auto& p : blk->kids)
        if (p->tag == "p") {
          const std::vector<double> pv = numbers(p->text);
          take(pv, 0, pv.size() / stride);
        }
    }
    out.push_back(static_cast<int>(sc.meshes.size()));
    sc.meshes.push_back(std::move(m));
  }
  return out;
}

void node_transform(const XNode* n, double* M) {
  identity(M);
  for (auto& k : n->kids) {
    double T[16];
    identity(T);
    const std::vector<double> v = numbers(k->text);
    if (k->tag == "matrix" && v.size() >= 16) {
      for (int i = 0; i < 16; i++) T[i] = v[i];
    } else if (k->tag == "translate" && v.size() >= 3) {
      T[3] = v[0]; T[7] = v[1]; T[11] = v[2];
    } else if (k->tag == "scale" && v.size() >= 3) {
      T[0] = v[0]; T[5] = v[1]; T[10] = v[2];
    } else if (k->tag == "rotate" && v.size() >= 4) {
      double ax = v[0], ay = v[1], az = v[2];
      const double len = std::sqrt(ax * ax + ay * ay + az * az);
      if (len > 0) { ax /= len; ay /= len; az /= len; }
      const double a = v[3] * M_PI / 180.0, c = std::cos(a), s = std::sin(a), C = 1 - c;
      const double R[9] = {c + ax * ax * C, ax * ay * C - az * s, ax * az * C + ay * s,
                           ay * ax * C + az * s, c + ay * ay * C, ay * az * C - ax * s,
                           az * ax * C - ay * s, az * ay * C + ax * s, c + az * az * C};
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) T[4 * i + j] = R[3 * i + j];
    } else {
      continue;
    }
    matmul(M, T, M);
  }
}

void collada_node(const XNode* xn, int parent, const std::map<std::string, const XNode*>& ids,
                  std::map<const XNode*, std::vector<int>>& geom_meshes, SceneData& sc, int depth) {
  if (depth > 256) throw LoadError("COLLADA: node hierarchy too deep");
  Node nd;
  nd.name = xn->get("name");
  nd.parent = parent;
  node_transform(xn, nd.M);
  for (auto& k : xn->kids)
    if (k->tag == "instance_geometry") {
      const XNode* g = ref(ids, k->get("url"));
      if (!g) continue;
      auto it = geom_meshes.find(g);
      if (it == geom_meshes.end()) it = geom_meshes.emplace(g, collada_geometry(g, ids, sc)).first;
      nd.meshes.insert(nd.meshes.end(), it->second.begin(), it->second.end());
    }
  const int id = static_cast<int>(sc.nodes.size());
  sc.nodes.push_back(nd);
  if (parent >= 0) sc.nodes[parent].children.push_back(id);
  for (auto& k : xn->kids) {
    if (k->tag == "node") collada_node(k.get(), id, ids, geom_meshes, sc, depth + 1);
    if (k->tag == "instance_node") {
      const XNode* r = ref(ids, k->get("url"));
      if (r) collada_node(r, id, ids, geom_meshes, sc, depth + 1);
    }
  }
}

void load_dae(const std::string& data, SceneData& sc) {
  auto doc = parse_xml(data);
  const XNode* col = doc->child("COLLADA");
  if (!col) throw LoadError("COLLADA: no <COLLADA> element");
  std::map<std::string, const XNode*> ids;
  collect_ids(col, ids);
  double unit = 1.0;
  if (const XNode* asset = col->child("asset"))
    if (const XNode* u = asset->child("unit")) {
      const std::string m = u->get("meter");
      if (!m.empty()) unit = std::stod(m);
    }
  const XNode* vs = nullptr;
  if (const XNode* scn = col->child("scene"))
    if (const XNode* ivs = scn->child("instance_visual_scene")) vs = ref(ids, ivs->get("url"));
  if (!vs)
    if (const XNode* lib = col->child("library_visual_scenes")) vs = lib->child("visual_scene");
  Node root;
  root.name = "root";
  identity(root.M);
  root.M[0] = root.M[5] = root.M[10] = unit;
  sc.nodes.push_back(root);
  std::map<const XNode*, std::vector<int>> geom_meshes;
  if (vs)
    for (auto& k : vs->kids)
      if (k->tag == "node") collada_node(k.get(), 0, ids, geom_meshes, sc, 0);
}

// ---- scene semantics -------------------------------------------------------------------------

// Polar decomposition of the 3x3 part (rotation R with det +1, symmetric S) and the scale
// vector diag(S), as Eigen's computeRotationScaling + diag in the reference's decompose().
void decompose(const double* M, double R[9], double s[3]) {
  double A[9] = {M[0], M[1], M[2], M[4], M[5], M[6], M[8], M[9], M[10]};
  // eigen-decomposition of A^T A by cyclic Jacobi
  double B[9];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      double t = 0;
      for (int k = 0; k < 3; k++) t += A[3 * k + i] * A[3 * k + j];
      B[3 * i + j] = t;
    }
  double V[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  for (int sweep = 0; sweep < 50; sweep++) {
    const double off = B[1] * B[1] + B[2] * B[2] + B[5] * B[5];
    if (off < 1e-30) break;
    for (int p = 0; p < 2; p++)
      for (int q = p + 1; q < 3; q++) {
        const double bpq = B[3 * p + q];
        if (std::fabs(bpq) < 1e-300) continue;
        const double theta = (B[3 * q + q] - B[3 * p + p]) / (2 * bpq);
        const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1));
        const double c = 1 / std::sqrt(t * t + 1), sn = t * c;
        for (int k = 0; k < 3; k++) {  // B = J^T B J
          const double bkp = B[3 * k + p], bkq = B[3 * k + q];
          B[3 * k + p] = c * bkp - sn * bkq;
          B[3 * k + q] = sn * bkp + c * bkq;
        }
        for (int k = 0; k < 3; k++) {
          const double bpk = B[3 * p + k], bqk = B[3 * q + k];
          B[3 * p + k] = c * bpk - sn * bqk;
          B[3 * q + k] = sn * bpk + c * bqk;
        }
        for (int k = 0; k < 3; k++) {
          const double vkp = V[3 * k + p], vkq = V[3 * k + q];
          V[3 * k + p] = c * vkp - sn * vkq;
          V[3 * k + q] = sn * vkp + c * vkq;
        }
      }
  }
  double sig[3];
  for (int i = 0; i < 3; i++) sig[i] = std::sqrt(std::max(B[4 * i], 0.0));
  const double detA = A[0] * (A[4] * A[8] - A[5] * A[7]) - A[1] * (A[3] * A[8] - A[5] * A[6]) + A[2] * (A[3] * A[7] - A[4] * A[6]);
  // U = A V diag(1/sig); a rotation needs det(U V^T) = +1: flip the smallest singular value
  int imin = 0;
  for (int i = 1; i < 3; i++) if (sig[i] < sig[imin]) imin = i;
  if (detA < 0) sig[imin] = -sig[imin];
  double U[9];
  for (int j = 0; j < 3; j++) {
    for (int i = 0; i < 3; i++) {
      double t = 0;
      for (int k = 0; k < 3; k++) t += A[3 * i + k] * V[3 * k + j];
      U[3 * i + j] = sig[j] != 0 ? t / sig[j] : (i == j ? 1.0 : 0.0);
    }
  }
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      double r = 0, sm = 0;
      for (int k = 0; k < 3; k++) {
        r += U[3 * i + k] * V[3 * j + k];
        sm += V[3 * i + k] * sig[k] * V[3 * j + k];
      }
      R[3 * i + j] = r;
      if (i == j) s[i] = sm;
    }
}

struct Flat {
  std::vector<float> V;
  std::vector<uint32_t> F;
  std::vector<uint32_t> offsets{0};
  std::vector<uint32_t> geom, obj;
};

// Appends mesh `m` transformed by M (decomposed as the reference does), with its ids.
// Returns false when the mesh has no triangle (the reference skips such meshes).
bool emit(const Mesh& m, const double* M, bool fan, uint32_t geom, uint32_t obj, Flat& out) {
  bool has_tri = false;
  for (size_t p = 0; p < m.n_poly(); p++) if (m.poly_start[p + 1] - m.poly_start[p] == 3) has_tri = true;
  if (!has_tri && !fan) return false;
  double R[9], s[3];
  decompose(M, R, s);
  const double t[3] = {M[3], M[7], M[11]};
  const size_t nv = m.v.size() / 3;
  const size_t base = out.V.size() / 3;
  if (base + nv > 0xFFFFFFFFull) throw LoadError("map too large");
  for (size_t i = 0; i < nv; i++) {
    const double x = s[0] * m.v[3 * i], y = s[1] * m.v[3 * i + 1], z = s[2] * m.v[3 * i + 2];
    for (int r = 0; r < 3; r++)
      out.V.push_back(static_cast<float>(R[3 * r] * x + R[3 * r + 1] * y + R[3 * r + 2] * z + t[r]));
  }
  const size_t f0 = out.F.size();
  for (size_t p = 0; p < m.n_poly(); p++) {
    const uint32_t a = m.poly_start[p], b = m.poly_start[p + 1], k = b - a;
    if (k < 3) continue;
    for (uint32_t j = 0; j < k; j++)
      if (m.poly[a + j] >= nv) throw LoadError("face index out of range in mesh '" + m.name + "'");
    const uint32_t ntri = fan ? k - 2 : 1;  // reference: the first three indices of a polygon
    for (uint32_t j = 0; j < ntri; j++)
      out.F.insert(out.F.end(), {static_cast<uint32_t>(base + m.poly[a]), static_cast<uint32_t>(base + m.poly[a + j + 1]),
                                 static_cast<uint32_t>(base + m.poly[a + j + 2])});
  }
  if (out.F.size() / 3 >= (1u << 28)) throw LoadError("map too large");
  if (out.F.size() == f0) return false;
  out.offsets.push_back(static_cast<uint32_t>(out.F.size() / 3));
  out.geom.push_back(geom);
  out.obj.push_back(obj);
  return true;
}

void preorder(const SceneData& sc, int n, std::vector<int>& out) {
  if (!sc.nodes[n].meshes.empty()) out.push_back(n);
  for (int c : sc.nodes[n].children) preorder(sc, c, out);
}

void global_transform(const SceneData& sc, int n, double* M) {
  identity(M);
  for (int it = n; it >= 0; it = sc.nodes[it].parent) matmul(sc.nodes[it].M, M, M);
}

Flat flatten(const SceneData& sc, bool fan) {
  Flat out;
  auto usable = [&](int mi) {
    const Mesh& m = sc.meshes[mi];
    for (size_t p = 0; p < m.n_poly(); p++) {
      const uint32_t k = m.poly_start[p + 1] - m.poly_start[p];
      if (k == 3 || (fan && k > 3)) return true;
    }
    return false;
  };
  std::vector<int> nodes;
  if (!sc.nodes.empty()) preorder(sc, 0, nodes);
  std::vector<char> referenced(sc.meshes.size(), 0);
  uint32_t top = 0;
  for (int n : nodes) {
    double M[16];
    global_transform(sc, n, M);
    std::vector<int> ms;
    for (int mi : sc.nodes[n].meshes) {
      if (mi < 0 || mi >= static_cast<int>(sc.meshes.size())) continue;
      referenced[mi] = 1;
      if (usable(mi)) ms.push_back(mi);
    }
    const uint32_t id = top++;
    if (ms.size() == 1) {
      emit(sc.meshes[ms[0]], M, fan, id, id, out);                  // frozen into a mesh
    } else {
      for (size_t j = 0; j < ms.size(); j++) emit(sc.meshes[ms[j]], M, fan, static_cast<uint32_t>(j), id, out);
    }
  }
  double I[16];
  identity(I);
  for (size_t mi = 0; mi < sc.meshes.size(); mi++)
    if (!referenced[mi] && usable(static_cast<int>(mi))) {
      const uint32_t id = top++;
      emit(sc.meshes[mi], I, fan, id, id, out);
    }
  return out;
}

// ---- Ouster sensor metadata (JSON) ------------------------------------------------------------
struct JVal {
  enum Kind { NUL, BOOL, NUM, STR, ARR, OBJ } kind = NUL;
  double num = 0.0;
  std::string str;
  std::vector<JVal> arr;
  std::vector<std::pair<std::string, JVal>> obj;
  const JVal* get(const std::string& k) const {
    for (auto& kv : obj) if (kv.first == k) return &kv.second;
    return nullptr;
  }
};

struct JParser {
  const std::string& s;
  size_t i = 0;
  void ws() { while (i < s.size() && std::isspace(static_cast<unsigned char>(s[i]))) i++; }
  [[noreturn]] void bad() { throw LoadError("JSON: syntax error at byte " + std::to_string(i)); }
  std::string string() {
    if (s[i] != '"') bad();
    std::string out;
    for (i++; i < s.size() && s[i] != '"'; i++) {
      if (s[i] == '\\' && i + 1 < s.size()) {
        i++;
        const char e = s[i];
        out += e == 'n' ? '\n' : e == 't' ? '\t' : e == 'u' ? (i += 4, '?') : e;
      } else {
        out += s[i];
      }
    }
    if (i >= s.size()) bad();
    i++;
    return out;
  }
  JVal value(int depth = 0) {
    if (depth > 64) bad();
    ws();
    if (i >= s.size()) bad();
    JVal v;
    const char c = s[i];
    if (c == '{') {
      v.kind = JVal::OBJ;
      i++;
      ws();
      if (s[i] == '}') { i++; return v; }
      for (;;) {
        ws();
        std::string k = string();
        ws();
        if (s[i] != ':') bad();
        i++;
        v.obj.emplace_back(k, value(depth + 1));
        ws();
        if (s[i] == ',') { i++; continue; }
        if (s[i] == '}') { i++; return v; }
        bad();
      }
    }
    if (c == '[') {
      v.kind = JVal::ARR;
      i++;
      ws();
      if (s[i] == ']') { i++; return v; }
      for (;;) {
        v.arr.push_back(value(depth + 1));
        ws();
        if (s[i] == ',') { i++; continue; }
        if (s[i] == ']') { i++; return v; }
        bad();
      }
    }
    if (c == '"') { v.kind = JVal::STR; v.str = string(); return v; }
    if (s.compare(i, 4, "true") == 0) { v.kind = JVal::BOOL; v.num = 1; i += 4; return v; }
    if (s.compare(i, 5, "false") == 0) { v.kind = JVal::BOOL; i += 5; return v; }
    if (s.compare(i, 4, "null") == 0) { i += 4; return v; }
    char* e = nullptr;
    v.num = std::strtod(s.c_str() + i, &e);
    if (e == s.c_str() + i) bad();
    v.kind = JVal::NUM;
    i = static_cast<size_t>(e - s.c_str());
    return v;
  }
};

std::vector<float> float_array(const JVal& root, const char* key) {
  const JVal* a = root.get(key);
  if (!a || a->kind != JVal::ARR) throw LoadError(std::string("Ouster metadata without '") + key + "'");
  std::vector<float> out;
  for (auto& v : a->arr) {
    if (v.kind != JVal::NUM) throw LoadError(std::string("Ouster metadata: non-numeric '") + key + "'");
    out.push_back(static_cast<float>(v.num));  // Json::Value::asFloat
  }
  return out;
}

}  // namespace

struct rmh_scene {
  Flat flat;
};

extern "C" {

int rmh_scene_load(const char* path, uint32_t flags, rmh_scene** out) {
  if (!path || !out) return rmh::fail(RMH_ERR_INVALID, "rmh_scene_load: NULL argument");
  *out = nullptr;
  try {
    const std::string p(path);
    const std::string data = read_file(p);
    const size_t dot = p.find_last_of('.');
    const std::string ext = dot == std::string::npos ? std::string() : lower(p.substr(dot + 1));
    SceneData sc;
    if (ext == "ply" || data.compare(0, 3, "ply") == 0) load_ply(data, sc);
    else if (ext == "obj") load_obj(data, sc);
    else if (ext == "stl") load_stl(data, sc);
    else if (ext == "dae" || data.find("<COLLADA") != std::string::npos) load_dae(data, sc);
    else return rmh::fail(RMH_ERR_UNSUPPORTED, "rmh_scene_load: unknown format of '" + p + "'");
    auto* s = new (std::nothrow) rmh_scene();
    if (!s) return rmh::fail(RMH_ERR_ALLOC, "rmh_scene_load: out of host memory");
    s->flat = flatten(sc, (flags & RMH_SCENE_POLYGONS_FAN) != 0);
    if (s->flat.F.empty()) {
      delete s;
      return rmh::fail(RMH_ERR_INVALID, "rmh_scene_load: '" + p + "' contains no triangles");
    }
    *out = s;
    return RMH_OK;
  } catch (const std::exception& e) {
    return rmh::fail(RMH_ERR_INVALID, std::string("rmh_scene_load: ") + e.what());
  }
}

int rmh_scene_counts(const rmh_scene* s, uint32_t* n_vertices, uint32_t* n_faces, uint32_t* n_meshes) {
  if (!s) return rmh::fail(RMH_ERR_INVALID, "rmh_scene_counts: NULL scene");
  if (n_vertices) *n_vertices = static_cast<uint32_t>(s->flat.V.size() / 3);
  if (n_faces) *n_faces = static_cast<uint32_t>(s->flat.F.size() / 3);
  if (n_meshes) *n_meshes = static_cast<uint32_t>(s->flat.geom.size());
  return RMH_OK;
}

int rmh_scene_copy(const rmh_scene* s, float* vertices, uint32_t* faces, uint32_t* mesh_face_offsets,
                   uint32_t* geom_ids, uint32_t* object_ids) {
  if (!s) return rmh::fail(RMH_ERR_INVALID, "rmh_scene_copy: NULL scene");
  const Flat& f = s->flat;
  if (vertices) std::memcpy(vertices, f.V.data(), sizeof(float) * f.V.size());
  if (faces) std::memcpy(faces, f.F.data(), sizeof(uint32_t) * f.F.size());
  if (mesh_face_offsets) std::memcpy(mesh_face_offsets, f.offsets.data(), sizeof(uint32_t) * f.offsets.size());
  if (geom_ids) std::memcpy(geom_ids, f.geom.data(), sizeof(uint32_t) * f.geom.size());
  if (object_ids) std::memcpy(object_ids, f.obj.data(), sizeof(uint32_t) * f.obj.size());
  return RMH_OK;
}

int rmh_scene_destroy(rmh_scene* s) {
  delete s;
  return RMH_OK;
}

// o1dn_from_ouster_meta_file (src/rmagine_ouster/src/types/ouster_sensors.cpp:8-80): width
// from "lidar_mode" ("<columns>x<Hz>"), one row per beam; the direction of (row v, column h)
// is (cos phi cos theta, cos phi sin theta, sin phi), phi = altitude[v] * DEG_TO_RAD_F,
// theta = 0 + (2 pi / W) * h + azimuth[v] * DEG_TO_RAD_F, in FP32 as there, with the
// correctly rounded cos/sin of the FP32 angles.
int rmh_ouster_o1dn(const char* path, uint32_t* width, uint32_t* height, float* dirs, size_t capacity) {
  if (!path || !width || !height) return rmh::fail(RMH_ERR_INVALID, "rmh_ouster_o1dn: NULL argument");
  try {
    const std::string data = read_file(path);
    JParser jp{data};
    const JVal root = jp.value();
    if (root.kind != JVal::OBJ) throw LoadError("Ouster metadata is not a JSON object");
    const JVal* mode = root.get("lidar_mode");
    if (!mode || mode->kind != JVal::STR) throw LoadError("Ouster metadata without 'lidar_mode'");
    const size_t x = mode->str.find('x');
    const long w = std::strtol(mode->str.substr(0, x).c_str(), nullptr, 10);
    if (x == std::string::npos || w <= 0 || w > (1 << 20)) throw LoadError("bad lidar_mode '" + mode->str + "'");
    const std::vector<float> alt = float_array(root, "beam_altitude_angles");
    const std::vector<float> az = float_array(root, "beam_azimuth_angles");
    if (alt.empty() || az.size() < alt.size()) throw LoadError("Ouster metadata: beam angle arrays do not match");
    const uint32_t W = static_cast<uint32_t>(w), H = static_cast<uint32_t>(alt.size());
    *width = W;
    *height = H;
    if (!dirs) return RMH_OK;
    if (capacity < 3ull * W * H) return rmh::fail(RMH_ERR_INVALID, "rmh_ouster_o1dn: dirs buffer too small");
    constexpr float kDegToRad = 0.017453292519943295f;
    const float hor_angle_inc = static_cast<float>((2.0 * M_PI) / static_cast<float>(W));
    for (uint32_t v = 0; v < H; v++) {
      const float phi = alt[v] * kDegToRad;
      const float cp = static_cast<float>(std::cos(static_cast<double>(phi)));
      const float sp = static_cast<float>(std::sin(static_cast<double>(phi)));
      for (uint32_t h = 0; h < W; h++) {
        const float theta_offset = az[v] * kDegToRad;
        const float theta = (0.0f + hor_angle_inc * static_cast<float>(h)) + theta_offset;
        float* d = dirs + 3 * (static_cast<size_t>(v) * W + h);
        d[0] = cp * static_cast<float>(std::cos(static_cast<double>(theta)));
        d[1] = cp * static_cast<float>(std::sin(static_cast<double>(theta)));
        d[2] = sp;
      }
    }
    return RMH_OK;
  } catch (const std::exception& e) {
    return rmh::fail(RMH_ERR_INVALID, std::string("rmh_ouster_o1dn: ") + e.what());
  }
}

}  // extern "C"
