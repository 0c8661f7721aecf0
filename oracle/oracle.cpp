// ORACLE -- test infrastructure, not product code.
//
// CPU restatement of Rmagine's Embree simulation path, used only by tests/, by
// __graft_entry__.smoke() and by bench.py's cpu_baseline leg, always as the checker or the
// reported CPU baseline, never as the thing measured or shipped.  The product path is
// librmh.so (rmagine_amd/csrc), which never calls into this file.
//
// What it restates (uos/rmagine v2.4.2):
//   simulate loops            src/rmagine_embree/include/rmagine/simulation/SphereSimulatorEmbree.tcc:50-254
//                             PinholeSimulatorEmbree.tcc (same body), O1DnSimulatorEmbree.tcc:50-254
//                             (ray origin Tsm*orig, point = d*t + orig), OnDnSimulatorEmbree.tcc
//   ray directions            src/rmagine_core/include/rmagine/types/sensor_models.h:170-175 (sphere),
//                             :256-277 (pinhole), :376-385 (O1Dn), :460-470 (OnDn)
//   pose algebra              src/rmagine_core/include/rmagine/math/types/Transform.tcc:24-62,
//                             Quaternion.tcc:19-75, Vector3.tcc:150-218
//   Tsb default identity      src/rmagine_embree/src/simulation/SimulatorEmbree.cpp:7-11
//   invalid values            sensor_models.h:59-63 (range.max + 1), NaN points/normals, UINT_MAX ids
//   geom/object ids           EmbreeScene::add -> rtcAttachGeometry order (src/rmagine_embree/src/map/embree/EmbreeScene.cpp:54-71);
//                             object id = instID if valid else geomID (SphereSimulatorEmbree.tcc:176-187)
//
// rtcIntersect1 itself (Embree 4, a third-party library absent here) is restated as an
// exact closest-hit query: the watertight ray/triangle predicate of Woop, Benthin and
// Wald (JCGT 2013) evaluated in IEEE double without contraction, over a binned-SAH BVH
// whose double-precision slab test is padded to be conservative (or over all triangles
// in brute-force mode).  Ties at equal distance go to the lower global face index.  The
// per-ray ambiguity flags mark rays whose outcome an FP32 intersector (Embree's) may
// legitimately decide differently: the hit lies within 1e-6 (barycentric) of a triangle
// edge, a triangle is missed by less than 1e-6, a second triangle is within a relative
// 1e-6 in distance, or the distance is within a relative 1e-6 of range.min / range.max.
//
// Spherical directions use the FP32 angle of the reference and the correctly rounded
// FP32 cosine/sine (FP64 evaluation rounded once), which is what cosf/sinf return up to
// their last-bit error.
//
// Parity pins for this file: the reference's known-answer rays on the unit cube
// (tests/embree/simulation_{spherical,o1dn,pinhole,ondn,single_ray}.cpp), the
// brute-force mode as a check of the BVH mode, and the independent NumPy
// Moller-Trumbore closest-hit and closed-form plane distances in tests/.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <thread>
#include <vector>

namespace {

// ---------------------------------------------------------------------------------------
// FP32 pose / ray algebra in the reference's operation order (compiled with
// -ffp-contract=off).
struct Vec { float x, y, z; };
struct Quat { float x, y, z, w; };
struct Tform { Quat R; Vec t; };

Quat qmult(const Quat& a, const Quat& b) {  // Quaternion.tcc:35-42
  return {a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y,
          a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x,
          a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w,
          a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z};
}
Quat qinv(const Quat& q) { return {-q.x, -q.y, -q.z, q.w}; }  // Quaternion.tcc:19-24
Vec qrotate(const Quat& q, const Vec& p) {                    // Quaternion.tcc:70-75
  const Quat P{p.x, p.y, p.z, 0.0f};
  const Quat r = qmult(qmult(q, P), qinv(q));
  return {r.x, r.y, r.z};
}
Vec vadd(const Vec& a, const Vec& b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
Vec vmul(const Vec& a, float s) { return {a.x * s, a.y * s, a.z * s}; }
float vdot(const Vec& a, const Vec& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
Vec vnormalize(const Vec& a) {  // Vector3.tcc:180-211
  const float n = std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z);
  return {a.x / n, a.y / n, a.z / n};
}
Tform tmult(const Tform& A, const Tform& B) {  // Transform.tcc:34-42
  Tform r;
  r.t = qrotate(A.R, B.t);
  r.R = qmult(A.R, B.R);
  r.t = vadd(r.t, A.t);
  return r;
}
Vec tapply(const Tform& T, const Vec& v) { return vadd(qrotate(T.R, v), T.t); }  // Transform.tcc:57-62

Tform load_tform(const float* p) {
  Tform T;
  T.R = {p[0], p[1], p[2], p[3]};
  T.t = {p[4], p[5], p[6]};
  return T;
}

// ---------------------------------------------------------------------------------------
// Mesh + BVH
struct Box { double lo[3], hi[3]; };

struct BNode {
  Box box;
  int left = -1, right = -1;   // interior children
  int first = 0, count = 0;    // leaf range in `order` (count > 0 => leaf)
};

struct Map {
  std::vector<float> V;
  std::vector<uint32_t> F;
  std::vector<uint32_t> geom, local, obj, mesh;  // per face: geom id, id within its mesh, object id, mesh index
  std::vector<double> nrm;     // unit normal (FP64), 3 per face
  std::vector<int> order;
  std::vector<BNode> nodes;
  uint32_t nf = 0;
};

void tri_box(const Map& m, uint32_t f, Box& b) {
  for (int a = 0; a < 3; a++) {
    double lo = 1e300, hi = -1e300;
    for (int k = 0; k < 3; k++) {
      const double x = m.V[3 * m.F[3 * f + k] + a];
      lo = std::min(lo, x);
      hi = std::max(hi, x);
    }
    b.lo[a] = lo;
    b.hi[a] = hi;
  }
}

double box_area(const Box& b) {
  const double dx = b.hi[0] - b.lo[0], dy = b.hi[1] - b.lo[1], dz = b.hi[2] - b.lo[2];
  return dx * dy + dy * dz + dz * dx;
}

void box_grow(Box& a, const Box& b) {
  for (int k = 0; k < 3; k++) {
    a.lo[k] = std::min(a.lo[k], b.lo[k]);
    a.hi[k] = std::max(a.hi[k], b.hi[k]);
  }
}

Box empty_box() {
  Box b;
  for (int k = 0; k < 3; k++) { b.lo[k] = 1e300; b.hi[k] = -1e300; }
  return b;
}

// Binned SAH, 16 bins per axis, leaves of at most 4 triangles.
int build_rec(Map& m, std::vector<Box>& fb, std::vector<double>& cen, int first, int count, int depth) {
  const int id = static_cast<int>(m.nodes.size());
  m.nodes.emplace_back();
  Box b = empty_box(), cb = empty_box();
  for (int i = first; i < first + count; i++) {
    const int f = m.order[i];
    box_grow(b, fb[f]);
    for (int k = 0; k < 3; k++) {
      cb.lo[k] = std::min(cb.lo[k], cen[3 * f + k]);
      cb.hi[k] = std::max(cb.hi[k], cen[3 * f + k]);
    }
  }
  m.nodes[id].box = b;
  if (count <= 4 || depth > 60) {
    m.nodes[id].first = first;
    m.nodes[id].count = count;
    return id;
  }
  constexpr int NB = 16;
  int best_axis = -1, best_split = -1;
  double best_cost = 1e300;
  for (int a = 0; a < 3; a++) {
    const double ext = cb.hi[a] - cb.lo[a];
    if (!(ext > 0)) continue;
    Box bb[NB];
    int bc[NB] = {0};
    for (int i = 0; i < NB; i++) bb[i] = empty_box();
    for (int i = first; i < first + count; i++) {
      const int f = m.order[i];
      int bin = static_cast<int>((cen[3 * f + a] - cb.lo[a]) / ext * NB);
      bin = std::min(std::max(bin, 0), NB - 1);
      bc[bin]++;
      box_grow(bb[bin], fb[f]);
    }
    double la[NB], ra[NB];
    int lc[NB], rc[NB];
    Box acc = empty_box();
    int c = 0;
    for (int i = 0; i < NB; i++) {
      box_grow(acc, bb[i]);
      c += bc[i];
      la[i] = c ? box_area(acc) : 0.0;
      lc[i] = c;
    }
    acc = empty_box();
    c = 0;
    for (int i = NB - 1; i >= 0; i--) {
      box_grow(acc, bb[i]);
      c += bc[i];
      ra[i] = c ? box_area(acc) : 0.0;
      rc[i] = c;
    }
    for (int s = 0; s < NB - 1; s++) {
      if (lc[s] == 0 || rc[s + 1] == 0) continue;
      const double cost = la[s] * lc[s] + ra[s + 1] * rc[s + 1];
      if (cost < best_cost) { best_cost = cost; best_axis = a; best_split = s; }
    }
  }
  int mid;
  if (best_axis < 0) {
    mid = first + count / 2;  // all centroids equal: split by index
  } else {
    const double ext = cb.hi[best_axis] - cb.lo[best_axis];
    auto it = std::partition(m.order.begin() + first, m.order.begin() + first + count, [&](int f) {
      int bin = static_cast<int>((cen[3 * f + best_axis] - cb.lo[best_axis]) / ext * NB);
      bin = std::min(std::max(bin, 0), NB - 1);
      return bin <= best_split;
    });
    mid = static_cast<int>(it - m.order.begin());
    if (mid == first || mid == first + count) mid = first + count / 2;
  }
  const int l = build_rec(m, fb, cen, first, mid - first, depth + 1);
  const int r = build_rec(m, fb, cen, mid, first + count - mid, depth + 1);
  m.nodes[id].left = l;
  m.nodes[id].right = r;
  return id;
}

// ---------------------------------------------------------------------------------------
// Exact watertight predicate in FP64 (Woop, Benthin, Wald 2013), no contraction.
struct DRay {
  double o[3];
  double d[3];
  double inv[3];
  int kx, ky, kz;
  double Sx, Sy, Sz;
};

void make_dray(const Vec& o, const Vec& d, DRay& r) {
  const float ax = std::fabs(d.x), ay = std::fabs(d.y), az = std::fabs(d.z);
  int kz = (ax > ay) ? (ax > az ? 0 : 2) : (ay > az ? 1 : 2);
  int kx = kz == 2 ? 0 : kz + 1;
  int ky = kx == 2 ? 0 : kx + 1;
  const float dkz = kz == 0 ? d.x : (kz == 1 ? d.y : d.z);
  if (dkz < 0.0f) std::swap(kx, ky);
  r.o[0] = o.x; r.o[1] = o.y; r.o[2] = o.z;
  r.d[0] = d.x; r.d[1] = d.y; r.d[2] = d.z;
  for (int k = 0; k < 3; k++) r.inv[k] = 1.0 / r.d[k];
  r.kx = kx; r.ky = ky; r.kz = kz;
  r.Sx = r.d[kx] / r.d[kz];
  r.Sy = r.d[ky] / r.d[kz];
  r.Sz = 1.0 / r.d[kz];
}

struct TriRes { double t; double bmin; bool inside; };

// returns false when the ray is parallel to the plane (det == 0) or NaN
bool woop(const Map& m, uint32_t f, const DRay& r, TriRes& res) {
  double P[3][3];
  for (int k = 0; k < 3; k++) {
    const uint32_t vi = m.F[3 * f + k];
    P[k][0] = (double)m.V[3 * vi + r.kx] - r.o[r.kx];
    P[k][1] = (double)m.V[3 * vi + r.ky] - r.o[r.ky];
    P[k][2] = (double)m.V[3 * vi + r.kz] - r.o[r.kz];
  }
  const double Ax = P[0][0] - r.Sx * P[0][2], Ay = P[0][1] - r.Sy * P[0][2];
  const double Bx = P[1][0] - r.Sx * P[1][2], By = P[1][1] - r.Sy * P[1][2];
  const double Cx = P[2][0] - r.Sx * P[2][2], Cy = P[2][1] - r.Sy * P[2][2];
  const double U = Cx * By - Cy * Bx;
  const double Vv = Ax * Cy - Ay * Cx;
  const double W = Bx * Ay - By * Ax;
  const double det = U + Vv + W;
  res.inside = !((U < 0.0 || Vv < 0.0 || W < 0.0) && (U > 0.0 || Vv > 0.0 || W > 0.0));
  if (det == 0.0 || !std::isfinite(det)) return false;
  const double T = U * (r.Sz * P[0][2]) + Vv * (r.Sz * P[1][2]) + W * (r.Sz * P[2][2]);
  res.t = T / det;
  res.bmin = std::min(U / det, std::min(Vv / det, W / det));
  return std::isfinite(res.t);
}

bool box_hit(const Box& b, const DRay& r, double tmax) {
  double t0 = 0.0, t1 = tmax;
  for (int k = 0; k < 3; k++) {
    const double pad = 1e-9 * (std::fabs(b.lo[k]) + std::fabs(b.hi[k])) + 1e-12;
    double a = (b.lo[k] - pad - r.o[k]) * r.inv[k];
    double c = (b.hi[k] + pad - r.o[k]) * r.inv[k];
    if (std::isnan(a) || std::isnan(c)) {  // d == 0 and origin on the slab plane: inside
      continue;
    }
    if (a > c) std::swap(a, c);
    t0 = std::max(t0, a);
    t1 = std::min(t1, c);
    if (t0 > t1 * (1.0 + 1e-12)) return false;
  }
  return true;
}

constexpr double kBand = 1e-6;

struct Query {
  double t;          // closest distance (range.max when none)
  int32_t face;      // global face index, -1 for none
  double bmin;       // min barycentric of the winner
  double near_t;     // smallest t of a near-edge / near-miss candidate other than the winner
  double second_t;   // smallest t among other hits
};

void consider(const Map& m, uint32_t f, const DRay& r, double tfar, Query& q) {
  TriRes tr;
  if (!woop(m, f, r, tr)) return;
  if (!(tr.t >= 0.0) || tr.t > tfar * (1.0 + 2 * kBand)) return;
  if (tr.inside && tr.t <= tfar) {
    const bool better = tr.t < q.t || (tr.t == q.t && (q.face < 0 || static_cast<int32_t>(f) < q.face));
    if (better) {
      if (q.face >= 0) q.second_t = std::min(q.second_t, q.t);
      q.t = tr.t;
      q.face = static_cast<int32_t>(f);
      q.bmin = tr.bmin;
    } else {
      q.second_t = std::min(q.second_t, tr.t);
    }
  } else if (tr.bmin >= -kBand) {
    q.near_t = std::min(q.near_t, tr.t);
  }
}

void closest(const Map& m, const DRay& r, double tfar, bool brute, Query& q) {
  q.t = tfar;
  q.face = -1;
  q.bmin = 0;
  q.near_t = 1e300;
  q.second_t = 1e300;
  if (m.nf == 0) return;
  if (brute) {
    for (uint32_t f = 0; f < m.nf; f++) consider(m, f, r, tfar, q);
    return;
  }
  int stack[128];
  int sp = 0;
  stack[sp++] = 0;
  while (sp) {
    const BNode& n = m.nodes[stack[--sp]];
    const double cull = (q.face >= 0 ? q.t : tfar) * (1.0 + 4 * kBand);
    if (!box_hit(n.box, r, cull)) continue;
    if (n.count > 0) {
      for (int i = n.first; i < n.first + n.count; i++) consider(m, static_cast<uint32_t>(m.order[i]), r, tfar, q);
    } else {
      stack[sp++] = n.right;
      stack[sp++] = n.left;
    }
  }
}

// ---------------------------------------------------------------------------------------
struct Model {
  int kind;
  uint32_t W, H;
  float rmin, rmax;
  // spherical
  float phi_min, phi_inc, th_min, th_inc;
  std::vector<float> cphi, sphi, cth, sth;
  // pinhole
  float f0, f1, c0, c1;
  // O1Dn / OnDn
  Vec orig;
  const float* dirs;
  const float* origs;
};

void ray_of(const Model& M, uint32_t vid, uint32_t hid, Vec& ds, Vec& os) {
  const uint32_t lid = vid * M.W + hid;
  switch (M.kind) {
    case 0:
      ds = {M.cphi[vid] * M.cth[hid], M.cphi[vid] * M.sth[hid], M.sphi[vid]};
      os = {0.0f, 0.0f, 0.0f};
      break;
    case 1: {
      const float pX = (static_cast<float>(hid) - M.c0) / M.f0;
      const float pY = (static_cast<float>(vid) - M.c1) / M.f1;
      const Vec d = vnormalize({pX, pY, 1.0f});
      ds = {d.z, -d.x, -d.y};
      os = {0.0f, 0.0f, 0.0f};
      break;
    }
    case 2:
      ds = {M.dirs[3 * lid], M.dirs[3 * lid + 1], M.dirs[3 * lid + 2]};
      os = M.orig;
      break;
    default:
      ds = {M.dirs[3 * lid], M.dirs[3 * lid + 1], M.dirs[3 * lid + 2]};
      os = {M.origs[3 * lid], M.origs[3 * lid + 1], M.origs[3 * lid + 2]};
      break;
  }
}

}  // namespace

extern "C" {

void* oracle_map_build(const float* V, uint32_t nv, const uint32_t* F, uint32_t nf, const uint32_t* mesh_off,
                       uint32_t n_meshes) {
  Map* m = new Map();
  m->V.assign(V, V + 3 * static_cast<size_t>(nv));
  m->F.assign(F, F + 3 * static_cast<size_t>(nf));
  m->nf = nf;
  m->geom.resize(nf);
  m->local.resize(nf);
  m->obj.resize(nf);
  m->mesh.resize(nf);
  std::vector<uint32_t> off;
  if (mesh_off && n_meshes > 0) off.assign(mesh_off, mesh_off + n_meshes + 1);
  else off = {0u, nf};
  for (size_t g = 0; g + 1 < off.size(); g++)
    for (uint32_t f = off[g]; f < off[g + 1]; f++) {
      m->geom[f] = static_cast<uint32_t>(g);
      m->obj[f] = static_cast<uint32_t>(g);
      m->mesh[f] = static_cast<uint32_t>(g);
      m->local[f] = f - off[g];
    }
  m->nrm.resize(3 * static_cast<size_t>(nf));
  std::vector<Box> fb(nf);
  std::vector<double> cen(3 * static_cast<size_t>(nf));
  for (uint32_t f = 0; f < nf; f++) {
    const float* a = &m->V[3 * F[3 * f]];
    const float* b = &m->V[3 * F[3 * f + 1]];
    const float* c = &m->V[3 * F[3 * f + 2]];
    const double e1[3] = {(double)b[0] - a[0], (double)b[1] - a[1], (double)b[2] - a[2]};
    const double e2[3] = {(double)c[0] - a[0], (double)c[1] - a[1], (double)c[2] - a[2]};
    const double n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
    const double len = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    for (int k = 0; k < 3; k++) m->nrm[3 * f + k] = len > 0 ? n[k] / len : 0.0;
    tri_box(*m, f, fb[f]);
    for (int k = 0; k < 3; k++) cen[3 * f + k] = 0.5 * (fb[f].lo[k] + fb[f].hi[k]);
  }
  m->order.resize(nf);
  for (uint32_t f = 0; f < nf; f++) m->order[f] = static_cast<int>(f);
  if (nf > 0) build_rec(*m, fb, cen, 0, static_cast<int>(nf), 0);
  return m;
}

void oracle_map_free(void* m) { delete static_cast<Map*>(m); }

// Per-mesh geom / object ids (NULL keeps); the object id is instID[0] for a hit through an
// instance, else the geomID (SphereSimulatorEmbree.tcc:175-185).
void oracle_map_set_mesh_ids(void* mp, const uint32_t* geom_ids, const uint32_t* object_ids) {
  Map& m = *static_cast<Map*>(mp);
  for (uint32_t f = 0; f < m.nf; f++) {
    if (geom_ids) m.geom[f] = geom_ids[m.mesh[f]];
    if (object_ids) m.obj[f] = object_ids[m.mesh[f]];
  }
}

// model_f: spherical {phi.min, phi.inc, phi.size, theta.min, theta.inc, theta.size, rmin, rmax}
//          pinhole   {width, height, rmin, rmax, f0, f1, c0, c1}
//          O1Dn/OnDn {width, height, rmin, rmax, ox, oy, oz, 0}
// (sizes passed as floats holding exact integers)
int oracle_simulate(void* mp, int kind, const float* model_f, const float* dirs, const float* origs,
                    const float* Tsb8, const float* Tbm8, uint32_t n_poses,
                    uint8_t* hits, float* ranges, float* points, float* normals,
                    uint32_t* face_ids, uint32_t* geom_ids, uint32_t* object_ids,
                    uint8_t* flags, int n_threads, int brute) {
  const Map& m = *static_cast<const Map*>(mp);
  Model M{};
  M.kind = kind;
  if (kind == 0) {
    M.phi_min = model_f[0]; M.phi_inc = model_f[1]; M.H = static_cast<uint32_t>(model_f[2]);
    M.th_min = model_f[3]; M.th_inc = model_f[4]; M.W = static_cast<uint32_t>(model_f[5]);
    M.rmin = model_f[6]; M.rmax = model_f[7];
    M.cphi.resize(M.H); M.sphi.resize(M.H); M.cth.resize(M.W); M.sth.resize(M.W);
    for (uint32_t i = 0; i < M.H; i++) {
      const float a = M.phi_min + static_cast<float>(i) * M.phi_inc;  // DiscreteInterval::getValue
      M.cphi[i] = static_cast<float>(std::cos(static_cast<double>(a)));
      M.sphi[i] = static_cast<float>(std::sin(static_cast<double>(a)));
    }
    for (uint32_t j = 0; j < M.W; j++) {
      const float a = M.th_min + static_cast<float>(j) * M.th_inc;
      M.cth[j] = static_cast<float>(std::cos(static_cast<double>(a)));
      M.sth[j] = static_cast<float>(std::sin(static_cast<double>(a)));
    }
  } else {
    M.W = static_cast<uint32_t>(model_f[0]);
    M.H = static_cast<uint32_t>(model_f[1]);
    M.rmin = model_f[2]; M.rmax = model_f[3];
    if (kind == 1) { M.f0 = model_f[4]; M.f1 = model_f[5]; M.c0 = model_f[6]; M.c1 = model_f[7]; }
    else { M.orig = {model_f[4], model_f[5], model_f[6]}; M.dirs = dirs; M.origs = origs; }
  }
  const uint64_t per = static_cast<uint64_t>(M.W) * M.H;
  const uint64_t total = per * n_poses;
  const Tform Tsb = load_tform(Tsb8);
  if (n_threads <= 0) n_threads = static_cast<int>(std::max(1u, std::thread::hardware_concurrency()));
  std::atomic<uint64_t> next{0};
  const uint64_t chunk = 1024;
  auto worker = [&]() {
    for (;;) {
      const uint64_t c0 = next.fetch_add(chunk);
      if (c0 >= total) break;
      const uint64_t c1 = std::min(total, c0 + chunk);
      for (uint64_t g = c0; g < c1; g++) {
        const uint32_t pid = static_cast<uint32_t>(g / per);
        const uint32_t lid = static_cast<uint32_t>(g - static_cast<uint64_t>(pid) * per);
        const uint32_t vid = lid / M.W, hid = lid - vid * M.W;
        const Tform Tsm = tmult(load_tform(Tbm8 + 8 * static_cast<size_t>(pid)), Tsb);
        Vec ds, os;
        ray_of(M, vid, hid, ds, os);
        const Vec dm = qrotate(Tsm.R, ds);
        const Vec om = (kind <= 1) ? Tsm.t : tapply(Tsm, os);
        DRay r;
        make_dray(om, dm, r);
        Query q;
        closest(m, r, static_cast<double>(M.rmax), brute != 0, q);
        uint8_t fl = 0;
        if (q.face >= 0) {
          const float t = static_cast<float>(q.t);
          const uint32_t f = static_cast<uint32_t>(q.face);
          if (q.bmin < kBand) fl |= 1;
          if (q.near_t <= q.t * (1.0 + kBand) + 1e-12) fl |= 2;
          if (q.second_t <= q.t * (1.0 + kBand) + 1e-12) fl |= 4;
          if (std::fabs(q.t - M.rmin) <= kBand * std::max(1.0, q.t) ||
              std::fabs(q.t - M.rmax) <= kBand * std::max(1.0, q.t)) fl |= 8;
          if (hits) hits[g] = (t >= M.rmin) ? 1 : 0;
          if (ranges) ranges[g] = t;
          if (points) {
            Vec p = vmul(ds, t);
            if (kind >= 2) p = vadd(p, os);
            points[3 * g] = p.x; points[3 * g + 1] = p.y; points[3 * g + 2] = p.z;
          }
          if (normals) {
            const Vec n0{(float)m.nrm[3 * f], (float)m.nrm[3 * f + 1], (float)m.nrm[3 * f + 2]};
            Vec n = qrotate(qinv(Tsm.R), n0);  // Tms.R = Tsm.R^-1
            if (vdot(ds, n) > 0.0f) n = vmul(n, -1.0f);
            n = vnormalize(n);
            normals[3 * g] = n.x; normals[3 * g + 1] = n.y; normals[3 * g + 2] = n.z;
          }
          if (face_ids) face_ids[g] = m.local[f];
          if (geom_ids) geom_ids[g] = m.geom[f];
          if (object_ids) object_ids[g] = m.obj[f];
        } else {
          if (q.near_t <= M.rmax * (1.0 + kBand)) fl |= 2;
          const float qn = std::numeric_limits<float>::quiet_NaN();
          if (hits) hits[g] = 0;
          if (ranges) ranges[g] = static_cast<float>(static_cast<double>(M.rmax) + 1.0);
          if (points) { points[3 * g] = qn; points[3 * g + 1] = qn; points[3 * g + 2] = qn; }
          if (normals) { normals[3 * g] = qn; normals[3 * g + 1] = qn; normals[3 * g + 2] = qn; }
          if (face_ids) face_ids[g] = 0xFFFFFFFFu;
          if (geom_ids) geom_ids[g] = 0xFFFFFFFFu;
          if (object_ids) object_ids[g] = 0xFFFFFFFFu;
        }
        if (flags) flags[g] = fl;
      }
    }
  };
  std::vector<std::thread> th;
  for (int i = 1; i < n_threads; i++) th.emplace_back(worker);
  worker();
  for (auto& t : th) t.join();
  return 0;
}

}  // extern "C"
