// Device code shared by librmh's traversal kernels (plain simulate, batched simulate,
// pose adjoint): FP32 ray generation in the reference's operation order, the FP64
// watertight ray/triangle test, BVH traversal with an LDS stack, and the fused attribute
// epilogue.  References for each piece are given in rmh_simulate.hip.
#pragma once

#include "rmh_fastdiv.hpp"
#include "rmh_internal.hpp"

namespace rmh {

// Model parameters of one sensor as the kernels see them.
struct RayGen {
  int kind;
  uint32_t W, H, ray_count;
  float range_min, range_max;
  const float2* lut;                 // spherical direction table: H rows then W columns
  float pf0, pf1, pc0, pc1;          // pinhole
  V3 orig;                           // O1Dn origin (sensor frame)
  const rmh_vec3* dirs;              // O1Dn / OnDn, in traversal order
  const rmh_vec3* origs;             // OnDn, in traversal order
  const uint32_t* order;             // O1Dn / OnDn: buffer index of each traversal slot
  uint32_t grid_w, grid_h;           // image that is cut into tiles: W x H, or (W*H) x 1 for ray lists
};

inline RayGen make_raygen(const rmh_sensor* s) {
  RayGen g;
  g.kind = s->kind;
  g.W = s->width;
  g.H = s->height;
  g.ray_count = s->width * s->height;
  g.range_min = s->range.min;
  g.range_max = s->range.max;
  g.lut = s->lut;
  g.pf0 = s->pinhole.f[0];
  g.pf1 = s->pinhole.f[1];
  g.pc0 = s->pinhole.c[0];
  g.pc1 = s->pinhole.c[1];
  g.orig = v3(s->orig.x, s->orig.y, s->orig.z);
  g.dirs = s->dirs;
  g.origs = s->origs;
  g.order = s->order;
  const bool list = s->kind == RMH_MODEL_O1DN || s->kind == RMH_MODEL_ONDN;
  g.grid_w = list ? g.ray_count : g.W;
  g.grid_h = list ? 1u : g.H;
  return g;
}

inline MapData make_mapdata(const rmh_map* m) {
  MapData d;
  d.tris = m->tris;
  d.normals = m->normals;
  d.nodes = m->nodes;
  d.root = m->root;
  d.n_faces = m->n_faces;
  d.nodes4 = m->nodes4;
  d.root4 = m->root4;
  return d;
}

// Traversal variant of a map, which the launchers turn into template arguments: the
// 4-wide tree when the map has one (stack of 3 entries per wide level), else the binary
// tree (one entry per level).  Stack classes 32 / 64 / 128; -1 = deeper than 128.
struct TravSel {
  int stack;
  bool wide;
};

inline TravSel trav_select(const rmh_map* m) {
  const uint32_t need = m->nodes4 ? 3u * m->info.wide_levels : m->info.max_depth;
  TravSel t;
  t.wide = m->nodes4 != nullptr;
  t.stack = need <= 32 ? 32 : (need <= 64 ? 64 : (need <= kMaxBinaryDepth ? 128 : -1));
  return t;
}

struct TriRay {
  double okx, oky, okz;
  double Sx, Sy, Sz;
  int perm;  // 2 kz + swap (one register; the axes are derived at the leaves)
};

__device__ __forceinline__ TriRay make_tri_ray(const V3& o, const V3& d) {
  TriRay r;
  const float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
  int kz = (ax > ay) ? (ax > az ? 0 : 2) : (ay > az ? 1 : 2);
  int kx = kz == 2 ? 0 : kz + 1;
  int ky = kx == 2 ? 0 : kx + 1;
  const float dkz = kz == 0 ? d.x : (kz == 1 ? d.y : d.z);
  if (dkz < 0.0f) { const int t = kx; kx = ky; ky = t; }
  const double dd[3] = {(double)d.x, (double)d.y, (double)d.z};
  const double oo[3] = {(double)o.x, (double)o.y, (double)o.z};
  r.perm = 2 * kz + (dkz < 0.0f ? 1 : 0);
  r.Sx = dd[kx] / dd[kz];
  r.Sy = dd[ky] / dd[kz];
  r.Sz = 1.0 / dd[kz];
  r.okx = oo[kx]; r.oky = oo[ky]; r.okz = oo[kz];
  return r;
}

// Axis permutation of the watertight test as lane masks: kz is the dominant axis of the
// direction, kx and ky follow it cyclically and are swapped when d[kz] < 0.  Choosing a
// component by an index (v[k]) compiles to a loop over the distinct indices of the wave
// with branches over the execution mask; two selects on masks computed once per leaf are
// two v_cndmask instructions per component.
struct TriAxes {
  bool x0, x1, y0, y1, z0, z1;  // kx == 0, kx == 1, ...
};

__device__ __forceinline__ TriAxes tri_axes(int perm) {
  const int kz = perm >> 1;
  const int kx0 = kz == 2 ? 0 : kz + 1, ky0 = kx0 == 2 ? 0 : kx0 + 1;
  const int kx = (perm & 1) ? ky0 : kx0, ky = (perm & 1) ? kx0 : ky0;
  return TriAxes{kx == 0, kx == 1, ky == 0, ky == 1, kz == 0, kz == 1};
}

__device__ __forceinline__ float pick(const float4& v, bool c0, bool c1) { return c0 ? v.x : (c1 ? v.y : v.z); }

// Watertight test in FP64.  Returns t >= 0 of the hit or a negative value for no hit.
__device__ __forceinline__ double tri_test(const TriRay& r, const TriAxes& k, const float4& v0, const float4& v1,
                                           const float4& v2) {
  const double Akx = (double)pick(v0, k.x0, k.x1) - r.okx, Aky = (double)pick(v0, k.y0, k.y1) - r.oky;
  const double Akz = (double)pick(v0, k.z0, k.z1) - r.okz;
  const double Bkx = (double)pick(v1, k.x0, k.x1) - r.okx, Bky = (double)pick(v1, k.y0, k.y1) - r.oky;
  const double Bkz = (double)pick(v1, k.z0, k.z1) - r.okz;
  const double Ckx = (double)pick(v2, k.x0, k.x1) - r.okx, Cky = (double)pick(v2, k.y0, k.y1) - r.oky;
  const double Ckz = (double)pick(v2, k.z0, k.z1) - r.okz;
  const double Ax = Akx - r.Sx * Akz, Ay = Aky - r.Sy * Akz;
  const double Bx = Bkx - r.Sx * Bkz, By = Bky - r.Sy * Bkz;
  const double Cx = Ckx - r.Sx * Ckz, Cy = Cky - r.Sy * Ckz;
  const double U = Cx * By - Cy * Bx;
  const double V = Ax * Cy - Ay * Cx;
  const double W = Bx * Ay - By * Ax;
  // non-short-circuit form: each triple of compares becomes one FP64 min or max and one
  // compare; the short-circuit form also took canonicalising moves and a branch
  const bool neg = (U < 0.0) | (V < 0.0) | (W < 0.0);
  const bool pos = (U > 0.0) | (V > 0.0) | (W > 0.0);
  if (neg & pos) return -1.0;
  const double det = U + V + W;
  if (det == 0.0) return -1.0;
  const double T = U * (r.Sz * Akz) + V * (r.Sz * Bkz) + W * (r.Sz * Ckz);
  return T / det;  // NaN or negative => no hit (callers test t >= 0)
}

struct Hit {
  double t;
  uint32_t slot;  // Morton slot of the hit triangle, 0xFFFFFFFF = none
  uint32_t face;  // global face index (tie break)
};

// The triangles of leaf reference `leaf` against the ray; keeps the nearest hit, equal
// distances going to the lower global face index.
__device__ __forceinline__ void test_leaf(const MapData& m, const TriRay& tr, int leaf, Hit& h) {
  const TriAxes ax = tri_axes(tr.perm);
  const uint32_t e = ~static_cast<uint32_t>(leaf);
  const uint32_t first = e >> 3, cnt = (e & 7u) + 1u;
  for (uint32_t k = 0; k < cnt; k++) {
    const size_t s = static_cast<size_t>(first + k) * 3;
    const float4 v0 = m.tris[s], v1 = m.tris[s + 1], v2 = m.tris[s + 2];
    const double t = tri_test(tr, ax, v0, v1, v2);
    if (t >= 0.0 && t <= h.t) {
      const uint32_t face = __float_as_uint(v2.w);
      if (t < h.t || face < h.face) {
        h.t = t;
        h.face = face;
        h.slot = first + k;
      }
    }
  }
}

// Traversal stack.  The top of every lane's stack lives in LDS, LDS rows of 8 bytes per
// lane (one column per lane); deeper entries go to a private per-lane array, which measured
// rays rarely reach (a lane holds one entry per level only where both children were hit).
// 10 rows = 5 KB of LDS per wave: 32 resident waves (8 per SIMD, the most a CU holds) fill
// the 160 KB of a CU exactly (terrain 9.44 -> 9.38 ms against 8 rows).
//   CULL (classic walks, and the run-time mode of batched kernels): an entry is a subtree
//     reference and the entry distance of its box
//     (the slab test's lower bound, taken at the push), one row.  A pop skips entries whose
//     box starts beyond the closest hit found since the push: every triangle below lies
//     farther than that hit, so the subtree can neither win nor tie, and its node is never
//     loaded (terrain: 5.7 % of node visits; 10,000 particles in a room: 7.4 %).
//   otherwise (speculative walks): an entry is a reference, two per row.  Their random
//     rays keep deeper stacks and could skip 2 % of visits: the deeper LDS stack pays more.
constexpr int kLdsStack = 10;

// Element of the per-lane spill array behind the LDS stack, addressed as private memory
// (scratch instructions).  Through a generic pointer the accesses were flat instructions,
// which count against both the vector-memory and the LDS counters, and a flat access that
// may still be in flight makes the compiler wait for every outstanding vector load before
// the first node word is used.
typedef __attribute__((address_space(5))) volatile int PrivInt;

// pushes a subtree (and, when culling, its entry distance)
template <int STACK, int BLOCK, int LDS, bool CULL>
__device__ __forceinline__ void push_ref(int2 (*stk)[BLOCK], PrivInt* spill, int& sp, int ref, float t) {
  if constexpr (CULL) {
    if (sp < LDS) {
      stk[sp][threadIdx.x] = make_int2(ref, __float_as_int(t));
    } else if (sp < STACK) {
      spill[2 * (sp - LDS)] = ref;
      spill[2 * (sp - LDS) + 1] = __float_as_int(t);
    }
  } else {
    int (*s1)[BLOCK] = reinterpret_cast<int (*)[BLOCK]>(stk);
    if (sp < 2 * LDS) s1[sp][threadIdx.x] = ref;
    else if (sp < STACK) spill[sp - 2 * LDS] = ref;
  }
  sp++;
}

// Node words read through the constant address space, which the compiler serves with
// scalar loads when the address is wave-uniform.
typedef __attribute__((address_space(4))) const float CFloat;
typedef __attribute__((address_space(1))) const char GChar;
typedef __attribute__((address_space(1))) const float GFloat;
typedef __attribute__((address_space(1))) const int GInt;

// Slab test.  A slab value is fma(plane, 1/d, -(o +- P) / d) with a per-ray margin
//   P = 2^-21 (max|o| + range_max) + 1e-30.
// Written out, a computed value is t' = (plane - o -+ P'') / d * (1 + r), where
//   * P'' lies in [0.74 P, 1.26 P]: o +- P and its product with 1/d are rounded, at most
//     2^-23 (|o| + P) <= P / 4 in position (the lo plane takes o + P, the hi plane o - P,
//     so both planes move outward);
//   * |r| <= 1.2e-7: the roundings of the reciprocal and of the fma.
// Replacing a direction component below 2^-64 by +-2^-64 moves the ray by at most
// range_max 2^-64, far below P.  Let the exact ray meet the box at some t <= range_max in
// [0, tbf] (tbf: an FP32 bound at least the closest hit found so far).  On each axis the near plane's exact crossing is <= t, so its moved
// plane crosses at <= t - P''/|d| and the computed near value is <= t - P''/|d| + 1.2e-7 t
// <= t, because P''/|d| >= 0.74 P >= 3.5e-7 range_max (|d| <= 1, t <= range_max); in the
// same way every computed far value is >= t.  So tmin <= tmax holds as computed, with no
// relative slack, and the box is never culled; boxes grow by about 2^-21 range_max
// (0.05 mm at 100 m).  SlabRay holds the per-ray constants of that test.
struct SlabRay {
  float idx, idy, idz;   // reciprocal direction (components clamped away from zero)
  float lx, ly, lz;      // -(o + P) / d: offsets of the lo planes
  float hx, hy, hz;      // -(o - P) / d: offsets of the hi planes
};

// max(v, 0) and min(v, hi) of slab values as one v_med3_f32 each.  fminf/fmaxf would first
// canonicalise an operand the compiler cannot prove canonical (tbf, which lives across the
// node loop), one more VALU instruction per node visit.  The slab values are never NaN
// (finite planes, finite clamped reciprocals), where med3 and min/max agree.
__device__ __forceinline__ float clamp_lo0(float v) { return __builtin_amdgcn_fmed3f(v, 0.0f, 3.40282347e38f); }
__device__ __forceinline__ float clamp_hi(float v, float hi) { return __builtin_amdgcn_fmed3f(v, hi, -3.40282347e38f); }

// Entry/exit distances of both children of one node.  OCT >= 0: the sign of every
// direction component is known for the whole wave (bit 0/1/2 set = x/y/z negative), so
// the near and far plane of each axis are known and min/max per axis are not needed.
// The values are the same as the generic form: with lo <= hi and (o + P) >= (o - P), the
// lo-plane value is below the hi-plane value whenever 1/d > 0 (monotone rounding).
template <int OCT>
__device__ __forceinline__ void slab2(const SlabRay& r, float tbf, const float4& na, const float4& nb,
                                      const float4& nc, float& tmin0, float& tmax0, float& tmin1, float& tmax1) {
  const float x00 = __fmaf_rn(na.x, r.idx, r.lx), x01 = __fmaf_rn(na.y, r.idx, r.hx);
  const float y00 = __fmaf_rn(na.z, r.idy, r.ly), y01 = __fmaf_rn(na.w, r.idy, r.hy);
  const float z00 = __fmaf_rn(nc.x, r.idz, r.lz), z01 = __fmaf_rn(nc.y, r.idz, r.hz);
  const float x10 = __fmaf_rn(nb.x, r.idx, r.lx), x11 = __fmaf_rn(nb.y, r.idx, r.hx);
  const float y10 = __fmaf_rn(nb.z, r.idy, r.ly), y11 = __fmaf_rn(nb.w, r.idy, r.hy);
  const float z10 = __fmaf_rn(nc.z, r.idz, r.lz), z11 = __fmaf_rn(nc.w, r.idz, r.hz);
  if constexpr (OCT < 0) {
    tmin0 = fmaxf(fmaxf(fminf(x00, x01), fminf(y00, y01)), clamp_lo0(fminf(z00, z01)));
    tmax0 = fminf(fminf(fmaxf(x00, x01), fmaxf(y00, y01)), clamp_hi(fmaxf(z00, z01), tbf));
    tmin1 = fmaxf(fmaxf(fminf(x10, x11), fminf(y10, y11)), clamp_lo0(fminf(z10, z11)));
    tmax1 = fminf(fminf(fmaxf(x10, x11), fmaxf(y10, y11)), clamp_hi(fmaxf(z10, z11), tbf));
  } else {
    constexpr bool nx = OCT & 1, ny = OCT & 2, nz = OCT & 4;
    tmin0 = fmaxf(fmaxf(nx ? x01 : x00, ny ? y01 : y00), clamp_lo0(nz ? z01 : z00));
    tmax0 = fminf(fminf(nx ? x00 : x01, ny ? y00 : y01), clamp_hi(nz ? z00 : z01, tbf));
    tmin1 = fmaxf(fmaxf(nx ? x11 : x10, ny ? y11 : y10), clamp_lo0(nz ? z11 : z10));
    tmax1 = fminf(fminf(nx ? x10 : x11, ny ? y10 : y11), clamp_hi(nz ? z10 : z11, tbf));
  }
}

// An FP32 bound at least as large as a non-negative FP64 distance, for the culling tests:
// the nearest float plus one ulp (two instructions, where the correctly rounded-up
// conversion takes about fourteen).  It exceeds the rounded-up value by at most one ulp,
// so the tests keep at most the boxes that start within that ulp beyond the hit.
__device__ __forceinline__ float f32_above(double t) { return __int_as_float(__float_as_int(static_cast<float>(t)) + 1); }

constexpr int kDone = 0x7fffffff;

// Walks interior nodes from `ref` until a leaf is reached (ref < 0) or the stack runs
// dry (ref = kDone).
// Speculation modes of trace(): never (classic while-while), always, or per a wave-uniform
// run-time flag (one kernel for several sensor kinds: it takes the speculative loop form,
// with the walk's leaf holding switched by the flag; a classic form in the same kernel as
// well costs registers and made a batched kernel 50 % slower).  Stack entries carry entry
// distances (pop skipping) in every mode but kSpecOn.
constexpr int kSpecOff = 0, kSpecOn = 1, kSpecRuntime = 2;

// The culling bound after a leaf test.  The run-time mode (one batched kernel for all sensor
// kinds) keeps the rounded-up conversion: with the cheaper bound its register allocation
// came out slower (batched workload 19.26 -> 19.39 ms), while the terrain ran 0.7 % faster.
template <int SPEC>
__device__ __forceinline__ float leaf_bound(double t) {
  if constexpr (SPEC == kSpecRuntime) return __double2float_ru(t);
  else return f32_above(t);
}

// Per-lane traversal counts of the diagnostic build of trace() (COUNT = true): node
// visits through the vector and the scalar path, leaves, triangle tests, and (on the
// first active lane only) wave iterations of the node loop and how many were wave-uniform.
struct VisitCount {
  uint32_t vec_nodes, uni_nodes, leaves, tris, wave_iters, wave_uni_iters;
};

__device__ __forceinline__ void count_visit(VisitCount* vc, bool uniform) {
  if (uniform) vc->uni_nodes++;
  else vc->vec_nodes++;
  const uint64_t act = __ballot(1);
  if (threadIdx.x % 64 == static_cast<uint32_t>(__ffsll(static_cast<unsigned long long>(act)) - 1)) {
    vc->wave_iters++;
    if (uniform) vc->wave_uni_iters++;
  }
}

// Pops the next subtree of a lane, kDone when its stack runs empty.  When culling, entries
// whose box the slab test would no longer accept against tbf (an FP32 bound at least the
// current closest hit: entry <= tbf) are skipped: skipping is the slab test of the entry's parent
// repeated with the shorter interval (the stored entry is the computed near value, at
// most the distance of any point of the box the ray meets within the old interval).
template <int STACK, int BLOCK, int LDS, bool CULL>
__device__ __forceinline__ int pop_ref(int2 (*stk)[BLOCK], PrivInt* spill, int& sp, float tbf) {
  if constexpr (!CULL) {
    if (sp == 0) return kDone;
    --sp;
    if (sp < 2 * LDS) return reinterpret_cast<int (*)[BLOCK]>(stk)[sp][threadIdx.x];
    return spill[sp - 2 * LDS];
  }
  while (sp > 0) {
    --sp;
    int2 e;
    if (sp < LDS) {
      e = stk[sp][threadIdx.x];
    } else {
      e.x = spill[2 * (sp - LDS)];
      e.y = spill[2 * (sp - LDS) + 1];
    }
    if (__int_as_float(e.y) <= tbf) return e.x;
  }
  return kDone;
}

// SPEC == kSpecOn (or kSpecRuntime with spec): speculative walk (Aila and Laine, "Understanding the Efficiency of
// Ray Traversal on GPUs", HPG 2009): a lane that reaches its first leaf holds it in `leaf`
// and keeps walking; the wave leaves the loop once an eighth of its walking lanes hold a
// leaf (on 2,000 x 32k random directions with the current kernel: 20.7 ms for an eighth,
// 20.8 for a sixth, 21.0 for a twelfth, 21.3 for a quarter, 21.4 for a sixteenth, 22.0 for
// a third; on earlier kernels 25.9 ms when waiting for all lanes, 24.1 for half, and
// 34.5 ms without speculation).  This fills
// the node iterations of incoherent waves; coherent image tiles reach their leaves
// together and lose by it (5 %: a held leaf shrinks the ray's interval later), so they
// walk without it.
template <int OCT, int STACK, int BLOCK, int SPEC, bool COUNT, int LDS>
__device__ __forceinline__ void walk(const MapData& m, const SlabRay& r, float tbf, int2 (*stk)[BLOCK],
                                     PrivInt* spill, int& ref, int& sp, int& leaf, bool spec, VisitCount* vc) {
  while (ref >= 0 && ref != kDone) {
    float tmin0, tmax0, tmin1, tmax1;
    int2 nr;
    const int r0 = __builtin_amdgcn_readfirstlane(ref);
    if constexpr (COUNT) count_visit(vc, __ballot(ref != r0) == 0);
    if (__ballot(ref != r0) == 0) {
      // every active lane is at the same node (45% of visits on the terrain): read it
      // through the scalar cache, which keeps the vector memory return path free; the
      // slab code sits in both branches so that it reads the node words from SGPRs
      const CFloat* cf = (const CFloat*)(m.nodes) + static_cast<size_t>(r0) * 16;
      __builtin_amdgcn_s_setprio(0);
      slab2<OCT>(r, tbf, make_float4(cf[0], cf[1], cf[2], cf[3]), make_float4(cf[4], cf[5], cf[6], cf[7]),
                 make_float4(cf[8], cf[9], cf[10], cf[11]), tmin0, tmax0, tmin1, tmax1);
      nr = make_int2(__float_as_int(cf[12]), __float_as_int(cf[13]));
    } else {
      // 32-bit byte offset from the map's base: one shift, and the loads take the
      // scalar-base + vector-offset form
      const uint32_t off = static_cast<uint32_t>(ref) << 6;
      const GFloat* gf = (const GFloat*)((const GChar*)(m.nodes) + off);
      const float4 na = make_float4(gf[0], gf[1], gf[2], gf[3]);
      const float4 nb = make_float4(gf[4], gf[5], gf[6], gf[7]);
      const float4 nc = make_float4(gf[8], gf[9], gf[10], gf[11]);
      // only the two child references of the last 16 bytes are used; hiding the
      // offset's alignment keeps the compiler from widening this read to 16 bytes, and
      // adding the 48 after the barrier lets it go into the load's immediate offset
      uint32_t roff = off;
      asm volatile("" : "+v"(roff));
      __builtin_assume((roff & 7u) == 0u);
      const GInt* gr = (const GInt*)((const GChar*)(m.nodes) + roff + 48u);
      nr = make_int2(gr[0], gr[1]);
      __builtin_amdgcn_s_setprio(0);
      slab2<OCT>(r, tbf, na, nb, nc, tmin0, tmax0, tmin1, tmax1);
    }
    // Wave priority: the slab tests run at priority 0, everything else (the choice of the
    // next node, stack work, the loads that follow, and the leaf tests) at 1, so that a wave
    // that has its box results issues its next dependent loads ahead of waves still testing
    // boxes (terrain 8.58 -> 8.53 ms, room 0.124 -> 0.119 ms; the leaf tests at priority 0 or
    // 2, or the whole walk at 0, were 1-5 % slower).
    __builtin_amdgcn_s_setprio(1);
    const bool h0 = tmin0 <= tmax0;
    const bool h1 = tmin1 <= tmax1;
    if (h0 && h1) {
      int near = nr.x, far = nr.y;
      float tfar = tmin1;
      if (tmin1 < tmin0) { near = nr.y; far = nr.x; tfar = tmin0; }
      push_ref<STACK, BLOCK, LDS, SPEC != kSpecOn>(stk, spill, sp, far, tfar);
      ref = near;
    } else if (h0) {
      ref = nr.x;
    } else if (h1) {
      ref = nr.y;
    } else {
      ref = pop_ref<STACK, BLOCK, LDS, SPEC != kSpecOn>(stk, spill, sp, tbf);
    }
    if (SPEC == kSpecOn || (SPEC == kSpecRuntime && spec)) {
      if (ref < 0 && leaf == 0) {
        leaf = ref;
        ref = pop_ref<STACK, BLOCK, LDS, SPEC != kSpecOn>(stk, spill, sp, tbf);
      }
      // an eighth for the ray-list kernels (the ray-list launch of a batch split in two
      // included: 19.10 -> 19.27 ms there, against ICP 21.38 -> 20.74 ms); the batched
      // run-time mode, whose chunks mix sensor kinds, keeps a quarter (with an eighth: the
      // batched workload's one launch 19.27 -> 19.48 ms)
      constexpr int kHoldShare = SPEC == kSpecOn ? 8 : 4;
      if (__popcll(__ballot(leaf != 0)) * kHoldShare >= __popcll(__ballot(1))) break;
    }
  }
}

// Entry distances and hit flags of the (up to) four children of a 4-wide node, from its
// words w[0..9] (Node4).  Each grid plane is computed exactly (fma(q, 2^e, origin) is an
// FP32 number), then goes through the binary slab test's formula, so the test is
// conservative in the same way for boxes that contain the binary tree's boxes.
template <int OCT>
__device__ __forceinline__ void slab4(const SlabRay& r, float tbf, uint32_t w0, uint32_t w1, uint32_t w2, uint32_t w3,
                                      uint32_t qlx, uint32_t qhx, uint32_t qly, uint32_t qhy, uint32_t qlz,
                                      uint32_t qhz, float (&tmin)[4], bool (&hit)[4]) {
  const float ox = __uint_as_float(w0), oy = __uint_as_float(w1), oz = __uint_as_float(w2);
  const float sx = __uint_as_float((w3 & 0xffu) << 23);
  const float sy = __uint_as_float(((w3 >> 8) & 0xffu) << 23);
  const float sz = __uint_as_float(((w3 >> 16) & 0xffu) << 23);
  const uint32_t nch = w3 >> 24;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const float lx = __fmaf_rn(static_cast<float>((qlx >> (8 * k)) & 0xffu), sx, ox);
    const float hx = __fmaf_rn(static_cast<float>((qhx >> (8 * k)) & 0xffu), sx, ox);
    const float ly = __fmaf_rn(static_cast<float>((qly >> (8 * k)) & 0xffu), sy, oy);
    const float hy = __fmaf_rn(static_cast<float>((qhy >> (8 * k)) & 0xffu), sy, oy);
    const float lz = __fmaf_rn(static_cast<float>((qlz >> (8 * k)) & 0xffu), sz, oz);
    const float hz = __fmaf_rn(static_cast<float>((qhz >> (8 * k)) & 0xffu), sz, oz);
    const float x0 = __fmaf_rn(lx, r.idx, r.lx), x1 = __fmaf_rn(hx, r.idx, r.hx);
    const float y0 = __fmaf_rn(ly, r.idy, r.ly), y1 = __fmaf_rn(hy, r.idy, r.hy);
    const float z0 = __fmaf_rn(lz, r.idz, r.lz), z1 = __fmaf_rn(hz, r.idz, r.hz);
    float t0, t1;
    if constexpr (OCT < 0) {
      t0 = fmaxf(fmaxf(fminf(x0, x1), fminf(y0, y1)), clamp_lo0(fminf(z0, z1)));
      t1 = fminf(fminf(fmaxf(x0, x1), fmaxf(y0, y1)), clamp_hi(fmaxf(z0, z1), tbf));
    } else {
      constexpr bool nx = OCT & 1, ny = OCT & 2, nz = OCT & 4;
      t0 = fmaxf(fmaxf(nx ? x1 : x0, ny ? y1 : y0), clamp_lo0(nz ? z1 : z0));
      t1 = fminf(fminf(nx ? x0 : x1, ny ? y0 : y1), clamp_hi(nz ? z0 : z1, tbf));
    }
    tmin[k] = t0;
    hit[k] = t0 <= t1 && (k < 2 || static_cast<uint32_t>(k) < nch);
  }
}

__device__ __forceinline__ void cswap(uint32_t& a, uint32_t& b) {
  const uint32_t lo = min(a, b), hi = max(a, b);
  a = lo;
  b = hi;
}

__device__ __forceinline__ int pick_ref(uint32_t key, int r0, int r1, int r2, int r3) {
  const int a = (key & 1u) ? r1 : r0, b = (key & 1u) ? r3 : r2;
  return (key & 2u) ? b : a;
}

// The 4-wide form of walk(): the hit children of a node are ordered by entry distance
// (keys = distance bits with the child slot in the low 2 bits, sorted by a 5-exchange
// network), the nearest is taken next and the others are pushed farthest first.  The host
// picks STACK >= 3 x the number of wide levels (at most 3 entries are pushed per level).
template <int OCT, int STACK, int BLOCK, int LDS, bool CULL>
__device__ __forceinline__ void walk4(const MapData& m, const SlabRay& r, float tbf, int2 (*stk)[BLOCK],
                                      PrivInt* spill, int& ref, int& sp) {
  while (ref >= 0) {
    float tmin[4];
    bool hit[4];
    int c0, c1, c2, c3;
    const int r0 = __builtin_amdgcn_readfirstlane(ref);
    if (__ballot(ref != r0) == 0) {
      // wave-uniform node: scalar loads, node words in SGPRs (as in walk())
      typedef __attribute__((address_space(4))) const uint32_t CUint;
      const CUint* cw = (const CUint*)(m.nodes4) + static_cast<size_t>(r0) * 16;
      slab4<OCT>(r, tbf, cw[0], cw[1], cw[2], cw[3], cw[4], cw[5], cw[6], cw[7], cw[8], cw[9], tmin, hit);
      c0 = static_cast<int>(cw[10]); c1 = static_cast<int>(cw[11]);
      c2 = static_cast<int>(cw[12]); c3 = static_cast<int>(cw[13]);
    } else {
      typedef __attribute__((address_space(1))) const uint32_t GUint;
      const uint32_t off = static_cast<uint32_t>(ref) << 6;
      const GUint* gw = (const GUint*)((const GChar*)(m.nodes4) + off);
      const uint32_t w0 = gw[0], w1 = gw[1], w2 = gw[2], w3 = gw[3];
      const uint32_t w4 = gw[4], w5 = gw[5], w6 = gw[6], w7 = gw[7];
      const uint32_t w8 = gw[8], w9 = gw[9];
      c0 = static_cast<int>(gw[10]); c1 = static_cast<int>(gw[11]);
      // the last two references: an 8-byte read (see walk())
      uint32_t roff = off + 48u;
      asm volatile("" : "+v"(roff));
      __builtin_assume((roff & 7u) == 0u);
      const GInt* gr = (const GInt*)((const GChar*)(m.nodes4) + roff);
      c2 = gr[0];
      c3 = gr[1];
      slab4<OCT>(r, tbf, w0, w1, w2, w3, w4, w5, w6, w7, w8, w9, tmin, hit);
    }
    uint32_t k0 = hit[0] ? ((__float_as_uint(tmin[0]) & ~3u) | 0u) : 0xFFFFFFFFu;
    uint32_t k1 = hit[1] ? ((__float_as_uint(tmin[1]) & ~3u) | 1u) : 0xFFFFFFFFu;
    uint32_t k2 = hit[2] ? ((__float_as_uint(tmin[2]) & ~3u) | 2u) : 0xFFFFFFFFu;
    uint32_t k3 = hit[3] ? ((__float_as_uint(tmin[3]) & ~3u) | 3u) : 0xFFFFFFFFu;
    cswap(k0, k1);
    cswap(k2, k3);
    cswap(k0, k2);
    cswap(k1, k3);
    cswap(k1, k2);
    if (k0 == 0xFFFFFFFFu) {
      ref = pop_ref<STACK, BLOCK, LDS, CULL>(stk, spill, sp, tbf);
      if (ref == kDone) return;
      continue;
    }
    // a key with its two low bits cleared is a distance no larger than the child's entry
    // distance (or a negative one, which never skips)
    if (k3 != 0xFFFFFFFFu) push_ref<STACK, BLOCK, LDS, CULL>(stk, spill, sp, pick_ref(k3, c0, c1, c2, c3), __uint_as_float(k3 & ~3u));
    if (k2 != 0xFFFFFFFFu) push_ref<STACK, BLOCK, LDS, CULL>(stk, spill, sp, pick_ref(k2, c0, c1, c2, c3), __uint_as_float(k2 & ~3u));
    if (k1 != 0xFFFFFFFFu) push_ref<STACK, BLOCK, LDS, CULL>(stk, spill, sp, pick_ref(k1, c0, c1, c2, c3), __uint_as_float(k1 & ~3u));
    ref = pick_ref(k0, c0, c1, c2, c3);
  }
}

// Closest hit along o + t d, 0 <= t <= range_max.  While-while traversal: a lane walks
// interior nodes (near child first, far child onto its stack) until it reaches a leaf or
// runs out of work, then the lanes holding leaves test their triangles together; with
// SPEC == kSpecOn (or kSpecRuntime and spec) the walk is speculative (see walk()).  The
// stack never overflows:
// the host picks STACK >= the depth of the BVH.
//
// Waves whose 64 rays share one direction octant (almost all tiles) run the walk
// specialised for it; mixed waves run the generic one.
template <int STACK, int BLOCK, bool WIDE, int SPEC, bool COUNT = false, int LDS = kLdsStack>
__device__ __forceinline__ void trace(const MapData& m, const V3& o, const V3& d, float range_max,
                                      int2 (*stk)[BLOCK], Hit& h, bool spec = false, VisitCount* vc = nullptr) {
  h.t = (double)range_max;
  h.slot = 0xFFFFFFFFu;
  h.face = 0xFFFFFFFFu;
  if (m.n_faces == 0) return;
  constexpr float kMinDir = 0x1p-64f;
  SlabRay r;
  r.idx = 1.0f / (fabsf(d.x) > kMinDir ? d.x : copysignf(kMinDir, d.x));
  r.idy = 1.0f / (fabsf(d.y) > kMinDir ? d.y : copysignf(kMinDir, d.y));
  r.idz = 1.0f / (fabsf(d.z) > kMinDir ? d.z : copysignf(kMinDir, d.z));
  const float omax = fmaxf(fmaxf(fabsf(o.x), fabsf(o.y)), fabsf(o.z));
  const float P = (omax + range_max) * 0x1p-21f + 1e-30f;
  r.lx = -((o.x + P) * r.idx); r.hx = -((o.x - P) * r.idx);
  r.ly = -((o.y + P) * r.idy); r.hy = -((o.y - P) * r.idy);
  r.lz = -((o.z + P) * r.idz); r.hz = -((o.z - P) * r.idz);
  // direction octant, specialised when the whole wave shares it (almost all tiles)
  const int oct = (r.idx < 0.0f ? 1 : 0) | (r.idy < 0.0f ? 2 : 0) | (r.idz < 0.0f ? 4 : 0);
  const int oct0 = __builtin_amdgcn_readfirstlane(oct);
  const int woct = __ballot(oct != oct0) == 0 ? oct0 : -1;
  const TriRay tr = make_tri_ray(o, d);
  // volatile keeps the rare deep entries in scratch and stops the compiler from merging
  // the LDS and scratch pops into one flat load
  volatile int spill_mem[STACK > LDS ? 2 * (STACK - LDS) : 2];
  PrivInt* spill = (PrivInt*)spill_mem;
  float tbf = range_max;
  int ref = WIDE ? m.root4 : m.root;
  int sp = 0;
  int leaf = 0;  // leaf held by a speculative walk (< 0), 0 = none
  for (;;) {
    if constexpr (WIDE) {
      switch (woct) {
        case 0: walk4<0, STACK, BLOCK, LDS, SPEC != kSpecOn>(m, r, tbf, stk, spill, ref, sp); break;
        case 1: walk4<1, STACK, BLOCK, LDS, SPEC != kSpecOn>(m, r, tbf, stk, spill, ref, sp); break;
        case 2: walk4<2, STACK, BLOCK, LDS, SPEC != kSpecOn>(m, r, tbf, stk, spill, ref, sp); break;
        case 3: walk4<3, STACK, BLOCK, LDS, SPEC != kSpecOn>(m, r, tbf, stk, spill, ref, sp); break;
        case 4: walk4<4, STACK, BLOCK, LDS, SPEC != kSpecOn>(m, r, tbf, stk, spill, ref, sp); break;
        case 5: walk4<5, STACK, BLOCK, LDS, SPEC != kSpecOn>(m, r, tbf, stk, spill, ref, sp); break;
        case 6: walk4<6, STACK, BLOCK, LDS, SPEC != kSpecOn>(m, r, tbf, stk, spill, ref, sp); break;
        case 7: walk4<7, STACK, BLOCK, LDS, SPEC != kSpecOn>(m, r, tbf, stk, spill, ref, sp); break;
        default: walk4<-1, STACK, BLOCK, LDS, SPEC != kSpecOn>(m, r, tbf, stk, spill, ref, sp); break;
      }
    } else {
      switch (woct) {
        case 0: walk<0, STACK, BLOCK, SPEC, COUNT, LDS>(m, r, tbf, stk, spill, ref, sp, leaf, spec, vc); break;
        case 1: walk<1, STACK, BLOCK, SPEC, COUNT, LDS>(m, r, tbf, stk, spill, ref, sp, leaf, spec, vc); break;
        case 2: walk<2, STACK, BLOCK, SPEC, COUNT, LDS>(m, r, tbf, stk, spill, ref, sp, leaf, spec, vc); break;
        case 3: walk<3, STACK, BLOCK, SPEC, COUNT, LDS>(m, r, tbf, stk, spill, ref, sp, leaf, spec, vc); break;
        case 4: walk<4, STACK, BLOCK, SPEC, COUNT, LDS>(m, r, tbf, stk, spill, ref, sp, leaf, spec, vc); break;
        case 5: walk<5, STACK, BLOCK, SPEC, COUNT, LDS>(m, r, tbf, stk, spill, ref, sp, leaf, spec, vc); break;
        case 6: walk<6, STACK, BLOCK, SPEC, COUNT, LDS>(m, r, tbf, stk, spill, ref, sp, leaf, spec, vc); break;
        case 7: walk<7, STACK, BLOCK, SPEC, COUNT, LDS>(m, r, tbf, stk, spill, ref, sp, leaf, spec, vc); break;
        default: walk<-1, STACK, BLOCK, SPEC, COUNT, LDS>(m, r, tbf, stk, spill, ref, sp, leaf, spec, vc); break;
      }
    }
    if constexpr (SPEC == kSpecOff) {
      // classic while-while order: test the leaf the walk stopped at, then pop (popping
      // first would hold one more register through the FP64 tests: 5 % slower on the
      // terrain)
      if (ref == kDone) break;
      if constexpr (COUNT) { vc->leaves++; vc->tris += (~static_cast<uint32_t>(ref) & 7u) + 1u; }
      test_leaf(m, tr, ref, h);
      tbf = leaf_bound<SPEC>(h.t);
      ref = pop_ref<STACK, BLOCK, LDS, SPEC != kSpecOn>(stk, spill, sp, tbf);
      if (ref == kDone) break;
      continue;
    }
    // speculative: hold the leaf the walk stopped at (if it holds none) and pop the next
    // subtree, test the held leaf, then a leaf the walk stopped at after it
    if (ref < 0 && leaf == 0) {
      leaf = ref;
      ref = pop_ref<STACK, BLOCK, LDS, SPEC != kSpecOn>(stk, spill, sp, tbf);
    }
    while (leaf < 0) {
      if constexpr (COUNT) { vc->leaves++; vc->tris += (~static_cast<uint32_t>(leaf) & 7u) + 1u; }
      test_leaf(m, tr, leaf, h);
      tbf = leaf_bound<SPEC>(h.t);
      if (ref < 0) {
        leaf = ref;
        ref = pop_ref<STACK, BLOCK, LDS, SPEC != kSpecOn>(stk, spill, sp, tbf);
      } else {
        leaf = 0;
      }
    }
    if (ref == kDone) break;
  }
}

// Speculative walks for ray lists (O1Dn/OnDn: direction-sorted, but any directions),
// classic walks for image sensors (coherent tiles).
RMH_HD constexpr bool speculative_walk(int kind) { return kind == RMH_MODEL_O1DN || kind == RMH_MODEL_ONDN; }
template <int KIND>
constexpr int spec_mode() { return speculative_walk(KIND) ? kSpecOn : kSpecOff; }

// Ray tiles: each wave casts a th x tw block of one sensor image (th * tw = 64), which keeps
// the 64 rays of a wave close in direction.  Tiles are 4 rows x 16 columns: rays of one
// row (same elevation for a spinning lidar) end at similar ranges, so wide tiles diverge
// less than square ones (measured on the terrain: 4 x 16 is 7% faster than 8 x 8).
// O1Dn/OnDn images of height 1 use 1 x 64.
struct Tiling {
  uint32_t tw, th, tiles_w, tiles;  // tiles per pose image
  uint32_t tw_log2;                 // tw is a power of two
  FastDiv div_tiles, div_tiles_w;   // division by tiles and by tiles_w
};

constexpr uint32_t kDefaultTileRows = 4;

// rows: wanted tile height (a power of two <= 64), capped by the image height.
inline Tiling make_tiling(uint32_t W, uint32_t H, uint32_t rows = kDefaultTileRows) {
  Tiling t;
  t.th = 1;
  while (t.th * 2 <= rows && t.th * 2 <= H && t.th < 64) t.th *= 2;
  t.tw = 64u / t.th;
  t.tiles_w = (W + t.tw - 1) / t.tw;
  t.tiles = t.tiles_w * ((H + t.th - 1) / t.th);
  t.tw_log2 = 0;
  while ((1u << t.tw_log2) < t.tw) t.tw_log2++;
  t.div_tiles = make_fastdiv(t.tiles > 0 ? t.tiles : 1u);
  t.div_tiles_w = make_fastdiv(t.tiles_w > 0 ? t.tiles_w : 1u);
  return t;
}

// Tiling of a sensor's image as the kernels cut it (ray lists: one row in traversal order).
inline Tiling sensor_tiling(const rmh_sensor* s, uint32_t rows = kDefaultTileRows) {
  const bool list = s->kind == RMH_MODEL_O1DN || s->kind == RMH_MODEL_ONDN;
  return list ? make_tiling(s->width * s->height, 1u) : make_tiling(s->width, s->height, rows);
}

// (row, column) of lane `lane` in tile `tile`; false for lanes outside the image.
// Workgroups are dispatched round-robin over the 8 XCDs (workgroup b runs on XCD b % 8).
// Renumbering so that XCD x receives one contiguous range of (pose, tile) indices keeps
// the BVH nodes of neighbouring tiles in that XCD's L2.  A bijection on [0, n).
__device__ __forceinline__ uint32_t xcd_contiguous(uint32_t b, uint32_t n) {
  const uint32_t xcd = b & 7u, idx = b >> 3, q = n >> 3, r = n & 7u;
  return xcd < r ? xcd * (q + 1) + idx : r * (q + 1) + (xcd - r) * q + idx;
}

// Most one-wave workgroups per launch of the (pose, tile) kernels: a dispatch holds the grid
// size in work-items as a 32-bit count, so blocks x 64 stays below 2^31.
constexpr uint64_t kMaxTileBlocks = 1ull << 25;

// `tile` is wave-uniform in every caller, so its division stays on the scalar unit.
__device__ __forceinline__ bool tile_ray(const Tiling& t, uint32_t W, uint32_t H, uint32_t tile, uint32_t lane,
                                         uint32_t& vid, uint32_t& hid) {
  const uint32_t ty = fast_div(tile, t.div_tiles_w), tx = tile - ty * t.tiles_w;
  vid = ty * t.th + (lane >> t.tw_log2);
  hid = tx * t.tw + (lane & (t.tw - 1u));
  return vid < H && hid < W;
}

// Sensor-frame ray of tile coordinates (vid, hid) and its buffer index lid.  Spherical
// and pinhole images are tiled directly (lid = row * W + column); ray lists are tiled in
// traversal order (vid = 0, hid = slot) and lid comes from the order table.
template <int KIND>
__device__ __forceinline__ void gen_ray(const RayGen& a, uint32_t vid, uint32_t hid, uint32_t& lid, V3& ds, V3& os) {
  if constexpr (KIND == RMH_MODEL_SPHERICAL || KIND == RMH_MODEL_PINHOLE) lid = vid * a.W + hid;
  else lid = a.order[hid];
  if constexpr (KIND == RMH_MODEL_SPHERICAL) {
    // SphericalModel::getDirection (sensor_models.h:170-175)
    const float2 p = a.lut[vid];
    const float2 t = a.lut[a.H + hid];
    ds = v3(p.x * t.x, p.x * t.y, p.y);
    os = v3(0.0f, 0.0f, 0.0f);
  } else if constexpr (KIND == RMH_MODEL_PINHOLE) {
    // PinholeModel::getDirectionOptical / getDirection (sensor_models.h:256-277)
    const float pX = (static_cast<float>(hid) - a.pc0) / a.pf0;
    const float pY = (static_cast<float>(vid) - a.pc1) / a.pf1;
    const V3 dopt = normalize(v3(pX, pY, 1.0f));
    ds = v3(dopt.z, -dopt.x, -dopt.y);
    os = v3(0.0f, 0.0f, 0.0f);
  } else if constexpr (KIND == RMH_MODEL_O1DN) {
    // O1DnModel_::getDirection / getOrigin (sensor_models.h:376-385)
    const rmh_vec3 dd = a.dirs[hid];
    ds = v3(dd.x, dd.y, dd.z);
    os = a.orig;
  } else {
    // OnDnModel_::getDirection / getOrigin (sensor_models.h:460-470)
    const rmh_vec3 dd = a.dirs[hid];
    const rmh_vec3 oo = a.origs[hid];
    ds = v3(dd.x, dd.y, dd.z);
    os = v3(oo.x, oo.y, oo.z);
  }
}

// Ray origin in the map frame: Tsm.t for sphere/pinhole (SphereSimulatorEmbree.tcc:88-90),
// Tsm * orig for O1Dn/OnDn (O1DnSimulatorEmbree.tcc:86-87).
template <int KIND>
__device__ __forceinline__ V3 map_origin(const Tf& Tsm, const V3& os) {
  if constexpr (KIND == RMH_MODEL_SPHERICAL || KIND == RMH_MODEL_PINHOLE) return Tsm.t;
  else return tapply(Tsm, os);
}

// Output stores are nontemporal: the outputs are written once and not read again by the
// kernel, and as plain stores their lines (6.8 GB per terrain launch) displaced BVH nodes
// and triangles from the 4 MB per-XCD L2 (terrain: 10.04 -> 9.47 ms with streaming stores;
// the bytes written are the same).
// Ray lists (O1Dn, OnDn) are traced in direction-sorted order, so each lane writes to a
// scattered buffer index: there plain stores let the L2 merge the partial lines that
// neighbouring waves write before they go to HBM.
template <int KIND>
constexpr bool kStreamOut = !(KIND == RMH_MODEL_O1DN || KIND == RMH_MODEL_ONDN);

template <int KIND, class T>
__device__ __forceinline__ void stream_out(T* p, T v) {
  if constexpr (kStreamOut<KIND>) __builtin_nontemporal_store(v, p);
  else *p = v;
}
template <int KIND>
__device__ __forceinline__ void stream_out3(rmh_vec3* p, float x, float y, float z) {
  float* f = &p->x;
  stream_out<KIND>(f, x);
  stream_out<KIND>(f + 1, y);
  stream_out<KIND>(f + 2, z);
}

// Attribute epilogue (SphereSimulatorEmbree.tcc:104-249, O1DnSimulatorEmbree.tcc:132).
template <int KIND>
__device__ __forceinline__ void write_outputs(const rmh_outputs& out, uint64_t glob, const RayGen& g,
                                              const MapData& map, const Tf& Tsm, const V3& ds, const V3& os,
                                              const Hit& h) {
  if (h.slot != 0xFFFFFFFFu) {
    const float t = static_cast<float>(h.t);
    if (out.hits) stream_out<KIND>(&out.hits[glob], static_cast<uint8_t>((t >= g.range_min) ? 1 : 0));
    if (out.ranges) stream_out<KIND>(&out.ranges[glob], t);
    if (out.points) {
      V3 p = scale(ds, t);
      if constexpr (KIND == RMH_MODEL_O1DN || KIND == RMH_MODEL_ONDN) p = add(p, os);
      stream_out3<KIND>(&out.points[glob], p.x, p.y, p.z);
    }
    float4 nn = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (out.normals || out.object_ids) nn = map.normals[h.slot];  // w = object id
    if (out.normals) {
      // face normal into the sensor frame, flipped toward the sensor, normalised
      V3 n = qrot(qconj(Tsm.R), v3(nn.x, nn.y, nn.z));
      if (dot(ds, n) > 0.0f) n = scale(n, -1.0f);
      n = normalize(n);
      stream_out3<KIND>(&out.normals[glob], n.x, n.y, n.z);
    }
    if (out.face_ids || out.geom_ids || out.object_ids) {
      const size_t s = static_cast<size_t>(h.slot) * 3;
      const uint32_t face = __float_as_uint(map.tris[s].w);
      const uint32_t geom = __float_as_uint(map.tris[s + 1].w);
      if (out.face_ids) stream_out<KIND>(&out.face_ids[glob], face);
      if (out.geom_ids) stream_out<KIND>(&out.geom_ids[glob], geom);
      // object id: the instance id for meshes of a multi-mesh instance, else the geom id
      if (out.object_ids) stream_out<KIND>(&out.object_ids[glob], __float_as_uint(nn.w));
    }
  } else {
    const float qnan = __builtin_nanf("");
    if (out.hits) stream_out<KIND>(&out.hits[glob], static_cast<uint8_t>(0));
    if (out.ranges) stream_out<KIND>(&out.ranges[glob], g.range_max + 1.0f);
    if (out.points) stream_out3<KIND>(&out.points[glob], qnan, qnan, qnan);
    if (out.normals) stream_out3<KIND>(&out.normals[glob], qnan, qnan, qnan);
    if (out.face_ids) stream_out<KIND>(&out.face_ids[glob], 0xFFFFFFFFu);
    if (out.geom_ids) stream_out<KIND>(&out.geom_ids[glob], 0xFFFFFFFFu);
    if (out.object_ids) stream_out<KIND>(&out.object_ids[glob], 0xFFFFFFFFu);
  }
}

}  // namespace rmh
