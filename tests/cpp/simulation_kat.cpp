// C++ drop-in test of the *SimulatorHip API (runs on a HIP device).
//
// Each case follows one of the reference's simulation tests with the Hip simulator in
// place of the Embree/OptiX one: unit cube map, 100 identity poses, repeated simulate()
// calls, known-answer ray checked with the reference's tolerance 1e-4.
//   spherical   tests/embree/simulation_spherical.cpp:26-72 (range 0.500076)
//   o1dn        tests/embree/simulation_o1dn.cpp:30-78 (point (0.256732, 0.5, -0.1))
//   pinhole     tests/embree/simulation_pinhole.cpp:28-79 (point (0.5, 0.5, 0.375))
//   ondn        tests/embree/simulation_ondn.cpp:30-76 (range 0.5)
//   single ray  tests/embree/simulation_single_ray.cpp:24-70 (range 0.5)
// plus the error convention (no map -> std::runtime_error), and the batched, graph and
// pose-gradient forms (BatchSimulatorHip, HipGraph, poseGradHip; no reference counterpart)
// against the plain simulators.  Throws on failure.
#include <rmagine/map/HipMap.hpp>
#include <rmagine/noise/NoiseHip.hpp>
#include <rmagine/types/ouster_sensors.h>
#include <rmagine/simulation/BatchSimulatorHip.hpp>
#include <rmagine/simulation/O1DnSimulatorHip.hpp>
#include <rmagine/simulation/OnDnSimulatorHip.hpp>
#include <rmagine/simulation/PinholeSimulatorHip.hpp>
#include <rmagine/simulation/SphereSimulatorHip.hpp>

#include <cmath>
#include <cstring>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

using namespace rmagine;

static void require(bool ok, const std::string& what) {
  if (!ok) throw std::runtime_error("FAILED: " + what);
}

static HipMapPtr make_cube_map() {
  // genCube (src/rmagine_core/src/util/synthetic.cpp:183-221)
  const std::vector<Vector> v = {{-0.5f, -0.5f, -0.5f}, {-0.5f, -0.5f, 0.5f}, {-0.5f, 0.5f, 0.5f}, {-0.5f, 0.5f, -0.5f},
                                 {0.5f, 0.5f, -0.5f},   {0.5f, 0.5f, 0.5f},   {0.5f, -0.5f, 0.5f}, {0.5f, -0.5f, -0.5f}};
  const std::vector<Face> f = {{0, 1, 2}, {2, 3, 0}, {4, 5, 6}, {6, 7, 4}, {0, 7, 1}, {1, 7, 6},
                               {1, 6, 5}, {5, 2, 1}, {2, 5, 4}, {3, 2, 4}, {3, 4, 7}, {3, 7, 0}};
  Memory<Vector, RAM> V(v.size());
  Memory<Face, RAM> F(f.size());
  for (size_t i = 0; i < v.size(); i++) V[i] = v[i];
  for (size_t i = 0; i < f.size(); i++) F[i] = f[i];
  return std::make_shared<HipMap>(V, F);
}

static Memory<Transform, VRAM_HIP> identity_poses(size_t n) {
  Memory<Transform, RAM> T(n);
  for (size_t i = 0; i < n; i++) T[i] = Transform::Identity();
  Memory<Transform, VRAM_HIP> Td;
  Td = T;
  return Td;
}

static void test_spherical(const HipMapPtr& map) {
  SphereSimulatorHip sim;
  sim.setMap(map);
  const SphericalModel model = example_spherical();
  sim.setModel(model);
  IntAttrAll<VRAM_HIP> result;
  resize_memory_bundle<VRAM_HIP>(result, model.getWidth(), model.getHeight(), 100);
  const auto T = identity_poses(100);
  for (int i = 0; i < 20; i++) {
    sim.simulate(T, result);
    Memory<float, RAM> last_scan = result.ranges(model.size() * 99, model.size() * 100);
    const float range = last_scan[model.getBufferId(model.phi.size / 2, 0)];
    std::ostringstream ss;
    ss << "spherical range " << range << " vs 0.500076";
    require(std::fabs(range - 0.500076f) <= 1e-4f, ss.str());
  }
  // returned bundle + single host pose overloads
  auto res = sim.simulate<Bundle<Ranges<VRAM_HIP>>>(Transform::Identity());
  Memory<float, RAM> r = res.ranges;
  require(r.size() == model.size() && std::fabs(r[model.getBufferId(8, 0)] - 0.500076f) <= 1e-4f, "spherical single pose");
}

static void test_o1dn(const HipMapPtr& map) {
  O1DnSimulatorHip sim;
  sim.setMap(map);
  const O1DnModel model = example_o1dn();
  sim.setModel(model);
  IntAttrAll<VRAM_HIP> result;
  resize_memory_bundle<VRAM_HIP>(result, model.getWidth(), model.getHeight(), 100);
  const auto T = identity_poses(100);
  for (int i = 0; i < 100; i++) {
    sim.simulate(T, result);
    Memory<float, RAM> last_scan = result.ranges(model.size() * 99, model.size() * 100);
    const float range = last_scan[model.getBufferId(0, 0)];
    const Vector p = model.orig + model.dirs[0] * range;
    const float err = std::fabs(p.x - 0.256732f) + std::fabs(p.y - 0.5f) + std::fabs(p.z + 0.1f);
    require(err <= 1e-4f, "o1dn point");
  }
  Memory<Vector, RAM> normals = result.normals(0, 1);
  require(std::fabs(normals[0].y + 1.0f) < 1e-6f, "o1dn normal (0,-1,0)");
}

static void test_pinhole(const HipMapPtr& map) {
  PinholeSimulatorHip sim(map);
  const PinholeModel model = example_pinhole();
  sim.setModel(model);
  IntAttrAll<VRAM_HIP> result;
  resize_memory_bundle<VRAM_HIP>(result, model.getWidth(), model.getHeight(), 100);
  const auto T = identity_poses(100);
  for (int i = 0; i < 10; i++) {
    sim.simulate(T, result);
    Memory<float, RAM> last_scan = result.ranges(model.size() * 99, model.size() * 100);
    const float range = last_scan[model.getBufferId(0, 0)];
    const Vector p = model.getDirection(0, 0) * range;
    const float err = std::fabs(p.x - 0.5f) + std::fabs(p.y - 0.5f) + std::fabs(p.z - 0.375f);
    require(err <= 1e-4f, "pinhole point");
  }
}

static void test_ondn(const HipMapPtr& map) {
  OnDnSimulatorHip sim(map);
  const OnDnModel model = example_ondn();
  sim.setModel(model);
  IntAttrAll<VRAM_HIP> result;
  resize_memory_bundle<VRAM_HIP>(result, model.getWidth(), model.getHeight(), 100);
  const auto T = identity_poses(100);
  for (int i = 0; i < 100; i++) {
    sim.simulate(T, result);
    Memory<float, RAM> last_scan = result.ranges(model.size() * 99, model.size() * 100);
    require(std::fabs(last_scan[0] - 0.5f) <= 1e-4f, "ondn range");
  }
  // single ray (tests/embree/simulation_single_ray.cpp)
  OnDnModel one;
  one.width = 1;
  one.height = 1;
  one.range = {0.0f, 100.0f};
  one.dirs.resize(1);
  one.origs.resize(1);
  one.dirs[0] = {1.0f, 0.0f, 0.0f};
  one.origs[0] = {0.0f, 0.0f, 0.0f};
  sim.setModel(one);
  auto res = sim.simulate<IntAttrAll<VRAM_HIP>>(identity_poses(100));
  Memory<float, RAM> r = res.ranges;
  require(std::fabs(r[0] - 0.5f) <= 1e-4f && std::fabs(r[99] - 0.5f) <= 1e-4f, "single ray range");
  Memory<unsigned int, RAM> gid = res.geom_ids;
  require(gid[0] == 0u, "single mesh geom id 0");
}

static void test_no_map() {
  SphereSimulatorHip sim;
  sim.setModel(example_spherical());
  bool thrown = false;
  try {
    IntAttrAll<VRAM_HIP> res;
    sim.simulate(identity_poses(1), res);
  } catch (const std::runtime_error&) {
    thrown = true;
  }
  require(thrown, "simulate without a map throws");
}

// import_hip_map on the reference's own map files (copies under tests/golden/maps).
static void test_import(const std::string& dir) {
  // sphere.ply: icosphere of radius 1 (642 vertices, 1280 faces); every ray from the centre
  // hits, no farther than 1 and no nearer than the faces' distance to the centre (> 0.98)
  const HipMapPtr sph = import_hip_map(dir + "/sphere.ply");
  require(sph->info().n_faces == 1280 && sph->info().n_meshes == 1, "sphere.ply counts");
  SphereSimulatorHip sim(sph);
  sim.setModel(example_spherical());
  auto res = sim.simulate<IntAttrAll<VRAM_HIP>>(Transform::Identity());
  Memory<uint8_t, RAM> hits = res.hits;
  Memory<float, RAM> ranges = res.ranges;
  Memory<unsigned int, RAM> geom = res.geom_ids;
  for (size_t i = 0; i < ranges.size(); i++)
    require(hits[i] == 1 && ranges[i] > 0.98f && ranges[i] <= 1.0f + 1e-6f && geom[i] == 0, "sphere.ply ray");
  // the reference's OS0-128 metadata as an O1Dn model, cast from the sphere's centre
  const O1DnModel os0 = o1dn_from_ouster_meta_file(dir + "/../ouster/os0-128x1024x20-meta.json");
  require(os0.width == 1024 && os0.height == 128, "ouster model size");
  O1DnSimulatorHip osim(sph);
  osim.setModel(os0);
  auto ores = osim.simulate<Bundle<Hits<VRAM_HIP>, Ranges<VRAM_HIP>>>(Transform::Identity());
  Memory<uint8_t, RAM> oh = ores.hits;
  Memory<float, RAM> orr = ores.ranges;
  for (size_t i = 0; i < orr.size(); i++) require(oh[i] == 1 && orr[i] > 0.98f && orr[i] <= 1.0f + 1e-6f, "ouster ray");
  // two_cubes.dae: two single-mesh nodes, frozen into geometries 0 and 1 (node order)
  const HipMapPtr cubes = import_hip_map(dir + "/two_cubes.dae");
  require(cubes->info().n_meshes == 2 && cubes->info().n_faces == 24, "two_cubes.dae counts");
}

// GaussianNoiseHip on simulated ranges (GaussianNoiseCuda usage, e.g. apps' noise demos):
// stddev 0 makes the result exact: every range <= max_range moves by the mean.
static void test_noise(const HipMapPtr& map) {
  SphereSimulatorHip sim(map);
  sim.setModel(example_spherical());
  auto res = sim.simulate<Bundle<Ranges<VRAM_HIP>>>(Transform::Identity());
  Memory<float, RAM> before = res.ranges;
  GaussianNoiseHip::Options opt;
  opt.max_range = 50.0f;
  GaussianNoiseHip noise(0.25f, 0.0f, opt);
  noise.apply(res.ranges);
  Memory<float, RAM> after = res.ranges;
  for (size_t i = 0; i < after.size(); i++) {
    const float want = before[i] <= 50.0f ? before[i] + 0.25f : before[i];
    require(std::fabs(after[i] - want) <= 1e-6f, "noise shift");
  }
}

using HitsRanges = Bundle<Hits<VRAM_HIP>, Ranges<VRAM_HIP>>;

static bool same_results(const HitsRanges& a, const HitsRanges& b) {
  Memory<uint8_t, RAM> ha = a.hits, hb = b.hits;
  Memory<float, RAM> ra = a.ranges, rb = b.ranges;
  if (ha.size() != hb.size() || ra.size() != rb.size()) return false;
  for (size_t i = 0; i < ha.size(); i++)
    if (ha[i] != hb[i]) return false;
  return std::memcmp(ra.raw(), rb.raw(), sizeof(float) * ra.size()) == 0;
}

static void scramble(HitsRanges& b) {
  Memory<uint8_t, RAM> h(b.hits.size());
  Memory<float, RAM> r(b.ranges.size());
  for (size_t i = 0; i < h.size(); i++) { h[i] = 7; r[i] = -1.0f; }
  b.hits = h;
  b.ranges = r;
}

// Batched bundles equal each simulator's own simulate(), in one persistent launch and in
// two; a captured graph replays the same bytes; the pose loss of a scan against itself is 0.
static void test_batch(const HipMapPtr& map) {
  SphereSimulatorHip s1(map);
  s1.setModel(example_spherical());
  PinholeSimulatorHip s2(map);
  s2.setModel(example_pinhole());
  O1DnSimulatorHip s3(map);
  s3.setModel(example_o1dn());
  Memory<Transform, RAM> Th(5);
  for (size_t i = 0; i < Th.size(); i++) {
    Th[i] = Transform::Identity();
    Th[i].t = Vector{0.05f * i - 0.1f, 0.02f * i, -0.03f * i};
  }
  Memory<Transform, VRAM_HIP> T;
  T = Th;
  const SimulatorHip* sims[3] = {&s1, &s2, &s3};
  HitsRanges ref[3], got[3];
  for (int k = 0; k < 3; k++) {
    ref[k] = (k == 0 ? s1.simulate<HitsRanges>(T) : k == 1 ? s2.simulate<HitsRanges>(T) : s3.simulate<HitsRanges>(T));
    resize_memory_bundle<VRAM_HIP>(got[k], sims[k]->getWidth(), sims[k]->getHeight(), static_cast<unsigned>(T.size()));
  }
  s1.synchronize();
  BatchSimulatorHip batch(map);
  batch.add(s3, T, got[2]);  // a ray list before the image sensors
  batch.add(s1, T, got[0]);
  batch.add(s2, T, got[1]);
  batch.commit();
  require(batch.params().launches == 1, "batch: one persistent launch by default");
  require(batch.totalRays() == T.size() * (s1.getWidth() * s1.getHeight() + s2.getWidth() * s2.getHeight() +
                                           s3.getWidth() * s3.getHeight()),
          "batch: total rays");
  s3.setRayOrder(false);  // list order: same results
  for (uint32_t launches : {1u, 2u}) {
    BatchParamsHip p;
    p.launches = launches;
    batch.setParams(p);
    for (auto& g : got) scramble(g);
    batch.simulate();
    s1.synchronize();
    for (int k = 0; k < 3; k++) require(same_results(got[k], ref[k]), "batch entry equals its simulate()");
  }
  void* stream = nullptr;
  RM_RMH_CHECK(rmh_stream_create(&stream));
  {
    HipGraph graph = HipGraph::capture(stream, [&] { batch.simulate(stream); });
    for (int rep = 0; rep < 2; rep++) {
      for (auto& g : got) scramble(g);
      graph.launch(stream);
      RM_RMH_CHECK(rmh_stream_synchronize(stream));
      for (int k = 0; k < 3; k++) require(same_results(got[k], ref[k]), "graph replay equals simulate()");
    }
  }
  RM_RMH_CHECK(rmh_stream_destroy(stream));
  // pose loss of pose 0's own scan: only the FP32 rounding of the target ranges remains
  // (|e| < 1e-7 m per ray), every hit ray is used
  const size_t n = static_cast<size_t>(s1.getWidth()) * s1.getHeight();
  Memory<double, VRAM_HIP> out(8);
  poseGradHip(s1, T(0, 1), ref[0].ranges(0, n), ref[0].hits(0, n), false, 0.05f, out);
  s1.synchronize();
  Memory<double, RAM> o = out;
  Memory<uint8_t, RAM> h = ref[0].hits(0, n);
  size_t hits = 0;
  for (size_t i = 0; i < n; i++) hits += h[i] != 0;
  require(o[0] >= 0.0 && o[0] < 1e-10 && o[7] == static_cast<double>(hits) && hits > 0,
          "pose loss of a scan against itself");
}

int main(int argc, char** argv) {
  const HipMapPtr map = make_cube_map();
  test_spherical(map);
  test_o1dn(map);
  test_pinhole(map);
  test_ondn(map);
  test_no_map();
  test_noise(map);
  test_batch(map);
  // the same known answers through the 4-wide quantised node layout
  require(map->setLayout(true), "4-wide layout of the cube map");
  test_spherical(map);
  test_o1dn(map);
  test_pinhole(map);
  test_ondn(map);
  require(!map->setLayout(false), "back to the binary layout");
  if (argc > 1) test_import(argv[1]);
  std::cout << "simulation_kat: all C++ drop-in checks passed" << std::endl;
  return 0;
}
