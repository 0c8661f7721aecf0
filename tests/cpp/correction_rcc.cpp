// The reference's GPU RCC correction test (tests/optix/correction_rcc.cpp:1-207) on the Hip
// backend, written as a Rmagine user would after switching: make_simulator<SphericalModel,
// Hip>() instead of the OptiX simulator, VRAM_HIP instead of VRAM_CUDA, statistics_p2l over
// PointCloudView_<VRAM_HIP>, umeyama_transform, and the reference's acceptance bound
// (|t.z| and |yaw| of the remaining error <= 1e-3 after 10 x 5 iterations).  The map is the
// unit cube (OptixCube, src/rmagine_core/src/util/synthetic.cpp:183-221).  Beyond the
// reference test, the same run checks the other drop-ins it touches against each other:
// partial statistics (a view of 7 outputs) merged with CrossStatistics' formula
// (CrossStatistics.tcc:9-58) against the total, the batched device umeyama_transform
// against the host one, multNxN / invert / pack / setIdentity; and, where the reference test
// prints them side by side, the device statistics against the host ones on RAM copies of the
// clouds (correction_rcc.cpp:112-119,169-177).  Throws on failure.
#include <rmagine/map/HipMap.hpp>
#include <rmagine/math/memory_math_hip.hpp>
#include <rmagine/math/statistics.h>
#include <rmagine/math/statistics_hip.hpp>
#include <rmagine/simulation/SimulatorHipTraits.hpp>
#include <rmagine/types/PointCloud.hpp>

#include <cmath>
#include <iostream>
#include <sstream>
#include <vector>

namespace rm = rmagine;

static void require(bool ok, const std::string& what) {
  if (!ok) throw std::runtime_error("FAILED: " + what);
}

static rm::HipMapPtr make_map() {
  const std::vector<rm::Vector> v = {{-0.5f, -0.5f, -0.5f}, {-0.5f, -0.5f, 0.5f}, {-0.5f, 0.5f, 0.5f},
                                     {-0.5f, 0.5f, -0.5f},  {0.5f, 0.5f, -0.5f},   {0.5f, 0.5f, 0.5f},
                                     {0.5f, -0.5f, 0.5f},   {0.5f, -0.5f, -0.5f}};
  const std::vector<rm::Face> f = {{0, 1, 2}, {2, 3, 0}, {4, 5, 6}, {6, 7, 4}, {0, 7, 1}, {1, 7, 6},
                                   {1, 6, 5}, {5, 2, 1}, {2, 5, 4}, {3, 2, 4}, {3, 4, 7}, {3, 7, 0}};
  rm::Memory<rm::Vector, rm::RAM> V(v.size());
  rm::Memory<rm::Face, rm::RAM> F(f.size());
  for (size_t i = 0; i < v.size(); i++) V[i] = v[i];
  for (size_t i = 0; i < f.size(); i++) F[i] = f[i];
  return std::make_shared<rm::HipMap>(V, F);
}

// correction_rcc.cpp:52-67
static rm::SphericalModel define_sensor_model() {
  const float deg = static_cast<float>(M_PI / 180.0);
  rm::SphericalModel model;
  model.theta.min = -M_PI;
  model.theta.inc = 1.0 * deg;
  model.theta.size = 360;
  model.phi.min = -64.0 * deg;
  model.phi.inc = 4.0 * deg;
  model.phi.size = 32;
  model.range.min = 0.5;
  model.range.max = 130.0;
  return model;
}

// CrossStatistics_::add (CrossStatistics.tcc:9-31) in FP64
static rm::CrossStatisticsD merge(const rm::CrossStatisticsD& a, const rm::CrossStatistics& b32) {
  rm::CrossStatisticsD b;
  b.dataset_mean = {b32.dataset_mean.x, b32.dataset_mean.y, b32.dataset_mean.z};
  b.model_mean = {b32.model_mean.x, b32.model_mean.y, b32.model_mean.z};
  for (int k = 0; k < 9; k++) b.covariance.data[k] = b32.covariance.data[k];
  b.n_meas = b32.n_meas;
  const unsigned int n = a.n_meas + b.n_meas;
  if (n == 0) return a;
  const double w1 = double(a.n_meas) / n, w2 = double(b.n_meas) / n;
  rm::CrossStatisticsD r;
  r.dataset_mean = a.dataset_mean * w1 + b.dataset_mean * w2;
  r.model_mean = a.model_mean * w1 + b.model_mean * w2;
  const rm::Vector3_<double> da = a.dataset_mean - r.dataset_mean, db = b.dataset_mean - r.dataset_mean;
  const rm::Vector3_<double> ma = a.model_mean - r.model_mean, mb = b.model_mean - r.model_mean;
  const double dav[3] = {da.x, da.y, da.z}, dbv[3] = {db.x, db.y, db.z};
  const double mav[3] = {ma.x, ma.y, ma.z}, mbv[3] = {mb.x, mb.y, mb.z};
  for (int c = 0; c < 3; c++)
    for (int row = 0; row < 3; row++)
      r.covariance.data[c * 3 + row] = (a.covariance.data[c * 3 + row] + mav[row] * dav[c]) * w1 +
                                       (b.covariance.data[c * 3 + row] + mbv[row] * dbv[c]) * w2;
  r.n_meas = n;
  return r;
}

int main() {
  std::cout << "Correction Hip-RCC + GPU optimization" << std::endl;

  // make_simulator<SphericalModel, Hip> (Simulator.hpp:66-78, SimulatorHipTraits.hpp)
  rm::SimulatorPtr<rm::SphericalModel, rm::Hip> sim = rm::make_simulator<rm::SphericalModel, rm::Hip>();
  static_assert(std::is_same<decltype(sim), rm::SphereSimulatorHipPtr>::value, "traits pick SphereSimulatorHip");
  static_assert(std::is_same<rm::Simulator<rm::O1DnModel, rm::Hip>, rm::O1DnSimulatorHip>::value, "O1Dn traits");
  static_assert(std::is_same<rm::Simulator<rm::PinholeModel, rm::Hip>, rm::PinholeSimulatorHip>::value, "pinhole");
  static_assert(std::is_same<rm::Simulator<rm::OnDnModel, rm::Hip>, rm::OnDnSimulatorHip>::value, "OnDn traits");

  rm::HipMapPtr map = make_map();
  sim->setMap(map);
  auto sensor_model = define_sensor_model();
  sim->setModel(sensor_model);

  rm::Transform Tbm_gt = rm::Transform::Identity();
  rm::Memory<rm::Transform, rm::RAM> Tbm(1);
  Tbm[0] = Tbm_gt;

  // define what we want to simulate and pre-malloc all buffers
  rm::IntAttrAll<rm::VRAM_HIP> dataset;
  rm::resize_memory_bundle<rm::VRAM_HIP>(dataset, sensor_model.getWidth(), sensor_model.getHeight(), 1);
  sim->setTsb(rm::Transform::Identity());
  sim->simulate(Tbm, dataset);

  rm::PointCloudView_<rm::VRAM_HIP> cloud_dataset = {.points = dataset.points, .mask = dataset.hits};

  // only for the CPU comparison (correction_rcc.cpp:112-119)
  rm::PointCloud_<rm::RAM> cloud_dataset_data_ram = rm::transfer<rm::RAM>(cloud_dataset);
  rm::PointCloudView_<rm::RAM> cloud_dataset_ram = {.points = cloud_dataset_data_ram.points,
                                                    .mask = cloud_dataset_data_ram.mask};

  // MICP params (correction_rcc.cpp:132-139)
  const size_t n_outer = 10;
  const size_t n_inner = 5;
  rm::UmeyamaReductionConstraints params;
  params.max_dist = 100.0;
  params.dataset_id = 0;
  params.model_id = 0;

  rm::Transform Tbm_est = rm::Transform::Identity();
  Tbm_est.t.z = 0.1;
  Tbm_est.R = rm::EulerAngles{0.0, 0.0, 0.1};

  rm::IntAttrAll<rm::VRAM_HIP> model;
  rm::resize_memory_bundle<rm::VRAM_HIP>(model, sensor_model.getWidth(), sensor_model.getHeight(), 1);

  for (size_t i = 0; i < n_outer; i++) {
    // find RCC at the estimated pose of the robot
    sim->simulate(rm::make_view(Tbm_est), model);
    rm::PointCloudView_<rm::VRAM_HIP> cloud_model = {.points = model.points, .mask = model.hits,
                                                     .normals = model.normals};
    rm::PointCloud_<rm::RAM> cloud_model_ram_data = rm::transfer<rm::RAM>(cloud_model);
    rm::PointCloudView_<rm::RAM> cloud_model_ram = {.points = cloud_model_ram_data.points,
                                                    .mask = cloud_model_ram_data.mask,
                                                    .normals = cloud_model_ram_data.normals};
    rm::Transform Tpre = rm::Transform::Identity();
    for (size_t j = 0; j < n_inner; j++) {
      rm::CrossStatistics stats = rm::statistics_p2l(Tpre, cloud_dataset, cloud_model, params);
      rm::CrossStatistics stats2 = rm::statistics_p2l(Tpre, cloud_dataset_ram, cloud_model_ram, params);
      {
        // device (FP64 accumulation, FP64 gate) against host (FP32 gate, FP64 accumulation)
        require(stats.n_meas == stats2.n_meas, "GPU and CPU statistics count different correspondences");
        double err = 0.0, scale = 1e-6;
        for (int k = 0; k < 9; k++) {
          err = std::max(err, std::fabs(double(stats.covariance.data[k]) - stats2.covariance.data[k]));
          scale = std::max(scale, std::fabs(double(stats2.covariance.data[k])));
        }
        const rm::Vector dd = stats.dataset_mean - stats2.dataset_mean, dmm = stats.model_mean - stats2.model_mean;
        std::ostringstream ss;
        ss << "GPU vs CPU statistics: covariance differs by " << err << " (scale " << scale << "), means by "
           << dd.l2norm() << " / " << dmm.l2norm();
        require(err <= 1e-5 * scale && dd.l2norm() <= 1e-5f && dmm.l2norm() <= 1e-5f, ss.str());
      }
      if (j == 0) {
        // partial results over the reference's block partition, merged, equal the total
        rm::Memory<rm::CrossStatistics, rm::VRAM_HIP> parts_d(7);
        rm::statistics_p2l(Tpre, cloud_dataset, cloud_model, params, parts_d);
        rm::Memory<rm::CrossStatistics, rm::RAM> parts;
        parts = parts_d;
        rm::CrossStatisticsD acc{};
        for (size_t k = 0; k < parts.size(); k++) acc = merge(acc, parts[k]);
        require(acc.n_meas == stats.n_meas && stats.n_meas > 0, "partial statistics count");
        double err = 0.0, scale = 0.0;
        for (int k = 0; k < 9; k++) {
          err = std::max(err, std::fabs(acc.covariance.data[k] - stats.covariance.data[k]));
          scale = std::max(scale, std::fabs(double(stats.covariance.data[k])));
        }
        std::ostringstream ss;
        ss << "merged partial covariance differs by " << err << " (scale " << scale << ")";
        require(err <= 1e-5 * scale + 1e-7, ss.str());
        // the device's batched Umeyama equals the host one on the FP64 statistics
        rm::Memory<rm::CrossStatisticsD, rm::VRAM_HIP> s64(1);
        rm::statistics_p2l(Tpre, cloud_dataset, cloud_model, params, s64);
        rm::Memory<rm::Transform, rm::VRAM_HIP> Tu_d(1);
        rm::umeyama_transform(Tu_d, s64);
        rm::Memory<rm::Transform, rm::RAM> Tu;
        Tu = Tu_d;
        rm::Memory<rm::CrossStatisticsD, rm::RAM> s64h;
        s64h = s64;
        const rm::Transform Th = rm::umeyama_transform(s64h[0]);
        require(Tu[0].R.x == Th.R.x && Tu[0].R.y == Th.R.y && Tu[0].R.z == Th.R.z && Tu[0].R.w == Th.R.w &&
                    Tu[0].t.x == Th.t.x && Tu[0].t.y == Th.t.y && Tu[0].t.z == Th.t.z,
                "device and host Umeyama differ");
      }
      rm::Transform Tpre_next = rm::umeyama_transform(stats);
      Tpre = Tpre * Tpre_next;
    }
    // (base, t+1) -> (base, t) -> map
    Tbm_est = Tbm_est * Tpre;
    std::cout << i + 1 << ": t = (" << Tbm_est.t.x << ", " << Tbm_est.t.y << ", " << Tbm_est.t.z << ")" << std::endl;
  }

  // transform from gt to estimation base frame (correction_rcc.cpp:196-205)
  auto Tdiff = ~Tbm_est * Tbm_gt;
  const rm::EulerAngles Ediff = Tdiff.R;
  if (std::fabs(Tdiff.t.z) > 0.001 || std::fabs(Ediff.yaw) > 0.001) {
    std::stringstream ss;
    ss << "Unexpected Hip Correction RCC results! dz " << Tdiff.t.z << " dyaw " << Ediff.yaw;
    throw std::runtime_error(ss.str());
  }

  // memory_math drop-ins on device poses
  const size_t n = 64;
  rm::Memory<rm::Transform, rm::RAM> A(n);
  for (size_t i = 0; i < n; i++) {
    A[i].R = rm::EulerAngles{0.01f * i, -0.02f * i, 0.03f * i};
    A[i].t = {0.1f * i, -0.2f, 0.3f};
    A[i].stamp = static_cast<uint32_t>(i);
  }
  rm::Memory<rm::Transform, rm::VRAM_HIP> Ad(A);
  rm::Memory<rm::Transform, rm::VRAM_HIP> Ai = rm::invert(Ad);
  rm::Memory<rm::Transform, rm::VRAM_HIP> I_d = rm::multNxN(Ad, Ai);
  rm::Memory<rm::Transform, rm::RAM> I;
  I = I_d;
  for (size_t i = 0; i < n; i++) {
    const rm::Transform ref = A[i] * ~A[i];
    require(I[i].R.x == ref.R.x && I[i].R.w == ref.R.w && I[i].t.x == ref.t.x && I[i].t.z == ref.t.z,
            "multNxN(T, invert(T)) differs from the host product");
    require(std::fabs(I[i].R.w - 1.0f) < 1e-6f && std::fabs(I[i].t.x) < 1e-5f, "T * T^-1 is not the identity");
  }
  rm::Memory<rm::Transform, rm::VRAM_HIP> one(1);
  rm::setIdentity(one);
  rm::Memory<rm::Transform, rm::RAM> M1 = rm::multNx1(Ad, one);
  for (size_t i = 0; i < n; i++) require(M1[i].t.x == A[i].t.x && M1[i].R.z == A[i].R.z, "multNx1 with identity");
  rm::Memory<rm::Quaternion, rm::RAM> q(n);
  rm::Memory<rm::Vector, rm::RAM> t(n);
  for (size_t i = 0; i < n; i++) {
    q[i] = A[i].R;
    t[i] = A[i].t;
  }
  rm::Memory<rm::Quaternion, rm::VRAM_HIP> qd(q);
  rm::Memory<rm::Vector, rm::VRAM_HIP> td(t);
  rm::Memory<rm::Transform, rm::VRAM_HIP> P_d(Ad);
  rm::pack(qd, td, P_d);
  rm::Memory<rm::Transform, rm::RAM> P;
  P = P_d;
  for (size_t i = 0; i < n; i++)
    require(P[i].R.y == A[i].R.y && P[i].t.y == A[i].t.y && P[i].stamp == A[i].stamp, "pack keeps R, t and the stamp");

  // Memory::resize keeps the contents of RAM (Memory.tcc:126-137)
  rm::Memory<float, rm::RAM> grow(3);
  grow[0] = 1.f;
  grow[1] = 2.f;
  grow[2] = 3.f;
  grow.resize(5);
  require(grow.size() == 5 && grow[0] == 1.f && grow[2] == 3.f, "RAM resize keeps the contents");

  std::cout << "correction_rcc ok" << std::endl;
  return 0;
}
