// rtr_registration.hpp -- header-only C++ interface over librtr, shaped like
// the PCL registration surface the reference calls, so code written against
// pcl::IterativeClosestPoint<PointXYZ, PointXYZ> (RealTimeRobot/function.h:
// 111-123) ports by changing the namespace:
//
//   rtr::IterativeClosestPoint<rtr::PointXYZ, rtr::PointXYZ> icp;   // :112
//   icp.setInputCloud(mcloud);                                      // :113
//   icp.setInputTarget(cloud);                                      // :114
//   rtr::PointCloud<rtr::PointXYZ> Final;
//   icp.align(Final);                                               // :117
//   std::cout << icp.getFinalTransformation() << std::endl;         // :119
//   rtr::transformPointCloud(*mcloud, *transformed, icp.getFinalTransformation());  // :123
//
// Errors: like PCL, align() does not throw for registration failures (read
// hasConverged()); invalid use and device errors throw rtr::Error with the
// library's message.  Matrix4 is a plain column-major float/double 4x4 with
// operator() and stream output (no Eigen dependency).
#pragma once
#include <cfloat>
#include <cmath>
#include <cstring>
#include <memory>
#include <ostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "rtr.h"

namespace rtr {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error("librtr error " + std::to_string(c) + ": " + m), code(c) {}
};

inline void check(int rc) {
  if (rc != RTR_OK) throw Error(rc, rtr_last_error());
}

// pcl::PointXYZ layout: x, y, z plus one float of padding (16 bytes).
struct PointXYZ {
  float x = 0.f, y = 0.f, z = 0.f, pad_ = 1.f;
  PointXYZ() = default;
  PointXYZ(float x_, float y_, float z_) : x(x_), y(y_), z(z_) {}
};

template <class PointT>
struct PointCloud {
  using Ptr = std::shared_ptr<PointCloud>;
  using ConstPtr = std::shared_ptr<const PointCloud>;
  std::vector<PointT> points;
  uint32_t width = 0, height = 1;
  bool is_dense = true;
  size_t size() const { return points.size(); }
  void resize(size_t n) {
    points.resize(n);
    width = (uint32_t)n;
    height = 1;
  }
  PointT& operator[](size_t i) { return points[i]; }
  const PointT& operator[](size_t i) const { return points[i]; }
};

// Column-major 4x4 (Eigen::Matrix4f / Matrix4d memory layout).
template <class S>
struct Matrix4 {
  S m[16];
  Matrix4() { setIdentity(); }
  static Matrix4 Identity() { return Matrix4(); }
  void setIdentity() {
    for (int i = 0; i < 16; ++i) m[i] = (i % 5 == 0) ? S(1) : S(0);
  }
  S& operator()(int r, int c) { return m[c * 4 + r]; }
  S operator()(int r, int c) const { return m[c * 4 + r]; }
  template <class T>
  Matrix4<T> cast() const {
    Matrix4<T> o;
    for (int i = 0; i < 16; ++i) o.m[i] = (T)m[i];
    return o;
  }
};
using Matrix4f = Matrix4<float>;
using Matrix4d = Matrix4<double>;

template <class S>
inline std::ostream& operator<<(std::ostream& os, const Matrix4<S>& M) {
  for (int r = 0; r < 4; ++r) {
    for (int c = 0; c < 4; ++c) os << (c ? " " : "") << M(r, c);
    os << "\n";
  }
  return os;
}

// pcl::transformPointCloud: each row ((m0*x + m1*y) + m2*z) + m3 in float;
// the bottom row is not read.
template <class PointT, class S>
inline void transformPointCloud(const PointCloud<PointT>& in, PointCloud<PointT>& out, const Matrix4<S>& T) {
  const Matrix4f f = T.template cast<float>();
  PointCloud<PointT> tmp = in;
  for (size_t i = 0; i < in.points.size(); ++i) {
    const PointT& p = in.points[i];
    PointT& o = tmp.points[i];
    float a = f(0, 0) * p.x;
    a = a + f(0, 1) * p.y;
    a = a + f(0, 2) * p.z;
    o.x = a + f(0, 3);
    float b = f(1, 0) * p.x;
    b = b + f(1, 1) * p.y;
    b = b + f(1, 2) * p.z;
    o.y = b + f(1, 3);
    float c = f(2, 0) * p.x;
    c = c + f(2, 1) * p.y;
    c = c + f(2, 2) * p.z;
    o.z = c + f(2, 3);
  }
  out = std::move(tmp);
}

// pcl::IterativeClosestPoint<PointSource, PointTarget> (Scalar = float): the
// registration runs on the GPU through librtr.
template <class PointSource, class PointTarget, class Scalar = float>
class IterativeClosestPoint {
 public:
  using PointCloudSource = PointCloud<PointSource>;
  using PointCloudTarget = PointCloud<PointTarget>;
  using Matrix4 = rtr::Matrix4<Scalar>;

  explicit IterativeClosestPoint(int device = 0) {
    check(rtr_icp_create(&h_, device));
    check(rtr_params_default(&p_));
  }
  ~IterativeClosestPoint() {
    if (h_) rtr_icp_destroy(h_);
  }
  IterativeClosestPoint(const IterativeClosestPoint&) = delete;
  IterativeClosestPoint& operator=(const IterativeClosestPoint&) = delete;

  void setInputSource(const std::shared_ptr<const PointCloudSource>& cloud) {
    src_ = cloud;
    check(rtr_icp_set_input_source(h_, cloud->points.empty() ? nullptr : &cloud->points[0].x, cloud->points.size(),
                                   sizeof(PointSource)));
  }
  template <class P>
  void setInputSource(const std::shared_ptr<P>& cloud) {
    setInputSource(std::shared_ptr<const PointCloudSource>(cloud));
  }
  // deprecated PCL name used by the reference (function.h:113)
  template <class P>
  void setInputCloud(const std::shared_ptr<P>& cloud) {
    setInputSource(cloud);
  }
  template <class P>
  void setInputTarget(const std::shared_ptr<P>& cloud) {
    const PointCloudTarget& c = *cloud;
    check(rtr_icp_set_input_target(h_, c.points.empty() ? nullptr : &c.points[0].x, c.points.size(),
                                   sizeof(PointTarget)));
  }
  void setMaxCorrespondenceDistance(double d) {
    p_.max_correspondence_distance = d;
    check(rtr_icp_set_params(h_, &p_));
  }
  double getMaxCorrespondenceDistance() const { return p_.max_correspondence_distance; }
  void setMaximumIterations(int n) {
    p_.max_iterations = n;
    check(rtr_icp_set_params(h_, &p_));
  }
  int getMaximumIterations() const { return p_.max_iterations; }
  void setTransformationEpsilon(double e) {
    p_.transformation_epsilon = e;
    check(rtr_icp_set_params(h_, &p_));
  }
  void setTransformationRotationEpsilon(double e) {
    p_.rotation_epsilon = e;
    check(rtr_icp_set_params(h_, &p_));
  }
  void setEuclideanFitnessEpsilon(double e) {
    p_.euclidean_fitness_epsilon = e;
    check(rtr_icp_set_params(h_, &p_));
  }
  // Anderson acceleration of the iteration, memory 1..8 (0 = PCL's plain
  // iteration; no PCL counterpart)
  void setAndersonMemory(int m) {
    p_.anderson_memory = m;
    check(rtr_icp_set_params(h_, &p_));
  }
  // point-to-plane (pcl::IterativeClosestPointWithNormals) on the same handle
  void setUsePointToPlane(bool on) {
    p_.mode = on ? RTR_MODE_POINT_TO_PLANE : RTR_MODE_POINT_TO_POINT;
    check(rtr_icp_set_params(h_, &p_));
  }
  void computeTargetNormals(int k = 20, float vx = 0.f, float vy = 0.f, float vz = 0.f) {
    const float vp[3] = {vx, vy, vz};
    check(rtr_icp_compute_target_normals(h_, k, vp));
  }

  void align(PointCloudSource& output) { align(output, Matrix4::Identity()); }
  void align(PointCloudSource& output, const Matrix4& guess) {
    if (!src_) throw Error(RTR_ERR_NOT_READY, "no input source");
    double g[16], T[16];
    for (int i = 0; i < 16; ++i) g[i] = (double)guess.m[i];
    output = *src_;
    int conv = 0;
    check(rtr_icp_align(h_, g, output.points.empty() ? nullptr : &output.points[0].x, sizeof(PointSource), T,
                        &conv));
    for (int i = 0; i < 16; ++i) final_.m[i] = (Scalar)T[i];
    converged_ = conv != 0;
  }
  Matrix4 getFinalTransformation() const { return final_; }
  bool hasConverged() const { return converged_; }
  double getFitnessScore(double max_range = DBL_MAX) {
    double f = 0.0;
    check(rtr_icp_get_fitness_score(h_, max_range, &f));
    return f;
  }
  rtr_result getResult() {
    rtr_result r;
    check(rtr_icp_get_result(h_, &r));
    return r;
  }

 private:
  rtr_icp* h_ = nullptr;
  rtr_params p_;
  std::shared_ptr<const PointCloudSource> src_;
  Matrix4 final_;
  bool converged_ = false;
};

// Learned-init registration of many (model, scene) pairs per call: the
// reference's coarse alignment (keypoints, TDF yaw search, consensus:
// RealTimeRobot.cpp:45-105) followed by its ICP (function.h:111-123), as one
// batched pipeline (rtr_learned_*): network proposal, batched ICP refine,
// registration judge, multi-start fallback for the rejected pairs.
//
//   rtr::LearnedRegistration<rtr::PointXYZ> reg(256);
//   reg.loadWeights("proposer.f32");
//   reg.setInputClouds(models, scans);           // vectors of PointCloud::ConstPtr
//   reg.align();
//   rtr::Matrix4d T = reg.getFinalTransformation(i);
//   bool fell_back = reg.getRejected()[i];
template <class PointT>
class LearnedRegistration {
 public:
  using CloudConstPtr = std::shared_ptr<const PointCloud<PointT>>;

  explicit LearnedRegistration(int max_pairs, int device = 0, const rtr_params* icp = nullptr,
                               const rtr_learned_config* cfg = nullptr) {
    rtr_params p;
    if (!icp) {
      // the batched registration workloads' ICP settings
      check(rtr_params_default(&p));
      p.max_iterations = 50;
      p.max_correspondence_distance = 0.05;
      p.transformation_epsilon = 1e-8;
      icp = &p;
    }
    check(rtr_learned_create(&h_, device, max_pairs, icp, cfg));
  }
  ~LearnedRegistration() {
    if (h_) rtr_learned_destroy(h_);
  }
  LearnedRegistration(const LearnedRegistration&) = delete;
  LearnedRegistration& operator=(const LearnedRegistration&) = delete;

  void loadWeights(const std::string& path) { check(rtr_learned_load_weights(h_, path.c_str())); }
  void setWeights(const std::vector<float>& w) { check(rtr_learned_set_weights(h_, w.data())); }

  // packed copies of the clouds are kept by this object (the library reads
  // the rejected pairs' clouds from them)
  void setInputClouds(const std::vector<CloudConstPtr>& sources, const std::vector<CloudConstPtr>& targets) {
    if (sources.size() != targets.size()) throw Error(RTR_ERR_INVALID_ARG, "sources and targets must pair up");
    const size_t n = sources.size();
    src_.assign(n, {});
    tgt_.assign(n, {});
    std::vector<const float*> sp(n), tp(n);
    std::vector<size_t> ns(n), nt(n);
    for (size_t i = 0; i < n; ++i) {
      pack(*sources[i], src_[i]);
      pack(*targets[i], tgt_[i]);
      sp[i] = src_[i].data();
      tp[i] = tgt_[i].data();
      ns[i] = sources[i]->size();
      nt[i] = targets[i]->size();
    }
    check(rtr_learned_set_clouds(h_, (int)n, sp.data(), ns.data(), tp.data(), nt.data(), nullptr, nullptr));
    res_.assign(n, rtr_result());
    rej_.assign(n, 0);
  }
  void align() { check(rtr_learned_align(h_, res_.data(), nullptr, rej_.data())); }
  // from the caller's initial poses instead of the network's
  void align(const std::vector<Matrix4d>& proposals) {
    std::vector<double> g(16 * proposals.size());
    for (size_t i = 0; i < proposals.size(); ++i) std::memcpy(&g[16 * i], proposals[i].m, 16 * sizeof(double));
    check(rtr_learned_align_from(h_, g.data(), res_.data(), nullptr, rej_.data()));
  }
  Matrix4d getFinalTransformation(size_t i) const {
    Matrix4d T;
    std::memcpy(T.m, res_.at(i).T, sizeof(T.m));
    return T;
  }
  bool hasConverged(size_t i) const { return res_.at(i).converged != 0; }
  const std::vector<rtr_result>& getResults() const { return res_; }
  const std::vector<int>& getRejected() const { return rej_; }
  rtr_learned_stats getStats() {
    rtr_learned_stats s;
    check(rtr_learned_get_stats(h_, &s));
    return s;
  }

 private:
  static void pack(const PointCloud<PointT>& c, std::vector<float>& out) {
    out.resize(3 * c.size());
    for (size_t i = 0; i < c.size(); ++i) {
      out[3 * i] = c.points[i].x;
      out[3 * i + 1] = c.points[i].y;
      out[3 * i + 2] = c.points[i].z;
    }
  }
  rtr_learned* h_ = nullptr;
  std::vector<std::vector<float>> src_, tgt_;
  std::vector<rtr_result> res_;
  std::vector<int> rej_;
};

// Voxel-hash map: the merge of transformed clouds into the scene map
// (the overlay at RealTimeRobot.cpp:104-114, as a fused map).
class VoxelMap {
 public:
  VoxelMap(float voxel, size_t capacity, int device = 0) { check(rtr_map_create(&m_, device, voxel, capacity)); }
  ~VoxelMap() {
    if (m_) rtr_map_destroy(m_);
  }
  VoxelMap(const VoxelMap&) = delete;
  VoxelMap& operator=(const VoxelMap&) = delete;
  template <class PointT, class S>
  void integrate(const PointCloud<PointT>& c, const rtr::Matrix4<S>& T) {
    double d[16];
    for (int i = 0; i < 16; ++i) d[i] = (double)T.m[i];
    check(rtr_map_integrate(m_, c.points.empty() ? nullptr : &c.points[0].x, c.points.size(), sizeof(PointT), d));
  }
  size_t size() {
    size_t n = 0;
    check(rtr_map_size(m_, &n));
    return n;
  }
  PointCloud<PointXYZ> centroids() {
    const size_t n = size();
    std::vector<float> xyz(3 * (n ? n : 1));
    size_t out = 0;
    check(rtr_map_extract(m_, nullptr, xyz.data(), nullptr, n ? n : 1, &out, 1));
    PointCloud<PointXYZ> c;
    c.resize(out);
    for (size_t i = 0; i < out; ++i) c.points[i] = PointXYZ(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
    return c;
  }

 private:
  rtr_map* m_ = nullptr;
};

namespace io {
// pcl::io::loadPCDFile for the x/y/z fields
template <class PointT>
inline int loadPCDFile(const std::string& path, PointCloud<PointT>& cloud) {
  size_t n = 0;
  int hn = 0;
  int rc = rtr_pcd_info(path.c_str(), &n, &hn);
  if (rc) return -1;
  std::vector<float> xyz(3 * (n ? n : 1));
  size_t got = 0;
  rc = rtr_pcd_read(path.c_str(), xyz.data(), nullptr, n ? n : 1, &got);
  if (rc) return -1;
  cloud.resize(got);
  for (size_t i = 0; i < got; ++i) {
    cloud.points[i].x = xyz[3 * i];
    cloud.points[i].y = xyz[3 * i + 1];
    cloud.points[i].z = xyz[3 * i + 2];
  }
  return 0;
}
// pcl::io::savePCDFileASCII (8 significant digits)
template <class PointT>
inline int savePCDFileASCII(const std::string& path, const PointCloud<PointT>& cloud) {
  return rtr_pcd_write_ascii(path.c_str(), cloud.points.empty() ? nullptr : &cloud.points[0].x, cloud.points.size(),
                             sizeof(PointT)) == RTR_OK
             ? 0
             : -1;
}
}  // namespace io

}  // namespace rtr
