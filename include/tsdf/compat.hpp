// Types the reference API is spelled in: Eigen's Vector3f and PCL's
// PointXYZ / PointCloud.  When Eigen or PCL is installed the real types are
// used; otherwise layout-compatible stand-ins are declared (12-byte Vector3f,
// 16-byte PointXYZ with a padding float, a PointCloud with a points vector and
// a shared_ptr Ptr), which is all the integration path touches
// (tsdf_container.cuh:10, tsdf_handler.cu:977-981).
#pragma once

#include <cstddef>
#include <memory>
#include <vector>

#if defined(__has_include)
#if __has_include(<Eigen/Dense>) && !defined(TSDF_NO_EIGEN)
#include <Eigen/Dense>
#define TSDF_HAVE_EIGEN 1
#endif
#if __has_include(<pcl/point_types.h>) && !defined(TSDF_NO_PCL)
#include <pcl/point_cloud.h>
#include <pcl/point_types.h>
#define TSDF_HAVE_PCL 1
#endif
#endif

#ifdef TSDF_HAVE_EIGEN
typedef Eigen::Matrix<float, 3, 1> Vector3f;
#else
struct Vector3f {
  float v[3];
  Vector3f() : v{0.f, 0.f, 0.f} {}
  Vector3f(float x, float y, float z) : v{x, y, z} {}
  float& operator()(int i) { return v[i]; }
  float operator()(int i) const { return v[i]; }
  float& x() { return v[0]; }
  float& y() { return v[1]; }
  float& z() { return v[2]; }
  float x() const { return v[0]; }
  float y() const { return v[1]; }
  float z() const { return v[2]; }
  float* data() { return v; }
  const float* data() const { return v; }
};
#endif
static_assert(sizeof(Vector3f) == 12, "Vector3f must be three packed floats");

#ifndef TSDF_HAVE_PCL
namespace pcl {
struct alignas(16) PointXYZ {
  float x, y, z, pad;
  PointXYZ() : x(0.f), y(0.f), z(0.f), pad(1.f) {}
  PointXYZ(float x_, float y_, float z_) : x(x_), y(y_), z(z_), pad(1.f) {}
};
template <class PointT>
struct PointCloud {
  typedef std::shared_ptr<PointCloud<PointT>> Ptr;
  std::vector<PointT> points;
  std::size_t size() const { return points.size(); }
  void push_back(const PointT& p) { points.push_back(p); }
};
}  // namespace pcl
#endif
static_assert(sizeof(pcl::PointXYZ) == 16, "pcl::PointXYZ must be 16 bytes (x, y, z, pad)");
