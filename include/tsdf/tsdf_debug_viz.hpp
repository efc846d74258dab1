// Downstream consumer of the published TSDF, in the role of the debug node
// that the reference's README points to (README.md:80-85: debug_viz.cpp of
// the AirSim cinematography package, outside the reference tree).  That node
// subscribes to tsdf_package_msgs/Tsdf and, in its fetchTSDF callback, renders
// at most once every 5 s: every voxel of the message becomes a cube marker
// of the message's voxel_size at the voxel's position, green where
// sdf >= 0 and red where sdf < 0.
//
// Here the split runs on the GPU (tsdf_msg_markers over the message's
// Voxel records) and the marker lists are kept for the caller to draw or
// hand to a marker publisher; each list keeps the message's voxel order.
#pragma once

#include <cstdint>
#include <vector>

#include "compat.hpp"

namespace tsdf {

// tsdf_package_msgs/msg/Voxel.msg: float32 sdf, weight, x, y, z
struct VoxelMsg {
  float sdf, weight, x, y, z;
};
static_assert(sizeof(VoxelMsg) == 20, "Voxel.msg is five packed float32");

class TsdfDebugViz {
 public:
  // period_s: the least time between two renders (5 s in the debug node).
  explicit TsdfDebugViz(double period_s = 5.0) : period_s_(period_s) {}

  // The subscription callback for one Tsdf message (voxels[0..size), its
  // voxel_size, and its time stamp in seconds).  Renders when no message was
  // rendered yet or period_s has passed since the last render; returns
  // whether it rendered.  A failed GPU split is reported on stderr and leaves
  // the previous markers in place (returns false).
  bool fetchTSDF(const VoxelMsg* voxels, uint32_t size, float voxel_size, double stamp_s);

  const std::vector<Vector3f>& green() const { return green_; }  // sdf >= 0
  const std::vector<Vector3f>& red() const { return red_; }      // sdf < 0
  float cube_scale() const { return scale_; }                     // the rendered message's voxel_size
  int renders() const { return renders_; }

 private:
  double period_s_;
  double last_s_ = 0.0;
  bool rendered_ = false;
  int renders_ = 0;
  float scale_ = 0.f;
  std::vector<Vector3f> green_, red_;
};

}  // namespace tsdf
