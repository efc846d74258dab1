// Drives the drop-in tsdf::TSDFHandler exactly as the reference's ROS node
// does (initGlobalVars -> new TSDFHandler -> per scan
// processPointCloudAndUpdateVoxels with caller-owned output arrays of
// PUBLISH_VOXELS_MAX_SIZE entries and a zeroed count).  The other modes call
// the device-pointer stages that processPointCloudAndUpdateVoxels runs in the
// reference (tsdf_handler.cu:974-1016), on points, sensor position and point
// count copied to device memory:
//   members: allocateVoxelBlocksAndUpdateVoxels -> retrievePublishableVoxels
//            -> garbageCollectDistantBlocks
//   stages:  getVoxelBlocks -> allocateVoxelBlocks -> getVoxelsAndUpdate ->
//            retrievePublishableVoxels -> garbageCollectDistantBlocks
//   viz:     scan, then the published voxels as a Tsdf message (Voxel.msg
//            records, as the reference's Publisher builds it) handed to the
//            debug consumer tsdf::TsdfDebugViz, stamped 2 s apart
//
//   handler_dropin_test <in.bin> <out.bin> [scan|members|stages|viz]
// in.bin : int32 nscans, float params[5]; per scan: int32 n, float sensor[3], float points[n][4]
// out.bin: per scan: int32 count, then count x {float x, y, z, sdf, weight}; in viz mode then int32 rendered,
//          int32 n_green, int32 n_red, n_green x float[3], n_red x float[3] (the markers after the callback)
#include <hip/hip_runtime_api.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <vector>

#include "tsdf/tsdf_debug_viz.hpp"
#include "tsdf/tsdf_handler.hpp"

namespace {
// device copies of one scan's inputs, as the reference stages them
struct DeviceScan {
  pcl::PointXYZ* points = nullptr;
  Vector3f* lidar = nullptr;
  int* size = nullptr;
  bool load(const pcl::PointCloud<pcl::PointXYZ>& cloud, const Vector3f& s) {
    const int n = (int)cloud.size();
    if (hipMalloc((void**)&points, sizeof(pcl::PointXYZ) * (n > 0 ? n : 1)) != hipSuccess ||
        hipMalloc((void**)&lidar, sizeof(Vector3f)) != hipSuccess || hipMalloc((void**)&size, sizeof(int)) != hipSuccess)
      return false;
    if (n > 0 && hipMemcpy(points, cloud.points.data(), sizeof(pcl::PointXYZ) * n, hipMemcpyHostToDevice) != hipSuccess)
      return false;
    return hipMemcpy(lidar, &s, sizeof(Vector3f), hipMemcpyHostToDevice) == hipSuccess &&
           hipMemcpy(size, &n, sizeof(int), hipMemcpyHostToDevice) == hipSuccess;
  }
  ~DeviceScan() {
    (void)hipFree(points);
    (void)hipFree(lidar);
    (void)hipFree(size);
  }
};
}  // namespace

int main(int argc, char** argv) {
  if (argc != 3 && argc != 4) {
    std::fprintf(stderr, "usage: %s in.bin out.bin [scan|members|stages|viz]\n", argv[0]);
    return 2;
  }
  const int mode = argc == 4 ? (!std::strcmp(argv[3], "members") ? 1 : !std::strcmp(argv[3], "stages") ? 2 : 0) : 0;
  const bool viz_mode = argc == 4 && !std::strcmp(argv[3], "viz");
  tsdf::TsdfDebugViz viz;
  FILE* in = std::fopen(argv[1], "rb");
  FILE* out = std::fopen(argv[2], "wb");
  if (!in || !out) return 2;
  int nscans = 0;
  float params[5];
  if (std::fread(&nscans, 4, 1, in) != 1 || std::fread(params, 4, 5, in) != 5) return 2;
  tsdf::initGlobalVars(params[0], params[1], params[2], params[3], params[4]);
  tsdf::TSDFHandler* handler = new tsdf::TSDFHandler();
  std::vector<Vector3f> pos(PUBLISH_VOXELS_MAX_SIZE);
  std::vector<tsdf::Voxel> data(PUBLISH_VOXELS_MAX_SIZE);
  for (int k = 0; k < nscans; ++k) {
    int n = 0;
    float s[3];
    if (std::fread(&n, 4, 1, in) != 1 || std::fread(s, 4, 3, in) != 3) return 2;
    pcl::PointCloud<pcl::PointXYZ>::Ptr cloud(new pcl::PointCloud<pcl::PointXYZ>());
    cloud->points.resize(n);
    for (int i = 0; i < n; ++i) {
      float p[4];
      if (std::fread(p, 4, 4, in) != 4) return 2;
      cloud->points[i] = pcl::PointXYZ(p[0], p[1], p[2]);
    }
    Vector3f lidar(s[0], s[1], s[2]);
    int size = 0;
    if (mode == 0) {
      handler->processPointCloudAndUpdateVoxels(cloud, &lidar, pos.data(), &size, data.data());
    } else {
      DeviceScan d;
      if (!d.load(*cloud, lidar)) return 3;
      if (mode == 1) {
        handler->allocateVoxelBlocksAndUpdateVoxels(d.points, d.lidar, d.size, n, nullptr, nullptr);
      } else {
        int num_cuda_blocks = 0;
        handler->getVoxelBlocks(num_cuda_blocks, d.points, nullptr, nullptr, d.lidar, d.size);
        handler->allocateVoxelBlocks(nullptr, nullptr, nullptr, nullptr);
        handler->getVoxelsAndUpdate(num_cuda_blocks, d.points, d.lidar, d.size, nullptr, nullptr);
      }
      handler->retrievePublishableVoxels(d.lidar, pos.data(), &size, data.data(), nullptr, nullptr);
      handler->garbageCollectDistantBlocks(d.lidar, nullptr, nullptr);
    }
    std::fwrite(&size, 4, 1, out);
    for (int i = 0; i < size; ++i) {
      const float rec[5] = {pos[i](0), pos[i](1), pos[i](2), data[i].sdf, data[i].weight};
      std::fwrite(rec, 4, 5, out);
    }
    if (viz_mode) {
      std::vector<tsdf::VoxelMsg> msg(size);
      for (int i = 0; i < size; ++i) msg[i] = {data[i].sdf, data[i].weight, pos[i](0), pos[i](1), pos[i](2)};
      const int rendered = viz.fetchTSDF(msg.data(), (uint32_t)size, params[0], 2.0 * k) ? 1 : 0;
      const int ng = (int)viz.green().size(), nr = (int)viz.red().size();
      std::fwrite(&rendered, 4, 1, out);
      std::fwrite(&ng, 4, 1, out);
      std::fwrite(&nr, 4, 1, out);
      if (ng) std::fwrite(viz.green().data(), 12, ng, out);
      if (nr) std::fwrite(viz.red().data(), 12, nr, out);
    }
  }
  delete handler;
  std::fclose(in);
  std::fclose(out);
  std::printf("ok %d scans\n", nscans);
  return 0;
}
