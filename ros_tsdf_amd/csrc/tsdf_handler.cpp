// tsdf::TSDFHandler / tsdf::TSDFContainer / tsdf::initGlobalVars over the C
// ABI of the HIP engine: the reference's host API (tsdf_handler.cuh:41-81,
// tsdf_container.cuh:112-133) for callers such as the ROS2 node.
#include <cstddef>
#include <cstdio>
#include <cstdlib>
#include <mutex>

#include "tsdf/tsdf_debug_viz.hpp"
#include "tsdf/tsdf_handler.hpp"
#include "tsdf_hip.h"

static_assert(offsetof(tsdf::Voxel, sdf) == 0 && offsetof(tsdf::Voxel, weight) == 4, "Voxel is {float sdf, float weight}");

namespace {

// Process-wide parameters, as the reference keeps them in __constant__
// memory (tsdf_handler.cuh:23-37).  Defaults: the reference README's values.
tsdf_params g_params = {0.5f, 0.1f, 10000.0f, 2500.0f, 250000.0f};
bool g_params_set = false;

std::mutex g_mu;
tsdf_engine* g_engine = nullptr;
int g_handlers = 0;

tsdf_engine* engine() {
  if (g_engine) return g_engine;
  tsdf_engine_config c{};
  c.device = 0;
  c.state_format = TSDF_STATE_F32;
  const char* hb = std::getenv("TSDF_HEAP_BLOCKS");
  c.heap_blocks = hb ? std::atoll(hb) : tsdf::NUM_HEAP_BLOCKS;
  c.total_buckets = 1 << 20;
  c.max_batch_points = 1 << 17;
  if (tsdf_engine_create(&c, &g_engine) != TSDF_OK) {
    std::fprintf(stderr, "tsdf: engine creation failed: %s\n", tsdf_last_error());
    g_engine = nullptr;
  }
  return g_engine;
}

void report(const char* what) { std::fprintf(stderr, "tsdf: %s failed: %s\n", what, tsdf_last_error()); }

bool read_device_vec3(const Vector3f* d, float out[3]) {
  tsdf_engine* E = engine();
  if (!E || !d) return false;
  return tsdf_copy_to_host(E, out, d, sizeof(float) * 3) == TSDF_OK;
}

}  // namespace

namespace tsdf {

void initGlobalVars(float voxel_size, float truncation_distance, float max_weight, float publish_distance_squared,
                    float garbage_collect_distance_squared) {
  std::lock_guard<std::mutex> lk(g_mu);
  g_params = {voxel_size, truncation_distance, max_weight, publish_distance_squared, garbage_collect_distance_squared};
  g_params_set = true;
}

TSDFContainer::TSDFContainer() : map_(nullptr) {
  std::lock_guard<std::mutex> lk(g_mu);
  tsdf_engine* E = engine();
  if (!E) return;
  tsdf_map_config mc{};
  mc.buckets_log2 = 16;
  mc.max_chain = 64;
  if (tsdf_map_create(E, &g_params, &mc, &map_) != TSDF_OK) {
    report("map creation");
    map_ = nullptr;
  }
  g_handlers++;
}

TSDFContainer::~TSDFContainer() {
  std::lock_guard<std::mutex> lk(g_mu);
  if (map_) tsdf_map_destroy(map_);
  map_ = nullptr;
  if (--g_handlers == 0 && g_engine) {
    tsdf_engine_destroy(g_engine);
    g_engine = nullptr;
  }
}

HashTable* TSDFContainer::getCudaHashTable() { return reinterpret_cast<HashTable*>(map_); }
BlockHeap* TSDFContainer::getCudaBlockHeap() { return reinterpret_cast<BlockHeap*>(map_); }
::tsdf_map* TSDFContainer::map() { return map_; }

TSDFHandler::TSDFHandler() : tsdf_container(nullptr), publish_voxels_pos_d(nullptr), publish_voxels_data_d(nullptr) {}

TSDFHandler::~TSDFHandler() { delete tsdf_container; }

::tsdf_map* TSDFHandler::map() {
  if (!tsdf_container) tsdf_container = new TSDFContainer();
  return tsdf_container->map();
}

void TSDFHandler::processPointCloudAndUpdateVoxels(pcl::PointCloud<pcl::PointXYZ>::Ptr point_cloud,
                                                   Vector3f* lidar_position_h, Vector3f* publish_voxels_pos_h,
                                                   int* publish_voxels_size_h, Voxel* publish_voxels_data_h) {
  ::tsdf_map* m = map();
  if (!m || !point_cloud || !lidar_position_h) {
    std::fprintf(stderr, "tsdf: processPointCloudAndUpdateVoxels: invalid arguments or no map\n");
    if (publish_voxels_size_h) *publish_voxels_size_h = 0;
    return;
  }
  const float sensor[3] = {(*lidar_position_h)(0), (*lidar_position_h)(1), (*lidar_position_h)(2)};
  const int64_t n = (int64_t)point_cloud->size();
  // pcl::PointXYZ is x, y, z, pad: the engine's float4 input layout
  const float* pts = n > 0 ? reinterpret_cast<const float*>(&point_cloud->points[0]) : nullptr;
  int64_t count = 0;
  const int64_t cap = (publish_voxels_pos_h && publish_voxels_data_h) ? PUBLISH_VOXELS_MAX_SIZE : 0;
  // the caller's arrays are the engine's output layout (Vector3f: three packed
  // floats, Voxel: {sdf, weight}), so the voxels are copied into them directly
  float* pos = cap ? publish_voxels_pos_h[0].data() : nullptr;
  float* data = cap ? &publish_voxels_data_h[0].sdf : nullptr;
  int r = tsdf_process_scan(m, pts, n, sensor, pos, data, cap, &count);
  // the reference prints on heap overflow and carries on (tsdf_handler.cu:43-45): so does the
  // handler when blocks were refused; the published voxels are complete
  if (r == TSDF_ERR_DROPPED) {
    std::fprintf(stderr, "tsdf: %s\n", tsdf_last_error());
    r = TSDF_OK;
  }
  if (r != TSDF_OK) {
    report("processPointCloudAndUpdateVoxels");
    if (publish_voxels_size_h) *publish_voxels_size_h = 0;
    return;
  }
  const int64_t k = count < cap ? count : cap;
  if (count > cap) std::fprintf(stderr, "tsdf: %lld voxels to publish, %lld fit the output arrays\n", (long long)count, (long long)cap);
  if (publish_voxels_size_h) *publish_voxels_size_h = (int)k;
}

void TSDFHandler::allocateVoxelBlocksAndUpdateVoxels(pcl::PointXYZ* lidar_points_d, Vector3f* lidar_position_d,
                                                     int* lidar_points_size_d, int point_cloud_size_h, HashTable*,
                                                     BlockHeap*) {
  (void)lidar_points_size_d;
  ::tsdf_map* m = map();
  float s[3];
  if (!m || !read_device_vec3(lidar_position_d, s)) {
    report("allocateVoxelBlocksAndUpdateVoxels");
    return;
  }
  if (tsdf_integrate(m, reinterpret_cast<const float*>(lidar_points_d), point_cloud_size_h, s) != TSDF_OK ||
      tsdf_engine_synchronize(engine()) != TSDF_OK) {
    report("allocateVoxelBlocksAndUpdateVoxels");
    return;
  }
  tsdf_scan_stats st;
  if (tsdf_last_stats(engine(), &st) == TSDF_ERR_DROPPED) std::fprintf(stderr, "tsdf: %s\n", tsdf_last_error());
}

void TSDFHandler::getVoxelBlocks(int, pcl::PointXYZ*, Vector3f*, int*, Vector3f*, int*) {}

void TSDFHandler::allocateVoxelBlocks(Vector3f*, int*, HashTable*, BlockHeap*) {}

void TSDFHandler::getVoxelsAndUpdate(int&, pcl::PointXYZ* lidar_points_d, Vector3f* lidar_position_d,
                                     int* lidar_points_size_d, HashTable*, BlockHeap*) {
  int n = 0;
  tsdf_engine* E = engine();
  if (!E || !lidar_points_size_d || tsdf_copy_to_host(E, &n, lidar_points_size_d, sizeof(int)) != TSDF_OK) {
    report("getVoxelsAndUpdate");
    return;
  }
  allocateVoxelBlocksAndUpdateVoxels(lidar_points_d, lidar_position_d, lidar_points_size_d, n, nullptr, nullptr);
}

void TSDFHandler::retrievePublishableVoxels(Vector3f* lidar_position_d, Vector3f* publish_voxels_pos_h,
                                            int* publish_voxels_size_h, Voxel* publish_voxels_data_h, HashTable*,
                                            BlockHeap*) {
  ::tsdf_map* m = map();
  float s[3];
  if (!m || !read_device_vec3(lidar_position_d, s)) {
    report("retrievePublishableVoxels");
    return;
  }
  const int64_t cap = (publish_voxels_pos_h && publish_voxels_data_h) ? PUBLISH_VOXELS_MAX_SIZE : 0;
  float* pos = cap ? publish_voxels_pos_h[0].data() : nullptr;
  float* data = cap ? &publish_voxels_data_h[0].sdf : nullptr;
  int64_t count = 0;
  if (tsdf_publish(m, s, pos, data, cap, &count) != TSDF_OK) {
    report("retrievePublishableVoxels");
    return;
  }
  const int64_t k = count < cap ? count : cap;
  if (publish_voxels_size_h) *publish_voxels_size_h = (int)k;
}

void TSDFHandler::garbageCollectDistantBlocks(Vector3f* lidar_position_d, HashTable*, BlockHeap*) {
  ::tsdf_map* m = map();
  float s[3];
  if (!m || !read_device_vec3(lidar_position_d, s) || tsdf_gc(m, s, nullptr) != TSDF_OK)
    report("garbageCollectDistantBlocks");
}

bool TsdfDebugViz::fetchTSDF(const VoxelMsg* voxels, uint32_t size, float voxel_size, double stamp_s) {
  if (rendered_ && stamp_s - last_s_ < period_s_) return false;
  tsdf_engine* E = engine();
  if (!E) return false;
  std::vector<Vector3f> g(size), r(size);
  int64_t ng = 0, nr = 0;
  if (size > 0 &&
      tsdf_msg_markers(E, reinterpret_cast<const float*>(voxels), size, 0, reinterpret_cast<float*>(g.data()),
                       reinterpret_cast<float*>(r.data()), &ng, &nr) != TSDF_OK) {
    report("fetchTSDF");
    return false;
  }
  g.resize(ng);
  r.resize(nr);
  green_.swap(g);
  red_.swap(r);
  scale_ = voxel_size;
  last_s_ = stamp_s;
  rendered_ = true;
  ++renders_;
  return true;
}

}  // namespace tsdf
