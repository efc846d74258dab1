// Drop-in for tsdf_package/include/cuda/tsdf_handler.cuh: the same class,
// method and function signatures, implemented over the MI355X engine's C ABI
// (include/tsdf_hip.h) by ros_tsdf_amd/csrc/tsdf_handler.cpp.
//
// Behavioural contract (reference citations are tsdf_handler.cu lines):
//  * initGlobalVars sets the process-wide parameters (H:1289-1317); a handler
//    takes the values current at its first scan;
//  * processPointCloudAndUpdateVoxels integrates, publishes into the caller's
//    arrays and garbage-collects, in that order (H:1003-1007); the published
//    count starts at 0 (the reference adds to *publish_voxels_size_h,
//    H:1202) and at most PUBLISH_VOXELS_MAX_SIZE entries are written, with a
//    warning on stderr when more voxels qualified (the reference keeps
//    counting past its arrays, H:780-788);
//  * errors are reported on stderr and leave the map unchanged; nothing
//    calls exit() (the reference exits on CUDA errors, H:15-22).
#pragma once

#include "compat.hpp"
#include "tsdf_container.hpp"

// tsdf_handler.cuh:14-20
const int PUBLISH_VOXELS_MAX_SIZE = 400000;
const int GARBAGE_LINKED_LIST_BLOCKS_MAX_SIZE = 1000;

namespace tsdf {

// tsdf_handler.cuh:41-78
class TSDFHandler {
 public:
  TSDFHandler();
  ~TSDFHandler();

  void processPointCloudAndUpdateVoxels(pcl::PointCloud<pcl::PointXYZ>::Ptr point_cloud, Vector3f* lidar_position_h,
                                        Vector3f* publish_voxels_pos_h, int* publish_voxels_size_h,
                                        Voxel* publish_voxels_data_h);

  // Device-pointer stages of the reference.  lidar_points_d must be device
  // memory holding point_cloud_size_h points; lidar_position_d and
  // lidar_points_size_d device pointers as in the reference.
  void allocateVoxelBlocksAndUpdateVoxels(pcl::PointXYZ* lidar_points_d, Vector3f* lidar_position_d,
                                          int* lidar_points_size_d, int point_cloud_size_h, HashTable* hash_table_d,
                                          BlockHeap* block_heap_d);
  // Block discovery and allocation are fused into the update pass of the
  // engine; these two stages are accepted and do nothing.
  void getVoxelBlocks(int num_cuda_blocks, pcl::PointXYZ* lidar_points_d, Vector3f* point_cloud_voxel_blocks_d,
                      int* point_cloud_voxel_blocks_size_d, Vector3f* lidar_position_d, int* lidar_points_size_d);
  void allocateVoxelBlocks(Vector3f* lidar_points_d, int* point_cloud_voxel_blocks_size_d, HashTable* hash_table_d,
                           BlockHeap* block_heap_d);
  // Integrates the scan (block allocation included).
  void getVoxelsAndUpdate(int& num_cuda_blocks, pcl::PointXYZ* lidar_points_d, Vector3f* lidar_position_d,
                          int* lidar_points_size_d, HashTable* hash_table_d, BlockHeap* block_heap_d);
  void retrievePublishableVoxels(Vector3f* lidar_position_d, Vector3f* publish_voxels_pos_h, int* publish_voxels_size_h,
                                 Voxel* publish_voxels_data_h, HashTable* hash_table_d, BlockHeap* block_heap_d);
  void garbageCollectDistantBlocks(Vector3f* lidar_position_d, HashTable* hash_table_d, BlockHeap* block_heap_d);

 private:
  TSDFContainer* tsdf_container;
  Vector3f* publish_voxels_pos_d;  // unused: the engine sizes publish buffers by count
  Voxel* publish_voxels_data_d;    // unused
  ::tsdf_map* map();
};

// tsdf_handler.cuh:80-81
void initGlobalVars(float voxel_size_input, float truncation_distance_input, float max_weight_input,
                    float publish_distance_squared_input, float garbage_collect_distance_squared_input);

}  // namespace tsdf
