// Drop-in for tsdf_package/include/cuda/tsdf_container.cuh: the reference's
// map types and container, backed by the MI355X engine (include/tsdf_hip.h).
#pragma once

#include "compat.hpp"

struct tsdf_map;

namespace tsdf {

// Compile-time constants of the reference header (tsdf_container.cuh:18-38).
// The engine sizes its hash table and block heap at run time; these remain
// for source compatibility and as the defaults of the drop-in container.
const int VOXELS_PER_SIDE = 8;
const int VOXELS_PER_BLOCK = VOXELS_PER_SIDE * VOXELS_PER_SIDE * VOXELS_PER_SIDE;
const int HASH_ENTRIES_PER_BUCKET = 4;
const int NUM_BUCKETS = 131072;
const int HASH_TABLE_SIZE = HASH_ENTRIES_PER_BUCKET * NUM_BUCKETS;
const int NUM_HEAP_BLOCKS = 524288;
const int PRIME_ONE = 73856093;
const int PRIME_TWO = 19349669;
const int PRIME_THREE = 83492791;

// tsdf_container.cuh:40-46
struct Voxel {
  float sdf;
  float weight;
  Voxel() : sdf(0), weight(0) {}
};
static_assert(sizeof(Voxel) == 8, "Voxel is {float sdf, float weight}");

// Opaque here: the device layouts of the hash table and heap belong to the
// engine (GPU spatial hash with overflow chains, HBM-sized block heap).
struct HashTable;
struct BlockHeap;

// tsdf_container.cuh:112-133
class TSDFContainer {
 public:
  TSDFContainer();
  ~TSDFContainer();
  // Opaque handles standing for the reference's device pointers.
  HashTable* getCudaHashTable();
  BlockHeap* getCudaBlockHeap();
  // The engine map behind this container.
  ::tsdf_map* map();

 private:
  ::tsdf_map* map_;
};

}  // namespace tsdf
