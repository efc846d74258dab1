// Device data structures of the engine: the spatial hash (buckets of 8
// entries with overflow chains), the block heap and its free list, per-call
// counters, and the scratch layout of one integration pass.
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <stdint.h>

#include "geom.hpp"

namespace tsdfhip {

constexpr uint64_t kEmptyKey = ~0ull;  // packed keys use 63 bits, so never all ones
constexpr int kBucket = 8;             // entries per bucket (one 64-byte line of keys)
constexpr int kMaxScansPerBatch = 256;
constexpr int kPointIdxBits = 23;      // point index inside one pass
constexpr int kMaxPassPoints = (1 << kPointIdxBits) - 2;
constexpr int kPartSlots = 1024;       // statistics are summed into this many slots, then reduced

// Statistic slots (EngCtr::stats), mirrored by tsdf_scan_stats.
enum Stat {
  ST_POINTS = 0, ST_VALID, ST_SKIPPED, ST_BLOCK_CELLS, ST_UPDATES, ST_UNIQUE, ST_NEW_BLOCKS, ST_TRUNC_VOX,
  ST_TRUNC_BLK, ST_INDEX_DROPPED, ST_INSERT_FAIL, ST_HEAP_FAIL, ST_UPD_DROPPED, ST_PUBLISHED, ST_GC_REMOVED,
  ST_QUOTA_FAIL, ST_COUNT
};

// Statistics of an integration pass: every workgroup of a pass kernel writes
// its own row of counts (no atomics, and nothing to clear between passes),
// and the rows are summed on demand.  Each kernel has a region of rows, one
// per workgroup of its largest grid.
enum StatRegion { SR_BIN = 0, SR_SCAN, SR_FOLD_A, SR_FOLD_B, SR_COUNT };
constexpr int kStatCols = 8;  // counts per row
constexpr int kStatRows[SR_COUNT] = {65536, 256, 65536, 512};
__host__ __device__ constexpr int stat_region_base(int r) {
  int b = 0;
  for (int k = 0; k < r; ++k) b += kStatRows[k];
  return b;
}
constexpr int kStatRowsTotal = stat_region_base(SR_COUNT);

// Per-map constants, resident in device memory (one entry per map slot).
struct MapDev {
  float vs, T, maxW, pubD2, gcD2;
  float half_vs, BL, half_BL;
  float sc_v, sc_b;       // 1/voxel_size and 1/block_size, as getVolumeCenterFromPoint computes them
  float isc_v, isc_b;     // 1/sc_v and 1/sc_b when division by them may use div_by_scale's product form, else 0
  int32_t max_vox, max_blk;
  uint32_t home_mask;     // home buckets - 1 (power of two)
  int32_t bucket_base;    // global index of home bucket 0
  int32_t n_home, n_ovf;  // overflow buckets live at [bucket_base + n_home, + n_ovf)
  int32_t max_chain;
  int32_t slot;
  int64_t block_quota;    // 0 = none
};

// Per-map mutable counters.
struct MapCtr {
  int32_t ovf_top;    // entries on the overflow-bucket free stack
  int32_t pad0;
  int32_t blocks;     // heap blocks held
  int32_t pad;
};

// Engine-wide counters.
struct EngCtr {
  int32_t heap_top;   // entries on the heap free stack
  int32_t touched_n;  // blocks with updates in this pass
  int32_t fail_n;     // entries inserted in this pass that got no heap block
  int32_t sel_n;      // selected blocks (publish / export)
  int32_t freed_n;    // blocks freed by the current gc
  int32_t new_n;      // entries inserted in this pass, waiting for heap blocks
  int32_t fold_next;  // large-block fold items handed out beyond the first per workgroup
  int32_t bin_next;   // k_bin ray chunks handed out from the work queue beyond the first per workgroup
  uint32_t seg_top;   // records given out to the pass's block segments (k_seg, one atomic per tile)
  int32_t fold_n[8];  // touched blocks per fold size class
  unsigned long long stats[ST_COUNT];
};

struct ScanDev {
  int32_t map;        // map slot
  int32_t n;
  int32_t off;        // first point in the pass
  float s[3];
};

struct PassHdr {
  const float4* points;
  int32_t total;      // points in the pass
  int32_t nscans;
  int32_t stride;     // record stride (max MAX_VOXELS_TRAVERSED of the pass's maps)
  int32_t pad;
};

// Everything a kernel may touch, passed by value.
struct Dev {
  // engine tables
  const MapDev* maps;
  MapCtr* mctr;
  EngCtr* ctr;
  unsigned long long* keys;  // [entries]
  int32_t* vals;             // [entries] heap block of an entry, -1 none
  unsigned long long* blk_cnt;  // [entries] pass counter: voxel records in the high 32 bits
  uint32_t* blk_base;        // [entries] first record of the block's segment in this pass
  int32_t* next;             // [buckets] overflow link, -1 none
  int32_t* ovf_stack;        // [buckets] per-map overflow free stacks
  void* heap;                // [heap_blocks][512] voxel state
  int32_t* heap_free;        // [heap_blocks] free stack
  int32_t* freed;            // [heap_blocks] blocks freed by gc (to be zeroed)
  // pass scratch
  const PassHdr* hdr;
  const ScanDev* scans;
  uint32_t* pair_e;          // [stride][pcap] block entry of each (point, block) pair
  uint32_t* pair_r;          // [stride][pcap] record offset of the run inside its block
  uint8_t* pair_n;           // [pcap] pairs of each point
  int64_t pcap;
  uint32_t* touched;         // [entries]
  uint8_t* touched_map;      // [entries] map slot of each touched block
  uint32_t* tile_sums;       // [entries / 1024 + 2]
  uint32_t* fail_list;       // [entries] inserted entries without a heap block (removed after binning)
  uint32_t* new_list;        // [entries] entries inserted in this pass (heap blocks assigned after binning)
  uint8_t* new_map;          // [entries] their map slots
  int4* fold_desc;           // per size class c, from fold_base[c]: {record base, records, entry, map}
                             // (map | kDescDeferred: handed from k_fold_a to k_fold_b)
  unsigned long long* part;  // [ST_COUNT][kPartSlots] statistics partial sums (publish and gc)
  uint32_t* srows;           // [kStatRowsTotal][kStatCols] pass statistics, a row per workgroup
  int64_t fold_cap;
  int64_t fold_base[8];      // first descriptor of each size class
  unsigned long long* seg;   // [records] update records, contiguous per touched block
  int64_t seg_cap;
  uint32_t* sel;             // [entries]
  uint32_t* pub_cnt;         // [entries + 1]
};

// ---- atomics that bypass the per-CU L1 --------------------------------------
__device__ __forceinline__ unsigned long long ld_acq(const unsigned long long* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int32_t ld_acq(const int32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint32_t hash_key(unsigned long long k) {
  k ^= k >> 33;
  k *= 0xff51afd7ed558ccdull;
  k ^= k >> 33;
  k *= 0xc4ceb9fe1a85ec53ull;
  k ^= k >> 33;
  return (uint32_t)k;
}

// ---- voxel state codecs -----------------------------------------------------
// Per-map constants the packed formats need.
struct Codec {
  float T, inv_T;     // truncation distance and its reciprocal
  float wq, inv_wq;   // 65535 / max_weight and its reciprocal (fixed-point weight)
};
__device__ __forceinline__ Codec make_codec(const MapDev& m) {
  Codec c;
  c.T = m.T;
  c.inv_T = 1.0f / m.T;
  c.wq = 65535.0f / m.maxW;
  c.inv_wq = m.maxW / 65535.0f;
  return c;
}

// The reference's Voxel {float sdf, float weight}, 8 bytes.
struct StateF32 {
  static constexpr int kBytes = 8;
  __device__ static __forceinline__ void load(const void* heap, int64_t i, const Codec&, float& S, float& W) {
    const float2 v = reinterpret_cast<const float2*>(heap)[i];
    S = v.x;
    W = v.y;
  }
  __device__ static __forceinline__ void store(void* heap, int64_t i, const Codec&, float S, float W) {
    reinterpret_cast<float2*>(heap)[i] = make_float2(S, W);
  }
  // the state as a store and a load would leave it
  __device__ static __forceinline__ void round(const Codec&, float&, float&) {}
  __device__ static __forceinline__ bool nonzero(const void* heap, int64_t i) {
    return reinterpret_cast<const float2*>(heap)[i].y != 0.0f;
  }
  __device__ static __forceinline__ void zero(void* heap, int64_t i) { reinterpret_cast<float2*>(heap)[i] = make_float2(0.f, 0.f); }
};

// FP16 SDF normalised by the truncation distance plus FP16 weight, 4 bytes.
struct StateF16 {
  static constexpr int kBytes = 4;
  __device__ static __forceinline__ void load(const void* heap, int64_t i, const Codec& c, float& S, float& W) {
    const uint32_t v = reinterpret_cast<const uint32_t*>(heap)[i];
    S = __half2float(__ushort_as_half((unsigned short)(v & 0xFFFFu))) * c.T;
    W = __half2float(__ushort_as_half((unsigned short)(v >> 16)));
  }
  __device__ static __forceinline__ void store(void* heap, int64_t i, const Codec& c, float S, float W) {
    const unsigned short s = __half_as_ushort(__float2half_rn(S * c.inv_T));
    const unsigned short w = __half_as_ushort(__float2half_rn(W));
    reinterpret_cast<uint32_t*>(heap)[i] = (uint32_t)s | ((uint32_t)w << 16);
  }
  __device__ static __forceinline__ void round(const Codec& c, float& S, float& W) {
    S = __half2float(__float2half_rn(S * c.inv_T)) * c.T;
    W = __half2float(__float2half_rn(W));
  }
  __device__ static __forceinline__ bool nonzero(const void* heap, int64_t i) {
    return (reinterpret_cast<const uint32_t*>(heap)[i] >> 16) != 0u;
  }
  __device__ static __forceinline__ void zero(void* heap, int64_t i) { reinterpret_cast<uint32_t*>(heap)[i] = 0u; }
};

// FP16 SDF / truncation plus a 16-bit fixed-point weight: W = q * max_weight
// / 65535, a uniform step of max_weight / 65535 (0.015 at 1000) instead of
// FP16's 0.5 above 512.  4 bytes.
struct StateF16Q {
  static constexpr int kBytes = 4;
  __device__ static __forceinline__ void load(const void* heap, int64_t i, const Codec& c, float& S, float& W) {
    const uint32_t v = reinterpret_cast<const uint32_t*>(heap)[i];
    S = __half2float(__ushort_as_half((unsigned short)(v & 0xFFFFu))) * c.T;
    W = (float)(v >> 16) * c.inv_wq;
  }
  __device__ static __forceinline__ void store(void* heap, int64_t i, const Codec& c, float S, float W) {
    const unsigned short s = __half_as_ushort(__float2half_rn(S * c.inv_T));
    const float q = fminf(fmaxf(rintf(W * c.wq), 0.0f), 65535.0f);
    // a touched voxel keeps a nonzero weight
    const uint32_t w = W > 0.0f ? max((uint32_t)q, 1u) : 0u;
    reinterpret_cast<uint32_t*>(heap)[i] = (uint32_t)s | (w << 16);
  }
  __device__ static __forceinline__ void round(const Codec& c, float& S, float& W) {
    const float q = fminf(fmaxf(rintf(W * c.wq), 0.0f), 65535.0f);
    const uint32_t w = W > 0.0f ? max((uint32_t)q, 1u) : 0u;
    S = __half2float(__float2half_rn(S * c.inv_T)) * c.T;
    W = (float)w * c.inv_wq;
  }
  __device__ static __forceinline__ bool nonzero(const void* heap, int64_t i) {
    return (reinterpret_cast<const uint32_t*>(heap)[i] >> 16) != 0u;
  }
  __device__ static __forceinline__ void zero(void* heap, int64_t i) { reinterpret_cast<uint32_t*>(heap)[i] = 0u; }
};

}  // namespace tsdfhip
