/*
 * tsdf_hip.h — C ABI of the MI355X (gfx950) lidar -> TSDF integration engine.
 *
 * Plain C types and pointers only.  Every entry point returns a status code
 * (TSDF_OK or a negative TSDF_ERR_*); tsdf_last_error() returns a message for
 * the last failure on the calling thread.  Nothing here calls exit(): the
 * reference aborts the process on CUDA errors (tsdf_handler.cu:15-22) and
 * only prints on overflow (tsdf_handler.cu:43-45, :787, :857).
 *
 * Reference interfaces cited as  H: tsdf_package/src/cuda/tsdf_handler.cu,
 * HH: tsdf_package/include/cuda/tsdf_handler.cuh,
 * CH: tsdf_package/include/cuda/tsdf_container.cuh,
 * C:  tsdf_package/src/cuda/tsdf_container.cu.
 *
 * Threading: an engine owns one HIP stream; calls on one engine must not run
 * concurrently from several host threads.  Maps of one engine share its block
 * heap and hash-entry pool.
 */
#ifndef TSDF_HIP_H
#define TSDF_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define TSDF_OK 0
#define TSDF_ERR_INVALID_ARG (-1)
#define TSDF_ERR_HIP (-2)          /* a HIP runtime call failed */
#define TSDF_ERR_NO_MEMORY (-3)    /* device allocation failed */
#define TSDF_ERR_CAPACITY (-4)     /* pool / table / batch capacity exceeded */
#define TSDF_ERR_BUDGET (-5)       /* map admission refused by the heap budget */
#define TSDF_ERR_DEADLINE (-6)     /* serving: request finished after its deadline */
/* The work completed, but the last integration pass could not store every
 * update: blocks were refused because a hash chain or the overflow-bucket
 * pool was full, or the heap or the map's block quota was exhausted (the
 * reference only prints on heap overflow, tsdf_handler.cu:43-45).  Every
 * block that was stored holds all of its updates; the refused blocks' updates
 * are lost.  tsdf_last_stats fills in the counts (insert_failures,
 * heap_failures, quota_failures, updates_dropped) and returns this code. */
#define TSDF_ERR_DROPPED (-7)

/* Voxel state formats of an engine's block heap. */
#define TSDF_STATE_F32 0  /* {float sdf, float weight}: the reference Voxel, CH:40-46 */
#define TSDF_STATE_F16 1  /* {half sdf/truncation, half weight} packed in 4 bytes   */
#define TSDF_STATE_F16_QW 2 /* {half sdf/truncation, 16-bit fixed-point weight, step max_weight/65535} */

typedef struct tsdf_engine tsdf_engine;
typedef struct tsdf_map tsdf_map;

/* Integration parameters: the five arguments of tsdf::initGlobalVars
 * (HH:80-81, H:1289-1317).  Derived constants (half sizes, block size,
 * traversal bounds) are computed from them exactly as initGlobalVars does. */
typedef struct {
  float voxel_size;
  float truncation_distance;
  float max_weight;
  float publish_distance_squared;
  float garbage_collect_distance_squared;
} tsdf_params;

/* Engine sizing.  Replaces the compile-time NUM_HEAP_BLOCKS, NUM_BUCKETS and
 * HASH_ENTRIES_PER_BUCKET of CH:23-31 with run-time sizes chosen for HBM. */
typedef struct {
  int32_t device;            /* HIP device ordinal */
  int32_t state_format;      /* TSDF_STATE_F32 or TSDF_STATE_F16 */
  int64_t heap_blocks;       /* block heap capacity, blocks of 8^3 voxels */
  int64_t total_buckets;     /* hash buckets (8 entries each) shared by all maps */
  int64_t max_batch_points;  /* initial scratch capacity in points (grows on demand) */
} tsdf_engine_config;

/* Per-map hash layout and heap quota (0 in a field selects its default). */
typedef struct {
  int32_t buckets_log2;      /* home buckets = 2^buckets_log2 (default 16) */
  int32_t overflow_buckets;  /* overflow bucket pool (default home/4) */
  int32_t max_chain;         /* max overflow buckets chained to one home bucket (default 64) */
  int32_t reserved;
  int64_t block_quota;       /* max heap blocks this map may hold (default: no quota) */
} tsdf_map_config;

/* Counters of one integrate / batch call. */
typedef struct {
  int64_t points;            /* input points */
  int64_t valid;             /* points integrated */
  int64_t skipped;           /* non-finite, zero-range or out-of-range points */
  int64_t block_cells;       /* block-DDA cells emitted */
  int64_t voxel_updates;     /* voxel updates folded */
  int64_t unique_voxels;     /* distinct voxels updated */
  int64_t new_blocks;        /* blocks allocated */
  int64_t trunc_vox;         /* points whose voxel DDA hit MAX_VOXELS_TRAVERSED */
  int64_t trunc_blk;         /* points whose block DDA hit the block bound */
  int64_t index_dropped;     /* updates whose local voxel index fell outside 0..7 */
  int64_t insert_failures;   /* block inserts refused: chain or overflow pool full */
  int64_t heap_failures;     /* blocks refused: heap or quota exhausted */
  int64_t updates_dropped;   /* updates lost to insert/heap failures */
  int64_t published;         /* voxels published by the last publish call */
  int64_t gc_removed;        /* blocks removed by the last gc call */
  int64_t quota_failures;    /* subset of heap_failures caused by block_quota */
} tsdf_scan_stats;

/* One scan of a batch. */
typedef struct {
  int32_t map_index;         /* index into the maps array of the batch call */
  int32_t n_points;
  int64_t point_offset;      /* first point of this scan in the batch point array */
  float sensor[3];           /* lidar position, the reference's lidar_position_h */
  int32_t reserved;
} tsdf_scan_desc;

/* ---- errors --------------------------------------------------------------- */
const char* tsdf_last_error(void);

/* ---- engine ----------------------------------------------------------------
 * Replaces: TSDFContainer::TSDFContainer / ~TSDFContainer (C:11-26,
 * CH:112-133): allocation of the device hash table and block heap.  Unlike
 * the reference, one engine holds many maps. */
int tsdf_engine_create(const tsdf_engine_config* cfg, tsdf_engine** out);
int tsdf_engine_destroy(tsdf_engine* eng);
/* Blocks until all work queued on the engine stream has finished. */
int tsdf_engine_synchronize(tsdf_engine* eng);
/* The engine's HIP stream (hipStream_t) for callers that order their own work. */
void* tsdf_engine_stream(tsdf_engine* eng);
/* Free heap blocks right now (synchronises). */
int tsdf_engine_free_blocks(tsdf_engine* eng, int64_t* out);

/* ---- maps ------------------------------------------------------------------
 * Replaces: TSDFHandler::TSDFHandler (H:954-958) together with
 * initGlobalVars (H:1289-1317): a map owns its parameters instead of
 * process-global __constant__ symbols (HH:23-37).  cfg may be NULL. */
int tsdf_map_create(tsdf_engine* eng, const tsdf_params* params, const tsdf_map_config* cfg, tsdf_map** out);
/* Replaces ~TSDFHandler (H:960-964): returns the map's blocks to the heap. */
int tsdf_map_destroy(tsdf_map* map);
int tsdf_map_num_blocks(tsdf_map* map, int64_t* out);
/* Copies every block of the map to host memory: keys int32[n][3] (integer
 * block coordinates, floor(x / (8 * voxel_size))) and voxels float[n][512][2]
 * ({sdf, weight}, local index x + 8y + 64z as in H:507-512).  Writes at most
 * capacity blocks; *count receives the total.  Either output may be NULL. */
int tsdf_map_export(tsdf_map* map, int32_t* keys, float* voxels, int64_t capacity, int64_t* count);
/* No reference counterpart (the reference has no snapshot path): restores
 * blocks from host arrays in the tsdf_map_export layout, so that a map
 * exported and destroyed to free HBM can be rebuilt later.  Keys must be
 * distinct, each coordinate within +-8192 (the block range of the point
 * validity rule).  Each block is found or inserted in the map's hash, takes
 * a heap block when new (subject to the free heap and the map's block_quota)
 * and has all 512 voxels overwritten; in a packed state format the values
 * are re-encoded.  *imported receives the blocks stored, *failed those
 * refused for lack of heap, quota or hash space (either may be NULL).
 * Synchronous. */
int tsdf_map_import(tsdf_map* map, const int32_t* keys, const float* voxels, int64_t n, int64_t* imported,
                    int64_t* failed);

/* ---- per-scan path ---------------------------------------------------------
 * Replaces: TSDFHandler::allocateVoxelBlocksAndUpdateVoxels (H:1027-1060),
 * i.e. getVoxelBlocks (H:1071-1094), allocateVoxelBlocks (H:1103-1153) and
 * getVoxelsAndUpdate (H:1164-1182).  points_dev: device array of n
 * pcl::PointXYZ-layout float4 (x, y, z, pad), 16-byte aligned
 * (TSDF_ERR_INVALID_ARG otherwise).  Asynchronous on the engine stream; the
 * point buffer must stay valid until the work completes. */
int tsdf_integrate(tsdf_map* map, const float* points_dev, int64_t n, const float sensor[3]);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): tsdf_integrate with the pass scratch taken from a fixed
 * 96 MiB device array instead of the engine's scratch, which tsdf_integrate
 * grows with hipMalloc when a scan outgrows it.  The call allocates and frees
 * nothing.  The array holds the pair rows and update records of n points at
 * the map's MAX_VOXELS_TRAVERSED slots each (16 bytes per point and slot, 1
 * per point) and the fold descriptors of each block size class, 16 bytes
 * each, as many as the class can hold in a pass over the map: a 131,072-point
 * scan at voxel 0.1 m / truncation 0.4 m (32 slots) of a map of 2^17 home
 * buckets needs 92 MiB.  Larger calls return
 * TSDF_ERR_CAPACITY with nothing enqueued.  Each GPU holds one such array
 * for the life of the process, shared by every engine on it: calls on one
 * GPU must not overlap (on one engine they are ordered by its stream; calls
 * from engines with different streams on one GPU race unless the caller
 * orders them).  Results equal tsdf_integrate's.  Asynchronous. */
int tsdf_integrate_noalloc(tsdf_map* map, const float* points_dev, int64_t n, const float sensor[3]);

/* Replaces: TSDFHandler::retrievePublishableVoxels (H:1193-1224).  Writes
 * every voxel with weight != 0 in blocks whose centre lies within
 * sqrt(publish_distance_squared) of the sensor: positions float[n][3] and
 * data float[n][2] = {sdf, weight}, i.e. the reference's Vector3f[] and
 * Voxel[] output arrays (HH:49-51).  Counts first, so there is no fixed
 * capacity: *count receives the total, at most capacity entries are written.
 * Host pointers; synchronous. */
int tsdf_publish(tsdf_map* map, const float sensor[3], float* pos, float* data, int64_t capacity, int64_t* count);

/* Publishes straight into the tsdf_package_msgs/Voxel wire layout
 * (tsdf_package_msgs/msg/Voxel.msg: float32 sdf, weight, x, y, z; 20 bytes per
 * voxel), replacing the host push_back loop of Publisher::publish
 * (tsdf_package/src/publisher.cpp:13-25).  Host pointer; synchronous. */
int tsdf_publish_msg(tsdf_map* map, const float sensor[3], float* voxels_msg, int64_t capacity, int64_t* count);

/* Replaces: TSDFHandler::garbageCollectDistantBlocks (H:1232-1279).  Removes
 * every block farther than sqrt(garbage_collect_distance_squared) from the
 * sensor and returns its heap block to the free list.  removed may be NULL
 * (then the call is asynchronous). */
int tsdf_gc(tsdf_map* map, const float sensor[3], int64_t* removed);

/* Replaces: TSDFHandler::processPointCloudAndUpdateVoxels (H:974-1016): the
 * whole per-scan sequence integrate -> publish -> gc with host inputs and
 * outputs.  points_host: n float4 rows; pos/data as for tsdf_publish.
 * Synchronous.  Returns TSDF_ERR_DROPPED when the scan's integration dropped
 * updates; publish and gc still ran and the outputs are complete. */
int tsdf_process_scan(tsdf_map* map, const float* points_host, int64_t n, const float sensor[3],
                      float* pos, float* data, int64_t capacity, int64_t* count);

/* Counters of the last integrate/batch/publish/gc call on this engine
 * (synchronises).  Returns TSDF_ERR_DROPPED, with *out filled in, when the
 * last integration pass dropped updates; integration calls themselves are
 * asynchronous, so this (or tsdf_process_scan's status) is where a caller
 * learns of it. */
int tsdf_last_stats(tsdf_engine* eng, tsdf_scan_stats* out);

/* ---- batched path ------------------------------------------------------------
 * No reference counterpart (the reference integrates one scan per call).
 * Integrates nscans scans of possibly different maps in one pass of the
 * pipeline.  Scans of the same map are folded in array order, so the result
 * equals calling tsdf_integrate for each scan in order.  points_dev holds all
 * scans' points (scan i at point_offset).  Asynchronous. */
int tsdf_integrate_batch(tsdf_engine* eng, tsdf_map* const* maps, int32_t nmaps,
                         const tsdf_scan_desc* scans, int32_t nscans, const float* points_dev);

/* A batch pipeline captured once into a HIP graph for up to max_points
 * points and max_scans scans over a fixed set of maps (all of eng), then
 * replayed with new scan descriptors and point buffers.  The graph does not
 * keep its maps alive: a map may be destroyed while the graph exists, and a
 * later launch with a scan of that map returns TSDF_ERR_INVALID_ARG without
 * enqueueing anything (launches with scans of the remaining maps still run). */
typedef struct tsdf_batch_graph tsdf_batch_graph;
int tsdf_batch_graph_create(tsdf_engine* eng, tsdf_map* const* maps, int32_t nmaps,
                            int64_t max_points, int32_t max_scans, tsdf_batch_graph** out);
int tsdf_batch_graph_launch(tsdf_batch_graph* g, const tsdf_scan_desc* scans, int32_t nscans, const float* points_dev);
int tsdf_batch_graph_destroy(tsdf_batch_graph* g);

/* Launch-shape knobs of the pass, set by the autotuner.  A traversal chunk
 * is rays_per_lane x traverse_block consecutive points sharing one LDS block
 * table, i.e. 64 x rays_per_lane rays per wavefront. */
typedef struct {
  int32_t traverse_block;    /* threads per workgroup of the traversal kernel: 64, 128, 192 or 256 */
  int32_t traverse_grid_cap; /* max workgroups of the traversal kernel (1..65536) */
  int32_t fold_grid;         /* workgroups of the first fold launch, 1..65536 (its work items are spread over them) */
  int32_t rays_per_lane;     /* rays each traversal lane walks per chunk: 1..4 (0 reads as 1) */
  int32_t traverse_queue;    /* 1: the traversal kernel runs as a persistent grid of resident workgroups that take
                                chunks from a device work queue; 0: one workgroup per chunk up to traverse_grid_cap */
  int32_t reserved;          /* 0 */
} tsdf_launch_config;
int tsdf_engine_set_launch(tsdf_engine* eng, const tsdf_launch_config* cfg);
int tsdf_engine_get_launch(tsdf_engine* eng, tsdf_launch_config* cfg);

/* ---- timing -----------------------------------------------------------------
 * When enabled, each pipeline stage is bracketed by HIP events on the engine
 * stream.  tsdf_timing_read synchronises and returns per-stage totals (ms)
 * and launch counts since the last reset, for stage ids below. */
#define TSDF_STAGE_TRAVERSE 0
#define TSDF_STAGE_ASSIGN 1
#define TSDF_STAGE_SCAN 2
#define TSDF_STAGE_SCATTER 3
#define TSDF_STAGE_FOLD 4
#define TSDF_STAGE_PUBLISH 5
#define TSDF_STAGE_GC 6
#define TSDF_STAGE_TOTAL 7
#define TSDF_NUM_STAGES 8
int tsdf_timing_enable(tsdf_engine* eng, int on);
int tsdf_timing_read(tsdf_engine* eng, double* ms, int64_t* launches, int reset);

/* ---- serving ----------------------------------------------------------------
 * No reference counterpart.  A continuous-batching request loop: requests
 * (scans of resident maps) become eligible at their arrival time; while the
 * device works on earlier batches, newly arrived requests are gathered into
 * the next batch (at most max_batch_scans scans and max_batch_points points,
 * oldest first; a batch folds the scans of one map in arrival order, so each
 * map's scans keep their order).  Up to
 * max_inflight batches are queued on the device.  Arrival times are seconds
 * relative to the start of the run; the loop sleeps until arrivals.
 * latency_ms[i] receives completion minus arrival of request i, and
 * batch_of[i] the batch it ran in (either may be NULL). */
typedef struct {
  int32_t map_index;         /* index into the maps array */
  int32_t n_points;
  int64_t point_offset;      /* into points_dev */
  float sensor[3];
  float arrival_s;
} tsdf_request;

typedef struct {
  int32_t max_batch_scans;
  int32_t max_inflight;
  int64_t max_batch_points;
  double deadline_ms;        /* requests slower than this are counted as misses (0: none) */
} tsdf_serve_config;

typedef struct {
  int64_t requests, batches, deadline_misses;
  double wall_s;             /* first arrival to last completion */
  double busy_s;             /* summed device time of the batches */
  int64_t points;
} tsdf_serve_stats;

int tsdf_serve_run(tsdf_engine* eng, tsdf_map* const* maps, int32_t nmaps,
                   const tsdf_request* reqs, int32_t nreq, const float* points_dev,
                   const tsdf_serve_config* cfg, double* latency_ms, int32_t* batch_of,
                   tsdf_serve_stats* stats);

/* Map admission against a heap budget: succeeds when the engine's free heap
 * blocks minus the quotas already promised to admitted maps can cover
 * cfg->block_quota; otherwise returns TSDF_ERR_BUDGET without creating the map. */
int tsdf_map_admit(tsdf_engine* eng, const tsdf_params* params, const tsdf_map_config* cfg, tsdf_map** out);

/* ---- device memory plumbing ---------------------------------------------------
 * No reference counterpart: lets hosts without their own GPU runtime stage
 * point clouds in HBM.  stream-ordered on the engine stream. */
int tsdf_device_alloc(tsdf_engine* eng, int64_t bytes, void** out);
int tsdf_device_free(tsdf_engine* eng, void* ptr);
int tsdf_copy_to_device(tsdf_engine* eng, void* dst, const void* src, int64_t bytes);
int tsdf_copy_to_host(tsdf_engine* eng, void* dst, const void* src, int64_t bytes);
/* Device-to-device copy with a plain streaming kernel (bytes: multiple of 16,
 * both pointers 16-byte aligned); synchronous.  Also serves as the
 * known-size reference for HBM byte counters. */
int tsdf_device_copy(tsdf_engine* eng, void* dst, const void* src, int64_t bytes);
/* Page-locked host memory (no reference counterpart: the reference's caller
 * passes pageable arrays).  Host points and publish outputs placed here move
 * over PCIe at DMA speed instead of through the runtime's staging copies,
 * which for a large publish is most of the call.  The memory outlives the
 * engine; tsdf_host_free takes no engine so that it can be released after. */
int tsdf_host_alloc(tsdf_engine* eng, int64_t bytes, void** out);
int tsdf_host_free(void* ptr);

/* ---- input path ------------------------------------------------------------------
 * Replaces: Transformer::convertPointCloud2ToPointCloudXYZ
 * (tsdf_package/src/transformer.cpp:63-67, pcl_conversions::toPCL +
 * pcl::fromPCLPointCloud2) for the fields the integration reads: the x, y, z
 * fields of a sensor_msgs/PointCloud2 data block become the pcl::PointXYZ
 * float4 rows (x, y, z, 1) that tsdf_integrate takes, on the device.
 * Points are taken row by row (width * height of them); non-finite and
 * zero-range points are kept and later skipped by the integration, which
 * counts them in tsdf_scan_stats.skipped. */
#define TSDF_FIELD_FLOAT32 7 /* sensor_msgs/PointField FLOAT32, the only type PointXYZ fields match */
typedef struct {
  int64_t width, height;     /* PointCloud2 width and height */
  int64_t row_step;          /* bytes per row */
  int64_t data_bytes;        /* length of the data block (row_step * height in a well-formed message) */
  int32_t point_step;        /* bytes per point */
  int32_t x_offset, y_offset, z_offset; /* PointField offsets of x, y, z */
  int32_t datatype;          /* PointField datatype of x, y, z: TSDF_FIELD_FLOAT32 */
  int32_t is_bigendian;      /* PointCloud2 is_bigendian */
} tsdf_cloud2_layout;
/* cloud: the data block, in device memory when cloud_on_device is nonzero
 * (then asynchronous on the engine stream), else in host memory (copied
 * first; synchronous).  points_dev (16-byte aligned) receives width * height
 * float4 rows (capacity rows at most: TSDF_ERR_CAPACITY beyond); *n_points
 * their number.
 * A layout whose fields or rows reach past data_bytes, or fields of another
 * datatype, is refused with TSDF_ERR_INVALID_ARG. */
int tsdf_cloud2_to_points(tsdf_engine* eng, const tsdf_cloud2_layout* layout, const void* cloud, int cloud_on_device,
                          float* points_dev, int64_t capacity, int64_t* n_points);

/* ---- downstream consumer of the Tsdf message -------------------------------
 * Replaces: the message's debug consumer, debug_viz.cpp's fetchTSDF (referenced
 * by README.md:80-85, outside the reference tree), which draws every voxel of
 * a tsdf_package_msgs/Tsdf as a marker: green for sdf >= 0, red for sdf < 0.
 * voxels_msg: count Voxel.msg records {sdf, weight, x, y, z} (20 bytes each,
 * as tsdf_publish_msg writes them), in device memory when msg_on_device is
 * nonzero, else in host memory.  The records are split by the sign of sdf on
 * the GPU, each side keeping the message's order: nonneg_xyz receives the
 * positions of the sdf >= 0 voxels (-0.0 included, as the comparison has it),
 * neg_xyz those of the sdf < 0 voxels, float[count][3] each at most (host
 * pointers); *n_nonneg + *n_neg = count.  Synchronous. */
int tsdf_msg_markers(tsdf_engine* eng, const float* voxels_msg, int64_t count, int msg_on_device, float* nonneg_xyz,
                     float* neg_xyz, int64_t* n_nonneg, int64_t* n_neg);

/* Query of the derived constants initGlobalVars computes (H:1296-1315):
 * out_f = {half_voxel, block_size, half_block, block_eps, voxel_eps},
 * out_i = {MAX_VOXELS_TRAVERSED_PER_LIDAR_POINT, max_voxel_blocks_traversed}. */
int tsdf_derived_params(const tsdf_params* params, float* out_f, int32_t* out_i);

#ifdef __cplusplus
}
#endif
#endif /* TSDF_HIP_H */
