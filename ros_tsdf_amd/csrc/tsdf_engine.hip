// Host side of the engine and the C ABI declared in include/tsdf_hip.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/tsdf_hip.h"
#include "kernels.hpp"

using namespace tsdfhip;

namespace {

thread_local std::string g_last_error;

int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

#define HIPCHK(expr)                                                                           \
  do {                                                                                         \
    hipError_t _e = (expr);                                                                    \
    if (_e != hipSuccess) return fail(TSDF_ERR_HIP, "%s: %s (%s:%d)", #expr, hipGetErrorString(_e), __FILE__, __LINE__); \
  } while (0)

#define CHK(expr)              \
  do {                         \
    int _r = (expr);           \
    if (_r != TSDF_OK) return _r; \
  } while (0)

constexpr int kMaxMaps = 256;
constexpr int kRing = 16;          // staged descriptor slots
constexpr int kMaxStride = 128;    // largest MAX_VOXELS_TRAVERSED a map may have
constexpr int kTimingRing = 256;
// statistics an integration pass produces (publish and gc keep their own)
constexpr uint32_t kPassStats = ((1u << ST_COUNT) - 1u) & ~((1u << ST_PUBLISHED) | (1u << ST_GC_REMOVED));
// the pass statistics that mean updates were lost
constexpr uint32_t kDropStats = (1u << ST_INSERT_FAIL) | (1u << ST_HEAP_FAIL) | (1u << ST_UPD_DROPPED);

struct DescSlot {
  PassHdr hdr;
  ScanDev scans[kMaxScansPerBatch];
};

template <class T>
int dev_alloc(T** p, int64_t count) {
  *p = nullptr;
  if (count <= 0) count = 1;
  hipError_t e = hipMalloc((void**)p, sizeof(T) * (size_t)count);
  if (e != hipSuccess) {
    *p = nullptr;
    (void)hipGetLastError();
    return fail(TSDF_ERR_NO_MEMORY, "hipMalloc of %lld bytes failed: %s", (long long)(sizeof(T) * count), hipGetErrorString(e));
  }
  return TSDF_OK;
}

int grid_for(int64_t n, int block, int cap) {
  int64_t g = (n + block - 1) / block;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (int)g;
}

// Calls f with a value of the voxel-state codec type of format fmt.
template <class F>
auto with_state(int fmt, F&& f) {
  switch (fmt) {
    case TSDF_STATE_F16: return f(StateF16{});
    case TSDF_STATE_F16_QW: return f(StateF16Q{});
    default: return f(StateF32{});
  }
}

}  // namespace

// ---- pass scratch without allocation -------------------------------------------------------
// tsdf_integrate_noalloc takes its pass scratch (pair rows, run counts,
// update records, fold descriptors) from this file-scope device array: each
// GPU that loads the library holds one copy until the process ends, and
// every call on that GPU shares it.  Nothing in it needs to start at zero:
// a pass writes every pair row, run count, record and descriptor it reads.
// In host code the name stands for a host-side shadow, never a device
// address: each device's address is resolved once with hipGetSymbolAddress
// (with that device current, when the first engine is created on it) and
// only that address is passed to kernels.
constexpr int64_t kNoAllocScratchBytes = int64_t(96) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
bool g_scratch_resolved[kMaxDevices] = {};
void* g_scratch_addr[kMaxDevices] = {};
hipError_t g_scratch_status[kMaxDevices] = {};

// Resolves the current device's copy of scratch (once per device).
void resolve_scratch(int dev) {
  if (dev < 0 || dev >= kMaxDevices) return;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_resolved[dev]) return;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) (void)hipGetLastError();
  g_scratch_status[dev] = e;
  g_scratch_addr[dev] = e == hipSuccess ? p : nullptr;
  g_scratch_resolved[dev] = true;
}

// The device address of device dev's scratch, or null (with the reason in *err).
void* scratch_address(int dev, hipError_t* err) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (dev < 0 || dev >= kMaxDevices || !g_scratch_resolved[dev]) {
    *err = hipErrorNotInitialized;
    return nullptr;
  }
  *err = g_scratch_status[dev];
  return g_scratch_addr[dev];
}
}  // namespace

struct tsdf_batch_graph;

struct tsdf_map {
  tsdf_engine* eng = nullptr;
  int slot = -1;
  uint64_t id = 0;  // unique per engine: a slot reused by a later map gets a new id
  MapDev desc{};
  tsdf_params params{};
  int64_t quota_promised = 0;
};

struct tsdf_engine {
  int device = 0;
  int state_format = TSDF_STATE_F32;
  hipStream_t stream = nullptr;
  int64_t heap_blocks = 0, total_buckets = 0, entries = 0;

  // tables
  MapDev* maps_dev = nullptr;
  MapCtr* mctr = nullptr;
  EngCtr* ctr = nullptr;
  EngCtr* ctr_host = nullptr;  // pinned
  unsigned long long* keys = nullptr;
  int32_t* vals = nullptr;
  unsigned long long* blk_cnt = nullptr;
  uint32_t* blk_base = nullptr;
  int32_t* next = nullptr;
  int32_t* ovf_stack = nullptr;
  void* heap = nullptr;
  int32_t* heap_free = nullptr;
  int32_t* freed = nullptr;

  // pass scratch
  DescSlot* desc_dev = nullptr;  // the pass header and scan descriptors, one copy per pass
  DescSlot* stage = nullptr;  // pinned [kRing]
  hipEvent_t stage_ev[kRing] = {};
  bool stage_used[kRing] = {};
  int stage_next = 0;
  uint32_t* pair_e = nullptr;
  uint32_t* pair_r = nullptr;
  uint8_t* pair_n = nullptr;
  int64_t pcap = 0;
  int stride_cap = 0;
  unsigned long long* seg = nullptr;
  int64_t seg_cap = 0;
  uint32_t* touched = nullptr;
  uint32_t* tile_sums = nullptr;
  uint32_t* fail_list = nullptr;
  uint32_t* new_list = nullptr;
  uint8_t* new_map = nullptr;
  int4* fold_desc = nullptr;
  unsigned long long* part = nullptr;
  uint32_t* srows = nullptr;
  StatGrids last_grids{};  // the grids of the last pass's kernels, by statistics region
  int64_t fold_cap = 0;
  int64_t fold_base[8] = {};
  uint8_t* touched_map = nullptr;
  uint32_t* sel = nullptr;
  uint32_t* pub_cnt = nullptr;
  uint32_t* pub_off = nullptr;
  int scratch_gen = 0;

  // publish output (device)
  float* pub_buf = nullptr;
  int64_t pub_buf_floats = 0;
  // host points of tsdf_process_scan (device copy, grown on demand)
  float* in_buf = nullptr;
  int64_t in_buf_floats = 0;

  // maps
  tsdf_map* map_by_slot[kMaxMaps] = {};
  uint64_t next_map_id = 1;
  std::vector<std::pair<int64_t, int64_t>> free_ranges;  // bucket ranges [start, len)
  int64_t quota_promised = 0;

  // last call
  int64_t last_points = 0;

  // launch knobs
  tsdf_launch_config launch{256, 65536, 4096, 1, 0, 0};
  int cus = 0;           // compute units of the device
  int bin_resident = 0;  // k_bin workgroups of traverse_block threads resident at once on the device

  // timing
  bool timing = false;
  struct TimingSet {
    hipEvent_t ev[6];
    int kind;  // 0 pass, 1 publish, 2 gc
  };
  std::vector<TimingSet> tsets;
  std::deque<int> tpending;
  std::vector<int> tfree;
  double tms[TSDF_NUM_STAGES] = {};
  int64_t tcount[TSDF_NUM_STAGES] = {};

  Dev view() const {
    Dev D;
    D.maps = maps_dev;
    D.mctr = mctr;
    D.ctr = ctr;
    D.keys = keys;
    D.vals = vals;
    D.blk_cnt = blk_cnt;
    D.blk_base = blk_base;
    D.next = next;
    D.ovf_stack = ovf_stack;
    D.heap = heap;
    D.heap_free = heap_free;
    D.freed = freed;
    D.hdr = &desc_dev->hdr;
    D.scans = desc_dev->scans;
    D.pair_e = pair_e;
    D.pair_r = pair_r;
    D.pair_n = pair_n;
    D.pcap = pcap;
    D.touched = touched;
    D.tile_sums = tile_sums;
    D.fail_list = fail_list;
    D.new_list = new_list;
    D.new_map = new_map;
    D.fold_desc = fold_desc;
    D.part = part;
    D.srows = srows;
    D.fold_cap = fold_cap;
    for (int c = 0; c < 8; ++c) D.fold_base[c] = fold_base[c];
    D.touched_map = touched_map;
    D.seg = seg;
    D.seg_cap = seg_cap;
    D.sel = sel;
    D.pub_cnt = pub_cnt;
    return D;
  }
};

struct tsdf_batch_graph {
  tsdf_engine* eng = nullptr;
  StatGrids grids{};  // the captured pass's grids
  // the graph's maps by slot and id: a launch checks that each map it uses
  // still exists (a destroyed map's slot is either empty or holds a map with
  // another id), and never touches a tsdf_map that may have been freed
  std::vector<int> slots;
  std::vector<uint64_t> ids;
  int64_t max_points = 0;
  int max_scans = 0;
  int stride = 0;
  int gen = -1;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

namespace {

int set_device(tsdf_engine* E) {
  HIPCHK(hipSetDevice(E->device));
  return TSDF_OK;
}

// 1/sc when k / sc == fma(fma(-q0, sc, k), 1/sc, q0), q0 = k * (1/sc), for
// every integer k a valid point can produce (|k| <= 2^17, with margin),
// else 0: the map then divides.
float product_form_divisor(float sc) {
  const float isc = 1.0f / sc;
  if (!std::isfinite(isc) || isc == 0.0f) return 0.0f;
  for (int k = -(1 << 18); k <= (1 << 18); ++k) {
    const float x = (float)k;
    const float q0 = x * isc;
    const float r = std::fmaf(-q0, sc, x);
    if (std::fmaf(r, isc, q0) != x / sc) return 0.0f;
  }
  return isc;
}

void derive(const tsdf_params& p, MapDev& m) {
  // initGlobalVars, H:1289-1317
  m.vs = p.voxel_size;
  m.T = p.truncation_distance;
  m.maxW = p.max_weight;
  m.pubD2 = p.publish_distance_squared;
  m.gcD2 = p.garbage_collect_distance_squared;
  m.half_vs = p.voxel_size / 2;
  m.BL = p.voxel_size * kVoxelsPerSide;
  m.half_BL = m.BL / 2;
  m.sc_v = 1 / m.vs;
  m.sc_b = 1 / m.BL;
  m.isc_v = product_form_divisor(m.sc_v);
  m.isc_b = product_form_divisor(m.sc_b);
  m.max_vox = (int)(std::ceil(p.truncation_distance * 2 / p.voxel_size) * 4);
  m.max_blk = (int)(std::ceil(p.truncation_distance * 2 / m.BL) * 4);
}

int alloc_buckets(tsdf_engine* E, int64_t n, int64_t* base) {
  for (size_t i = 0; i < E->free_ranges.size(); ++i) {
    auto& r = E->free_ranges[i];
    if (r.second >= n) {
      *base = r.first;
      r.first += n;
      r.second -= n;
      if (r.second == 0) E->free_ranges.erase(E->free_ranges.begin() + i);
      return TSDF_OK;
    }
  }
  return fail(TSDF_ERR_CAPACITY, "hash bucket pool exhausted: %lld buckets requested", (long long)n);
}

void release_buckets(tsdf_engine* E, int64_t base, int64_t n) {
  E->free_ranges.push_back({base, n});
  std::sort(E->free_ranges.begin(), E->free_ranges.end());
  std::vector<std::pair<int64_t, int64_t>> merged;
  for (auto& r : E->free_ranges) {
    if (!merged.empty() && merged.back().first + merged.back().second == r.first) merged.back().second += r.second;
    else merged.push_back(r);
  }
  E->free_ranges.swap(merged);
}

int ensure_scratch(tsdf_engine* E, int64_t P, int stride) {
  if (P <= E->pcap && stride <= E->stride_cap) return TSDF_OK;
  HIPCHK(hipStreamSynchronize(E->stream));
  int64_t np = std::max<int64_t>(P, E->pcap);
  np = std::max<int64_t>(np, 1024);
  int ns = std::max(stride, E->stride_cap);
  (void)hipFree(E->pair_e);
  (void)hipFree(E->pair_r);
  (void)hipFree(E->pair_n);
  (void)hipFree(E->seg);
  E->pair_e = nullptr;
  E->pair_r = nullptr;
  E->pair_n = nullptr;
  E->seg = nullptr;
  E->pcap = 0;
  E->stride_cap = 0;
  CHK(dev_alloc(&E->pair_e, np * ns));
  CHK(dev_alloc(&E->pair_r, np * ns));
  CHK(dev_alloc(&E->pair_n, np));
  CHK(dev_alloc(&E->seg, np * ns));
  (void)hipFree(E->fold_desc);
  E->fold_desc = nullptr;
  // one region per size class, each as large as that class can get
  E->fold_cap = std::min<int64_t>(E->entries, np * ns);
  int64_t fold_total = 0;
  for (int c = 0; c < FC_COUNT; ++c) {
    E->fold_base[c] = fold_total;
    fold_total += fold_class_cap(c, E->fold_cap, np * ns);
  }
  CHK(dev_alloc(&E->fold_desc, fold_total));
  E->pcap = np;
  E->stride_cap = ns;
  E->seg_cap = np * ns;
  E->scratch_gen++;
  return TSDF_OK;
}

// ---- timing ---------------------------------------------------------------------
int timing_begin(tsdf_engine* E, int kind, int* id) {
  *id = -1;
  if (!E->timing) return TSDF_OK;
  if (E->tfree.empty()) {
    // drain the oldest pending set
    HIPCHK(hipStreamSynchronize(E->stream));
    while (!E->tpending.empty()) {
      const int s = E->tpending.front();
      E->tpending.pop_front();
      auto& T = E->tsets[s];
      float ms;
      if (T.kind == 0) {
        for (int k = 0; k < 5; ++k) {
          HIPCHK(hipEventElapsedTime(&ms, T.ev[k], T.ev[k + 1]));
          E->tms[k] += ms;
          E->tcount[k]++;
        }
        HIPCHK(hipEventElapsedTime(&ms, T.ev[0], T.ev[5]));
        E->tms[TSDF_STAGE_TOTAL] += ms;
        E->tcount[TSDF_STAGE_TOTAL]++;
      } else {
        HIPCHK(hipEventElapsedTime(&ms, T.ev[0], T.ev[1]));
        const int st = T.kind == 1 ? TSDF_STAGE_PUBLISH : TSDF_STAGE_GC;
        E->tms[st] += ms;
        E->tcount[st]++;
      }
      E->tfree.push_back(s);
    }
  }
  *id = E->tfree.back();
  E->tfree.pop_back();
  E->tsets[*id].kind = kind;
  return TSDF_OK;
}

int timing_mark(tsdf_engine* E, int id, int k) {
  if (id < 0) return TSDF_OK;
  HIPCHK(hipEventRecord(E->tsets[id].ev[k], E->stream));
  return TSDF_OK;
}

void timing_end(tsdf_engine* E, int id) {
  if (id >= 0) E->tpending.push_back(id);
}

// ---- pass -------------------------------------------------------------------------
// Stages descriptors through a pinned ring slot and copies them to the
// device descriptor buffer (stream ordered).
int stage_descs(tsdf_engine* E, const PassHdr& hdr, const ScanDev* scans, int nscans) {
  const int r = E->stage_next;
  E->stage_next = (r + 1) % kRing;
  if (E->stage_used[r]) HIPCHK(hipEventSynchronize(E->stage_ev[r]));
  DescSlot& S = E->stage[r];
  S.hdr = hdr;
  std::memcpy(S.scans, scans, sizeof(ScanDev) * nscans);
  const size_t bytes = offsetof(DescSlot, scans) + sizeof(ScanDev) * std::max(nscans, 1);
  HIPCHK(hipMemcpyAsync(E->desc_dev, &S, bytes, hipMemcpyHostToDevice, E->stream));
  HIPCHK(hipEventRecord(E->stage_ev[r], E->stream));
  E->stage_used[r] = true;
  return TSDF_OK;
}

// Enqueues the pass kernels sized for P points (grid sizes only; the
// kernels read the actual counts from device memory) over the pass scratch
// of D, whose pair rows and records have `stride` slots per point.  inl: the
// pass's descriptors as kernel arguments, or nscans = -1 when they were
// staged.
// gw: windows of voxel sums k_fold_a keeps in LDS (groups of one block
// folded per turn over its records): the most scans one map has in the pass,
// at most kFoldAWindows.
constexpr int kFoldAWindows = 4;
int enqueue_pass(tsdf_engine* E, const Dev& D, int64_t P, int64_t stride, int tid, const PassInline& inl, int gw) {
  hipStream_t st = E->stream;
  const tsdf_launch_config& L = E->launch;
  CHK(timing_mark(E, tid, 0));
  const int tb = L.traverse_block;
  BinSched bs;
  bs.rays_per_lane = std::max(1, L.rays_per_lane);
  bs.queue = L.traverse_queue;
  // queue: the resident workgroups (each keeps taking chunks); else one
  // workgroup per chunk up to the cap
  const int bg = grid_for(P, (int64_t)tb * bs.rays_per_lane, bs.queue ? E->bin_resident : L.traverse_grid_cap);
  if (bs.rays_per_lane > 1) hipLaunchKernelGGL(k_bin<true>, dim3(bg), dim3(tb), 0, st, D, inl, bs);
  else hipLaunchKernelGGL(k_bin<false>, dim3(bg), dim3(tb), 0, st, D, inl, bs);
  CHK(timing_mark(E, tid, 1));
  // (heap assignment runs in k_seg, inside the scan stage: the assign stage is empty)
  CHK(timing_mark(E, tid, 2));
  const int64_t tmax = std::min<int64_t>(E->entries, P * stride);
  // tiles of touched blocks are taken grid-stride by at most 256 workgroups
  // (the touched count is only known on the device; most passes have tens)
  const int sg = grid_for(tmax, kScanTile, 256);
  hipLaunchKernelGGL(k_seg, dim3(sg), dim3(256), 0, st, D);
  CHK(timing_mark(E, tid, 3));
  if (P > kEmitStagePoints) hipLaunchKernelGGL(k_emit<true>, dim3(grid_for(P, 256, 65536)), dim3(256), 0, st, D);
  else hipLaunchKernelGGL(k_emit<false>, dim3(grid_for(P, 256, 65536)), dim3(256), 0, st, D);
  CHK(timing_mark(E, tid, 4));
  const int fg = std::max(1, L.fold_grid);
  constexpr int kFoldBGrid = 512;
  with_state(E->state_format, [&](auto s) {
    using State = decltype(s);
    gw = std::min(std::max(gw, 1), kFoldAWindows);
    hipLaunchKernelGGL(k_fold_a<State>, dim3(fg), dim3(256), (size_t)fold_a_lds_bytes(gw), st, D, gw);
    hipLaunchKernelGGL(k_fold_b<State>, dim3(kFoldBGrid), dim3(kFoldBThreads), 0, st, D);
    return 0;
  });
  static_assert(kFoldBGrid <= kStatRows[SR_FOLD_B], "a statistics row per workgroup");
  E->last_grids.n[SR_BIN] = bg;
  E->last_grids.n[SR_SCAN] = sg;
  E->last_grids.n[SR_FOLD_A] = fg;
  E->last_grids.n[SR_FOLD_B] = kFoldBGrid;
  CHK(timing_mark(E, tid, 5));
  HIPCHK(hipGetLastError());
  return TSDF_OK;
}

// k_bin workgroups of `block` threads resident at once: per-CU occupancy
// times the CU count (the persistent grid of the queue mode).
int bin_residency(tsdf_engine* E, int block, int& out) {
  int per_cu = 0;
  // (both instances have the same register and LDS budget)
  HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(&k_bin<true>), block, 0));
  if (per_cu < 1) return fail(TSDF_ERR_HIP, "k_bin: no workgroup of %d threads fits a CU", block);
  out = per_cu * E->cus;
  return TSDF_OK;
}

int check_map(tsdf_map* m) {
  if (!m || !m->eng || m->slot < 0) return fail(TSDF_ERR_INVALID_ARG, "invalid map handle");
  return TSDF_OK;
}

int run_batch(tsdf_engine* E, tsdf_map* const* maps, int nmaps, const tsdf_scan_desc* scans, int nscans,
              const float* points_dev) {
  if (nscans < 0 || nscans > kMaxScansPerBatch)
    return fail(TSDF_ERR_CAPACITY, "a batch holds at most %d scans (got %d)", kMaxScansPerBatch, nscans);
  if (nscans > 0 && (!scans || !maps)) return fail(TSDF_ERR_INVALID_ARG, "null scans or maps");
  ScanDev sd[kMaxScansPerBatch];
  int64_t P = 0;
  int stride = 1;
  for (int i = 0; i < nscans; ++i) {
    const tsdf_scan_desc& s = scans[i];
    if (s.map_index < 0 || s.map_index >= nmaps) return fail(TSDF_ERR_INVALID_ARG, "scan %d: map_index %d out of range", i, s.map_index);
    tsdf_map* m = maps[s.map_index];
    CHK(check_map(m));
    if (m->eng != E) return fail(TSDF_ERR_INVALID_ARG, "scan %d: map belongs to another engine", i);
    if (s.n_points < 0 || s.point_offset < 0) return fail(TSDF_ERR_INVALID_ARG, "scan %d: negative size/offset", i);
    if (s.n_points > 0 && !points_dev) return fail(TSDF_ERR_INVALID_ARG, "null point buffer");
    for (int k = 0; k < 3; ++k)
      if (!std::isfinite(s.sensor[k])) return fail(TSDF_ERR_INVALID_ARG, "scan %d: non-finite sensor position", i);
    sd[i].map = m->slot;
    sd[i].n = s.n_points;
    sd[i].off = (int32_t)P;
    sd[i].s[0] = s.sensor[0];
    sd[i].s[1] = s.sensor[1];
    sd[i].s[2] = s.sensor[2];
    P += s.n_points;
    stride = std::max(stride, m->desc.max_vox);
  }
  if (P > kMaxPassPoints) return fail(TSDF_ERR_CAPACITY, "a pass holds at most %d points (got %lld)", kMaxPassPoints, (long long)P);
  if (P > 0 && reinterpret_cast<uintptr_t>(points_dev) % alignof(float4) != 0)
    return fail(TSDF_ERR_INVALID_ARG, "the point buffer must be 16-byte aligned (float4 rows)");
  CHK(set_device(E));
  CHK(ensure_scratch(E, std::max<int64_t>(P, 1), stride));
  // per-scan source offsets are folded into the point pointer: scans are
  // re-based so that pass point i of scan k reads points[point_offset_k + i - off_k]
  // by giving each scan its own base through a contiguous copy when offsets
  // are not already contiguous.
  bool contiguous = true;
  int64_t base0 = nscans > 0 ? scans[0].point_offset : 0;
  for (int i = 0; i < nscans; ++i)
    if (scans[i].point_offset != base0 + sd[i].off) contiguous = false;
  const float4* pts = reinterpret_cast<const float4*>(points_dev);
  if (!contiguous) {
    return fail(TSDF_ERR_INVALID_ARG, "scans of a batch must be contiguous and in order in the point buffer");
  }
  PassInline inl{};
  inl.hdr.points = pts + base0;
  inl.hdr.total = (int32_t)P;
  inl.hdr.nscans = nscans;
  inl.hdr.stride = stride;
  if (nscans <= kInlineScans) {
    inl.nscans = nscans;
    std::memcpy(inl.scans, sd, sizeof(ScanDev) * nscans);
  } else {
    inl.nscans = -1;
    CHK(stage_descs(E, inl.hdr, sd, nscans));
  }
  int tid;
  CHK(timing_begin(E, 0, &tid));
  // the most scans of one map in the pass
  int gw = 1;
  for (int i = 0; i < nscans; ++i) {
    int same = 0;
    for (int j = 0; j < nscans; ++j) same += sd[j].map == sd[i].map ? 1 : 0;
    gw = std::max(gw, same);
  }
  CHK(enqueue_pass(E, E->view(), P, E->stride_cap, tid, inl, gw));
  timing_end(E, tid);
  E->last_points = P;
  return TSDF_OK;
}

int read_ctr(tsdf_engine* E) {
  HIPCHK(hipMemcpyAsync(E->ctr_host, E->ctr, sizeof(EngCtr), hipMemcpyDeviceToHost, E->stream));
  HIPCHK(hipStreamSynchronize(E->stream));
  return TSDF_OK;
}

__global__ void k_begin_select(EngCtr* c, int stat_reset) {
  c->sel_n = 0;
  c->freed_n = 0;
  if (stat_reset >= 0) c->stats[stat_reset] = 0;
}

__global__ void k_store_total(const int32_t* n_ptr, const uint32_t* off, EngCtr* c, int stat) {
  c->stats[stat] = off[*n_ptr];
}

template <class State>
int publish_impl(tsdf_map* m, const float sensor[3], float* pos, float* data, int64_t cap, int64_t* count, int layout) {
  tsdf_engine* E = m->eng;
  CHK(set_device(E));
  for (int k = 0; k < 3; ++k)
    if (!std::isfinite(sensor[k])) return fail(TSDF_ERR_INVALID_ARG, "non-finite sensor position");
  Dev D = E->view();
  hipStream_t st = E->stream;
  int tid;
  CHK(timing_begin(E, 1, &tid));
  CHK(timing_mark(E, tid, 0));
  hipLaunchKernelGGL(k_begin_select, dim3(1), dim3(1), 0, st, E->ctr, (int)ST_PUBLISHED);
  const int64_t ne = (int64_t)(m->desc.n_home + m->desc.n_ovf) * kBucket;
  hipLaunchKernelGGL(k_select, dim3(grid_for(ne, kSelChunk, 4096)), dim3(256), 0, st, D, m->slot, sensor[0], sensor[1], sensor[2], 1);
  const int gq = grid_for(ne, 4, 2048);
  hipLaunchKernelGGL(k_pub_count<State>, dim3(gq), dim3(256), 0, st, D);
  CountPlain cp{E->pub_cnt};
  const int sg = grid_for(ne, kScanTile, 1024);
  hipLaunchKernelGGL(k_scan_tiles<CountPlain>, dim3(sg), dim3(256), 0, st, &E->ctr->sel_n, cp, E->tile_sums);
  hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(256), 0, st, &E->ctr->sel_n, E->tile_sums);
  hipLaunchKernelGGL(k_scan_apply<CountPlain>, dim3(sg), dim3(256), 0, st, &E->ctr->sel_n, cp, E->tile_sums, E->pub_off);
  hipLaunchKernelGGL(k_store_total, dim3(1), dim3(1), 0, st, &E->ctr->sel_n, E->pub_off, E->ctr, (int)ST_PUBLISHED);
  HIPCHK(hipGetLastError());
  CHK(read_ctr(E));
  const int64_t total = (int64_t)E->ctr_host->stats[ST_PUBLISHED];
  constexpr int fpv = 5;  // floats per voxel in either layout
  if (total * fpv > E->pub_buf_floats) {
    hipFree(E->pub_buf);
    E->pub_buf = nullptr;
    E->pub_buf_floats = 0;
    const int64_t want = std::max<int64_t>(total * fpv * 5 / 4, 1 << 20);
    CHK(dev_alloc(&E->pub_buf, want));
    E->pub_buf_floats = want;
  }
  float* dpos = E->pub_buf;
  float* ddata = E->pub_buf + total * 3;
  hipLaunchKernelGGL(k_pub_write<State>, dim3(gq), dim3(256), 0, st, D, m->slot, E->pub_off, dpos, ddata, total, layout);
  HIPCHK(hipGetLastError());
  CHK(timing_mark(E, tid, 1));
  timing_end(E, tid);
  const int64_t n = std::min(total, std::max<int64_t>(cap, 0));
  if (n > 0) {
    if (layout == 0) {
      if (pos) HIPCHK(hipMemcpyAsync(pos, dpos, sizeof(float) * 3 * n, hipMemcpyDeviceToHost, st));
      if (data) HIPCHK(hipMemcpyAsync(data, ddata, sizeof(float) * 2 * n, hipMemcpyDeviceToHost, st));
    } else {
      if (pos) HIPCHK(hipMemcpyAsync(pos, dpos, sizeof(float) * 5 * n, hipMemcpyDeviceToHost, st));
    }
  }
  HIPCHK(hipStreamSynchronize(st));
  if (count) *count = total;
  return TSDF_OK;
}

template <class State>
int gc_impl(tsdf_map* m, const float sensor[3], int mode, int64_t* removed) {
  tsdf_engine* E = m->eng;
  CHK(set_device(E));
  for (int k = 0; k < 3; ++k)
    if (!std::isfinite(sensor[k])) return fail(TSDF_ERR_INVALID_ARG, "non-finite sensor position");
  Dev D = E->view();
  hipStream_t st = E->stream;
  int tid;
  CHK(timing_begin(E, 2, &tid));
  CHK(timing_mark(E, tid, 0));
  hipLaunchKernelGGL(k_begin_select, dim3(1), dim3(1), 0, st, E->ctr, (int)ST_GC_REMOVED);
  hipLaunchKernelGGL(k_stats_zero, dim3(4), dim3(256), 0, st, D, 1u << ST_GC_REMOVED);
  const int64_t ne = (int64_t)(m->desc.n_home + m->desc.n_ovf) * kBucket;
  hipLaunchKernelGGL(k_gc_mark, dim3(grid_for(ne, kSelChunk, 4096)), dim3(256), 0, st, D, m->slot, sensor[0], sensor[1],
                     sensor[2], mode);
  hipLaunchKernelGGL(k_gc_zero<State>, dim3(1024), dim3(256), 0, st, D);
  hipLaunchKernelGGL(k_gc_chains, dim3(grid_for(m->desc.n_home, 256, 4096)), dim3(256), 0, st, D, m->slot);
  hipLaunchKernelGGL(k_stats_reduce, dim3(ST_COUNT), dim3(256), 0, st, D, 1u << ST_GC_REMOVED);
  HIPCHK(hipGetLastError());
  CHK(timing_mark(E, tid, 1));
  timing_end(E, tid);
  if (removed) {
    CHK(read_ctr(E));
    *removed = (int64_t)E->ctr_host->stats[ST_GC_REMOVED];
  }
  return TSDF_OK;
}

int init_map_tables(tsdf_engine* E, const MapDev& d) {
  hipStream_t st = E->stream;
  const int64_t nb = d.n_home + d.n_ovf;
  const int64_t e0 = (int64_t)d.bucket_base * kBucket;
  hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(nb * kBucket, 256, 4096)), dim3(256), 0, st, E->keys + e0, nb * kBucket, kEmptyKey);
  hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(nb * kBucket, 256, 4096)), dim3(256), 0, st, E->vals + e0, nb * kBucket, -1);
  hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(nb, 256, 4096)), dim3(256), 0, st, E->next + d.bucket_base, nb, -1);
  if (d.n_ovf > 0)
    hipLaunchKernelGGL(k_iota_desc, dim3(grid_for(d.n_ovf, 256, 4096)), dim3(256), 0, st,
                       E->ovf_stack + d.bucket_base + d.n_home, (int64_t)d.n_ovf, d.bucket_base + d.n_home);
  MapCtr c{d.n_ovf, 0, 0, 0};
  HIPCHK(hipMemcpyAsync(E->mctr + d.slot, &c, sizeof(MapCtr), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(E->maps_dev + d.slot, &d, sizeof(MapDev), hipMemcpyHostToDevice, st));
  HIPCHK(hipStreamSynchronize(st));
  HIPCHK(hipGetLastError());
  return TSDF_OK;
}

int validate_params(const tsdf_params* p) {
  if (!p) return fail(TSDF_ERR_INVALID_ARG, "null params");
  const float v[5] = {p->voxel_size, p->truncation_distance, p->max_weight, p->publish_distance_squared,
                      p->garbage_collect_distance_squared};
  for (float x : v)
    if (!std::isfinite(x)) return fail(TSDF_ERR_INVALID_ARG, "non-finite parameter");
  if (!(p->voxel_size > 0) || !(p->truncation_distance > 0) || !(p->max_weight > 0))
    return fail(TSDF_ERR_INVALID_ARG, "voxel_size, truncation_distance and max_weight must be positive");
  MapDev m{};
  derive(*p, m);
  if (m.max_vox > kMaxStride || m.max_vox > 255)
    return fail(TSDF_ERR_INVALID_ARG, "truncation_distance / voxel_size too large: %d voxels per point (max %d)", m.max_vox,
                kMaxStride);
  return TSDF_OK;
}

int map_create_impl(tsdf_engine* E, const tsdf_params* params, const tsdf_map_config* cfg, int64_t quota_promise,
                    tsdf_map** out) {
  if (!E || !out) return fail(TSDF_ERR_INVALID_ARG, "null engine or output");
  CHK(validate_params(params));
  CHK(set_device(E));
  tsdf_map_config c{16, 0, 64, 0, 0};
  if (cfg) {
    if (cfg->buckets_log2) c.buckets_log2 = cfg->buckets_log2;
    if (cfg->overflow_buckets) c.overflow_buckets = cfg->overflow_buckets;
    if (cfg->max_chain) c.max_chain = cfg->max_chain;
    c.block_quota = cfg->block_quota;
  }
  if (c.buckets_log2 < 1 || c.buckets_log2 > 28) return fail(TSDF_ERR_INVALID_ARG, "buckets_log2 out of range");
  const int64_t n_home = 1ll << c.buckets_log2;
  const int64_t n_ovf = c.overflow_buckets > 0 ? c.overflow_buckets : std::max<int64_t>(1, n_home / 4);
  if (c.max_chain < 0 || c.block_quota < 0 || n_ovf < 0) return fail(TSDF_ERR_INVALID_ARG, "negative map config");
  int slot = -1;
  for (int i = 0; i < kMaxMaps; ++i)
    if (!E->map_by_slot[i]) { slot = i; break; }
  if (slot < 0) return fail(TSDF_ERR_CAPACITY, "at most %d maps per engine", kMaxMaps);
  int64_t base;
  CHK(alloc_buckets(E, n_home + n_ovf, &base));
  tsdf_map* m = new tsdf_map();
  m->eng = E;
  m->slot = slot;
  m->params = *params;
  derive(*params, m->desc);
  m->desc.home_mask = (uint32_t)(n_home - 1);
  m->desc.bucket_base = (int32_t)base;
  m->desc.n_home = (int32_t)n_home;
  m->desc.n_ovf = (int32_t)n_ovf;
  m->desc.max_chain = c.max_chain;
  m->desc.slot = slot;
  m->desc.block_quota = c.block_quota;
  int r = init_map_tables(E, m->desc);
  if (r != TSDF_OK) {
    release_buckets(E, base, n_home + n_ovf);
    delete m;
    return r;
  }
  E->map_by_slot[slot] = m;
  m->id = E->next_map_id++;
  m->quota_promised = quota_promise;
  E->quota_promised += quota_promise;
  *out = m;
  return TSDF_OK;
}

// ---- serving -----------------------------------------------------------------------
using Clock = std::chrono::steady_clock;

}  // namespace

// =====================================================================================
extern "C" {

const char* tsdf_last_error(void) { return g_last_error.c_str(); }

int tsdf_derived_params(const tsdf_params* params, float* out_f, int32_t* out_i) {
  if (!params || !out_f || !out_i) return fail(TSDF_ERR_INVALID_ARG, "null argument");
  MapDev m{};
  derive(*params, m);
  out_f[0] = m.half_vs;
  out_f[1] = m.BL;
  out_f[2] = m.half_BL;
  out_f[3] = m.BL / 4;
  out_f[4] = m.vs / 4;
  out_i[0] = m.max_vox;
  out_i[1] = m.max_blk;
  return TSDF_OK;
}

int tsdf_engine_create(const tsdf_engine_config* cfg, tsdf_engine** out) {
  if (!out) return fail(TSDF_ERR_INVALID_ARG, "null output");
  *out = nullptr;
  tsdf_engine_config c{0, TSDF_STATE_F32, 1 << 18, 1 << 20, 1 << 17};
  if (cfg) {
    c.device = cfg->device;
    c.state_format = cfg->state_format;
    if (cfg->heap_blocks > 0) c.heap_blocks = cfg->heap_blocks;
    if (cfg->total_buckets > 0) c.total_buckets = cfg->total_buckets;
    if (cfg->max_batch_points > 0) c.max_batch_points = cfg->max_batch_points;
  }
  if (c.state_format != TSDF_STATE_F32 && c.state_format != TSDF_STATE_F16 && c.state_format != TSDF_STATE_F16_QW)
    return fail(TSDF_ERR_INVALID_ARG, "unknown state format %d", c.state_format);
  if (c.heap_blocks > (1ll << 30)) return fail(TSDF_ERR_INVALID_ARG, "heap_blocks too large");
  if (c.total_buckets * kBucket > (1ll << 31) - 1) return fail(TSDF_ERR_INVALID_ARG, "total_buckets too large");
  int ndev = 0;
  HIPCHK(hipGetDeviceCount(&ndev));
  if (c.device < 0 || c.device >= ndev) return fail(TSDF_ERR_INVALID_ARG, "device %d not present (%d devices)", c.device, ndev);
  tsdf_engine* E = new tsdf_engine();
  E->device = c.device;
  E->state_format = c.state_format;
  E->heap_blocks = c.heap_blocks;
  E->total_buckets = c.total_buckets;
  E->entries = c.total_buckets * kBucket;
  auto cleanup = [&](int r) {
    tsdf_engine_destroy(E);
    return r;
  };
  int r0;
  if (hipSetDevice(E->device) != hipSuccess) return cleanup(fail(TSDF_ERR_HIP, "hipSetDevice failed"));
  resolve_scratch(E->device);  // for tsdf_integrate_noalloc; a failure is reported by that call
  if (hipDeviceGetAttribute(&E->cus, hipDeviceAttributeMultiprocessorCount, E->device) != hipSuccess || E->cus < 1)
    return cleanup(fail(TSDF_ERR_HIP, "the compute-unit count of device %d is unavailable", E->device));
  if ((r0 = bin_residency(E, E->launch.traverse_block, E->bin_resident)) != TSDF_OK) return cleanup(r0);
  if (hipStreamCreateWithFlags(&E->stream, hipStreamNonBlocking) != hipSuccess)
    return cleanup(fail(TSDF_ERR_HIP, "hipStreamCreate failed"));
  int r;
  const int64_t entries = E->entries;
  if ((r = dev_alloc(&E->maps_dev, kMaxMaps)) || (r = dev_alloc(&E->mctr, kMaxMaps)) || (r = dev_alloc(&E->ctr, 1)) ||
      (r = dev_alloc(&E->keys, entries)) || (r = dev_alloc(&E->vals, entries)) || (r = dev_alloc(&E->blk_cnt, entries)) ||
      (r = dev_alloc(&E->blk_base, entries)) || (r = dev_alloc(&E->next, c.total_buckets)) ||
      (r = dev_alloc(&E->ovf_stack, c.total_buckets)) ||
      (r = dev_alloc(&E->heap_free, c.heap_blocks)) || (r = dev_alloc(&E->freed, c.heap_blocks)) ||
      (r = dev_alloc(&E->desc_dev, 1)) ||
      (r = dev_alloc(&E->touched, entries)) ||
      (r = dev_alloc(&E->tile_sums, entries / kScanTile + 2)) ||
      (r = dev_alloc(&E->fail_list, entries)) ||
      (r = dev_alloc(&E->new_list, entries)) || (r = dev_alloc(&E->new_map, entries)) ||
      (r = dev_alloc(&E->touched_map, entries)) || (r = dev_alloc(&E->part, ST_COUNT * kPartSlots)) || (r = dev_alloc(&E->srows, (int64_t)kStatRowsTotal * kStatCols)) ||
      (r = dev_alloc(&E->sel, entries)) || (r = dev_alloc(&E->pub_cnt, entries + 1)) ||
      (r = dev_alloc(&E->pub_off, entries + 1)))
    return cleanup(r);
  const int64_t vbytes = (int64_t)c.heap_blocks * kVoxelsPerBlock * (c.state_format == TSDF_STATE_F32 ? 8 : 4);
  {
    unsigned char* h = nullptr;
    if ((r = dev_alloc(&h, vbytes))) return cleanup(r);
    E->heap = h;
  }
  if (hipHostMalloc((void**)&E->ctr_host, sizeof(EngCtr), hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc((void**)&E->stage, sizeof(DescSlot) * kRing, hipHostMallocDefault) != hipSuccess)
    return cleanup(fail(TSDF_ERR_NO_MEMORY, "pinned host allocation failed"));
  for (int i = 0; i < kRing; ++i)
    if (hipEventCreateWithFlags(&E->stage_ev[i], hipEventDisableTiming) != hipSuccess)
      return cleanup(fail(TSDF_ERR_HIP, "hipEventCreate failed"));
  E->tsets.resize(kTimingRing);
  for (int i = 0; i < kTimingRing; ++i) {
    for (int k = 0; k < 6; ++k)
      if (hipEventCreate(&E->tsets[i].ev[k]) != hipSuccess) return cleanup(fail(TSDF_ERR_HIP, "hipEventCreate failed"));
    E->tfree.push_back(i);
  }
  hipStream_t st = E->stream;
  // zero counters and voxel state; keys empty; heap free stack 0..H-1
  if (hipMemsetAsync(E->ctr, 0, sizeof(EngCtr), st) != hipSuccess ||
      hipMemsetAsync(E->mctr, 0, sizeof(MapCtr) * kMaxMaps, st) != hipSuccess ||
      hipMemsetAsync(E->heap, 0, (size_t)vbytes, st) != hipSuccess ||
      hipMemsetAsync(E->blk_cnt, 0, sizeof(unsigned long long) * entries, st) != hipSuccess ||
      hipMemsetAsync(E->part, 0, sizeof(unsigned long long) * ST_COUNT * kPartSlots, st) != hipSuccess ||
      hipMemsetAsync(E->srows, 0, sizeof(uint32_t) * kStatRowsTotal * kStatCols, st) != hipSuccess ||
      hipMemsetAsync(E->keys, 0xFF, sizeof(unsigned long long) * entries, st) != hipSuccess ||
      hipMemsetAsync(E->vals, 0xFF, sizeof(int32_t) * entries, st) != hipSuccess ||
      hipMemsetAsync(E->next, 0xFF, sizeof(int32_t) * c.total_buckets, st) != hipSuccess)
    return cleanup(fail(TSDF_ERR_HIP, "device initialisation failed"));
  hipLaunchKernelGGL(k_iota_desc, dim3(grid_for(c.heap_blocks, 256, 4096)), dim3(256), 0, st, E->heap_free,
                     (int64_t)c.heap_blocks, 0);
  const int32_t top = (int32_t)c.heap_blocks;
  if (hipMemcpyAsync(&E->ctr->heap_top, &top, sizeof(int32_t), hipMemcpyHostToDevice, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return cleanup(fail(TSDF_ERR_HIP, "device initialisation failed"));
  E->free_ranges.push_back({0, c.total_buckets});
  r = ensure_scratch(E, c.max_batch_points, 32);
  if (r) return cleanup(r);
  *out = E;
  return TSDF_OK;
}

int tsdf_engine_destroy(tsdf_engine* E) {
  if (!E) return TSDF_OK;
  hipSetDevice(E->device);
  if (E->stream) hipStreamSynchronize(E->stream);
  for (int i = 0; i < kMaxMaps; ++i)
    if (E->map_by_slot[i]) {
      delete E->map_by_slot[i];
      E->map_by_slot[i] = nullptr;
    }
  void* ptrs[] = {E->maps_dev, E->mctr, E->ctr, E->keys, E->vals, E->blk_cnt, E->blk_base, E->next, E->ovf_stack,
                  E->heap, E->heap_free, E->freed, E->desc_dev, E->pair_e, E->pair_r, E->pair_n,
                  E->seg, E->touched, E->tile_sums, E->fail_list, E->new_list, E->new_map, E->fold_desc, E->part, E->srows,
                  E->touched_map, E->sel,
                  E->pub_cnt, E->pub_off, E->pub_buf, E->in_buf};
  for (void* p : ptrs)
    if (p) hipFree(p);
  if (E->ctr_host) hipHostFree(E->ctr_host);
  if (E->stage) hipHostFree(E->stage);
  for (int i = 0; i < kRing; ++i)
    if (E->stage_ev[i]) hipEventDestroy(E->stage_ev[i]);
  for (auto& T : E->tsets)
    for (int k = 0; k < 6; ++k)
      if (T.ev[k]) hipEventDestroy(T.ev[k]);
  if (E->stream) hipStreamDestroy(E->stream);
  delete E;
  return TSDF_OK;
}

int tsdf_engine_synchronize(tsdf_engine* E) {
  if (!E) return fail(TSDF_ERR_INVALID_ARG, "null engine");
  CHK(set_device(E));
  HIPCHK(hipStreamSynchronize(E->stream));
  return TSDF_OK;
}

void* tsdf_engine_stream(tsdf_engine* E) { return E ? (void*)E->stream : nullptr; }

int tsdf_engine_free_blocks(tsdf_engine* E, int64_t* out) {
  if (!E || !out) return fail(TSDF_ERR_INVALID_ARG, "null argument");
  CHK(set_device(E));
  CHK(read_ctr(E));
  *out = E->ctr_host->heap_top;
  return TSDF_OK;
}

int tsdf_map_create(tsdf_engine* E, const tsdf_params* params, const tsdf_map_config* cfg, tsdf_map** out) {
  return map_create_impl(E, params, cfg, 0, out);
}

int tsdf_map_admit(tsdf_engine* E, const tsdf_params* params, const tsdf_map_config* cfg, tsdf_map** out) {
  if (!E || !cfg || cfg->block_quota <= 0) return fail(TSDF_ERR_INVALID_ARG, "admission needs a positive block_quota");
  if (E->quota_promised + cfg->block_quota > E->heap_blocks)
    return fail(TSDF_ERR_BUDGET, "heap budget: %lld blocks promised, %lld requested, %lld in the heap",
                (long long)E->quota_promised, (long long)cfg->block_quota, (long long)E->heap_blocks);
  return map_create_impl(E, params, cfg, cfg->block_quota, out);
}

int tsdf_map_destroy(tsdf_map* m) {
  if (!m) return TSDF_OK;
  CHK(check_map(m));
  tsdf_engine* E = m->eng;
  const float z[3] = {0, 0, 0};
  int r = with_state(E->state_format, [&](auto s) { return gc_impl<decltype(s)>(m, z, 1, nullptr); });
  if (r) return r;
  HIPCHK(hipStreamSynchronize(E->stream));
  release_buckets(E, m->desc.bucket_base, (int64_t)m->desc.n_home + m->desc.n_ovf);
  E->quota_promised -= m->quota_promised;
  E->map_by_slot[m->slot] = nullptr;
  delete m;
  return TSDF_OK;
}

int tsdf_map_num_blocks(tsdf_map* m, int64_t* out) {
  CHK(check_map(m));
  if (!out) return fail(TSDF_ERR_INVALID_ARG, "null output");
  tsdf_engine* E = m->eng;
  CHK(set_device(E));
  MapCtr c;
  HIPCHK(hipMemcpyAsync(&c, E->mctr + m->slot, sizeof(MapCtr), hipMemcpyDeviceToHost, E->stream));
  HIPCHK(hipStreamSynchronize(E->stream));
  *out = c.blocks;
  return TSDF_OK;
}

int tsdf_map_export(tsdf_map* m, int32_t* keys, float* voxels, int64_t capacity, int64_t* count) {
  CHK(check_map(m));
  tsdf_engine* E = m->eng;
  CHK(set_device(E));
  Dev D = E->view();
  hipStream_t st = E->stream;
  hipLaunchKernelGGL(k_begin_select, dim3(1), dim3(1), 0, st, E->ctr, -1);
  const int64_t ne = (int64_t)(m->desc.n_home + m->desc.n_ovf) * kBucket;
  hipLaunchKernelGGL(k_select, dim3(grid_for(ne, kSelChunk, 4096)), dim3(256), 0, st, D, m->slot, 0.f, 0.f, 0.f, 0);
  HIPCHK(hipGetLastError());
  CHK(read_ctr(E));
  const int64_t n = E->ctr_host->sel_n;
  if (count) *count = n;
  const int64_t w = std::min(n, std::max<int64_t>(capacity, 0));
  if (w == 0 || (!keys && !voxels)) return TSDF_OK;
  int32_t* dk = nullptr;
  float2* dv = nullptr;
  CHK(dev_alloc(&dk, n * 3));
  int r = dev_alloc(&dv, n * kVoxelsPerBlock);
  if (r) {
    hipFree(dk);
    return r;
  }
  with_state(E->state_format, [&](auto s) {
    hipLaunchKernelGGL(k_export<decltype(s)>, dim3(grid_for(n, 1, 8192)), dim3(256), 0, st, D, m->slot, dk, dv);
    return 0;
  });
  hipError_t e1 = hipGetLastError();
  if (e1 == hipSuccess && keys) e1 = hipMemcpyAsync(keys, dk, sizeof(int32_t) * 3 * w, hipMemcpyDeviceToHost, st);
  if (e1 == hipSuccess && voxels)
    e1 = hipMemcpyAsync(voxels, dv, sizeof(float2) * kVoxelsPerBlock * w, hipMemcpyDeviceToHost, st);
  if (e1 == hipSuccess) e1 = hipStreamSynchronize(st);
  hipFree(dk);
  hipFree(dv);
  if (e1 != hipSuccess) return fail(TSDF_ERR_HIP, "export failed: %s", hipGetErrorString(e1));
  return TSDF_OK;
}

int tsdf_map_import(tsdf_map* m, const int32_t* keys, const float* voxels, int64_t n, int64_t* imported,
                    int64_t* failed) {
  CHK(check_map(m));
  if (n < 0 || (n > 0 && (!keys || !voxels))) return fail(TSDF_ERR_INVALID_ARG, "bad argument");
  if (imported) *imported = 0;
  if (failed) *failed = 0;
  if (n == 0) return TSDF_OK;
  // keys: inside the block range of the point validity rule, and distinct
  constexpr int32_t kMaxBlockCoord = 1 << 13;
  std::vector<uint64_t> packed((size_t)n);
  for (int64_t q = 0; q < n; ++q) {
    for (int k = 0; k < 3; ++k)
      if (keys[3 * q + k] < -kMaxBlockCoord || keys[3 * q + k] > kMaxBlockCoord)
        return fail(TSDF_ERR_INVALID_ARG, "block %lld: coordinate %d out of range", (long long)q, keys[3 * q + k]);
    packed[q] = pack_key(keys[3 * q], keys[3 * q + 1], keys[3 * q + 2]);
  }
  std::sort(packed.begin(), packed.end());
  if (std::adjacent_find(packed.begin(), packed.end()) != packed.end())
    return fail(TSDF_ERR_INVALID_ARG, "duplicate block keys");
  tsdf_engine* E = m->eng;
  CHK(set_device(E));
  hipStream_t st = E->stream;
  int32_t* dk = nullptr;
  float2* dv = nullptr;
  int* dc = nullptr;
  CHK(dev_alloc(&dk, n * 3));
  int r = dev_alloc(&dv, n * kVoxelsPerBlock);
  if (!r) r = dev_alloc(&dc, 2);
  hipError_t e1 = hipSuccess;
  int hc[2] = {0, 0};
  if (!r) {
    e1 = hipMemcpyAsync(dk, keys, sizeof(int32_t) * 3 * n, hipMemcpyHostToDevice, st);
    if (e1 == hipSuccess) e1 = hipMemcpyAsync(dv, voxels, sizeof(float2) * kVoxelsPerBlock * n, hipMemcpyHostToDevice, st);
    if (e1 == hipSuccess) e1 = hipMemsetAsync(dc, 0, sizeof(int) * 2, st);
    if (e1 == hipSuccess) {
      Dev D = E->view();
      with_state(E->state_format, [&](auto s) {
        hipLaunchKernelGGL(k_import<decltype(s)>, dim3(grid_for(n, 1, 8192)), dim3(256), 0, st, D, m->slot, dk, dv, n, dc);
        return 0;
      });
      e1 = hipGetLastError();
    }
    if (e1 == hipSuccess) e1 = hipMemcpyAsync(hc, dc, sizeof(hc), hipMemcpyDeviceToHost, st);
    if (e1 == hipSuccess) e1 = hipStreamSynchronize(st);
  }
  hipFree(dk);
  hipFree(dv);
  hipFree(dc);
  if (r) return r;
  if (e1 != hipSuccess) return fail(TSDF_ERR_HIP, "import failed: %s", hipGetErrorString(e1));
  if (imported) *imported = hc[0];
  if (failed) *failed = hc[1];
  return TSDF_OK;
}

int tsdf_integrate(tsdf_map* m, const float* points_dev, int64_t n, const float sensor[3]) {
  CHK(check_map(m));
  if (!sensor) return fail(TSDF_ERR_INVALID_ARG, "null sensor");
  if (n < 0 || n > kMaxPassPoints) return fail(TSDF_ERR_CAPACITY, "point count %lld out of range", (long long)n);
  tsdf_scan_desc s{};
  s.map_index = 0;
  s.n_points = (int32_t)n;
  s.point_offset = 0;
  s.sensor[0] = sensor[0];
  s.sensor[1] = sensor[1];
  s.sensor[2] = sensor[2];
  tsdf_map* maps[1] = {m};
  return run_batch(m->eng, maps, 1, &s, 1, points_dev);
}

int tsdf_integrate_batch(tsdf_engine* E, tsdf_map* const* maps, int32_t nmaps, const tsdf_scan_desc* scans, int32_t nscans,
                         const float* points_dev) {
  if (!E) return fail(TSDF_ERR_INVALID_ARG, "null engine");
  return run_batch(E, maps, nmaps, scans, nscans, points_dev);
}

int tsdf_integrate_noalloc(tsdf_map* m, const float* points_dev, int64_t n, const float sensor[3]) {
  CHK(check_map(m));
  if (!sensor) return fail(TSDF_ERR_INVALID_ARG, "null sensor");
  if (n < 0 || n > kMaxPassPoints) return fail(TSDF_ERR_CAPACITY, "point count %lld out of range", (long long)n);
  if (n > 0 && !points_dev) return fail(TSDF_ERR_INVALID_ARG, "null point buffer");
  if (n > 0 && reinterpret_cast<uintptr_t>(points_dev) % alignof(float4) != 0)
    return fail(TSDF_ERR_INVALID_ARG, "the point buffer must be 16-byte aligned (float4 rows)");
  for (int k = 0; k < 3; ++k)
    if (!std::isfinite(sensor[k])) return fail(TSDF_ERR_INVALID_ARG, "non-finite sensor position");
  tsdf_engine* E = m->eng;
  // the pass scratch layout, as ensure_scratch lays it out for the engine's
  // own scratch: pair rows and records with max_vox slots per point, and one
  // region of fold descriptors per size class, each as large as that class
  // can get in a pass over this map (no more blocks than the map's entries)
  const int64_t pc = std::max<int64_t>(n, 1);
  const int64_t stride = m->desc.max_vox;
  const int64_t slots = pc * stride;
  const int64_t fold_cap = std::min<int64_t>(((int64_t)m->desc.n_home + m->desc.n_ovf) * kBucket, slots);
  int64_t fold_base[FC_COUNT], fold_total = 0;
  for (int c = 0; c < FC_COUNT; ++c) {
    fold_base[c] = fold_total;
    fold_total += fold_class_cap(c, fold_cap, slots);
  }
  auto up = [](int64_t b) { return (b + 255) & ~int64_t(255); };
  const int64_t o_pe = 0;
  const int64_t o_pr = up(o_pe + (int64_t)sizeof(uint32_t) * slots);
  const int64_t o_seg = up(o_pr + (int64_t)sizeof(uint32_t) * slots);
  const int64_t o_fd = up(o_seg + (int64_t)sizeof(unsigned long long) * slots);
  const int64_t o_pn = up(o_fd + (int64_t)sizeof(int4) * fold_total);
  const int64_t bytes = up(o_pn + pc);
  if (bytes > kNoAllocScratchBytes)
    return fail(TSDF_ERR_CAPACITY,
                "tsdf_integrate_noalloc: %lld points at %lld voxels per point need %lld bytes of pass scratch, "
                "more than its fixed %lld",
                (long long)n, (long long)stride, (long long)bytes, (long long)kNoAllocScratchBytes);
  hipError_t se = hipSuccess;
  unsigned char* base = static_cast<unsigned char*>(scratch_address(E->device, &se));
  if (!base) return fail(TSDF_ERR_HIP, "the pass scratch of device %d is unavailable: %s", E->device, hipGetErrorString(se));
  CHK(set_device(E));
  Dev D = E->view();
  D.pair_e = reinterpret_cast<uint32_t*>(base + o_pe);
  D.pair_r = reinterpret_cast<uint32_t*>(base + o_pr);
  D.seg = reinterpret_cast<unsigned long long*>(base + o_seg);
  D.fold_desc = reinterpret_cast<int4*>(base + o_fd);
  D.fold_cap = fold_cap;
  for (int c = 0; c < FC_COUNT; ++c) D.fold_base[c] = fold_base[c];
  D.pair_n = base + o_pn;
  D.pcap = pc;
  D.seg_cap = slots;
  PassInline inl{};
  inl.hdr.points = reinterpret_cast<const float4*>(points_dev);
  inl.hdr.total = (int32_t)n;
  inl.hdr.nscans = 1;
  inl.hdr.stride = (int32_t)stride;
  inl.nscans = 1;
  inl.scans[0].map = m->slot;
  inl.scans[0].n = (int32_t)n;
  inl.scans[0].off = 0;
  for (int k = 0; k < 3; ++k) inl.scans[0].s[k] = sensor[k];
  int tid;
  CHK(timing_begin(E, 0, &tid));
  CHK(enqueue_pass(E, D, n, stride, tid, inl, 1));
  timing_end(E, tid);
  E->last_points = n;
  return TSDF_OK;
}

int tsdf_publish(tsdf_map* m, const float sensor[3], float* pos, float* data, int64_t capacity, int64_t* count) {
  CHK(check_map(m));
  if (!sensor) return fail(TSDF_ERR_INVALID_ARG, "null sensor");
  return with_state(m->eng->state_format,
                    [&](auto s) { return publish_impl<decltype(s)>(m, sensor, pos, data, capacity, count, 0); });
}

int tsdf_publish_msg(tsdf_map* m, const float sensor[3], float* voxels_msg, int64_t capacity, int64_t* count) {
  CHK(check_map(m));
  if (!sensor) return fail(TSDF_ERR_INVALID_ARG, "null sensor");
  return with_state(m->eng->state_format,
                    [&](auto s) { return publish_impl<decltype(s)>(m, sensor, voxels_msg, nullptr, capacity, count, 1); });
}

int tsdf_gc(tsdf_map* m, const float sensor[3], int64_t* removed) {
  CHK(check_map(m));
  if (!sensor) return fail(TSDF_ERR_INVALID_ARG, "null sensor");
  return with_state(m->eng->state_format, [&](auto s) { return gc_impl<decltype(s)>(m, sensor, 0, removed); });
}

int tsdf_process_scan(tsdf_map* m, const float* points_host, int64_t n, const float sensor[3], float* pos, float* data,
                      int64_t capacity, int64_t* count) {
  CHK(check_map(m));
  if (!sensor || (n > 0 && !points_host)) return fail(TSDF_ERR_INVALID_ARG, "null argument");
  if (n < 0 || n > kMaxPassPoints) return fail(TSDF_ERR_CAPACITY, "point count %lld out of range", (long long)n);
  tsdf_engine* E = m->eng;
  CHK(set_device(E));
  // the points go through an engine buffer kept between calls (the stream is
  // idle when it grows: every call ends synchronised)
  const int64_t want = std::max<int64_t>(n, 1) * 4;
  if (want > E->in_buf_floats) {
    HIPCHK(hipStreamSynchronize(E->stream));
    (void)hipFree(E->in_buf);
    E->in_buf = nullptr;
    E->in_buf_floats = 0;
    const int64_t grow = std::max<int64_t>(want, (int64_t)1 << 20);
    CHK(dev_alloc(&E->in_buf, grow));
    E->in_buf_floats = grow;
  }
  float* dpts = E->in_buf;
  int r = TSDF_OK;
  hipError_t e = hipMemcpyAsync(dpts, points_host, sizeof(float) * 4 * n, hipMemcpyHostToDevice, E->stream);
  if (e != hipSuccess) r = fail(TSDF_ERR_HIP, "H2D copy failed: %s", hipGetErrorString(e));
  if (!r) r = tsdf_integrate(m, dpts, n, sensor);
  if (!r) {
    // the pass's failure counters, reduced while the stream is still in order
    // (publish and gc keep their own statistics)
    hipLaunchKernelGGL(k_pass_stats, dim3(ST_COUNT), dim3(kPassStatsThreads), 0, E->stream, E->view(), E->last_grids, kDropStats);
    e = hipGetLastError();
    if (e != hipSuccess) r = fail(TSDF_ERR_HIP, "statistics reduction failed: %s", hipGetErrorString(e));
  }
  if (!r) r = tsdf_publish(m, sensor, pos, data, capacity, count);
  if (!r) r = tsdf_gc(m, sensor, nullptr);
  e = hipStreamSynchronize(E->stream);
  if (!r && e != hipSuccess) r = fail(TSDF_ERR_HIP, "stream synchronisation failed: %s", hipGetErrorString(e));
  if (!r) {
    // publish read the engine counters after the reduction above
    const unsigned long long* st = E->ctr_host->stats;
    if (st[ST_INSERT_FAIL] || st[ST_HEAP_FAIL] || st[ST_UPD_DROPPED])
      r = fail(TSDF_ERR_DROPPED, "the scan dropped %llu updates (%llu block inserts refused, %llu heap blocks refused)",
               st[ST_UPD_DROPPED], st[ST_INSERT_FAIL], st[ST_HEAP_FAIL]);
  }
  return r;
}

int tsdf_last_stats(tsdf_engine* E, tsdf_scan_stats* out) {
  if (!E || !out) return fail(TSDF_ERR_INVALID_ARG, "null argument");
  CHK(set_device(E));
  // the last pass's statistics stay in their workgroup rows until the next
  // pass overwrites them: sum them now
  hipLaunchKernelGGL(k_pass_stats, dim3(ST_COUNT), dim3(kPassStatsThreads), 0, E->stream, E->view(), E->last_grids, kPassStats);
  HIPCHK(hipGetLastError());
  CHK(read_ctr(E));
  const unsigned long long* s = E->ctr_host->stats;
  out->points = E->last_points;
  out->valid = (int64_t)s[ST_VALID];
  out->skipped = (int64_t)s[ST_SKIPPED];
  out->block_cells = (int64_t)s[ST_BLOCK_CELLS];
  out->voxel_updates = (int64_t)s[ST_UPDATES];
  out->unique_voxels = (int64_t)s[ST_UNIQUE];
  out->new_blocks = (int64_t)s[ST_NEW_BLOCKS];
  out->trunc_vox = (int64_t)s[ST_TRUNC_VOX];
  out->trunc_blk = (int64_t)s[ST_TRUNC_BLK];
  out->index_dropped = (int64_t)s[ST_INDEX_DROPPED];
  out->insert_failures = (int64_t)s[ST_INSERT_FAIL];
  out->heap_failures = (int64_t)s[ST_HEAP_FAIL];
  out->updates_dropped = (int64_t)s[ST_UPD_DROPPED];
  out->published = (int64_t)s[ST_PUBLISHED];
  out->gc_removed = (int64_t)s[ST_GC_REMOVED];
  out->quota_failures = (int64_t)s[ST_QUOTA_FAIL];
  if (out->insert_failures || out->heap_failures || out->updates_dropped)
    return fail(TSDF_ERR_DROPPED, "the last pass dropped %lld updates (%lld block inserts refused, %lld heap blocks refused)",
                (long long)out->updates_dropped, (long long)out->insert_failures, (long long)out->heap_failures);
  return TSDF_OK;
}

int tsdf_engine_set_launch(tsdf_engine* E, const tsdf_launch_config* cfg) {
  if (!E || !cfg) return fail(TSDF_ERR_INVALID_ARG, "null argument");
  if (cfg->traverse_block < 64 || cfg->traverse_block > 256 || (cfg->traverse_block % 64) != 0)
    return fail(TSDF_ERR_INVALID_ARG, "traverse_block must be 64, 128, 192 or 256");
  if (cfg->traverse_grid_cap < 1 || cfg->traverse_grid_cap > kStatRows[SR_BIN] || cfg->fold_grid < 1 ||
      cfg->fold_grid > kStatRows[SR_FOLD_A])
    return fail(TSDF_ERR_INVALID_ARG, "grid sizes out of range (1..65536)");
  if (cfg->rays_per_lane < 0 || cfg->rays_per_lane > kMaxRaysPerLane)
    return fail(TSDF_ERR_INVALID_ARG, "rays_per_lane must be 0..%d", kMaxRaysPerLane);
  if (cfg->traverse_queue != 0 && cfg->traverse_queue != 1) return fail(TSDF_ERR_INVALID_ARG, "traverse_queue must be 0 or 1");
  if (cfg->reserved != 0) return fail(TSDF_ERR_INVALID_ARG, "reserved must be 0");
  CHK(set_device(E));
  int resident = 0;
  CHK(bin_residency(E, cfg->traverse_block, resident));
  E->launch = *cfg;
  if (E->launch.rays_per_lane == 0) E->launch.rays_per_lane = 1;
  E->bin_resident = resident;
  return TSDF_OK;
}

int tsdf_engine_get_launch(tsdf_engine* E, tsdf_launch_config* cfg) {
  if (!E || !cfg) return fail(TSDF_ERR_INVALID_ARG, "null argument");
  *cfg = E->launch;
  return TSDF_OK;
}

int tsdf_timing_enable(tsdf_engine* E, int on) {
  if (!E) return fail(TSDF_ERR_INVALID_ARG, "null engine");
  E->timing = on != 0;
  return TSDF_OK;
}

int tsdf_timing_read(tsdf_engine* E, double* ms, int64_t* launches, int reset) {
  if (!E) return fail(TSDF_ERR_INVALID_ARG, "null engine");
  CHK(set_device(E));
  // force a drain of pending sets
  HIPCHK(hipStreamSynchronize(E->stream));
  while (!E->tpending.empty()) {
    const int s = E->tpending.front();
    E->tpending.pop_front();
    auto& T = E->tsets[s];
    float v;
    if (T.kind == 0) {
      for (int k = 0; k < 5; ++k) {
        HIPCHK(hipEventElapsedTime(&v, T.ev[k], T.ev[k + 1]));
        E->tms[k] += v;
        E->tcount[k]++;
      }
      HIPCHK(hipEventElapsedTime(&v, T.ev[0], T.ev[5]));
      E->tms[TSDF_STAGE_TOTAL] += v;
      E->tcount[TSDF_STAGE_TOTAL]++;
    } else {
      HIPCHK(hipEventElapsedTime(&v, T.ev[0], T.ev[1]));
      const int st = T.kind == 1 ? TSDF_STAGE_PUBLISH : TSDF_STAGE_GC;
      E->tms[st] += v;
      E->tcount[st]++;
    }
    E->tfree.push_back(s);
  }
  for (int k = 0; k < TSDF_NUM_STAGES; ++k) {
    if (ms) ms[k] = E->tms[k];
    if (launches) launches[k] = E->tcount[k];
    if (reset) {
      E->tms[k] = 0;
      E->tcount[k] = 0;
    }
  }
  return TSDF_OK;
}

// ---- graphs --------------------------------------------------------------------------
static int graph_capture(tsdf_batch_graph* g) {
  tsdf_engine* E = g->eng;
  if (g->exec) {
    hipGraphExecDestroy(g->exec);
    g->exec = nullptr;
  }
  if (g->graph) {
    hipGraphDestroy(g->graph);
    g->graph = nullptr;
  }
  CHK(ensure_scratch(E, g->max_points, g->stride));
  HIPCHK(hipStreamSynchronize(E->stream));
  const bool t = E->timing;
  E->timing = false;
  HIPCHK(hipStreamBeginCapture(E->stream, hipStreamCaptureModeThreadLocal));
  PassInline staged{};
  staged.nscans = -1;  // a replay's descriptors are copied in before each launch
  int r = enqueue_pass(E, E->view(), g->max_points, E->stride_cap, -1, staged, g->max_scans);
  hipGraph_t graph = nullptr;
  hipError_t e = hipStreamEndCapture(E->stream, &graph);
  E->timing = t;
  if (r) {
    if (graph) hipGraphDestroy(graph);
    return r;
  }
  if (e != hipSuccess) return fail(TSDF_ERR_HIP, "graph capture failed: %s", hipGetErrorString(e));
  g->graph = graph;
  g->grids = E->last_grids;
  HIPCHK(hipGraphInstantiate(&g->exec, graph, nullptr, nullptr, 0));
  g->gen = E->scratch_gen;
  return TSDF_OK;
}

int tsdf_batch_graph_create(tsdf_engine* E, tsdf_map* const* maps, int32_t nmaps, int64_t max_points, int32_t max_scans,
                            tsdf_batch_graph** out) {
  if (!E || !maps || nmaps <= 0 || !out) return fail(TSDF_ERR_INVALID_ARG, "null argument");
  if (max_points <= 0 || max_points > kMaxPassPoints) return fail(TSDF_ERR_CAPACITY, "max_points out of range");
  if (max_scans <= 0 || max_scans > kMaxScansPerBatch) return fail(TSDF_ERR_CAPACITY, "max_scans out of range");
  CHK(set_device(E));
  tsdf_batch_graph* g = new tsdf_batch_graph();
  g->eng = E;
  g->max_points = max_points;
  g->max_scans = max_scans;
  g->stride = 1;
  for (int i = 0; i < nmaps; ++i) {
    int r = check_map(maps[i]);
    if (!r && maps[i]->eng != E) r = fail(TSDF_ERR_INVALID_ARG, "map %d belongs to another engine", i);
    if (r) {
      delete g;
      return r;
    }
    g->slots.push_back(maps[i]->slot);
    g->ids.push_back(maps[i]->id);
    g->stride = std::max(g->stride, maps[i]->desc.max_vox);
  }
  int r = graph_capture(g);
  if (r) {
    tsdf_batch_graph_destroy(g);
    return r;
  }
  *out = g;
  return TSDF_OK;
}

int tsdf_batch_graph_launch(tsdf_batch_graph* g, const tsdf_scan_desc* scans, int32_t nscans, const float* points_dev) {
  if (!g) return fail(TSDF_ERR_INVALID_ARG, "null graph");
  tsdf_engine* E = g->eng;
  if (nscans < 0 || nscans > g->max_scans) return fail(TSDF_ERR_CAPACITY, "too many scans for this graph");
  if (nscans > 0 && !scans) return fail(TSDF_ERR_INVALID_ARG, "null scans");
  CHK(set_device(E));
  ScanDev sd[kMaxScansPerBatch];
  int64_t P = 0;
  for (int i = 0; i < nscans; ++i) {
    const tsdf_scan_desc& s = scans[i];
    if (s.map_index < 0 || s.map_index >= (int)g->slots.size()) return fail(TSDF_ERR_INVALID_ARG, "map_index out of range");
    const int slot = g->slots[s.map_index];
    if (!E->map_by_slot[slot] || E->map_by_slot[slot]->id != g->ids[s.map_index])
      return fail(TSDF_ERR_INVALID_ARG, "scan %d: map %d of the graph was destroyed", i, s.map_index);
    if (s.n_points < 0 || s.point_offset < 0) return fail(TSDF_ERR_INVALID_ARG, "scan %d: negative size/offset", i);
    if (s.point_offset != (nscans ? scans[0].point_offset : 0) + P)
      return fail(TSDF_ERR_INVALID_ARG, "scans of a batch must be contiguous and in order in the point buffer");
    for (int k = 0; k < 3; ++k)
      if (!std::isfinite(s.sensor[k])) return fail(TSDF_ERR_INVALID_ARG, "non-finite sensor position");
    sd[i].map = slot;
    sd[i].n = s.n_points;
    sd[i].off = (int32_t)P;
    sd[i].s[0] = s.sensor[0];
    sd[i].s[1] = s.sensor[1];
    sd[i].s[2] = s.sensor[2];
    P += s.n_points;
  }
  if (P > g->max_points) return fail(TSDF_ERR_CAPACITY, "batch of %lld points exceeds the graph's %lld", (long long)P,
                                     (long long)g->max_points);
  if (P > 0 && !points_dev) return fail(TSDF_ERR_INVALID_ARG, "null point buffer");
  if (P > 0 && reinterpret_cast<uintptr_t>(points_dev) % alignof(float4) != 0)
    return fail(TSDF_ERR_INVALID_ARG, "the point buffer must be 16-byte aligned (float4 rows)");
  if (g->gen != E->scratch_gen) CHK(graph_capture(g));
  PassHdr hdr{};
  hdr.points = reinterpret_cast<const float4*>(points_dev) + (nscans ? scans[0].point_offset : 0);
  hdr.total = (int32_t)P;
  hdr.nscans = nscans;
  hdr.stride = g->stride;
  CHK(stage_descs(E, hdr, sd, nscans));
  HIPCHK(hipGraphLaunch(g->exec, E->stream));
  E->last_points = P;
  E->last_grids = g->grids;
  return TSDF_OK;
}

int tsdf_batch_graph_destroy(tsdf_batch_graph* g) {
  if (!g) return TSDF_OK;
  if (g->eng) {
    hipSetDevice(g->eng->device);
    hipStreamSynchronize(g->eng->stream);
  }
  if (g->exec) hipGraphExecDestroy(g->exec);
  if (g->graph) hipGraphDestroy(g->graph);
  delete g;
  return TSDF_OK;
}

// ---- device memory plumbing ------------------------------------------------------------
int tsdf_device_alloc(tsdf_engine* E, int64_t bytes, void** out) {
  if (!E || !out || bytes < 0) return fail(TSDF_ERR_INVALID_ARG, "bad argument");
  CHK(set_device(E));
  unsigned char* p = nullptr;
  CHK(dev_alloc(&p, bytes));
  *out = p;
  return TSDF_OK;
}

int tsdf_device_free(tsdf_engine* E, void* ptr) {
  if (!E) return fail(TSDF_ERR_INVALID_ARG, "null engine");
  CHK(set_device(E));
  HIPCHK(hipStreamSynchronize(E->stream));
  if (ptr) HIPCHK(hipFree(ptr));
  return TSDF_OK;
}

int tsdf_host_alloc(tsdf_engine* E, int64_t bytes, void** out) {
  if (!E || !out || bytes < 0) return fail(TSDF_ERR_INVALID_ARG, "bad argument");
  *out = nullptr;
  CHK(set_device(E));
  void* p = nullptr;
  const hipError_t e = hipHostMalloc(&p, (size_t)std::max<int64_t>(bytes, 1), hipHostMallocDefault);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return fail(TSDF_ERR_NO_MEMORY, "hipHostMalloc of %lld bytes failed: %s", (long long)bytes, hipGetErrorString(e));
  }
  *out = p;
  return TSDF_OK;
}

int tsdf_host_free(void* ptr) {
  if (ptr) HIPCHK(hipHostFree(ptr));
  return TSDF_OK;
}

int tsdf_copy_to_device(tsdf_engine* E, void* dst, const void* src, int64_t bytes) {
  if (!E || (bytes > 0 && (!dst || !src)) || bytes < 0) return fail(TSDF_ERR_INVALID_ARG, "bad argument");
  CHK(set_device(E));
  HIPCHK(hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyHostToDevice, E->stream));
  HIPCHK(hipStreamSynchronize(E->stream));
  return TSDF_OK;
}

int tsdf_copy_to_host(tsdf_engine* E, void* dst, const void* src, int64_t bytes) {
  if (!E || (bytes > 0 && (!dst || !src)) || bytes < 0) return fail(TSDF_ERR_INVALID_ARG, "bad argument");
  CHK(set_device(E));
  HIPCHK(hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyDeviceToHost, E->stream));
  HIPCHK(hipStreamSynchronize(E->stream));
  return TSDF_OK;
}

int tsdf_device_copy(tsdf_engine* E, void* dst, const void* src, int64_t bytes) {
  if (!E || bytes < 0 || (bytes % 16) != 0 || (bytes > 0 && (!dst || !src)))
    return fail(TSDF_ERR_INVALID_ARG, "bad argument (bytes must be a multiple of 16)");
  if ((reinterpret_cast<uintptr_t>(src) | reinterpret_cast<uintptr_t>(dst)) % alignof(float4) != 0)
    return fail(TSDF_ERR_INVALID_ARG, "source and destination must be 16-byte aligned");
  CHK(set_device(E));
  const int64_t n = bytes / 16;
  hipLaunchKernelGGL(k_copy, dim3(grid_for(n, 256, 8192)), dim3(256), 0, E->stream, reinterpret_cast<const float4*>(src),
                     reinterpret_cast<float4*>(dst), n);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(E->stream));
  return TSDF_OK;
}

// ---- downstream consumer -------------------------------------------------------------------
int tsdf_msg_markers(tsdf_engine* E, const float* msg, int64_t n, int msg_on_device, float* nonneg, float* neg,
                     int64_t* n_nonneg, int64_t* n_neg) {
  if (!E || !n_nonneg || !n_neg || n < 0 || (n > 0 && (!msg || !nonneg || !neg)))
    return fail(TSDF_ERR_INVALID_ARG, "bad argument");
  if (n > ((int64_t)1 << 31) - kMsgTile) return fail(TSDF_ERR_CAPACITY, "a message holds at most 2^31 voxels here");
  *n_nonneg = *n_neg = 0;
  if (n == 0) return TSDF_OK;
  CHK(set_device(E));
  hipStream_t st = E->stream;
  const int tiles = (int)((n + kMsgTile - 1) / kMsgTile);
  float* dmsg = nullptr;
  float* dout = nullptr;
  uint32_t* dcnt = nullptr;
  auto release = [&](int r) {
    (void)hipStreamSynchronize(st);
    if (!msg_on_device) (void)hipFree(dmsg);
    (void)hipFree(dout);
    (void)hipFree(dcnt);
    return r;
  };
  int r;
  if (!msg_on_device && (r = dev_alloc(&dmsg, 5 * n))) return release(r);
  if ((r = dev_alloc(&dout, 3 * n)) || (r = dev_alloc(&dcnt, tiles))) return release(r);
  const float* src = msg_on_device ? msg : dmsg;
  if (!msg_on_device && hipMemcpyAsync(dmsg, msg, sizeof(float) * 5 * n, hipMemcpyHostToDevice, st) != hipSuccess)
    return release(fail(TSDF_ERR_HIP, "H2D copy of the message failed"));
  // the two lists share one device buffer, the sdf >= 0 list first: the
  // sdf < 0 list starts where the count puts the end of the first
  hipLaunchKernelGGL(k_msg_count, dim3(tiles), dim3(256), 0, st, src, n, dcnt);
  std::vector<uint32_t> cnt(tiles);
  if (hipMemcpyAsync(cnt.data(), dcnt, sizeof(uint32_t) * tiles, hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return release(fail(TSDF_ERR_HIP, "marker count failed"));
  int64_t np = 0;
  for (uint32_t c : cnt) np += c;
  hipLaunchKernelGGL(k_msg_split, dim3(tiles), dim3(256), 0, st, src, n, dcnt, dout, dout + 3 * np);
  if (hipGetLastError() != hipSuccess ||
      hipMemcpyAsync(nonneg, dout, sizeof(float) * 3 * np, hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipMemcpyAsync(neg, dout + 3 * np, sizeof(float) * 3 * (n - np), hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return release(fail(TSDF_ERR_HIP, "marker split failed"));
  *n_nonneg = np;
  *n_neg = n - np;
  return release(TSDF_OK);
}

// ---- input path --------------------------------------------------------------------------
int tsdf_cloud2_to_points(tsdf_engine* E, const tsdf_cloud2_layout* L, const void* cloud, int cloud_on_device,
                          float* points_dev, int64_t capacity, int64_t* n_points) {
  if (!E || !L || !n_points) return fail(TSDF_ERR_INVALID_ARG, "null argument");
  if (L->width < 0 || L->height < 0 || L->point_step <= 0 || L->row_step < 0 || L->data_bytes < 0)
    return fail(TSDF_ERR_INVALID_ARG, "negative cloud size or non-positive point_step");
  if (L->datatype != TSDF_FIELD_FLOAT32)
    return fail(TSDF_ERR_INVALID_ARG, "x/y/z datatype %d: only FLOAT32 (7) fields convert to pcl::PointXYZ", L->datatype);
  if (L->is_bigendian != 0 && L->is_bigendian != 1) return fail(TSDF_ERR_INVALID_ARG, "is_bigendian must be 0 or 1");
  const int32_t off[3] = {L->x_offset, L->y_offset, L->z_offset};
  for (int k = 0; k < 3; ++k)
    if (off[k] < 0 || (int64_t)off[k] + 4 > L->point_step)
      return fail(TSDF_ERR_INVALID_ARG, "field offset %d does not fit in point_step %d", off[k], L->point_step);
  if (L->width > 0 && L->height > 1 && L->row_step < L->width * (int64_t)L->point_step)
    return fail(TSDF_ERR_INVALID_ARG, "row_step %lld below width * point_step", (long long)L->row_step);
  if (L->width > 0 && L->height > 0 && L->width > (int64_t{1} << 31) / L->height)
    return fail(TSDF_ERR_INVALID_ARG, "cloud of more than 2^31 points");
  if (L->height > 1 && L->height - 1 > L->data_bytes / std::max<int64_t>(L->row_step, 1))
    return fail(TSDF_ERR_INVALID_ARG, "rows reach past the %lld-byte data block", (long long)L->data_bytes);
  const int64_t n = L->width * L->height;
  *n_points = n;
  if (n == 0) return TSDF_OK;
  if (n > capacity) return fail(TSDF_ERR_CAPACITY, "cloud of %lld points, capacity %lld", (long long)n, (long long)capacity);
  if (!cloud || !points_dev) return fail(TSDF_ERR_INVALID_ARG, "null cloud or output");
  if (reinterpret_cast<uintptr_t>(points_dev) % alignof(float4) != 0)
    return fail(TSDF_ERR_INVALID_ARG, "the output must be 16-byte aligned (float4 rows)");
  // bytes read: up to the end of the last point of the last row
  const int64_t extent = (L->height - 1) * L->row_step + L->width * (int64_t)L->point_step;
  if (extent > L->data_bytes)
    return fail(TSDF_ERR_INVALID_ARG, "layout reads %lld bytes of a %lld-byte data block", (long long)extent,
                (long long)L->data_bytes);
  CHK(set_device(E));
  const uint8_t* src = static_cast<const uint8_t*>(cloud);
  uint8_t* staged = nullptr;
  if (!cloud_on_device) {
    CHK(dev_alloc(&staged, extent));
    const hipError_t ce = hipMemcpyAsync(staged, cloud, (size_t)extent, hipMemcpyHostToDevice, E->stream);
    if (ce != hipSuccess) {
      hipFree(staged);
      return fail(TSDF_ERR_HIP, "H2D copy failed: %s", hipGetErrorString(ce));
    }
    src = staged;
  }
  const bool aligned = (reinterpret_cast<uintptr_t>(src) % 4) == 0 && (L->point_step % 4) == 0 &&
                       (L->row_step % 4) == 0 && (off[0] % 4) == 0 && (off[1] % 4) == 0 && (off[2] % 4) == 0;
  const dim3 grid(grid_for(n, 256, 8192));
  if (aligned)
    hipLaunchKernelGGL(k_cloud2<true>, grid, dim3(256), 0, E->stream, src, L->width, n, L->row_step, L->point_step,
                       off[0], off[1], off[2], L->is_bigendian, reinterpret_cast<float4*>(points_dev));
  else
    hipLaunchKernelGGL(k_cloud2<false>, grid, dim3(256), 0, E->stream, src, L->width, n, L->row_step, L->point_step,
                       off[0], off[1], off[2], L->is_bigendian, reinterpret_cast<float4*>(points_dev));
  hipError_t ke = hipGetLastError();
  if (staged) {
    const hipError_t se = hipStreamSynchronize(E->stream);
    if (ke == hipSuccess) ke = se;
    hipFree(staged);
  }
  if (ke != hipSuccess) return fail(TSDF_ERR_HIP, "cloud decode failed: %s", hipGetErrorString(ke));
  return TSDF_OK;
}

// ---- serving ---------------------------------------------------------------------------
int tsdf_serve_run(tsdf_engine* E, tsdf_map* const* maps, int32_t nmaps, const tsdf_request* reqs, int32_t nreq,
                   const float* points_dev, const tsdf_serve_config* cfg, double* latency_ms, int32_t* batch_of,
                   tsdf_serve_stats* stats) {
  if (!E || !maps || nmaps <= 0 || (nreq > 0 && !reqs) || !cfg) return fail(TSDF_ERR_INVALID_ARG, "null argument");
  for (int i = 0; i < nmaps; ++i) {
    CHK(check_map(maps[i]));
    if (maps[i]->eng != E) return fail(TSDF_ERR_INVALID_ARG, "map %d belongs to another engine", i);
  }
  const int max_scans = std::max(1, std::min<int>(cfg->max_batch_scans, kMaxScansPerBatch));
  const int max_inflight = std::max(1, std::min(cfg->max_inflight, kRing - 2));
  const int64_t max_pts = cfg->max_batch_points > 0 ? std::min<int64_t>(cfg->max_batch_points, kMaxPassPoints) : kMaxPassPoints;
  for (int i = 0; i < nreq; ++i) {
    if (reqs[i].map_index < 0 || reqs[i].map_index >= nmaps) return fail(TSDF_ERR_INVALID_ARG, "request %d: bad map", i);
    if (reqs[i].n_points < 0 || reqs[i].n_points > max_pts) return fail(TSDF_ERR_INVALID_ARG, "request %d: bad size", i);
    if (reqs[i].point_offset < 0) return fail(TSDF_ERR_INVALID_ARG, "request %d: negative point offset", i);
    if (reqs[i].n_points > 0 && !points_dev) return fail(TSDF_ERR_INVALID_ARG, "null point buffer");
    if (i > 0 && reqs[i].arrival_s < reqs[i - 1].arrival_s)
      return fail(TSDF_ERR_INVALID_ARG, "requests must be sorted by arrival time");
  }
  CHK(set_device(E));
  // Requests of one batch must be contiguous in the pass's point space; they
  // come from arbitrary places of points_dev, so each batch gathers them
  // into a staging buffer with device-to-device copies.
  const int64_t stage_pts = max_pts;
  float* gather[kRing] = {};
  std::vector<hipEvent_t> evs(max_inflight, nullptr);
  auto release = [&]() {
    for (auto& e : evs)
      if (e) hipEventDestroy(e);
    for (int i = 0; i < max_inflight; ++i)
      if (gather[i]) hipFree(gather[i]);
  };
  int rc = TSDF_OK;
  for (int i = 0; i < max_inflight && rc == TSDF_OK; ++i) rc = dev_alloc(&gather[i], stage_pts * 4);
  for (auto& e : evs)
    if (rc == TSDF_OK && hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess)
      rc = fail(TSDF_ERR_HIP, "hipEventCreate failed");
  if (rc != TSDF_OK) {
    release();
    return rc;
  }
  struct Inflight {
    hipEvent_t done;
    std::vector<int> reqs;
    int buf;
  };
  std::vector<int> free_bufs;
  for (int i = 0; i < max_inflight; ++i) free_bufs.push_back(i);
  std::deque<Inflight> inflight;
  std::deque<int> pending;
  int next_req = 0, done = 0;
  int64_t batches = 0, misses = 0, points = 0;
  const auto t0 = Clock::now();
  auto now_s = [&]() { return std::chrono::duration<double>(Clock::now() - t0).count(); };
  double first_arrival = nreq > 0 ? reqs[0].arrival_s : 0.0, last_done = 0.0;
  std::vector<tsdf_scan_desc> sd;
  std::vector<tsdf_map*> bmaps(maps, maps + nmaps);
  while (done < nreq && rc == TSDF_OK) {
    const double t = now_s();
    while (next_req < nreq && reqs[next_req].arrival_s <= t) pending.push_back(next_req++);
    // retire finished batches
    while (!inflight.empty()) {
      hipError_t q = hipEventQuery(inflight.front().done);
      if (q == hipErrorNotReady) break;
      if (q != hipSuccess) {
        rc = fail(TSDF_ERR_HIP, "batch failed: %s", hipGetErrorString(q));
        break;
      }
      const double td = now_s();
      for (int r : inflight.front().reqs) {
        const double lat = (td - reqs[r].arrival_s) * 1e3;
        if (latency_ms) latency_ms[r] = lat;
        if (cfg->deadline_ms > 0 && lat > cfg->deadline_ms) misses++;
        done++;
      }
      last_done = td;
      free_bufs.push_back(inflight.front().buf);
      inflight.pop_front();
    }
    if (rc) break;
    if (!pending.empty() && !free_bufs.empty()) {
      // admit arrived requests into the next batch, oldest first
      sd.clear();
      std::vector<int> taken;
      int64_t P = 0;
      std::deque<int> rest;
      for (int r : pending) {
        if ((int)taken.size() < max_scans && P + reqs[r].n_points <= max_pts) {
          tsdf_scan_desc s{};
          s.map_index = reqs[r].map_index;
          s.n_points = reqs[r].n_points;
          s.point_offset = P;
          s.sensor[0] = reqs[r].sensor[0];
          s.sensor[1] = reqs[r].sensor[1];
          s.sensor[2] = reqs[r].sensor[2];
          sd.push_back(s);
          taken.push_back(r);
          P += reqs[r].n_points;
        } else {
          rest.push_back(r);
        }
      }
      pending.swap(rest);
      const int b = free_bufs.back();
      free_bufs.pop_back();
      for (size_t k = 0; k < taken.size(); ++k) {
        const tsdf_request& q = reqs[taken[k]];
        if (q.n_points > 0) {
          hipError_t e = hipMemcpyAsync(gather[b] + 4 * sd[k].point_offset, points_dev + 4 * q.point_offset,
                                        sizeof(float) * 4 * q.n_points, hipMemcpyDeviceToDevice, E->stream);
          if (e != hipSuccess) rc = fail(TSDF_ERR_HIP, "gather copy failed: %s", hipGetErrorString(e));
        }
        if (batch_of) batch_of[taken[k]] = (int32_t)batches;
      }
      if (rc) break;
      rc = run_batch(E, bmaps.data(), nmaps, sd.data(), (int)sd.size(), gather[b]);
      if (rc) break;
      Inflight f;
      f.done = evs[b];
      f.reqs = taken;
      f.buf = b;
      const hipError_t re = hipEventRecord(f.done, E->stream);
      if (re != hipSuccess) {
        rc = fail(TSDF_ERR_HIP, "hipEventRecord failed: %s", hipGetErrorString(re));
        break;
      }
      inflight.push_back(f);
      batches++;
      points += P;
      continue;
    }
    // nothing to launch: wait briefly for an arrival or a completion
    if (inflight.empty() && pending.empty() && next_req < nreq) {
      const double wait = reqs[next_req].arrival_s - now_s();
      if (wait > 2e-4) std::this_thread::sleep_for(std::chrono::microseconds((int64_t)((wait - 1e-4) * 1e6)));
    } else {
      std::this_thread::yield();
    }
  }
  hipStreamSynchronize(E->stream);
  release();
  if (stats) {
    stats->requests = nreq;
    stats->batches = batches;
    stats->deadline_misses = misses;
    stats->wall_s = last_done - first_arrival;
    stats->busy_s = 0;
    stats->points = points;
  }
  return rc;
}

}  // extern "C"
