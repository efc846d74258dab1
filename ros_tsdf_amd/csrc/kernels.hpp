// Kernels of the integration pipeline.  One pass integrates the scans of a
// batch (a single scan on the per-scan path):
//
//   k_bin            per point: truncation segment, block DDA and voxel DDA;
//                    find-or-insert every block in the map's spatial hash
//                    (lock-free) once per workgroup; each run of voxels in
//                    one block gets its record offset in that block; new
//                    entries join the pass's new-entry list
//   k_seg            heap blocks for the new entries (one pop per wave);
//                    record segment offset and 16-byte fold descriptor of
//                    every touched block, filed by size class (a scan over
//                    the touched list, one segment atomic per tile)
//   k_emit           per point: second voxel-DDA walk writing each update
//                    record (voxel, point, sdf) to its block's contiguous
//                    record segment at the offset reserved in k_bin
//   k_fold_a         blocks of 129..1024 records (a workgroup each: per-voxel
//                    sums in LDS), blocks of 65..128 (a wave each) and of
//                    <= 64 / 32 / 16 records, 1 / 2 / 4 per wave (bitonic
//                    sort across lanes; each voxel's first record sums the
//                    voxel's updates)
//   k_fold_b         larger blocks by per-voxel sums, and the ordered fold of
//                    blocks with a group that crosses max_weight
// Every voxel's state is read once and written once per pass; its updates of
// each scan are applied as one weighted mean (see "the fold of one voxel's
// updates in a pass").  Pass statistics are per-workgroup rows, summed on
// demand (k_pass_stats).
//
// publish: k_select (radius test per entry) -> k_pub_count (wave per block)
//          -> scan -> k_pub_write (wave ballot compaction)
// gc:      k_gc_mark -> k_gc_zero -> k_gc_chains
#pragma once
#include <type_traits>

#include "device.hpp"

namespace tsdfhip {

constexpr int kFoldThreads = 256;
constexpr int kCountTileS = 1024; // records per LDS counting tile of k_fold_a's workgroup tier (2 x 8 KB)
constexpr int kCountTile = 4096;  // records per LDS counting tile of k_fold_b (2 x 32 KB)
constexpr int kFoldCap = 8192;    // records of k_fold_b's LDS buffer (64 KB: the counting tile pair)
constexpr int kScanTile = 1024;  // elements per workgroup in the scan kernels

// Size classes of touched blocks (records per block): the fold kernel and
// the number of blocks folded together differ per class.
enum FoldClass { FC_16 = 0, FC_32, FC_64, FC_WGL, FC_WG, FC_WG2, FC_HUGE, FC_COUNT };
// Set by k_fold_a in the map field of a descriptor it hands to k_fold_b (a
// group of the block would cross max_weight and the block holds more records
// than k_fold_a's ordered tile, kOrdTileA): a plain store, where a shared
// list would put hundreds of returning atomics on one counter.  k_seg writes
// every descriptor of a pass afresh, without it.
constexpr int32_t kDescDeferred = 1 << 30;
constexpr int kWgSplit = 128;  // FC_WGL: 65..128 records (a wave each), FC_WG: 129..kCountTileS
static_assert(FC_COUNT <= (int)(sizeof(EngCtr::fold_n) / sizeof(int32_t)), "fold class counters");

// Descriptors a class can need when a pass holds at most `blocks` touched
// blocks and `records` records: every block of class c holds at least lo
// records (the class's lower size bound).
__host__ __device__ constexpr int64_t fold_class_cap(int c, int64_t blocks, int64_t records) {
  const int64_t lo = c == FC_16 ? 1 : c == FC_32 ? 17 : c == FC_64 ? 33 : c == FC_WGL ? 65 : c == FC_WG ? kWgSplit + 1 : c == FC_WG2 ? kCountTileS + 1 : kCountTile + 1;
  const int64_t r = records / lo + 1;
  return r < blocks ? r : blocks;
}

__host__ __device__ constexpr int fold_class(int n) {
  return n <= 16 ? FC_16 : n <= 32 ? FC_32 : n <= 64 ? FC_64 : n <= kWgSplit ? FC_WGL : n <= kCountTileS ? FC_WG : n <= kCountTile ? FC_WG2 : FC_HUGE;
}

// ---- wave / workgroup helpers -------------------------------------------------
__device__ __forceinline__ int lane_id() { return __lane_id(); }

// Inclusive scan across the wave (lane i gets v[0] + ... + v[i]) in DPP
// moves between registers, with no LDS round trip: Hillis-Steele steps
// inside each row of 16 lanes (row_shr 1, 2, 4, 8; a lane whose source is
// outside its row adds 0), then the total of row 0 into row 1 and of row 2
// into row 3 (row_bcast15 on rows 1 and 3), then the total of rows 0-1 into
// rows 2 and 3 (row_bcast31).  Rows a step leaves out add 0.
template <int kCtrl, int kRowMask>
__device__ __forceinline__ uint32_t dpp_u32(uint32_t x) {
  return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, kCtrl, kRowMask, 0xf, true);
}
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t x) {
  x += dpp_u32<0x111, 0xf>(x);  // row_shr:1
  x += dpp_u32<0x112, 0xf>(x);  // row_shr:2
  x += dpp_u32<0x114, 0xf>(x);  // row_shr:4
  x += dpp_u32<0x118, 0xf>(x);  // row_shr:8
  x += dpp_u32<0x142, 0xa>(x);  // row_bcast:15
  x += dpp_u32<0x143, 0xc>(x);  // row_bcast:31
  return x;
}
template <int kCtrl, int kRowMask>
__device__ __forceinline__ unsigned long long dpp_u64(unsigned long long x) {
  const uint32_t lo = dpp_u32<kCtrl, kRowMask>((uint32_t)x), hi = dpp_u32<kCtrl, kRowMask>((uint32_t)(x >> 32));
  return ((unsigned long long)hi << 32) | lo;
}
__device__ __forceinline__ unsigned long long wave_incl_scan64(unsigned long long x) {
  x += dpp_u64<0x111, 0xf>(x);
  x += dpp_u64<0x112, 0xf>(x);
  x += dpp_u64<0x114, 0xf>(x);
  x += dpp_u64<0x118, 0xf>(x);
  x += dpp_u64<0x142, 0xa>(x);
  x += dpp_u64<0x143, 0xc>(x);
  return x;
}

// Lane i's copy of x from lane i ^ J (J a power of two below 64), every lane
// active, in VALU lane moves instead of an LDS permute: DPP quad_perm for 1
// and 2, a quad mirror then row_half_mirror for 4 (i ^ 3 ^ 7), row_ror:8 for
// 8, and gfx950's permlane16/32 swaps for 16 and 32.  A swap leaves each
// lane its own value in one result and lane i ^ J's in the other, so their
// sum minus x is the latter (exact in unsigned arithmetic).
template <int J>
__device__ __forceinline__ uint32_t xor_lane(uint32_t x) {
  if constexpr (J == 1) return dpp_u32<0xB1, 0xf>(x);
  else if constexpr (J == 2) return dpp_u32<0x4E, 0xf>(x);
  else if constexpr (J == 4) return dpp_u32<0x141, 0xf>(dpp_u32<0x1B, 0xf>(x));
  else if constexpr (J == 8) return dpp_u32<0x128, 0xf>(x);
  else if constexpr (J == 16) {
    const auto r = __builtin_amdgcn_permlane16_swap(x, x, false, false);
    return r[0] + r[1] - x;
  } else {
    static_assert(J == 32, "a lane distance below 64");
    const auto r = __builtin_amdgcn_permlane32_swap(x, x, false, false);
    return r[0] + r[1] - x;
  }
}
template <int J>
__device__ __forceinline__ unsigned long long xor_lane64(unsigned long long v) {
  return ((unsigned long long)xor_lane<J>((uint32_t)(v >> 32)) << 32) | xor_lane<J>((uint32_t)v);
}
// The same with j known only after unrolling (the switch folds away).
__device__ __forceinline__ unsigned long long xor_lane64(unsigned long long v, int j) {
  switch (j) {
    case 1: return xor_lane64<1>(v);
    case 2: return xor_lane64<2>(v);
    case 4: return xor_lane64<4>(v);
    case 8: return xor_lane64<8>(v);
    case 16: return xor_lane64<16>(v);
    default: return xor_lane64<32>(v);
  }
}
// Lane i's copy of x from lane i + 1 (wave_shl:1; the last lane gets 0), and
// from lane i - 1 (wave_shr:1; lane 0 gets 0).
__device__ __forceinline__ uint32_t next_lane(uint32_t x) { return dpp_u32<0x130, 0xf>(x); }
__device__ __forceinline__ float next_lane(float x) { return __uint_as_float(next_lane(__float_as_uint(x))); }
__device__ __forceinline__ unsigned long long next_lane64(unsigned long long v) {
  return ((unsigned long long)next_lane((uint32_t)(v >> 32)) << 32) | next_lane((uint32_t)v);
}
__device__ __forceinline__ unsigned long long prev_lane64(unsigned long long v) {
  return ((unsigned long long)dpp_u32<0x138, 0xf>((uint32_t)(v >> 32)) << 32) | dpp_u32<0x138, 0xf>((uint32_t)v);
}

// All lanes of the wave must call this.  Returns the slot of this lane's item
// (or -1 when pred is false) in a list whose length is *ctr.
__device__ __forceinline__ int wave_append(int32_t* ctr, bool pred) {
  const unsigned long long m = __ballot(pred);
  if (m == 0) return -1;
  const int lane = lane_id();
  const int leader = __ffsll((long long)m) - 1;
  int base = 0;
  if (lane == leader) base = atomicAdd(ctr, (int)__popcll(m));
  base = __builtin_amdgcn_readlane(base, leader);
  const int rank = (int)__popcll(m & ((1ull << lane) - 1ull));
  return pred ? base + rank : -1;
}

// Sums over the wave (every lane active; every lane gets the sum).
__device__ __forceinline__ unsigned long long wave_sum(unsigned long long v) {
  const unsigned long long x = wave_incl_scan64(v);
  return ((unsigned long long)(uint32_t)__builtin_amdgcn_readlane((int)(x >> 32), 63) << 32) |
         (uint32_t)__builtin_amdgcn_readlane((int)x, 63);
}

// Statistics: each wave adds its sum into one of kPartSlots partial slots
// (spread by wave, so the adds do not pile up on one address); the pass's
// last kernel reduces the slots into EngCtr::stats.
__device__ __forceinline__ void stat_add(const Dev& D, int stat, unsigned long long v) {
  const unsigned long long s = wave_sum(v);
  if (lane_id() == 0 && s) {
    const int slot = (blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) & (kPartSlots - 1);
    atomicAdd(&D.part[stat * kPartSlots + slot], s);
  }
}


__device__ __forceinline__ uint32_t wave_sum32(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_readlane((int)wave_incl_scan(v), 63);
}


// The workgroup's row of pass statistics in region `region`: each thread's
// counts v[0..N) summed over the workgroup.  Every workgroup writes its row,
// zeros included, so rows need no clearing.  All threads must call this.
template <int N>
__device__ __forceinline__ void write_stat_row(const Dev& D, int region, const uint32_t (&v)[N]) {
  static_assert(N <= kStatCols, "row width");
  __shared__ uint32_t wsum[16][kStatCols];  // up to 16 waves (1024 threads)
  const int w = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < N; ++k) {
    const uint32_t x = wave_sum32(v[k]);
    if (lane_id() == 0) wsum[w][k] = x;
  }
  __syncthreads();
  if ((int)threadIdx.x < N) {
    uint32_t t = 0;
    for (int k = 0; k < (int)(blockDim.x >> 6); ++k) t += wsum[k][threadIdx.x];
    D.srows[(int64_t)(stat_region_base(region) + blockIdx.x) * kStatCols + threadIdx.x] = t;
  }
}

// EngCtr::stats of the last pass: the rows of each kernel's workgroups
// summed (grid sizes of that pass's launches, by region), one workgroup per
// statistic in mask.
struct StatGrids {
  int32_t n[SR_COUNT];
};
constexpr int kPassStatsThreads = 1024;
__global__ void __launch_bounds__(kPassStatsThreads) k_pass_stats(Dev D, StatGrids g, uint32_t mask) {
  __shared__ unsigned long long sh[kPassStatsThreads / 64];
  const int st = blockIdx.x;
  if (st >= ST_COUNT || !(mask & (1u << st))) return;
  // (region, column) pairs that hold the statistic
  int reg[4] = {-1, -1, -1, -1}, col[4] = {0, 0, 0, 0};
  switch (st) {
    case ST_VALID: reg[0] = SR_BIN; col[0] = 0; break;
    case ST_SKIPPED: reg[0] = SR_BIN; col[0] = 1; break;
    case ST_BLOCK_CELLS: reg[0] = SR_BIN; col[0] = 2; break;
    case ST_UPDATES: reg[0] = SR_BIN; col[0] = 3; break;
    case ST_TRUNC_VOX: reg[0] = SR_BIN; col[0] = 4; break;
    case ST_TRUNC_BLK: reg[0] = SR_BIN; col[0] = 5; break;
    case ST_INSERT_FAIL: reg[0] = SR_BIN; col[0] = 6; break;
    case ST_UPD_DROPPED: reg[0] = SR_BIN; col[0] = 7; reg[1] = SR_FOLD_A; col[1] = 1; reg[2] = SR_FOLD_B; col[2] = 1;
      break;
    case ST_NEW_BLOCKS: reg[0] = SR_SCAN; col[0] = 0; break;
    case ST_HEAP_FAIL: reg[0] = SR_SCAN; col[0] = 1; break;
    case ST_QUOTA_FAIL: reg[0] = SR_SCAN; col[0] = 2; break;
    case ST_UNIQUE: reg[0] = SR_FOLD_A; col[0] = 0; reg[1] = SR_FOLD_B; col[1] = 0; break;
    default: break;  // ST_INDEX_DROPPED: no kernel counts it (valid points never produce it)
  }
  unsigned long long v = 0;
  for (int q = 0; q < 4; ++q) {
    if (reg[q] < 0) continue;
    const uint32_t* r = D.srows + (int64_t)stat_region_base(reg[q]) * kStatCols + col[q];
    const int n = g.n[reg[q]];
    // four rows per thread per step, loaded before they are added (a
    // 65536-row region is 16 steps of 1024 threads)
    for (int x0 = threadIdx.x; x0 < n; x0 += 4 * kPassStatsThreads) {
      uint32_t t[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int x = x0 + u * kPassStatsThreads;
        t[u] = x < n ? r[(int64_t)x * kStatCols] : 0u;
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) v += t[u];
    }
  }
  v = wave_sum(v);
  if (lane_id() == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long tot = 0;
    for (int w = 0; w < kPassStatsThreads / 64; ++w) tot += sh[w];
    D.ctr->stats[st] = tot;
  }
}

// Zeroes the partial slots of the statistics selected by mask.
__global__ void __launch_bounds__(256) k_stats_zero(Dev D, uint32_t mask) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < ST_COUNT * kPartSlots; i += gridDim.x * blockDim.x)
    if (mask & (1u << (i / kPartSlots))) D.part[i] = 0;
}

// EngCtr::stats[s] = sum of its partial slots, for the statistics in mask;
// one workgroup per statistic.
__global__ void __launch_bounds__(256) k_stats_reduce(Dev D, uint32_t mask) {
  __shared__ unsigned long long sh[4];
  const int st = blockIdx.x;
  if (st >= ST_COUNT || !(mask & (1u << st))) return;
  const unsigned long long* p = D.part + st * kPartSlots;
  unsigned long long v = 0;
#pragma unroll
  for (int j = 0; j < kPartSlots / 256; ++j) v += p[threadIdx.x + 256 * j];
  v = wave_sum(v);
  if (lane_id() == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) D.ctr->stats[st] = sh[0] + sh[1] + sh[2] + sh[3];
}

// ---- spatial hash ---------------------------------------------------------------
// Find-or-insert of a packed block key in the map's chain of buckets.
// Lock-free: the whole chain is searched first; if the key is absent the
// first empty slot (in chain order) is claimed by CAS.  Slots only go from
// empty to full while binning runs (keys leave the table only afterwards:
// in gc, and for entries that got no heap block), so every inserter of one
// key walks the same sequence of empty slots and the first CAS winner is
// seen by the others: no duplicates.
// A full chain is extended at its tail: the link is claimed first (CAS -1 ->
// kLinkPending), so exactly one inserter pops an overflow bucket and
// publishes it (or kLinkExhausted when the pool is empty); the others wait
// for the published link and continue there.  An inserter therefore fails
// only when no inserter of any key could extend that chain (chain limit or
// empty pool), never because another one was extending it at that moment.
// Returns the global entry index, or -1 on such a failure.  *inserted is set
// when this call created the entry.
// Reads are plain loads: a stale read can only show an older value of a slot
// (empty instead of filled) or of a link (none instead of set); the CAS that
// every insert and extension goes through then fails and reports the current
// value, so staleness costs a retry, never a duplicate.
constexpr int kLinkNone = -1;       // chain ends here
constexpr int kLinkPending = -2;    // an inserter is extending the chain here
constexpr int kLinkExhausted = -3;  // the overflow pool was empty (reset by gc)
constexpr int kLinkSpin = 1 << 16;  // bound on waiting for a pending link
__device__ inline int find_or_insert(const Dev& D, const MapDev& m, unsigned long long key, bool& inserted) {
  inserted = false;
  int g = m.bucket_base + (int)(hash_key(key) & m.home_mask);
  int s0 = 0;
  int depth = 0;  // overflow buckets between the home bucket and bucket g
  for (int guard = 0; guard < 4096; ++guard) {
    // search from (g, s0) to the end of the chain, noting the first empty slot
    int fe = -1, fe_depth = 0;
    int cg = g, cs = s0, cd = depth, tail = g, tail_depth = depth;
    for (;;) {
      // one 64-byte bucket in four 16-byte loads; plain loads may see an
      // older (empty) value of a slot another wave just filled, which only
      // costs a failed CAS below
      const ulonglong2* bk = reinterpret_cast<const ulonglong2*>(D.keys + (int64_t)cg * kBucket);
      const ulonglong2 v0 = bk[0], v1 = bk[1], v2 = bk[2], v3 = bk[3];
      const int nx = D.next[cg];
      const unsigned long long kk[kBucket] = {v0.x, v0.y, v1.x, v1.y, v2.x, v2.y, v3.x, v3.y};
#pragma unroll
      for (int s = 0; s < kBucket; ++s) {
        if (s < cs) continue;
        if (kk[s] == key) return cg * kBucket + s;
        if (kk[s] == kEmptyKey && fe < 0) {
          fe = cg * kBucket + s;
          fe_depth = cd;
        }
      }
      if (nx < 0) {
        tail = cg;
        tail_depth = cd;
        break;
      }
      cg = nx;
      cs = 0;
      if (++cd > m.n_ovf) return -1;  // a chain never holds more than the pool
    }
    if (fe >= 0) {
      const unsigned long long old = atomicCAS(&D.keys[fe], kEmptyKey, key);
      if (old == kEmptyKey) {
        inserted = true;
        return fe;
      }
      if (old == key) return fe;
      // lost the slot to another key: resume the search just after it
      g = fe / kBucket;
      s0 = fe % kBucket + 1;
      depth = fe_depth;
      continue;
    }
    // no empty slot: extend the chain at its tail
    if (tail_depth >= m.max_chain) return -1;
    int link = atomicCAS(&D.next[tail], kLinkNone, kLinkPending);
    if (link == kLinkNone) {
      // this lane owns the extension: pop a bucket and publish the link
      MapCtr* mc = &D.mctr[m.slot];
      const int idx = atomicSub(&mc->ovf_top, 1) - 1;
      if (idx < 0) {
        atomicAdd(&mc->ovf_top, 1);
        link = kLinkExhausted;
      } else {
        link = D.ovf_stack[m.bucket_base + m.n_home + idx];
      }
      atomicExch(&D.next[tail], link);
    }
    // another lane's extension: wait until its link is published
    for (int w = 0; link == kLinkPending && w < kLinkSpin; ++w) link = ld_acq(&D.next[tail]);
    if (link < 0) return -1;
    g = link;
    s0 = 0;
    depth = tail_depth + 1;
  }
  return -1;
}

// ---- pass kernels -----------------------------------------------------------------
// A pass of at most kInlineScans scans brings its header and scan
// descriptors as k_bin arguments; k_bin's first workgroup stores them for
// the pass's later kernels, so no host-to-device copy and no separate launch
// precede the pass (a short first launch left the GPU idle while the host
// issued k_bin).  nscans < 0: the descriptors were copied into place already
// (larger batches, and graph replays, whose captured arguments would be
// stale).
// The pass counters need no launch to clear them either: each is cleared
// after its last reader, by one thread of a later kernel (touched_n, new_n,
// fail_n and bin_next by k_emit, the fold counters by the next pass's
// k_bin), and pass statistics are per-workgroup rows that every pass
// overwrites.
constexpr int kInlineScans = 32;
struct PassInline {
  PassHdr hdr;
  int32_t nscans;
  ScanDev scans[kInlineScans];
};

// The scan of pass point i, from the inline descriptors or the stored ones.
__device__ __forceinline__ ScanDev scan_of(const Dev& D, const PassInline& inl, int ns, int i) {
  int lo = 0, hi = ns - 1;
  if (inl.nscans >= 0) {
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (inl.scans[mid].off <= i) lo = mid;
      else hi = mid - 1;
    }
    return inl.scans[lo];
  }
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (D.scans[mid].off <= i) lo = mid;
    else hi = mid - 1;
  }
  return D.scans[lo];
}

__device__ __forceinline__ int find_scan(const ScanDev* scans, int ns, int i) {
  int lo = 0, hi = ns - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (scans[mid].off <= i) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Heap blocks for new entries (free-list pop; pushes only happen in gc,
// never concurrently with pops, so the stack needs no ABA guard).  All
// lanes of the wave call this; lanes with active == false only take part in
// the ballots.  Both counters are updated once per wave: the per-map block
// count once per distinct map in the wave, the heap stack once; lanes past
// a quota or past the bottom of the stack hand their share back.  An entry
// that gets no block is filed on the fail list (its key leaves the table
// after binning) and its vals[] is -1.
// Returns kAssignNone for inactive lanes, else kAssignOk, kAssignHeapFail or
// kAssignQuotaFail.
constexpr int kAssignNone = -1, kAssignOk = 0, kAssignHeapFail = 1, kAssignQuotaFail = 2;
__device__ __forceinline__ int assign_heap(const Dev& D, bool active, uint32_t e, int ms) {
  const int lane = lane_id();
  const unsigned long long below = (1ull << lane) - 1ull;
  int held = 0;  // blocks the map held before this lane's block
  unsigned long long pending = __ballot(active);
  while (pending) {
    const int leader = __ffsll((long long)pending) - 1;
    const int lm = __builtin_amdgcn_readlane(ms, leader);
    const unsigned long long grp = __ballot(active && ms == lm);
    int base = 0;
    if (lane == leader) base = atomicAdd(&D.mctr[lm].blocks, (int)__popcll(grp));
    base = __builtin_amdgcn_readlane(base, leader);
    if (grp >> lane & 1ull) held = base + (int)__popcll(grp & below);
    pending &= ~grp;
  }
  bool ok = active;
  bool quota_fail = false;
  if (active) {
    const int64_t quota = D.maps[ms].block_quota;
    if (quota > 0 && held >= quota) {
      atomicSub(&D.mctr[ms].blocks, 1);
      ok = false;
      quota_fail = true;
    }
  }
  const unsigned long long want = __ballot(ok);
  const int cnt = (int)__popcll(want);
  const int leader = want ? __ffsll((long long)want) - 1 : 0;
  int top = 0;
  if (want && lane == leader) {
    top = atomicSub(&D.ctr->heap_top, cnt);
    const int failed = min(cnt, cnt - max(top, 0));
    if (failed > 0) atomicAdd(&D.ctr->heap_top, failed);
  }
  top = __builtin_amdgcn_readlane(top, leader);
  const int idx = top - cnt + (int)__popcll(want & below);
  if (ok && idx < 0) {
    atomicSub(&D.mctr[ms].blocks, 1);
    ok = false;
  }
  if (!active) return kAssignNone;
  if (ok) {
    D.vals[e] = D.heap_free[idx];
    return kAssignOk;
  }
  D.vals[e] = -1;
  D.fail_list[atomicAdd(&D.ctr->fail_n, 1)] = e;
  return quota_fail ? kAssignQuotaFail : kAssignHeapFail;
}

// Binning: per point, the truncation segment, the block DDA (every cell is
// allocated, H:255-272 and H:421-500) and the voxel DDA (H:675-710) walked
// for its cells only.  Along a walk every axis is monotone, so the voxels of
// one block form one run.  Neighbouring rays share blocks, so each workgroup
// first gathers its points' blocks in an LDS table: a block is looked up (or
// inserted) in the map's spatial hash once per workgroup, and reserves the
// workgroup's records in the block with one atomic.  Each run then learns
// its record offset in the block, stored with its entry for k_emit.  A full
// LDS table falls back to per-run global atomics.
constexpr int kBinSlots = 1024;
constexpr int kBlkReg = 4;  // block-walk cells a lane keeps in registers (MAX_BLK of the configs)
constexpr int kPairLds = 4;  // runs per ray whose pairs k_bin keeps in LDS
constexpr uint32_t kPairFail = 0xFFFFFFFFu;
constexpr uint32_t kPairDirect = 0x80000000u;

__device__ __forceinline__ unsigned long long lds_block_key(int map_slot, const int32_t kb[3]) {
  // 16 bits per axis suffice under the point validity rule (|block| < 2^14)
  return ((unsigned long long)(uint32_t)map_slot << 48) | ((unsigned long long)(kb[0] & 0xFFFF) << 32) |
         ((unsigned long long)(kb[1] & 0xFFFF) << 16) | (unsigned long long)(kb[2] & 0xFFFF);
}

// Slot of an LDS-table key: the 48 coordinate bits and the map folded into
// 32 bits and mixed with one multiply (the table only spreads the few
// hundred blocks of one workgroup, so a light hash suffices).
__device__ __forceinline__ uint32_t lds_hash(unsigned long long k) {
  const uint32_t x = (uint32_t)k ^ (uint32_t)(k >> 29) ^ (uint32_t)(k >> 48) * 0x9E3779B1u;
  return (x * 0x85EBCA6Bu) >> (32 - 10);
}
static_assert(kBinSlots == 1024, "lds_hash yields 10 bits");

// Slot of key k in the workgroup's table, inserting it when absent (-1 when
// the table is full).  *added is set when this call inserted it.  Probes
// follow the triangular sequence h, h+1, h+3, h+6, ... (every slot of the
// power-of-two table once), which breaks up the runs that the blocks of
// neighbouring rays form under linear probing.  Each probe reads the slot
// first: most lookups find their key, and many lanes reading one slot is a
// broadcast where a compare-and-swap would serialise them.
__device__ __forceinline__ int lds_find_or_add(unsigned long long* tkey, unsigned long long k, bool* added = nullptr) {
  uint32_t h = lds_hash(k);
  for (int probe = 0; probe < 64; ++probe) {
    const unsigned long long cur = tkey[h];
    if (cur == k) return (int)h;
    if (cur == kEmptyKey) {
      const unsigned long long old = atomicCAS(&tkey[h], kEmptyKey, k);
      if (old == kEmptyKey && added) *added = true;
      if (old == kEmptyKey || old == k) return (int)h;
    }
    h = (h + (uint32_t)probe + 1u) & (kBinSlots - 1);
  }
  return -1;
}
// The same lookup for the block walks, which mostly insert: each probe is
// one compare-and-swap, whose returned value also answers a lookup (one LDS
// round trip per probe instead of a read and then the swap).
__device__ __forceinline__ int lds_add(unsigned long long* tkey, unsigned long long k, bool* added) {
  uint32_t h = lds_hash(k);
  for (int probe = 0; probe < 64; ++probe) {
    const unsigned long long old = atomicCAS(&tkey[h], kEmptyKey, k);
    if (old == kEmptyKey) {
      *added = true;
      return (int)h;
    }
    if (old == k) return (int)h;
    h = (h + (uint32_t)probe + 1u) & (kBinSlots - 1);
  }
  return -1;
}

// The packed map key and map slot of an LDS-table key.
__device__ __forceinline__ unsigned long long table_map_key(unsigned long long k, int& map_slot) {
  map_slot = (int)(k >> 48);
  const int32_t k0 = (int32_t)(int16_t)((k >> 32) & 0xFFFF), k1 = (int32_t)(int16_t)((k >> 16) & 0xFFFF),
                k2 = (int32_t)(int16_t)(k & 0xFFFF);
  return pack_key(k0, k1, k2);
}

// The home bucket of a key, read ahead of its lookup: its index and keys.
struct HomePeek {
  int g;
  ulonglong2 v0, v1, v2, v3;
};
__device__ __forceinline__ HomePeek peek_home(const Dev& D, const MapDev& m, unsigned long long key) {
  HomePeek h;
  h.g = m.bucket_base + (int)(hash_key(key) & m.home_mask);
  const ulonglong2* bk = reinterpret_cast<const ulonglong2*>(D.keys + (int64_t)h.g * kBucket);
  h.v0 = bk[0];
  h.v1 = bk[1];
  h.v2 = bk[2];
  h.v3 = bk[3];
  return h;
}
// Entry of key when the peeked bucket holds it, else -1.  Slots only go
// from empty to full while binning runs, so a key seen in an earlier read
// of the bucket is still there; one not seen takes the full lookup.
__device__ __forceinline__ int peek_find(const HomePeek& h, unsigned long long key) {
  const unsigned long long kk[kBucket] = {h.v0.x, h.v0.y, h.v1.x, h.v1.y, h.v2.x, h.v2.y, h.v3.x, h.v3.y};
  int e = -1;
#pragma unroll
  for (int s = kBucket - 1; s >= 0; --s)
    if (kk[s] == key) e = h.g * kBucket + s;
  return e;
}

// How k_bin's workgroups take the pass's rays.  A chunk is rays_per_lane x
// blockDim consecutive points, which share one LDS block table (more rays
// per table: more of their blocks deduplicated before the global lookups
// and record reservations, fewer workgroups).  Workgroup w takes chunk w,
// then chunk w + gridDim (queue == 0), or the next one from the pass's work
// queue (queue == 1: a grid of resident workgroups, persistent for the
// launch, each taking chunks until none is left).
constexpr int kMaxRaysPerLane = 4;
struct BinSched {
  int32_t rays_per_lane;  // 1..kMaxRaysPerLane
  int32_t queue;
};

// 168 VGPRs at most: three waves per SIMD where the grid has them (batched
// passes; a 131k-point scan launches two per SIMD whatever the budget), with
// a handful of spills off the walks.  kMulti: rays_per_lane may exceed 1
// (one ray per lane compiles without the per-ray loops and their registers).
template <bool kMulti>
__global__ void __launch_bounds__(256, 3) k_bin(Dev D, PassInline inl, BinSched bs) {
  __shared__ unsigned long long tkey[kBinSlots];
  __shared__ int tent[kBinSlots];
  __shared__ uint32_t trec[kBinSlots];
  __shared__ uint32_t tbase[kBinSlots];
  __shared__ uint32_t lnew[kBinSlots];
  __shared__ uint8_t lnew_map[kBinSlots];
  __shared__ uint32_t ltouch[kBinSlots];
  __shared__ uint8_t ltouch_map[kBinSlots];
  __shared__ uint16_t locc[kBinSlots];
  // one ray per lane: the (entry or slot, record offset) pairs of a ray's
  // first kPairLds runs, kept in LDS from the voxel walk to phase C; later
  // runs go through the pair rows in global memory
  __shared__ uint2 lpair[kMulti ? 1 : kPairLds][kMulti ? 1 : 256];
  __shared__ int n_new, n_touch, g_touch, g_new, n_occ, next_chunk;
  const PassHdr hdr = inl.nscans >= 0 ? inl.hdr : *D.hdr;
  const int P = hdr.total;
  if (blockIdx.x == 0) {
    // the previous pass's fold is over: clear its counters; store the inline
    // descriptors for k_emit
    const int t = threadIdx.x;
    if (t == 0) {
      D.ctr->fold_next = 0;
      D.ctr->fail_n = 0;  // (its last reader, k_fold_b, is over)
      D.ctr->seg_top = 0;  // (its last user, k_seg, is over)
    }
    if (t < FC_COUNT) D.ctr->fold_n[t] = 0;
    if (inl.nscans >= 0) {
      if (t == 0) *const_cast<PassHdr*>(D.hdr) = inl.hdr;
      for (int k = t; k < inl.nscans; k += blockDim.x) const_cast<ScanDev*>(D.scans)[k] = inl.scans[k];
    }
  }
  // per-thread counts: ST_VALID, ST_SKIPPED, ST_BLOCK_CELLS, ST_UPDATES,
  // ST_TRUNC_VOX, ST_TRUNC_BLK, ST_INSERT_FAIL, ST_UPD_DROPPED (the columns of
  // the SR_BIN rows)
  uint32_t st[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  // a new entry joins the workgroup's list, which goes to the pass's
  // new-entry list with one atomic at the end of the workgroup's points
  // (heap blocks are assigned by k_seg), or joins the pass's list at
  // once when the workgroup's list is full
  auto note_new = [&](int e, int ms) {
    const int at = atomicAdd(&n_new, 1);
    if (at < kBinSlots) {
      lnew[at] = (uint32_t)e;
      lnew_map[at] = ms;
    } else {
      const int g = atomicAdd(&D.ctr->new_n, 1);
      D.new_list[g] = (uint32_t)e;
      D.new_map[g] = (uint8_t)ms;
    }
  };
  // global lookup of table slot x (phase B)
  auto resolve_slot = [&](int x, int e) {
    const unsigned long long k = tkey[x];
    int ms;
    const unsigned long long key = table_map_key(k, ms);
    if (e < 0) {
      bool ins;
      e = find_or_insert(D, D.maps[ms], key, ins);
      if (e >= 0 && ins) note_new(e, ms);
    }
    tent[x] = e;
    if (e < 0) {
      st[6]++;
      return;
    }
    const uint32_t nrec = trec[x];
    if (nrec > 0) {
      const unsigned long long old = atomicAdd(&D.blk_cnt[e], (unsigned long long)nrec << 32);
      tbase[x] = (uint32_t)(old >> 32);
      if (old == 0) {
        const int at = atomicAdd(&n_touch, 1);
        ltouch[at] = (uint32_t)e;
        ltouch_map[at] = (uint8_t)ms;
      }
    }
  };
  const int64_t stride = D.pcap;
  const int R = kMulti ? bs.rays_per_lane : 1;
  const int chunk = R * (int)blockDim.x;
  const int nchunks = (P + chunk - 1) / chunk;
  for (int c = blockIdx.x; c < nchunks;) {
    const int base = c * chunk;
    for (int x = threadIdx.x; x < kBinSlots; x += blockDim.x) {
      tkey[x] = kEmptyKey;
      trec[x] = 0;
    }
    if (threadIdx.x == 0) {
      n_new = 0;
      n_touch = 0;
      n_occ = 0;
    }
    __syncthreads();
    // ---- phase A1: block walks; blocks gathered in the LDS table ----
    // The ray of sub-step r is point base + r * blockDim + threadIdx.x.  Its
    // scan, map, segment and validity come from setup(); the last ray's stay
    // in registers for its voxel walk, the others' are set up again there.
    bool walk = false;
    ScanDev sd;
    MapDev m;
    float p[3], s[3], a[3], b[3];
    auto setup = [&](int i) {
      sd = scan_of(D, inl, hdr.nscans, i);
      m = D.maps[sd.map];
      const float4 q = hdr.points[i];
      p[0] = q.x;
      p[1] = q.y;
      p[2] = q.z;
      s[0] = sd.s[0];
      s[1] = sd.s[1];
      s[2] = sd.s[2];
      walk = point_valid(p, s, m.vs);
      if (walk) truncation_segment(p, s, m.T, a, b);
    };
    // the last ray's first kBlkReg block-walk cells and their table slots
    // (-1: not in the table), so that its voxel walk's runs find their slots
    // without probing the table
    int32_t bk0[kBlkReg], bk1[kBlkReg], bk2[kBlkReg];
    int bsl[kBlkReg];
    for (int r = 0; r < R; ++r) {
      const int i = base + r * (int)blockDim.x + (int)threadIdx.x;
#pragma unroll
      for (int u = 0; u < kBlkReg; ++u) {
        bk0[u] = bk1[u] = bk2[u] = 0;
        bsl[u] = -1;
      }
      int nbc = 0;
      walk = false;
      if (i < P) {
        setup(i);
        if (!walk) {
          st[1]++;
        } else {
          st[0]++;
          bool tb = false;
          const int nb = walk_cells(a, b, m.BL, m.sc_b, m.isc_b, m.half_BL, m.max_blk, tb, [&](int32_t k0, int32_t k1, int32_t k2) {
            const int32_t kb[3] = {k0, k1, k2};
            bool added = false;
            const int slot = lds_add(tkey, lds_block_key(sd.map, kb), &added);
            if (added) locc[atomicAdd(&n_occ, 1)] = (uint16_t)slot;  // the table's list
            if (slot < 0) {
              bool ins;
              const int e = find_or_insert(D, m, pack_key(k0, k1, k2), ins);
              if (e < 0) st[6]++;
              else if (ins) note_new(e, m.slot);
            }
#pragma unroll
            for (int u = 0; u < kBlkReg; ++u)
              if (u == nbc) {
                bk0[u] = k0;
                bk1[u] = k1;
                bk2[u] = k2;
                bsl[u] = slot;
              }
            ++nbc;
            return true;
          });
          st[2] += nb;
          st[5] += tb ? 1 : 0;
        }
      }
    }
    __syncthreads();
    // ---- the blocks the block walks listed: their home buckets are read now
    // and used after the voxel walks, which hide the memory latency ----
    const int n_listed = n_occ;
    // (one bucket per thread: a second one held through the walks would cost
    // the registers of the third wave per SIMD)
    HomePeek pk0;
    unsigned long long key0 = kEmptyKey;
    if ((int)threadIdx.x < n_listed) {
      int ms;
      key0 = table_map_key(tkey[locc[threadIdx.x]], ms);
      pk0 = peek_home(D, D.maps[ms], key0);
    }
    // ---- phase A2: voxel walks, the last ray's first ----
    int njr[kMaxRaysPerLane];
    for (int q = 0; q < R; ++q) {
      const int r = q == 0 ? R - 1 : q - 1;
      const int i = base + r * (int)blockDim.x + (int)threadIdx.x;
      if (q == 1) {
        // the cached cells were the last ray's
#pragma unroll
        for (int u = 0; u < kBlkReg; ++u) bsl[u] = -1;
      }
      if (q > 0) {
        walk = false;
        if (i < P) setup(i);
      }
      int nj = 0;
      if (walk) {
        // A run (the voxels of one block, consecutive along the walk) gets
        // its record offset in the block's LDS slot, or, past a full table,
        // straight from the global counter.  A block the block walk did not
        // list joins the table's list here.
        auto close_run = [&](const int32_t kb[3], int rn) {
          int slot = -1;
#pragma unroll
          for (int u = 0; u < kBlkReg; ++u)
            if (bsl[u] >= 0 && bk0[u] == kb[0] && bk1[u] == kb[1] && bk2[u] == kb[2]) slot = bsl[u];
          bool added = false;
          if (slot < 0) slot = lds_find_or_add(tkey, lds_block_key(sd.map, kb), &added);
          auto put = [&](uint32_t pe, uint32_t pr) {
            if (!kMulti && nj < kPairLds) {
              lpair[kMulti ? 0 : nj][threadIdx.x] = make_uint2(pe, pr);
            } else {
              D.pair_e[(int64_t)nj * stride + i] = pe;
              D.pair_r[(int64_t)nj * stride + i] = pr;
            }
          };
          if (slot >= 0) {
            if (added) locc[atomicAdd(&n_occ, 1)] = (uint16_t)slot;
            const uint32_t local = atomicAdd(&trec[slot], (uint32_t)rn);
            put((uint32_t)slot, local | ((uint32_t)rn << 24));
          } else {
            bool ins;
            const int e = find_or_insert(D, m, pack_key(kb[0], kb[1], kb[2]), ins);
            if (e < 0) {
              st[6]++;
              st[7] += rn;
              put(kPairFail, 0u);
            } else {
              if (ins) note_new(e, m.slot);
              const unsigned long long old = atomicAdd(&D.blk_cnt[e], (unsigned long long)rn << 32);
              if (old == 0) {
                const int t = atomicAdd(&D.ctr->touched_n, 1);
                D.touched[t] = (uint32_t)e;
                D.touched_map[t] = (uint8_t)m.slot;
              }
              put(kPairDirect | (uint32_t)e, (uint32_t)(old >> 32));
              st[3] += rn;
            }
          }
          ++nj;
        };
        // The voxel walk notes only which cells enter a new block (walk_runs);
        // the table work for its runs follows the walk, so the lanes of a wave
        // do it together rather than at scattered steps of their walks.
        bool tv = false;
        walk_runs(a, b, m.vs, m.sc_v, m.isc_v, m.half_vs, m.max_vox, tv, [&](int32_t b0, int32_t b1, int32_t b2, int rn) {
          const int32_t kb[3] = {b0, b1, b2};
          close_run(kb, rn);
        });
        st[4] += tv ? 1 : 0;
      }
      if (i < P) D.pair_n[i] = (uint8_t)nj;
#pragma unroll
      for (int u = 0; u < kMaxRaysPerLane; ++u)
        if (u == r) njr[u] = nj;
    }
    __syncthreads();
    // ---- phase B: one global lookup and one counter atomic per listed block
    // (the first per thread from the bucket read ahead) ----
    for (int y = threadIdx.x, rr = 0; y < n_occ; y += blockDim.x, ++rr) {
      int e = -1;
      if (y < n_listed && rr == 0) e = peek_find(pk0, key0);
      resolve_slot(locc[y], e);
    }
    __syncthreads();
    const int nn = min(n_new, kBinSlots);
    // The workgroup's places in the pass's touched and new-entry lists: the
    // atomics are issued now and their results read after phase C, which
    // does not need them, so their round trip overlaps it.
    int gt = 0, gn = 0, nc = 0;
    if (threadIdx.x == 0) {
      gt = n_touch ? atomicAdd(&D.ctr->touched_n, n_touch) : 0;
      gn = nn ? atomicAdd(&D.ctr->new_n, nn) : 0;
      nc = bs.queue ? (int)gridDim.x + atomicAdd(&D.ctr->bin_next, 1) : c + (int)gridDim.x;
    }
    // ---- phase C: final (entry, record offset) of every run ----
    for (int r = 0; r < R; ++r) {
      const int i = base + r * (int)blockDim.x + (int)threadIdx.x;
      int nj = 0;
#pragma unroll
      for (int u = 0; u < kMaxRaysPerLane; ++u)
        if (u == r) nj = njr[u];
      if (i >= P) continue;
      for (int j = 0; j < nj; ++j) {
        const bool inl = !kMulti && j < kPairLds;
        uint32_t v, pr;
        if (inl) {
          const uint2 q = lpair[kMulti ? 0 : j][threadIdx.x];
          v = q.x;
          pr = q.y;
        } else {
          v = D.pair_e[(int64_t)j * stride + i];
          pr = D.pair_r[(int64_t)j * stride + i];
        }
        uint32_t fe, fr = pr;
        if (v == kPairFail || (v & kPairDirect)) {
          fe = v == kPairFail ? kPairFail : v & ~kPairDirect;
        } else {
          const uint32_t rn = pr >> 24;
          const int e = tent[v];
          if (e < 0) {
            fe = kPairFail;
            st[7] += rn;
          } else {
            fe = (uint32_t)e;
            fr = tbase[v] + (pr & 0xFFFFFFu);
            st[3] += rn;
          }
        }
        D.pair_e[(int64_t)j * stride + i] = fe;
        if (fe != kPairFail) D.pair_r[(int64_t)j * stride + i] = fr;
      }
    }
    if (threadIdx.x == 0) {
      g_touch = gt;
      g_new = gn;
      next_chunk = nc;
    }
    __syncthreads();
    for (int x = threadIdx.x; x < n_touch; x += blockDim.x) {
      D.touched[g_touch + x] = ltouch[x];
      D.touched_map[g_touch + x] = ltouch_map[x];
    }
    for (int x = threadIdx.x; x < nn; x += blockDim.x) {
      D.new_list[g_new + x] = lnew[x];
      D.new_map[g_new + x] = lnew_map[x];
    }
    c = next_chunk;
    __syncthreads();
  }
  write_stat_row(D, SR_BIN, st);
}

// ---- exclusive scan over a device-sized list -------------------------------------
// count(i) is read through a functor; out[i] = sum of counts before i,
// out[n] = total.  Three kernels: tile sums, scan of tile sums, tile scan.
struct CountTouchedRecords {
  const uint32_t* list;
  const unsigned long long* cnt;
  __device__ uint32_t operator()(int i) const { return (uint32_t)(cnt[list[i]] >> 32); }
};
struct CountPlain {
  const uint32_t* cnt;
  __device__ uint32_t operator()(int i) const { return cnt[i]; }
};

// Workgroup exclusive scans (up to 1024 threads, wave64): inclusive wave
// scan, then a scan of the wave totals.  sh holds kScanSh entries: the
// waves' offsets, then the total.
constexpr int kScanSh = 17;
// NW: the workgroup's waves when known at compile time (0: from blockDim).
template <int NW = 0>
__device__ __forceinline__ uint32_t block_scan_excl(uint32_t v, uint32_t* sh, uint32_t& total) {
  const int lane = lane_id();
  const int w = threadIdx.x >> 6;
  const uint32_t x = wave_incl_scan(v);
  if (lane == 63) sh[w] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t acc = 0;
    const int nw = NW > 0 ? NW : (int)(blockDim.x >> 6);
#pragma unroll
    for (int k = 0; k < (NW > 0 ? NW : 16); ++k) {
      if (NW == 0 && k >= nw) break;
      const uint32_t t = sh[k];
      sh[k] = acc;
      acc += t;
    }
    sh[kScanSh - 1] = acc;
  }
  __syncthreads();
  total = sh[kScanSh - 1];
  const uint32_t r = x - v + sh[w];
  __syncthreads();
  return r;
}

__device__ __forceinline__ unsigned long long block_scan_excl64(unsigned long long v, unsigned long long* sh,
                                                                 unsigned long long& total) {
  const int lane = lane_id();
  const int w = threadIdx.x >> 6;
  const unsigned long long x = wave_incl_scan64(v);
  if (lane == 63) sh[w] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long acc = 0;
    for (int k = 0; k < (int)(blockDim.x >> 6); ++k) {
      const unsigned long long t = sh[k];
      sh[k] = acc;
      acc += t;
    }
    sh[kScanSh - 1] = acc;
  }
  __syncthreads();
  total = sh[kScanSh - 1];
  const unsigned long long r = x - v + sh[w];
  __syncthreads();
  return r;
}

template <class Count>
__global__ void __launch_bounds__(256) k_scan_tiles(const int32_t* n_ptr, Count cnt, uint32_t* tile_sums) {
  __shared__ uint32_t sh[kScanSh];
  const int n = ld_acq(n_ptr);
  const int ntiles = (n + kScanTile - 1) / kScanTile;
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    uint32_t v = 0;
    for (int k = 0; k < kScanTile / 256; ++k) {
      const int i = t * kScanTile + k * 256 + threadIdx.x;
      if (i < n) v += cnt(i);
    }
    uint32_t tot;
    block_scan_excl(v, sh, tot);
    if (threadIdx.x == 0) tile_sums[t] = tot;
  }
}

__global__ void __launch_bounds__(256) k_scan_top(const int32_t* n_ptr, uint32_t* tile_sums) {
  __shared__ uint32_t sh[kScanSh];
  const int n = ld_acq(n_ptr);
  const int ntiles = (n + kScanTile - 1) / kScanTile;
  uint32_t carry = 0;
  for (int base = 0; base < ntiles; base += 256) {
    const int i = base + threadIdx.x;
    const uint32_t v = i < ntiles ? tile_sums[i] : 0;
    uint32_t tot;
    const uint32_t ex = block_scan_excl(v, sh, tot);
    if (i < ntiles) tile_sums[i] = carry + ex;
    carry += tot;
  }
  if (threadIdx.x == 0) tile_sums[ntiles] = carry;
}

template <class Count>
__global__ void __launch_bounds__(256) k_scan_apply(const int32_t* n_ptr, Count cnt, const uint32_t* tile_sums,
                                                     uint32_t* out) {
  __shared__ uint32_t sh[kScanSh];
  const int n = ld_acq(n_ptr);
  const int ntiles = (n + kScanTile - 1) / kScanTile;
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    // each thread owns 4 consecutive elements
    const int i0 = t * kScanTile + threadIdx.x * 4;
    uint32_t c[4];
    uint32_t v = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      c[k] = (i0 + k < n) ? cnt(i0 + k) : 0u;
      v += c[k];
    }
    uint32_t tot;
    uint32_t ex = block_scan_excl(v, sh, tot) + tile_sums[t];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      if (i0 + k < n) out[i0 + k] = ex;
      ex += c[k];
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) out[n] = tile_sums[ntiles];
}

// The pass scan, one launch after binning:
//  * heap blocks for the pass's new entries, 256 per workgroup and chunk
//    (taken from the last workgroup down, so the tile workgroups get them
//    last); off the free stack with one pop, and one per-map count, per wave;
//  * record segments and fold descriptors of the touched blocks, per tile of
//    the touched list: a workgroup scan of the tile's record counts, and one
//    atomic on the pass's segment counter for the tile's total, which places
//    the tile's segments after those of the tiles served before it (in any
//    order of tiles: segments need only be disjoint, and each descriptor
//    holds its block's record base).  Each block's 16-byte descriptor
//    {record base, records, entry, map} is filed under its size class with
//    one counter atomic per class and workgroup, and its pass counter is
//    reset.  The descriptor holds the entry, not its heap block: the heap
//    blocks of new entries are being assigned beside it, and the fold reads
//    them.
__global__ void __launch_bounds__(256) k_seg(Dev D) {
  __shared__ uint32_t sh[kScanSh];
  __shared__ int lcnt[FC_COUNT], gbase[FC_COUNT];
  __shared__ uint32_t carry_sh;
  const int n = ld_acq(&D.ctr->touched_n);
  const int nn = ld_acq(&D.ctr->new_n);  // (read with the other counts: one round trip)
  // heap failures include quota failures, which are also counted apart
  // (columns of the SR_SCAN rows: ST_NEW_BLOCKS, ST_HEAP_FAIL, ST_QUOTA_FAIL)
  uint32_t st[3] = {0, 0, 0};
  const int nchunks = (nn + 255) / 256;
  for (int c = gridDim.x - 1 - blockIdx.x; c < nchunks; c += gridDim.x) {
    const int x = c * 256 + threadIdx.x;
    const bool active = x < nn;
    const int r = assign_heap(D, active, active ? D.new_list[x] : 0u, active ? (int)D.new_map[x] : -1);
    st[0] += r == kAssignOk ? 1 : 0;
    st[1] += (r == kAssignHeapFail || r == kAssignQuotaFail) ? 1 : 0;
    st[2] += r == kAssignQuotaFail ? 1 : 0;
  }
  const int ntiles = (n + kScanTile - 1) / kScanTile;
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    // the tile's loads first: touched entries, then their counters and maps
    const int i0 = t * kScanTile + threadIdx.x * 4;
    uint32_t e4[4], cn[4];
    uint8_t mp4[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) e4[k] = i0 + k < n ? D.touched[i0 + k] : 0u;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      cn[k] = 0;
      mp4[k] = 0;
      if (i0 + k < n) {
        cn[k] = (uint32_t)(D.blk_cnt[e4[k]] >> 32);
        mp4[k] = D.touched_map[i0 + k];
      }
    }
    if (threadIdx.x < FC_COUNT) lcnt[threadIdx.x] = 0;
    uint32_t tot;
    uint32_t ex = block_scan_excl(cn[0] + cn[1] + cn[2] + cn[3], sh, tot);
    // the tile's segments go to the next free records of the pass, in any
    // order of tiles: each block's descriptor holds its own record base
    if (threadIdx.x == 0) carry_sh = tot ? atomicAdd(&D.ctr->seg_top, tot) : 0u;
    __syncthreads();
    ex += carry_sh;
    int cls[4], rank[4];
    int4 dsc[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      cls[k] = -1;
      rank[k] = 0;
      dsc[k] = make_int4(0, 0, 0, 0);
      if (i0 + k < n) {
        D.blk_base[e4[k]] = ex;
        D.blk_cnt[e4[k]] = 0;
        dsc[k] = make_int4((int)ex, (int)cn[k], (int)e4[k], (int)mp4[k]);
        cls[k] = fold_class((int)cn[k]);
        rank[k] = atomicAdd(&lcnt[cls[k]], 1);
      }
      ex += cn[k];
    }
    __syncthreads();
    if (threadIdx.x < FC_COUNT)
      gbase[threadIdx.x] = lcnt[threadIdx.x] ? atomicAdd(&D.ctr->fold_n[threadIdx.x], lcnt[threadIdx.x]) : 0;
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 4; ++k)
      if (cls[k] >= 0) D.fold_desc[D.fold_base[cls[k]] + gbase[cls[k]] + rank[k]] = dsc[k];
    __syncthreads();
  }
  write_stat_row(D, SR_SCAN, st);
}

// Lanes of one wave exchange data through LDS: the wave's LDS requests
// complete in order, so waiting for the wave's own LDS traffic and keeping
// the compiler from moving memory accesses across the point is enough; no
// global-memory fence (and no wait on outstanding stores) is involved.
__device__ __forceinline__ void wave_sync() {
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0)
  __builtin_amdgcn_wave_barrier();
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

// ---- records ---------------------------------------------------------------------------
// Record of one voxel update: local voxel (9 bits) | point index in the pass
// (23 bits) | FP32 signed distance.  Sorting records as integers orders them
// by voxel, then by point: the reference's per-voxel fold order.
__device__ __forceinline__ unsigned long long make_rec(int local, int point, float d) {
  return ((unsigned long long)local << 55) | ((unsigned long long)point << 32) | (unsigned long long)__float_as_uint(d);
}

// Record offset of run r of point i, or -1 when its block could not be stored
// (a pass holds fewer than 2^31 records: kMaxPassPoints times at most 128).
constexpr int kEmitRuns = 6;
__device__ __forceinline__ int32_t run_offset(const Dev& D, int r, int i) {
  const uint32_t e = D.pair_e[(int64_t)r * D.pcap + i];
  return e == kPairFail ? -1 : (int32_t)(D.blk_base[e] + D.pair_r[(int64_t)r * D.pcap + i]);
}

// Second walk of every point's voxel DDA: each voxel's update record
// (getDistanceUpdate, H:562-573) is written to its block's record segment at
// the offset its run reserved in k_bin, so every block's records end up
// contiguous.
// A wave's lanes walk different rays, so the records of one walk step go to
// 64 unrelated places, and a store of them touches 64 cache lines.  The
// records are therefore staged in LDS, kEmitChunk walk steps at a time,
// lane-major: slot lane * kEmitChunk + step.  Consecutive slots are then
// consecutive records of one run (contiguous in the segment), and the wave
// writes the staged slots out 64 at a time, each store covering the runs of
// a few lanes.  Large passes, whose stores bound the kernel, stage
// (kStage); a single scan's pass, two waves per SIMD and bound by latency,
// stores each record as it is made.
constexpr int kEmitChunk = 8;
constexpr int kEmitStagePoints = 1 << 18;  // passes of more points use the staged variant
template <bool kStage>
__global__ void __launch_bounds__(256) k_emit(Dev D) {
  __shared__ unsigned long long srec[kStage ? 4 : 1][kStage ? 64 * kEmitChunk : 1];
  __shared__ int32_t sdst[kStage ? 4 : 1][kStage ? 64 * kEmitChunk : 1];
  const int lane = lane_id();
  unsigned long long* wrec = srec[kStage ? threadIdx.x >> 6 : 0];
  int32_t* wdst = sdst[kStage ? threadIdx.x >> 6 : 0];
  const PassHdr hdr = *D.hdr;
  const int P = hdr.total;
  const int64_t stride = D.pcap;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    // k_seg, their last reader, is over
    D.ctr->touched_n = 0;
    D.ctr->new_n = 0;
    D.ctr->bin_next = 0;  // (k_bin's work queue)
  }
  // run rows that exist in the pair arrays (they hold hdr.stride rows)
  const int rows = min(kEmitRuns, hdr.stride);
  const bool wide = P > kEmitStagePoints;  // a pass of more than two 131k-point scans
  // whole waves walk the grid, so that the lanes of a wave stage and write
  // out their records together
  for (int wbase = blockIdx.x * blockDim.x + (threadIdx.x & ~63); wbase < P; wbase += gridDim.x * blockDim.x) {
    const int i = wbase + lane;
    const bool have = i < P;
    // absolute record offsets of the point's runs, loaded before the walk
    // (the first kEmitRuns in registers; later runs, which need a walk far
    // longer than the truncation band of the workloads, are looked up when
    // reached).  The run rows are requested together with the run count, so
    // the blocks' record bases are one dependent load away -- in a large
    // (batched) pass only the first three rows: a point has 2-3 runs, and
    // the unused rows would be most of the kernel's reads.
    const int nj = have ? D.pair_n[i] : 0;
    uint32_t pe[kEmitRuns], pr[kEmitRuns];
#pragma unroll
    for (int r = 0; r < kEmitRuns; ++r) {
      pe[r] = kPairFail;
      pr[r] = 0;
      if (have && r < rows && (!wide || r < 3 || r < nj)) {
        pe[r] = D.pair_e[(int64_t)r * stride + i];
        pr[r] = D.pair_r[(int64_t)r * stride + i];
      }
    }
    int32_t rq[kEmitRuns];  // record offsets of the current run and the next ones
#pragma unroll
    for (int r = 0; r < kEmitRuns; ++r)
      rq[r] = r < nj && pe[r] != kPairFail ? (int32_t)(D.blk_base[pe[r]] + pr[r]) : -1;
    const int ic = have ? i : 0;
    const ScanDev sd = D.scans[find_scan(D.scans, hdr.nscans, ic)];
    const MapDev m = D.maps[sd.map];
    const float4 q = hdr.points[ic];
    const float p[3] = {q.x, q.y, q.z};
    const float s[3] = {sd.s[0], sd.s[1], sd.s[2]};
    float a[3], b[3];
    truncation_segment(p, s, m.T, a, b);
    // the walk of k_bin's voxel DDA (walk_cells' cells: at most max_vox,
    // the end cell included); a cell that enters a new block starts the next
    // run (runs are in walk order in the pair rows), and the run's record of
    // cell n goes to rq[0] + n - (the run's first cell)
    int j = 0;   // current run
    int rs = 0;  // its first cell
    int n = 0;   // cells so far
    bool alive = nj > 0 && m.max_vox > 0;
    // kProduct: every lane's map forms k / scale by the checked product (a
    // product and two FMAs per axis, no branch); otherwise each axis picks
    // its map's form.  The choice is made once per wave.
    auto walk = [&](auto product) {
      constexpr bool kProduct = decltype(product)::value;
      CellWalk w(a, b, m.vs, m.sc_v, m.isc_v, m.half_vs);
      auto record = [&](unsigned long long& rec) -> int32_t {
        if (rq[0] < 0) return -1;
        float vc[3];
#pragma unroll
        for (int u = 0; u < 3; ++u)
          vc[u] = kProduct ? product_centre(w.k[u], m.sc_v, m.isc_v, m.half_vs)
                           : cell_centre_of(w.k[u], m.sc_v, m.isc_v, m.half_vs);
        rec = make_rec(local_index(w.k), i, signed_distance(p, s, vc));
        return rq[0] + n - rs;
      };
      auto next_run = [&]() {
        ++j;
        rs = n + 1;
#pragma unroll
        for (int r = 0; r + 1 < kEmitRuns; ++r) rq[r] = rq[r + 1];
        rq[kEmitRuns - 1] = j + kEmitRuns - 1 < nj ? run_offset(D, j + kEmitRuns - 1, i) : -1;
      };
      if (!kStage) {
        // each turn: the current cell's record is stored, then the walk moves on
        for (; alive && n < m.max_vox; ++n) {
          const bool last = w.at_end();
          unsigned long long rec;
          const int32_t dst = record(rec);
          if (dst >= 0) D.seg[(uint32_t)dst] = rec;
          bool x[3];
          if (last || !w.advance(x)) break;
          if (x[0] || x[1] || x[2]) next_run();
        }
        return;
      }
      while (__ballot(alive)) {
#pragma unroll
        for (int st = 0; st < kEmitChunk; ++st) {
          int32_t dst = -1;
          unsigned long long rec = 0;
          if (alive) {
            const bool last = w.at_end();
            dst = record(rec);
            bool x[3];
            if (last || !w.advance(x)) alive = false;
            else if (x[0] || x[1] || x[2]) next_run();
            if (++n >= m.max_vox) alive = false;
          }
          wrec[lane * kEmitChunk + st] = rec;
          wdst[lane * kEmitChunk + st] = dst;
        }
        wave_sync();
#pragma unroll
        for (int r = 0; r < kEmitChunk; ++r) {
          const int32_t dst = wdst[r * 64 + lane];
          if (dst >= 0) D.seg[(uint32_t)dst] = wrec[r * 64 + lane];
        }
        wave_sync();
      }
    };
    if (__ballot(m.isc_v == 0.0f) == 0) walk(std::true_type{});
    else walk(std::false_type{});
  }
}

// ---- per-block fold ---------------------------------------------------------------------

__device__ __forceinline__ int next_pow2(int c) {
  int n = 2;
  while (n < c) n <<= 1;
  return n;
}

// H:637-654 with the update's weight w = 1/(|d|+1) and w*d precomputed:
// S = (W*S + w*d)/(W+w); W = min(W+w, maxW).  These are values (the FP64
// oracle bounds them within a tolerance), so the quotients use the hardware
// reciprocal (1 ulp) rather than the correctly rounded division sequence.
__device__ __forceinline__ void fold_step(float& S, float& W, float w, float wd, float maxW) {
  const float Wn = W + w;
  S = (W * S + wd) * __builtin_amdgcn_rcpf(Wn);
  W = fminf(Wn, maxW);
}
__device__ __forceinline__ void weight_of(float d, float& w, float& wd) {
  w = __builtin_amdgcn_rcpf(fabsf(d) + 1.0f);
  wd = w * d;
}

__device__ __forceinline__ int rec_voxel(unsigned long long r) { return (int)(r >> 55); }
__device__ __forceinline__ int rec_point(unsigned long long r) { return (int)((r >> 32) & ((1u << kPointIdxBits) - 1u)); }
__device__ __forceinline__ float rec_dist(unsigned long long r) { return __uint_as_float((uint32_t)r); }

// ---- the fold of one voxel's updates in a pass ---------------------------------------------
// Update by update (H:637-654), S = (W S + w d)/(W + w) and W = min(W + w, maxW).
// While W + w stays within maxW, n updates in a row give the weighted mean
//   S_n = (W_0 S_0 + sum w_i d_i) / (W_0 + sum w_i),   W_n = W_0 + sum w_i,
// whatever their order.  A pass therefore folds the updates of each (voxel,
// scan) pair -- a group -- as one such mean: each update's w and w d (FP32,
// as fold_step forms them) become fixed-point integers with 21 fractional
// bits (rounded to nearest: at most 2^-22 off, far below the fold's
// tolerance), their sums are exact and independent of the order in which
// updates arrive, and the mean is formed once.  A group whose weight would
// cross maxW is folded in ascending point index: as one mean up to the update
// that would cross it, then update by update.  Groups are applied in scan
// order and the state is rounded to its stored format between them, so a
// batch of scans gives the maps that the same scans give one pass each.
// 2^21: one unit of weight.  Both terms of an update are below 2^21 in
// magnitude, so the sums of up to 1024 updates fit 32-bit accumulators.
constexpr float kFixOne = 2097152.0f;
constexpr int kFixSum32 = 1024;  // updates whose sums fit 32 bits

__device__ __forceinline__ void fixed_terms(float d, uint32_t& wq, int32_t& wdq) {
  float w, wd;
  weight_of(d, w, wd);
  wq = max(__float2uint_rn(w * kFixOne), 1u);  // (an update always carries weight)
  wdq = __float2int_rn(wd * kFixOne);
}
// Fixed-point weight that a state of weight W can take before it reaches maxW.
__device__ __forceinline__ unsigned long long weight_room(float W, float maxW) {
  const double r = ((double)maxW - (double)W) * (double)kFixOne;
  return r <= 0.0 ? 0ull : r >= 1.8e19 ? ~0ull : (unsigned long long)r;
}
// A fixed-point sum as FP32: its high and low words converted and joined
// by one fused multiply-add.  A negative sum is converted by its magnitude:
// its two's-complement low word lies near 2^32, where FP32 keeps only 24 of
// its bits.
__device__ __forceinline__ float fixed_to_float(unsigned long long v) {
  return __builtin_fmaf((float)(uint32_t)(v >> 32), 4294967296.0f, (float)(uint32_t)v) * (1.0f / kFixOne);
}
__device__ __forceinline__ float fixed_to_float(long long v) {
  const float f = fixed_to_float((unsigned long long)(v < 0 ? -v : v));
  return v < 0 ? -f : f;
}
// One group's sums applied as a weighted mean (the caller has checked the
// room), with the FP32 operations and hardware reciprocal of fold_step; wf
// is fixed_to_float(sw).
__device__ __forceinline__ void apply_sums(float& S, float& W, float wf, long long swd) {
  const float Wn = W + wf;
  S = (W * S + fixed_to_float(swd)) * __builtin_amdgcn_rcpf(Wn);
  W = Wn;
}
__device__ __forceinline__ void apply_sums(float& S, float& W, unsigned long long sw, long long swd) {
  apply_sums(S, W, fixed_to_float(sw), swd);
}
// Whether a group of fixed-point weight sw (wf as FP32) fits below maxW:
// sw <= weight_room(W, maxW).  The FP32 test first: W + wf within 0.999 maxW
// (its rounding is far below that margin) implies it, so the exact test in
// FP64 only runs near maxW.
__device__ __forceinline__ bool group_fits(float W, unsigned long long sw, float wf, float maxW) {
  return W + wf <= 0.999f * maxW || sw <= weight_room(W, maxW);
}

// The pass's scan starts (soff[ns] = points in the pass) and maps, staged in
// LDS by the fold workgroups: a record's group is the scan of its point.
struct PassScans {
  int32_t off[kMaxScansPerBatch + 1];
  uint8_t map[kMaxScansPerBatch];
  int32_t ns;
};
__device__ __forceinline__ void stage_scans(const Dev& D, PassScans& ps) {
  const PassHdr hdr = *D.hdr;
  for (int s = threadIdx.x; s < hdr.nscans; s += blockDim.x) {
    const ScanDev sd = D.scans[s];
    ps.off[s] = sd.off;
    ps.map[s] = (uint8_t)sd.map;
  }
  if (threadIdx.x == 0) {
    ps.off[hdr.nscans] = hdr.total;
    ps.ns = hdr.nscans;
  }
}
// The scan of point p (scans are contiguous and in pass order), searched
// from scan `from` on (p lies in it or after it): a few steps forward, where
// the next group of a map usually is, else by bisection.
__device__ __forceinline__ int scan_in_lds(const PassScans& ps, int p, int from) {
  int s = from;
#pragma unroll
  for (int k = 0; k < 4; ++k)
    if (ps.off[s + 1] <= p) ++s;
  if (ps.off[s + 1] > p) return s;
  int lo = s, hi = ps.ns - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (ps.off[mid] <= p) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// A fold kernel's counts for the pass statistics: voxels folded, and
// records of blocks whose entry got no heap block (dropped).
struct FoldTally {
  unsigned long long uniq = 0;
  uint32_t dropped = 0;
};
// A fold descriptor with its entry (d.z) replaced by the entry's heap block,
// -1 when the entry got none.
__device__ __forceinline__ int4 resolve_desc(const Dev& D, int4 d) {
  d.z = D.vals[d.z];
  return d;
}

// The ordered fold of one voxel's updates, fed in ascending point index
// (possibly over several calls, as the point rounds do).
template <class State>
struct OrderedFold {
  float S, W;
  unsigned long long room, sw;  // the open group: room left below maxW, its pending weight sum
  long long swd;
  int32_t gend;                 // first point past the open group (-1: none open yet)
  int32_t gscan;                // its scan
  bool seq;                     // the open group crossed maxW: update by update from here on

  __device__ __forceinline__ void begin(float S0, float W0) {
    S = S0;
    W = W0;
    room = 0;
    sw = 0;
    swd = 0;
    gend = -1;
    gscan = 0;
    seq = false;
  }
  __device__ __forceinline__ void close() {
    if (!seq && sw != 0) apply_sums(S, W, sw, swd);
    sw = 0;
    swd = 0;
  }
  __device__ __forceinline__ void add(unsigned long long rec, float maxW, const Codec& cd, const PassScans& ps) {
    uint32_t wq;
    int32_t wdq;
    fixed_terms(rec_dist(rec), wq, wdq);
    add_terms(rec, wq, wdq, maxW, cd, ps);
  }
  // opens the group of point p: the previous one is applied and rounded to
  // the stored format
  __device__ __forceinline__ void open_group(int p, float maxW, const Codec& cd, const PassScans& ps) {
    if (gend >= 0) {
      close();
      State::round(cd, S, W);
    }
    if (ps.ns != 1) gscan = scan_in_lds(ps, p, gscan);
    gend = ps.off[gscan + 1];
    room = weight_room(W, maxW);
    seq = false;
  }
  // an update of a group past maxW: the pending sums applied, then update by update
  __device__ __forceinline__ void step_seq(unsigned long long rec, float maxW) {
    if (!seq) {
      close();
      seq = true;
    }
    float w, wd;
    weight_of(rec_dist(rec), w, wd);
    fold_step(S, W, w, wd, maxW);
  }
  // the same with the record's fixed-point terms formed by the caller
  __device__ __forceinline__ void add_terms(unsigned long long rec, uint32_t wq, int32_t wdq, float maxW,
                                            const Codec& cd, const PassScans& ps) {
    if (rec_point(rec) >= gend) open_group(rec_point(rec), maxW, cd, ps);
    if (!seq && sw + wq <= room) {
      sw += wq;
      swd += wdq;
      return;
    }
    step_seq(rec, maxW);
  }
};

// The groups of one block's map: bnd[j] is the start of the map's j-th scan
// in the pass; returns their number G.  Every wave counts them (each reads
// the pass's scans from LDS); wave 0 files the starts.
__device__ __forceinline__ int map_groups(const PassScans& ps, int m, int32_t* bnd) {
  int G = 0;
  const int lane = lane_id();
  const bool filer = threadIdx.x < 64;
  for (int s0 = 0; s0 < ps.ns; s0 += 64) {
    const int s = s0 + lane;
    const bool mine = s < ps.ns && ps.map[s] == m;
    const unsigned long long b = __ballot(mine);
    if (filer && mine) bnd[G + (int)__popcll(b & ((1ull << lane) - 1ull))] = ps.off[s];
    G += (int)__popcll(b);
  }
  return G;
}
// Group (rank of the scan within the block's map) of point p; b1..b3 are
// bnd[1..3] (past every point where G is smaller), bnd in LDS for G > 4.
__device__ __forceinline__ int group_of(int p, int G, int32_t b1, int32_t b2, int32_t b3, const int32_t* bnd) {
  if (G <= 4) return (p >= b1 ? 1 : 0) + (p >= b2 ? 1 : 0) + (p >= b3 ? 1 : 0);
  int lo = 0, hi = G - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (bnd[mid] <= p) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// One block by sums: each voxel's group sums in LDS accumulators (acc: gw
// windows of 512 {weight, weighted distance} pairs, 64-bit when kWide, else
// 32-bit for blocks of at most kFixSum32 records; groups wb .. wb + gw - 1
// per turn over the block's records), then each voxel's groups applied in
// order by the voxel's thread.  A block of at most 4 NT records keeps them
// in registers across the turns; a larger one reads them again per turn.
// Returns true, with nothing stored, when some group would cross maxW: the
// block then takes an ordered fold.  d.z is the heap block (resolve_desc).
// bnd: LDS for the map's group starts (kMaxScansPerBatch entries); flag: an
// LDS word.
template <class State, int NT, bool kWide>
__device__ bool fold_sum_block(const Dev& D, const int4 d, void* acc, int gw, const PassScans& ps,
                               int32_t* bnd, int* flag, FoldTally& tally) {
  using Pair = typename std::conditional<kWide, ulonglong2, uint2>::type;
  Pair* const accp = reinterpret_cast<Pair*>(acc);
  constexpr int kVpt = NT >= kVoxelsPerBlock ? 1 : kVoxelsPerBlock / NT;  // voxels per thread
  const int t = threadIdx.x;
  const int c = d.y;
  const unsigned long long* sg = D.seg + (uint32_t)d.x;
  const int G = ps.ns == 1 ? 1 : map_groups(ps, d.w, bnd);
  const MapDev& m = D.maps[d.w];
  const float maxW = m.maxW;
  const Codec cd = make_codec(m);
  float S[kVpt], W[kVpt];
  bool seen[kVpt];
#pragma unroll
  for (int u = 0; u < kVpt; ++u) {
    S[u] = 0.f;
    W[u] = 0.f;
    seen[u] = false;
  }
  const bool held = c <= 4 * NT;
  unsigned long long r[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) r[u] = held && t + u * NT < c ? sg[t + u * NT] : 0ull;
  if (d.z < 0) {  // the entry got no heap block: the block's records are dropped
    if (t == 0) tally.dropped += (uint32_t)c;
    return false;
  }
  // every voxel state of the block is requested now, with the records: a
  // block of this tier touches a voxel in nearly every 64-byte line of its
  // states, and the loads then overlap the sums instead of following them
#pragma unroll
  for (int u = 0; u < kVpt; ++u) {
    const int l = t + NT * u;
    if (l < kVoxelsPerBlock) State::load(D.heap, (int64_t)d.z * kVoxelsPerBlock + l, cd, S[u], W[u]);
  }
  for (int wb = 0;; wb += gw) {
    const int nw = min(gw, G - wb);
    for (int x = t; x < nw * kVoxelsPerBlock; x += NT) accp[x] = Pair{0, 0};
    __syncthreads();
    // (every wave has read the previous block's flag: it passed this barrier)
    if (t == 0) *flag = 0;
    const int32_t b1 = G > 1 ? bnd[1] : INT32_MAX, b2 = G > 2 ? bnd[2] : INT32_MAX, b3 = G > 3 ? bnd[3] : INT32_MAX;
    for (int x0 = t; x0 < c; x0 += 4 * NT) {
      if (!held) {
#pragma unroll
        for (int u = 0; u < 4; ++u) r[u] = x0 + u * NT < c ? sg[x0 + u * NT] : 0ull;
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        if (x0 + u * NT >= c) continue;
        const int g = (G == 1 ? 0 : group_of(rec_point(r[u]), G, b1, b2, b3, bnd)) - wb;
        if ((unsigned)g >= (unsigned)nw) continue;
        uint32_t wq;
        int32_t wdq;
        fixed_terms(rec_dist(r[u]), wq, wdq);
        Pair* a = accp + g * kVoxelsPerBlock + rec_voxel(r[u]);
        if constexpr (kWide) {
          atomicAdd(&a->x, (unsigned long long)wq);
          atomicAdd(&a->y, (unsigned long long)(long long)wdq);
        } else {
          atomicAdd(&a->x, wq);
          atomicAdd(&a->y, (uint32_t)wdq);
        }
      }
    }
    __syncthreads();
    // each voxel's groups of this turn, in order
    bool clamp = false;
#pragma unroll
    for (int u = 0; u < kVpt; ++u) {
      const int l = t + NT * u;
      if (l >= kVoxelsPerBlock) continue;
      for (int g = 0; g < nw; ++g) {
        const Pair ap = accp[g * kVoxelsPerBlock + l];
        const unsigned long long sw = ap.x;
        const long long swd = kWide ? (long long)ap.y : (long long)(int32_t)ap.y;
        if (sw == 0) continue;
        if (seen[u]) State::round(cd, S[u], W[u]);
        seen[u] = true;
        const float wf = fixed_to_float(sw);
        if (group_fits(W[u], sw, wf, maxW)) apply_sums(S[u], W[u], wf, swd);
        else clamp = true;
      }
    }
    if (clamp) *flag = 1;
    __syncthreads();
    if (*flag) return true;
    if (wb + gw >= G) break;
  }
#pragma unroll
  for (int u = 0; u < kVpt; ++u) {
    const int l = t + NT * u;
    if (l < kVoxelsPerBlock && seen[u]) {
      State::store(D.heap, (int64_t)d.z * kVoxelsPerBlock + l, cd, S[u], W[u]);
      tally.uniq++;
    }
  }
  return false;
}

// The ordered fold of the run rin[b..e) of one voxel's records (ascending point).
template <class State>
__device__ __forceinline__ void fold_run(OrderedFold<State>& of, const unsigned long long* rin, int b, int e,
                                         float maxW, const Codec& cd, const PassScans& ps) {
  for (int i = b; i < e; ++i) of.add(rin[i], maxW, cd, ps);
}


// Blocks of at most G records, 64/G blocks per wave: each group of G lanes
// holds one block's records, bitonic-sorts them with shuffles, and every
// run of one voxel is folded by its first lane pulling the run's weight
// terms from the following lanes.
// One wave folds the class-G descriptors [q_begin, q_end), 64/G at a time.
template <class State, int G>
__device__ __forceinline__ void fold_small_range(const Dev& D, int q_begin, int q_end, const PassScans& ps,
                                                 FoldTally& tally) {
  constexpr int kPer = 64 / G;
  constexpr int kCls = G == 16 ? FC_16 : G == 32 ? FC_32 : FC_64;
  const int lane = lane_id();
  const int gl = lane & (G - 1);
  const int grp = lane / G;
  const int n_desc = q_end;
  const int4* desc = D.fold_desc + D.fold_base[kCls];
  for (int q0 = q_begin; q0 < n_desc; q0 += kPer) {
    const int q = q0 + grp;
    int4 d = make_int4(0, 0, 0, 0);
    if (q < n_desc) d = resolve_desc(D, desc[q]);
    const int n = d.y;
    unsigned long long v = gl < n ? D.seg[(uint32_t)d.x + gl] : ~0ull;
#pragma unroll
    for (int k = 2; k <= G; k <<= 1) {
#pragma unroll
      for (int j = k >> 1; j > 0; j >>= 1) {
        const unsigned long long o = xor_lane64(v, j);
        const bool keep_min = ((gl & j) == 0) == ((gl & k) == 0);
        v = keep_min ? (o < v ? o : v) : (o > v ? o : v);
      }
    }
    const unsigned long long prev = prev_lane64(v);
    // (an entry without a heap block: the block's records are dropped)
    const bool live = d.z >= 0;
    if (!live && gl == 0) tally.dropped += (uint32_t)n;
    const bool head = live && gl < n && (gl == 0 || rec_voxel(prev) != rec_voxel(v));
    const unsigned long long heads = __ballot(head);
    const int gend = grp * G + n;  // first lane past this group's records
    const unsigned long long above = lane == 63 ? 0ull : (heads >> (lane + 1)) << (lane + 1);
    const int nh = above ? __ffsll((long long)above) - 1 : 64;
    const int len = head ? min(nh, gend) - lane : 0;
    uint32_t wq;
    int32_t wdq;
    fixed_terms(rec_dist(v), wq, wdq);
    const MapDev& m = D.maps[d.w];
    const float maxW = m.maxW;
    const Codec cd = make_codec(m);
    const int64_t vi = (int64_t)d.z * kVoxelsPerBlock + rec_voxel(v);
    float S0 = 0.f, W0 = 0.f;
    if (head) State::load(D.heap, vi, cd, S0, W0);
    OrderedFold<State> of;
    of.begin(S0, W0);
    uint32_t maxlen = (uint32_t)len;
    maxlen = max(maxlen, xor_lane<32>(maxlen));
    maxlen = max(maxlen, xor_lane<16>(maxlen));
    maxlen = max(maxlen, xor_lane<8>(maxlen));
    maxlen = max(maxlen, xor_lane<4>(maxlen));
    maxlen = max(maxlen, xor_lane<2>(maxlen));
    maxlen = max(maxlen, xor_lane<1>(maxlen));
    // step r: lane i holds the record and terms of lane i + r
    unsigned long long vr = v;
    uint32_t wqr = wq;
    int32_t wdqr = wdq;
    // the common turn (a pending group takes the update) is straight-line
    // code; group changes and updates past maxW go through branches that a
    // wave enters only when one of its lanes needs them
    for (int r = 0; r < (int)maxlen; ++r) {
      const bool act = r < len;
      const int p = rec_point(vr);
      const bool chg = act && p >= of.gend;
      if (__ballot(chg)) {
        if (chg) of.open_group(p, maxW, cd, ps);
      }
      const unsigned long long nsw = of.sw + wqr;
      const bool fits = act && !of.seq && nsw <= of.room;
      of.sw = fits ? nsw : of.sw;
      of.swd = fits ? of.swd + wdqr : of.swd;
      const bool slow = act && !fits;
      if (__ballot(slow)) {
        if (slow) of.step_seq(vr, maxW);
      }
      vr = next_lane64(vr);
      wqr = next_lane(wqr);
      wdqr = (int32_t)next_lane((uint32_t)wdqr);
    }
    if (head) {
      of.close();
      State::store(D.heap, vi, cd, of.S, of.W);
      tally.uniq++;
    }
  }
}

// One block of 65..64 R records, one wave (R records per lane):
// the records are bitonic-sorted in registers across the wave (element
// r * 64 + lane), which orders them by voxel, then by point; the sorted
// records go to the wave's LDS strip (64 R entries), and the lane holding
// the first record of each voxel folds that voxel's run from the strip.
// No workgroup barrier: the wave only waits for its own LDS traffic.
template <class State, int R>
__device__ __forceinline__ void fold_wave_block(const Dev& D, const int4 d, unsigned long long* strip,
                                                const PassScans& ps, FoldTally& tally) {
  constexpr int N = 64 * R;
  const int lane = lane_id();
  const int n = d.y;
  const unsigned long long* sg = D.seg + (uint32_t)d.x;
  unsigned long long v[R];
#pragma unroll
  for (int r = 0; r < R; ++r) v[r] = 64 * r + lane < n ? sg[64 * r + lane] : ~0ull;
  if (d.z < 0) {  // the entry got no heap block: the block's records are dropped
    if (lane == 0) tally.dropped += (uint32_t)n;
    return;
  }
#pragma unroll
  for (int k = 2; k <= N; k <<= 1) {
#pragma unroll
    for (int j = k >> 1; j > 0; j >>= 1) {
      if (j >= 64) {
        // partner in another register of this lane; the direction depends
        // on the register only (k > 64)
#pragma unroll
        for (int r = 0; r < R; ++r) {
          if (r & (j >> 6)) continue;
          const int r2 = r | (j >> 6);
          const bool up = ((64 * r) & k) == 0;
          const unsigned long long a = v[r], b = v[r2];
          const bool sw = up ? a > b : a < b;
          v[r] = sw ? b : a;
          v[r2] = sw ? a : b;
        }
      } else {
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const unsigned long long o = xor_lane64(v[r], j);
          const bool up = (((64 * r) | lane) & k) == 0;
          const bool keep_min = ((lane & j) == 0) == up;
          v[r] = keep_min ? (o < v[r] ? o : v[r]) : (o > v[r] ? o : v[r]);
        }
      }
    }
  }
#pragma unroll
  for (int r = 0; r < R; ++r) strip[64 * r + lane] = v[r];
  wave_sync();
  const MapDev& m = D.maps[d.w];
  const float maxW = m.maxW;
  const Codec cd = make_codec(m);
  bool head[R];
  float S[R], W[R];
  // heads and their voxels' states first (the state loads overlap each other)
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int e = 64 * r + lane;
    head[r] = e < n && (e == 0 || rec_voxel(strip[e - 1]) != rec_voxel(v[r]));
    S[r] = 0.f;
    W[r] = 0.f;
    if (head[r]) State::load(D.heap, (int64_t)d.z * kVoxelsPerBlock + rec_voxel(v[r]), cd, S[r], W[r]);
  }
#pragma unroll
  for (int r = 0; r < R; ++r) {
    if (!head[r]) continue;
    const int l = rec_voxel(v[r]);
    int x = 64 * r + lane;
    unsigned long long cur = v[r];
    OrderedFold<State> of;
    of.begin(S[r], W[r]);
    for (;;) {
      of.add(cur, maxW, cd, ps);
      if (++x >= n) break;
      cur = strip[x];
      if (rec_voxel(cur) != l) break;
    }
    of.close();
    State::store(D.heap, (int64_t)d.z * kVoxelsPerBlock + l, cd, of.S, of.W);
    tally.uniq++;
  }
}

__device__ __forceinline__ void bitonic_sort_lds(unsigned long long* a, int n) {
  for (int k = 2; k <= n; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int x = threadIdx.x; x < (n >> 1); x += blockDim.x) {
        const int i = ((x & ~(j - 1)) << 1) | (x & (j - 1));
        const int ip = i + j;
        const bool up = (i & k) == 0;
        const unsigned long long u = a[i], v = a[ip];
        if ((u > v) == up) {
          a[i] = v;
          a[ip] = u;
        }
      }
      __syncthreads();
    }
  }
}

// One block of 65..TILE records, one workgroup of NT threads: a counting
// sort by voxel in LDS (histogram, scan, scatter) groups every voxel's
// updates; each record is then placed by its rank (by point) inside its
// voxel, and each voxel's thread folds its updates in point order.  rin/rout
// hold TILE records.
template <class State, int TILE, int NT, class StartT = uint32_t>
__device__ __forceinline__ void fold_count_block(const Dev& D, const int4 d, unsigned long long* rin,
                                                 unsigned long long* rout, uint32_t* cnt, StartT* start,
                                                 uint32_t* sh, const PassScans& ps, FoldTally& tally) {
  static_assert(NT == 256 || NT == 512 || NT == 1024, "workgroup of 256, 512 or 1024 threads");
  constexpr int kPer = TILE / NT;                                         // records per thread
  constexpr int kVpt = NT >= kVoxelsPerBlock ? 1 : kVoxelsPerBlock / NT;  // voxels per thread
  const int c = d.y;
  const unsigned long long* sg = D.seg + (uint32_t)d.x;
  for (int l = threadIdx.x; l < kVoxelsPerBlock; l += NT) cnt[l] = 0;
  // all of this thread's records are requested before any is used (one
  // memory latency per block, not one per record)
  unsigned long long rr[kPer];
#pragma unroll
  for (int u = 0; u < kPer; ++u) {
    const int x = threadIdx.x + NT * u;
    rr[u] = x < c ? sg[x] : 0ull;
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < kPer; ++u) {
    const int x = threadIdx.x + NT * u;
    if (x < c) {
      rin[x] = rr[u];
      atomicAdd(&cnt[rec_voxel(rr[u])], 1u);
    }
  }
  __syncthreads();
  // the states of this thread's voxels (l = t + NT u) are loaded now, so
  // that the loads overlap the sort below
  const MapDev& m = D.maps[d.w];
  const float maxW = m.maxW;
  const Codec cd = make_codec(m);
  float S[kVpt], W[kVpt];
  bool h[kVpt];
#pragma unroll
  for (int u = 0; u < kVpt; ++u) {
    const int l = threadIdx.x + NT * u;
    h[u] = l < kVoxelsPerBlock && cnt[l] != 0;
    S[u] = 0.f;
    W[u] = 0.f;
    if (h[u]) State::load(D.heap, (int64_t)d.z * kVoxelsPerBlock + l, cd, S[u], W[u]);
  }
  __syncthreads();
  {
    // the scan takes voxels kVpt t .. kVpt t + kVpt - 1 per thread
    const int lb = kVpt * threadIdx.x;
    const bool own = lb < kVoxelsPerBlock;
    uint32_t cv[kVpt], sum = 0;
#pragma unroll
    for (int u = 0; u < kVpt; ++u) {
      cv[u] = own ? cnt[lb + u] : 0u;
      sum += cv[u];
    }
    uint32_t tot;
    uint32_t run = block_scan_excl<NT / 64>(sum, sh, tot);
    if (own) {
#pragma unroll
      for (int u = 0; u < kVpt; ++u) {
        start[lb + u] = run;
        run += cv[u];
        cnt[lb + u] = 0;
      }
    }
    if (threadIdx.x == 0) start[kVoxelsPerBlock] = tot;
  }
  __syncthreads();
  for (int x = threadIdx.x; x < c; x += NT) {
    const unsigned long long r = rin[x];
    const int l = rec_voxel(r);
    rout[start[l] + atomicAdd(&cnt[l], 1u)] = r;
  }
  __syncthreads();
  // point order inside each voxel: a record's place is the number of records
  // of its voxel that sort before it (records of one voxel are distinct, a
  // walk visits a cell once)
  // (eight independent LDS reads per step keep the loop from waiting on
  // each read in turn)
  for (int x = threadIdx.x; x < c; x += NT) {
    const unsigned long long r = rout[x];
    const int l = rec_voxel(r);
    const int b0 = (int)start[l], b1 = (int)start[l + 1];
    int rank = 0;
    int j = b0;
    const uint32_t* hw = reinterpret_cast<const uint32_t*>(rout) + 1;
    const uint32_t rh = (uint32_t)(r >> 32);
    for (; j + 8 <= b1; j += 8) {
      uint32_t v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) v[u] = hw[2 * (j + u)];
#pragma unroll
      for (int u = 0; u < 8; ++u) rank += v[u] < rh ? 1 : 0;
    }
    for (; j < b1; ++j) rank += hw[2 * j] < rh ? 1 : 0;
    rin[b0 + rank] = r;
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < kVpt; ++u) {
    const int l = threadIdx.x + NT * u;
    if (l >= kVoxelsPerBlock || !h[u]) continue;
    OrderedFold<State> of;
    of.begin(S[u], W[u]);
    fold_run(of, rin, (int)start[l], (int)start[l + 1], maxW, cd, ps);
    of.close();
    State::store(D.heap, (int64_t)d.z * kVoxelsPerBlock + l, cd, of.S, of.W);
    tally.uniq++;
  }
  __syncthreads();
}

// One block beyond the counting tile, folded in voxel ranges: the voxel
// counts of the whole block are taken first (a histogram over its records,
// read from the segment), and a voxel belongs to range start / span, where
// start is the number of the block's records in the voxels before it and
// span = kCountTile - (the block's largest voxel count).  While no voxel
// holds more than kRangeSpan records, span >= kRangeSpan, the ranges follow
// each other without gaps and each holds fewer than span + max = kCountTile
// records, so each range is sorted and folded like a counting tile, its
// records taken from the segment once per range (the widest span that fits
// keeps those passes few); voxel l's state is folded by thread l in its
// range.  Returns false (with nothing folded) for a block outside those
// bounds, which then takes the point rounds of fold_point_rounds.
constexpr int kRangeSpan = kCountTile / 2;
constexpr int kMaxRanges = 512;
template <class State, int NT>
__device__ __forceinline__ bool fold_ranged_block(const Dev& D, const int4 d, unsigned long long* rin,
                                                  unsigned long long* rout, uint32_t* cnt, uint32_t* start,
                                                  uint32_t* rstart, uint32_t* sh, const PassScans& ps,
                                                  FoldTally& tally) {
  static_assert(NT >= kVoxelsPerBlock, "a thread per voxel");
  static_assert(2 * kRangeSpan <= kCountTile, "a range fits the counting tile");
  __shared__ int too_big, nranges_sh, maxc_sh;
  const int c = d.y;
  const unsigned long long* sg = D.seg + (uint32_t)d.x;
  const int t = threadIdx.x;
  if (t < kVoxelsPerBlock) cnt[t] = 0;
  if (t == 0) {
    too_big = c / kRangeSpan + 1 > kMaxRanges ? 1 : 0;
    maxc_sh = 0;
  }
  __syncthreads();
  for (int x = t; x < c; x += NT) atomicAdd(&cnt[rec_voxel(sg[x])], 1u);
  __syncthreads();
  const uint32_t mine = t < kVoxelsPerBlock ? cnt[t] : 0u;
  if (mine > (uint32_t)kRangeSpan) too_big = 1;
  if (mine != 0) atomicMax(&maxc_sh, (int)mine);
  uint32_t tot;
  const uint32_t ex = block_scan_excl<NT / 64>(mine, sh, tot);  // (its barriers publish too_big and maxc_sh)
  if (too_big) return false;
  const uint32_t span = (uint32_t)(kCountTile - maxc_sh);  // >= kRangeSpan
  // voxel t: its start, range, state.  The ranges are those of the voxels
  // with records: 0 up to the range of the last such voxel, with no gap
  // (consecutive voxels' starts differ by at most the largest count <= span).
  const int rg = (int)(ex / span);
  if (t < kVoxelsPerBlock) start[t] = ex;
  if (t == 0) start[kVoxelsPerBlock] = tot;
  if (mine != 0 && ex + mine == tot) nranges_sh = rg + 1;
  const bool h = t < kVoxelsPerBlock && mine != 0;
  __syncthreads();
  // range r starts at the start of its first voxel: the voxel whose range
  // differs from the previous voxel's
  const int nranges = nranges_sh;
  if (t < kVoxelsPerBlock && (t == 0 || (int)(start[t - 1] / span) != rg) && ex < tot) rstart[rg] = ex;
  if (t == 0) rstart[nranges] = tot;
  __syncthreads();
  for (int r = 0; r < nranges; ++r) {
    const uint32_t base = rstart[r];
    const uint32_t lo = (uint32_t)r * span;
    const int nr = (int)(rstart[r + 1] - base);
    if (t < kVoxelsPerBlock) cnt[t] = 0;
    __syncthreads();
    for (int x = t; x < c; x += NT) {
      const unsigned long long rec = sg[x];
      const int l = rec_voxel(rec);
      const uint32_t sl = start[l];
      if (sl - lo < span) rout[sl - base + atomicAdd(&cnt[l], 1u)] = rec;  // (unsigned: sl >= lo too)
    }
    __syncthreads();
    // point order inside each voxel (as in fold_count_block)
    for (int x = t; x < nr; x += NT) {
      const unsigned long long rec = rout[x];
      const int l = rec_voxel(rec);
      const int b0 = (int)(start[l] - base), b1 = (int)(start[l + 1] - base);
      int rank = 0;
      int j = b0;
      const uint32_t* hw = reinterpret_cast<const uint32_t*>(rout) + 1;
      const uint32_t rh = (uint32_t)(rec >> 32);
      for (; j + 8 <= b1; j += 8) {
        uint32_t v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = hw[2 * (j + u)];
#pragma unroll
        for (int u = 0; u < 8; ++u) rank += v[u] < rh ? 1 : 0;
      }
      for (; j < b1; ++j) rank += hw[2 * j] < rh ? 1 : 0;
      rin[b0 + rank] = rec;
    }
    __syncthreads();
    if (h && rg == r) {
      // the voxel's state is read and written in its range's fold only
      const MapDev& m = D.maps[d.w];
      const Codec cd = make_codec(m);
      float S, W;
      State::load(D.heap, (int64_t)d.z * kVoxelsPerBlock + t, cd, S, W);
      OrderedFold<State> of;
      of.begin(S, W);
      fold_run(of, rin, (int)(ex - base), (int)(start[t + 1] - base), m.maxW, cd, ps);
      of.close();
      State::store(D.heap, (int64_t)d.z * kVoxelsPerBlock + t, cd, of.S, of.W);
      tally.uniq++;
    }
    __syncthreads();
  }
  return true;
}

// One block that no voxel-range split fits: some voxel holds more than
// kRangeSpan records (any block size).  The block is folded in rounds of
// consecutive point indices [lo, hi), each holding at most kCountTile
// records; one point has at most 22 records in a block (the longest
// monotone cell path through 8^3 voxels), so every round advances.  A round
// chooses hi with two histogram levels over the point index (bins of 1024
// points, then the single points of the bin where the tile fills up),
// gathers its records into the tile and sorts them as integers, which
// orders them by voxel, then by point.  Thread l keeps voxel l's state in
// registers from the first round to the last, so each voxel's updates are
// folded in point order across the rounds.  The block's records are read
// three times per round: this path is for inputs that pile thousands of
// updates of one pass onto single voxels, which scans of a moving lidar do
// not produce.  rin holds kCountTile records; rout's memory holds the
// histograms.
constexpr int kRoundBinShift = 10;
constexpr int kRoundBins = 1 << (kPointIdxBits - kRoundBinShift);  // first-level bins over the pass's points
template <class State, int NT>
__device__ void fold_point_rounds(const Dev& D, const int4 d, unsigned long long* rin, unsigned long long* rout,
                                  uint32_t* cnt, uint32_t* start, uint32_t* sh, const PassScans& ps,
                                  FoldTally& tally) {
  static_assert(NT == (1 << kRoundBinShift), "a thread per point of a second-level bin");
  static_assert(NT >= kVoxelsPerBlock, "a thread per voxel");
  static_assert(kRoundBins * sizeof(uint32_t) <= kCountTile * sizeof(unsigned long long), "histogram fits in rout");
  static_assert(kRoundBins % NT == 0, "first-level bins per thread");
  __shared__ int lo_sh, hi_sh, bin_sh, base_sh, ntile_sh;
  const int c = d.y;
  const unsigned long long* sg = D.seg + (uint32_t)d.x;
  const int t = threadIdx.x;
  uint32_t* hist = reinterpret_cast<uint32_t*>(rout);
  // the voxels with records, and their states
  if (t < kVoxelsPerBlock) cnt[t] = 0;
  if (t == 0) lo_sh = 0;
  __syncthreads();
  for (int x = t; x < c; x += NT) atomicAdd(&cnt[rec_voxel(sg[x])], 1u);
  __syncthreads();
  const MapDev& m = D.maps[d.w];
  const Codec cd = make_codec(m);
  const bool h = t < kVoxelsPerBlock && cnt[t] != 0;
  float S = 0.f, W = 0.f;
  if (h) State::load(D.heap, (int64_t)d.z * kVoxelsPerBlock + t, cd, S, W);
  OrderedFold<State> of;  // (kept across the rounds)
  of.begin(S, W);
  constexpr int kPer = kRoundBins / NT;
  for (;;) {
    // level 1: records of points >= lo, per bin of 1024 points
    for (int x = t; x < kRoundBins; x += NT) hist[x] = 0;
    __syncthreads();
    const int lo = lo_sh;
    for (int x = t; x < c; x += NT) {
      const int p = rec_point(sg[x]);
      if (p >= lo) atomicAdd(&hist[p >> kRoundBinShift], 1u);
    }
    __syncthreads();
    uint32_t v[kPer], sum = 0;
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      v[u] = hist[t * kPer + u];
      sum += v[u];
    }
    uint32_t tot;
    uint32_t run = block_scan_excl<NT / 64>(sum, sh, tot);
    if (tot == 0) break;  // (tot is the same for every thread)
    if (t == 0) {
      bin_sh = -1;
      hi_sh = 1 << kPointIdxBits;  // past every point
    }
    __syncthreads();
    if (tot > (uint32_t)kCountTile) {
      // the bin in which the tile fills up: exactly one
#pragma unroll
      for (int u = 0; u < kPer; ++u) {
        if (run <= (uint32_t)kCountTile && run + v[u] > (uint32_t)kCountTile) {
          bin_sh = t * kPer + u;
          base_sh = (int)run;
        }
        run += v[u];
      }
    }
    __syncthreads();
    const int B = bin_sh;
    if (B >= 0) {
      // level 2: the records of each point of bin B; hi is the first point
      // whose records no longer fit
      const int b0 = B << kRoundBinShift;
      hist[t] = 0;
      __syncthreads();
      for (int x = t; x < c; x += NT) {
        const int p = rec_point(sg[x]);
        if (p >= lo && (p >> kRoundBinShift) == B) atomicAdd(&hist[p - b0], 1u);
      }
      __syncthreads();
      const uint32_t w = hist[t];
      uint32_t tot2;
      const uint32_t e = block_scan_excl<NT / 64>(w, sh, tot2) + (uint32_t)base_sh;
      if (e <= (uint32_t)kCountTile && e + w > (uint32_t)kCountTile) hi_sh = b0 + t;
      __syncthreads();
    }
    const int hi = hi_sh;
    // the round's records, then each voxel's run in the sorted tile
    if (t < kVoxelsPerBlock) cnt[t] = 0;
    if (t == 0) ntile_sh = 0;
    __syncthreads();
    for (int x = t; x < c; x += NT) {
      const unsigned long long r = sg[x];
      const int p = rec_point(r);
      if (p >= lo && p < hi) {
        rin[atomicAdd(&ntile_sh, 1)] = r;
        atomicAdd(&cnt[rec_voxel(r)], 1u);
      }
    }
    __syncthreads();
    const int nr = ntile_sh;
    const int n2 = next_pow2(nr);
    for (int x = nr + t; x < n2; x += NT) rin[x] = ~0ull;
    {
      const uint32_t cv = t < kVoxelsPerBlock ? cnt[t] : 0u;
      uint32_t tot3;
      const uint32_t ex = block_scan_excl<NT / 64>(cv, sh, tot3);  // (its barriers also order the padding)
      if (t < kVoxelsPerBlock) start[t] = ex;
      if (t == 0) start[kVoxelsPerBlock] = tot3;
    }
    bitonic_sort_lds(rin, n2);
    if (h) fold_run(of, rin, (int)start[t], (int)start[t + 1], m.maxW, cd, ps);
    if (t == 0) lo_sh = hi;
    __syncthreads();
  }
  if (h) {
    of.close();
    State::store(D.heap, (int64_t)d.z * kVoxelsPerBlock + t, cd, of.S, of.W);
    tally.uniq++;
  }
  __syncthreads();
}

// k_fold_a's dynamic LDS: gw windows of 32-bit voxel sums (4 KB each) and
// the map's group starts after them for the sum path; for the ordered fold of
// a block of at most kOrdTileA records, its counting tile pair, counts,
// voxel starts and scan scratch over the same bytes (the sum path is over by
// then).  The wave tier's strips use the first 4 KB.
constexpr int kOrdTileA = 768;
constexpr int kOrdCntA = 2 * kOrdTileA * 8;                        // byte offset of the counts
constexpr int kOrdStartA = kOrdCntA + kVoxelsPerBlock * 4;         // of the voxel starts (16-bit)
constexpr int kOrdShA = kOrdStartA + ((kVoxelsPerBlock + 1) * 2 + 3) / 4 * 4;  // of the scan scratch
constexpr int kOrdBytesA = kOrdShA + kScanSh * 4;
__host__ __device__ constexpr int fold_a_lds_bytes(int gw) {
  return gw * kVoxelsPerBlock * 8 + kMaxScansPerBatch * 4 > kOrdBytesA ? gw * kVoxelsPerBlock * 8 + kMaxScansPerBatch * 4
                                                                         : kOrdBytesA;
}
static_assert(kOrdTileA % 256 == 0 && kOrdTileA <= kCountTileS, "a whole number of records per thread");

// Fold, first launch.  Work items, in this order: in a pass of one scan, the
// blocks of kCountTileS+1..kCountTile records (a workgroup each, by 64-bit
// sums in one window; a block of which some group would cross max_weight is
// marked kDescDeferred for k_fold_b); the blocks of 129..1024 records (a
// workgroup each, by 32-bit sums; a block of which some group would cross
// max_weight is folded in order in place when it holds at most kOrdTileA
// records, else marked kDescDeferred), those of 65..128 (a wave each), then
// chunks of the small classes (one group of 64/G blocks per wave).  LDS: fold_a_lds_bytes(gw) of dynamic LDS and the pass's
// scans; at most 64 VGPRs per thread, so eight workgroups share a CU.
// Workgroup q takes items q, q + grid, ...; with a grid larger than what fits
// on the chip at once, the dispatcher starts a new workgroup whenever one
// finishes, which balances the uneven items.
template <class State>
__global__ void __launch_bounds__(256, 8) k_fold_a(Dev D, int gw) {
  extern __shared__ __align__(16) unsigned char fold_a_lds[];
  __shared__ PassScans ps;
  __shared__ int flag;
  uint2* const acc = reinterpret_cast<uint2*>(fold_a_lds);
  int32_t* const bnd = reinterpret_cast<int32_t*>(fold_a_lds + gw * kVoxelsPerBlock * sizeof(uint2));
  unsigned long long* const rin = reinterpret_cast<unsigned long long*>(fold_a_lds);
  unsigned long long* const rout = rin + kOrdTileA;
  uint32_t* const cnt = reinterpret_cast<uint32_t*>(fold_a_lds + kOrdCntA);
  uint16_t* const ostart = reinterpret_cast<uint16_t*>(fold_a_lds + kOrdStartA);
  uint32_t* const osh = reinterpret_cast<uint32_t*>(fold_a_lds + kOrdShA);
  FoldTally tally;
  stage_scans(D, ps);
  const int n_wg2 = ld_acq(&D.ctr->fold_n[FC_WG2]);
  const int n_wg = ld_acq(&D.ctr->fold_n[FC_WG]);
  const int n_wgl = ld_acq(&D.ctr->fold_n[FC_WGL]);
  const int n64 = ld_acq(&D.ctr->fold_n[FC_64]);
  const int n32 = ld_acq(&D.ctr->fold_n[FC_32]);
  const int n16 = ld_acq(&D.ctr->fold_n[FC_16]);
  const int waves = blockDim.x >> 6, wave = threadIdx.x >> 6;
  // blocks per wave-level item: one block of 65..128 records per wave; one
  // group of 64/G small blocks per wave
  const int kl = waves, k64 = waves * 1, k32 = waves * 2, k16 = waves * 4;
  const int cl = (n_wgl + kl - 1) / kl;
  const int c64 = (n64 + k64 - 1) / k64, c32 = (n32 + k32 - 1) / k32, c16 = (n16 + k16 - 1) / k16;
  const int4* dwg2 = D.fold_desc + D.fold_base[FC_WG2];
  const int4* dwg = D.fold_desc + D.fold_base[FC_WG];
  const int4* dwgl = D.fold_desc + D.fold_base[FC_WGL];
  static_assert(4 * kWgSplit * sizeof(unsigned long long) <= kVoxelsPerBlock * sizeof(uint2),
                "the wave strips fit the first window of sums");
  static_assert(kCountTileS <= kFixSum32, "the workgroup tier's sums fit 32 bits");
  static_assert(kVoxelsPerBlock * sizeof(ulonglong2) <= kOrdBytesA, "one window of 64-bit sums fits");
  __syncthreads();
  // a pass of one scan (one group per voxel: one window of sums) folds its
  // blocks of kCountTileS+1..kCountTile records here as well, first, as its
  // longest items; k_fold_b, which takes them otherwise, makes the same test
  const int nb = ps.ns == 1 ? n_wg2 : 0;
  const int total = nb + n_wg + cl + c64 + c32 + c16;
  for (int it0 = blockIdx.x; it0 < total; it0 += gridDim.x) {
    if (it0 < nb) {
      const int4 d0 = dwg2[it0];
      if (fold_sum_block<State, 256, true>(D, resolve_desc(D, d0), acc, 1, ps, bnd, &flag, tally) &&
          threadIdx.x == 0)
        D.fold_desc[D.fold_base[FC_WG2] + it0].w = d0.w | kDescDeferred;
      continue;
    }
    const int it = it0 - nb;
    if (it < n_wg) {
      const int4 d0 = dwg[it];
      const int4 d = resolve_desc(D, d0);
      if (fold_sum_block<State, 256, false>(D, d, acc, gw, ps, bnd, &flag, tally)) {
        if (d.y <= kOrdTileA)
          fold_count_block<State, kOrdTileA, 256, uint16_t>(D, d, rin, rout, cnt, ostart, osh, ps, tally);
        else if (threadIdx.x == 0)
          D.fold_desc[D.fold_base[FC_WG] + it].w = d0.w | kDescDeferred;
      }
    } else if (it < n_wg + cl) {
      const int q = (it - n_wg) * kl + wave;
      if (q < n_wgl) {
        const int4 d = resolve_desc(D, dwgl[q]);
        unsigned long long* strip = reinterpret_cast<unsigned long long*>(fold_a_lds) + wave * kWgSplit;
        fold_wave_block<State, kWgSplit / 64>(D, d, strip, ps, tally);
      }
      __syncthreads();  // the strips live in the sums' LDS, which the next item may use
    } else {
      int c = it - n_wg - cl;
      if (c < c64) {
        const int q = c * k64 + wave;
        fold_small_range<State, 64>(D, q, min(q + 1, n64), ps, tally);
      } else if ((c -= c64) < c32) {
        const int q = c * k32 + wave * 2;
        fold_small_range<State, 32>(D, q, min(q + 2, n32), ps, tally);
      } else {
        c -= c32;
        const int q = c * k16 + wave * 4;
        fold_small_range<State, 16>(D, q, min(q + 4, n16), ps, tally);
      }
    }
  }
  const uint32_t row[2] = {(uint32_t)tally.uniq, tally.dropped};
  write_stat_row(D, SR_FOLD_A, row);
}

// Fold, second launch (which also takes the pass's failed inserts out of the
// table): blocks of kCountTileS+1..kCountTile records (in a pass of more than
// one scan; k_fold_a takes them in a pass of one) and the larger ones, by
// sums (a 1024-thread workgroup each, the 64 KB buffer as eight windows of
// voxel sums), and the blocks k_fold_a marked kDescDeferred.  A block of
// which some group would cross max_weight takes the ordered fold, with the
// buffer as the counting tile pair: a counting sort by voxel and rank
// placement by point, in voxel ranges that fit the tile for the larger
// blocks, or in point rounds when a voxel holds more than a range.
// Eight waves per SIMD (two workgroups, and their 2 x 68 KB of LDS, per CU)
// cap the kernel at 64 VGPRs.
// Items are handed out one at a time from a pass counter, largest first
// (the ranged blocks, then the kCountTile class):
// blocks differ in cost by an order of magnitude, and a fixed stride over
// them left some workgroups with several of the largest while the rest had
// finished.  A pass without ranged blocks (a single scan's: fewer items than
// workgroups) folds item w in workgroup w and touches the counter only for
// items beyond the grid.
constexpr int kFoldBThreads = 1024;
template <class State>
__global__ void __launch_bounds__(kFoldBThreads, 8) k_fold_b(Dev D) {
  static_assert(2 * kCountTile <= kFoldCap, "counting tiles share the sort buffer");
  __shared__ unsigned long long buf[kFoldCap];
  __shared__ uint32_t cnt[kVoxelsPerBlock];
  __shared__ uint32_t start[kVoxelsPerBlock + 1];
  __shared__ uint32_t sh[kScanSh];
  __shared__ uint32_t rstart[kMaxRanges + 1];
  __shared__ PassScans ps;
  __shared__ int32_t bnd[kMaxScansPerBatch];
  __shared__ int next_item, flag, n_marked;
  __shared__ int marked[kFoldBThreads];
  constexpr int kSumWindows = (int)(kFoldCap * sizeof(unsigned long long) / (kVoxelsPerBlock * sizeof(ulonglong2)));
  FoldTally tally;
  stage_scans(D, ps);
  const int n_wg2 = ld_acq(&D.ctr->fold_n[FC_WG2]);
  const int n_huge = ld_acq(&D.ctr->fold_n[FC_HUGE]);
  const int n_wg1 = ld_acq(&D.ctr->fold_n[FC_WG]);
  const int4* desc = D.fold_desc + D.fold_base[FC_WG2];
  const int4* hdesc = D.fold_desc + D.fold_base[FC_HUGE];
  const int4* wdesc = D.fold_desc + D.fold_base[FC_WG];
  __syncthreads();
  // (k_fold_a folds the FC_WG2 blocks of a one-scan pass)
  const int n_wg = ps.ns == 1 ? 0 : n_wg2;
  auto fold_item = [&](int k) {
    if (k < n_huge) {
      const int4 d = resolve_desc(D, hdesc[k]);
      if (fold_sum_block<State, kFoldBThreads, true>(D, d, buf, kSumWindows, ps, bnd, &flag, tally) &&
          !fold_ranged_block<State, kFoldBThreads>(D, d, buf, buf + kCountTile, cnt, start, rstart, sh, ps, tally))
        fold_point_rounds<State, kFoldBThreads>(D, d, buf, buf + kCountTile, cnt, start, sh, ps, tally);
    } else {
      const int4 d = resolve_desc(D, desc[k - n_huge]);
      if (fold_sum_block<State, kFoldBThreads, true>(D, d, buf, kSumWindows, ps, bnd, &flag, tally))
        fold_count_block<State, kCountTile, kFoldBThreads>(D, d, buf, buf + kCountTile, cnt, start, sh, ps, tally);
    }
  };
  // entries inserted in this pass that got no heap block leave the table
  // (binning, which may insert keys, is over)
  const int nf = ld_acq(&D.ctr->fail_n);
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < nf; q += gridDim.x * blockDim.x)
    D.keys[D.fail_list[q]] = kEmptyKey;
  const int total = n_huge + n_wg;
  int q0 = 0;  // the first item taken from the counter
  if (n_huge == 0) {
    if ((int)blockIdx.x < total) fold_item((int)blockIdx.x);
    q0 = (int)gridDim.x;
  }
  // with ranged blocks present every item comes from the counter
  while (q0 < total) {
    if (threadIdx.x == 0) next_item = q0 + atomicAdd(&D.ctr->fold_next, 1);
    __syncthreads();
    const int k = next_item;
    __syncthreads();
    if (k >= total) break;
    fold_item(k);
  }
  // The blocks k_fold_a marked, folded in order: the FC_WG list and, in a
  // pass of one scan, the FC_WG2 list, as one index range, dealt out by a
  // stride of the grid (workgroup w checks w, w + grid, ... with one load per
  // thread), so the few marked blocks of a pass spread over the workgroups.
  const int nm = n_wg1 + (ps.ns == 1 ? n_wg2 : 0);
  for (int r0 = 0; r0 < nm; r0 += (int)gridDim.x * kFoldBThreads) {
    if (threadIdx.x == 0) n_marked = 0;
    __syncthreads();
    const int x = r0 + (int)blockIdx.x + (int)threadIdx.x * (int)gridDim.x;
    if (x < nm && ((x < n_wg1 ? wdesc[x].w : desc[x - n_wg1].w) & kDescDeferred)) marked[atomicAdd(&n_marked, 1)] = x;
    __syncthreads();
    const int nk = n_marked;
    for (int j = 0; j < nk; ++j) {
      const int y = marked[j];
      int4 d = y < n_wg1 ? wdesc[y] : desc[y - n_wg1];
      d.w &= ~kDescDeferred;
      fold_count_block<State, kCountTile, kFoldBThreads>(D, resolve_desc(D, d), buf, buf + kCountTile, cnt, start, sh,
                                                          ps, tally);
    }
    __syncthreads();
  }
  const uint32_t row[2] = {(uint32_t)tally.uniq, tally.dropped};
  write_stat_row(D, SR_FOLD_B, row);
}

// ---- publish ---------------------------------------------------------------------------
// Selects the map's occupied entries; mode 0: all, 1: within the publish
// radius of the sensor (H:812).
// Each workgroup takes kSelChunk consecutive entries, gathers its selection
// in LDS and appends it with one atomic (a per-wave append would put tens of
// thousands of atomics on one counter).
constexpr int kSelChunk = 2048;

__global__ void __launch_bounds__(256) k_select(Dev D, int map_slot, float sx, float sy, float sz, int mode) {
  __shared__ uint32_t lsel[kSelChunk];
  __shared__ int ln, lbase;
  const MapDev m = D.maps[map_slot];
  const int64_t e0 = (int64_t)m.bucket_base * kBucket;
  const int64_t ne = (int64_t)(m.n_home + m.n_ovf) * kBucket;
  const float s[3] = {sx, sy, sz};
  for (int64_t c0 = (int64_t)blockIdx.x * kSelChunk; c0 < ne; c0 += (int64_t)gridDim.x * kSelChunk) {
    if (threadIdx.x == 0) ln = 0;
    __syncthreads();
    for (int u = 0; u < kSelChunk; u += 256) {
      const int64_t x = c0 + u + threadIdx.x;
      bool pred = false;
      if (x < ne) {
        const unsigned long long k = D.keys[e0 + x];
        if (k != kEmptyKey && D.vals[e0 + x] >= 0) {
          if (mode == 0) {
            pred = true;
          } else {
            int32_t kb[3];
            unpack_key(k, kb);
            float c[3];
            block_center(kb, m.BL, c);
            pred = dist2(c, s) <= m.pubD2;
          }
        }
      }
      const int at = wave_append(&ln, pred);
      if (pred) lsel[at] = (uint32_t)(e0 + x);
    }
    __syncthreads();
    if (threadIdx.x == 0) lbase = ln ? atomicAdd(&D.ctr->sel_n, ln) : 0;
    __syncthreads();
    for (int i = threadIdx.x; i < ln; i += blockDim.x) D.sel[lbase + i] = lsel[i];
    __syncthreads();
  }
}

template <class State>
__global__ void __launch_bounds__(256) k_pub_count(Dev D) {
  const int n = ld_acq(&D.ctr->sel_n);
  const int wpb = blockDim.x >> 6;
  for (int q = blockIdx.x * wpb + (threadIdx.x >> 6); q < n; q += gridDim.x * wpb) {
    const int slot = D.vals[D.sel[q]];
    uint32_t c = 0;
    for (int r = 0; r < kVoxelsPerBlock / 64; ++r) {
      c += State::nonzero(D.heap, (int64_t)slot * kVoxelsPerBlock + r * 64 + lane_id()) ? 1u : 0u;
    }
    const unsigned long long s = wave_sum(c);
    if (lane_id() == 0) D.pub_cnt[q] = (uint32_t)s;
  }
}

// layout 0: pos float[n][3] + data float[n][2]; 1: Voxel.msg {sdf, weight, x, y, z}
template <class State>
__global__ void __launch_bounds__(256) k_pub_write(Dev D, int map_slot, const uint32_t* off, float* pos, float* data,
                                                   int64_t cap, int layout) {
  const MapDev m = D.maps[map_slot];
  const int n = ld_acq(&D.ctr->sel_n);
  const int wpb = blockDim.x >> 6;
  const Codec cd = make_codec(m);
  for (int q = blockIdx.x * wpb + (threadIdx.x >> 6); q < n; q += gridDim.x * wpb) {
    const uint32_t e = D.sel[q];
    const int slot = D.vals[e];
    int32_t kb[3];
    unpack_key(D.keys[e], kb);
    float c[3];
    block_center(kb, m.BL, c);
    const float bl[3] = {c[0] - m.half_BL, c[1] - m.half_BL, c[2] - m.half_BL};
    int64_t o = off[q];
    for (int r = 0; r < kVoxelsPerBlock / 64; ++r) {
      const int l = r * 64 + lane_id();
      const int64_t vi = (int64_t)slot * kVoxelsPerBlock + l;
      float S, W;
      State::load(D.heap, vi, cd, S, W);
      const bool pred = W != 0.0f;
      const unsigned long long mask = __ballot(pred);
      const int64_t at = o + (int64_t)__popcll(mask & ((1ull << lane_id()) - 1ull));
      if (pred && at < cap) {
        float v[3];
        voxel_position(l, bl, m.vs, m.half_vs, v);
        if (layout == 0) {
          pos[at * 3 + 0] = v[0];
          pos[at * 3 + 1] = v[1];
          pos[at * 3 + 2] = v[2];
          data[at * 2 + 0] = S;
          data[at * 2 + 1] = W;
        } else {
          float* mm = pos + at * 5;
          mm[0] = S;
          mm[1] = W;
          mm[2] = v[0];
          mm[3] = v[1];
          mm[4] = v[2];
        }
      }
      o += (int64_t)__popcll(mask);
    }
  }
}

// ---- export (test/diagnostic dump of a map) -----------------------------------------------
template <class State>
__global__ void __launch_bounds__(256) k_export(Dev D, int map_slot, int32_t* keys_out, float2* vox_out) {
  const MapDev m = D.maps[map_slot];
  const int n = ld_acq(&D.ctr->sel_n);
  const Codec cd = make_codec(m);
  for (int q = blockIdx.x; q < n; q += gridDim.x) {
    const uint32_t e = D.sel[q];
    const int slot = D.vals[e];
    if (threadIdx.x == 0) unpack_key(D.keys[e], keys_out + 3 * (int64_t)q);
    for (int l = threadIdx.x; l < kVoxelsPerBlock; l += blockDim.x) {
      float S, W;
      State::load(D.heap, (int64_t)slot * kVoxelsPerBlock + l, cd, S, W);
      vox_out[(int64_t)q * kVoxelsPerBlock + l] = make_float2(S, W);
    }
  }
}

// ---- import (snapshot restore: the inverse of k_export) -------------------------------------
// Block q of the snapshot is found or inserted in the map's hash; a new entry
// takes a heap block under the same pool and quota rules as integration
// (an entry that gets none leaves the table again).  Then all 512 voxels of
// the block are overwritten.  cnt[0] counts stored blocks, cnt[1] refused ones.
// Keys are distinct (checked by the host), so no two workgroups share an entry.
template <class State>
__global__ void __launch_bounds__(256) k_import(Dev D, int map_slot, const int32_t* keys, const float2* vox, int64_t n,
                                                int* cnt) {
  __shared__ int slot_sh;
  const MapDev m = D.maps[map_slot];
  const Codec cd = make_codec(m);
  for (int64_t q = blockIdx.x; q < n; q += gridDim.x) {
    if (threadIdx.x == 0) {
      int slot = -1;
      bool ins = false;
      const int e = find_or_insert(D, m, pack_key(keys[3 * q], keys[3 * q + 1], keys[3 * q + 2]), ins);
      if (e >= 0 && !ins) {
        slot = D.vals[e];
      } else if (e >= 0) {
        const int held = atomicAdd(&D.mctr[map_slot].blocks, 1);
        bool ok = !(m.block_quota > 0 && held >= m.block_quota);
        if (ok) {
          const int idx = atomicSub(&D.ctr->heap_top, 1) - 1;
          if (idx < 0) {
            atomicAdd(&D.ctr->heap_top, 1);
            ok = false;
          } else {
            slot = D.heap_free[idx];
          }
        }
        if (ok) {
          D.vals[e] = slot;
        } else {
          atomicSub(&D.mctr[map_slot].blocks, 1);
          D.vals[e] = -1;
          atomicExch(&D.keys[e], kEmptyKey);
        }
      }
      atomicAdd(&cnt[slot >= 0 ? 0 : 1], 1);
      slot_sh = slot;
    }
    __syncthreads();
    const int slot = slot_sh;
    if (slot >= 0)
      for (int l = threadIdx.x; l < kVoxelsPerBlock; l += blockDim.x) {
        const float2 v = vox[q * kVoxelsPerBlock + l];
        State::store(D.heap, (int64_t)slot * kVoxelsPerBlock + l, cd, v.x, v.y);
      }
    __syncthreads();
  }
}

// ---- downstream consumer: markers of a Tsdf message --------------------------------------
// A stable split of Voxel.msg records {sdf, weight, x, y, z} by the sign of
// sdf.  k_msg_count: records of each tile of kMsgTile with sdf >= 0;
// k_msg_split: the tile's offset is the sum of the counts of the tiles
// before it (read by the workgroup itself: a message has a few hundred
// tiles), each record's rank inside its tile comes from a workgroup scan
// over the tile in message order, and the record's position goes to the
// sdf >= 0 list at the rank, or to the sdf < 0 list at its index minus it.
constexpr int kMsgTile = 4096;
__global__ void __launch_bounds__(256) k_msg_count(const float* __restrict__ msg, int64_t n, uint32_t* tile_cnt) {
  __shared__ uint32_t sh[kScanSh];
  const int64_t base = (int64_t)blockIdx.x * kMsgTile;
  uint32_t c = 0;
  for (int k = threadIdx.x; k < kMsgTile; k += blockDim.x) {
    const int64_t e = base + k;
    if (e < n && msg[5 * e] >= 0.f) ++c;
  }
  uint32_t tot;
  (void)block_scan_excl(c, sh, tot);
  if (threadIdx.x == 0) tile_cnt[blockIdx.x] = tot;
}
__global__ void __launch_bounds__(256) k_msg_split(const float* __restrict__ msg, int64_t n, const uint32_t* tile_cnt,
                                                   float* nonneg, float* neg) {
  __shared__ uint32_t sh[kScanSh];
  __shared__ unsigned long long carry_sh;
  // the tile's offset
  unsigned long long part = 0;
  for (int t = threadIdx.x; t < (int)blockIdx.x; t += blockDim.x) part += tile_cnt[t];
  part = wave_sum(part);
  if (threadIdx.x == 0) carry_sh = 0;
  __syncthreads();
  if (lane_id() == 0) atomicAdd(&carry_sh, part);
  __syncthreads();
  unsigned long long g = carry_sh;  // records with sdf >= 0 before this chunk
  const int64_t base = (int64_t)blockIdx.x * kMsgTile;
  for (int k0 = 0; k0 < kMsgTile; k0 += blockDim.x) {
    const int64_t e = base + k0 + threadIdx.x;
    const bool in = e < n;
    const bool pos = in && msg[5 * e] >= 0.f;
    uint32_t tot;
    const uint32_t r = block_scan_excl(pos ? 1u : 0u, sh, tot);
    if (in) {
      const unsigned long long before = g + r;  // sdf >= 0 records before e
      float* dst = pos ? nonneg + 3 * before : neg + 3 * ((unsigned long long)e - before);
      dst[0] = msg[5 * e + 2];
      dst[1] = msg[5 * e + 3];
      dst[2] = msg[5 * e + 4];
    }
    g += tot;
  }
}

// ---- garbage collection -----------------------------------------------------------------
// mode 0: remove blocks beyond the gc radius (H:836-868); mode 1: remove all
// (map teardown).
// Removes the map's blocks beyond the GC radius (mode 0) or all of them
// (mode 1).  Each workgroup takes kSelChunk consecutive entries and pushes
// its freed heap blocks with one atomic per counter.
__global__ void __launch_bounds__(256) k_gc_mark(Dev D, int map_slot, float sx, float sy, float sz, int mode) {
  __shared__ int32_t lfree[kSelChunk];
  __shared__ int ln, ltop, lfb;
  const MapDev m = D.maps[map_slot];
  const int64_t e0 = (int64_t)m.bucket_base * kBucket;
  const int64_t ne = (int64_t)(m.n_home + m.n_ovf) * kBucket;
  const float s[3] = {sx, sy, sz};
  unsigned long long removed = 0;
  for (int64_t c0 = (int64_t)blockIdx.x * kSelChunk; c0 < ne; c0 += (int64_t)gridDim.x * kSelChunk) {
    if (threadIdx.x == 0) ln = 0;
    __syncthreads();
    for (int u = 0; u < kSelChunk; u += 256) {
      const int64_t x = c0 + u + threadIdx.x;
      int slot = -1;
      if (x < ne) {
        const unsigned long long k = D.keys[e0 + x];
        if (k != kEmptyKey) {
          bool drop = true;
          if (mode == 0) {
            int32_t kb[3];
            unpack_key(k, kb);
            float c[3];
            block_center(kb, m.BL, c);
            drop = !(dist2(c, s) <= m.gcD2);
          }
          if (drop) {
            slot = D.vals[e0 + x];
            D.keys[e0 + x] = kEmptyKey;
            D.vals[e0 + x] = -1;
            removed++;
          }
        }
      }
      const int at = wave_append(&ln, slot >= 0);
      if (slot >= 0) lfree[at] = slot;
    }
    __syncthreads();
    if (threadIdx.x == 0 && ln) {
      ltop = atomicAdd(&D.ctr->heap_top, ln);
      lfb = atomicAdd(&D.ctr->freed_n, ln);
      atomicSub(&D.mctr[map_slot].blocks, ln);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < ln; i += blockDim.x) {
      D.heap_free[ltop + i] = lfree[i];
      D.freed[lfb + i] = lfree[i];
    }
    __syncthreads();
  }
  stat_add(D, ST_GC_REMOVED, removed);
}

template <class State>
__global__ void __launch_bounds__(256) k_gc_zero(Dev D) {
  const int n = ld_acq(&D.ctr->freed_n);
  for (int q = blockIdx.x; q < n; q += gridDim.x) {
    const int slot = D.freed[q];
    for (int l = threadIdx.x; l < kVoxelsPerBlock; l += blockDim.x) State::zero(D.heap, (int64_t)slot * kVoxelsPerBlock + l);
  }
}

// Unlink overflow buckets that gc emptied and return them to the map's stack.
__global__ void __launch_bounds__(256) k_gc_chains(Dev D, int map_slot) {
  const MapDev m = D.maps[map_slot];
  MapCtr* mc = &D.mctr[map_slot];
  for (int h = blockIdx.x * blockDim.x + threadIdx.x; h < m.n_home; h += gridDim.x * blockDim.x) {
    int prev = m.bucket_base + h;
    int cur = D.next[prev];
    if (cur == kLinkExhausted) D.next[prev] = kLinkNone;  // buckets may be back in the pool
    for (int guard = 0; cur >= 0 && guard <= m.n_ovf; ++guard) {
      const int nx = D.next[cur];
      bool empty = true;
      for (int s = 0; s < kBucket; ++s) empty = empty && (D.keys[(int64_t)cur * kBucket + s] == kEmptyKey);
      if (empty) {
        D.next[prev] = nx;
        D.next[cur] = -1;
        const int idx = atomicAdd(&mc->ovf_top, 1);
        D.ovf_stack[m.bucket_base + m.n_home + idx] = cur;
      } else {
        prev = cur;
      }
      cur = nx;
    }
    if (cur == kLinkExhausted) D.next[prev] = kLinkNone;
  }
}


// ---- input path: sensor_msgs/PointCloud2 -> pcl::PointXYZ layout -------------------
// pcl::fromPCLPointCloud2 (called by Transformer::convertPointCloud2ToPointCloudXYZ,
// tsdf_package/src/transformer.cpp:63-67) copies the FLOAT32 x, y, z fields
// of every point into PointXYZ {x, y, z, 1}.  Point i of a width x height
// cloud sits at row i / width, column i % width: byte row * row_step +
// column * point_step.  Non-finite values are copied as they are.
__device__ __forceinline__ float cloud_field(const uint8_t* p, bool swap, bool aligned) {
  uint32_t u;
  if (aligned) u = *reinterpret_cast<const uint32_t*>(p);
  else u = (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24);
  if (swap) u = __builtin_bswap32(u);
  return __uint_as_float(u);
}
template <bool Aligned>
__global__ void __launch_bounds__(256) k_cloud2(const uint8_t* __restrict__ cloud, int64_t width, int64_t n, int64_t row_step,
                                                int point_step, int xo, int yo, int zo, int swap, float4* __restrict__ out) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = i / width, c = i - r * width;
    const uint8_t* q = cloud + r * row_step + c * point_step;
    out[i] = make_float4(cloud_field(q + xo, swap, Aligned), cloud_field(q + yo, swap, Aligned),
                         cloud_field(q + zo, swap, Aligned), 1.0f);
  }
}

// ---- plain device copy (also the reference point for HBM byte counters) ------------
__global__ void __launch_bounds__(256) k_copy(const float4* __restrict__ src, float4* __restrict__ dst, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

// ---- initialisation ----------------------------------------------------------------------
__global__ void k_fill_u64(unsigned long long* p, int64_t n, unsigned long long v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_fill_i32(int32_t* p, int64_t n, int32_t v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_iota_desc(int32_t* p, int64_t n, int32_t first) {
  // p[i] = first + n - 1 - i : a stack whose top pops `first` first
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = first + (int32_t)(n - 1 - i);
}
__global__ void k_zero_ctr(EngCtr* c) {
  if (threadIdx.x < ST_COUNT) c->stats[threadIdx.x] = 0;
  if (threadIdx.x == 0) {
    c->sel_n = 0;
    c->freed_n = 0;
  }
}

}  // namespace tsdfhip
