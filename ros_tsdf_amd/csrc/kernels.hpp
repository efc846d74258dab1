// Kernels of the integration pipeline.  One pass integrates the scans of a
// batch (a single scan on the per-scan path):
//
//   k_pass_begin     zero the pass counters
//   k_bin            per point: truncation segment, block DDA and voxel DDA;
//                    find-or-insert every block in the map's spatial hash
//                    (lock-free) once per workgroup; each run of voxels in
//                    one block gets its record offset in that block
//   k_assign         pop a heap block from the free stack for each new entry
//   k_scan_*         exclusive scan of the per-block record counts
//   k_emit           per point: second voxel-DDA walk writing each update
//                    record (voxel, point, sdf) to its block's contiguous
//                    record segment at the offset reserved in k_bin
//   k_fold_prep      one 16-byte descriptor per touched block, filed by size
//   k_fold_small     blocks of <= 16 / 32 / 64 records, 4 / 2 / 1 per wave:
//                    sort by (voxel, point) across lanes with shuffles; each
//                    voxel's first record folds the voxel's updates in point
//                    order (one read + one write per voxel)
//   k_fold_count     65..4096 records, one workgroup: counting sort by voxel
//                    in LDS, per-voxel insertion sort by point, fold
//   k_fold_huge      larger blocks: sorted LDS tiles merged per voxel
//
// publish: k_select (radius test per entry) -> k_pub_count (wave per block)
//          -> scan -> k_pub_write (wave ballot compaction)
// gc:      k_gc_mark -> k_gc_zero -> k_gc_chains -> k_gc_orphans
#pragma once
#include "device.hpp"

namespace tsdfhip {

constexpr int kFoldThreads = 256;
constexpr int kCountTileS = 1024; // records per LDS tile of the small k_fold_count (2 x 8 KB)
constexpr int kCountTile = 4096;  // records per LDS tile of the large k_fold_count (2 x 32 KB)
constexpr int kFoldCap = 8192;    // records per LDS sort tile of k_fold_huge (64 KB)
constexpr int kMaxChunks = 32;    // sort tiles per block in k_fold_huge
constexpr int kScanTile = 1024;  // elements per workgroup in the scan kernels

// ---- wave / workgroup helpers -------------------------------------------------
__device__ __forceinline__ int lane_id() { return __lane_id(); }

// All lanes of the wave must call this.  Returns the slot of this lane's item
// (or -1 when pred is false) in a list whose length is *ctr.
__device__ __forceinline__ int wave_append(int32_t* ctr, bool pred) {
  const unsigned long long m = __ballot(pred);
  if (m == 0) return -1;
  const int lane = lane_id();
  const int leader = __ffsll((long long)m) - 1;
  int base = 0;
  if (lane == leader) base = atomicAdd(ctr, (int)__popcll(m));
  base = __shfl(base, leader);
  const int rank = (int)__popcll(m & ((1ull << lane) - 1ull));
  return pred ? base + rank : -1;
}

__device__ __forceinline__ unsigned long long wave_sum(unsigned long long v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
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

template <int N>
__device__ __forceinline__ void flush_stats(const Dev& D, const int* which, const unsigned long long (&v)[N]) {
#pragma unroll
  for (int i = 0; i < N; ++i) stat_add(D, which[i], v[i]);
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
// empty to full while a pass runs (deletion happens in gc, a separate pass),
// so every inserter of one key walks the same sequence of empty slots and
// the first CAS winner is seen by the others: no duplicates.  A full chain
// is extended by popping an overflow bucket and CAS-ing the tail link; the
// loser of that race files its bucket on the orphan list (recycled by gc).
// Returns the global entry index, or -1 when the chain limit or the overflow
// pool is exhausted.  *inserted is set when this call created the entry.
// Reads are plain loads: a stale read can only show an older value of a slot
// (empty instead of filled) or of a link (none instead of set); the CAS that
// every insert and extension goes through then fails and reports the current
// value, so staleness costs a retry, never a duplicate.
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
    MapCtr* mc = &D.mctr[m.slot];
    const int idx = atomicSub(&mc->ovf_top, 1) - 1;
    if (idx < 0) {
      atomicAdd(&mc->ovf_top, 1);
      return -1;
    }
    const int nb = D.ovf_stack[m.bucket_base + m.n_home + idx];
    const int prev = atomicCAS(&D.next[tail], -1, nb);
    if (prev == -1) {
      g = nb;
    } else {
      const int o = atomicAdd(&mc->orphan_n, 1);
      D.orphans[m.bucket_base + m.n_home + o] = nb;
      g = prev;
    }
    s0 = 0;
    depth = tail_depth + 1;
  }
  return -1;
}

// ---- pass kernels -----------------------------------------------------------------
__global__ void k_pass_begin(Dev D) {
  const int t = threadIdx.x;
  if (t == 0) {
    D.ctr->touched_n = 0;
    D.ctr->new_n = 0;
    D.ctr->big_n = 0;
  }
  if (t < 6) D.ctr->fold_n[t] = 0;
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

__device__ __forceinline__ void note_new(const Dev& D, int e, int map_slot) {
  const int q = atomicAdd(&D.ctr->new_n, 1);
  D.new_list[q] = (uint32_t)e;
  D.new_map[q] = map_slot;
}

// Binning: per point, the truncation segment, the block DDA (every cell is
// allocated, H:255-272 and H:421-500) and the voxel DDA (H:675-710) walked
// for its cells only.  Along a walk every axis is monotone, so the voxels of
// one block form one run.  Neighbouring rays share blocks, so each workgroup
// first gathers its points' blocks in an LDS table: a block is looked up (or
// inserted) in the map's spatial hash once per workgroup, and reserves the
// workgroup's records in the block with one 64-bit atomic (runs in the low
// half, records in the high half).  Each run then learns its record offset
// in the block, stored with its entry for k_emit.  A full LDS table falls
// back to per-run global atomics.
constexpr int kBinSlots = 1024;
constexpr uint32_t kPairFail = 0xFFFFFFFFu;
constexpr uint32_t kPairDirect = 0x80000000u;

__device__ __forceinline__ unsigned long long lds_block_key(int map_slot, const int32_t kb[3]) {
  // 16 bits per axis suffice under the point validity rule (|block| < 2^14)
  return ((unsigned long long)(uint32_t)map_slot << 48) | ((unsigned long long)(kb[0] & 0xFFFF) << 32) |
         ((unsigned long long)(kb[1] & 0xFFFF) << 16) | (unsigned long long)(kb[2] & 0xFFFF);
}

__device__ __forceinline__ int lds_find_or_add(unsigned long long* tkey, unsigned long long k) {
  uint32_t h = hash_key(k) & (kBinSlots - 1);
  for (int probe = 0; probe < 64; ++probe) {
    const unsigned long long cur = tkey[h];
    if (cur == k) return (int)h;
    if (cur == kEmptyKey) {
      const unsigned long long old = atomicCAS(&tkey[h], kEmptyKey, k);
      if (old == kEmptyKey || old == k) return (int)h;
    }
    h = (h + 1) & (kBinSlots - 1);
  }
  return -1;
}

__global__ void __launch_bounds__(256) k_bin(Dev D) {
  __shared__ unsigned long long tkey[kBinSlots];
  __shared__ int tent[kBinSlots];
  __shared__ uint32_t trec[kBinSlots];
  __shared__ uint32_t trun[kBinSlots];
  __shared__ uint32_t tbase[kBinSlots];
  __shared__ uint32_t lnew[kBinSlots];
  __shared__ int lnew_map[kBinSlots];
  __shared__ uint32_t ltouch[kBinSlots];
  __shared__ uint8_t ltouch_map[kBinSlots];
  __shared__ uint16_t locc[kBinSlots];
  __shared__ int n_new, n_touch, g_new, g_touch, n_occ;
  const PassHdr hdr = *D.hdr;
  const int P = hdr.total;
  unsigned long long st[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  const int which[8] = {ST_VALID, ST_SKIPPED, ST_BLOCK_CELLS, ST_UPDATES,
                        ST_TRUNC_VOX, ST_TRUNC_BLK, ST_INSERT_FAIL, ST_UPD_DROPPED};
  const int64_t stride = D.pcap;
  for (int base = blockIdx.x * blockDim.x; base < P; base += gridDim.x * blockDim.x) {
    for (int x = threadIdx.x; x < kBinSlots; x += blockDim.x) {
      tkey[x] = kEmptyKey;
      trec[x] = 0;
      trun[x] = 0;
    }
    if (threadIdx.x == 0) {
      n_new = 0;
      n_touch = 0;
      n_occ = 0;
    }
    __syncthreads();
    // ---- phase A: walks; blocks gathered in the LDS table ----
    const int i = base + threadIdx.x;
    int nj = 0;
    if (i < P) {
      const ScanDev sd = D.scans[find_scan(D.scans, hdr.nscans, i)];
      const MapDev m = D.maps[sd.map];
      const float4 q = hdr.points[i];
      const float p[3] = {q.x, q.y, q.z};
      const float s[3] = {sd.s[0], sd.s[1], sd.s[2]};
      if (!point_valid(p, s, m.vs)) {
        st[1]++;
      } else {
        st[0]++;
        float a[3], b[3];
        truncation_segment(p, s, m.T, a, b);
        bool tb = false;
        const int nb = walk_cells(a, b, m.BL, m.sc_b, m.half_BL, m.max_blk, tb, [&](int32_t k0, int32_t k1, int32_t k2) {
          const int32_t kb[3] = {k0, k1, k2};
          if (lds_find_or_add(tkey, lds_block_key(sd.map, kb)) < 0) {
            bool ins;
            const int e = find_or_insert(D, m, pack_key(k0, k1, k2), ins);
            if (e < 0) st[6]++;
            else if (ins) note_new(D, e, m.slot);
          }
          return true;
        });
        st[2] += nb;
        st[5] += tb ? 1 : 0;
        unsigned long long rkey = kEmptyKey;
        int32_t rk[3] = {0, 0, 0};
        int rn = 0;
        auto close_run = [&]() {
          if (rn == 0) return;
          const int slot = lds_find_or_add(tkey, rkey);
          if (slot >= 0) {
            const uint32_t local = atomicAdd(&trec[slot], (uint32_t)rn);
            atomicAdd(&trun[slot], 1u);
            D.pair_e[(int64_t)nj * stride + i] = (uint32_t)slot;
            D.pair_r[(int64_t)nj * stride + i] = local | ((uint32_t)rn << 24);
          } else {
            bool ins;
            const int e = find_or_insert(D, m, pack_key(rk[0], rk[1], rk[2]), ins);
            if (e < 0) {
              st[6]++;
              st[7] += rn;
              D.pair_e[(int64_t)nj * stride + i] = kPairFail;
            } else {
              if (ins) note_new(D, e, m.slot);
              const unsigned long long old = atomicAdd(&D.blk_cnt[e], ((unsigned long long)rn << 32) | 1ull);
              if (old == 0) {
                const int t = atomicAdd(&D.ctr->touched_n, 1);
                D.touched[t] = (uint32_t)e;
                D.touched_map[t] = (uint8_t)m.slot;
                D.touch_idx[e] = (uint32_t)t;
              }
              D.pair_e[(int64_t)nj * stride + i] = kPairDirect | (uint32_t)e;
              D.pair_r[(int64_t)nj * stride + i] = (uint32_t)(old >> 32);
              st[3] += rn;
            }
          }
          ++nj;
          rn = 0;
        };
        bool tv = false;
        walk_cells(a, b, m.vs, m.sc_v, m.half_vs, m.max_vox, tv, [&](int32_t k0, int32_t k1, int32_t k2) {
          const int32_t kb[3] = {block_of(k0), block_of(k1), block_of(k2)};
          const unsigned long long key = lds_block_key(sd.map, kb);
          if (key != rkey) {
            close_run();
            rkey = key;
            rk[0] = kb[0];
            rk[1] = kb[1];
            rk[2] = kb[2];
          }
          ++rn;
          return true;
        });
        close_run();
        st[4] += tv ? 1 : 0;
      }
      D.pair_n[i] = (uint8_t)nj;
    }
    __syncthreads();
    // ---- phase B: one global lookup and one counter atomic per gathered block ----
    for (int x0 = 0; x0 < kBinSlots; x0 += blockDim.x) {
      const int x = x0 + threadIdx.x;
      const bool occ = tkey[x] != kEmptyKey;
      const int at = wave_append(&n_occ, occ);
      if (occ) locc[at] = (uint16_t)x;
    }
    __syncthreads();
    for (int y = threadIdx.x; y < n_occ; y += blockDim.x) {
      const int x = locc[y];
      const unsigned long long k = tkey[x];
      const int ms = (int)(k >> 48);
      const MapDev& m = D.maps[ms];
      const int32_t k0 = (int32_t)(int16_t)((k >> 32) & 0xFFFF), k1 = (int32_t)(int16_t)((k >> 16) & 0xFFFF),
                    k2 = (int32_t)(int16_t)(k & 0xFFFF);
      bool ins;
      const int e = find_or_insert(D, m, pack_key(k0, k1, k2), ins);
      tent[x] = e;
      if (e < 0) {
        st[6]++;
        continue;
      }
      if (ins) {
        const int at = atomicAdd(&n_new, 1);
        lnew[at] = (uint32_t)e;
        lnew_map[at] = ms;
      }
      const uint32_t nrec = trec[x];
      if (nrec > 0) {
        const unsigned long long old = atomicAdd(&D.blk_cnt[e], ((unsigned long long)nrec << 32) | trun[x]);
        tbase[x] = (uint32_t)(old >> 32);
        if (old == 0) {
          const int at = atomicAdd(&n_touch, 1);
          ltouch[at] = (uint32_t)e;
          ltouch_map[at] = (uint8_t)ms;
        }
      }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      g_new = n_new ? atomicAdd(&D.ctr->new_n, n_new) : 0;
      g_touch = n_touch ? atomicAdd(&D.ctr->touched_n, n_touch) : 0;
    }
    __syncthreads();
    for (int x = threadIdx.x; x < n_new; x += blockDim.x) {
      D.new_list[g_new + x] = lnew[x];
      D.new_map[g_new + x] = lnew_map[x];
    }
    for (int x = threadIdx.x; x < n_touch; x += blockDim.x) {
      const uint32_t e = ltouch[x];
      D.touched[g_touch + x] = e;
      D.touched_map[g_touch + x] = ltouch_map[x];
      D.touch_idx[e] = (uint32_t)(g_touch + x);
    }
    // ---- phase C: final (entry, record offset) of every run ----
    if (i < P) {
      for (int j = 0; j < nj; ++j) {
        const uint32_t v = D.pair_e[(int64_t)j * stride + i];
        if (v == kPairFail || (v & kPairDirect)) {
          if (v != kPairFail) D.pair_e[(int64_t)j * stride + i] = v & ~kPairDirect;
          continue;
        }
        const uint32_t pr = D.pair_r[(int64_t)j * stride + i];
        const uint32_t rn = pr >> 24;
        const int e = tent[v];
        if (e < 0) {
          D.pair_e[(int64_t)j * stride + i] = kPairFail;
          st[7] += rn;
        } else {
          D.pair_e[(int64_t)j * stride + i] = (uint32_t)e;
          D.pair_r[(int64_t)j * stride + i] = tbase[v] + (pr & 0xFFFFFFu);
          st[3] += rn;
        }
      }
    }
    __syncthreads();
  }
  flush_stats(D, which, st);
}

// Heap block for every entry inserted by the pass (free-list pop; pushes only
// happen in gc, never concurrently with pops, so the stack needs no ABA
// guard).  Both counters are updated once per wave: the per-map block count
// once per distinct map in the wave, the heap stack once; lanes past a quota
// or past the bottom of the stack hand their share back.
__global__ void __launch_bounds__(256) k_assign(Dev D) {
  const int n = ld_acq(&D.ctr->new_n);
  unsigned long long st[3] = {0, 0, 0};
  const int which[3] = {ST_NEW_BLOCKS, ST_HEAP_FAIL, ST_QUOTA_FAIL};
  const int lane = lane_id();
  const unsigned long long below = (1ull << lane) - 1ull;
  for (int q0 = blockIdx.x * blockDim.x; q0 < n; q0 += gridDim.x * blockDim.x) {
    const int q = q0 + threadIdx.x;
    const bool active = q < n;
    uint32_t e = 0;
    int ms = -1;
    if (active) {
      e = D.new_list[q];
      ms = D.new_map[q];
    }
    // per-map block count: one atomic per distinct map in the wave
    int held = 0;  // blocks the map held before this lane's block
    unsigned long long pending = __ballot(active);
    while (pending) {
      const int leader = __ffsll((long long)pending) - 1;
      const int lm = __shfl(ms, leader);
      const unsigned long long grp = __ballot(active && ms == lm);
      int base = 0;
      if (lane == leader) base = atomicAdd(&D.mctr[lm].blocks, (int)__popcll(grp));
      base = __shfl(base, leader);
      if (grp >> lane & 1ull) held = base + (int)__popcll(grp & below);
      pending &= ~grp;
    }
    bool ok = active;
    if (active) {
      const int64_t quota = D.maps[ms].block_quota;
      if (quota > 0 && held >= quota) {
        atomicSub(&D.mctr[ms].blocks, 1);
        ok = false;
        st[2]++;
      }
    }
    // heap pops: one atomic per wave
    const unsigned long long want = __ballot(ok);
    const int cnt = (int)__popcll(want);
    const int leader = want ? __ffsll((long long)want) - 1 : 0;
    int top = 0;
    if (want && lane == leader) {
      top = atomicSub(&D.ctr->heap_top, cnt);
      const int failed = min(cnt, cnt - max(top, 0));
      if (failed > 0) atomicAdd(&D.ctr->heap_top, failed);
    }
    top = __shfl(top, leader);
    const int idx = top - cnt + (int)__popcll(want & below);
    if (ok && idx < 0) {
      atomicSub(&D.mctr[ms].blocks, 1);
      ok = false;
    }
    if (active) {
      if (ok) {
        D.vals[e] = D.heap_free[idx];
        st[0]++;
      } else {
        D.vals[e] = -1;
        D.keys[e] = kEmptyKey;
        st[1]++;
      }
    }
  }
  flush_stats(D, which, st);
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

__device__ __forceinline__ uint32_t block_scan_excl(uint32_t v, uint32_t* sh, uint32_t& total) {
  // 256 threads, wave64: inclusive wave scan, then scan of wave totals.
  const int lane = lane_id();
  const int w = threadIdx.x >> 6;
  uint32_t x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(x, o);
    if (lane >= o) x += y;
  }
  if (lane == 63) sh[w] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t acc = 0;
    for (int k = 0; k < (int)(blockDim.x >> 6); ++k) {
      const uint32_t t = sh[k];
      sh[k] = acc;
      acc += t;
    }
    sh[15] = acc;
  }
  __syncthreads();
  total = sh[15];
  const uint32_t r = x - v + sh[w];
  __syncthreads();
  return r;
}

template <class Count>
__global__ void __launch_bounds__(256) k_scan_tiles(const int32_t* n_ptr, Count cnt, uint32_t* tile_sums) {
  __shared__ uint32_t sh[16];
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
  __shared__ uint32_t sh[16];
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
  __shared__ uint32_t sh[16];
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

// ---- records ---------------------------------------------------------------------------
// Record of one voxel update: local voxel (9 bits) | point index in the pass
// (23 bits) | FP32 signed distance.  Sorting records as integers orders them
// by voxel, then by point: the reference's per-voxel fold order.
__device__ __forceinline__ unsigned long long make_rec(int local, int point, float d) {
  return ((unsigned long long)local << 55) | ((unsigned long long)point << 32) | (unsigned long long)__float_as_uint(d);
}

// Second walk of every point's voxel DDA: each voxel's update record
// (getDistanceUpdate, H:562-573) is written to its block's record segment at
// the offset its run reserved in k_bin, so every block's records end up
// contiguous.
__global__ void __launch_bounds__(256) k_emit(Dev D) {
  const PassHdr hdr = *D.hdr;
  const int P = hdr.total;
  const int64_t stride = D.pcap;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < P; i += gridDim.x * blockDim.x) {
    if (D.pair_n[i] == 0) continue;
    const ScanDev sd = D.scans[find_scan(D.scans, hdr.nscans, i)];
    const MapDev m = D.maps[sd.map];
    const float4 q = hdr.points[i];
    const float p[3] = {q.x, q.y, q.z};
    const float s[3] = {sd.s[0], sd.s[1], sd.s[2]};
    float a[3], b[3];
    truncation_segment(p, s, m.T, a, b);
    unsigned long long rkey = kEmptyKey;
    int j = -1;
    int64_t at = -1;
    bool tv;
    walk_cells(a, b, m.vs, m.sc_v, m.half_vs, m.max_vox, tv, [&](int32_t k0, int32_t k1, int32_t k2) {
      const unsigned long long key = pack_key(block_of(k0), block_of(k1), block_of(k2));
      if (key != rkey) {
        rkey = key;
        ++j;
        const uint32_t e = D.pair_e[(int64_t)j * stride + i];
        at = e == 0xFFFFFFFFu ? -1 : (int64_t)D.seg_off[D.touch_idx[e]] + D.pair_r[(int64_t)j * stride + i];
      }
      if (at >= 0) {
        const int32_t kk[3] = {k0, k1, k2};
        const float vc[3] = {cell_centre_of(k0, m.sc_v, m.half_vs), cell_centre_of(k1, m.sc_v, m.half_vs),
                             cell_centre_of(k2, m.sc_v, m.half_vs)};
        D.seg[at++] = make_rec(local_index(kk), i, signed_distance(p, s, vc));
      }
      return true;
    });
  }
}

// ---- per-block fold ---------------------------------------------------------------------
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

__device__ __forceinline__ int next_pow2(int c) {
  int n = 2;
  while (n < c) n <<= 1;
  return n;
}

// H:637-654 with the update's weight w = 1/(|d|+1) and w*d precomputed:
// S = (W*S + w*d)/(W+w); W = min(W+w, maxW).
__device__ __forceinline__ void fold_step(float& S, float& W, float w, float wd, float maxW) {
  const float Wn = W + w;
  S = (W * S + wd) / Wn;
  W = fminf(Wn, maxW);
}
__device__ __forceinline__ void weight_of(float d, float& w, float& wd) {
  w = 1.0f / (fabsf(d) + 1.0f);
  wd = w * d;
}

__device__ __forceinline__ int rec_voxel(unsigned long long r) { return (int)(r >> 55); }

// Size classes of touched blocks (records per block).
enum FoldClass { FC_16 = 0, FC_32, FC_64, FC_WG, FC_WG2, FC_HUGE, FC_COUNT };

// One descriptor per touched block with its records: everything the fold
// needs in one 16-byte load.
__global__ void __launch_bounds__(256) k_fold_prep(Dev D) {
  __shared__ int lcnt[FC_COUNT], gbase[FC_COUNT];
  const int nt = ld_acq(&D.ctr->touched_n);
  unsigned long long dropped = 0;
  for (int b0 = blockIdx.x * blockDim.x; b0 < nt; b0 += gridDim.x * blockDim.x) {
    if (threadIdx.x < FC_COUNT) lcnt[threadIdx.x] = 0;
    __syncthreads();
    const int t = b0 + threadIdx.x;
    int cls = -1, rank = 0;
    int4 d = make_int4(0, 0, 0, 0);
    if (t < nt) {
      const uint32_t e = D.touched[t];
      const uint32_t base = D.seg_off[t];
      const int n = (int)(D.seg_off[t + 1] - base);
      const int slot = D.vals[e];
      D.blk_cnt[e] = 0;
      if (slot < 0) {
        dropped += n;
      } else {
        d = make_int4((int)base, n, slot, (int)D.touched_map[t]);
        cls = n <= 16 ? FC_16 : n <= 32 ? FC_32 : n <= 64 ? FC_64 : n <= kCountTileS ? FC_WG : n <= kCountTile ? FC_WG2 : FC_HUGE;
        rank = atomicAdd(&lcnt[cls], 1);
      }
    }
    __syncthreads();
    if (threadIdx.x < FC_COUNT) gbase[threadIdx.x] = lcnt[threadIdx.x] ? atomicAdd(&D.ctr->fold_n[threadIdx.x], lcnt[threadIdx.x]) : 0;
    __syncthreads();
    if (cls >= 0) D.fold_desc[(int64_t)cls * D.fold_cap + gbase[cls] + rank] = d;
    __syncthreads();
  }
  stat_add(D, ST_UPD_DROPPED, dropped);
}

// Blocks of at most G records, 64/G blocks per wave: each group of G lanes
// holds one block's records, bitonic-sorts them with shuffles, and every
// run of one voxel is folded by its first lane pulling the run's weight
// terms from the following lanes.
template <class State, int G>
__global__ void __launch_bounds__(256) k_fold_small(Dev D) {
  constexpr int kPer = 64 / G;
  const int lane = lane_id();
  const int gl = lane & (G - 1);
  const int grp = lane / G;
  const int n_desc = ld_acq(&D.ctr->fold_n[G == 16 ? FC_16 : G == 32 ? FC_32 : FC_64]);
  const int4* desc = D.fold_desc + (int64_t)(G == 16 ? FC_16 : G == 32 ? FC_32 : FC_64) * D.fold_cap;
  const int gw = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const int nw = gridDim.x * (blockDim.x >> 6);
  unsigned long long uniq = 0;
  for (int q0 = gw * kPer; q0 < n_desc; q0 += nw * kPer) {
    const int q = q0 + grp;
    int4 d = make_int4(0, 0, 0, 0);
    if (q < n_desc) d = desc[q];
    const int n = d.y;
    unsigned long long v = gl < n ? D.seg[(uint32_t)d.x + gl] : ~0ull;
#pragma unroll
    for (int k = 2; k <= G; k <<= 1) {
#pragma unroll
      for (int j = k >> 1; j > 0; j >>= 1) {
        const unsigned long long o = __shfl_xor(v, j);
        const bool keep_min = ((gl & j) == 0) == ((gl & k) == 0);
        v = keep_min ? (o < v ? o : v) : (o > v ? o : v);
      }
    }
    const unsigned long long prev = __shfl_up(v, 1);
    const bool head = gl < n && (gl == 0 || rec_voxel(prev) != rec_voxel(v));
    const unsigned long long heads = __ballot(head);
    const int gend = grp * G + n;  // first lane past this group's records
    const unsigned long long above = lane == 63 ? 0ull : (heads >> (lane + 1)) << (lane + 1);
    const int nh = above ? __ffsll((long long)above) - 1 : 64;
    const int len = head ? min(nh, gend) - lane : 0;
    float w, wd;
    weight_of(__uint_as_float((uint32_t)v), w, wd);
    const MapDev& m = D.maps[d.w];
    const float maxW = m.maxW, Tt = m.T, invT = 1.0f / m.T;
    const int64_t vi = (int64_t)d.z * kVoxelsPerBlock + rec_voxel(v);
    float S = 0.f, W = 0.f;
    if (head) State::load(D.heap, vi, invT, Tt, S, W);
    int maxlen = len;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) maxlen = max(maxlen, __shfl_xor(maxlen, o));
    for (int r = 0; r < maxlen; ++r) {
      const float wr = __shfl(w, lane + r);
      const float wdr = __shfl(wd, lane + r);
      if (r < len) fold_step(S, W, wr, wdr, maxW);
    }
    if (head) {
      State::store(D.heap, vi, invT, Tt, S, W);
      uniq++;
    }
  }
  stat_add(D, ST_UNIQUE, uniq);
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

// Blocks of 65..kCountTile records, one workgroup each: a counting sort by
// voxel in LDS (histogram, scan, scatter) groups every voxel's updates;
// each voxel's thread then orders its few updates by point (insertion sort:
// a voxel receives at most a few dozen updates per pass) and folds them.
template <class State, int TILE, int CLS>
__global__ void __launch_bounds__(256) k_fold_count(Dev D) {
  __shared__ unsigned long long rin[TILE];
  __shared__ unsigned long long rout[TILE];
  __shared__ uint32_t cnt[kVoxelsPerBlock];
  __shared__ uint32_t start[kVoxelsPerBlock + 1];
  __shared__ uint32_t sh[16];
  const int n_desc = ld_acq(&D.ctr->fold_n[CLS]);
  const int4* desc = D.fold_desc + (int64_t)CLS * D.fold_cap;
  unsigned long long uniq = 0;
  for (int q = blockIdx.x; q < n_desc; q += gridDim.x) {
    const int4 d = desc[q];
    const int c = d.y;
    const unsigned long long* sg = D.seg + (uint32_t)d.x;
    for (int l = threadIdx.x; l < kVoxelsPerBlock; l += blockDim.x) cnt[l] = 0;
    __syncthreads();
    for (int x = threadIdx.x; x < c; x += blockDim.x) {
      const unsigned long long r = sg[x];
      rin[x] = r;
      atomicAdd(&cnt[rec_voxel(r)], 1u);
    }
    __syncthreads();
    {
      const int l0 = 2 * threadIdx.x;
      const uint32_t c0 = cnt[l0], c1 = cnt[l0 + 1];
      uint32_t tot;
      const uint32_t ex = block_scan_excl(c0 + c1, sh, tot);
      start[l0] = ex;
      start[l0 + 1] = ex + c0;
      cnt[l0] = 0;
      cnt[l0 + 1] = 0;
      if (threadIdx.x == 0) start[kVoxelsPerBlock] = tot;
    }
    __syncthreads();
    for (int x = threadIdx.x; x < c; x += blockDim.x) {
      const unsigned long long r = rin[x];
      const int l = rec_voxel(r);
      rout[start[l] + atomicAdd(&cnt[l], 1u)] = r;
    }
    __syncthreads();
    const MapDev& m = D.maps[d.w];
    const float maxW = m.maxW, Tt = m.T, invT = 1.0f / m.T;
    for (int l = threadIdx.x; l < kVoxelsPerBlock; l += blockDim.x) {
      const int b0 = (int)start[l], b1 = (int)start[l + 1];
      if (b1 == b0) continue;
      for (int i = b0 + 1; i < b1; ++i) {
        const unsigned long long key = rout[i];
        int j = i - 1;
        while (j >= b0 && rout[j] > key) {
          rout[j + 1] = rout[j];
          --j;
        }
        rout[j + 1] = key;
      }
      const int64_t vi = (int64_t)d.z * kVoxelsPerBlock + l;
      float S, W;
      State::load(D.heap, vi, invT, Tt, S, W);
      for (int i = b0; i < b1; ++i) {
        float w, wd;
        weight_of(__uint_as_float((uint32_t)rout[i]), w, wd);
        fold_step(S, W, w, wd, maxW);
      }
      State::store(D.heap, vi, invT, Tt, S, W);
      uniq++;
    }
    __syncthreads();
  }
  stat_add(D, ST_UNIQUE, uniq);
}

// Blocks beyond the counting tile: records sorted in LDS tiles written back
// in place, then every thread merges its voxel's runs across the tiles in point
// order.
template <class State>
__global__ void __launch_bounds__(kFoldThreads) k_fold_huge(Dev D) {
  __shared__ unsigned long long sk[kFoldCap];
  const int n_huge = ld_acq(&D.ctr->fold_n[FC_HUGE]);
  const int4* desc = D.fold_desc + (int64_t)FC_HUGE * D.fold_cap;
  unsigned long long uniq = 0, dropped = 0;
  for (int qb = blockIdx.x; qb < n_huge; qb += gridDim.x) {
    const int4 d = desc[qb];
    const int c = d.y;
    const int slot = d.z;
    const int nch = (c + kFoldCap - 1) / kFoldCap;
    __syncthreads();
    if (nch > kMaxChunks) {
      if (threadIdx.x == 0) dropped += c;
      continue;
    }
    unsigned long long* sg = D.seg + (uint32_t)d.x;
    const MapDev& m = D.maps[d.w];
    const float maxW = m.maxW, Tt = m.T, invT = 1.0f / m.T;
    for (int ch = 0; ch < nch; ++ch) {
      const int len = min(kFoldCap, c - ch * kFoldCap);
      const int n2 = next_pow2(len);
      for (int x = threadIdx.x; x < n2; x += blockDim.x) sk[x] = x < len ? sg[ch * kFoldCap + x] : ~0ull;
      __syncthreads();
      bitonic_sort_lds(sk, n2);
      for (int x = threadIdx.x; x < len; x += blockDim.x) sg[ch * kFoldCap + x] = sk[x];
      __threadfence();
      __syncthreads();
    }
    for (int l = threadIdx.x; l < kVoxelsPerBlock; l += blockDim.x) {
      int cur[kMaxChunks], end[kMaxChunks];
      bool any = false;
      for (int ch = 0; ch < nch; ++ch) {
        const int lo0 = ch * kFoldCap, hi0 = min(c, lo0 + kFoldCap);
        int lo = lo0, hi = hi0;
        while (lo < hi) {
          const int mid = (lo + hi) >> 1;
          if (rec_voxel(ld_acq(&sg[mid])) < l) lo = mid + 1;
          else hi = mid;
        }
        int lo2 = lo;
        hi = hi0;
        while (lo2 < hi) {
          const int mid = (lo2 + hi) >> 1;
          if (rec_voxel(ld_acq(&sg[mid])) < l + 1) lo2 = mid + 1;
          else hi = mid;
        }
        cur[ch] = lo;
        end[ch] = lo2;
        any = any || (lo2 > lo);
      }
      if (!any) continue;
      const int64_t vi = (int64_t)slot * kVoxelsPerBlock + l;
      float S, W;
      State::load(D.heap, vi, invT, Tt, S, W);
      for (;;) {
        int best = -1;
        unsigned long long bv = ~0ull;
        for (int ch = 0; ch < nch; ++ch) {
          if (cur[ch] < end[ch]) {
            const unsigned long long v = ld_acq(&sg[cur[ch]]);
            if (best < 0 || v < bv) {
              best = ch;
              bv = v;
            }
          }
        }
        if (best < 0) break;
        float w, wd;
        weight_of(__uint_as_float((uint32_t)bv), w, wd);
        fold_step(S, W, w, wd, maxW);
        cur[best]++;
      }
      State::store(D.heap, vi, invT, Tt, S, W);
      uniq++;
    }
  }
  stat_add(D, ST_UNIQUE, uniq);
  stat_add(D, ST_UPD_DROPPED, dropped);
}

// ---- publish ---------------------------------------------------------------------------
// Selects the map's occupied entries; mode 0: all, 1: within the publish
// radius of the sensor (H:812).
__global__ void __launch_bounds__(256) k_select(Dev D, int map_slot, float sx, float sy, float sz, int mode) {
  const MapDev m = D.maps[map_slot];
  const int64_t e0 = (int64_t)m.bucket_base * kBucket;
  const int64_t ne = (int64_t)(m.n_home + m.n_ovf) * kBucket;
  const float s[3] = {sx, sy, sz};
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < ne; base += (int64_t)gridDim.x * blockDim.x) {
    const int64_t x = base + threadIdx.x;
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
    const int slot = wave_append(&D.ctr->sel_n, pred);
    if (pred) D.sel[slot] = (uint32_t)(e0 + x);
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
  const float Tt = m.T, invT = 1.0f / m.T;
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
      State::load(D.heap, vi, invT, Tt, S, W);
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
  const float Tt = m.T, invT = 1.0f / m.T;
  for (int q = blockIdx.x; q < n; q += gridDim.x) {
    const uint32_t e = D.sel[q];
    const int slot = D.vals[e];
    if (threadIdx.x == 0) unpack_key(D.keys[e], keys_out + 3 * (int64_t)q);
    for (int l = threadIdx.x; l < kVoxelsPerBlock; l += blockDim.x) {
      float S, W;
      State::load(D.heap, (int64_t)slot * kVoxelsPerBlock + l, invT, Tt, S, W);
      vox_out[(int64_t)q * kVoxelsPerBlock + l] = make_float2(S, W);
    }
  }
}

// ---- garbage collection -----------------------------------------------------------------
// mode 0: remove blocks beyond the gc radius (H:836-868); mode 1: remove all
// (map teardown).
__global__ void __launch_bounds__(256) k_gc_mark(Dev D, int map_slot, float sx, float sy, float sz, int mode) {
  const MapDev m = D.maps[map_slot];
  const int64_t e0 = (int64_t)m.bucket_base * kBucket;
  const int64_t ne = (int64_t)(m.n_home + m.n_ovf) * kBucket;
  const float s[3] = {sx, sy, sz};
  unsigned long long removed = 0;
  for (int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; x < ne; x += (int64_t)gridDim.x * blockDim.x) {
    const unsigned long long k = D.keys[e0 + x];
    if (k == kEmptyKey) continue;
    bool far = true;
    if (mode == 0) {
      int32_t kb[3];
      unpack_key(k, kb);
      float c[3];
      block_center(kb, m.BL, c);
      far = !(dist2(c, s) <= m.gcD2);
    }
    if (!far) continue;
    const int slot = D.vals[e0 + x];
    D.keys[e0 + x] = kEmptyKey;
    D.vals[e0 + x] = -1;
    if (slot >= 0) {
      const int idx = atomicAdd(&D.ctr->heap_top, 1);
      D.heap_free[idx] = slot;
      const int f = atomicAdd(&D.ctr->freed_n, 1);
      D.freed[f] = slot;
      atomicSub(&D.mctr[map_slot].blocks, 1);
    }
    removed++;
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
  }
}

__global__ void k_gc_orphans(Dev D, int map_slot) {
  const MapDev m = D.maps[map_slot];
  MapCtr* mc = &D.mctr[map_slot];
  const int n = mc->orphan_n;
  for (int o = threadIdx.x; o < n; o += blockDim.x) {
    const int b = D.orphans[m.bucket_base + m.n_home + o];
    const int idx = atomicAdd(&mc->ovf_top, 1);
    D.ovf_stack[m.bucket_base + m.n_home + idx] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) mc->orphan_n = 0;
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
