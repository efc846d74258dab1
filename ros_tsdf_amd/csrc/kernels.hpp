// Kernels of the integration pipeline.  One pass integrates the scans of a
// batch (a single scan on the per-scan path):
//
//   k_pass_begin   zero the pass counters
//   k_traverse     per point: truncation segment, block DDA and voxel DDA;
//                  find-or-insert every block in the map's spatial hash
//                  (lock-free), count updates per block, write one record
//                  (entry, local voxel, point, sdf) per traversed voxel
//   k_assign       pop a heap block from the free stack for each new entry
//   k_scan_*       exclusive scan of the per-block update counts
//   k_scatter      move records into per-block segments
//   k_fold         one workgroup per touched block: sort its records by
//                  (voxel, point) in LDS, then fold each voxel's updates in
//                  point order into the voxel state (one read + one write)
//   k_fold_big     the same for blocks whose segment exceeds one LDS tile
//
// publish: k_select (radius test per entry) -> k_pub_count (wave per block)
//          -> scan -> k_pub_write (wave ballot compaction)
// gc:      k_gc_mark -> k_gc_zero -> k_gc_chains -> k_gc_orphans
#pragma once
#include "device.hpp"

namespace tsdfhip {

constexpr int kFoldThreads = 256;
constexpr int kFoldCap = 4096;   // records per LDS tile (32 KB)
constexpr int kMaxChunks = 32;   // tiles per block handled by k_fold_big
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

// Adds per-thread counters into the engine statistics: one atomic per wave.
template <int N>
__device__ __forceinline__ void flush_stats(EngCtr* ctr, const int* which, const unsigned long long (&v)[N]) {
#pragma unroll
  for (int i = 0; i < N; ++i) {
    const unsigned long long s = wave_sum(v[i]);
    if (lane_id() == 0 && s) atomicAdd(&ctr->stats[which[i]], s);
  }
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
      for (int s = cs; s < kBucket; ++s) {
        const unsigned long long k = ld_acq(&D.keys[(int64_t)cg * kBucket + s]);
        if (k == key) return cg * kBucket + s;
        if (k == kEmptyKey && fe < 0) {
          fe = cg * kBucket + s;
          fe_depth = cd;
        }
      }
      const int nx = ld_acq(&D.next[cg]);
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
  if (t < ST_COUNT) D.ctr->stats[t] = 0;
  if (t == 0) {
    D.ctr->touched_n = 0;
    D.ctr->new_n = 0;
    D.ctr->big_n = 0;
  }
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

__global__ void __launch_bounds__(256) k_traverse(Dev D) {
  const PassHdr hdr = *D.hdr;
  const int P = hdr.total;
  unsigned long long st[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  const int which[9] = {ST_VALID, ST_SKIPPED, ST_BLOCK_CELLS, ST_UPDATES, ST_TRUNC_VOX,
                        ST_TRUNC_BLK, ST_INDEX_DROPPED, ST_INSERT_FAIL, ST_UPD_DROPPED};
  const int64_t stride = D.pcap;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < P; i += gridDim.x * blockDim.x) {
    const ScanDev sd = D.scans[find_scan(D.scans, hdr.nscans, i)];
    const MapDev m = D.maps[sd.map];
    const float4 q = hdr.points[i];
    const float p[3] = {q.x, q.y, q.z};
    const float s[3] = {sd.s[0], sd.s[1], sd.s[2]};
    int nrec = 0;
    if (!point_valid(p, s, m.vs)) {
      st[1]++;
      D.rec_n[i] = 0;
      continue;
    }
    st[0]++;
    float a[3], b[3];
    truncation_segment(p, s, m.T, a, b);
    // block pass: allocate every block cell (H:255-272, H:421-500)
    bool tb = false;
    const int nb = traverse(a, b, m.BL, m.max_blk, tb, [&](int32_t k0, int32_t k1, int32_t k2, float, float, float) {
      bool ins;
      const int e = find_or_insert(D, m, pack_key(k0, k1, k2), ins);
      if (e < 0) {
        st[7]++;
      } else if (ins) {
        const int q2 = atomicAdd(&D.ctr->new_n, 1);
        D.new_list[q2] = (uint32_t)e;
        D.new_map[q2] = m.slot;
      }
    });
    st[2] += nb;
    st[5] += tb ? 1 : 0;
    // voxel pass: one record per traversed voxel (H:675-710, H:604-665)
    unsigned long long last_key = kEmptyKey;
    int last_e = -1;
    bool tv = false;
    traverse(a, b, m.vs, m.max_vox, tv, [&](int32_t, int32_t, int32_t, float c0, float c1, float c2) {
      const float vc[3] = {c0, c1, c2};
      int32_t kb[3];
      int local;
      if (!voxel_to_block(vc, m.vs, m.BL, m.half_BL, kb, local)) {
        st[6]++;
        return;
      }
      const unsigned long long key = pack_key(kb[0], kb[1], kb[2]);
      int e;
      if (key == last_key) {
        e = last_e;
      } else {
        bool ins;
        e = find_or_insert(D, m, key, ins);
        if (e < 0) {
          st[7]++;
        } else if (ins) {
          const int q2 = atomicAdd(&D.ctr->new_n, 1);
          D.new_list[q2] = (uint32_t)e;
          D.new_map[q2] = m.slot;
        }
        last_key = key;
        last_e = e;
      }
      if (e < 0) {
        st[8]++;
        return;
      }
      const float d = signed_distance(p, s, vc);
      const uint32_t r = atomicAdd(&D.blk_cnt[e], 1u);
      if (r == 0) {
        const int t = atomicAdd(&D.ctr->touched_n, 1);
        D.touched[t] = (uint32_t)e;
        D.touch_idx[e] = (uint32_t)t;
      }
      const unsigned long long kd = ((unsigned long long)local << 55) | ((unsigned long long)i << 32) |
                                    (unsigned long long)__float_as_uint(d);
      D.rec_e[(int64_t)nrec * stride + i] = (uint32_t)e;
      D.rec_kd[(int64_t)nrec * stride + i] = kd;
      ++nrec;
    });
    st[3] += nrec;
    st[4] += tv ? 1 : 0;
    D.rec_n[i] = (uint8_t)nrec;
  }
  flush_stats(D.ctr, which, st);
}

// Heap block for every entry inserted by the pass (free-list pop; pushes only
// happen in gc, never concurrently with pops, so the stack needs no ABA guard).
__global__ void __launch_bounds__(256) k_assign(Dev D) {
  const int n = ld_acq(&D.ctr->new_n);
  unsigned long long st[3] = {0, 0, 0};
  const int which[3] = {ST_NEW_BLOCKS, ST_HEAP_FAIL, ST_QUOTA_FAIL};
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
    const uint32_t e = D.new_list[q];
    const int ms = D.new_map[q];
    const MapDev& m = D.maps[ms];
    bool ok = true;
    if (m.block_quota > 0) {
      const int b = atomicAdd(&D.mctr[ms].blocks, 1);
      if (b >= m.block_quota) {
        atomicSub(&D.mctr[ms].blocks, 1);
        ok = false;
        st[2]++;
      }
    } else {
      atomicAdd(&D.mctr[ms].blocks, 1);
    }
    int slot = -1;
    if (ok) {
      const int idx = atomicSub(&D.ctr->heap_top, 1) - 1;
      if (idx < 0) {
        atomicAdd(&D.ctr->heap_top, 1);
        atomicSub(&D.mctr[ms].blocks, 1);
        ok = false;
      } else {
        slot = D.heap_free[idx];
      }
    }
    if (ok) {
      D.vals[e] = slot;
      st[0]++;
    } else {
      D.vals[e] = -1;
      D.keys[e] = kEmptyKey;
      st[1]++;
    }
  }
  flush_stats(D.ctr, which, st);
}

// ---- exclusive scan over a device-sized list -------------------------------------
// count(i) is read through a functor; out[i] = sum of counts before i,
// out[n] = total.  Three kernels: tile sums, scan of tile sums, tile scan.
struct CountTouched {
  const uint32_t* list;
  const uint32_t* cnt;
  __device__ uint32_t operator()(int i) const { return cnt[list[i]]; }
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

// ---- scatter + fold ------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_scatter(Dev D) {
  const int P = D.hdr->total;
  const int64_t stride = D.pcap;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < P; i += gridDim.x * blockDim.x) {
    const int n = D.rec_n[i];
    for (int j = 0; j < n; ++j) {
      const uint32_t e = D.rec_e[(int64_t)j * stride + i];
      const unsigned long long kd = D.rec_kd[(int64_t)j * stride + i];
      const uint32_t t = D.touch_idx[e];
      const uint32_t r = atomicSub(&D.blk_cnt[e], 1u) - 1u;
      D.seg[D.seg_off[t] + r] = kd;
    }
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

// first index in [lo, hi) whose record has local voxel >= l
__device__ __forceinline__ int lower_local(const unsigned long long* a, int lo, int hi, int l) {
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if ((int)(a[mid] >> 55) < l) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

__device__ __forceinline__ void fold_one(float& S, float& W, float d, float maxW) {
  // H:637-654: w = 1/(|d|+1); S = (W*S + w*d)/(W+w); W = min(W+w, maxW)
  const float w = 1.0f / (fabsf(d) + 1.0f);
  const float wd = w * d;
  const float Wn = W + w;
  S = (W * S + wd) / Wn;
  W = fminf(Wn, maxW);
}

__device__ __forceinline__ int next_pow2(int c) {
  int n = 2;
  while (n < c) n <<= 1;
  return n;
}

template <class State>
__global__ void __launch_bounds__(kFoldThreads) k_fold(Dev D) {
  __shared__ unsigned long long sk[kFoldCap];
  const int nt = ld_acq(&D.ctr->touched_n);
  unsigned long long uniq = 0, dropped = 0;
  for (int t = blockIdx.x; t < nt; t += gridDim.x) {
    const uint32_t e = D.touched[t];
    const uint32_t base = D.seg_off[t];
    const int c = (int)(D.seg_off[t + 1] - base);
    const int slot = D.vals[e];
    if (slot < 0) {
      if (threadIdx.x == 0) dropped += c;
      continue;
    }
    if (c > kFoldCap) {
      if (threadIdx.x == 0) {
        const int b = atomicAdd(&D.ctr->big_n, 1);
        D.big_list[b] = (uint32_t)t;
      }
      continue;
    }
    const int n2 = next_pow2(c);
    for (int x = threadIdx.x; x < n2; x += blockDim.x) sk[x] = x < c ? D.seg[base + x] : ~0ull;
    __syncthreads();
    bitonic_sort_lds(sk, n2);
    // Map constants: the map owning entry e is the map of the scan of any of
    // its records; every record of a block belongs to one map.
    const int pidx = (int)((sk[0] >> 32) & ((1u << kPointIdxBits) - 1));
    const MapDev& m = D.maps[D.scans[find_scan(D.scans, D.hdr->nscans, pidx)].map];
    const float maxW = m.maxW, Tt = m.T, invT = 1.0f / m.T;
    for (int l = threadIdx.x; l < kVoxelsPerBlock; l += blockDim.x) {
      const int lo = lower_local(sk, 0, c, l);
      const int hi = lower_local(sk, lo, c, l + 1);
      if (hi > lo) {
        const int64_t vi = (int64_t)slot * kVoxelsPerBlock + l;
        float S, W;
        State::load(D.heap, vi, invT, Tt, S, W);
        for (int r = lo; r < hi; ++r) fold_one(S, W, __uint_as_float((uint32_t)sk[r]), maxW);
        State::store(D.heap, vi, invT, Tt, S, W);
        uniq++;
      }
    }
    __syncthreads();
  }
  const unsigned long long su = wave_sum(uniq);
  if (lane_id() == 0 && su) atomicAdd(&D.ctr->stats[ST_UNIQUE], su);
  if (threadIdx.x == 0 && dropped) atomicAdd(&D.ctr->stats[ST_UPD_DROPPED], dropped);
}

// Blocks with more records than one LDS tile: sort each tile in LDS and write
// it back, then every thread merges its voxel's runs across the tiles in
// point order.
template <class State>
__global__ void __launch_bounds__(kFoldThreads) k_fold_big(Dev D) {
  __shared__ unsigned long long sk[kFoldCap];
  const int nb = ld_acq(&D.ctr->big_n);
  unsigned long long uniq = 0, dropped = 0;
  for (int q = blockIdx.x; q < nb; q += gridDim.x) {
    const uint32_t t = D.big_list[q];
    const uint32_t e = D.touched[t];
    const uint32_t base = D.seg_off[t];
    const int c = (int)(D.seg_off[t + 1] - base);
    const int slot = D.vals[e];
    const int nch = (c + kFoldCap - 1) / kFoldCap;
    if (nch > kMaxChunks) {
      if (threadIdx.x == 0) dropped += c;
      continue;
    }
    for (int ch = 0; ch < nch; ++ch) {
      const int len = min(kFoldCap, c - ch * kFoldCap);
      const int n2 = next_pow2(len);
      for (int x = threadIdx.x; x < n2; x += blockDim.x) sk[x] = x < len ? D.seg[base + ch * kFoldCap + x] : ~0ull;
      __syncthreads();
      bitonic_sort_lds(sk, n2);
      for (int x = threadIdx.x; x < len; x += blockDim.x) D.seg[base + ch * kFoldCap + x] = sk[x];
      __threadfence();
      __syncthreads();
    }
    const unsigned long long* sg = D.seg + base;
    const int pidx = (int)((ld_acq(&sg[0]) >> 32) & ((1u << kPointIdxBits) - 1));
    const MapDev& m = D.maps[D.scans[find_scan(D.scans, D.hdr->nscans, pidx)].map];
    const float maxW = m.maxW, Tt = m.T, invT = 1.0f / m.T;
    for (int l = threadIdx.x; l < kVoxelsPerBlock; l += blockDim.x) {
      int cur[kMaxChunks], end[kMaxChunks];
      bool any = false;
      for (int ch = 0; ch < nch; ++ch) {
        const int lo0 = ch * kFoldCap, hi0 = min(c, lo0 + kFoldCap);
        int lo = lo0, hi = hi0;
        while (lo < hi) {
          const int mid = (lo + hi) >> 1;
          if ((int)(ld_acq(&sg[mid]) >> 55) < l) lo = mid + 1;
          else hi = mid;
        }
        int lo2 = lo;
        hi = hi0;
        while (lo2 < hi) {
          const int mid = (lo2 + hi) >> 1;
          if ((int)(ld_acq(&sg[mid]) >> 55) < l + 1) lo2 = mid + 1;
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
            if (best < 0 || v < bv) { best = ch; bv = v; }
          }
        }
        if (best < 0) break;
        fold_one(S, W, __uint_as_float((uint32_t)bv), maxW);
        cur[best]++;
      }
      State::store(D.heap, vi, invT, Tt, S, W);
      uniq++;
    }
    __syncthreads();
  }
  const unsigned long long su = wave_sum(uniq);
  if (lane_id() == 0 && su) atomicAdd(&D.ctr->stats[ST_UNIQUE], su);
  if (threadIdx.x == 0 && dropped) atomicAdd(&D.ctr->stats[ST_UPD_DROPPED], dropped);
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
  const unsigned long long s2 = wave_sum(removed);
  if (lane_id() == 0 && s2) atomicAdd(&D.ctr->stats[ST_GC_REMOVED], s2);
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
