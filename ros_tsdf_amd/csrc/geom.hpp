// Device-side FP32 geometry of the integration path.
//
// Every expression keeps the evaluation order of the reference
// (tsdf_package/src/cuda/tsdf_handler.cu, "H:") so that cell keys, voxel
// centres and the discrete decisions (DDA steps, sign of the SDF, publish and
// GC radius tests) are bit-identical to an IEEE FP32 evaluation on any
// machine.  This translation unit is compiled with -ffp-contract=off, so no
// product is fused into an add, and FP32 '/' and sqrtf are correctly rounded.
// pow(x, 2) of the reference is x * x.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tsdfhip {

constexpr int kVoxelsPerSide = 8;    // VOXELS_PER_SIDE, tsdf_container.cuh:19
constexpr int kVoxelsPerBlock = 512; // VOXELS_PER_BLOCK, tsdf_container.cuh:21

// Integer cell coordinates are packed 21 bits per axis with a 2^20 bias.
// The point validity rule (|coord| <= 2^16 voxels) keeps every key in range.
__host__ __device__ inline uint64_t pack_key(int32_t x, int32_t y, int32_t z) {
  const uint64_t b = 1u << 20;
  return ((uint64_t)(uint32_t)(x + (int32_t)b) << 42) | ((uint64_t)(uint32_t)(y + (int32_t)b) << 21) |
         (uint64_t)(uint32_t)(z + (int32_t)b);
}
__host__ __device__ inline void unpack_key(uint64_t k, int32_t* o) {
  const int32_t b = 1 << 20;
  o[0] = (int32_t)((k >> 42) & 0x1FFFFF) - b;
  o[1] = (int32_t)((k >> 21) & 0x1FFFFF) - b;
  o[2] = (int32_t)(k & 0x1FFFFF) - b;
}

// floorFun + getVolumeCenterFromPoint (H:65-85) for one axis: returns the
// cell centre floor(x * s) / s + L/2 with s = 1/L, and the integer floor.
__device__ __forceinline__ float cell_center(float x, float L, int32_t& k) {
  const float scale = 1.0f / L;
  const float half = L / 2.0f;
  const float f = floorf(x * scale);
  k = (int32_t)f;
  return f / scale + half;
}

// x / sc for a cell scale sc = fl(1/L): correctly rounded, like the
// reference's division.  With isc = fl(1/sc) the quotient is formed by a
// product and one FMA correction (Markstein); the host enables that form for
// a map only after checking it against true division for every integer cell
// index the validity rule allows (isc = 0 selects the division itself).
__device__ __forceinline__ float div_by_scale(float x, float sc, float isc) {
  if (isc == 0.0f) return x / sc;
  const float q0 = x * isc;
  const float r = __builtin_fmaf(-q0, sc, x);
  return __builtin_fmaf(r, isc, q0);
}

// Point validity (the engine's rule for inputs the reference cannot handle:
// H:168 never terminates on NaN, H:119 divides by a zero range).
__device__ __forceinline__ bool point_valid(const float p[3], const float s[3], float vs) {
  const float lim = vs * 65536.0f;
  bool ok = true;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    ok = ok && isfinite(p[i]) && isfinite(s[i]) && fabsf(p[i]) <= lim && fabsf(s[i]) <= lim;
  }
  if (!ok) return false;
  const float v0 = p[0] - s[0], v1 = p[1] - s[1], v2 = p[2] - s[2];
  const float m = sqrtf(v0 * v0 + v1 * v1 + v2 * v2);
  return m > 0.0f;
}

// getTruncationLineEndPoints (H:112-133): the segment p -+ T * (p - s)/|p - s|,
// ordered so that a is the end nearer the sensor.
__device__ __forceinline__ void truncation_segment(const float p[3], const float s[3], float T, float a[3], float b[3]) {
  const float v0 = p[0] - s[0], v1 = p[1] - s[1], v2 = p[2] - s[2];
  const float m = sqrtf(v0 * v0 + v1 * v1 + v2 * v2);
  const float n[3] = {v0 / m, v1 / m, v2 / m};
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const float t = T * n[i];
    a[i] = p[i] - t;
    b[i] = p[i] + t;
  }
  const float a0 = a[0] - s[0], a1 = a[1] - s[1], a2 = a[2] - s[2];
  const float b0 = b[0] - s[0], b1 = b[1] - s[1], b2 = b[2] - s[2];
  const float da = a0 * a0 + a1 * a1 + a2 * a2;
  const float db = b0 * b0 + b1 * b1 + b2 * b2;
  if (db < da) {
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const float t = a[i];
      a[i] = b[i];
      b[i] = t;
    }
  }
}

// floor(x) as an integer, in one instruction (v_cvt_flr_i32_f32): the same
// value as (int)floorf(x) for every x the point validity rule allows
// (|x| < 2^31; NaN gives 0 both ways).
__device__ __forceinline__ int32_t floor_int(float x) {
  int32_t r;
  asm("v_cvt_flr_i32_f32 %0, %1" : "=v"(r) : "v"(x));
  return r;
}

// traverseVolume (H:146-245) for the cell keys: steps a point from a towards
// b in increments of +-L, one or more axes at a time by the reference's tMax
// comparison tree, and follows the integer cell k = floor(x * s), s = 1/L.
// The reference compares cell centres with tolerance L/4: "centres equal"
// (loop end, |c - c_b| < L/4 on every axis) and its precision guard (a step
// that moved no axis by a full cell, or some axis by more than one) are
// tests on k alone, because centres of adjacent cells are L apart up to FP32
// rounding that the point validity rule keeps far below L/4.  Only the setup
// (tMax from the centre of a's cell, tDelta) needs FP32 divisions.
struct CellWalk {
  float step[3], tMax[3], tDelta[3], cur[3];
  int32_t k[3], ke[3];
  float scale;

  __device__ __forceinline__ CellWalk(const float a[3], const float b[3], float L, float sc, float isc, float half)
      : scale(sc) {
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const float v = b[i] - a[i];
      step[i] = v > 0 ? L : -1.0f * L;
      const float f = floorf(a[i] * sc);
      const float c0 = div_by_scale(f, sc, isc) + half;
      tMax[i] = fabsf(v < 0 ? (c0 - half - a[i]) / v : (c0 + half - a[i]) / v);
      tDelta[i] = fabsf(L / v);
      cur[i] = a[i];
      k[i] = (int32_t)f;
      ke[i] = floor_int(b[i] * sc);
    }
  }

  __device__ __forceinline__ bool at_end() const { return k[0] == ke[0] && k[1] == ke[1] && k[2] == ke[2]; }

  // Moves to the next cell by the reference's comparison tree (H:175-227):
  //   t0 < t1:  x if t0 < t2, z if t0 > t2, else x and z
  //   t0 > t1:  y if t1 < t2, z if t1 > t2, else y and z
  //   else:     z if t2 < t0, x and y if t2 > t0, else all three
  // ("else" also covers comparisons with NaN, which are all false).  Written
  // with its six distinct comparisons: x steps unless t0 > t1 or t0 > t2; z
  // steps unless (t0 > t1 ? t1 : t0) < t2; y steps when t0 > t1 and not
  // t1 > t2, or when t0, t1 are tied (neither < nor >) and not t0 > t2.
  // Returns false, with the walk unchanged, when the precision guard ends
  // it; otherwise xing[i] tells whether axis i entered a new block of 8 cells
  // (with |kn - k| <= 1, the block index k >> 3 changes exactly when k ^ kn
  // has a bit above the low three).
  __device__ __forceinline__ bool advance(bool xing[3]) {
    const bool gt01 = tMax[0] > tMax[1], lt01 = tMax[0] < tMax[1];
    const bool gt02 = tMax[0] > tMax[2];
    const bool gt12 = tMax[1] > tMax[2];
    const bool sx = !gt01 && !gt02;
    const bool sy = (gt01 && !gt12) || (!lt01 && !gt01 && !gt02);
    const bool sz = !((gt01 ? tMax[1] : tMax[0]) < tMax[2]);
    const bool st[3] = {sx, sy, sz};
    int32_t kn[3], d[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      cur[i] = st[i] ? cur[i] + step[i] : cur[i];
      tMax[i] = st[i] ? tMax[i] + tDelta[i] : tMax[i];
      kn[i] = floor_int(cur[i] * scale);
      d[i] = kn[i] - k[i];
    }
    // the guard: no axis moved a cell, or one moved more than one
    const uint32_t far = max(max((uint32_t)(d[0] + 1), (uint32_t)(d[1] + 1)), (uint32_t)(d[2] + 1));
    if ((d[0] | d[1] | d[2]) == 0 || far > 2u) return false;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      xing[i] = (uint32_t)(kn[i] ^ k[i]) > 7u;
      k[i] = kn[i];
    }
    return true;
  }

  // advance() with no early return, for loops that end at the first false:
  // the step is applied whatever the guard says (a walk the guard ends is not
  // used again), so the loop body has a single exit and no nested branches.
  __device__ __forceinline__ bool advance_always(bool xing[3]) {
    const bool gt01 = tMax[0] > tMax[1], lt01 = tMax[0] < tMax[1];
    const bool gt02 = tMax[0] > tMax[2];
    const bool gt12 = tMax[1] > tMax[2];
    const bool sx = !gt01 && !gt02;
    const bool sy = (gt01 && !gt12) || (!lt01 && !gt01 && !gt02);
    const bool sz = !((gt01 ? tMax[1] : tMax[0]) < tMax[2]);
    const bool st[3] = {sx, sy, sz};
    int32_t kn[3], d[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      cur[i] = st[i] ? cur[i] + step[i] : cur[i];
      tMax[i] = st[i] ? tMax[i] + tDelta[i] : tMax[i];
      kn[i] = floor_int(cur[i] * scale);
      d[i] = kn[i] - k[i];
    }
    const uint32_t far = max(max((uint32_t)(d[0] + 1), (uint32_t)(d[1] + 1)), (uint32_t)(d[2] + 1));
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      xing[i] = (uint32_t)(kn[i] ^ k[i]) > 7u;
      k[i] = kn[i];
    }
    return (d[0] | d[1] | d[2]) != 0 && far <= 2u;
  }
};

// The cells of traverseVolume, in order: at most max_cells are emitted and
// trunc reports the cut.  emit(k0, k1, k2) receives each cell and returns
// false to stop the walk early (the caller has seen every cell it needs).
template <class Emit>
__device__ __forceinline__ int walk_cells(const float a[3], const float b[3], float L, float scale, float iscale,
                                          float half, int max_cells, bool& trunc, Emit&& emit) {
  trunc = false;
  CellWalk w(a, b, L, scale, iscale, half);
  int n = 0;
  // each turn emits the current cell (the end cell included) and steps on;
  // the step is computed whether or not the walk goes on
  for (;;) {
    if (n == max_cells) {
      trunc = true;
      break;
    }
    const bool last = w.at_end();
    const bool go = emit(w.k[0], w.k[1], w.k[2]);
    ++n;
    bool x[3];
    const bool moved = w.advance_always(x);
    if (last || !go || !moved) break;
  }
  return n;
}

// The same walk (voxel cells, blocks of 8^3) summarised as runs: every axis
// moves monotonically, so the cells of one block are consecutive and the
// walk is a sequence of runs, one per block it passes through.
// on_run(kb0, kb1, kb2, len) is called for each run in walk order.  Per cell
// the walk only notes, in three bit masks over a window of 32 cells, which
// axes entered a new block; the runs are read off the masks when the window
// fills and at the end.  Returns the number of cells (as walk_cells).
template <class OnRun>
__device__ __forceinline__ int walk_runs(const float a[3], const float b[3], float L, float scale, float iscale,
                                         float half, int max_cells, bool& trunc, OnRun&& on_run) {
  trunc = false;
  CellWalk w(a, b, L, scale, iscale, half);
  int32_t blk[3] = {w.k[0] >> 3, w.k[1] >> 3, w.k[2] >> 3};  // block of the open run
  const int32_t dir[3] = {w.step[0] > 0 ? 1 : -1, w.step[1] > 0 ? 1 : -1, w.step[2] > 0 ? 1 : -1};
  int open = 0;  // cells of the open run before the window
  int n = 0;     // cells so far
  // the runs ending among window cells [base, base + len): cm[i] bit j tells
  // that window cell j entered a new block along axis i
  auto drain = [&](const uint32_t cm[3], int len) {
    uint32_t m = cm[0] | cm[1] | cm[2];
    if (len < 32) m &= (1u << len) - 1u;
    int prev = 0;
    while (m) {
      const int p = __builtin_ctz(m);
      if (open + p - prev > 0) on_run(blk[0], blk[1], blk[2], open + p - prev);
#pragma unroll
      for (int i = 0; i < 3; ++i) blk[i] += (cm[i] >> p) & 1u ? dir[i] : 0;
      open = 0;
      prev = p;
      m &= m - 1u;
    }
    open += len - prev;
  };
  if (max_cells <= 32) {
    // every cell fits one window: a loop with no exit but the walk's own end
    // Each turn computes the next step whatever happens and has one exit: the
    // cell limit (the end cell too: walk_cells cuts it as well), the end cell,
    // or the precision guard.
    uint32_t cm[3] = {0u, 0u, 0u};
    for (;;) {
      const bool cut = n == max_cells;
      const bool last = w.at_end();
      bool x[3];
      const bool moved = w.advance_always(x);
      trunc = trunc || cut;
      n += cut ? 0 : 1;
      const bool go = !cut && !last && moved;
      // now at cell n; a cell past max_cells (n = 32) is never counted
      const uint32_t bit = go && n < 32 ? 1u << n : 0u;
#pragma unroll
      for (int i = 0; i < 3; ++i) cm[i] |= x[i] ? bit : 0u;
      if (!go) break;
    }
    drain(cm, n);
    if (open > 0) on_run(blk[0], blk[1], blk[2], open);
    return n;
  }
  bool more = true;
  bool x[3] = {false, false, false};  // block entries of the cell that opens the window
  for (int base = 0; more; base = n) {
    // bit j of cm[i]: window cell j entered a new block along axis i
    uint32_t cm[3] = {x[0] ? 1u : 0u, x[1] ? 1u : 0u, x[2] ? 1u : 0u};
    // each turn: the current cell n is counted, then the walk moves on
    for (;;) {
      if (n == max_cells) {  // (the end cell too: walk_cells cuts it as well)
        trunc = true;
        more = false;
        break;
      }
      const bool last = w.at_end();
      ++n;
      if (last || !w.advance(x)) {
        more = false;
        break;
      }
      if (n - base == 32) break;  // cell n opens the next window
      const uint32_t bit = 1u << (n - base);
#pragma unroll
      for (int i = 0; i < 3; ++i) cm[i] = x[i] ? cm[i] | bit : cm[i];
    }
    drain(cm, n - base);
  }
  if (open > 0) on_run(blk[0], blk[1], blk[2], open);
  return n;
}

// Centre of cell k: getVolumeCenterFromPoint (H:76-85) with floor(x * s)
// replaced by the k it produced.
__device__ __forceinline__ float cell_centre_of(int32_t k, float scale, float iscale, float half) {
  return div_by_scale((float)k, scale, iscale) + half;
}
// The same for a map whose product form is enabled (iscale != 0).
__device__ __forceinline__ float product_centre(int32_t k, float scale, float iscale, float half) {
  const float x = (float)k;
  const float q0 = x * iscale;
  const float r = __builtin_fmaf(-q0, scale, x);
  return __builtin_fmaf(r, iscale, q0) + half;
}

// Block and local index of voxel k.  The reference derives them from the
// voxel centre in FP32 (H:612, H:622-626, H:507-512): floor(c / (8 L)) and
// floor((c - block corner) / L).  With centres (k + 1/2) L up to rounding,
// those floors are floor(k / 8) and k mod 8 exactly (the fractional parts are
// at least 1/16 away from an integer, far above FP32 error at valid points).
__device__ __forceinline__ int32_t block_of(int32_t k) { return k >> 3; }
__device__ __forceinline__ int local_index(const int32_t k[3]) {
  return (k[0] & 7) + 8 * (k[1] & 7) + 64 * (k[2] & 7);
}

// getDistanceUpdate (H:562-573): |p - x|, negated when (p - s).(p - x) < 0.
// The sign is a discrete decision and is evaluated exactly as the reference
// writes it; the magnitude is a value (compared within a tolerance), so it
// takes the hardware square root (1 ulp) instead of the correctly rounded
// sequence.
__device__ __forceinline__ float signed_distance(const float p[3], const float s[3], const float x[3]) {
  const float ps0 = p[0] - s[0], ps1 = p[1] - s[1], ps2 = p[2] - s[2];
  const float px0 = p[0] - x[0], px1 = p[1] - x[1], px2 = p[2] - x[2];
  const float mag = __builtin_amdgcn_sqrtf(px0 * px0 + px1 * px1 + px2 * px2);
  const float dot = ps0 * px0 + ps1 * px1 + ps2 * px2;
  return dot < 0 ? -mag : mag;
}

// Block centre from its integer key (the value getVolumeCenterFromPoint
// produced when it derived that key).
__device__ __forceinline__ void block_center(const int32_t k[3], float BL, float c[3]) {
  const float scale = 1.0f / BL;
  const float half = BL / 2.0f;
#pragma unroll
  for (int i = 0; i < 3; ++i) c[i] = (float)k[i] / scale + half;
}

// withinDistanceSquared (H:749-757) with a = block centre, b = sensor.
__device__ __forceinline__ float dist2(const float a[3], const float b[3]) {
  const float d0 = b[0] - a[0], d1 = b[1] - a[1], d2 = b[2] - a[2];
  return d0 * d0 + d1 * d1 + d2 * d2;
}

// getVoxelCoordinatesFromLocalVoxelIndex (H:520-534).
__device__ __forceinline__ void voxel_position(int idx, const float bl[3], float vs, float half_vs, float out[3]) {
  const float z = (float)(idx / (kVoxelsPerSide * kVoxelsPerSide));
  idx = (int)((float)idx - z * (float)kVoxelsPerSide * (float)kVoxelsPerSide);
  const float y = (float)(idx / kVoxelsPerSide);
  idx = (int)((float)idx - y * (float)kVoxelsPerSide);
  const float x = (float)idx;
  out[0] = x * vs + half_vs + bl[0];
  out[1] = y * vs + half_vs + bl[1];
  out[2] = z * vs + half_vs + bl[2];
}

}  // namespace tsdfhip
