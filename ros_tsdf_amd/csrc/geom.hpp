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

// traverseVolume (H:146-245) for the cell keys: steps a point from a towards
// b in increments of +-L, one or more axes at a time by the reference's tMax
// comparison tree, and follows the integer cell k = floor(x * s), s = 1/L.
// The reference compares cell centres with tolerance L/4: "centres equal"
// (loop end, |c - c_b| < L/4 on every axis) and its precision guard (a step
// that moved no axis by a full cell, or some axis by more than one) are
// tests on k alone, because centres of adjacent cells are L apart up to FP32
// rounding that the point validity rule keeps far below L/4.  Only the setup
// (tMax from the centre of a's cell, tDelta) needs FP32 divisions.
// At most max_cells cells are emitted; trunc reports the cut.
// emit(k0, k1, k2) receives each cell and returns false to stop the walk
// early (the caller has seen every cell it needs).
template <class Emit>
__device__ __forceinline__ int walk_cells(const float a[3], const float b[3], float L, float scale, float iscale,
                                          float half, int max_cells, bool& trunc, Emit&& emit) {
  trunc = false;
  const float v[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
  float step[3], tMax[3], tDelta[3], cur[3];
  int32_t k[3], ke[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    step[i] = v[i] > 0 ? L : -1.0f * L;
    const float f = floorf(a[i] * scale);
    const float c0 = div_by_scale(f, scale, iscale) + half;
    tMax[i] = fabsf(v[i] < 0 ? (c0 - half - a[i]) / v[i] : (c0 + half - a[i]) / v[i]);
    tDelta[i] = fabsf(L / v[i]);
    cur[i] = a[i];
    k[i] = (int32_t)f;
    ke[i] = (int32_t)floorf(b[i] * scale);
  }
  int n = 0;
  while (!(k[0] == ke[0] && k[1] == ke[1] && k[2] == ke[2])) {
    if (n == max_cells) {
      trunc = true;
      return n;
    }
    const bool go = emit(k[0], k[1], k[2]);
    ++n;
    if (!go) return n;
    // The reference's comparison tree (H:175-227), written without branches:
    //   t0 < t1:  x if t0 < t2, z if t0 > t2, else x and z
    //   t0 > t1:  y if t1 < t2, z if t1 > t2, else y and z
    //   else:     z if t2 < t0, x and y if t2 > t0, else all three
    // ("else" also covers comparisons with NaN, which are all false).
    const bool A = tMax[0] < tMax[1], B = tMax[0] > tMax[1], E = !A && !B;
    const bool sx = (A && !(tMax[0] > tMax[2])) || (E && !(tMax[2] < tMax[0]));
    const bool sy = (B && !(tMax[1] > tMax[2])) || (E && !(tMax[2] < tMax[0]));
    const bool sz = (A && !(tMax[0] < tMax[2])) || (B && !(tMax[1] < tMax[2])) || (E && !(tMax[2] > tMax[0]));
    cur[0] = sx ? cur[0] + step[0] : cur[0];
    tMax[0] = sx ? tMax[0] + tDelta[0] : tMax[0];
    cur[1] = sy ? cur[1] + step[1] : cur[1];
    tMax[1] = sy ? tMax[1] + tDelta[1] : tMax[1];
    cur[2] = sz ? cur[2] + step[2] : cur[2];
    tMax[2] = sz ? tMax[2] + tDelta[2] : tMax[2];
    int32_t kn[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) kn[i] = (int32_t)floorf(cur[i] * scale);
    const int32_t d0 = abs(kn[0] - k[0]), d1 = abs(kn[1] - k[1]), d2 = abs(kn[2] - k[2]);
    if ((d0 == 0 && d1 == 0 && d2 == 0) || d0 > 1 || d1 > 1 || d2 > 1) return n;
    k[0] = kn[0];
    k[1] = kn[1];
    k[2] = kn[2];
  }
  if (n == max_cells) {
    trunc = true;
    return n;
  }
  emit(k[0], k[1], k[2]);
  return n + 1;
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
