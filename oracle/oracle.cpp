// ORACLE — test infrastructure only.
//
// A single-threaded CPU restatement of ros_tsdf's per-scan lidar -> TSDF
// integration (tsdf_package/src/cuda/tsdf_handler.cu, "H:" below), plus a
// multithreaded variant that produces bit-identical results.  Only the test
// suite, the smoke check and the benchmark's CPU-baseline leg load this
// library; the HIP engine never links or calls it.
//
// Arithmetic contract (the engine restates the same contract independently):
//   * geometry (segment end points, DDA, cell centres, local voxel index,
//     the sign test and the publish/GC distance tests) is IEEE FP32 evaluated
//     left to right exactly as the reference writes it, with no FMA
//     contraction (build with -ffp-contract=off) and pow(x,2) read as x*x;
//   * the SDF magnitude, the weight increment and the weighted-mean fold are
//     FP64 here (the engine computes them in FP32 and is compared within a
//     tolerance);
//   * each voxel is folded in ascending point index, scans in call order.
//
// Deliberate deviations from the reference (each also applied by the engine):
//   D1 exact block de-duplication (H:717-741 keeps a stale j+1'th entry);
//   D2 heap slots come from a free list (H:335, H:375 ignore free_blocks);
//   D3 no 1000-entry cap on GC of chained entries (H:853-858);
//   D4 blocks of every updated voxel are allocated (H:620/629 index
//      blocks[-1] when the block DDA missed a voxel's block);
//   D5 deterministic per-voxel fold order (H:643-664 use lock order): here
//      ascending point index; the engine sums each scan's updates of a voxel
//      exactly, which is this order's value in exact arithmetic until the
//      voxel reaches max_weight, and folds in this order from there;
//   D6 each DDA emits at most MAX_*_TRAVERSED cells (H:1311, H:1315) and
//      invalid points (non-finite, zero range, |coord| > 2^16 voxels) are
//      skipped (H:168 never terminates on NaN);
//   D7 the publish counter starts at 0 and has no cap (H:1202, HH:16);
//   D8 pow(x,2) is x*x (H:118, H:125-126, H:542, H:752);
//   D9 the bucket hash (H:53-57) is not reproduced: it is not observable;
//   D10 a voxel whose local index falls outside 0..7 on an axis is dropped
//       and counted (H:509 asserts instead); it does not occur at sane
//       coordinates.
//
// Parity pins for every function below are listed in tests/test_oracle_*.py
// (closed forms, exhaustive FP32 sweeps, exact-rational brute force).

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <unordered_map>
#include <algorithm>
#include <thread>
#include <bitset>

namespace {

constexpr int VPS = 8;          // VOXELS_PER_SIDE, tsdf_container.cuh:19
constexpr int VPB = 512;        // VOXELS_PER_BLOCK, tsdf_container.cuh:21

struct Params {
  float vs, T, maxW, pubD2, gcD2;
  // derived exactly as initGlobalVars does (H:1289-1317)
  float half_vs, BL, half_BL, blk_eps, vox_eps;
  int max_vox, max_blk;
};

static Params derive(const float* in) {
  Params p;
  p.vs = in[0]; p.T = in[1]; p.maxW = in[2]; p.pubD2 = in[3]; p.gcD2 = in[4];
  p.half_vs = p.vs / 2;                                   // H:1296
  p.BL = p.vs * VPS;                                      // H:1299
  p.half_BL = p.BL / 2;                                   // H:1302
  p.blk_eps = p.BL / 4;                                   // H:1305
  p.vox_eps = p.vs / 4;                                   // H:1308
  p.max_vox = (int)(std::ceil(p.T * 2 / p.vs) * 4);       // H:1311
  p.max_blk = (int)(std::ceil(p.T * 2 / p.BL) * 4);       // H:1315
  return p;
}

struct Key { int32_t k[3]; };
static inline uint64_t pack(const int32_t k[3]) {
  const uint64_t b = 1u << 20;
  return ((uint64_t)(k[0] + b) << 42) | ((uint64_t)(k[1] + b) << 21) | (uint64_t)(k[2] + b);
}
static inline void unpack(uint64_t key, int32_t k[3]) {
  const int64_t b = 1 << 20;
  k[0] = (int32_t)((int64_t)((key >> 42) & 0x1FFFFF) - b);
  k[1] = (int32_t)((int64_t)((key >> 21) & 0x1FFFFF) - b);
  k[2] = (int32_t)((int64_t)(key & 0x1FFFFF) - b);
}

struct Block {
  double sdf[VPB];
  double w[VPB];
  Block() { std::memset(sdf, 0, sizeof(sdf)); std::memset(w, 0, sizeof(w)); }
};

// floorFun + getVolumeCenterFromPoint (H:65-85): c = floor(x*s)/s + h, s = 1/L.
static inline float cell_center(float x, float L, int32_t* k) {
  float scale = 1 / L;
  float half = L / 2;
  float f = std::floor(x * scale);
  if (k) *k = (int32_t)f;
  return f / scale + half;
}

// getTruncationLineEndPoints (H:112-133).
static void segment(const float p[3], const float s[3], float T, float a[3], float b[3]) {
  float v[3] = {p[0] - s[0], p[1] - s[1], p[2] - s[2]};
  float m = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  float n[3] = {v[0] / m, v[1] / m, v[2] / m};
  for (int i = 0; i < 3; ++i) {
    float t = T * n[i];
    a[i] = p[i] - t;
    b[i] = p[i] + t;
  }
  float da0 = a[0] - s[0], da1 = a[1] - s[1], da2 = a[2] - s[2];
  float db0 = b[0] - s[0], db1 = b[1] - s[1], db2 = b[2] - s[2];
  float da = da0 * da0 + da1 * da1 + da2 * da2;
  float db = db0 * db0 + db1 * db1 + db2 * db2;
  if (db < da) {
    for (int i = 0; i < 3; ++i) { float t = a[i]; a[i] = b[i]; b[i] = t; }
  }
}

// D6 validity rule; the engine applies the same predicate.
static bool point_valid(const float p[3], const float s[3], float vs) {
  const float lim = vs * 65536.0f;
  for (int i = 0; i < 3; ++i) {
    if (!std::isfinite(p[i]) || !std::isfinite(s[i])) return false;
    if (std::fabs(p[i]) > lim || std::fabs(s[i]) > lim) return false;
  }
  float v[3] = {p[0] - s[0], p[1] - s[1], p[2] - s[2]};
  float m = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  return m > 0.0f;
}

struct Cell { int32_t k[3]; float c[3]; };

// traverseVolume (H:146-245): Amanatides-Woo stepping of a point by +-L,
// re-deriving the cell centre after every step.  Emits at most max_cells
// cells (D6).  Returns the number of cells; *trunc = 1 when the bound cut it.
static int traverse(const float a[3], const float b[3], float L, int max_cells, Cell* out, int* trunc) {
  *trunc = 0;
  const float u[3] = {a[0], a[1], a[2]};
  const float v[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
  float half = L / 2;
  const float eps = L / 4;
  const float Lpe = L + eps;
  const float Lme = L - eps;
  float step[3];
  for (int i = 0; i < 3; ++i) step[i] = v[i] > 0 ? L : -1 * L;
  float c0[3];
  for (int i = 0; i < 3; ++i) c0[i] = cell_center(a[i], L, nullptr);
  float tMax[3], tDelta[3];
  for (int i = 0; i < 3; ++i) {
    tMax[i] = std::fabs(v[i] < 0 ? (c0[i] - half - u[i]) / v[i] : (c0[i] + half - u[i]) / v[i]);
    tDelta[i] = std::fabs(L / v[i]);
  }
  float cur[3] = {a[0], a[1], a[2]};
  Cell cc, ce;
  for (int i = 0; i < 3; ++i) {
    cc.c[i] = cell_center(cur[i], L, &cc.k[i]);
    ce.c[i] = cell_center(b[i], L, &ce.k[i]);
  }
  int n = 0;
  while (!(std::fabs(cc.c[0] - ce.c[0]) < eps && std::fabs(cc.c[1] - ce.c[1]) < eps &&
           std::fabs(cc.c[2] - ce.c[2]) < eps)) {
    if (n == max_cells) { *trunc = 1; return n; }
    out[n++] = cc;
    bool sx = false, sy = false, sz = false;
    if (tMax[0] < tMax[1]) {
      if (tMax[0] < tMax[2]) sx = true;
      else if (tMax[0] > tMax[2]) sz = true;
      else { sx = true; sz = true; }
    } else if (tMax[0] > tMax[1]) {
      if (tMax[1] < tMax[2]) sy = true;
      else if (tMax[1] > tMax[2]) sz = true;
      else { sy = true; sz = true; }
    } else {
      if (tMax[2] < tMax[0]) sz = true;
      else if (tMax[2] > tMax[0]) { sx = true; sy = true; }
      else { sx = true; sy = true; sz = true; }
    }
    if (sx) { cur[0] += step[0]; tMax[0] += tDelta[0]; }
    if (sy) { cur[1] += step[1]; tMax[1] += tDelta[1]; }
    if (sz) { cur[2] += step[2]; tMax[2] += tDelta[2]; }
    Cell prev = cc;
    for (int i = 0; i < 3; ++i) cc.c[i] = cell_center(cur[i], L, &cc.k[i]);
    float d0 = std::fabs(prev.c[0] - cc.c[0]);
    float d1 = std::fabs(prev.c[1] - cc.c[1]);
    float d2 = std::fabs(prev.c[2] - cc.c[2]);
    if ((d0 < Lme && d1 < Lme && d2 < Lme) || (d0 > Lpe || d1 > Lpe || d2 > Lpe)) return n;
  }
  if (n == max_cells) { *trunc = 1; return n; }
  out[n++] = cc;
  return n;
}

// Voxel centre -> (block key, local index): H:612 and H:622-626, H:507-512.
// Returns false (D10) if a floor falls outside 0..7.
static bool voxel_to_block(const Params& P, const float vc[3], int32_t kb[3], int* local) {
  float f[3];
  for (int i = 0; i < 3; ++i) {
    float bc = cell_center(vc[i], P.BL, &kb[i]);
    float bl = bc - P.half_BL;
    float d = vc[i] - bl;
    d = d / P.vs;
    f[i] = std::floor(d);
    if (!(f[i] >= 0.0f && f[i] <= 7.0f)) return false;
  }
  float idx = f[0] + (f[1] * VPS) + (f[2] * VPS * VPS);
  *local = (int)idx;
  return true;
}

// withinDistanceSquared (H:749-757): diff = b - a, dist2 = d0^2 + d1^2 + d2^2.
static inline float dist2(const float a[3], const float b[3]) {
  float d0 = b[0] - a[0], d1 = b[1] - a[1], d2 = b[2] - a[2];
  return d0 * d0 + d1 * d1 + d2 * d2;
}

// Block centre from its integer key: the same float expression as
// getVolumeCenterFromPoint with floor(x*s) replaced by the key it produced.
static inline void block_center(const Params& P, const int32_t k[3], float c[3]) {
  float scale = 1 / P.BL;
  float half = P.BL / 2;
  for (int i = 0; i < 3; ++i) c[i] = (float)k[i] / scale + half;
}

// getVoxelCoordinatesFromLocalVoxelIndex (H:520-534).
static inline void voxel_pos(const Params& P, int idx, const float bl[3], float out[3]) {
  float z = idx / (VPS * VPS);
  idx -= z * VPS * VPS;
  float y = idx / VPS;
  idx -= y * VPS;
  float x = idx;
  out[0] = x * P.vs + P.half_vs + bl[0];
  out[1] = y * P.vs + P.half_vs + bl[1];
  out[2] = z * P.vs + P.half_vs + bl[2];
}

// One voxel update record of one point (H:604-665 with FP64 value path).
struct Rec {
  uint64_t bkey;
  int32_t local;
  double d;
};

// getDistanceUpdate (H:562-573): sign from the FP32 dot product, magnitude FP64.
static inline double signed_distance(const float p[3], const float s[3], const float x[3]) {
  float ps0 = p[0] - s[0], ps1 = p[1] - s[1], ps2 = p[2] - s[2];
  float px0 = p[0] - x[0], px1 = p[1] - x[1], px2 = p[2] - x[2];
  float dot = ps0 * px0 + ps1 * px1 + ps2 * px2;
  double e0 = (double)p[0] - (double)x[0];
  double e1 = (double)p[1] - (double)x[1];
  double e2 = (double)p[2] - (double)x[2];
  double mag = std::sqrt(e0 * e0 + e1 * e1 + e2 * e2);
  return dot < 0 ? -mag : mag;
}

struct ScanStats {
  int64_t n_points, n_valid, n_skipped;
  int64_t n_block_cells, n_voxel_updates, n_unique_voxels;
  int64_t n_new_blocks, n_trunc_vox, n_trunc_blk, n_index_dropped;
  int64_t n_vox_blocks_outside_blockdda;  // D4 diagnostic
  int64_t n_blocks_after;
};

struct PointWork {
  std::vector<uint64_t> blocks;   // block DDA keys of this point
  std::vector<Rec> recs;          // voxel records in DDA order
  int valid = 0, tb = 0, tv = 0, dropped = 0;
};

static void process_point(const Params& P, const float p[3], const float s[3], PointWork& w,
                          std::vector<Cell>& buf) {
  w.blocks.clear(); w.recs.clear();
  w.valid = 0; w.tb = w.tv = 0; w.dropped = 0;
  if (!point_valid(p, s, P.vs)) return;
  w.valid = 1;
  float a[3], b[3];
  segment(p, s, P.T, a, b);
  int cap = std::max(P.max_blk, P.max_vox);
  if ((int)buf.size() < cap) buf.resize(cap);
  int nb = traverse(a, b, P.BL, P.max_blk, buf.data(), &w.tb);
  for (int i = 0; i < nb; ++i) w.blocks.push_back(pack(buf[i].k));
  int nv = traverse(a, b, P.vs, P.max_vox, buf.data(), &w.tv);
  for (int i = 0; i < nv; ++i) {
    int32_t kb[3]; int local;
    if (!voxel_to_block(P, buf[i].c, kb, &local)) { w.dropped++; continue; }
    Rec r;
    r.bkey = pack(kb);
    r.local = local;
    r.d = signed_distance(p, s, buf[i].c);
    w.recs.push_back(r);
  }
}

static inline void fold(double& S, double& W, double d, double maxW) {
  double w = 1.0 / (std::fabs(d) + 1.0);
  double Wn = W + w;
  S = (W * S + w * d) / Wn;
  W = std::min(Wn, maxW);
}

struct Map {
  Params P;
  std::unordered_map<uint64_t, Block*> blocks;
  ~Map() { for (auto& kv : blocks) delete kv.second; }
  Block* get_or_create(uint64_t k, int64_t* created) {
    auto it = blocks.find(k);
    if (it != blocks.end()) return it->second;
    Block* b = new Block();
    blocks.emplace(k, b);
    if (created) (*created)++;
    return b;
  }
};

static void finish_stats(Map* m, ScanStats* st) { st->n_blocks_after = (int64_t)m->blocks.size(); }

}  // namespace

extern "C" {

void* or_map_create(const float* params5) {
  Map* m = new Map();
  m->P = derive(params5);
  return m;
}

void or_map_destroy(void* h) { delete (Map*)h; }

void or_derived(const float* params5, float* out_f, int* out_i) {
  Params P = derive(params5);
  out_f[0] = P.half_vs; out_f[1] = P.BL; out_f[2] = P.half_BL; out_f[3] = P.blk_eps; out_f[4] = P.vox_eps;
  out_i[0] = P.max_vox; out_i[1] = P.max_blk;
}

// Unit-level entry points for the pins.
void or_segment(const float* p, const float* s, float T, float* a, float* b) { segment(p, s, T, a, b); }

float or_cell_center(float x, float L, int32_t* k) { return cell_center(x, L, k); }

int or_point_valid(const float* p, const float* s, float vs) { return point_valid(p, s, vs) ? 1 : 0; }

int or_traverse(const float* a, const float* b, float L, int max_cells, int32_t* keys, float* centers, int* trunc) {
  std::vector<Cell> buf(std::max(max_cells, 1));
  int n = traverse(a, b, L, max_cells, buf.data(), trunc);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < 3; ++j) { keys[3 * i + j] = buf[i].k[j]; centers[3 * i + j] = buf[i].c[j]; }
  return n;
}

int or_voxel_to_block(const float* params5, const float* vc, int32_t* kb, int* local) {
  Params P = derive(params5);
  return voxel_to_block(P, vc, kb, local) ? 1 : 0;
}

double or_signed_distance(const float* p, const float* s, const float* x) { return signed_distance(p, s, x); }

void or_fold(double* S, double* W, double d, double maxW) { fold(*S, *W, d, maxW); }

// Integrate one scan (H:1003 allocateVoxelBlocksAndUpdateVoxels): points are
// pcl::PointXYZ-layout float4 (x, y, z, pad).  Single-threaded.
int or_integrate(void* h, const float* pts, int64_t n, const float* sensor, ScanStats* st) {
  Map* m = (Map*)h;
  const Params& P = m->P;
  ScanStats z; std::memset(&z, 0, sizeof(z));
  z.n_points = n;
  PointWork w; std::vector<Cell> buf;
  std::unordered_map<uint64_t, char> blockdda;
  std::unordered_map<uint64_t, std::bitset<VPB>> touched_vox;
  for (int64_t i = 0; i < n; ++i) {
    process_point(P, pts + 4 * i, sensor, w, buf);
    if (!w.valid) { z.n_skipped++; continue; }
    z.n_valid++;
    z.n_trunc_blk += w.tb; z.n_trunc_vox += w.tv; z.n_index_dropped += w.dropped;
    z.n_block_cells += (int64_t)w.blocks.size();
    for (uint64_t k : w.blocks) { m->get_or_create(k, &z.n_new_blocks); blockdda[k] = 1; }
  }
  for (int64_t i = 0; i < n; ++i) {
    process_point(P, pts + 4 * i, sensor, w, buf);
    if (!w.valid) continue;
    for (const Rec& r : w.recs) {
      Block* b = m->get_or_create(r.bkey, &z.n_new_blocks);
      if (!blockdda.count(r.bkey)) z.n_vox_blocks_outside_blockdda++;
      fold(b->sdf[r.local], b->w[r.local], r.d, (double)P.maxW);
      z.n_voxel_updates++;
      touched_vox[r.bkey].set(r.local);
    }
  }
  for (auto& kv : touched_vox) z.n_unique_voxels += (int64_t)kv.second.count();
  finish_stats(m, &z);
  if (st) *st = z;
  return 0;
}

// Multithreaded variant, bit-identical to or_integrate: threads traverse
// contiguous point ranges; records are partitioned by block key and each
// partition is folded by one thread in ascending (point, DDA) order.
int or_integrate_mt(void* h, const float* pts, int64_t n, const float* sensor, int nthreads, ScanStats* st) {
  Map* m = (Map*)h;
  const Params P = m->P;
  if (nthreads < 1) nthreads = 1;
  ScanStats z; std::memset(&z, 0, sizeof(z));
  z.n_points = n;
  const int T = nthreads;
  // Phase A: traversal.  per[t][q] = records of chunk t routed to partition q.
  std::vector<std::vector<std::vector<Rec>>> per(T, std::vector<std::vector<Rec>>(T));
  std::vector<std::vector<uint64_t>> blk(T);
  std::vector<ScanStats> zs(T);
  auto part_of = [T](uint64_t k) { return (int)((k * 0x9E3779B97F4A7C15ull) >> 40) % T; };
  {
    std::vector<std::thread> th;
    for (int t = 0; t < T; ++t) {
      th.emplace_back([&, t]() {
        int64_t lo = n * t / T, hi = n * (t + 1) / T;
        ScanStats& s = zs[t]; std::memset(&s, 0, sizeof(s));
        PointWork w; std::vector<Cell> buf;
        for (int64_t i = lo; i < hi; ++i) {
          process_point(P, pts + 4 * i, sensor, w, buf);
          if (!w.valid) { s.n_skipped++; continue; }
          s.n_valid++;
          s.n_trunc_blk += w.tb; s.n_trunc_vox += w.tv; s.n_index_dropped += w.dropped;
          s.n_block_cells += (int64_t)w.blocks.size();
          for (uint64_t k : w.blocks) blk[t].push_back(k);
          for (const Rec& r : w.recs) per[t][part_of(r.bkey)].push_back(r);
        }
      });
    }
    for (auto& x : th) x.join();
  }
  for (int t = 0; t < T; ++t) {
    z.n_valid += zs[t].n_valid; z.n_skipped += zs[t].n_skipped;
    z.n_trunc_blk += zs[t].n_trunc_blk; z.n_trunc_vox += zs[t].n_trunc_vox;
    z.n_index_dropped += zs[t].n_index_dropped; z.n_block_cells += zs[t].n_block_cells;
  }
  // Phase B: block allocation (serial; a set union).
  std::unordered_map<uint64_t, char> blockdda;
  for (int t = 0; t < T; ++t)
    for (uint64_t k : blk[t]) { m->get_or_create(k, &z.n_new_blocks); blockdda[k] = 1; }
  for (int t = 0; t < T; ++t)
    for (int q = 0; q < T; ++q)
      for (const Rec& r : per[t][q]) {
        if (!blockdda.count(r.bkey)) z.n_vox_blocks_outside_blockdda++;
        m->get_or_create(r.bkey, &z.n_new_blocks);
      }
  // Phase C: ordered fold per partition.
  std::vector<int64_t> upd(T, 0), uniq(T, 0);
  {
    std::vector<std::thread> th;
    for (int q = 0; q < T; ++q) {
      th.emplace_back([&, q]() {
        std::unordered_map<uint64_t, std::bitset<VPB>> tv;
        for (int t = 0; t < T; ++t)
          for (const Rec& r : per[t][q]) {
            Block* b = m->blocks.find(r.bkey)->second;
            fold(b->sdf[r.local], b->w[r.local], r.d, (double)P.maxW);
            upd[q]++;
            tv[r.bkey].set(r.local);
          }
        for (auto& kv : tv) uniq[q] += (int64_t)kv.second.count();
      });
    }
    for (auto& x : th) x.join();
  }
  for (int q = 0; q < T; ++q) { z.n_voxel_updates += upd[q]; z.n_unique_voxels += uniq[q]; }
  finish_stats(m, &z);
  if (st) *st = z;
  return 0;
}

// retrievePublishableVoxels (H:801-825, H:767-790): every voxel with
// weight != 0 in a block whose centre is within sqrt(pubD2) of the sensor.
// Output order: blocks by ascending packed key, voxels by local index.
// Returns the total count; writes at most cap entries.
int64_t or_publish(void* h, const float* sensor, float* pos, double* sdf, double* w, int64_t cap) {
  Map* m = (Map*)h;
  const Params& P = m->P;
  std::vector<uint64_t> keys;
  keys.reserve(m->blocks.size());
  for (auto& kv : m->blocks) keys.push_back(kv.first);
  std::sort(keys.begin(), keys.end());
  int64_t cnt = 0;
  for (uint64_t key : keys) {
    int32_t k[3]; unpack(key, k);
    float c[3]; block_center(P, k, c);
    if (!(dist2(c, sensor) <= P.pubD2)) continue;
    Block* b = m->blocks[key];
    float bl[3] = {c[0] - P.half_BL, c[1] - P.half_BL, c[2] - P.half_BL};
    for (int l = 0; l < VPB; ++l) {
      if (b->w[l] != 0) {
        if (cnt < cap) {
          voxel_pos(P, l, bl, pos + 3 * cnt);
          sdf[cnt] = b->sdf[l];
          w[cnt] = b->w[l];
        }
        cnt++;
      }
    }
  }
  return cnt;
}

// garbageCollectDistantBlocks (H:836-952): remove every block whose centre is
// farther than sqrt(gcD2) from the sensor.  Returns the number removed.
int64_t or_gc(void* h, const float* sensor) {
  Map* m = (Map*)h;
  const Params& P = m->P;
  int64_t removed = 0;
  for (auto it = m->blocks.begin(); it != m->blocks.end();) {
    int32_t k[3]; unpack(it->first, k);
    float c[3]; block_center(P, k, c);
    if (!(dist2(c, sensor) <= P.gcD2)) {
      delete it->second;
      it = m->blocks.erase(it);
      removed++;
    } else {
      ++it;
    }
  }
  return removed;
}

int64_t or_num_blocks(void* h) { return (int64_t)((Map*)h)->blocks.size(); }

// Dump the map sorted by packed key: keys int32[n][3], sdf/w double[n][512].
int64_t or_export(void* h, int32_t* keys, double* sdf, double* w, int64_t cap) {
  Map* m = (Map*)h;
  std::vector<uint64_t> ks;
  for (auto& kv : m->blocks) ks.push_back(kv.first);
  std::sort(ks.begin(), ks.end());
  int64_t n = std::min<int64_t>(cap, (int64_t)ks.size());
  for (int64_t i = 0; i < n; ++i) {
    unpack(ks[i], keys + 3 * i);
    Block* b = m->blocks[ks[i]];
    if (sdf) std::memcpy(sdf + VPB * i, b->sdf, sizeof(b->sdf));
    if (w) std::memcpy(w + VPB * i, b->w, sizeof(b->w));
  }
  return (int64_t)ks.size();
}

int or_stats_size() { return (int)sizeof(ScanStats); }

}  // extern "C"
