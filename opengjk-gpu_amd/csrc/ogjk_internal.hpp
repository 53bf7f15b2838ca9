// Internal launch interface between the C ABI (ogjk_api.hip) and the kernels
// (ogjk_kernels.hip).  Host-side only.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "opengjk/common.h"

namespace ogjk {

// Where a pair's two bodies come from.
struct SceneArgs {       // SoA float scene (or double for the plain path)
  const void* x;
  const void* y;
  const void* z;
  const int64_t* off;
  const int32_t* cnt;
  const float* r1;       // may be null for the double scene
  int32_t n_bodies;
  const int32_t* pairs;  // 2 per pair, or null: pair i = (2i, 2i+1)
  // optional vertex clusters (float scene only): bodies with at least
  // cluster_min vertices use the culled support sweep
  const float* cx = nullptr;
  const float* cy = nullptr;
  const float* cz = nullptr;
  const int32_t* corig = nullptr;
  const int64_t* cslot = nullptr;
  const float* cbounds = nullptr;
  const float* csbounds = nullptr;  // supercluster bounds; null: the scene has none
  int32_t cluster_min = 0x7fffffff;
  const double* xyz = nullptr;  // plain path: FP64 AoS pool instead of x/y/z
};

struct PolyArgs {                // gkPolytope descriptors, AoS coordinates
  const gkPolytope* a;           // pair i: (a[i], b[i]) when pairs is null
  const gkPolytope* b;
  const gkCollisionPair* pairs;  // else (a[pairs[i].idx1], a[pairs[i].idx2])
  int32_t n_pool;
};

struct OutArgs {
  gkSimplex* simplices;  // GJK: out; EPA: in/out; null with contacts
  float* distances;      // GJK: out; EPA: in/out; null with contacts
  float* normals;        // EPA: out (3 per pair), may be null
  double* f64;           // optional FP64 verification record, may be null
  float* contacts = nullptr;  // optional compact record, OGJK_CONTACT_FLOATS per pair
};

// Floats per pair in the compact contact record (ogjk_contact).
constexpr int OGJK_CONTACT_FLOATS = 10;

// Doubles per pair in the optional FP64 verification record.
constexpr int OGJK_F64_REC = 24;

// Launch shape.  Pairs are bucketed by max(nA, nB): bucket b holds sizes
// <= max_verts[b] (the last bucket is open-ended) and runs GJK on
// gjk_lanes[b] (8, 16, 32, 64) lanes per pair and EPA on epa_lanes[b] (16,
// 32, 64).  fused 1 runs GJK+EPA as one kernel per pair instead of the
// GJK -> worklist -> persistent EPA pipeline (bucket 0's EPA lanes); fused 2
// runs the plain-FP64 path in the reference's shape (a GJK kernel then an EPA
// kernel, one group per pair each; GJK lanes may be 1) and other paths as 0.
constexpr int kMaxBuckets = 4;
struct LaunchConfig {
  int nb = 1;
  int max_verts[kMaxBuckets] = {0x7fffffff, 0x7fffffff, 0x7fffffff, 0x7fffffff};
  int gjk_lanes[kMaxBuckets] = {8, 8, 8, 8};
  int epa_lanes[kMaxBuckets] = {32, 32, 32, 32};
  // waves per block (1, 2, 4) of the GJK kernel and of the persistent EPA
  // kernel (a class's launch takes the first bucket of the class)
  int gjk_waves[kMaxBuckets] = {1, 1, 1, 1};
  int epa_waves[kMaxBuckets] = {4, 4, 4, 4};
  int fused = 0;
};

// Per-launch copy of the bucket table passed to the kernels.
struct BucketPlan {
  int nb;
  int max_verts[kMaxBuckets];
  int epa_class[kMaxBuckets];  // 0, 1, 2 for EPA on 16, 32, 64 lanes
};
int set_config(const LaunchConfig& c);  // -1 if invalid
LaunchConfig get_config();

// Device workspace a GJK+EPA launch over n pairs needs: 8 KiB of counters,
// the bucketed pair list (4 B per pair), the per-block bucket counts of the
// partition (4 B per bucket per 256 pairs) and one 192-B EPA job per pair,
// each piece 256-B aligned.
constexpr size_t kWsCounterBytes = 8192;
constexpr size_t kEpaJobBytes = 192;
constexpr size_t ws_up256(size_t b) { return (b + 255) & ~size_t(255); }
constexpr size_t workspace_bytes_for(int n) {
  return kWsCounterBytes + ws_up256(4 * (size_t)(n > 0 ? n : 0)) +
         ws_up256(4 * (size_t)kMaxBuckets * (size_t)(((n > 0 ? n : 0) + 255) / 256)) +
         (size_t)(n > 0 ? n : 0) * kEpaJobBytes;
}
size_t workspace_bytes(int n);

// ws: device workspace of workspace_bytes(n) (used by mode 3 only).
hipError_t launch_scene(int mode, const SceneArgs& s, const OutArgs& o, int n, void* ws,
                        hipStream_t st);
hipError_t launch_scene_f64(int mode, const SceneArgs& s, const OutArgs& o, int n, void* ws,
                            hipStream_t st);
hipError_t launch_poly(int mode, const PolyArgs& p, const OutArgs& o, int n, void* ws,
                       hipStream_t st);
// Scene bodies under per-pair poses (14 floats per pair: quaternion w, x, y,
// z and translation of the first body, then of the second); s.pairs required.
// Posed scene bodies: poses per pair side (14 floats per pair) or, by_body,
// per scene body (7 floats per body).
hipError_t launch_posed(int mode, const SceneArgs& s, const float* poses, bool by_body,
                        const OutArgs& o, int n, void* ws, hipStream_t st);
// Vertex clusters (see ogjk_cluster.hip): slots needed, and the build.
int64_t cluster_slots(int64_t n_verts, int32_t n_bodies);
int64_t cluster_supers(int64_t n_slots);
hipError_t build_clusters(const float* x, const float* y, const float* z, const int64_t* off,
                          const int32_t* cnt, int32_t n_bodies, float* cx, float* cy, float* cz,
                          int32_t* orig, int64_t* body_slot, float* bounds, int64_t n_slots,
                          float* sbounds, int64_t* body_super, int64_t n_supers,
                          int64_t* n_supers_used, int32_t min_build, hipStream_t st);
hipError_t launch_pack_f64_aos(const float* aos, int64_t n_verts, double* xyz, hipStream_t st);
hipError_t launch_pack_f64(const float* aos, int64_t n_verts, double* x, double* y, double* z,
                           hipStream_t st);
hipError_t launch_pack(const float* aos, const int64_t* off, const int32_t* cnt, int n_bodies,
                       int64_t n_verts, float* x, float* y, float* z, float* r1, hipStream_t st);

}  // namespace ogjk
