// opengjk-amd kernels for gfx950: one group of LPP lanes per pair, block of
// 256 threads.  GJK and EPA run fused in one kernel when both are asked for,
// so a pair's vertices are swept from L1/L2 by the same wave throughout.
#include <type_traits>

#include "ogjk_epa.hpp"
#include "ogjk_internal.hpp"

// Minimum resident waves per SIMD the compiler must allow (caps VGPRs at
// 512 / waves).  Chosen by measurement on MI355X.
#ifndef OGJK_GJK_WAVES
#define OGJK_GJK_WAVES 3
#endif
// Clustered and posed bodies carry more live state through GJK.
#ifndef OGJK_GJK_WAVES_WIDE
#define OGJK_GJK_WAVES_WIDE 3
#endif
#ifndef OGJK_GJK_WAVES_POSED
#define OGJK_GJK_WAVES_POSED 3
#endif
// EPA: float scenes swept whole run faster at 3 waves/SIMD even though the
// compiler then spills 80 B/lane (100K x 64: 1.22 -> 1.12 ms); the 64-lane
// kernels of clustered scenes fit 3 waves without spilling (the group is the
// wave, so the pair's descriptors are scalar); narrower clustered kernels
// (200+ B/lane at 3), posed and the other sources stay at 2.
#ifndef OGJK_EPA_WAVES
#define OGJK_EPA_WAVES 2
#endif
#ifndef OGJK_EPA_WAVES_FLOAT
#define OGJK_EPA_WAVES_FLOAT 3
#endif

namespace ogjk {

// Dynamic LDS of the GJK staging area and of the persistent EPA kernel, sized
// per launch for the block's groups (waves per block are a launch parameter).
extern __shared__ __attribute__((aligned(16))) unsigned char ogjk_lds[];

// ----------------------------------------------------------------- sources

// CL: 0 no clusters (the kernels contain no culled sweep at all), 1 clusters,
// 2 clusters and superclusters (float scenes only).
template <class T, int CL = 0>
struct SceneSrc {
  const T* x;
  const T* y;
  const T* z;
  const int64_t* off;
  const int32_t* cnt;
  const float* r1;
  int32_t nb;
  const int32_t* pairs;
  ClusterView cv;        // float scene: vertex clusters, or null bounds
  const int64_t* cslot;  // first cluster slot of each body
  int32_t cluster_min;   // bodies with at least this many vertices are culled
  using BodyT =
      std::conditional_t<CL == 2, CBody2, std::conditional_t<CL == 1, CBody, Body<T, 1>>>;
  static constexpr int kGjkWaves = CL ? OGJK_GJK_WAVES_WIDE : OGJK_GJK_WAVES;
  template <int LPP>
  static constexpr int kEpaWaves = std::is_same_v<T, float> && (CL == 0 || (CL == 1 && LPP == 64))
                                       ? OGJK_EPA_WAVES_FLOAT
                                       : OGJK_EPA_WAVES;
  // the EPA kernel of a superclustered scene is faster with the pair's
  // descriptors in vector registers (10K x 100K: 4.7 vs 5.3 ms)
  static constexpr bool kEpaScalarPair = CL != 2;
  // GJK stages small pairs in LDS (float scenes without clusters)
  static constexpr bool kStage = std::is_same_v<T, float> && CL == 0;

  // max(nA, nB) of pair i, or 0 for an invalid pair (it goes to bucket 0)
  OGJK_DEV int pair_size(int i) const {
    const int a = pairs ? pairs[2 * i] : 2 * i, b = pairs ? pairs[2 * i + 1] : 2 * i + 1;
    if (a < 0 || a >= nb || b < 0 || b >= nb) return 0;
    return cnt[a] > cnt[b] ? cnt[a] : cnt[b];
  }

  template <int LPP>
  OGJK_DEV bool load(int i, int lane, BodyT& A, BodyT& B) const {
    int a, b;
    if (pairs) {
      a = pairs[2 * i];
      b = pairs[2 * i + 1];
    } else {
      a = 2 * i;
      b = 2 * i + 1;
    }
    if (a < 0 || a >= nb || b < 0 || b >= nb) return false;
    const int64_t oa = off[a], ob = off[b];
    const int na = cnt[a], nb_ = cnt[b];
    const float ra = r1 ? r1[a] : 0.0f, rb = r1 ? r1[b] : 0.0f;
    if constexpr (CL != 0) {
      A = BodyT{
          CBody{{x + oa, y + oa, z + oa, na, ra}, cv, na >= cluster_min ? cslot[a] : int64_t(-1)}};
      B = BodyT{CBody{
          {x + ob, y + ob, z + ob, nb_, rb}, cv, nb_ >= cluster_min ? cslot[b] : int64_t(-1)}};
    } else {
      A = BodyT{x + oa, y + oa, z + oa, na, ra};
      B = BodyT{x + ob, y + ob, z + ob, nb_, rb};
    }
    return true;
  }
};

// Scene bodies under per-pair rigid poses: pair i = (pairs[2i], pairs[2i+1])
// placed by poses[14i .. 14i+6] and poses[14i+7 .. 14i+13].
struct PoseSrc {
  const float* x;
  const float* y;
  const float* z;
  const int64_t* off;
  const int32_t* cnt;
  const float* r1;
  int32_t nb;
  const int32_t* pairs;
  const float* poses;
  int32_t by_body;  // poses[7 b] per body b, else poses[14 i], [14 i + 7] per pair side
  using BodyT = PBody;
  static constexpr int kGjkWaves = OGJK_GJK_WAVES_POSED;
  template <int LPP>
  static constexpr int kEpaWaves = OGJK_EPA_WAVES;
  static constexpr bool kEpaScalarPair = false;  // posed pairs: measured no faster
  static constexpr bool kStage = false;

  OGJK_DEV int pair_size(int i) const {
    const int a = pairs[2 * i], b = pairs[2 * i + 1];
    if (a < 0 || a >= nb || b < 0 || b >= nb) return 0;
    return cnt[a] > cnt[b] ? cnt[a] : cnt[b];
  }

  template <int LPP>
  OGJK_DEV bool load(int i, int lane, BodyT& A, BodyT& B) const {
    const int a = pairs[2 * i], b = pairs[2 * i + 1];
    if (a < 0 || a >= nb || b < 0 || b >= nb) return false;
    const int64_t oa = off[a], ob = off[b];
    const float* pa = by_body ? poses + 7 * (int64_t)a : poses + 14 * (int64_t)i;
    const float* pb = by_body ? poses + 7 * (int64_t)b : poses + 14 * (int64_t)i + 7;
    A = PBody{{x + oa, y + oa, z + oa, cnt[a], r1[a]}, pa};
    B = PBody{{x + ob, y + ob, z + ob, cnt[b], r1[b]}, pb};
    return true;
  }
};

struct PolySrc {
  const gkPolytope* a;
  const gkPolytope* b;
  const gkCollisionPair* pairs;
  int32_t npool;
  using BodyT = Body<float, 3>;
  static constexpr int kGjkWaves = OGJK_GJK_WAVES;
  template <int LPP>
  static constexpr int kEpaWaves = OGJK_EPA_WAVES;
  static constexpr bool kEpaScalarPair = false;
  static constexpr bool kStage = false;

  OGJK_DEV int pair_size(int i) const {
    const gkPolytope *pa, *pb;
    if (pairs) {
      const int ia = pairs[i].idx1, ib = pairs[i].idx2;
      if (ia < 0 || ia >= npool || ib < 0 || ib >= npool) return 0;
      pa = a + ia;
      pb = a + ib;
    } else {
      pa = a + i;
      pb = b + i;
    }
    return pa->numpoints > pb->numpoints ? pa->numpoints : pb->numpoints;
  }

  template <int LPP>
  OGJK_DEV bool load(int i, int lane, BodyT& A, BodyT& B) const {
    const gkPolytope* pa;
    const gkPolytope* pb;
    if (pairs) {
      const int ia = pairs[i].idx1, ib = pairs[i].idx2;
      if (ia < 0 || ia >= npool || ib < 0 || ib >= npool) return false;
      pa = a + ia;
      pb = a + ib;
    } else {
      pa = a + i;
      pb = b + i;
    }
    const float* ca = pa->coord;
    const float* cb = pb->coord;
    A = BodyT{ca, ca + 1, ca + 2, pa->numpoints, 0.0f};
    B = BodyT{cb, cb + 1, cb + 2, pb->numpoints, 0.0f};
    if (A.n < 1 || B.n < 1 || ca == nullptr || cb == nullptr) return false;
    A.r1 = body_r1<LPP>(A, lane);
    B.r1 = body_r1<LPP>(B, lane);
    return true;
  }
};

// Plain-FP64 bodies in one AoS pool [x0, y0, z0, x1, ...] (the reference's
// gkPolytope coordinate layout at FP64, GJK/common.h:68-78); body b is
// [off[b], off[b] + cnt[b]).  Comparison path only.
struct Aos64Src {
  const double* xyz;
  const int64_t* off;
  const int32_t* cnt;
  int32_t nb;
  const int32_t* pairs;
  using BodyT = Body<double, 3>;
  static constexpr int kGjkWaves = OGJK_GJK_WAVES;
  template <int LPP>
  static constexpr int kEpaWaves = OGJK_EPA_WAVES;
  static constexpr bool kEpaScalarPair = false;
  static constexpr bool kStage = false;

  OGJK_DEV int pair_size(int i) const {
    const int a = pairs ? pairs[2 * i] : 2 * i, b = pairs ? pairs[2 * i + 1] : 2 * i + 1;
    if (a < 0 || a >= nb || b < 0 || b >= nb) return 0;
    return cnt[a] > cnt[b] ? cnt[a] : cnt[b];
  }

  template <int LPP>
  OGJK_DEV bool load(int i, int lane, BodyT& A, BodyT& B) const {
    const int a = pairs ? pairs[2 * i] : 2 * i, b = pairs ? pairs[2 * i + 1] : 2 * i + 1;
    if (a < 0 || a >= nb || b < 0 || b >= nb) return false;
    const double* pa = xyz + 3 * off[a];
    const double* pb = xyz + 3 * off[b];
    A = BodyT{pa, pa + 1, pa + 2, cnt[a], 0.0f};
    B = BodyT{pb, pb + 1, pb + 2, cnt[b], 0.0f};
    return true;
  }
};

// Sources of the plain-FP64 comparison path (the only ones the reference's
// split launch shape is built for).
template <class Src>
constexpr bool kPlainSrc = std::is_same_v<Src, Aos64Src> ||
                           std::is_same_v<typename Src::BodyT, Body<double, 1>>;

// ----------------------------------------------------------------- records

OGJK_DEV void store_invalid(const OutArgs& o, int i, bool epa) {
  const double nan = __builtin_nan("");
  if (o.contacts) {
    float* c = o.contacts + (int64_t)i * OGJK_CONTACT_FLOATS;
#pragma unroll
    for (int k = 0; k < OGJK_CONTACT_FLOATS; k++) c[k] = (float)nan;
  }
  if (o.distances) o.distances[i] = (float)nan;
  if (o.simplices) o.simplices[i].nvrtx = 0;
  if (epa && o.normals) {
    o.normals[3 * i + 0] = (float)nan;
    o.normals[3 * i + 1] = (float)nan;
    o.normals[3 * i + 2] = (float)nan;
  }
  if (o.f64) {
    for (int k = 0; k < OGJK_F64_REC; k++) o.f64[(int64_t)i * OGJK_F64_REC + k] = nan;
  }
}

OGJK_DEV void store(const OutArgs& o, int i, const Simplex& s, const Wit& w, double dist, D3 n,
                    bool epa) {
  if (o.contacts) {
    // distance, normal (zero without EPA), witness on A, witness on B
    float2* c = (float2*)(o.contacts + (int64_t)i * OGJK_CONTACT_FLOATS);
    c[0] = float2{(float)dist, epa ? (float)n.x : 0.0f};
    c[1] = float2{epa ? (float)n.y : 0.0f, epa ? (float)n.z : 0.0f};
    c[2] = float2{(float)w.w0.x, (float)w.w0.y};
    c[3] = float2{(float)w.w0.z, (float)w.w1.x};
    c[4] = float2{(float)w.w1.y, (float)w.w1.z};
  }
  if (o.simplices) {
    gkSimplex* r = o.simplices + i;
    r->nvrtx = s.n;
#pragma unroll
    for (int k = 0; k < 4; k++) {
      r->vrtx[k][0] = (float)s.p[k].x;
      r->vrtx[k][1] = (float)s.p[k].y;
      r->vrtx[k][2] = (float)s.p[k].z;
      r->vrtx_idx[k][0] = s.ia[k];
      r->vrtx_idx[k][1] = s.ib[k];
    }
    r->witnesses[0][0] = (float)w.w0.x;
    r->witnesses[0][1] = (float)w.w0.y;
    r->witnesses[0][2] = (float)w.w0.z;
    r->witnesses[1][0] = (float)w.w1.x;
    r->witnesses[1][1] = (float)w.w1.y;
    r->witnesses[1][2] = (float)w.w1.z;
  }
  if (o.distances) o.distances[i] = (float)dist;
  if (epa && o.normals) {
    o.normals[3 * i + 0] = (float)n.x;
    o.normals[3 * i + 1] = (float)n.y;
    o.normals[3 * i + 2] = (float)n.z;
  }
  if (o.f64) {
    double* f = o.f64 + (int64_t)i * OGJK_F64_REC;
    f[0] = dist;
    f[1] = w.w0.x; f[2] = w.w0.y; f[3] = w.w0.z;
    f[4] = w.w1.x; f[5] = w.w1.y; f[6] = w.w1.z;
    f[7] = n.x; f[8] = n.y; f[9] = n.z;
#pragma unroll
    for (int k = 0; k < 4; k++) {
      f[10 + 3 * k] = s.p[k].x;
      f[11 + 3 * k] = s.p[k].y;
      f[12 + 3 * k] = s.p[k].z;
    }
    f[22] = 0.0;
    f[23] = 0.0;
  }
}

// EPA-only input: the simplex written by a previous GJK call.  Its vertices
// are exact differences A[ia] - B[ib]; the FP32 record rounds them, so each
// one whose rounding matches is rebuilt in FP64 from its indices (any other
// value is taken as given).  Out-of-range indices are clamped to vertex 0.
template <class BT>
OGJK_DEV void load_input(const OutArgs& o, int i, const BT& A,
                         const BT& B, Simplex& s, Wit& w, double& dist) {
  const gkSimplex* r = o.simplices + i;
  s.n = r->nvrtx;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    int a = r->vrtx_idx[k][0], b = r->vrtx_idx[k][1];
    a = (a >= 0 && a < A.n) ? a : 0;
    b = (b >= 0 && b < B.n) ? b : 0;
    s.ia[k] = a;
    s.ib[k] = b;
    const D3 g{(double)r->vrtx[k][0], (double)r->vrtx[k][1], (double)r->vrtx[k][2]};
    const D3 e = sub(A.d(a), B.d(b));
    const bool same = (float)e.x == r->vrtx[k][0] && (float)e.y == r->vrtx[k][1] &&
                      (float)e.z == r->vrtx[k][2];
    s.p[k] = sel(same, e, g);
  }
  w.w0 = D3{(double)r->witnesses[0][0], (double)r->witnesses[0][1], (double)r->witnesses[0][2]};
  w.w1 = D3{(double)r->witnesses[1][0], (double)r->witnesses[1][1], (double)r->witnesses[1][2]};
  dist = (double)o.distances[i];
}

// ----------------------------------------------------------------- kernels

// One penetrating pair handed from the GJK kernel to the EPA kernel: the FP64
// simplex and witnesses exactly as GJK left them.
struct EpaJob {
  double p[4][3];
  double w0[3], w1[3];
  double dist;
  int ia[4], ib[4];
  int n, pair;
};
static_assert(sizeof(EpaJob) == kEpaJobBytes, "EpaJob layout");

OGJK_DEV void job_store(EpaJob* j, int pair, const Simplex& s, const Wit& w, double dist) {
#pragma unroll
  for (int k = 0; k < 4; k++) {
    j->p[k][0] = s.p[k].x;
    j->p[k][1] = s.p[k].y;
    j->p[k][2] = s.p[k].z;
    j->ia[k] = s.ia[k];
    j->ib[k] = s.ib[k];
  }
  j->w0[0] = w.w0.x; j->w0[1] = w.w0.y; j->w0[2] = w.w0.z;
  j->w1[0] = w.w1.x; j->w1[1] = w.w1.y; j->w1[2] = w.w1.z;
  j->dist = dist;
  j->n = s.n;
  j->pair = pair;
}

OGJK_DEV int job_load(const EpaJob* j, Simplex& s, Wit& w, double& dist) {
#pragma unroll
  for (int k = 0; k < 4; k++) {
    s.p[k] = D3{j->p[k][0], j->p[k][1], j->p[k][2]};
    s.ia[k] = j->ia[k];
    s.ib[k] = j->ib[k];
  }
  w.w0 = D3{j->w0[0], j->w0[1], j->w0[2]};
  w.w1 = D3{j->w1[0], j->w1[1], j->w1[2]};
  dist = j->dist;
  s.n = j->n;
  return j->pair;
}

// Fused per-pair kernel: GJK and/or EPA for pair = group index.  Used for
// EPA on caller-provided simplices (MODE 2) and as the single-kernel variant
// of GJK+EPA (MODE 3).
template <int LPP, int MODE, class Src>
__global__ void __launch_bounds__(256) k_pairs(Src src, OutArgs out, int n) {
  constexpr int GPB = 256 / LPP;
  constexpr bool kEpa = (MODE & 2) != 0;
  __shared__ EpaShared sh[kEpa ? GPB : 1];
  const int lane = threadIdx.x & (LPP - 1);
  const int g = threadIdx.x / LPP;
  const int pair = blockIdx.x * GPB + g;
  if (pair >= n) return;

  typename Src::BodyT A, B;
  const bool ok = src.template load<LPP>(pair, lane, A, B) && A.n > 0 && B.n > 0;
  if (!ok) {
    if (lane == 0) store_invalid(out, pair, kEpa);
    return;
  }
  Simplex s = {};
  Wit w = {};
  double dist = 0.0;
  D3 normal{0.0, 0.0, 0.0};
  if constexpr ((MODE & 1) != 0) {
    dist = gjk<LPP>(A, B, s, w, lane);
  } else {
    load_input(out, pair, A, B, s, w, dist);
  }
  if constexpr (kEpa) epa<LPP>(A, B, s, dist, w, normal, sh[kEpa ? g : 0], lane);
  if (lane == 0) store(out, pair, s, w, dist, normal, kEpa);
}

// ------------------------------------------------- size-bucketed pipeline
//
// Workspace (see workspace_bytes): 8 KiB of counters, then the pair list
// (int32, bucket b at [start_b, start_b + count_b)), the per-block bucket
// counts of the partition, then the EPA jobs (one 192-B record per
// penetrating pair, grouped by EPA lane class).  Counters, in 4-byte words:
//   [0..3]   pairs per bucket      [4..7]  first list slot of each bucket
//   [8..10]  first job slot of each EPA class
//            (written once per launch by k_scan_blocks, read-only after)
//   [32 * (1 + cls)]  jobs emitted for EPA class cls (one 128-B line each)
//   [kClaimBase + (q*8 + s)*32]  claim cursor of slice s of queue q
//            (q 0..3: GJK buckets, q 4..6: EPA classes)
// Words that take atomics never share a 128-B line with words other waves
// read: a device-scope atomic and a load on one line serialise across XCDs.
enum { kCntBucket = 0, kBucketStart = 4, kClassStart = 8, kJobLine = 32 };

OGJK_DEV unsigned* job_counter(unsigned* c, int cls) { return c + kJobLine * (1 + cls); }

OGJK_DEV int bucket_of(const BucketPlan& p, int size) {
  int b = 0;
#pragma unroll
  for (int k = 0; k < kMaxBuckets - 1; k++)
    if (k < p.nb - 1 && size > p.max_verts[k]) b = k + 1;
  return b;
}

// Device-scope atomics on one address serialise across the XCDs, so every
// counter is bumped once per wave: ballot, one atomicAdd of the popcount by
// the lowest lane, broadcast, and a per-lane prefix.  Returns this lane's slot
// (meaningful where pred).  Call with the wave's active lanes converged.
OGJK_DEV unsigned wave_fetch_add(unsigned* ctr, bool pred) {
  const uint64_t m = __ballot(pred);
  if (m == 0) return 0u;
  const int leader = (int)__builtin_ctzll(m);
  unsigned base = 0u;
  if (lane_id() == leader) base = atomicAdd(ctr, (unsigned)__popcll(m));
  base = (unsigned)__shfl((int)base, leader, 64);
  return base + (unsigned)__popcll(m & ((1ull << lane_id()) - 1ull));
}

// Work-claim cursors.  One device-scope atomic address serialises every
// claim, so each queue (4 GJK buckets, 3 EPA classes) is split into kSlices
// contiguous slices with one cursor each, on its own 128-B line.  A wave
// starts on the slice of its block and moves on when a slice runs dry.
constexpr int kSlices = 8;
constexpr int kLineU = 32;  // uints per 128-B line
constexpr int kClaimBase = 256;
constexpr int kQueueEpa = 4;
constexpr size_t kCounterBytes = kWsCounterBytes;  // = (kClaimBase + 7*kSlices*kLineU) * 4
static_assert(kCounterBytes == (kClaimBase + 7 * kSlices * kLineU) * 4, "counter block");

// Claim the next item of [0, count) for every lane with `want` (group
// leaders).  Returns the item, or 0xffffffff when all slices are exhausted.
// Wave-uniform state: slice, tried.
OGJK_DEV unsigned claim(unsigned* c, int queue, unsigned count, int& slice, int& tried,
                        bool want) {
  unsigned result = 0xffffffffu;
  bool need = want;
  slice = __builtin_amdgcn_readfirstlane(slice);
  tried = __builtin_amdgcn_readfirstlane(tried);
  while (__ballot(need) != 0 && tried < kSlices) {
    const unsigned lo = (unsigned)(((uint64_t)count * slice) / kSlices);
    const unsigned hi = (unsigned)(((uint64_t)count * (slice + 1)) / kSlices);
    unsigned* cur = &c[kClaimBase + (queue * kSlices + slice) * kLineU];
    // a drained slice is seen with a load; only live slices take an atomic
    const bool live =
        __builtin_amdgcn_readfirstlane(__atomic_load_n(cur, __ATOMIC_RELAXED)) < hi - lo;
    const unsigned k = live ? wave_fetch_add(cur, need) : 0xffffffffu;
    if (need && live && lo + k < hi) {
      result = lo + k;
      need = false;
    }
    if (__ballot(need) != 0) {
      slice = (slice + 1) & (kSlices - 1);
      tried++;
    }
  }
  return result;
}

// LDS staging for GJK: a pair whose two bodies hold at most CAP vertices
// together is copied once into the group's LDS slice (x | y | z, CAP floats
// each) and its bodies are pointed there, so the ~10 support sweeps of its
// GJK read LDS instead of L2 (11% faster on 64-vertex pairs; EPA, bound by its
// reductions rather than its loads, gains nothing).  Group-uniform decision.
constexpr int kStageGjk = 128;

template <int LPP, int CAP, class BT>
OGJK_DEV void stage_pair(BT& A, BT& B, float* lds, int lane) {
  wave_sync();
  if (A.n + B.n <= CAP) {
    float* lx = lds;
    float* ly = lds + CAP;
    float* lz = lds + 2 * CAP;
    for (int i = lane; i < A.n; i += LPP) {
      lx[i] = A.x[i];
      ly[i] = A.y[i];
      lz[i] = A.z[i];
    }
    for (int i = lane; i < B.n; i += LPP) {
      lx[A.n + i] = B.x[i];
      ly[A.n + i] = B.y[i];
      lz[A.n + i] = B.z[i];
    }
    A.x = lx;
    A.y = ly;
    A.z = lz;
    B.x = lx + A.n;
    B.y = ly + A.n;
    B.z = lz + A.n;
  }
  wave_sync();
}

// Stable partition of the pairs by bucket, in three kernels: per-block
// counts, one exclusive scan over blocks, scatter.  The list keeps the input
// order inside each bucket, so concurrent waves work on neighbouring pairs
// (and neighbouring pages) as in the unbucketed launch.
template <class Src>
__global__ void __launch_bounds__(256) k_classify(Src src, int n, BucketPlan plan, unsigned* blk,
                                                  int nblk) {
  __shared__ unsigned wc[4][kMaxBuckets];
  const int i = blockIdx.x * 256 + threadIdx.x;
  const int b = i < n ? bucket_of(plan, src.pair_size(i)) : -1;
  const int w = threadIdx.x >> 6;
  for (int k = 0; k < plan.nb; k++) {
    const unsigned cnt = (unsigned)__popcll(__ballot(b == k));
    if ((threadIdx.x & 63) == 0) wc[w][k] = cnt;
  }
  __syncthreads();
  if (threadIdx.x < plan.nb) {
    const int k = threadIdx.x;
    blk[(size_t)k * nblk + blockIdx.x] = wc[0][k] + wc[1][k] + wc[2][k] + wc[3][k];
  }
}

// One block: exclusive scan of each bucket's block counts; bucket totals go
// to the counters.
__global__ void __launch_bounds__(1024) k_scan_blocks(unsigned* blk, int nblk, BucketPlan plan,
                                                      unsigned* c) {
  const int nb = plan.nb;
  __shared__ unsigned wsum[16];
  const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int k = 0; k < nb; k++) {
    unsigned* a = blk + (size_t)k * nblk;
    unsigned carry = 0;
    for (int base = 0; base < nblk; base += 1024) {
      const int i = base + threadIdx.x;
      const unsigned v = i < nblk ? a[i] : 0u;
      unsigned x = v;
#pragma unroll
      for (int d = 1; d < 64; d <<= 1) {
        const unsigned y = __shfl_up(x, d, 64);
        if (l >= d) x += y;
      }
      if (l == 63) wsum[w] = x;
      __syncthreads();
      unsigned woff = 0, total = 0;
      for (int q = 0; q < 16; q++) {
        woff += q < w ? wsum[q] : 0u;
        total += wsum[q];
      }
      if (i < nblk) a[i] = carry + woff + x - v;
      carry += total;
      __syncthreads();
    }
    if (threadIdx.x == 0) c[kCntBucket + k] = carry;
  }
  if (threadIdx.x == 0) {
    unsigned start = 0;
    for (int k = 0; k < nb; k++) {
      c[kBucketStart + k] = start;
      start += c[kCntBucket + k];
    }
    for (int cls = 0; cls < 3; cls++) {
      unsigned cs = 0;
      for (int k = 0; k < nb; k++)
        if (plan.epa_class[k] < cls) cs += c[kCntBucket + k];
      c[kClassStart + cls] = cs;
    }
  }
}

template <class Src>
__global__ void __launch_bounds__(256) k_scatter(Src src, int n, BucketPlan plan,
                                                 const unsigned* c, const unsigned* blk, int nblk,
                                                 int* list) {
  __shared__ unsigned wc[4][kMaxBuckets];
  const int i = blockIdx.x * 256 + threadIdx.x;
  const int b = i < n ? bucket_of(plan, src.pair_size(i)) : -1;
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  unsigned rank = 0;
  for (int k = 0; k < plan.nb; k++) {
    const uint64_t m = __ballot(b == k);
    if (l == 0) wc[w][k] = (unsigned)__popcll(m);
    if (b == k) rank = (unsigned)__popcll(m & ((1ull << l) - 1ull));
  }
  __syncthreads();
  if (b >= 0) {
    unsigned off = c[kBucketStart + b] + blk[(size_t)b * nblk + blockIdx.x] + rank;
    for (int q = 0; q < w; q++) off += wc[q][b];
    list[off] = i;
  }
}

// GJK over one bucket, one group per list entry and one wave per block: the
// hardware dispatcher balances pairs of different sizes wave by wave, and no
// wave waits for the slowest pair of a larger block.  Blocks past the bucket's
// count (the grid is sized for every pair) return at once.  With EMIT (GJK
// then EPA) separated pairs are finished here (contact normal from the
// witnesses, GJK/cpu/EPA.c:370-373) and each penetrating pair is appended to
// its EPA class's job list with its FP64 state.  With a single bucket the
// list is the identity and is not materialised.
template <int LPP, bool EMIT, class Src>
__global__ void __launch_bounds__(256, Src::kGjkWaves)
    k_gjk(Src src, OutArgs out, int n, BucketPlan plan, int bucket, unsigned* c, const int* list,
          EpaJob* jobs) {
  const int GPB = (int)blockDim.x / LPP;  // pairs per block: 1, 2 or 4 waves of groups
  const int lane = threadIdx.x & (LPP - 1);
  // the partition's results are read-only here (plain, cached loads)
  const bool multi = plan.nb > 1;
  const unsigned count =
      __builtin_amdgcn_readfirstlane(multi ? c[kCntBucket + bucket] : (unsigned)n);
  unsigned j = blockIdx.x * GPB + threadIdx.x / LPP;
  // a 64-lane group is the wave: say so, and the pair's descriptors load
  // into scalar registers instead of taking vector registers
  if constexpr (LPP == 64) j = __builtin_amdgcn_readfirstlane(j);
  if (j >= count) return;
  const int pair = multi ? list[__builtin_amdgcn_readfirstlane(c[kBucketStart + bucket]) + j]
                         : (int)j;
  typename Src::BodyT A, B;
  const bool ok = src.template load<LPP>(pair, lane, A, B) && A.n > 0 && B.n > 0;
  if (!ok) {
    if (lane == 0) store_invalid(out, pair, EMIT);
    return;
  }
  if constexpr (Src::kStage) {
    float* stg = (float*)ogjk_lds + (threadIdx.x / LPP) * (3 * kStageGjk);
    stage_pair<LPP, kStageGjk>(A, B, stg, lane);
  }
  Simplex s = {};
  Wit w = {};
  const double dist = gjk<LPP>(A, B, s, w, lane);
  const bool pen = EMIT && !(dist > kEps64);
  const int cls = plan.epa_class[bucket];
  const unsigned k = wave_fetch_add(job_counter(c, cls), lane == 0 && pen);
  if (lane != 0) return;
  if (pen) {
    job_store(jobs + (multi ? c[kClassStart + cls] : 0u) + k, pair, s, w, dist);
  } else {
    const D3 normal = EMIT ? contact_normal(w.w0, w.w1) : D3{0.0, 0.0, 0.0};
    store(out, pair, s, w, dist, normal, EMIT);
  }
}

// EPA over one class's job list, persistent: the groups of a wave claim jobs
// together from sliced cursors until the list is exhausted.
template <int LPP, class Src>
__global__ void __launch_bounds__(256, Src::template kEpaWaves<LPP>)
    k_epa_list(Src src, OutArgs out, BucketPlan plan, int cls, unsigned* c, const EpaJob* jobs) {
  const int GPB = (int)blockDim.x / LPP;  // groups per block: 1, 2 or 4 waves of them
  EpaShared* sh = (EpaShared*)ogjk_lds;
  const int lane = threadIdx.x & (LPP - 1);
  const int g = threadIdx.x / LPP;
  const unsigned njobs =
      __builtin_amdgcn_readfirstlane(__atomic_load_n(job_counter(c, cls), __ATOMIC_RELAXED));
  if (blockIdx.x * GPB >= njobs) return;
  const EpaJob* mine =
      jobs + __builtin_amdgcn_readfirstlane(plan.nb > 1 ? c[kClassStart + cls] : 0u);
  int slice = blockIdx.x & (kSlices - 1), tried = 0;
  for (;;) {
    unsigned j = claim(c, kQueueEpa + cls, njobs, slice, tried, lane == 0);
    j = (unsigned)shfl_i<LPP>((int)j, 0);
    if (j >= njobs) break;
    Simplex s = {};
    Wit w = {};
    double dist;
    int pair = job_load(mine + j, s, w, dist);
    // a 64-lane group is the wave: the pair's descriptors can live in scalar
    // registers
    if constexpr (LPP == 64 && Src::kEpaScalarPair) pair = __builtin_amdgcn_readfirstlane(pair);
    typename Src::BodyT A, B;
    src.template load<LPP>(pair, lane, A, B);  // validated by the GJK kernel
    D3 normal{0.0, 0.0, 0.0};
    epa<LPP>(A, B, s, dist, w, normal, sh[g], lane);
    if (lane == 0) store(out, pair, s, w, dist, normal, true);
  }
}

// The reference's launch shape (GJK/gpu/openGJK.cu:2968-3004): a GJK kernel
// over every pair, then an EPA kernel over every pair, one group of LPP lanes
// per pair in each, non-persistent (LPP = 1 is one thread per pair).  The
// FP64 state of a penetrating pair passes through jobs[pair]; separated
// pairs are finished by the GJK kernel and their job is marked n = 0, so
// their EPA group returns at once, as the reference's EPA kernel does.
template <int LPP, bool EMIT, class Src>
__global__ void __launch_bounds__(256) k_gjk_all(Src src, OutArgs out, int n, EpaJob* jobs) {
  constexpr int GPB = 256 / LPP;
  const int lane = threadIdx.x & (LPP - 1);
  const int pair = blockIdx.x * GPB + threadIdx.x / LPP;
  if (pair >= n) return;
  typename Src::BodyT A, B;
  const bool ok = src.template load<LPP>(pair, lane, A, B) && A.n > 0 && B.n > 0;
  if (!ok) {
    if (lane == 0) {
      store_invalid(out, pair, EMIT);
      if (EMIT) jobs[pair].n = 0;
    }
    return;
  }
  Simplex s = {};
  Wit w = {};
  const double dist = gjk<LPP>(A, B, s, w, lane);
  if (lane != 0) return;
  if (EMIT && !(dist > kEps64)) {
    job_store(jobs + pair, pair, s, w, dist);
  } else {
    if (EMIT) jobs[pair].n = 0;
    store(out, pair, s, w, dist, EMIT ? contact_normal(w.w0, w.w1) : D3{0.0, 0.0, 0.0}, EMIT);
  }
}

template <int LPP, class Src>
__global__ void __launch_bounds__(256) k_epa_all(Src src, OutArgs out, int n, const EpaJob* jobs) {
  constexpr int GPB = 256 / LPP;
  __shared__ EpaShared sh[GPB];
  const int lane = threadIdx.x & (LPP - 1);
  const int g = threadIdx.x / LPP;
  const int pair = blockIdx.x * GPB + g;
  if (pair >= n || jobs[pair].n == 0) return;
  Simplex s = {};
  Wit w = {};
  double dist;
  job_load(jobs + pair, s, w, dist);
  typename Src::BodyT A, B;
  src.template load<LPP>(pair, lane, A, B);  // validated by the GJK kernel
  D3 normal{0.0, 0.0, 0.0};
  epa<LPP>(A, B, s, dist, w, normal, sh[g], lane);
  if (lane == 0) store(out, pair, s, w, dist, normal, true);
}

__global__ void __launch_bounds__(256) k_zero(unsigned* c) {
  for (unsigned i = threadIdx.x; i < kCounterBytes / 4; i += 256) c[i] = 0u;
}

// Element-wise kernels over pools of any size run grid-stride over a capped
// grid: one dispatch covers at most 2^32 work-items.
constexpr unsigned kStrideBlocks = 65536;

static unsigned stride_blocks(int64_t n) {
  const int64_t b = (n + 255) / 256;
  return (unsigned)(b < (int64_t)kStrideBlocks ? b : (int64_t)kStrideBlocks);
}

// AoS pool -> SoA arrays, one thread per vertex.
__global__ void __launch_bounds__(256) k_pack(const float* __restrict__ aos, int64_t nv,
                                              float* __restrict__ x, float* __restrict__ y,
                                              float* __restrict__ z) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < nv; i += (int64_t)gridDim.x * 256) {
    x[i] = aos[3 * i + 0];
    y[i] = aos[3 * i + 1];
    z[i] = aos[3 * i + 2];
  }
}

__global__ void __launch_bounds__(256) k_pack_f64(const float* __restrict__ aos, int64_t nv,
                                                  double* __restrict__ x, double* __restrict__ y,
                                                  double* __restrict__ z) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < nv; i += (int64_t)gridDim.x * 256) {
    x[i] = (double)aos[3 * i + 0];
    y[i] = (double)aos[3 * i + 1];
    z[i] = (double)aos[3 * i + 2];
  }
}

__global__ void __launch_bounds__(256) k_pack_f64_aos(const float* __restrict__ aos, int64_t n,
                                                      double* __restrict__ xyz) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
    xyz[i] = (double)aos[i];
}

// Per-body max L1 norm, one wave per body.
__global__ void __launch_bounds__(256) k_body_r1(const float* __restrict__ x,
                                                 const float* __restrict__ y,
                                                 const float* __restrict__ z,
                                                 const int64_t* __restrict__ off,
                                                 const int32_t* __restrict__ cnt, int nb,
                                                 float* __restrict__ r1) {
  const int lane = threadIdx.x & 63;
  const int b = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (b >= nb) return;
  const int64_t o = off[b];
  const Body<float, 1> body{x + o, y + o, z + o, cnt[b], 0.0f};
  const float r = body_r1<64>(body, lane);
  if (lane == 0) r1[b] = r;
}

// ----------------------------------------------------------------- launchers

static LaunchConfig g_cfg;

static bool lanes_ok(int l, bool epa) {
  return (!epa && l == 8) || l == 16 || l == 32 || l == 64;
}

static bool waves_ok(int w) { return w == 1 || w == 2 || w == 4; }

int set_config(const LaunchConfig& c) {
  if (c.nb < 1 || c.nb > kMaxBuckets || c.fused < 0 || c.fused > 2) return -1;
  if (c.fused == 2 && c.nb != 1) return -1;
  for (int b = 0; b < c.nb; b++) {
    if (!waves_ok(c.gjk_waves[b]) || !waves_ok(c.epa_waves[b])) return -1;
    // one lane per pair: GJK of the split shape only
    const bool one = c.fused == 2 && c.gjk_lanes[b] == 1;
    if (!(one || lanes_ok(c.gjk_lanes[b], false)) || !lanes_ok(c.epa_lanes[b], true)) return -1;
    if (b > 0 && c.max_verts[b] <= c.max_verts[b - 1] && b < c.nb - 1) return -1;
  }
  g_cfg = c;
  return 0;
}
LaunchConfig get_config() { return g_cfg; }

static int epa_class_of_lanes(int lanes) { return lanes == 16 ? 0 : (lanes == 32 ? 1 : 2); }

static BucketPlan plan_of(const LaunchConfig& c) {
  BucketPlan p{};
  p.nb = c.nb;
  for (int b = 0; b < kMaxBuckets; b++) {
    p.max_verts[b] = b < c.nb ? c.max_verts[b] : 0x7fffffff;
    p.epa_class[b] = b < c.nb ? epa_class_of_lanes(c.epa_lanes[b]) : 0;
  }
  return p;
}

static size_t up256(size_t b) { return (b + 255) & ~size_t(255); }

// Workspace pieces: counters | pair list | per-block bucket counts | EPA jobs.
static size_t ws_list_off() { return kCounterBytes; }
static int ws_blocks(int n) { return (n + 255) / 256; }
static size_t ws_blk_off(int n) { return kCounterBytes + up256(4 * (size_t)n); }
static size_t ws_jobs_off(int n) {
  return ws_blk_off(n) + up256(4 * (size_t)kMaxBuckets * (size_t)ws_blocks(n));
}

size_t workspace_bytes(int n) {
  const int m = n > 0 ? n : 0;
  return ws_jobs_off(m) + (size_t)m * sizeof(EpaJob);  // == workspace_bytes_for(m)
}
static_assert(workspace_bytes_for(65536) == kCounterBytes + 4 * 65536 + 4 * kMaxBuckets * 256 +
                                               65536 * sizeof(EpaJob),
              "workspace layout");

template <int LPP, int MODE, class Src>
static hipError_t launch_pairs(const Src& src, const OutArgs& o, int n, hipStream_t st) {
  constexpr int GPB = 256 / LPP;
  const unsigned blocks = (unsigned)((n + GPB - 1) / GPB);
  hipLaunchKernelGGL((k_pairs<LPP, MODE, Src>), dim3(blocks), dim3(256), 0, st, src, o, n);
  return hipGetLastError();
}

// Resident grid for a persistent launch over at most n items.
static unsigned persistent_blocks(int n, int gpb, unsigned cap = 4096u) {
  const unsigned want = (unsigned)((n + gpb - 1) / gpb);
  return want < cap ? want : cap;
}

// wpb: waves per block (1, 2, 4).  The grid covers every pair (blocks past
// the bucket's count return at once).
template <int LPP, bool EMIT, class Src>
static hipError_t launch_gjk(const Src& src, const OutArgs& o, int n, const BucketPlan& p, int b,
                             int wpb, void* ws, hipStream_t st) {
  unsigned* c = (unsigned*)ws;
  const int* list = (const int*)((char*)ws + ws_list_off());
  EpaJob* jobs = (EpaJob*)((char*)ws + ws_jobs_off(n));
  const int gpb = 64 * wpb / LPP;
  const size_t lds = Src::kStage ? (size_t)gpb * 3 * kStageGjk * sizeof(float) : 0;
  hipLaunchKernelGGL((k_gjk<LPP, EMIT, Src>), dim3((unsigned)((n + gpb - 1) / gpb)),
                     dim3(64 * wpb), lds, st, src, o, n, p, b, c, list, jobs);
  return hipGetLastError();
}

// Persistent: at most 16,384 waves (four per SIMD of the chip), whatever the
// block size.
template <int LPP, class Src>
static hipError_t launch_epa_list(const Src& src, const OutArgs& o, int n, const BucketPlan& p,
                                  int cls, int wpb, void* ws, hipStream_t st) {
  unsigned* c = (unsigned*)ws;
  const EpaJob* jobs = (const EpaJob*)((char*)ws + ws_jobs_off(n));
  const int gpb = 64 * wpb / LPP;
  hipLaunchKernelGGL((k_epa_list<LPP, Src>), dim3(persistent_blocks(n, gpb, 16384u / wpb)),
                     dim3(64 * wpb), (size_t)gpb * sizeof(EpaShared), st, src, o, p, cls, c,
                     jobs);
  return hipGetLastError();
}

template <bool EMIT, class Src>
static hipError_t gjk_by_lanes(int lanes, const Src& src, const OutArgs& o, int n,
                               const BucketPlan& p, int b, int wpb, void* ws, hipStream_t st) {
  switch (lanes) {
    case 8: return launch_gjk<8, EMIT>(src, o, n, p, b, wpb, ws, st);
    case 16: return launch_gjk<16, EMIT>(src, o, n, p, b, wpb, ws, st);
    case 32: return launch_gjk<32, EMIT>(src, o, n, p, b, wpb, ws, st);
    default: return launch_gjk<64, EMIT>(src, o, n, p, b, wpb, ws, st);
  }
}

template <int MODE, class Src>
static hipError_t pairs_by_lanes(const Src& src, const OutArgs& o, int n, hipStream_t st) {
  switch (g_cfg.epa_lanes[0]) {
    case 16: return launch_pairs<16, MODE>(src, o, n, st);
    case 32: return launch_pairs<32, MODE>(src, o, n, st);
    default: return launch_pairs<64, MODE>(src, o, n, st);
  }
}

// GJK (mode 1) or GJK -> EPA (mode 3) over size buckets; every step is a
// device-driven kernel, so the sequence can be captured in a graph.
template <class Src>
static hipError_t launch_bucketed(int mode, const Src& src, const OutArgs& o, int n, void* ws,
                                  hipStream_t st) {
  const LaunchConfig cfg = g_cfg;
  const BucketPlan p = plan_of(cfg);
  unsigned* c = (unsigned*)ws;
  // GJK alone over a single bucket reads and writes no counter (no
  // partition, no EPA jobs): then there is nothing to clear
  if (mode == 3 || p.nb > 1) hipLaunchKernelGGL(k_zero, dim3(1), dim3(256), 0, st, c);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  if (p.nb > 1) {
    const int blocks = ws_blocks(n);
    unsigned* blk = (unsigned*)((char*)ws + ws_blk_off(n));
    hipLaunchKernelGGL((k_classify<Src>), dim3(blocks), dim3(256), 0, st, src, n, p, blk, blocks);
    hipLaunchKernelGGL(k_scan_blocks, dim3(1), dim3(1024), 0, st, blk, blocks, p, c);
    hipLaunchKernelGGL((k_scatter<Src>), dim3(blocks), dim3(256), 0, st, src, n, p, c, blk, blocks,
                       (int*)((char*)ws + ws_list_off()));
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  for (int b = 0; b < p.nb; b++) {
    const int w = cfg.gjk_waves[b];
    e = mode == 3 ? gjk_by_lanes<true>(cfg.gjk_lanes[b], src, o, n, p, b, w, ws, st)
                  : gjk_by_lanes<false>(cfg.gjk_lanes[b], src, o, n, p, b, w, ws, st);
    if (e != hipSuccess) return e;
  }
  if (mode != 3) return hipSuccess;
  bool done[3] = {false, false, false};
  for (int b = 0; b < p.nb; b++) {
    const int cls = p.epa_class[b];
    if (done[cls]) continue;
    done[cls] = true;
    // one launch per EPA class, shaped by the first bucket of the class
    const int w = cfg.epa_waves[b];
    switch (cfg.epa_lanes[b]) {
      case 16: e = launch_epa_list<16>(src, o, n, p, cls, w, ws, st); break;
      case 32: e = launch_epa_list<32>(src, o, n, p, cls, w, ws, st); break;
      default: e = launch_epa_list<64>(src, o, n, p, cls, w, ws, st); break;
    }
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

// EL = 0: GJK only (one launch).
template <int GL, int EL, class Src>
static hipError_t launch_split_lanes(const Src& src, const OutArgs& o, int n, void* ws,
                                     hipStream_t st) {
  EpaJob* jobs = (EpaJob*)((char*)ws + ws_jobs_off(n));
  hipLaunchKernelGGL((k_gjk_all<GL, (EL > 0), Src>),
                     dim3((unsigned)((n + 256 / GL - 1) / (256 / GL))), dim3(256), 0, st, src, o,
                     n, jobs);
  if constexpr (EL > 0) {
    hipLaunchKernelGGL((k_epa_all<EL, Src>), dim3((unsigned)((n + 256 / EL - 1) / (256 / EL))),
                       dim3(256), 0, st, src, o, n, (const EpaJob*)jobs);
  }
  return hipGetLastError();
}

template <int GL, class Src>
static hipError_t launch_split_epa(int el, const Src& src, const OutArgs& o, int n, void* ws,
                                   hipStream_t st) {
  switch (el) {
    case 0: return launch_split_lanes<GL, 0>(src, o, n, ws, st);
    case 16: return launch_split_lanes<GL, 16>(src, o, n, ws, st);
    case 32: return launch_split_lanes<GL, 32>(src, o, n, ws, st);
    default: return launch_split_lanes<GL, 64>(src, o, n, ws, st);
  }
}

// GJK (mode 1) or GJK then EPA (mode 3) in the reference's launch shape
// (plain-FP64 sources only).
template <class Src>
static hipError_t launch_split(int mode, const Src& src, const OutArgs& o, int n, void* ws,
                               hipStream_t st) {
  if constexpr (kPlainSrc<Src>) {
    const int el = mode == 3 ? g_cfg.epa_lanes[0] : 0;
    switch (g_cfg.gjk_lanes[0]) {
      case 1: return launch_split_epa<1>(el, src, o, n, ws, st);
      case 8: return launch_split_epa<8>(el, src, o, n, ws, st);
      case 16: return launch_split_epa<16>(el, src, o, n, ws, st);
      case 32: return launch_split_epa<32>(el, src, o, n, ws, st);
      default: return launch_split_epa<64>(el, src, o, n, ws, st);
    }
  } else {
    return hipErrorInvalidValue;
  }
}

template <class Src>
static hipError_t launch_any(int mode, const Src& src, const OutArgs& o, int n, void* ws,
                             hipStream_t st) {
  if (n <= 0) return hipSuccess;
  if (mode == 2) return pairs_by_lanes<2>(src, o, n, st);
  if (mode != 1 && mode != 3) return hipErrorInvalidValue;
  // the split shape is the comparison path's; elsewhere it runs the pipeline
  if (g_cfg.fused == 2 && kPlainSrc<Src>) return launch_split(mode, src, o, n, ws, st);
  if (mode == 3 && g_cfg.fused == 1) return pairs_by_lanes<3>(src, o, n, st);
  return launch_bucketed(mode, src, o, n, ws, st);
}

hipError_t launch_scene(int mode, const SceneArgs& s, const OutArgs& o, int n, void* ws,
                        hipStream_t st) {
  const ClusterView cv{s.cx, s.cy, s.cz, s.corig, (const ClusterBound*)s.cbounds,
                       (const ClusterBound*)s.csbounds};
  if (s.cbounds && s.csbounds) {
    const SceneSrc<float, 2> src{(const float*)s.x, (const float*)s.y, (const float*)s.z, s.off,
                                 s.cnt, s.r1, s.n_bodies, s.pairs, cv, s.cslot, s.cluster_min};
    return launch_any(mode, src, o, n, ws, st);
  }
  if (s.cbounds) {
    const SceneSrc<float, 1> src{(const float*)s.x, (const float*)s.y, (const float*)s.z, s.off,
                                 s.cnt, s.r1, s.n_bodies, s.pairs, cv, s.cslot, s.cluster_min};
    return launch_any(mode, src, o, n, ws, st);
  }
  const SceneSrc<float> src{(const float*)s.x, (const float*)s.y, (const float*)s.z, s.off,
                            s.cnt, s.r1, s.n_bodies, s.pairs, cv, nullptr, 0x7fffffff};
  return launch_any(mode, src, o, n, ws, st);
}

hipError_t launch_scene_f64(int mode, const SceneArgs& s, const OutArgs& o, int n, void* ws,
                            hipStream_t st) {
  if (s.xyz) {
    const Aos64Src src{s.xyz, s.off, s.cnt, s.n_bodies, s.pairs};
    return launch_any(mode, src, o, n, ws, st);
  }
  const SceneSrc<double> src{(const double*)s.x, (const double*)s.y, (const double*)s.z, s.off,
                             s.cnt, nullptr, s.n_bodies, s.pairs, ClusterView{}, nullptr,
                             0x7fffffff};
  return launch_any(mode, src, o, n, ws, st);
}

hipError_t launch_posed(int mode, const SceneArgs& s, const float* poses, bool by_body,
                        const OutArgs& o, int n, void* ws, hipStream_t st) {
  const PoseSrc src{(const float*)s.x, (const float*)s.y, (const float*)s.z, s.off, s.cnt, s.r1,
                    s.n_bodies, s.pairs, poses, by_body ? 1 : 0};
  return launch_any(mode, src, o, n, ws, st);
}

hipError_t launch_poly(int mode, const PolyArgs& p, const OutArgs& o, int n, void* ws,
                       hipStream_t st) {
  const PolySrc src{p.a, p.b, p.pairs, p.n_pool};
  return launch_any(mode, src, o, n, ws, st);
}

hipError_t launch_pack_f64(const float* aos, int64_t n_verts, double* x, double* y, double* z,
                           hipStream_t st) {
  if (n_verts <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_pack_f64, dim3(stride_blocks(n_verts)), dim3(256), 0, st, aos, n_verts, x,
                     y, z);
  return hipGetLastError();
}

hipError_t launch_pack_f64_aos(const float* aos, int64_t n_verts, double* xyz, hipStream_t st) {
  if (n_verts <= 0) return hipSuccess;
  const int64_t n = 3 * n_verts;
  hipLaunchKernelGGL(k_pack_f64_aos, dim3(stride_blocks(n)), dim3(256), 0, st, aos, n, xyz);
  return hipGetLastError();
}

hipError_t launch_pack(const float* aos, const int64_t* off, const int32_t* cnt, int n_bodies,
                       int64_t n_verts, float* x, float* y, float* z, float* r1, hipStream_t st) {
  if (n_verts > 0) {
    hipLaunchKernelGGL(k_pack, dim3(stride_blocks(n_verts)), dim3(256), 0, st, aos, n_verts, x, y,
                       z);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  if (n_bodies > 0 && r1) {
    const unsigned blocks = (unsigned)((n_bodies + 3) / 4);
    hipLaunchKernelGGL(k_body_r1, dim3(blocks), dim3(256), 0, st, x, y, z, off, cnt, n_bodies, r1);
    return hipGetLastError();
  }
  return hipSuccess;
}

}  // namespace ogjk
