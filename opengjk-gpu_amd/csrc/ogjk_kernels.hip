// opengjk-amd kernels for gfx950: one group of LPP lanes per pair, block of
// 256 threads.  GJK and EPA run fused in one kernel when both are asked for,
// so a pair's vertices are swept from L1/L2 by the same wave throughout.
#include "ogjk_epa.hpp"
#include "ogjk_internal.hpp"

// Minimum resident waves per SIMD the compiler must allow (caps VGPRs at
// 512 / waves).  Chosen by measurement on MI355X.
#ifndef OGJK_GJK_WAVES
#define OGJK_GJK_WAVES 3
#endif
#ifndef OGJK_EPA_WAVES
#define OGJK_EPA_WAVES 2
#endif

namespace ogjk {

// ----------------------------------------------------------------- sources

template <class T>
struct SceneSrc {
  const T* x;
  const T* y;
  const T* z;
  const int64_t* off;
  const int32_t* cnt;
  const float* r1;
  int32_t nb;
  const int32_t* pairs;
  using BodyT = Body<T, 1>;

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
    A = BodyT{x + oa, y + oa, z + oa, cnt[a], r1 ? r1[a] : 0.0f};
    B = BodyT{x + ob, y + ob, z + ob, cnt[b], r1 ? r1[b] : 0.0f};
    return true;
  }
};

struct PolySrc {
  const gkPolytope* a;
  const gkPolytope* b;
  const gkCollisionPair* pairs;
  int32_t npool;
  using BodyT = Body<float, 3>;

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

// ----------------------------------------------------------------- records

OGJK_DEV void store_invalid(const OutArgs& o, int i, bool epa) {
  const double nan = __builtin_nan("");
  o.distances[i] = (float)nan;
  o.simplices[i].nvrtx = 0;
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
  o.distances[i] = (float)dist;
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
template <class T, int STRIDE>
OGJK_DEV void load_input(const OutArgs& o, int i, const Body<T, STRIDE>& A,
                         const Body<T, STRIDE>& B, Simplex& s, Wit& w, double& dist) {
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
static_assert(sizeof(EpaJob) == 192, "EpaJob layout");

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

// GJK for every pair.  With EMIT (GJK then EPA), separated pairs are finished
// here (contact normal from the witnesses, GJK/cpu/EPA.c:370-373) and each
// penetrating pair is appended to the EPA worklist with its FP64 state.
template <int LPP, bool EMIT, class Src>
__global__ void __launch_bounds__(256, OGJK_GJK_WAVES) k_gjk(Src src, OutArgs out, int n, unsigned* counters,
                                             EpaJob* jobs) {
  constexpr int GPB = 256 / LPP;
  const int lane = threadIdx.x & (LPP - 1);
  const int pair = blockIdx.x * GPB + (int)(threadIdx.x / LPP);
  if (pair >= n) return;
  typename Src::BodyT A, B;
  const bool ok = src.template load<LPP>(pair, lane, A, B) && A.n > 0 && B.n > 0;
  if (!ok) {
    if (lane == 0) store_invalid(out, pair, EMIT);
    return;
  }
  Simplex s = {};
  Wit w = {};
  const double dist = gjk<LPP>(A, B, s, w, lane);
  if (lane != 0) return;
  if (EMIT && !(dist > kEps64)) {
    const unsigned j = atomicAdd(&counters[0], 1u);
    job_store(jobs + j, pair, s, w, dist);
    return;
  }
  const D3 normal = EMIT ? contact_normal(w.w0, w.w1) : D3{0.0, 0.0, 0.0};
  store(out, pair, s, w, dist, normal, EMIT);
}

// EPA over the worklist, persistent: each group claims jobs from an atomic
// cursor until the list (length counters[0]) is exhausted.
template <int LPP, class Src>
__global__ void __launch_bounds__(256, OGJK_EPA_WAVES) k_epa_list(Src src, OutArgs out, unsigned* counters,
                                                  const EpaJob* jobs) {
  constexpr int GPB = 256 / LPP;
  __shared__ EpaShared sh[GPB];
  const int lane = threadIdx.x & (LPP - 1);
  const int g = threadIdx.x / LPP;
  const unsigned njobs = __atomic_load_n(&counters[0], __ATOMIC_RELAXED);
  for (;;) {
    unsigned j = 0;
    if (lane == 0) j = atomicAdd(&counters[1], 1u);
    j = (unsigned)shfl_i<LPP>((int)j, 0);
    if (j >= njobs) break;
    Simplex s = {};
    Wit w = {};
    double dist;
    const int pair = job_load(jobs + j, s, w, dist);
    typename Src::BodyT A, B;
    src.template load<LPP>(pair, lane, A, B);  // validated by the GJK kernel
    D3 normal{0.0, 0.0, 0.0};
    epa<LPP>(A, B, s, dist, w, normal, sh[g], lane);
    if (lane == 0) store(out, pair, s, w, dist, normal, true);
  }
}

__global__ void k_zero(unsigned* counters) {
  if (threadIdx.x < 2) counters[threadIdx.x] = 0u;
}

// AoS pool -> SoA arrays, one thread per vertex.
__global__ void __launch_bounds__(256) k_pack(const float* __restrict__ aos, int64_t nv,
                                              float* __restrict__ x, float* __restrict__ y,
                                              float* __restrict__ z) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nv) return;
  x[i] = aos[3 * i + 0];
  y[i] = aos[3 * i + 1];
  z[i] = aos[3 * i + 2];
}

__global__ void __launch_bounds__(256) k_pack_f64(const float* __restrict__ aos, int64_t nv,
                                                  double* __restrict__ x, double* __restrict__ y,
                                                  double* __restrict__ z) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nv) return;
  x[i] = (double)aos[3 * i + 0];
  y[i] = (double)aos[3 * i + 1];
  z[i] = (double)aos[3 * i + 2];
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

int set_config(const LaunchConfig& c) {
  const bool gok = c.gjk_lanes == 8 || c.gjk_lanes == 16 || c.gjk_lanes == 32 || c.gjk_lanes == 64;
  const bool eok = c.epa_lanes == 16 || c.epa_lanes == 32 || c.epa_lanes == 64;
  if (!gok || !eok) return -1;
  g_cfg = c;
  return 0;
}
LaunchConfig get_config() { return g_cfg; }

size_t workspace_bytes(int n) { return 256 + (size_t)(n > 0 ? n : 0) * sizeof(EpaJob); }

template <int LPP, int MODE, class Src>
static hipError_t launch_pairs(const Src& src, const OutArgs& o, int n, hipStream_t st) {
  constexpr int GPB = 256 / LPP;
  const unsigned blocks = (unsigned)((n + GPB - 1) / GPB);
  hipLaunchKernelGGL((k_pairs<LPP, MODE, Src>), dim3(blocks), dim3(256), 0, st, src, o, n);
  return hipGetLastError();
}

template <int LPP, bool EMIT, class Src>
static hipError_t launch_gjk(const Src& src, const OutArgs& o, int n, void* ws, hipStream_t st) {
  constexpr int GPB = 256 / LPP;
  const unsigned blocks = (unsigned)((n + GPB - 1) / GPB);
  unsigned* counters = (unsigned*)ws;
  EpaJob* jobs = (EpaJob*)((char*)ws + 256);
  hipLaunchKernelGGL((k_gjk<LPP, EMIT, Src>), dim3(blocks), dim3(256), 0, st, src, o, n, counters,
                     jobs);
  return hipGetLastError();
}

template <int LPP, class Src>
static hipError_t launch_epa_list(const Src& src, const OutArgs& o, int n, void* ws,
                                  hipStream_t st) {
  constexpr int GPB = 256 / LPP;
  // enough resident blocks to fill the device; the groups pull jobs until done
  const unsigned want = (unsigned)((n + GPB - 1) / GPB);
  const unsigned blocks = want < 4096u ? want : 4096u;
  unsigned* counters = (unsigned*)ws;
  const EpaJob* jobs = (const EpaJob*)((char*)ws + 256);
  hipLaunchKernelGGL((k_epa_list<LPP, Src>), dim3(blocks), dim3(256), 0, st, src, o, counters, jobs);
  return hipGetLastError();
}

template <class Src>
static hipError_t gjk_by_lanes(bool emit, const Src& src, const OutArgs& o, int n, void* ws,
                               hipStream_t st) {
  switch (g_cfg.gjk_lanes) {
    case 8: return emit ? launch_gjk<8, true>(src, o, n, ws, st) : launch_gjk<8, false>(src, o, n, ws, st);
    case 16: return emit ? launch_gjk<16, true>(src, o, n, ws, st) : launch_gjk<16, false>(src, o, n, ws, st);
    case 32: return emit ? launch_gjk<32, true>(src, o, n, ws, st) : launch_gjk<32, false>(src, o, n, ws, st);
    default: return emit ? launch_gjk<64, true>(src, o, n, ws, st) : launch_gjk<64, false>(src, o, n, ws, st);
  }
}

template <int MODE, class Src>
static hipError_t pairs_by_lanes(const Src& src, const OutArgs& o, int n, hipStream_t st) {
  switch (g_cfg.epa_lanes) {
    case 16: return launch_pairs<16, MODE>(src, o, n, st);
    case 32: return launch_pairs<32, MODE>(src, o, n, st);
    default: return launch_pairs<64, MODE>(src, o, n, st);
  }
}

template <class Src>
static hipError_t epa_list_by_lanes(const Src& src, const OutArgs& o, int n, void* ws,
                                    hipStream_t st) {
  switch (g_cfg.epa_lanes) {
    case 16: return launch_epa_list<16>(src, o, n, ws, st);
    case 32: return launch_epa_list<32>(src, o, n, ws, st);
    default: return launch_epa_list<64>(src, o, n, ws, st);
  }
}

template <class Src>
static hipError_t launch_any(int mode, const Src& src, const OutArgs& o, int n, void* ws,
                             hipStream_t st) {
  if (n <= 0) return hipSuccess;
  if (mode == 1) return gjk_by_lanes(false, src, o, n, ws, st);
  if (mode == 2) return pairs_by_lanes<2>(src, o, n, st);
  if (mode != 3) return hipErrorInvalidValue;
  if (g_cfg.fused) return pairs_by_lanes<3>(src, o, n, st);
  hipLaunchKernelGGL(k_zero, dim3(1), dim3(64), 0, st, (unsigned*)ws);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  e = gjk_by_lanes(true, src, o, n, ws, st);
  if (e != hipSuccess) return e;
  return epa_list_by_lanes(src, o, n, ws, st);
}

hipError_t launch_scene(int mode, const SceneArgs& s, const OutArgs& o, int n, void* ws,
                        hipStream_t st) {
  const SceneSrc<float> src{(const float*)s.x, (const float*)s.y, (const float*)s.z, s.off,
                            s.cnt, s.r1, s.n_bodies, s.pairs};
  return launch_any(mode, src, o, n, ws, st);
}

hipError_t launch_scene_f64(int mode, const SceneArgs& s, const OutArgs& o, int n, void* ws,
                            hipStream_t st) {
  const SceneSrc<double> src{(const double*)s.x, (const double*)s.y, (const double*)s.z, s.off,
                             s.cnt, nullptr, s.n_bodies, s.pairs};
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
  const unsigned blocks = (unsigned)((n_verts + 255) / 256);
  hipLaunchKernelGGL(k_pack_f64, dim3(blocks), dim3(256), 0, st, aos, n_verts, x, y, z);
  return hipGetLastError();
}

hipError_t launch_pack(const float* aos, const int64_t* off, const int32_t* cnt, int n_bodies,
                       int64_t n_verts, float* x, float* y, float* z, float* r1, hipStream_t st) {
  if (n_verts > 0) {
    const unsigned blocks = (unsigned)((n_verts + 255) / 256);
    hipLaunchKernelGGL(k_pack, dim3(blocks), dim3(256), 0, st, aos, n_verts, x, y, z);
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
