// Vertex clusters for culled support sweeps: per body, the vertices are
// copied in Hilbert order of their direction from the body's centre, padded
// to a multiple of kCluster slots, and every run of kCluster slots gets a
// bounding sphere and box.  Built once per scene; the support sweep that reads it
// is in ogjk_core.hpp (support(const CBody&, ...)).
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstdlib>
#include <vector>

#include "ogjk_core.hpp"
#include "ogjk_internal.hpp"

namespace ogjk {

namespace {

// Bodies of at most this many vertices are built by one block each, sorted in
// LDS (at most 256 threads x 16 keys), in size classes of at most 64, 256,
// 768, 1024, 1536, 2048, 3072 and 4096 vertices (a block sorts its class's
// maximum, padding included).
constexpr int kLdsBodyMax = 4096;
constexpr int kClasses = 8;

OGJK_DEV int body_class(int n) {
  return n <= 64     ? 0
         : n <= 256  ? 1
         : n <= 768  ? 2
         : n <= 1024 ? 3
         : n <= 1536 ? 4
         : n <= 2048 ? 5
         : n <= 3072 ? 6
                     : 7;
}

// Per body: its padded slot count when it is built (n >= min_build), else 0,
// with psize[n_bodies] = 0 closing the scan; and its list: one of the
// one-block classes (n <= lds_max) or kClasses, the batched bodies, as
// lists[k * n_bodies + i].  One atomic per list per block; the order inside a
// list is arbitrary (a body's output depends only on its own slots).
__global__ void __launch_bounds__(256) k_body_classes(const int32_t* __restrict__ cnt,
                                                      int n_bodies, int min_build, int lds_max,
                                                      unsigned* __restrict__ counts,
                                                      int* __restrict__ lists,
                                                      int64_t* __restrict__ psize) {
  __shared__ unsigned hist[kClasses + 1], base[kClasses + 1];
  if (threadIdx.x <= kClasses) hist[threadIdx.x] = 0u;
  __syncthreads();
  const int b = blockIdx.x * 256 + threadIdx.x;
  int k = -1;
  unsigned r = 0;
  if (b < n_bodies) {
    const int n = cnt[b];
    const bool built = n > 0 && n >= min_build;
    psize[b] = built ? ((int64_t)n + kCluster - 1) / kCluster * kCluster : 0;
    if (built) {
      k = n <= lds_max ? body_class(n) : kClasses;
      r = atomicAdd(&hist[k], 1u);
    }
  } else if (b == n_bodies) {
    psize[b] = 0;
  }
  __syncthreads();
  if (threadIdx.x <= kClasses)
    base[threadIdx.x] = hist[threadIdx.x] ? atomicAdd(&counts[threadIdx.x], hist[threadIdx.x]) : 0u;
  __syncthreads();
  if (k >= 0) lists[(size_t)k * n_bodies + base[k] + r] = b;
}

// body_slot[b]: first slot of a built body (the exclusive scan of psize), -1
// for the others; body_super[b] = -1 (k_set_super marks the superclustered).
__global__ void __launch_bounds__(256) k_body_slots(const int64_t* __restrict__ psize,
                                                    const int64_t* __restrict__ pos, int n_bodies,
                                                    int64_t* __restrict__ body_slot,
                                                    int64_t* __restrict__ body_super) {
  const int b = blockIdx.x * 256 + threadIdx.x;
  if (b >= n_bodies) return;
  body_slot[b] = psize[b] > 0 ? pos[b] : -1;
  body_super[b] = -1;
}

// Count and first slot of each batched body: out[2q], out[2q + 1].
__global__ void __launch_bounds__(256) k_gather_large(const int32_t* __restrict__ cnt,
                                                      const int64_t* __restrict__ body_slot,
                                                      const int* __restrict__ large, int n_large,
                                                      int64_t* __restrict__ out) {
  const int q = blockIdx.x * 256 + threadIdx.x;
  if (q >= n_large) return;
  out[2 * q] = cnt[large[q]];
  out[2 * q + 1] = body_slot[large[q]];
}

// body_super[bs[2q]] = bs[2q + 1]: first supercluster of each superclustered body.
__global__ void __launch_bounds__(256) k_set_super(const int64_t* __restrict__ bs, int n,
                                                   int64_t* __restrict__ body_super) {
  const int q = blockIdx.x * 256 + threadIdx.x;
  if (q < n) body_super[bs[2 * q]] = bs[2 * q + 1];
}

// Sort keys: Hilbert index on a 2^G x 2^G grid over the octahedral map of
// directions.  Bodies of at most kLdsBodyMax vertices take G = 8 (256^2
// cells, about 500 per cluster of a 4,096-vertex body, and two 8-bit passes
// of the block sort), larger ones G = 12 (4096^2 cells, finer than a cluster
// of bodies of 10^5 vertices; their device sort runs over kKeyBits bits).
// The grid depends only on the body's size, so both build paths key a body
// alike; vertices sharing a cell keep their input order.
constexpr int kGridBitsSmall = 8, kGridBitsLarge = 12;
constexpr int kKeyBits = 2 * kGridBitsLarge;
// The one-block sort packs a body-local index (< kLdsBodyMax) below the key.
constexpr int kIndexBits = 12;

// Index of (x, y) on the Hilbert curve over the 2^G x 2^G grid.
template <int G>
OGJK_DEV uint32_t hilbert_key(uint32_t x, uint32_t y) {
  constexpr uint32_t n = 1u << G;
  uint32_t d = 0;
#pragma unroll
  for (uint32_t s = n / 2; s > 0; s >>= 1) {
    const uint32_t rx = (x & s) ? 1u : 0u, ry = (y & s) ? 1u : 0u;
    d += s * s * ((3u * rx) ^ ry);
    if (ry == 0) {
      if (rx == 1) {
        x = n - 1 - x;
        y = n - 1 - y;
      }
      const uint32_t t = x;
      x = y;
      y = t;
    }
  }
  return d;
}

// Key of the direction v (any length): its cell on the octahedral map of
// the 2^G x 2^G grid, then that cell's Hilbert index.  (An 8-bit key, one
// radix pass, built C2's clusters 2.5 ms faster and its queries 1 ms slower.)
template <int G>
OGJK_DEV uint32_t direction_key(float vx, float vy, float vz) {
  const float l1 = fabsf(vx) + fabsf(vy) + fabsf(vz);
  float u = 0.0f, w = 0.0f;
  if (l1 > 0.0f) {
    // the key only groups nearby directions: the hardware reciprocal is enough
    const float inv = __builtin_amdgcn_rcpf(l1);
    u = vx * inv;
    w = vy * inv;
    if (vz < 0.0f) {
      const float uu = (1.0f - fabsf(w)) * (u >= 0.0f ? 1.0f : -1.0f);
      const float ww = (1.0f - fabsf(u)) * (w >= 0.0f ? 1.0f : -1.0f);
      u = uu;
      w = ww;
    }
  }
  constexpr float half = 0.5f * (float)(1u << G) - 0.5f, top = (float)((1u << G) - 1);
  const float fu = (u + 1.0f) * half, fw = (w + 1.0f) * half;
  const uint32_t qx = fu > 0.0f ? (uint32_t)fminf(fu, top) : 0u;  // NaN -> 0
  const uint32_t qy = fw > 0.0f ? (uint32_t)fminf(fw, top) : 0u;
  return hilbert_key<G>(qx, qy);
}

OGJK_DEV float block_min(float v, float* red) {
  for (int m = 32; m >= 1; m >>= 1) v = fminf(v, __shfl_xor(v, m, 64));
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  return fminf(fminf(red[0], red[1]), fminf(red[2], red[3]));
}

OGJK_DEV float block_max(float v, float* red) { return -block_min(-v, red); }

// One block per body (blist[b0 + block]): sort key of every vertex = Hilbert index of its
// direction from the centre of the body's bounding box (octahedral map), so
// that consecutive vertices lie in the same narrow cone of directions, the
// set a support query in nearby directions reaches; value = body-local index.
__global__ void __launch_bounds__(256) k_order_keys(const float* __restrict__ x,
                                                    const float* __restrict__ y,
                                                    const float* __restrict__ z,
                                                    const int64_t* __restrict__ off,
                                                    const int32_t* __restrict__ cnt,
                                                    const int* __restrict__ blist, int b0,
                                                    const int* __restrict__ kbeg,
                                                    uint32_t* __restrict__ keys,
                                                    int* __restrict__ vals) {
  __shared__ float red[4];
  const int b = blist[b0 + blockIdx.x];
  const int n = cnt[b];
  if (n <= 0) return;
  const int64_t o = off[b];
  float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int i = threadIdx.x; i < n; i += 256) {
    const float p[3] = {x[o + i], y[o + i], z[o + i]};
#pragma unroll
    for (int k = 0; k < 3; k++) {
      lo[k] = fminf(lo[k], p[k]);
      hi[k] = fmaxf(hi[k], p[k]);
    }
  }
  float c[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    lo[k] = block_min(lo[k], red);
    hi[k] = block_max(hi[k], red);
    c[k] = 0.5f * lo[k] + 0.5f * hi[k];
  }
  const int base = kbeg[blockIdx.x];
  const bool small = n <= kLdsBodyMax;  // uniform over the block
  for (int i = threadIdx.x; i < n; i += 256) {
    const float vx = x[o + i] - c[0], vy = y[o + i] - c[1], vz = z[o + i] - c[2];
    keys[base + i] = small ? direction_key<kGridBitsSmall>(vx, vy, vz)
                           : direction_key<kGridBitsLarge>(vx, vy, vz);
    vals[base + i] = i;
  }
}

// The value of lane CTRL-permuted within each 16-lane row (DPP: CTRL 0xB1
// pairs lanes i and i ^ 1, 0x4E lanes i and i ^ 2, 0x141 mirrors each 8-lane
// half, 0x140 each row).
template <int CTRL>
OGJK_DEV float row_perm(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}

// op over the 32-lane slice (kCluster lanes), in every lane of it, with the
// whole wave active: four DPP steps inside each 16-lane row, then
// v_permlane16_swap exchanges the slice's two rows.  No LDS traffic.
template <class Op>
OGJK_DEV float slice_reduce(float v, Op op) {
  static_assert(kCluster == 32, "a slice is two 16-lane rows");
  v = op(v, row_perm<0xB1>(v));
  v = op(v, row_perm<0x4E>(v));
  v = op(v, row_perm<0x141>(v));
  v = op(v, row_perm<0x140>(v));
  const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(v), __float_as_uint(v), false, false);
  return op(__uint_as_float(r[0]), __uint_as_float(r[1]));
}

// op over the whole wave, in every lane: the slice reduction, then
// v_permlane32_swap exchanges the wave's two halves.
template <class Op>
OGJK_DEV float wave_reduce(float v, Op op) {
  v = slice_reduce(v, op);
  const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(v), __float_as_uint(v), false, false);
  return op(__uint_as_float(r[0]), __uint_as_float(r[1]));
}

// Bounds of one cluster held by a kCluster-lane slice, lane t holding slot t
// (valid: a vertex, not padding), with the whole wave active: centre c =
// midpoint of the bounding box (rounded to float), radius r = max |x - c| and
// half extents h_k = max |x_k - c_k|, computed in FP64 and rounded up to
// float.  Every lane of the slice returns the bound.
OGJK_DEV ClusterBound slice_bound(bool valid, const float (&p)[3]) {
  float lo[3], hi[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    lo[k] = valid ? p[k] : INFINITY;
    hi[k] = valid ? p[k] : -INFINITY;
  }
  const auto fmin_ = [](float a, float b) { return fminf(a, b); };
  const auto fmax_ = [](float a, float b) { return fmaxf(a, b); };
#pragma unroll
  for (int k = 0; k < 3; k++) {
    lo[k] = slice_reduce(lo[k], fmin_);
    hi[k] = slice_reduce(hi[k], fmax_);
  }
  float ctr[3];
  double e[3], r2 = 0.0;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    ctr[k] = 0.5f * lo[k] + 0.5f * hi[k];
    e[k] = valid ? fabs((double)p[k] - (double)ctr[k]) : 0.0;
  }
  if (valid) r2 = e[0] * e[0] + e[1] * e[1] + e[2] * e[2];
  // Each lane rounds its own distances up to float, then the slice takes the
  // maximum: the correctly rounded square root, the conversion to float and
  // the step up are all monotone, so this is bit for bit the rounded-up
  // maximum, with 32-bit rather than 64-bit exchanges.
  float rr = nextafterf((float)sqrt(r2), INFINITY), hh[3];
#pragma unroll
  for (int k = 0; k < 3; k++) hh[k] = nextafterf((float)e[k], INFINITY);
  rr = slice_reduce(rr, fmax_);
#pragma unroll
  for (int k = 0; k < 3; k++) hh[k] = slice_reduce(hh[k], fmax_);
  ClusterBound cb;
  cb.c = make_float4(ctr[0], ctr[1], ctr[2], rr);
  cb.h = make_float4(hh[0], hh[1], hh[2], 0.0f);
  return cb;
}

// Minimum of lo[k] and maximum of hi[k] (k = 0..2) over a block of BLOCK
// threads (a multiple of the wave), in every thread, with the whole block
// active: each wave reduces with DPP and lane permutes, then the waves' six
// values pass once through red (6 * BLOCK / 64 floats) behind one barrier.
template <int BLOCK>
OGJK_DEV void block_box(float (&lo)[3], float (&hi)[3], float* red) {
  const auto fmin_ = [](float a, float b) { return fminf(a, b); };
  const auto fmax_ = [](float a, float b) { return fmaxf(a, b); };
#pragma unroll
  for (int k = 0; k < 3; k++) {
    lo[k] = wave_reduce(lo[k], fmin_);
    hi[k] = wave_reduce(hi[k], fmax_);
  }
  if constexpr (BLOCK > 64) {
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
      for (int k = 0; k < 3; k++) {
        red[6 * w + k] = lo[k];
        red[6 * w + 3 + k] = hi[k];
      }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 3; k++) {
      float a = red[k], b = red[3 + k];
#pragma unroll
      for (int u = 1; u < BLOCK / 64; u++) {
        a = fminf(a, red[6 * u + k]);
        b = fmaxf(b, red[6 * u + 3 + k]);
      }
      lo[k] = a;
      hi[k] = b;
    }
  }
}

// Bodies of at most BLOCK * IPT vertices: one block per body (blist) builds
// its whole part of the cluster copy -- the same keys as k_order_keys, a
// stable sort in LDS, the copy in sorted order and every cluster's bound --
// which is exactly what the batched path (keys, device segmented sort,
// scatter with bounds) produces, in one pass over the body.  Small
// bodies take one-wave blocks (no block barriers in the reductions).
template <int BLOCK, int IPT>
__global__ void __launch_bounds__(BLOCK) k_cluster_body(
    const float* __restrict__ x, const float* __restrict__ y, const float* __restrict__ z,
    const int64_t* __restrict__ off, const int32_t* __restrict__ cnt,
    const int64_t* __restrict__ slot, const int* __restrict__ blist, float* __restrict__ cx,
    float* __restrict__ cy, float* __restrict__ cz, int32_t* __restrict__ orig,
    ClusterBound* __restrict__ bounds) {
  // Keys only (the index rides below the key), ranked a wave at a time by
  // matching; the input is warp-striped -- item j of lane ln in wave wv is
  // element (wv * IPT + j) * 64 + ln -- which is index order, so no exchange
  // precedes the sort.
  using Sort = ::rocprim::block_radix_sort<uint32_t, BLOCK, IPT>;
  static_assert(kLdsBodyMax <= (1 << kIndexBits), "a body-local index fits below the key");
  __shared__ typename Sort::storage_type sort_storage;
  __shared__ float red[6 * (BLOCK / 64)];
  const int b = blist[blockIdx.x];
  const int n = cnt[b];  // 0 < n <= BLOCK * IPT: the host's classes
  const int64_t o = off[b], s = slot[b];
  const int t = threadIdx.x, wv = t >> 6, ln = t & 63;

  // bounding-box centre, then the keys, the coordinates kept in registers
  // between the two (measured faster than loading them again, even at 16
  // items a thread)
  float px[IPT], py[IPT], pz[IPT];
  float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
#pragma unroll
  for (int j = 0; j < IPT; j++) {
    const int i = (wv * IPT + j) * 64 + ln;
    if (i < n) {
      const float p[3] = {x[o + i], y[o + i], z[o + i]};
      px[j] = p[0];
      py[j] = p[1];
      pz[j] = p[2];
#pragma unroll
      for (int k = 0; k < 3; k++) {
        lo[k] = fminf(lo[k], p[k]);
        hi[k] = fmaxf(hi[k], p[k]);
      }
    }
  }
  float c[3];
  block_box<BLOCK>(lo, hi, red);
#pragma unroll
  for (int k = 0; k < 3; k++) c[k] = 0.5f * lo[k] + 0.5f * hi[k];
  // item (key << kIndexBits) | index, sorted on the key bits only: the index
  // travels with its key and no value array goes through the sort
  uint32_t key[IPT];
#pragma unroll
  for (int j = 0; j < IPT; j++) {
    const int i = (wv * IPT + j) * 64 + ln;
    key[j] = 0xffffffffu;  // padding sorts after every vertex (stable: after equal keys)
    if (i < n) {
      key[j] = direction_key<kGridBitsSmall>(px[j] - c[0], py[j] - c[1], pz[j] - c[2])
                   << kIndexBits |
               (uint32_t)i;
    }
  }
  // a stable sort keeps index order among equal keys, as the segmented sort does
  Sort().sort_warp_striped_to_striped(key, sort_storage, kIndexBits,
                                      kIndexBits + 2 * kGridBitsSmall);

  // rank j*BLOCK + t: the copy (coalesced) and, per 32-lane slice, one cluster
  const int padded = (n + kCluster - 1) / kCluster * kCluster;
#pragma unroll
  for (int j = 0; j < IPT; j++) {
    const int r = j * BLOCK + t;
    if (j * BLOCK >= padded) break;  // uniform over the block
    const bool valid = r < n;
    const int l = (int)(key[j] & ((1u << kIndexBits) - 1u));
    const float p[3] = {valid ? x[o + l] : 0.0f, valid ? y[o + l] : 0.0f,
                        valid ? z[o + l] : 0.0f};
    if (r < padded) {
      cx[s + r] = p[0];
      cy[s + r] = p[1];
      cz[s + r] = p[2];
      orig[s + r] = valid ? l : -1;
    }
    const ClusterBound cb = slice_bound(valid, p);
    if ((t & (kCluster - 1)) == 0 && r < padded) bounds[(s + r) / kCluster] = cb;
  }
}

// The batch's padded slots, one thread each.  Padded position p of the batch
// (its bodies blist[b0 + q] in order, each padded to a multiple of kCluster;
// body q from pbeg[q], q by search) is slot i = p - pbeg[q] of that body: it
// gets the vertex sorted to the body's rank i (segment q of `sorted` starts
// at kbeg[q]), or padding (zeros, orig -1) past the body's count.  Bodies
// start on cluster boundaries, so each 32-lane slice holds one whole cluster
// and writes its bound as well.  Blocks are remapped so each XCD (blocks i,
// i + 8, ...) takes one contiguous eighth of the positions: an XCD then
// gathers from about one body at a time, whose vertices stay in its L2 (one
// block per body gathered from many bodies at once and missed in L2 on nearly
// every load).
__global__ void __launch_bounds__(256) k_cluster_scatter(
    const float* __restrict__ x, const float* __restrict__ y, const float* __restrict__ z,
    const int64_t* __restrict__ off, const int64_t* __restrict__ slot,
    const int* __restrict__ blist, int b0, int nb, const int* __restrict__ kbeg,
    const int* __restrict__ pbeg, int ptotal, const int* __restrict__ sorted,
    float* __restrict__ cx, float* __restrict__ cy, float* __restrict__ cz,
    int32_t* __restrict__ orig, ClusterBound* __restrict__ bounds) {
  // the grid is a multiple of 8 blocks, so this is a bijection of [0, grid)
  const unsigned per = gridDim.x / 8u;
  const unsigned lb = (blockIdx.x & 7u) * per + (blockIdx.x >> 3);
  const int64_t p0 = (int64_t)lb * 256;
  // last q with pbeg[q] <= pos, searched in [lo, nb - 1] (pbeg[lo] <= pos)
  auto last_start = [&](int lo, int64_t pos) {
    int hi = nb - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (pbeg[mid] <= pos) lo = mid;
      else hi = mid - 1;
    }
    return lo;
  };
  // one search for the block's first position; a block rarely spans more
  // than one body, so each thread steps on from there (a few steps, then a
  // search when the block crosses many small bodies)
  __shared__ int q_first;
  if (threadIdx.x == 0) q_first = last_start(0, p0);
  __syncthreads();
  const int64_t p = p0 + threadIdx.x;
  const bool in = p < ptotal;  // ptotal is a multiple of kCluster: whole slices in or out
  bool valid = false;
  float v[3] = {0.0f, 0.0f, 0.0f};
  int64_t dst = 0;
  if (in) {
    int q = q_first;
    for (int k = 0; k < 4 && q + 1 < nb && pbeg[q + 1] <= p; k++) q++;
    if (q + 1 < nb && pbeg[q + 1] <= p) q = last_start(q + 1, p);
    const int b = blist[b0 + q];
    const int i = (int)(p - pbeg[q]);
    valid = i < kbeg[q + 1] - kbeg[q];
    int l = -1;
    if (valid) {
      l = sorted[kbeg[q] + i];
      const int64_t o = off[b];
      v[0] = x[o + l];
      v[1] = y[o + l];
      v[2] = z[o + l];
    }
    dst = slot[b] + i;
    cx[dst] = v[0];
    cy[dst] = v[1];
    cz[dst] = v[2];
    orig[dst] = l;
  }
  // every lane reaches the slice's reduction
  const ClusterBound cb = slice_bound(valid, v);
  if (in && (threadIdx.x & (kCluster - 1)) == 0) bounds[dst / kCluster] = cb;
}


// One wave per supercluster (first slot, slot count <= kSuper*kCluster), four
// per block: sphere and box bounds of its vertices, made as a cluster's are
// (each lane rounds its own FP64 distances up to float before the maximum).
__global__ void __launch_bounds__(256) k_super_bounds(const float* __restrict__ cx,
                                                      const float* __restrict__ cy,
                                                      const float* __restrict__ cz,
                                                      const int32_t* __restrict__ orig,
                                                      const int64_t* __restrict__ first,
                                                      const int32_t* __restrict__ count,
                                                      int64_t n_supers,
                                                      ClusterBound* __restrict__ sbounds) {
  const int64_t q = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (q >= n_supers) return;  // the whole wave: the reductions below need every lane
  const int lane = threadIdx.x & 63;
  const int64_t s0 = first[q];
  const int n = count[q];
  const auto fmin_ = [](float a, float b) { return fminf(a, b); };
  const auto fmax_ = [](float a, float b) { return fmaxf(a, b); };
  float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int i = lane; i < n; i += 64) {
    if (orig[s0 + i] < 0) continue;
    const float p[3] = {cx[s0 + i], cy[s0 + i], cz[s0 + i]};
#pragma unroll
    for (int k = 0; k < 3; k++) {
      lo[k] = fminf(lo[k], p[k]);
      hi[k] = fmaxf(hi[k], p[k]);
    }
  }
  float ctr[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    lo[k] = wave_reduce(lo[k], fmin_);
    hi[k] = wave_reduce(hi[k], fmax_);
    ctr[k] = 0.5f * lo[k] + 0.5f * hi[k];
  }
  double r2 = 0.0, e[3] = {0.0, 0.0, 0.0};
  for (int i = lane; i < n; i += 64) {
    if (orig[s0 + i] < 0) continue;
    const double ex = fabs((double)cx[s0 + i] - (double)ctr[0]);
    const double ey = fabs((double)cy[s0 + i] - (double)ctr[1]);
    const double ez = fabs((double)cz[s0 + i] - (double)ctr[2]);
    r2 = fmax(r2, ex * ex + ey * ey + ez * ez);
    e[0] = fmax(e[0], ex);
    e[1] = fmax(e[1], ey);
    e[2] = fmax(e[2], ez);
  }
  const float rr = wave_reduce(nextafterf((float)sqrt(r2), INFINITY), fmax_);
  float hh[3];
#pragma unroll
  for (int k = 0; k < 3; k++) hh[k] = wave_reduce(nextafterf((float)e[k], INFINITY), fmax_);
  if (lane == 0) {
    ClusterBound cb;
    cb.c = make_float4(ctr[0], ctr[1], ctr[2], rr);
    cb.h = make_float4(hh[0], hh[1], hh[2], 0.0f);
    sbounds[q] = cb;
  }
}

// Each body's first supercluster index goes into the spare word of its first
// cluster's bound (read by the culled support query).
__global__ void __launch_bounds__(256) k_mark_super(const int32_t* __restrict__ cnt,
                                                    const int64_t* __restrict__ slot,
                                                    const int64_t* __restrict__ super,
                                                    int n_bodies, ClusterBound* __restrict__ b) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n_bodies || cnt[i] <= 0 || slot[i] < 0) return;
  b[slot[i] / kCluster].h.w = __builtin_bit_cast(float, (int)super[i]);
}

}  // namespace

// Superclusters a scene of n_slots cluster slots can need: only bodies of
// more than kSuperMinClusters clusters have them (ceil(nc / kSuper) each).
int64_t cluster_supers(int64_t n_slots) {
  return 2 * (n_slots / (kSuper * kCluster)) + 1;
}

int64_t cluster_slots(int64_t n_verts, int32_t n_bodies) {
  return n_verts + (int64_t)(kCluster - 1) * (n_bodies > 0 ? n_bodies : 0);
}

// Builds the clusters of every body of at least min_build vertices; the
// others get body_slot -1 and no slots (queries sweep them whole).  The slot
// layout (a scan of the padded sizes) and the per-class body lists are made
// on the device; only the batched bodies (more than lds_max vertices) come
// back to the host, which cuts them into batches of at most kBatch padded slots
// so the sort's scratch stays bounded.
hipError_t build_clusters(const float* x, const float* y, const float* z, const int64_t* off,
                          const int32_t* cnt, int32_t n_bodies, float* cx, float* cy, float* cz,
                          int32_t* orig, int64_t* body_slot, float* bounds, int64_t n_slots,
                          float* sbounds, int64_t* body_super, int64_t n_supers,
                          int64_t* n_supers_used, int32_t min_build, hipStream_t st) {
  constexpr int64_t kBatch = int64_t(1) << 27;
  if (n_bodies <= 0) return hipSuccess;
  if (n_bodies == 0x7fffffff) return hipErrorInvalidValue;  // n_bodies + 1 scan entries
  min_build = std::max(min_build, 1);
  // Bodies of at most lds_max vertices are built whole by one block each
  // (k_cluster_body, eight size classes); larger ones go through batches of
  // keys, a device segmented sort and a scatter.  OGJK_CLUSTER_LDS_MAX=0 sends
  // every body through the batches (the two paths give identical arrays).
  int lds_max = kLdsBodyMax;
  if (const char* v = getenv("OGJK_CLUSTER_LDS_MAX"))
    lds_max = std::max(0, std::min(atoi(v), kLdsBodyMax));

  uint32_t *keys = nullptr, *keys2 = nullptr;
  int *vals = nullptr, *vals2 = nullptr, *kbeg = nullptr, *pbeg = nullptr, *blist = nullptr,
      *lbig = nullptr;
  unsigned* ccount = nullptr;
  int64_t *psize = nullptr, *pos = nullptr, *dgath = nullptr;
  void* tmp = nullptr;
  size_t tmp_bytes = 0;
  auto cleanup = [&]() {
    (void)hipFree(keys);
    (void)hipFree(keys2);
    (void)hipFree(vals);
    (void)hipFree(vals2);
    (void)hipFree(kbeg);
    (void)hipFree(pbeg);
    (void)hipFree(blist);
    (void)hipFree(lbig);
    (void)hipFree(ccount);
    (void)hipFree(psize);
    (void)hipFree(pos);
    (void)hipFree(dgath);
    (void)hipFree(tmp);
  };
  auto fail_with = [&](hipError_t err) {
    cleanup();
    return err;
  };
  hipError_t e;

  // slot layout and body lists
  const int nb1 = n_bodies + 1;
  size_t scan_bytes = 0;
  if ((e = hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, psize, pos, nb1, st)) != hipSuccess)
    return fail_with(e);
  tmp_bytes = std::max<size_t>(scan_bytes, 1);
  if ((e = hipMalloc(&blist, sizeof(int) * (size_t)(kClasses + 1) * n_bodies)) != hipSuccess ||
      (e = hipMalloc(&ccount, sizeof(unsigned) * (kClasses + 1))) != hipSuccess ||
      (e = hipMalloc(&psize, sizeof(int64_t) * nb1)) != hipSuccess ||
      (e = hipMalloc(&pos, sizeof(int64_t) * nb1)) != hipSuccess ||
      (e = hipMalloc(&tmp, tmp_bytes)) != hipSuccess ||
      (e = hipMemsetAsync(ccount, 0, sizeof(unsigned) * (kClasses + 1), st)) != hipSuccess)
    return fail_with(e);
  hipLaunchKernelGGL(k_body_classes, dim3((unsigned)((nb1 + 255) / 256)), dim3(256), 0, st, cnt,
                     n_bodies, min_build, lds_max, ccount, blist, psize);
  if ((e = hipcub::DeviceScan::ExclusiveSum(tmp, scan_bytes, psize, pos, nb1, st)) != hipSuccess)
    return fail_with(e);
  hipLaunchKernelGGL(k_body_slots, dim3((unsigned)((n_bodies + 255) / 256)), dim3(256), 0, st,
                     psize, pos, n_bodies, body_slot, body_super);
  unsigned hcount[kClasses + 1] = {};
  int64_t total = 0;
  if ((e = hipGetLastError()) != hipSuccess ||
      (e = hipMemcpyAsync(hcount, ccount, sizeof(hcount), hipMemcpyDeviceToHost, st)) !=
          hipSuccess ||
      (e = hipMemcpyAsync(&total, pos + n_bodies, sizeof(int64_t), hipMemcpyDeviceToHost, st)) !=
          hipSuccess ||
      (e = hipStreamSynchronize(st)) != hipSuccess)
    return fail_with(e);
  if (total > n_slots) return fail_with(hipErrorInvalidValue);

  // the one-block bodies
  {
    ClusterBound* cb = (ClusterBound*)bounds;
    auto launch = [&](auto kernel, int k, int block) {
      if (hcount[k] > 0)
        hipLaunchKernelGGL(kernel, dim3(hcount[k]), dim3(block), 0, st, x, y, z, off, cnt,
                           body_slot, blist + (size_t)k * n_bodies, cx, cy, cz, orig, cb);
    };
    launch(k_cluster_body<64, 1>, 0, 64);
    launch(k_cluster_body<64, 4>, 1, 64);
    launch(k_cluster_body<256, 3>, 2, 256);
    launch(k_cluster_body<256, 4>, 3, 256);
    launch(k_cluster_body<256, 6>, 4, 256);
    launch(k_cluster_body<256, 8>, 5, 256);  // 512-thread blocks measured slower
    launch(k_cluster_body<256, 12>, 6, 256);
    launch(k_cluster_body<256, 16>, 7, 256);
    if ((e = hipGetLastError()) != hipSuccess) return fail_with(e);
  }

  // the batched bodies, in index order (the supercluster numbering follows it)
  const int n_large = (int)hcount[kClasses];
  std::vector<int> large(n_large);
  std::vector<int64_t> lcs(2 * (size_t)n_large);  // count, first slot
  if (n_large > 0) {
    if ((e = hipMemcpyAsync(large.data(), blist + (size_t)kClasses * n_bodies,
                            sizeof(int) * n_large, hipMemcpyDeviceToHost, st)) != hipSuccess ||
        (e = hipStreamSynchronize(st)) != hipSuccess)
      return fail_with(e);
    std::sort(large.begin(), large.end());
    if ((e = hipMalloc(&lbig, sizeof(int) * n_large)) != hipSuccess ||
        (e = hipMalloc(&dgath, sizeof(int64_t) * 2 * n_large)) != hipSuccess ||
        (e = hipMemcpyAsync(lbig, large.data(), sizeof(int) * n_large, hipMemcpyHostToDevice,
                            st)) != hipSuccess)
      return fail_with(e);
    hipLaunchKernelGGL(k_gather_large, dim3((unsigned)((n_large + 255) / 256)), dim3(256), 0, st,
                       cnt, body_slot, lbig, n_large, dgath);
    if ((e = hipGetLastError()) != hipSuccess ||
        (e = hipMemcpyAsync(lcs.data(), dgath, sizeof(int64_t) * 2 * n_large,
                            hipMemcpyDeviceToHost, st)) != hipSuccess ||
        (e = hipStreamSynchronize(st)) != hipSuccess)
      return fail_with(e);
    for (int q = 0; q < n_large; q++)
      if (lcs[2 * q] > kBatch) return fail_with(hipErrorInvalidValue);
  }
  // superclusters of the bodies with more than kSuperMinClusters clusters:
  // all batched (kSuperMinClusters * kCluster >= kLdsBodyMax >= lds_max)
  static_assert(kSuperMinClusters * kCluster >= kLdsBodyMax, "superclustered bodies are batched");
  std::vector<int64_t> sfirst, sbody;  // sbody: (body, first supercluster) pairs
  std::vector<int32_t> scount;
  for (int q = 0; q < n_large; q++) {
    const int64_t nc = (lcs[2 * q] + kCluster - 1) / kCluster;
    if (nc <= kSuperMinClusters) continue;
    sbody.push_back(large[q]);
    sbody.push_back((int64_t)sfirst.size());
    for (int64_t c0 = 0; c0 < nc; c0 += kSuper) {
      sfirst.push_back(lcs[2 * q + 1] + c0 * kCluster);
      scount.push_back((int32_t)(std::min<int64_t>(kSuper, nc - c0) * kCluster));
    }
  }
  if ((int64_t)sfirst.size() > n_supers || (int64_t)sfirst.size() > 0x7fffffff)
    return fail_with(hipErrorInvalidValue);
  *n_supers_used = (int64_t)sfirst.size();
  if (!sbody.empty()) {
    // dgath (2 * n_large entries) holds the sbody pairs: one per batched body at most
    const int ns = (int)(sbody.size() / 2);
    if ((e = hipMemcpyAsync(dgath, sbody.data(), sizeof(int64_t) * sbody.size(),
                            hipMemcpyHostToDevice, st)) != hipSuccess)
      return fail_with(e);
    hipLaunchKernelGGL(k_set_super, dim3((unsigned)((ns + 255) / 256)), dim3(256), 0, st, dgath,
                       ns, body_super);
  }

  // batches of consecutive batched bodies (positions in `large`)
  std::vector<int> bstart{0};
  std::vector<int> kbeg_h;  // batch-local key offset of every batched body, batch after batch
  std::vector<int64_t> bverts;
  {
    // a batch holds at most kBatch padded slots (so at most kBatch vertices),
    // or one body alone
    int64_t acc = 0, pacc = 0;  // vertices, padded slots
    for (int q = 0; q < n_large; q++) {
      const int64_t n = lcs[2 * q], pn = (n + kCluster - 1) / kCluster * kCluster;
      if (pacc > 0 && pacc + pn > kBatch) {
        bstart.push_back(q);
        bverts.push_back(acc);
        acc = pacc = 0;
      }
      kbeg_h.push_back((int)acc);
      acc += n;
      pacc += pn;
    }
    bverts.push_back(acc);
    bstart.push_back(n_large);
  }
  int64_t maxv = 0;
  int maxb = 0;
  for (size_t i = 0; i + 1 < bstart.size(); i++) {
    maxv = std::max(maxv, bverts[i]);
    maxb = std::max(maxb, bstart[i + 1] - bstart[i]);
  }
  maxv = std::max<int64_t>(maxv, 1);
  if (!large.empty() &&
      ((e = hipMalloc(&keys, 4 * maxv)) != hipSuccess || (e = hipMalloc(&keys2, 4 * maxv)) ||
       (e = hipMalloc(&vals, 4 * maxv)) || (e = hipMalloc(&vals2, 4 * maxv)) ||
       (e = hipMalloc(&kbeg, sizeof(int) * ((size_t)maxb + 1))) ||
       (e = hipMalloc(&pbeg, sizeof(int) * ((size_t)maxb + 1)))))
    return fail_with(e);
  for (size_t i = 0; i + 1 < bstart.size() && !large.empty(); i++) {
    const int b0 = bstart[i], nb = bstart[i + 1] - b0;
    const int nv = (int)bverts[i];
    if (nb == 0) continue;
    // segment offsets: kbeg[0..nb], kbeg[nb] = nv (end of the last body); and
    // the same over the bodies padded to whole clusters, pbeg
    std::vector<int> seg(kbeg_h.begin() + b0, kbeg_h.begin() + b0 + nb);
    seg.push_back(nv);
    std::vector<int> pseg(nb + 1, 0);
    for (int q = 0; q < nb; q++)
      pseg[q + 1] = pseg[q] + (int)((lcs[2 * (b0 + q)] + kCluster - 1) / kCluster * kCluster);
    if ((e = hipMemcpyAsync(kbeg, seg.data(), sizeof(int) * (nb + 1), hipMemcpyHostToDevice,
                            st)) != hipSuccess ||
        (e = hipMemcpyAsync(pbeg, pseg.data(), sizeof(int) * (nb + 1), hipMemcpyHostToDevice,
                            st)) != hipSuccess) {
      cleanup();
      return e;
    }
    hipLaunchKernelGGL(k_order_keys, dim3(nb), dim3(256), 0, st, x, y, z, off, cnt, lbig, b0, kbeg,
                       keys, vals);
    size_t need = 0;
    e = hipcub::DeviceSegmentedRadixSort::SortPairs(nullptr, need, keys, keys2, vals, vals2, nv,
                                                    nb, kbeg, kbeg + 1, 0, kKeyBits, st);
    if (e != hipSuccess) {
      cleanup();
      return e;
    }
    if (need > tmp_bytes) {
      (void)hipFree(tmp);
      tmp = nullptr;
      if ((e = hipMalloc(&tmp, need)) != hipSuccess) {
        cleanup();
        return e;
      }
      tmp_bytes = need;
    }
    e = hipcub::DeviceSegmentedRadixSort::SortPairs(tmp, need, keys, keys2, vals, vals2, nv, nb,
                                                    kbeg, kbeg + 1, 0, kKeyBits, st);
    if (e != hipSuccess) {
      cleanup();
      return e;
    }
    const int ptotal = pseg[nb];
    const unsigned sblocks = ((unsigned)((ptotal + 255) / 256) + 7u) / 8u * 8u;  // a multiple of 8
    hipLaunchKernelGGL(k_cluster_scatter, dim3(sblocks), dim3(256), 0, st, x, y, z, off, body_slot,
                       lbig, b0, nb, kbeg, pbeg, ptotal, vals2, cx, cy, cz, orig,
                       (ClusterBound*)bounds);
    // the host vectors of this batch die at the end of the iteration
    if ((e = hipStreamSynchronize(st)) != hipSuccess) {
      cleanup();
      return e;
    }
  }
  hipLaunchKernelGGL(k_mark_super, dim3((unsigned)((n_bodies + 255) / 256)), dim3(256), 0, st, cnt,
                     body_slot, body_super, n_bodies, (ClusterBound*)bounds);
  if ((e = hipGetLastError()) != hipSuccess) {
    cleanup();
    return e;
  }
  if (!sfirst.empty()) {
    int64_t* d_first = nullptr;
    int32_t* d_count = nullptr;
    const size_t ns = sfirst.size();
    if ((e = hipMalloc(&d_first, 8 * ns)) == hipSuccess &&
        (e = hipMalloc(&d_count, 4 * ns)) == hipSuccess &&
        (e = hipMemcpyAsync(d_first, sfirst.data(), 8 * ns, hipMemcpyHostToDevice, st)) ==
            hipSuccess &&
        (e = hipMemcpyAsync(d_count, scount.data(), 4 * ns, hipMemcpyHostToDevice, st)) ==
            hipSuccess) {
      hipLaunchKernelGGL(k_super_bounds, dim3((unsigned)((ns + 3) / 4)), dim3(256), 0, st, cx, cy,
                         cz, orig, d_first, d_count, (int64_t)ns, (ClusterBound*)sbounds);
      e = hipGetLastError();
      if (e == hipSuccess) e = hipStreamSynchronize(st);
    }
    (void)hipFree(d_first);
    (void)hipFree(d_count);
    if (e != hipSuccess) {
      cleanup();
      return e;
    }
  }
  e = hipStreamSynchronize(st);
  cleanup();
  return e;
}

}  // namespace ogjk
