// opengjk-amd device core: mixed-precision support sweep, GJK and EPA for one
// pair on a group of LPP lanes.
//
// Precision design
//   * Vertices are stored as FP32.  The O(n) support sweep runs in FP32
//     (explicit FMAs) over an exactly power-of-two-scaled copy of the
//     direction, tracking each lane's best and second-best value.
//   * A rigorous bound on |FP32 dot - FP64 dot| (16 ulp of the direction's
//     max-norm times the body's max L1 vertex norm) defines a candidate
//     window below the FP32 maximum.  Only candidates are re-evaluated in
//     FP64 with the reference's association, so the selected vertex is exactly
//     the one the FP64 reference support selects, ties included.
//   * Everything that is O(1) per iteration (simplex, sub-distance regions,
//     termination, witnesses, EPA faces and horizon) runs in FP64 in the
//     reference's operation order, without FMA contraction.
//   Result: outputs match the FP64 CPU reference bit for bit before the final
//   rounding to the gkFloat ABI.
#pragma once

#include <type_traits>

#include "ogjk_math.hpp"

namespace ogjk {

// ------------------------------------------------------------------ bodies

// Vertex i of a body lives at x[i*STRIDE], y[i*STRIDE], z[i*STRIDE]:
// STRIDE 1 is the SoA scene, STRIDE 3 reads caller AoS coordinates in place.
// T is the storage type: float (the product path) or double (plain FP64).
template <class T, int STRIDE>
struct Body {
  using Elem = T;
  const T* x;
  const T* y;
  const T* z;
  int n;
  float r1;  // >= max_i |x_i|+|y_i|+|z_i| (used by the float sweep only)
  OGJK_DEV T rx(int i) const { return x[(int64_t)i * STRIDE]; }
  OGJK_DEV T ry(int i) const { return y[(int64_t)i * STRIDE]; }
  OGJK_DEV T rz(int i) const { return z[(int64_t)i * STRIDE]; }
  OGJK_DEV double c(int i, int k) const { return (double)(k == 0 ? rx(i) : (k == 1 ? ry(i) : rz(i))); }
  OGJK_DEV D3 d(int i) const { return D3{(double)rx(i), (double)ry(i), (double)rz(i)}; }
};

// Max L1 vertex norm over the group (FP32, rounded up by a relative 2^-22).
template <int LPP, class T, int STRIDE>
OGJK_DEV float body_r1(const Body<T, STRIDE>& b, int lane) {
  float m = 0.0f;
  for (int i = lane; i < b.n; i += LPP) {
    const float s = fabsf((float)b.rx(i)) + fabsf((float)b.ry(i)) + fabsf((float)b.rz(i));
    m = fmaxf(m, s);
  }
  m = group_max_f<LPP>(m);
  return m * (1.0f + 0x1.0p-22f);
}

// --------------------------------------------------------------- support

// Result of a support query: the selected vertex index and its coordinates
// (exact FP64 conversion of the stored values).
struct Sup {
  int i;
  D3 p;
};

// One support query.  has_inc selects the reference's rule:
//   has_inc: GJK (GJK/cpu/openGJK.c:615-639): the incumbent support point
//            keeps its place unless some vertex is strictly better; then the
//            lowest index attaining the maximum wins.
//   !has_inc: EPA (GJK/cpu/EPA.c:310-335): running max starts at -1e10, the
//            lowest index attaining the maximum wins; index -1 if none beats
//            -1e10.
struct Query {
  D3 dir;
  bool has_inc;
  int inc_idx;
  D3 inc_pt;
};

// A lane's FP64-exact candidate after the sweep: value, index, point.
struct Cand {
  double v;
  int i;
  D3 p;
};

// FP32 sweep state of one lane.
struct Sweep32 {
  float dx, dy, dz;
  float best, second, bx, by, bz;
  int bi;
};

// Scale the direction by an exact power of two so its max-norm is in
// [0.5, 1): the FP32 copy then carries a relative error <= 2^-24 per
// component whatever the magnitude of dir (tiny GJK directions included).
OGJK_DEV Sweep32 sweep32_begin(D3 dir) {
  const double mx = fmax(fabs(dir.x), fmax(fabs(dir.y), fabs(dir.z)));
  int e = 0;
  (void)frexp(mx, &e);
  const double sc = ldexp(1.0, -e);
  Sweep32 s;
  s.dx = (float)(dir.x * sc);
  s.dy = (float)(dir.y * sc);
  s.dz = (float)(dir.z * sc);
  s.best = s.second = -INFINITY;
  s.bx = s.by = s.bz = 0.0f;
  s.bi = -1;
  return s;
}

template <int STRIDE>
OGJK_DEV void sweep32_step(const Body<float, STRIDE>& b, int i, Sweep32& s) {
  const float px = b.rx(i), py = b.ry(i), pz = b.rz(i);
  const float d = fmaf(px, s.dx, fmaf(py, s.dy, pz * s.dz));
  s.second = fmaxf(s.second, fminf(s.best, d));
  if (d > s.best) {
    s.best = d;
    s.bi = i;
    s.bx = px;
    s.by = py;
    s.bz = pz;
  }
}

// Candidates within T of the FP32 maximum, re-ranked in FP64 with the
// reference's association.  `multi`: some lane of the group has more than one
// candidate, so every lane rescans its vertices inside the window.
template <int LPP, int STRIDE>
OGJK_DEV Cand candidates32(const Body<float, STRIDE>& b, const Sweep32& s, float T, bool multi,
                           D3 dir, int lane) {
  Cand c{-INFINITY, 0x7fffffff, D3{0.0, 0.0, 0.0}};
  if (!multi) {
    if (s.bi >= 0 && s.best >= T) {
      c.p = D3{(double)s.bx, (double)s.by, (double)s.bz};
      c.v = dot(c.p, dir);
      c.i = s.bi;
    }
  } else {
    for (int i = lane; i < b.n; i += LPP) {
      const float px = b.rx(i), py = b.ry(i), pz = b.rz(i);
      const float d = fmaf(px, s.dx, fmaf(py, s.dy, pz * s.dz));
      if (d >= T) {
        const D3 p{(double)px, (double)py, (double)pz};
        const double v = dot(p, dir);
        if (v > c.v) {  // ascending i within the lane keeps the lowest index
          c.v = v;
          c.i = i;
          c.p = p;
        }
      }
    }
  }
  if (!(c.v == c.v)) c.v = -INFINITY;
  return c;
}

// Group argmax of the lanes' candidates (lowest index on ties), then the
// acceptance rule of the query and the winner's point.  Usual case: one lane
// holds the only candidate and a single parallel shuffle fetches it.
// STRIDED: lane l swept indices l, l+LPP, ..., so the winner's index names
// its lane; otherwise the lane holding the winning index is found by ballot.
template <int LPP, bool STRIDED = true>
OGJK_DEV Sup finish1(Cand c, const Query& q) {
  const uint64_t holders = group_ballot<LPP>(c.v > -INFINITY);
  int src;
  if (__popcll(holders) == 1) {
    src = (int)__builtin_ctzll(holders);
    c.v = shfl_d<LPP>(c.v, src);
    c.i = shfl_i<LPP>(c.i, src);
  } else {
    const int mine = c.i;
    group_argmax_d<LPP>(c.v, c.i);
    if constexpr (STRIDED) {
      src = c.i & (LPP - 1);
    } else {
      src = (int)__builtin_ctzll(group_ballot<LPP>(mine == c.i) | (1ull << 63));
    }
  }
  if (q.has_inc && !(c.v > dot(q.inc_pt, q.dir))) return Sup{q.inc_idx, q.inc_pt};
  if (!q.has_inc && !(c.v > -1e10)) return Sup{-1, D3{0.0, 0.0, 0.0}};
  return Sup{c.i, D3{shfl_d<LPP>(c.p.x, src), shfl_d<LPP>(c.p.y, src), shfl_d<LPP>(c.p.z, src)}};
}

// Support point of one body (FP32 sweep, FP64 re-rank of the window).
template <int LPP, int STRIDE>
OGJK_DEV Sup support(const Body<float, STRIDE>& A, const Query& q, int lane) {
  Sweep32 sw = sweep32_begin(q.dir);
  for (int i = lane; i < A.n; i += LPP) sweep32_step(A, i, sw);
  const float m = group_max_f<LPP>(sw.best);
  // |D32 - scale*D64| <= 16u * r1 (u = 2^-24): candidates lie within 2x that.
  const float t = m - 0x1.0p-19f * A.r1;
  const bool multi = group_any<LPP>(sw.second >= t);
  return finish1<LPP>(candidates32<LPP>(A, sw, t, multi, q.dir, lane), q);
}

// ------------------------------------------------ clustered support sweep
//
// A scene body may carry a second copy of its vertices ordered by direction
// from the body's centre (Hilbert order of the octahedral map), cut into
// clusters of kCluster slots, each with a bounding sphere (float centre c,
// radius r >= max |x - c|) and box (half extents h about c).  A support
// query then
//   1. bounds every cluster: ub = c.d + min(r|d|, h.|d|) >= max of x.d,
//   2. sweeps the cluster with the largest ub: m0 = its FP32 maximum,
//   3. sweeps only the clusters with ub >= m0 - 2^-15 r1.
// Rounding: |fl(ub) - ub| <= 2^-17 r1 and |D32 - D| <= 2^-20 r1 (d scaled to
// max-norm in [0.5,1), |c|_1 <= 3 r1, |h|_1 <= 3 r1, r <= 2 r1), so every skipped vertex
// lies more than 2^-16 r1 below the maximum: it can neither be the FP64
// argmax nor tie with it, and the selection is the full sweep's.  Indices
// reported are the body-local input indices (the copy stores them).
//
// Bodies of more than kSuper clusters also carry superclusters (kSuper
// consecutive clusters, bounded the same way from their vertices): the query
// bounds the superclusters first and only opens those whose bound reaches
// the threshold, so a 100,000-vertex body reads ~100 supercluster bounds per
// query instead of ~3,000 cluster bounds.  A supercluster's bound covers
// every vertex of its clusters, so skipping it is exact by the same argument.
constexpr int kCluster = 32;
constexpr int kSuper = 32;
constexpr int kSuperMinClusters = 4 * kSuper;  // superclusters above 4,096 vertices

// Bounds of one cluster: centre c and radius r (c.w) of a sphere, and half
// extents h of the box around c, all rounded outwards.
struct ClusterBound {
  float4 c;
  float4 h;
};

struct ClusterView {
  const float* x;  // cluster-ordered vertices of every body
  const float* y;
  const float* z;
  const int32_t* orig;          // body-local input index of each slot
  const ClusterBound* bounds;   // per cluster
  const ClusterBound* sbounds;  // per supercluster
};

struct CBody : Body<float, 1> {
  ClusterView cv;
  int64_t slot;  // first slot of the body (a multiple of kCluster); < 0: no culling
};

// A clustered body of a scene that has superclusters (two-level queries).
struct CBody2 : CBody {};

// The body's first supercluster (< 0: none) is kept in the spare word of its
// first cluster's bound, so a body carries no extra field through GJK.
OGJK_DEV int first_super(const ClusterBound* body_bounds) {
  return __builtin_bit_cast(int, body_bounds[0].h.w);
}

// c.d + min(r |d|, h.|d|) >= max over the cluster of x.d
OGJK_DEV float cluster_ub(const ClusterBound& b, float dx, float dy, float dz, float dn) {
  const float cd = fmaf(b.c.x, dx, fmaf(b.c.y, dy, b.c.z * dz));
  const float box = fmaf(b.h.x, fabsf(dx), fmaf(b.h.y, fabsf(dy), b.h.z * fabsf(dz)));
  return cd + fminf(b.c.w * dn, box);
}

// Calls f(k) for every body-local slot k of the clusters c0 <= c < c0 + ncl
// whose bound reaches L (cb: the body's cluster bounds), each slot on exactly
// one lane, in the same order on every call.
template <int LPP, class F>
OGJK_DEV void visit_range(const CBody& A, const ClusterBound* cb, int c0, int ncl, float dx,
                          float dy, float dz, float dn, float L, int lane, F&& f) {
  for (int base = 0; base < ncl; base += LPP) {
    const int c = base + lane;
    const bool live = c < ncl && !(cluster_ub(cb[c0 + c], dx, dy, dz, dn) < L);
    uint64_t m = group_ballot<LPP>(live);
    while (m != 0) {
      if constexpr (LPP >= kCluster) {
        // LPP/kCluster clusters per round, one per kCluster-lane slice
        constexpr int PER = LPP / kCluster;
        int cid = -1;
#pragma unroll
        for (int j = 0; j < PER; j++) {
          if (m != 0) {
            if (j == lane / kCluster) cid = (int)__builtin_ctzll(m);
            m &= m - 1;
          }
        }
        if (cid >= 0) {
          const int k = (c0 + base + cid) * kCluster + (lane & (kCluster - 1));
          if (k < A.n) f(k);
        }
      } else {
        const int cid = (int)__builtin_ctzll(m);
        m &= m - 1;
#pragma unroll
        for (int t = lane; t < kCluster; t += LPP) {
          const int k = (c0 + base + cid) * kCluster + t;
          if (k < A.n) f(k);
        }
      }
    }
  }
}

// All surviving clusters of the body: through its superclusters when it has
// them (sb non-null), in supercluster then cluster order; without, the body
// is a single unit of nc clusters.  One loop for both keeps one copy of the
// cluster visit per call site.
template <int LPP, bool TWO, class F>
OGJK_DEV void visit_clusters(const CBody& A, const ClusterBound* cb, int nc,
                             const ClusterBound* sb, float dx, float dy, float dz, float dn,
                             float L, int lane, F&& f) {
  if constexpr (!TWO) {
    visit_range<LPP>(A, cb, 0, nc, dx, dy, dz, dn, L, lane, f);
    return;
  }
  const int nunits = sb != nullptr ? (nc + kSuper - 1) / kSuper : 1;
  for (int base = 0; base < nunits; base += LPP) {
    const int u = base + lane;
    const bool live = u < nunits && (sb == nullptr || !(cluster_ub(sb[u], dx, dy, dz, dn) < L));
    uint64_t m = group_ballot<LPP>(live);
    while (m != 0) {
      const int u0 = base + (int)__builtin_ctzll(m);
      m &= m - 1;
      const int c0 = sb != nullptr ? u0 * kSuper : 0;
      const int ncl = sb != nullptr ? min(kSuper, nc - c0) : nc;
      visit_range<LPP>(A, cb, c0, ncl, dx, dy, dz, dn, L, lane, f);
    }
  }
}

// Index (within [0, n)) of the entry of b with the largest bound.
template <int LPP>
OGJK_DEV int best_bound(const ClusterBound* b, int n, float dx, float dy, float dz, float dn,
                        int lane) {
  float bub = -INFINITY;
  int bi = 0;
  for (int c = lane; c < n; c += LPP) {
    const float ub = cluster_ub(b[c], dx, dy, dz, dn);
    if (ub > bub) {
      bub = ub;
      bi = c;
    }
  }
  const float gub = group_max_f<LPP>(bub);
  return shfl_i<LPP>(bi, (int)__builtin_ctzll(group_ballot<LPP>(bub == gub) | (1ull << 63)));
}

// The clustered query through superclusters (TWO) or, without TWO, as one
// level (used by groups of fewer than 64 lanes; see support(CBody) below).
template <int LPP, bool TWO>
OGJK_DEV Sup support_clustered(const CBody& A, const Query& q, int lane) {
  if (A.slot < 0) return support<LPP>(static_cast<const Body<float, 1>&>(A), q, lane);
  Sweep32 sw = sweep32_begin(q.dir);
  const float dx = sw.dx, dy = sw.dy, dz = sw.dz;
  const float dn = sqrtf(fmaf(dx, dx, fmaf(dy, dy, dz * dz))) * (1.0f + 0x1.0p-20f);
  const int nc = (A.n + kCluster - 1) / kCluster;
  const ClusterBound* cb = A.cv.bounds + A.slot / kCluster;
  const float* sx = A.cv.x + A.slot;
  const float* sy = A.cv.y + A.slot;
  const float* sz = A.cv.z + A.slot;

  const ClusterBound* sb = nullptr;
  if constexpr (TWO) {
    const int sup = first_super(cb);
    if (sup >= 0) sb = A.cv.sbounds + sup;
  }

  // 1-2: the cluster with the largest bound (inside the supercluster with the
  // largest bound, when there are superclusters) gives a lower bound of the max
  int c0 = 0, ncl = nc;
  if (TWO && sb != nullptr) {
    c0 = best_bound<LPP>(sb, (nc + kSuper - 1) / kSuper, dx, dy, dz, dn, lane) * kSuper;
    ncl = min(kSuper, nc - c0);
  }
  const int bc = c0 + best_bound<LPP>(cb + c0, ncl, dx, dy, dz, dn, lane);
  float m0 = -INFINITY;
#pragma unroll
  for (int t = lane; t < kCluster; t += LPP) {
    const int k = bc * kCluster + t;
    if (k < A.n) m0 = fmaxf(m0, fmaf(sx[k], dx, fmaf(sy[k], dy, sz[k] * dz)));
  }
  const float L = group_max_f<LPP>(m0) - 0x1.0p-15f * A.r1;

  // 3: sweep the surviving clusters
  visit_clusters<LPP, TWO>(A, cb, nc, sb, dx, dy, dz, dn, L, lane, [&](int k) {
    const float px = sx[k], py = sy[k], pz = sz[k];
    const float d = fmaf(px, dx, fmaf(py, dy, pz * dz));
    sw.second = fmaxf(sw.second, fminf(sw.best, d));
    if (d > sw.best) {
      sw.best = d;
      sw.bi = k;
      sw.bx = px;
      sw.by = py;
      sw.bz = pz;
    }
  });
  const float M = group_max_f<LPP>(sw.best);
  const float T = M - 0x1.0p-19f * A.r1;
  const bool multi = group_any<LPP>(sw.second >= T);

  // FP64 re-rank of the window; ties go to the lowest input index
  const int32_t* orig = A.cv.orig + A.slot;
  Cand c{-INFINITY, 0x7fffffff, D3{0.0, 0.0, 0.0}};
  if (!multi) {
    if (sw.bi >= 0 && sw.best >= T) {
      c.p = D3{(double)sw.bx, (double)sw.by, (double)sw.bz};
      c.v = dot(c.p, q.dir);
      c.i = orig[sw.bi];
    }
  } else {
    visit_clusters<LPP, TWO>(A, cb, nc, sb, dx, dy, dz, dn, L, lane, [&](int k) {
      const float px = sx[k], py = sy[k], pz = sz[k];
      const float d = fmaf(px, dx, fmaf(py, dy, pz * dz));
      if (d >= T) {
        const D3 p{(double)px, (double)py, (double)pz};
        const double v = dot(p, q.dir);
        const int oi = orig[k];
        if (v > c.v || (v == c.v && oi < c.i)) {
          c.v = v;
          c.i = oi;
          c.p = p;
        }
      }
    });
  }
  if (!(c.v == c.v)) c.v = -INFINITY;
  return finish1<LPP, false>(c, q);
}

// Calls f(k) for every slot of the clusters base + b for the set bits b of
// m (a group ballot), each slot on exactly one lane.
template <int LPP, class F>
OGJK_DEV void visit_mask(const CBody& A, int base, uint64_t m, int lane, F&& f) {
  while (m != 0) {
    if constexpr (LPP >= kCluster) {
      constexpr int PER = LPP / kCluster;
      int cid = -1;
#pragma unroll
      for (int j = 0; j < PER; j++) {
        if (m != 0) {
          if (j == lane / kCluster) cid = (int)__builtin_ctzll(m);
          m &= m - 1;
        }
      }
      if (cid >= 0) {
        const int k = (base + cid) * kCluster + (lane & (kCluster - 1));
        if (k < A.n) f(k);
      }
    } else {
      const int cid = (int)__builtin_ctzll(m);
      m &= m - 1;
#pragma unroll
      for (int t = lane; t < kCluster; t += LPP) {
        const int k = (base + cid) * kCluster + t;
        if (k < A.n) f(k);
      }
    }
  }
}

// One-level clustered query, for scenes without superclusters (the cluster
// build gives superclusters to every body of more than kSuperMinClusters
// clusters, so here a body has at most that many: K per lane).  Same
// selection as support_clustered<LPP, false>, with fewer dependent memory
// round trips: each lane keeps the bounds of its clusters in registers from
// the first pass (the survivor test reads no bounds again), the best
// cluster's sweep is the first step of the window sweep (it is not swept
// twice), and each lane's input index is fetched before the final
// reductions instead of after them.
template <int LPP>
OGJK_DEV Sup support_clustered1(const CBody& A, const Query& q, int lane) {
  if (A.slot < 0) return support<LPP>(static_cast<const Body<float, 1>&>(A), q, lane);
  constexpr int K = (kSuperMinClusters + LPP - 1) / LPP;
  Sweep32 sw = sweep32_begin(q.dir);
  const float dx = sw.dx, dy = sw.dy, dz = sw.dz;
  const float dn = sqrtf(fmaf(dx, dx, fmaf(dy, dy, dz * dz))) * (1.0f + 0x1.0p-20f);
  const int nc = (A.n + kCluster - 1) / kCluster;
  const ClusterBound* cb = A.cv.bounds + A.slot / kCluster;
  const float* sx = A.cv.x + A.slot;
  const float* sy = A.cv.y + A.slot;
  const float* sz = A.cv.z + A.slot;

  // 1: every cluster's bound; the largest (lowest cluster on ties) is swept
  // first.  Bounds past K per lane are recomputed in step 3.
  float ub[K];
  float bub = -INFINITY;
  int bi = 0;
#pragma unroll
  for (int j = 0; j < K; j++) {
    const int c = lane + j * LPP;
    ub[j] = c < nc ? cluster_ub(cb[c], dx, dy, dz, dn) : -INFINITY;
    if (ub[j] > bub) {
      bub = ub[j];
      bi = c;
    }
  }
  for (int c = lane + K * LPP; c < nc; c += LPP) {
    const float u = cluster_ub(cb[c], dx, dy, dz, dn);
    if (u > bub) {
      bub = u;
      bi = c;
    }
  }
  const float gub = group_max_f<LPP>(bub);
  const int bc = shfl_i<LPP>(bi, (int)__builtin_ctzll(group_ballot<LPP>(bub == gub) | (1ull << 63)));
  // A group of two clusters' width sweeps the cluster with the second largest
  // bound in step 2 as well, on its upper half (idle otherwise): the threshold
  // is then the maximum over both clusters (it can only rise), and step 3 has
  // one cluster fewer to visit.
  constexpr bool kPair = LPP == 2 * kCluster;
  int bc2 = -1;
  if constexpr (kPair) {
    float sub = -INFINITY;
    int si = 0;
#pragma unroll
    for (int j = 0; j < K; j++) {
      const int c = lane + j * LPP;
      if (c != bc && ub[j] > sub) {
        sub = ub[j];
        si = c;
      }
    }
    for (int c = lane + K * LPP; c < nc; c += LPP) {
      const float u = c != bc ? cluster_ub(cb[c], dx, dy, dz, dn) : -INFINITY;
      if (u > sub) {
        sub = u;
        si = c;
      }
    }
    const float g2 = group_max_f<LPP>(sub);
    if (g2 > -INFINITY)
      bc2 = shfl_i<LPP>(si, (int)__builtin_ctzll(group_ballot<LPP>(sub == g2) | (1ull << 63)));
  }
  // the slot of cluster bc (lower half) or bc2 (upper half) this lane sweeps
  const int k2 = kPair ? ((lane < kCluster ? bc : bc2) * kCluster + (lane & (kCluster - 1))) : -1;
  const bool has2 = kPair && (lane < kCluster || bc2 >= 0) && k2 < A.n;

  auto sweep = [&](int k) {
    const float px = sx[k], py = sy[k], pz = sz[k];
    const float d = fmaf(px, dx, fmaf(py, dy, pz * dz));
    sw.second = fmaxf(sw.second, fminf(sw.best, d));
    if (d > sw.best) {
      sw.best = d;
      sw.bi = k;
      sw.bx = px;
      sw.by = py;
      sw.bz = pz;
    }
  };
  // 2: the best cluster (and the second, see above): the FP32 maximum m0 of
  // what was swept gives the threshold L
  if constexpr (kPair) {
    if (has2) sweep(k2);
  } else {
#pragma unroll
    for (int t = lane; t < kCluster; t += LPP) {
      const int k = bc * kCluster + t;
      if (k < A.n) sweep(k);
    }
  }
  const float L = group_max_f<LPP>(sw.best) - 0x1.0p-15f * A.r1;

  // 3: the other clusters whose bound reaches L (the visit order does not
  // change the selection: ties are broken by input index below)
  auto others = [&](auto&& f) {
#pragma unroll
    for (int j = 0; j < K; j++) {
      const int c = lane + j * LPP;
      visit_mask<LPP>(A, j * LPP,
                      group_ballot<LPP>(c < nc && c != bc && c != bc2 && !(ub[j] < L)), lane, f);
    }
    for (int base = K * LPP; base < nc; base += LPP) {
      const int c = base + lane;
      const bool live =
          c < nc && c != bc && c != bc2 && !(cluster_ub(cb[c], dx, dy, dz, dn) < L);
      visit_mask<LPP>(A, base, group_ballot<LPP>(live), lane, f);
    }
  };
  others(sweep);

  const int32_t* orig = A.cv.orig + A.slot;
  const int oi = sw.bi >= 0 ? orig[sw.bi] : 0x7fffffff;
  const float M = group_max_f<LPP>(sw.best);
  const float T = M - 0x1.0p-19f * A.r1;
  const bool multi = group_any<LPP>(sw.second >= T);

  // FP64 re-rank of the window; ties go to the lowest input index
  Cand c{-INFINITY, 0x7fffffff, D3{0.0, 0.0, 0.0}};
  if (!multi) {
    if (sw.bi >= 0 && sw.best >= T) {
      c.p = D3{(double)sw.bx, (double)sw.by, (double)sw.bz};
      c.v = dot(c.p, q.dir);
      c.i = oi;
    }
  } else {
    auto rank = [&](int k) {
      const float px = sx[k], py = sy[k], pz = sz[k];
      const float d = fmaf(px, dx, fmaf(py, dy, pz * dz));
      if (d >= T) {
        const D3 p{(double)px, (double)py, (double)pz};
        const double v = dot(p, q.dir);
        const int o = orig[k];
        if (v > c.v || (v == c.v && o < c.i)) {
          c.v = v;
          c.i = o;
          c.p = p;
        }
      }
    };
    if constexpr (kPair) {
      if (has2) rank(k2);
    } else {
#pragma unroll
      for (int t = lane; t < kCluster; t += LPP) {
        const int k = bc * kCluster + t;
        if (k < A.n) rank(k);
      }
    }
    others(rank);
  }
  if (!(c.v == c.v)) c.v = -INFINITY;
  return finish1<LPP, false>(c, q);
}

// The one-level query with cached bounds for whole-wave groups; narrower
// groups would keep 128 / LPP bounds per lane, which costs the small-hull
// kernels of the same scene more in spills than it saves.
template <int LPP>
OGJK_DEV Sup support(const CBody& A, const Query& q, int lane) {
  if constexpr (LPP >= 64) {
    return support_clustered1<LPP>(A, q, lane);
  } else {
    return support_clustered<LPP, false>(A, q, lane);
  }
}
template <int LPP>
OGJK_DEV Sup support(const CBody2& A, const Query& q, int lane) {
  return support_clustered<LPP, true>(A, q, lane);
}

// ------------------------------------------------------------ posed bodies
//
// A body placed by a rigid pose: world vertex = R x + t, R from a unit
// quaternion (w, x, y, z) and t stored as 7 floats.  World coordinates are
// computed in FP64, each component as ((R0 x + R1 y) + R2 z) + t, the same
// arithmetic the checker uses to build its world vertices.  A support query
// sweeps the FP32 local vertices along R^T d and re-ranks its candidates on
// their FP64 world coordinates, so the selection is that of a sweep over the
// FP64 world vertices.

struct Pose {
  double r[9];
  double t[3];
  OGJK_DEV static Pose load(const float* p) {
    const double w = p[0], x = p[1], y = p[2], z = p[3];
    Pose P;
    P.r[0] = 1.0 - 2.0 * (y * y + z * z);
    P.r[1] = 2.0 * (x * y - w * z);
    P.r[2] = 2.0 * (x * z + w * y);
    P.r[3] = 2.0 * (x * y + w * z);
    P.r[4] = 1.0 - 2.0 * (x * x + z * z);
    P.r[5] = 2.0 * (y * z - w * x);
    P.r[6] = 2.0 * (x * z - w * y);
    P.r[7] = 2.0 * (y * z + w * x);
    P.r[8] = 1.0 - 2.0 * (x * x + y * y);
    P.t[0] = p[4];
    P.t[1] = p[5];
    P.t[2] = p[6];
    return P;
  }
  OGJK_DEV D3 world(double x, double y, double z) const {
    return D3{((r[0] * x + r[1] * y) + r[2] * z) + t[0], ((r[3] * x + r[4] * y) + r[5] * z) + t[1],
              ((r[6] * x + r[7] * y) + r[8] * z) + t[2]};
  }
  OGJK_DEV D3 to_local(D3 d) const {  // R^T d
    return D3{(r[0] * d.x + r[3] * d.y) + r[6] * d.z, (r[1] * d.x + r[4] * d.y) + r[7] * d.z,
              (r[2] * d.x + r[5] * d.y) + r[8] * d.z};
  }
};

struct PBody : Body<float, 1> {
  const float* pose;  // 7 floats: quaternion w, x, y, z, then translation
  OGJK_DEV D3 d(int i) const { return Pose::load(pose).world(rx(i), ry(i), rz(i)); }
  OGJK_DEV double c(int i, int k) const {
    const D3 p = d(i);
    return k == 0 ? p.x : (k == 1 ? p.y : p.z);
  }
};

// Local FP32 sweep along R^T d.  Besides the FP32 error (16u r1 in scaled
// units), each candidate's FP64 world value may differ from the real
// (R x + t).d by at most u64 (45 r1 + 15 |t|_1) scaled, < 2^-46 (r1 + |t|_1):
// the window widens by twice that.
template <int LPP>
OGJK_DEV Sup support(const PBody& A, const Query& q, int lane) {
  const Pose P = Pose::load(A.pose);
  Sweep32 sw = sweep32_begin(P.to_local(q.dir));
  for (int i = lane; i < A.n; i += LPP) sweep32_step(A, i, sw);
  const float m = group_max_f<LPP>(sw.best);
  const float t1 = (float)(fabs(P.t[0]) + fabs(P.t[1]) + fabs(P.t[2]));
  const float T = m - (0x1.0p-19f * A.r1 + 0x1.0p-45f * (A.r1 + t1));
  const bool multi = group_any<LPP>(sw.second >= T);
  Cand c{-INFINITY, 0x7fffffff, D3{0.0, 0.0, 0.0}};
  if (!multi) {
    if (sw.bi >= 0 && sw.best >= T) {
      c.p = P.world(sw.bx, sw.by, sw.bz);
      c.v = dot(c.p, q.dir);
      c.i = sw.bi;
    }
  } else {
    for (int i = lane; i < A.n; i += LPP) {
      const float px = A.rx(i), py = A.ry(i), pz = A.rz(i);
      const float d = fmaf(px, sw.dx, fmaf(py, sw.dy, pz * sw.dz));
      if (d >= T) {
        const D3 p = P.world(px, py, pz);
        const double v = dot(p, q.dir);
        if (v > c.v) {  // ascending i within the lane keeps the lowest index
          c.v = v;
          c.i = i;
          c.p = p;
        }
      }
    }
  }
  if (!(c.v == c.v)) c.v = -INFINITY;
  return finish1<LPP>(c, q);
}

// Field-wise value selects (a select between two structs' addresses would
// put them in scratch memory).
template <class T, int STRIDE>
OGJK_DEV Body<T, STRIDE> pick(bool c, const Body<T, STRIDE>& a, const Body<T, STRIDE>& b) {
  return Body<T, STRIDE>{c ? a.x : b.x, c ? a.y : b.y, c ? a.z : b.z, c ? a.n : b.n,
                         c ? a.r1 : b.r1};
}
OGJK_DEV CBody pick(bool c, const CBody& a, const CBody& b) {
  return CBody{pick(c, static_cast<const Body<float, 1>&>(a), static_cast<const Body<float, 1>&>(b)),
               a.cv, c ? a.slot : b.slot};
}
OGJK_DEV CBody2 pick(bool c, const CBody2& a, const CBody2& b) {
  return CBody2{pick(c, static_cast<const CBody&>(a), static_cast<const CBody&>(b))};
}
OGJK_DEV PBody pick(bool c, const PBody& a, const PBody& b) {
  return PBody{pick(c, static_cast<const Body<float, 1>&>(a), static_cast<const Body<float, 1>&>(b)),
               c ? a.pose : b.pose};
}
OGJK_DEV Query pick(bool c, const Query& a, const Query& b) {
  return Query{sel(c, a.dir, b.dir), c ? a.has_inc : b.has_inc, c ? a.inc_idx : b.inc_idx,
               sel(c, a.inc_pt, b.inc_pt)};
}

// The support points of both bodies, one query after the other through one
// copy of the sweep code (two sweeps in one loop keep two sweep states live;
// two inlined copies double the code the register allocator sees).
template <int LPP>
OGJK_DEV Sup swap_halves(const Sup& r) {
  Sup o;
  o.i = xchg<LPP / 2>(r.i);
  o.p = D3{xchg<LPP / 2>(r.p.x), xchg<LPP / 2>(r.p.y), xchg<LPP / 2>(r.p.z)};
  return o;
}

// Split the group between the two queries for bodies swept whole (small
// hulls: the reduction chains dominate); clustered bodies keep every lane on
// one query at a time (large hulls: the sweep dominates; splitting was 16%
// slower on the uneven 1M-pair workload, 8% faster on 64-vertex pairs).
template <class BT>
constexpr bool kSplitQueries = !std::is_base_of_v<CBody, BT>;
// Posed bodies in a 64-lane group (the wave) also query one after the other:
// each query's pose is then wave-uniform and stays in scalar registers (the
// posed GJK kernel's scratch 60 -> 12 B/lane; 10,000 posed pairs 0.186 ->
// 0.181 ms).
template <int LPP, class BT>
constexpr bool kSplit = kSplitQueries<BT> && !(LPP == 64 && std::is_same_v<BT, PBody>);

template <int LPP, class BT>
OGJK_DEV void support_pair(const BT& A, const Query& qa, const BT& B, const Query& qb, int lane,
                           Sup& ra, Sup& rb) {
  if constexpr (LPP >= 8 && kSplit<LPP, BT>) {
    // the lower half of the group answers A's query, the upper half B's, at
    // the same time; then each half takes the other's result
    const bool hi = lane >= LPP / 2;
    const Sup r = support<LPP / 2>(pick(hi, B, A), pick(hi, qb, qa), lane & (LPP / 2 - 1));
    // each half's result is uniform over the half, so any lane of the other
    // half (the exchange partner) holds it
    const Sup o = swap_halves<LPP>(r);
    ra = Sup{hi ? o.i : r.i, sel(hi, o.p, r.p)};
    rb = Sup{hi ? r.i : o.i, sel(hi, r.p, o.p)};
    return;
  }
#pragma clang loop unroll(disable)
  for (int w = 0; w < 2; w++) {
    const bool second = w != 0;
    const Sup r = support<LPP>(pick(second, B, A), pick(second, qb, qa), lane);
    if (second) {
      rb = r;
    } else {
      ra = r;
    }
  }
}

// Plain FP64 path over double storage: the same selection rules computed
// directly in FP64 (the comparison baseline; no FP32 pass).
template <int LPP, int STRIDE>
OGJK_DEV void support_pair(const Body<double, STRIDE>& A, const Query& qa,
                           const Body<double, STRIDE>& B, const Query& qb, int lane, Sup& ra,
                           Sup& rb) {
  Cand ca{-INFINITY, 0x7fffffff, D3{0.0, 0.0, 0.0}}, cb = ca;
  const int nmax = A.n > B.n ? A.n : B.n;
  for (int i = lane; i < nmax; i += LPP) {
    if (i < A.n) {
      const D3 p = A.d(i);
      const double v = dot(p, qa.dir);
      if (v > ca.v) ca = Cand{v, i, p};
    }
    if (i < B.n) {
      const D3 p = B.d(i);
      const double v = dot(p, qb.dir);
      if (v > cb.v) cb = Cand{v, i, p};
    }
  }
  if (!(ca.v == ca.v)) ca.v = -INFINITY;
  if (!(cb.v == cb.v)) cb.v = -INFINITY;
  ra = finish1<LPP>(ca, qa);
  rb = finish1<LPP>(cb, qb);
}

// --------------------------------------------------------------- simplex

struct Simplex {
  int n;
  D3 p[4];
  int ia[4], ib[4];
};

// Runtime-indexed access to the simplex.  LLVM would turn a select between
// two loads of the simplex into a load from a selected address, which forces
// the whole simplex into scratch memory; reading every slot and passing the
// values through an empty asm statement keeps the selects on registers.
OGJK_DEV double keep_reg(double v) {
  asm volatile("" : "+v"(v));
  return v;
}
OGJK_DEV int keep_reg(int v) {
  asm volatile("" : "+v"(v));
  return v;
}
OGJK_DEV D3 keep_reg(D3 v) { return D3{keep_reg(v.x), keep_reg(v.y), keep_reg(v.z)}; }

OGJK_DEV D3 sp(const Simplex& s, int k) {
  const D3 p0 = keep_reg(s.p[0]), p1 = keep_reg(s.p[1]), p2 = keep_reg(s.p[2]),
           p3 = keep_reg(s.p[3]);
  return sel(k == 0, p0, sel(k == 1, p1, sel(k == 2, p2, p3)));
}
OGJK_DEV int sa(const Simplex& s, int k) {
  const int a0 = keep_reg(s.ia[0]), a1 = keep_reg(s.ia[1]), a2 = keep_reg(s.ia[2]),
            a3 = keep_reg(s.ia[3]);
  return k == 0 ? a0 : (k == 1 ? a1 : (k == 2 ? a2 : a3));
}
OGJK_DEV int sb(const Simplex& s, int k) {
  const int b0 = keep_reg(s.ib[0]), b1 = keep_reg(s.ib[1]), b2 = keep_reg(s.ib[2]),
            b3 = keep_reg(s.ib[3]);
  return k == 0 ? b0 : (k == 1 ? b1 : (k == 2 ? b2 : b3));
}
// Write slot k: every slot is rewritten with a value select (no address select).
OGJK_DEV void sset(Simplex& s, int k, D3 p, int a, int b) {
#pragma unroll
  for (int t = 0; t < 4; t++) {
    const bool h = (k == t);
    s.p[t] = sel(h, p, keep_reg(s.p[t]));
    s.ia[t] = h ? a : keep_reg(s.ia[t]);
    s.ib[t] = h ? b : keep_reg(s.ib[t]);
  }
}
template <int DST, int SRC>
OGJK_DEV void smove(Simplex& s) {
  s.p[DST] = s.p[SRC];
  s.ia[DST] = s.ia[SRC];
  s.ib[DST] = s.ib[SRC];
}

struct Pt {  // a simplex vertex with its pair of body indices
  D3 p;
  int a, b;
};
OGJK_DEV Pt spt(const Simplex& s, int k) { return Pt{sp(s, k), sa(s, k), sb(s, k)}; }

// select_1ik / 1ij / 1jk (GJK/cpu/openGJK.c:53-78): {newest, a, b}
OGJK_DEV void keep3(Simplex& s, const Pt& a, const Pt& b) {
  s.n = 3;
  smove<2, 3>(s);
  s.p[1] = a.p; s.ia[1] = a.a; s.ib[1] = a.b;
  s.p[0] = b.p; s.ia[0] = b.a; s.ib[0] = b.b;
}
// select_1i / 1j / 1k (GJK/cpu/openGJK.c:80-99): {newest, a}
OGJK_DEV void keep2(Simplex& s, const Pt& a) {
  s.n = 2;
  smove<1, 3>(s);
  s.p[0] = a.p; s.ia[0] = a.a; s.ib[0] = a.b;
}

// ------------------------------------------------------- predicates (FP64)

// determinant: GJK/cpu/openGJK.c:167-173
OGJK_DEV double det3(D3 p, D3 q, D3 r) {
  return p.x * ((q.y * r.z) - (r.y * q.z)) - p.y * (q.x * r.z - r.x * q.z) +
         p.z * (q.x * r.y - r.x * q.y);
}
// projectOnLine: GJK/cpu/openGJK.c:183-196
OGJK_DEV D3 proj_line(D3 p, D3 q) {
  const D3 pq = sub(p, q);
  const double t = dot(p, pq) / dot(pq, pq);
  return D3{p.x - pq.x * t, p.y - pq.y * t, p.z - pq.z * t};
}
// projectOnPlane: GJK/cpu/openGJK.c:198-217
OGJK_DEV D3 proj_plane(D3 p, D3 q, D3 r) {
  const D3 n = cross(sub(p, q), sub(p, r));
  const double t = dot(n, p) / dot(n, n);
  return D3{n.x * t, n.y * t, n.z * t};
}
// hff1: GJK/cpu/openGJK.c:219-230 (sum starts at 0, as in the reference)
OGJK_DEV int hff1(D3 p, D3 q) {
  double t = 0;
  t += (p.x * p.x - p.x * q.x);
  t += (p.y * p.y - p.y * q.y);
  t += (p.z * p.z - p.z * q.z);
  return t > 0;
}
// hff2: GJK/cpu/openGJK.c:232-248
OGJK_DEV int hff2(D3 p, D3 q, D3 r) {
  const D3 pq = sub(q, p);
  const D3 nt = cross(pq, sub(r, p));
  const D3 n = cross(pq, nt);
  return dot(p, n) < 0;
}
// hff3: GJK/cpu/openGJK.c:250-255
OGJK_DEV int hff3(D3 p, D3 q, D3 r) { return dot(p, cross(q, r)) <= 0; }

// ------------------------------------------------------- S1D / S2D / S3D

// S1D: GJK/cpu/openGJK.c:257-268
OGJK_DEV void S1D(Simplex& s, D3& v) {
  if (hff1(s.p[1], s.p[0])) {
    v = proj_line(s.p[1], s.p[0]);
  } else {
    v = s.p[1];
    s.n = 1;
    smove<0, 1>(s);
  }
}

// S2D: GJK/cpu/openGJK.c:270-313
OGJK_DEV void S2D(Simplex& s, D3& v) {
  const D3 s1 = s.p[2], s2 = s.p[1], s3 = s.p[0];
  const int h12 = hff1(s1, s2);
  const int h13 = hff1(s1, s3);
  if (h12) {
    if (!hff2(s1, s2, s3)) {
      if (h13) {
        if (!hff2(s1, s3, s2)) {
          v = proj_plane(s1, s2, s3);
        } else {
          v = proj_line(s1, s3);
          s.n = 2;
          smove<1, 2>(s);
        }
      } else {
        v = proj_plane(s1, s2, s3);
      }
    } else {
      v = proj_line(s1, s2);
      s.n = 2;
      smove<0, 2>(s);
    }
  } else if (h13) {
    if (!hff2(s1, s3, s2)) {
      v = proj_plane(s1, s2, s3);
    } else {
      v = proj_line(s1, s3);
      s.n = 2;
      smove<1, 2>(s);
    }
  } else {
    v = s.p[2];
    s.n = 1;
    smove<0, 2>(s);
  }
}

// S3D: GJK/cpu/openGJK.c:315-613.  (k, i, j) name the three older vertices
// with k opposite the face under test, as in the reference.
OGJK_DEV void S3D(Simplex& s, D3& v) {
  const Pt P1{s.p[3], s.ia[3], s.ib[3]};
  const D3 s1 = P1.p, s2 = s.p[2], s3 = s.p[1], s4 = s.p[0];
  const Pt P2{s2, s.ia[2], s.ib[2]};
  const Pt P3{s3, s.ia[1], s.ib[1]};
  const D3 s1s2 = sub(s2, s1), s1s3 = sub(s3, s1), s1s4 = sub(s4, s1);

  const int h1_2 = hff1(s1, s2);
  const int h1_1 = hff1(s1, s3);
  const int h1_0 = hff1(s1, s4);
  const int line3 = h1_1, line4 = h1_0;
  const int dot_total = h1_2 + line3 + line4;

  if (dot_total == 0) {
    v = s1;
    s.n = 1;
    s.p[0] = s1; s.ia[0] = P1.a; s.ib[0] = P1.b;
    return;
  }

  const double det134 = det3(s1s3, s1s4, s1s2);
  const int sss = (det134 <= 0);
  int p2 = hff3(s1, s3, s4) - sss;
  p2 = p2 * p2;
  int p3 = hff3(s1, s4, s2) - sss;
  p3 = p3 * p3;
  int p4 = hff3(s1, s2, s3) - sss;
  p4 = p4 * p4;
  const int planes = p2 + p3 + p4;

  if (planes == 3) {
    v = D3{0.0, 0.0, 0.0};
    s.n = 4;
    return;
  }
  if (planes == 2) {
    s.n = 3;
    if (!p2) {
      smove<2, 3>(s);
    } else if (!p3) {
      s.p[1] = P2.p; s.ia[1] = P2.a; s.ib[1] = P2.b;
      smove<2, 3>(s);
    } else if (!p4) {
      s.p[0] = P3.p; s.ia[0] = P3.a; s.ib[0] = P3.b;
      s.p[1] = P2.p; s.ia[1] = P2.a; s.ib[1] = P2.b;
      smove<2, 3>(s);
    }
    S2D(s, v);
    return;
  }
  if (planes == 1) {
    s.n = 3;
    int k, i, j;
    if (p2) {
      k = 2; i = 1; j = 0;
    } else if (p3) {
      k = 1; i = 0; j = 2;
    } else {
      k = 0; i = 2; j = 1;
    }
    const Pt si = spt(s, i), sj = spt(s, j), sk = spt(s, k);
    const int hk = (k == 2) ? h1_2 : (k == 1 ? h1_1 : h1_0);
    const int hi = (i == 2) ? h1_2 : (i == 1 ? h1_1 : h1_0);
    const int hj = (j == 2) ? h1_2 : (j == 1 ? h1_1 : h1_0);
    if (dot_total == 1) {
      if (hk) {
        if (!hff2(s1, sk.p, si.p)) {
          keep3(s, si, sk);
          v = proj_plane(s1, si.p, sk.p);
        } else if (!hff2(s1, sk.p, sj.p)) {
          keep3(s, sj, sk);
          v = proj_plane(s1, sj.p, sk.p);
        } else {
          keep2(s, sk);
          v = proj_line(s1, sk.p);
        }
      } else if (hi) {
        if (!hff2(s1, si.p, sk.p)) {
          keep3(s, si, sk);
          v = proj_plane(s1, si.p, sk.p);
        } else {
          keep2(s, si);
          v = proj_line(s1, si.p);
        }
      } else {
        if (!hff2(s1, sj.p, sk.p)) {
          keep3(s, sj, sk);
          v = proj_plane(s1, sj.p, sk.p);
        } else {
          keep2(s, sj);
          v = proj_line(s1, sj.p);
        }
      }
    } else if (dot_total == 2) {
      if (hi) {
        if (!hff2(s1, sk.p, si.p)) {
          if (!hff2(s1, si.p, sk.p)) {
            keep3(s, si, sk);
            v = proj_plane(s1, si.p, sk.p);
          } else {
            keep2(s, sk);
            v = proj_line(s1, sk.p);
          }
        } else {
          if (!hff2(s1, sk.p, sj.p)) {
            keep3(s, sj, sk);
            v = proj_plane(s1, sj.p, sk.p);
          } else {
            keep2(s, sk);
            v = proj_line(s1, sk.p);
          }
        }
      } else if (hj) {
        if (!hff2(s1, sk.p, sj.p)) {
          if (!hff2(s1, sj.p, sk.p)) {
            keep3(s, sj, sk);
            v = proj_plane(s1, sj.p, sk.p);
          } else {
            keep2(s, sj);
            v = proj_line(s1, sj.p);
          }
        } else {
          if (!hff2(s1, sk.p, si.p)) {
            keep3(s, si, sk);
            v = proj_plane(s1, si.p, sk.p);
          } else {
            keep2(s, sk);
            v = proj_line(s1, sk.p);
          }
        }
      }
      // else: no region; v and the 3-vertex prefix stay (openGJK.c:494-496)
    } else {  // dot_total == 3
      const int h2ik = hff2(s1, si.p, sk.p);
      const int h2jk = hff2(s1, sj.p, sk.p);
      const int h2ki = hff2(s1, sk.p, si.p);
      const int h2kj = hff2(s1, sk.p, sj.p);
      if (h2ki == 1 && h2kj == 1) {
        keep2(s, sk);
        v = proj_line(s1, sk.p);
      } else if (h2ki) {
        if (h2jk) {
          keep2(s, sj);
          v = proj_line(s1, sj.p);
        } else {
          keep3(s, sj, sk);
          v = proj_plane(s1, sk.p, sj.p);
        }
      } else {
        if (h2ik) {
          keep2(s, si);
          v = proj_line(s1, si.p);
        } else {
          keep3(s, si, sk);
          v = proj_plane(s1, sk.p, si.p);
        }
      }
    }
    return;
  }
  // planes == 0: origin outside all three faces (openGJK.c:538-608)
  if (dot_total == 1) {
    int k, i, j;
    if (line3) {
      k = 2; i = 1; j = 0;
    } else if (line4) {
      k = 1; i = 0; j = 2;
    } else {
      k = 0; i = 2; j = 1;
    }
    const Pt si = spt(s, i), sj = spt(s, j), sk = spt(s, k);
    if (!hff2(s1, si.p, sj.p)) {
      keep3(s, si, sj);
      v = proj_plane(s1, si.p, sj.p);
    } else if (!hff2(s1, si.p, sk.p)) {
      keep3(s, si, sk);
      v = proj_plane(s1, si.p, sk.p);
    } else {
      keep2(s, si);
      v = proj_line(s1, si.p);
    }
  } else if (dot_total == 2) {
    s.n = 3;
    int k, i, j;
    if (!line3) {
      k = 2; i = 1; j = 0;
    } else if (!line4) {
      k = 1; i = 0; j = 2;
    } else {
      k = 0; i = 2; j = 1;
    }
    const Pt si = spt(s, i), sj = spt(s, j), sk = spt(s, k);
    if (!hff2(s1, sj.p, sk.p)) {
      if (!hff2(s1, sk.p, sj.p)) {
        keep3(s, sj, sk);
        v = proj_plane(s1, sj.p, sk.p);
      } else if (!hff2(s1, sk.p, si.p)) {
        keep3(s, si, sk);
        v = proj_plane(s1, sk.p, si.p);
      } else {
        keep2(s, sk);
        v = proj_line(s1, sk.p);
      }
    } else if (!hff2(s1, sj.p, si.p)) {
      keep3(s, si, sj);
      v = proj_plane(s1, si.p, sj.p);
    } else {
      keep2(s, sj);
      v = proj_line(s1, sj.p);
    }
  }
  // dot_total == 3 with planes == 0: nothing (the reference has no branch)
}

// subalgorithm: GJK/cpu/openGJK.c:641-656
OGJK_DEV void subalgorithm(Simplex& s, D3& v) {
  if (s.n == 4)
    S3D(s, v);
  else if (s.n == 3)
    S2D(s, v);
  else if (s.n == 2)
    S1D(s, v);
}

// ------------------------------------------------------- witnesses W*D

struct Wit {
  D3 w0, w1;
};

template <class BT>
OGJK_DEV void W0D(const BT& A, const BT& B, const Simplex& s, Wit& w) {
  w.w0 = A.d(s.ia[0]);
  w.w1 = B.d(s.ib[0]);
}

// Blend of original vertices: sum_k (coord_k * a_k), left to right.
template <class BT>
OGJK_DEV double blend_c(const BT& b, const int* idx, const double* a, int m, int c) {
  double acc = 0.0;
  for (int k = 0; k < m; k++) {
    const int i = idx[k];
    const double x = b.c(i, c);
    acc = (k == 0) ? x * a[k] : acc + x * a[k];
  }
  return acc;
}

// W1D: GJK/cpu/openGJK.c:668-700 (det == 0 falls through, as the reference)
template <class BT>
OGJK_DEV void W1D(const BT& A, const BT& B, const Simplex& s, Wit& w) {
  const D3 p = s.p[0], q = s.p[1];
  const D3 pq = sub(q, p), po = neg(p);
  const double det = dot(pq, pq);
  if (det == 0.0) W0D(A, B, s, w);
  const double a1 = dot(pq, po) / det;
  const double a[2] = {1.0 - a1, a1};
  const int ia[2] = {s.ia[0], s.ia[1]}, ib[2] = {s.ib[0], s.ib[1]};
  w.w0 = D3{blend_c(A, ia, a, 2, 0), blend_c(A, ia, a, 2, 1), blend_c(A, ia, a, 2, 2)};
  w.w1 = D3{blend_c(B, ib, a, 2, 0), blend_c(B, ib, a, 2, 1), blend_c(B, ib, a, 2, 2)};
}

// W2D: GJK/cpu/openGJK.c:702-795
template <class BT>
OGJK_DEV void W2D(const BT& A, const BT& B, Simplex& s, Wit& w) {
  const D3 p = s.p[0], q = s.p[1], r = s.p[2];
  const D3 pq = sub(q, p), pr = sub(r, p), po = neg(p);
  const double T00 = dot(pq, pq), T01 = dot(pq, pr), T11 = dot(pr, pr);
  const double det = T00 * T11 - T01 * T01;
  if (det == 0.0) W1D(A, B, s, w);
  const double b0 = dot(pq, po), b1 = dot(pr, po);
  const double I00 = T11 / det, I01 = -T01 / det, I11 = T00 / det;
  const double a1 = I00 * b0 + I01 * b1;
  const double a2 = I01 * b0 + I11 * b1;
  const double a0 = 1.0 - a1 - a2;
  if (a0 < kEps64) {
    s.n = 2;
    smove<0, 2>(s);
    W1D(A, B, s, w);
  } else if (a1 < kEps64) {
    s.n = 2;
    smove<1, 2>(s);
    W1D(A, B, s, w);
  } else if (a2 < kEps64) {
    s.n = 2;
    W1D(A, B, s, w);
  }
  const double a[3] = {a0, a1, a2};
  const int ia[3] = {s.ia[0], s.ia[1], s.ia[2]}, ib[3] = {s.ib[0], s.ib[1], s.ib[2]};
  w.w0 = D3{blend_c(A, ia, a, 3, 0), blend_c(A, ia, a, 3, 1), blend_c(A, ia, a, 3, 2)};
  w.w1 = D3{blend_c(B, ib, a, 3, 0), blend_c(B, ib, a, 3, 1), blend_c(B, ib, a, 3, 2)};
}

// W3D: GJK/cpu/openGJK.c:797-929
template <class BT>
OGJK_DEV void W3D(const BT& A, const BT& B, Simplex& s, Wit& w) {
  const D3 p = s.p[0], q = s.p[1], r = s.p[2], t = s.p[3];
  const D3 pq = sub(q, p), pr = sub(r, p), ps = sub(t, p), po = neg(p);
  const double T00 = dot(pq, pq), T01 = dot(pq, pr), T02 = dot(pq, ps);
  const double T11 = dot(pr, pr), T12 = dot(pr, ps), T22 = dot(ps, ps);
  const double det00 = T11 * T22 - T12 * T12;
  const double det01 = T01 * T22 - T02 * T12;
  const double det02 = T01 * T12 - T02 * T11;
  const double det = T00 * det00 - T01 * det01 + T02 * det02;
  if (det == 0.0) W2D(A, B, s, w);
  const double b0 = dot(pq, po), b1 = dot(pr, po), b2 = dot(ps, po);
  const double det11 = T00 * T22 - T02 * T02;
  const double det12 = T00 * T12 - T01 * T02;
  const double det22 = T00 * T11 - T01 * T01;
  const double I00 = det00 / det, I01 = -det01 / det, I02 = det02 / det;
  const double I11 = det11 / det, I12 = -det12 / det, I22 = det22 / det;
  const double a1 = I00 * b0 + I01 * b1 + I02 * b2;
  const double a2 = I01 * b0 + I11 * b1 + I12 * b2;
  const double a3 = I02 * b0 + I12 * b1 + I22 * b2;
  const double a0 = 1.0 - a1 - a2 - a3;
  if (a0 < kEps64) {
    s.n = 3;
    smove<0, 3>(s);
    W2D(A, B, s, w);
  } else if (a1 < kEps64) {
    s.n = 3;
    smove<1, 3>(s);
    W2D(A, B, s, w);
  } else if (a2 < kEps64) {
    s.n = 3;
    smove<2, 3>(s);
    W2D(A, B, s, w);
  } else if (a3 < kEps64) {
    s.n = 3;
    W2D(A, B, s, w);
  }
  const double a[4] = {a0, a1, a2, a3};
  const int ia[4] = {s.ia[0], s.ia[1], s.ia[2], s.ia[3]};
  const int ib[4] = {s.ib[0], s.ib[1], s.ib[2], s.ib[3]};
  w.w0 = D3{blend_c(A, ia, a, 4, 0), blend_c(A, ia, a, 4, 1), blend_c(A, ia, a, 4, 2)};
  w.w1 = D3{blend_c(B, ib, a, 4, 0), blend_c(B, ib, a, 4, 1), blend_c(B, ib, a, 4, 2)};
}

// compute_witnesses: GJK/cpu/openGJK.c:931-950
template <class BT>
OGJK_DEV void witnesses(const BT& A, const BT& B, Simplex& s, Wit& w) {
  if (s.n == 4)
    W3D(A, B, s, w);
  else if (s.n == 3)
    W2D(A, B, s, w);
  else if (s.n == 2)
    W1D(A, B, s, w);
  else if (s.n == 1)
    W0D(A, B, s, w);
}

// ------------------------------------------------------------- GJK driver

// compute_minimum_distance: GJK/cpu/openGJK.c:952-1056.  Returns ||v||.
template <int LPP, class BT>
OGJK_DEV double gjk(const BT& A, const BT& B, Simplex& s, Wit& w, int lane) {
  const double eps_rel = kEps64 * 1e4;  // (gkFloat)DBL_EPSILON * 1e4f
  const double eps_tot = kEps64 * 1e2;
  const double eps_rel2 = eps_rel * eps_rel;

  D3 pa = A.d(0), pb = B.d(0);
  int ia = 0, ib = 0;
  D3 v = sub(pa, pb);
  s.n = 1;
  s.p[0] = v;
  s.ia[0] = 0;
  s.ib[0] = 0;
  double norm2Wmax = 0.0;
  unsigned k = 0;
  do {
    k++;
    Sup sA, sB;
    support_pair<LPP>(A, Query{neg(v), true, ia, pa}, B, Query{v, true, ib, pb}, lane, sA, sB);
    ia = sA.i;
    pa = sA.p;
    ib = sB.i;
    pb = sB.p;
    const D3 wv = sub(pa, pb);
    const double gap = norm2(v) - dot(v, wv);
    if (gap <= eps_rel * norm2(v) || gap < eps_tot) break;
    if (norm2(v) < eps_rel2) break;

    sset(s, s.n, wv, ia, ib);
    s.n++;
    subalgorithm(s, v);

#pragma unroll
    for (int jj = 0; jj < 4; jj++) {
      if (jj < s.n) {
        const double tn = norm2(s.p[jj]);
        if (tn > norm2Wmax) norm2Wmax = tn;
      }
    }
    if (norm2(v) <= eps_tot * eps_tot * norm2Wmax) break;
  } while (s.n != 4 && k != 25u);

  witnesses(A, B, s, w);
  return sqrt(norm2(v));
}

}  // namespace ogjk
