// opengjk-amd device EPA for one pair on a group of LPP lanes (LPP <= 64).
//
// Layout: the polytope's vertex pool (<= 132 FP64 points and their body
// indices), the horizon edge lists and the input witnesses live in LDS; the
// 128 face slots are spread over the group's lanes, slot = q*LPP + lane,
// FPL = 128/LPP faces per lane held in registers (normal, plane distance,
// packed vertex ids).  Nothing else of the pair stays in registers through
// the expansion loop: the starting tetrahedron is slots 0-3 of the vertex
// pool and is read back from there when the loop ends.
//
// The reference EPA (GJK/cpu/EPA.c:362-864) is sequential; its discrete
// choices are reproduced exactly with ballots:
//   * closest face: min distance, lowest slot on ties (EPA.c:608-617);
//   * horizon: edges of the faces removed this iteration, in slot order and
//     (v0v1, v1v2, v2v0) order; an edge that occurs more than once is dropped
//     (the pairwise cancellation of EPA.c:746-759 drops every occurrence);
//   * new faces take the lowest free slots in horizon order (EPA.c:762-775);
//   * a face's normal is computed when it is created and a degenerate face is
//     retired at the start of the next iteration, which is when the reference
//     recomputes it (EPA.c:600-604) -- until then it still occupies its slot.
#pragma once

#include "ogjk_core.hpp"

namespace ogjk {

constexpr int kEpaMaxFaces = 128;
constexpr int kEpaMaxVerts = kEpaMaxFaces + 4;

struct EpaShared {
  double vx[kEpaMaxVerts], vy[kEpaMaxVerts], vz[kEpaMaxVerts];
  int via[kEpaMaxVerts], vib[kEpaMaxVerts];
  // edges of the faces removed this iteration: directed edge v1 | v2 << 8 in
  // the low half, its direction-free key in the high half
  unsigned edir[kEpaMaxFaces * 3];
  unsigned short surv[kEpaMaxFaces * 3];  // surviving horizon edges, in order
  double w_in[6];                         // the witnesses EPA started from
};

OGJK_DEV void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// Directed edge (a, b) packed as a | b << 8, with its direction-free key
// (min | max << 8) in the high 16 bits.
OGJK_DEV unsigned edge_word(int a, int b) {
  const unsigned lo = a < b ? (unsigned)(a | (b << 8)) : (unsigned)(b | (a << 8));
  return (unsigned)(a | (b << 8)) | (lo << 16);
}

OGJK_DEV D3 lds_v(const EpaShared& sh, int i) { return D3{sh.vx[i], sh.vy[i], sh.vz[i]}; }

// set_contact_normal: GJK/cpu/EPA.c:350-360
OGJK_DEV D3 contact_normal(D3 w1, D3 w2) {
  const D3 d = sub(w2, w1);
  const double n = sqrt(norm2(d));
  if (n > kEps64) return D3{d.x / n, d.y / n, d.z / n};
  return D3{1.0, 0.0, 0.0};
}

// compute_barycentric_origin: GJK/cpu/EPA.c:238-304
OGJK_DEV void bary_origin(D3 v0, D3 v1, D3 v2, double& a0, double& a1, double& a2) {
  const D3 e0 = sub(v1, v0), e1 = sub(v2, v0);
  const double d00 = dot(e0, e0), d01 = dot(e0, e1), d11 = dot(e1, e1);
  const double d20 = -dot(v0, e0), d21 = -dot(v0, e1);
  const double denom = d00 * d11 - d01 * d01;
  if (fabs(denom) < kEps64) {
    a0 = a1 = a2 = 1.0 / 3.0;
    return;
  }
  const double inv = 1.0 / denom;
  const double u = (d11 * d20 - d01 * d21) * inv;
  const double v = (d00 * d21 - d01 * d20) * inv;
  const double w = 1.0 - u - v;
  if (w < 0) {
    const D3 e12 = sub(v2, v1);
    double t = -dot(v1, e12) / dot(e12, e12);
    t = fmax(0.0, fmin(1.0, t));
    a0 = 0;
    a1 = 1.0 - t;
    a2 = t;
  } else if (u < 0) {
    double t = -dot(v0, e1) / dot(e1, e1);
    t = fmax(0.0, fmin(1.0, t));
    a0 = 1.0 - t;
    a1 = 0;
    a2 = t;
  } else if (v < 0) {
    double t = -dot(v0, e0) / dot(e0, e0);
    t = fmax(0.0, fmin(1.0, t));
    a0 = 1.0 - t;
    a1 = t;
    a2 = 0;
  } else {
    a0 = w;
    a1 = u;
    a2 = v;
  }
}

template <int FPL>
struct Faces {
  double nx[FPL], ny[FPL], nz[FPL], dist[FPL];
  int verts[FPL];  // v0 | v1 << 8 | v2 << 16
  unsigned valid;  // bit q
  unsigned pend;   // bit q: degenerate, retire at the next iteration start
};

// A built face: unit normal, plane distance, packed vertex ids, and whether
// it is degenerate (retired at the next iteration start).
struct NewFace {
  D3 n;
  double d;
  int packed;
  bool ok;
};

// A new face (v0, v1, v2): winding fixed against the centroid
// (GJK/cpu/EPA.c:204-231, 791-819), then its normal and plane distance
// (compute_face_normal_distance, GJK/cpu/EPA.c:92-129).  The winding test's
// cross product is reused for the plane: with no contraction, cross(p2 - p0,
// p1 - p0) is exactly -cross(p1 - p0, p2 - p0) (products commute exactly,
// x - y = -(y - x)), so the swapped face's normal is the negated one, bit for
// bit.
OGJK_DEV NewFace build_face(const EpaShared& sh, int v0, int v1, int v2, D3 centroid) {
  const D3 p0 = lds_v(sh, v0);
  D3 n = cross(sub(lds_v(sh, v1), p0), sub(lds_v(sh, v2), p0));
  const bool swap = dot(n, sub(centroid, p0)) > 0;
  NewFace f;
  f.packed = swap ? (v0 | (v2 << 8) | (v1 << 16)) : (v0 | (v1 << 8) | (v2 << 16));
  if (swap) n = neg(n);
  const double nsq = norm2(n);
  f.ok = nsq > kEps64 * kEps64;
  if (f.ok) {
    const double nn = sqrt(nsq);
    n = D3{n.x / nn, n.y / nn, n.z / nn};
    f.d = dot(n, p0);
    if (f.d < 0) {
      n = neg(n);
      f.d = -f.d;
    }
  } else {
    f.d = 1e10;
  }
  f.n = n;
  return f;
}

template <int FPL>
OGJK_DEV void put_face(Faces<FPL>& F, int q, const NewFace& f) {
  F.nx[q] = f.n.x;
  F.ny[q] = f.n.y;
  F.nz[q] = f.n.z;
  F.dist[q] = f.d;
  F.verts[q] = f.packed;
  F.valid |= 1u << q;
  if (!f.ok) F.pend |= 1u << q;
}

// Exclusive rank of set flags in slot order (slot = q*LPP + lane).
template <int LPP, int FPL>
OGJK_DEV int slot_rank(const bool (&flag)[FPL], int (&rank)[FPL]) {
  int base = 0;
#pragma unroll
  for (int q = 0; q < FPL; q++) {
    const uint64_t gb = group_ballot<LPP>(flag[q]);
    rank[q] = base + group_prefix<LPP>(gb);
    base += __popcll(gb);
  }
  return base;
}

// Closest valid face with distance in [0, 1e10), lowest slot on ties.
template <int LPP, int FPL>
OGJK_DEV int closest_face(const Faces<FPL>& F, int lane, double& cdist) {
  double bv = INFINITY;
  int bi = 0x7fffffff;
#pragma unroll
  for (int q = 0; q < FPL; q++) {
    const double d = F.dist[q];
    const bool ok = ((F.valid >> q) & 1u) && d >= 0.0 && d < 1e10;
    if (ok && d < bv) {
      bv = d;
      bi = q * LPP + lane;
    }
  }
  group_argmin_d<LPP>(bv, bi);
  cdist = bv;
  return bi == 0x7fffffff ? -1 : bi;
}

template <int LPP, int FPL>
OGJK_DEV void fetch_face(const Faces<FPL>& F, int slot, D3& n, int& verts) {
  const int q = slot / LPP, src = slot % LPP;
  // value selects over laundered registers (see keep_reg): no address select
  double x = keep_reg(F.nx[0]), y = keep_reg(F.ny[0]), z = keep_reg(F.nz[0]);
  int vv = keep_reg(F.verts[0]);
#pragma unroll
  for (int t = 1; t < FPL; t++) {
    const bool h = (q == t);
    x = h ? keep_reg(F.nx[t]) : x;
    y = h ? keep_reg(F.ny[t]) : y;
    z = h ? keep_reg(F.nz[t]) : z;
    vv = h ? keep_reg(F.verts[t]) : vv;
  }
  n = D3{shfl_d<LPP>(x, src), shfl_d<LPP>(y, src), shfl_d<LPP>(z, src)};
  verts = shfl_i<LPP>(vv, src);
}

// support_epa (GJK/cpu/EPA.c:307-344): farthest vertex of A along dir and of
// B along -dir; returns the Minkowski point A[ia] - B[ib].
template <int LPP, class BT>
OGJK_DEV D3 support_epa(const BT& A, const BT& B, D3 dir, int& ia, int& ib, int lane) {
  const D3 zero{0.0, 0.0, 0.0};
  Sup a, b;
  support_pair<LPP>(A, Query{dir, false, 0, zero}, B, Query{neg(dir), false, 0, zero}, lane, a,
                    b);
  // The reference leaves its output unset when no vertex beats -1e10; such
  // coordinates are out of any physical range, fall back to vertex 0.
  if (a.i < 0) a = Sup{0, A.d(0)};
  if (b.i < 0) b = Sup{0, B.d(0)};
  ia = a.i;
  ib = b.i;
  return sub(a.p, b.p);
}

// Is p farther than eps from every simplex vertex (GJK/cpu/EPA.c:453-463)?
OGJK_DEV bool is_new_point(const Simplex& s, D3 p) {
  const double eps_sq = kEps64 * kEps64;
  bool fresh = true;
#pragma unroll
  for (int t = 0; t < 4; t++) {
    if (t < s.n) {
      const D3 q = s.p[t];
      const double dx = p.x - q.x, dy = p.y - q.y, dz = p.z - q.z;
      if (dx * dx + dy * dy + dz * dz < eps_sq) fresh = false;
    }
  }
  return fresh;
}

template <class BT>
OGJK_DEV void zero_depth(const BT& A, const BT& B, int ia, int ib,
                         double& distance, Wit& w, D3& normal) {
  distance = 0.0;
  w.w0 = A.d(ia);
  w.w1 = B.d(ib);
  normal = contact_normal(w.w0, w.w1);
}

template <class BT>
OGJK_DEV void face_result(const BT& A, const BT& B, const EpaShared& sh,
                          int verts, D3 n, double cdist, double& distance, Wit& w, D3& normal) {
  const int v0 = verts & 0xff, v1 = (verts >> 8) & 0xff, v2 = (verts >> 16) & 0xff;
  double a[3];
  bary_origin(lds_v(sh, v0), lds_v(sh, v1), lds_v(sh, v2), a[0], a[1], a[2]);
  const int ia[3] = {sh.via[v0], sh.via[v1], sh.via[v2]};
  const int ib[3] = {sh.vib[v0], sh.vib[v1], sh.vib[v2]};
  w.w0 = D3{blend_c(A, ia, a, 3, 0), blend_c(A, ia, a, 3, 1), blend_c(A, ia, a, 3, 2)};
  w.w1 = D3{blend_c(B, ib, a, 3, 0), blend_c(B, ib, a, 3, 1), blend_c(B, ib, a, 3, 2)};
  normal = n;
  distance = -cdist;
}

// computeCollisionInformation for one pair (GJK/cpu/EPA.c:362-864).
template <int LPP, class BT>
OGJK_DEV void epa(const BT& A, const BT& B, Simplex& s, double& distance, Wit& w, D3& normal,
                  EpaShared& sh, int lane) {
  constexpr int FPL = kEpaMaxFaces / LPP;
  static_assert(FPL * LPP == kEpaMaxFaces, "LPP must divide 128");

  if (distance > kEps64) {
    normal = contact_normal(w.w0, w.w1);
    return;
  }

  // -------- grow a degenerate GJK simplex to a tetrahedron (EPA.c:376-583)
  if (s.n != 4) {
    int ia, ib;
    if (s.n == 1) {
      const D3 nv = support_epa<LPP>(A, B, s.p[0], ia, ib, lane);
      if (is_new_point(s, nv)) {
        sset(s, 1, nv, ia, ib);
        s.n = 2;
      } else {
        zero_depth(A, B, ia, ib, distance, w, normal);
        return;
      }
    }
    if (s.n == 2) {
      const D3 edge = sub(s.p[1], s.p[0]);
      D3 axis{1.0, 0.0, 0.0};
      const double en = sqrt(norm2(edge));
      if (en > kEps64 && fabs(edge.x) > (double)0.9f * en) axis = D3{0.0, 1.0, 0.0};
      D3 dir = cross(edge, axis);
      if (norm2(dir) < kEps64) {
        axis = D3{0.0, 0.0, 1.0};
        dir = cross(edge, axis);
      }
      const D3 nv = support_epa<LPP>(A, B, dir, ia, ib, lane);
      if (is_new_point(s, nv)) {
        sset(s, 2, nv, ia, ib);
        s.n = 3;
      } else {
        zero_depth(A, B, ia, ib, distance, w, normal);
        return;
      }
    }
    if (s.n == 3) {
      D3 dir = cross(sub(s.p[1], s.p[0]), sub(s.p[2], s.p[0]));
      D3 nv = support_epa<LPP>(A, B, dir, ia, ib, lane);
      if (is_new_point(s, nv)) {
        sset(s, 3, nv, ia, ib);
        s.n = 4;
      } else {
        dir = neg(dir);
      }
      if (s.n == 3) {
        nv = support_epa<LPP>(A, B, dir, ia, ib, lane);
        if (is_new_point(s, nv)) {
          sset(s, 3, nv, ia, ib);
          s.n = 4;
        } else {
          zero_depth(A, B, ia, ib, distance, w, normal);
          return;
        }
      }
    }
    if (s.n != 4) {
      int best = s.n > 0 ? s.n - 1 : 0;
      best = best > 3 ? 3 : best;
      int ba = sa(s, best), bb = sb(s, best);
      ba = (ba >= 0 && ba < A.n) ? ba : 0;  // sanitise caller-provided indices
      bb = (bb >= 0 && bb < B.n) ? bb : 0;
      zero_depth(A, B, ba, bb, distance, w, normal);
      return;
    }
  }

  // -------- initial tetrahedron (init_epa_polytope, EPA.c:144-235)
  if (lane < 4) {
    const D3 p = sp(s, lane);
    sh.vx[lane] = p.x;
    sh.vy[lane] = p.y;
    sh.vz[lane] = p.z;
    sh.via[lane] = sa(s, lane);
    sh.vib[lane] = sb(s, lane);
  }
  if (lane == 0) {
    sh.w_in[0] = w.w0.x;
    sh.w_in[1] = w.w0.y;
    sh.w_in[2] = w.w0.z;
    sh.w_in[3] = w.w1.x;
    sh.w_in[4] = w.w1.y;
    sh.w_in[5] = w.w1.z;
  }
  wave_sync();
  D3 centroid{0.0, 0.0, 0.0};
#pragma unroll
  for (int i = 0; i < 4; i++) {
    centroid.x += s.p[i].x * 0.25;
    centroid.y += s.p[i].y * 0.25;
    centroid.z += s.p[i].z * 0.25;
  }
  Faces<FPL> F;
  F.valid = 0;
  F.pend = 0;
#pragma unroll
  for (int q = 0; q < FPL; q++) {
    F.nx[q] = F.ny[q] = F.nz[q] = 0.0;
    F.dist[q] = 0.0;
    F.verts[q] = 0;
  }
  if (lane < 4) {
    // faces (0,1,2) (0,3,1) (0,2,3) (1,3,2)
    const int a = lane == 3 ? 1 : 0;
    const int b = lane == 0 ? 1 : (lane == 1 ? 3 : (lane == 2 ? 2 : 3));
    const int c = lane == 0 ? 2 : (lane == 1 ? 1 : (lane == 2 ? 3 : 2));
    put_face(F, 0, build_face(sh, a, b, c, centroid));
  }

  const double tol = kEps64 * 1e2;
  int nverts = 4;
  int iteration = 0;
  bool wrote = false;  // w, distance and normal hold a face's result
  while (iteration < 64 && nverts < kEpaMaxVerts - 1) {
    iteration++;
    F.valid &= ~F.pend;  // the reference's per-iteration recompute retires them
    F.pend = 0;

    double cdist;
    const int cf = closest_face<LPP>(F, lane, cdist);
    if (cf < 0) break;
    D3 n;
    int cverts;
    fetch_face<LPP>(F, cf, n, cverts);

    int ia, ib;
    const D3 nv = support_epa<LPP>(A, B, n, ia, ib, lane);
    const double improvement = dot(n, nv) - cdist;
    if (improvement < tol) {
      face_result(A, B, sh, cverts, n, cdist, distance, w, normal);
      wrote = true;
      break;
    }

    bool dup = false;
    {
      const double eps_sq = kEps64 * kEps64;
      for (int t = lane; t < nverts; t += LPP) {
        const double dx = nv.x - sh.vx[t], dy = nv.y - sh.vy[t], dz = nv.z - sh.vz[t];
        if (dx * dx + dy * dy + dz * dz < eps_sq) dup = true;
      }
    }
    if (group_any<LPP>(dup)) {
      face_result(A, B, sh, cverts, n, cdist, distance, w, normal);
      wrote = true;
      break;
    }

    // add the vertex and update the running-mean centroid (EPA.c:686-699)
    const int nid = nverts;
    if (lane == 0) {
      sh.vx[nid] = nv.x;
      sh.vy[nid] = nv.y;
      sh.vz[nid] = nv.z;
      sh.via[nid] = ia;
      sh.vib[nid] = ib;
    }
    nverts++;
    const double inv_n = 1.0 / (double)nverts;
    centroid.x += (nv.x - centroid.x) * inv_n;
    centroid.y += (nv.y - centroid.y) * inv_n;
    centroid.z += (nv.z - centroid.z) * inv_n;
    wave_sync();

    // faces visible from the new vertex (is_face_visible, EPA.c:132-140)
    bool vis[FPL];
#pragma unroll
    for (int q = 0; q < FPL; q++) {
      vis[q] = false;
      if ((F.valid >> q) & 1u) {
        const D3 p0 = lds_v(sh, F.verts[q] & 0xff);
        const D3 fn{F.nx[q], F.ny[q], F.nz[q]};
        vis[q] = dot(fn, sub(nv, p0)) > kEps64;
      }
    }
    int vrank[FPL];
    const int nvis = slot_rank<LPP, FPL>(vis, vrank);
#pragma unroll
    for (int q = 0; q < FPL; q++) {
      if (vis[q]) {
        F.valid &= ~(1u << q);
        const int pv = F.verts[q];
        const int x0 = pv & 0xff, x1 = (pv >> 8) & 0xff, x2 = (pv >> 16) & 0xff;
        const int e = 3 * vrank[q];
        sh.edir[e + 0] = edge_word(x0, x1);
        sh.edir[e + 1] = edge_word(x1, x2);
        sh.edir[e + 2] = edge_word(x2, x0);
      }
    }
    wave_sync();

    // horizon = edges that occur exactly once, kept in edge order
    const int E = 3 * nvis;
    int nsurv = 0;
    for (int base = 0; base < E; base += LPP) {
      const int e = base + lane;
      bool keep = false;
      unsigned mine = 0u;
      if (e < E) {
        mine = sh.edir[e];
        const unsigned key = mine >> 16;
        int c = 0;
#pragma unroll 8
        for (int f = 0; f < E; f++) c += (sh.edir[f] >> 16) == key;
        keep = (c == 1);
      }
      const uint64_t gb = group_ballot<LPP>(keep);
      if (keep) sh.surv[nsurv + group_prefix<LPP>(gb)] = (unsigned short)(mine & 0xffffu);
      nsurv += __popcll(gb);
    }
    wave_sync();

    // fan new faces into the lowest free slots (EPA.c:761-825)
    bool fr[FPL];
    int frank[FPL];
#pragma unroll
    for (int q = 0; q < FPL; q++) fr[q] = !((F.valid >> q) & 1u);
    slot_rank<LPP, FPL>(fr, frank);
#pragma unroll
    for (int q = 0; q < FPL; q++) {
      if (fr[q] && frank[q] < nsurv) {
        const int ed = sh.surv[frank[q]];
        put_face(F, q, build_face(sh, ed & 0xff, (ed >> 8) & 0xff, nid, centroid));
      }
    }
  }

  if (iteration >= 64) {  // cap fallback (EPA.c:828-863)
    F.valid &= ~F.pend;
    F.pend = 0;
    double cdist;
    const int cf = closest_face<LPP>(F, lane, cdist);
    if (cf >= 0) {
      D3 n;
      int cverts;
      fetch_face<LPP>(F, cf, n, cverts);
      face_result(A, B, sh, cverts, n, cdist, distance, w, normal);
      wrote = true;
    }
  }
  // the starting tetrahedron (the simplex this call returns) and, when no
  // face result was written, the witnesses EPA started from
  s.n = 4;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    s.p[k] = lds_v(sh, k);
    s.ia[k] = sh.via[k];
    s.ib[k] = sh.vib[k];
  }
  if (!wrote) {
    w.w0 = D3{sh.w_in[0], sh.w_in[1], sh.w_in[2]};
    w.w1 = D3{sh.w_in[3], sh.w_in[4], sh.w_in[5]};
  }
}

}  // namespace ogjk
