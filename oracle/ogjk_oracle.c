/*
 * Oracle (test infrastructure only): CPU restatement of the reference
 * openGJK GJK and EPA.  See ogjk_oracle.h for scope and provenance.
 *
 * Every function cites the reference lines it follows.  The arithmetic is
 * kept in the reference's order and association, with C's usual conversions
 * (e.g. `1.0 - a1` is evaluated in double even when oreal is float, exactly as
 * GJK/cpu/openGJK.c:689 does), so that the FP32 instantiation is bit-identical
 * to the reference built as shipped.  Build with -ffp-contract=off.
 */
#include "ogjk_oracle.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdbool.h>
#include <stdlib.h>
#include <string.h>

#if OGJK_FP64
#define O_EPS DBL_EPSILON
#define O_SQRT sqrt
#define O_FABS fabs
#define O_FMAX fmax
#define O_FMIN fmin
#else
#define O_EPS FLT_EPSILON
#define O_SQRT sqrtf
#define O_FABS fabsf
#define O_FMAX fmaxf
#define O_FMIN fminf
#endif

/* openGJK.c:41-42 and EPA.c:46 */
#define EPS_REL ((oreal)O_EPS * (oreal)1e4f)
#define EPS_TOT ((oreal)O_EPS * (oreal)1e2f)

/* EPA.c:43-44 */
#define EPA_MAX_FACES 128
#define EPA_MAX_VERTS (EPA_MAX_FACES + 4)

/* ---------------------------------------------------------------- vectors */

static inline oreal v_dot(const oreal* a, const oreal* b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
static inline oreal v_norm2(const oreal* a) {
  return a[0] * a[0] + a[1] * a[1] + a[2] * a[2];
}
static inline void v_copy(oreal* d, const oreal* a) {
  d[0] = a[0];
  d[1] = a[1];
  d[2] = a[2];
}
static inline void v_sub(oreal* d, const oreal* a, const oreal* b) {
  d[0] = a[0] - b[0];
  d[1] = a[1] - b[1];
  d[2] = a[2] - b[2];
}
/* crossProduct: openGJK.c:175-181, EPA.c:81-87 */
static inline void v_cross(const oreal* a, const oreal* b, oreal* c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}

/* ----------------------------------------------- GJK predicates (a11) */

/* determinant: openGJK.c:167-173 */
static inline oreal det3(const oreal* p, const oreal* q, const oreal* r) {
  return p[0] * ((q[1] * r[2]) - (r[1] * q[2])) - p[1] * (q[0] * r[2] - r[0] * q[2]) +
         p[2] * (q[0] * r[1] - r[0] * q[1]);
}

/* projectOnLine: openGJK.c:183-196 */
static inline void proj_line(const oreal* p, const oreal* q, oreal* v) {
  oreal pq[3];
  v_sub(pq, p, q);
  const oreal t = v_dot(p, pq) / v_dot(pq, pq);
  for (int i = 0; i < 3; i++) v[i] = p[i] - pq[i] * t;
}

/* projectOnPlane: openGJK.c:198-217 */
static inline void proj_plane(const oreal* p, const oreal* q, const oreal* r, oreal* v) {
  oreal n[3], pq[3], pr[3];
  v_sub(pq, p, q);
  v_sub(pr, p, r);
  v_cross(pq, pr, n);
  const oreal t = v_dot(n, p) / v_dot(n, n);
  for (int i = 0; i < 3; i++) v[i] = n[i] * t;
}

/* hff1: openGJK.c:219-230 */
static inline int hff1(const oreal* p, const oreal* q) {
  oreal t = 0;
  for (int i = 0; i < 3; i++) t += (p[i] * p[i] - p[i] * q[i]);
  return t > 0;
}

/* hff2: openGJK.c:232-248 */
static inline int hff2(const oreal* p, const oreal* q, const oreal* r) {
  oreal nt[3], n[3], pq[3], pr[3];
  v_sub(pq, q, p);
  v_sub(pr, r, p);
  v_cross(pq, pr, nt);
  v_cross(pq, nt, n);
  return v_dot(p, n) < 0;
}

/* hff3: openGJK.c:250-255 */
static inline int hff3(const oreal* p, const oreal* q, const oreal* r) {
  oreal n[3];
  v_cross(q, r, n);
  return v_dot(p, n) <= 0;
}

/* ----------------------------------------------- simplex slot helpers */

static inline void slot_set(osimplex* s, int k, const oreal* p, const int* idx) {
  v_copy(s->vrtx[k], p);
  s->vrtx_idx[k][0] = idx[0];
  s->vrtx_idx[k][1] = idx[1];
}
static inline void slot_move(osimplex* s, int dst, int src) {
  v_copy(s->vrtx[dst], s->vrtx[src]);
  s->vrtx_idx[dst][0] = s->vrtx_idx[src][0];
  s->vrtx_idx[dst][1] = s->vrtx_idx[src][1];
}

/* Reduced simplex {newest, a, b} as 3 slots (select_1ik/1ij/1jk,
 * openGJK.c:53-78): slot2 <- slot3, slot1 <- a, slot0 <- b. */
static inline void keep3(osimplex* s, const oreal* a, const int* ia, const oreal* b,
                         const int* ib) {
  s->nvrtx = 3;
  slot_move(s, 2, 3);
  slot_set(s, 1, a, ia);
  slot_set(s, 0, b, ib);
}
/* Reduced simplex {newest, a} (select_1i/1j/1k, openGJK.c:80-99). */
static inline void keep2(osimplex* s, const oreal* a, const int* ia) {
  s->nvrtx = 2;
  slot_move(s, 1, 3);
  slot_set(s, 0, a, ia);
}

/* ----------------------------------------------- sub-distance S1D/S2D/S3D */

/* S1D: openGJK.c:257-268 */
static void S1D(osimplex* s, oreal* v) {
  if (hff1(s->vrtx[1], s->vrtx[0])) {
    proj_line(s->vrtx[1], s->vrtx[0], v);
  } else {
    v_copy(v, s->vrtx[1]); /* S1Dregion1, openGJK.c:118-127 */
    s->nvrtx = 1;
    slot_move(s, 0, 1);
  }
}

/* S2D: openGJK.c:270-313 (regions: openGJK.c:129-154) */
static void S2D(osimplex* s, oreal* v) {
  const oreal* s1 = s->vrtx[2];
  const oreal* s2 = s->vrtx[1];
  const oreal* s3 = s->vrtx[0];
  const int h12 = hff1(s1, s2);
  const int h13 = hff1(s1, s3);

  if (h12) {
    if (!hff2(s1, s2, s3)) {
      if (h13) {
        if (!hff2(s1, s3, s2)) {
          proj_plane(s1, s2, s3, v);
        } else {
          proj_line(s1, s3, v);
          s->nvrtx = 2; /* region 13 */
          slot_move(s, 1, 2);
        }
      } else {
        proj_plane(s1, s2, s3, v);
      }
    } else {
      proj_line(s1, s2, v);
      s->nvrtx = 2; /* region 12 */
      slot_move(s, 0, 2);
    }
  } else if (h13) {
    if (!hff2(s1, s3, s2)) {
      proj_plane(s1, s2, s3, v);
    } else {
      proj_line(s1, s3, v);
      s->nvrtx = 2; /* region 13 */
      slot_move(s, 1, 2);
    }
  } else {
    v_copy(v, s->vrtx[2]); /* region 1 */
    s->nvrtx = 1;
    slot_move(s, 0, 2);
  }
}

/* S3D: openGJK.c:315-613.  The reference's i/j/k naming of the remaining
 * three vertices is kept: slot k is the vertex opposite the face under test. */
static void S3D(osimplex* s, oreal* v) {
  oreal s1[3], s2[3], s3[3], s4[3], s1s2[3], s1s3[3], s1s4[3];
  oreal si[3], sj[3], sk[3];
  int s1i[2], s2i[2], s3i[2], sii[2], sji[2], ski[2];
  int i, j, k;

  v_copy(s1, s->vrtx[3]);
  s1i[0] = s->vrtx_idx[3][0];
  s1i[1] = s->vrtx_idx[3][1];
  v_copy(s2, s->vrtx[2]);
  s2i[0] = s->vrtx_idx[2][0];
  s2i[1] = s->vrtx_idx[2][1];
  v_copy(s3, s->vrtx[1]);
  s3i[0] = s->vrtx_idx[1][0];
  s3i[1] = s->vrtx_idx[1][1];
  v_copy(s4, s->vrtx[0]);
  v_sub(s1s2, s2, s->vrtx[3]);
  v_sub(s1s3, s3, s->vrtx[3]);
  v_sub(s1s4, s4, s->vrtx[3]);

  int h1[3];
  h1[2] = hff1(s1, s2);
  h1[1] = hff1(s1, s3);
  h1[0] = hff1(s1, s4);
  const int line3 = hff1(s1, s3);
  const int line4 = hff1(s1, s4);
  const int dot_total = hff1(s1, s2) + line3 + line4;

  if (dot_total == 0) { /* S3Dregion1, openGJK.c:156-165 */
    v_copy(v, s1);
    s->nvrtx = 1;
    slot_set(s, 0, s1, s1i);
    return;
  }

  const oreal det134 = det3(s1s3, s1s4, s1s2);
  const int sss = (det134 <= 0);
  int p2 = hff3(s1, s3, s4) - sss;
  p2 = p2 * p2;
  int p3 = hff3(s1, s4, s2) - sss;
  p3 = p3 * p3;
  int p4 = hff3(s1, s2, s3) - sss;
  p4 = p4 * p4;

#define LOAD_IJK()                \
  do {                            \
    v_copy(si, s->vrtx[i]);       \
    sii[0] = s->vrtx_idx[i][0];   \
    sii[1] = s->vrtx_idx[i][1];   \
    v_copy(sj, s->vrtx[j]);       \
    sji[0] = s->vrtx_idx[j][0];   \
    sji[1] = s->vrtx_idx[j][1];   \
    v_copy(sk, s->vrtx[k]);       \
    ski[0] = s->vrtx_idx[k][0];   \
    ski[1] = s->vrtx_idx[k][1];   \
  } while (0)

  switch (p2 + p3 + p4) {
    case 3: /* origin inside the tetrahedron (openGJK.c:47-51) */
      v[0] = 0;
      v[1] = 0;
      v[2] = 0;
      s->nvrtx = 4;
      break;

    case 2: /* one face faces the origin: drop the opposite vertex, then S2D */
      s->nvrtx = 3;
      if (!p2) {
        slot_move(s, 2, 3);
      } else if (!p3) {
        slot_set(s, 1, s2, s2i);
        slot_move(s, 2, 3);
      } else if (!p4) {
        slot_set(s, 0, s3, s3i);
        slot_set(s, 1, s2, s2i);
        slot_move(s, 2, 3);
      }
      S2D(s, v);
      break;

    case 1: /* two faces face the origin (openGJK.c:397-536) */
      s->nvrtx = 3;
      if (p2) {
        k = 2; i = 1; j = 0;
      } else if (p3) {
        k = 1; i = 0; j = 2;
      } else {
        k = 0; i = 2; j = 1;
      }
      LOAD_IJK();
      if (dot_total == 1) {
        if (h1[k]) {
          if (!hff2(s1, sk, si)) {
            keep3(s, si, sii, sk, ski);
            proj_plane(s1, si, sk, v);
          } else if (!hff2(s1, sk, sj)) {
            keep3(s, sj, sji, sk, ski);
            proj_plane(s1, sj, sk, v);
          } else {
            keep2(s, sk, ski);
            proj_line(s1, sk, v);
          }
        } else if (h1[i]) {
          if (!hff2(s1, si, sk)) {
            keep3(s, si, sii, sk, ski);
            proj_plane(s1, si, sk, v);
          } else {
            keep2(s, si, sii);
            proj_line(s1, si, v);
          }
        } else {
          if (!hff2(s1, sj, sk)) {
            keep3(s, sj, sji, sk, ski);
            proj_plane(s1, sj, sk, v);
          } else {
            keep2(s, sj, sji);
            proj_line(s1, sj, v);
          }
        }
      } else if (dot_total == 2) {
        if (h1[i]) {
          if (!hff2(s1, sk, si)) {
            if (!hff2(s1, si, sk)) {
              keep3(s, si, sii, sk, ski);
              proj_plane(s1, si, sk, v);
            } else {
              keep2(s, sk, ski);
              proj_line(s1, sk, v);
            }
          } else {
            if (!hff2(s1, sk, sj)) {
              keep3(s, sj, sji, sk, ski);
              proj_plane(s1, sj, sk, v);
            } else {
              keep2(s, sk, ski);
              proj_line(s1, sk, v);
            }
          }
        } else if (h1[j]) {
          if (!hff2(s1, sk, sj)) {
            if (!hff2(s1, sj, sk)) {
              keep3(s, sj, sji, sk, ski);
              proj_plane(s1, sj, sk, v);
            } else {
              keep2(s, sj, sji);
              proj_line(s1, sj, v);
            }
          } else {
            if (!hff2(s1, sk, si)) {
              keep3(s, si, sii, sk, ski);
              proj_plane(s1, si, sk, v);
            } else {
              keep2(s, sk, ski);
              proj_line(s1, sk, v);
            }
          }
        }
        /* else: no region selected; the reference leaves v and the
         * 3-vertex prefix untouched (openGJK.c:494-496) */
      } else if (dot_total == 3) {
        const int h2ik = hff2(s1, si, sk);
        const int h2jk = hff2(s1, sj, sk);
        const int h2ki = hff2(s1, sk, si);
        const int h2kj = hff2(s1, sk, sj);
        if (h2ki == 1 && h2kj == 1) {
          keep2(s, sk, ski);
          proj_line(s1, sk, v);
        } else if (h2ki) {
          if (h2jk) {
            keep2(s, sj, sji);
            proj_line(s1, sj, v);
          } else {
            keep3(s, sj, sji, sk, ski);
            proj_plane(s1, sk, sj, v);
          }
        } else {
          if (h2ik) {
            keep2(s, si, sii);
            proj_line(s1, si, v);
          } else {
            keep3(s, si, sii, sk, ski);
            proj_plane(s1, sk, si, v);
          }
        }
      }
      break;

    case 0: /* origin outside all three faces (openGJK.c:538-608) */
      if (dot_total == 1) {
        if (line3) {
          k = 2; i = 1; j = 0;
        } else if (line4) {
          k = 1; i = 0; j = 2;
        } else {
          k = 0; i = 2; j = 1;
        }
        LOAD_IJK();
        if (!hff2(s1, si, sj)) {
          keep3(s, si, sii, sj, sji);
          proj_plane(s1, si, sj, v);
        } else if (!hff2(s1, si, sk)) {
          keep3(s, si, sii, sk, ski);
          proj_plane(s1, si, sk, v);
        } else {
          keep2(s, si, sii);
          proj_line(s1, si, v);
        }
      } else if (dot_total == 2) {
        s->nvrtx = 3;
        if (!line3) {
          k = 2; i = 1; j = 0;
        } else if (!line4) {
          k = 1; i = 0; j = 2;
        } else {
          k = 0; i = 2; j = 1;
        }
        LOAD_IJK();
        if (!hff2(s1, sj, sk)) {
          if (!hff2(s1, sk, sj)) {
            keep3(s, sj, sji, sk, ski);
            proj_plane(s1, sj, sk, v);
          } else if (!hff2(s1, sk, si)) {
            keep3(s, si, sii, sk, ski);
            proj_plane(s1, sk, si, v);
          } else {
            keep2(s, sk, ski);
            proj_line(s1, sk, v);
          }
        } else if (!hff2(s1, sj, si)) {
          keep3(s, si, sii, sj, sji);
          proj_plane(s1, si, sj, v);
        } else {
          keep2(s, sj, sji);
          proj_line(s1, sj, v);
        }
      }
      break;
    default:
      break;
  }
#undef LOAD_IJK
}

/* subalgorithm: openGJK.c:641-656 */
static void subalgorithm(osimplex* s, oreal* v) {
  switch (s->nvrtx) {
    case 4: S3D(s, v); break;
    case 3: S2D(s, v); break;
    case 2: S1D(s, v); break;
    default: break;
  }
}

/* support: openGJK.c:615-639.  Incumbent is the previous support point;
 * strict '>' so the lowest index among strict improvers wins. */
static void support(opoly* b, const oreal* v) {
  oreal maxs = v_dot(b->s, v);
  int better = -1;
  for (int i = 0; i < b->numpoints; ++i) {
    const oreal d = v_dot(&b->coord[i * 3], v);
    if (d > maxs) {
      maxs = d;
      better = i;
    }
  }
  if (better != -1) {
    v_copy(b->s, &b->coord[better * 3]);
    b->s_idx = better;
  }
}

/* ----------------------------------------------- witnesses W0D..W3D */

static inline const oreal* vtx(const opoly* b, int i) { return &b->coord[i * 3]; }

/* W0D: openGJK.c:658-666 */
static void W0D(const opoly* b1, const opoly* b2, osimplex* s) {
  const oreal* w00 = vtx(b1, s->vrtx_idx[0][0]);
  const oreal* w01 = vtx(b2, s->vrtx_idx[0][1]);
  for (int t = 0; t < 3; t++) {
    s->witnesses[0][t] = w00[t];
    s->witnesses[1][t] = w01[t];
  }
}

/* W1D: openGJK.c:668-700 (det == 0 falls through, as in the reference) */
static void W1D(const opoly* b1, const opoly* b2, osimplex* s) {
  oreal pq[3], po[3];
  const oreal* p = s->vrtx[0];
  const oreal* q = s->vrtx[1];
  for (int t = 0; t < 3; t++) {
    pq[t] = q[t] - p[t];
    po[t] = -p[t];
  }
  const oreal det = v_dot(pq, pq);
  if (det == 0.0) W0D(b1, b2, s);
  const oreal a1 = v_dot(pq, po) / det;
  const oreal a0 = (oreal)(1.0 - (double)a1);
  const oreal* w00 = vtx(b1, s->vrtx_idx[0][0]);
  const oreal* w01 = vtx(b2, s->vrtx_idx[0][1]);
  const oreal* w10 = vtx(b1, s->vrtx_idx[1][0]);
  const oreal* w11 = vtx(b2, s->vrtx_idx[1][1]);
  for (int t = 0; t < 3; t++) {
    s->witnesses[0][t] = w00[t] * a0 + w10[t] * a1;
    s->witnesses[1][t] = w01[t] * a0 + w11[t] * a1;
  }
}

/* W2D: openGJK.c:702-795 */
static void W2D(const opoly* b1, const opoly* b2, osimplex* s) {
  oreal pq[3], pr[3], po[3];
  const oreal* p = s->vrtx[0];
  const oreal* q = s->vrtx[1];
  const oreal* r = s->vrtx[2];
  for (int t = 0; t < 3; t++) {
    pq[t] = q[t] - p[t];
    pr[t] = r[t] - p[t];
    po[t] = -p[t];
  }
  const oreal T00 = v_dot(pq, pq);
  const oreal T01 = v_dot(pq, pr);
  const oreal T11 = v_dot(pr, pr);
  const oreal det = T00 * T11 - T01 * T01;
  if (det == 0.0) W1D(b1, b2, s);
  const oreal b0 = v_dot(pq, po);
  const oreal bb1 = v_dot(pr, po);
  const oreal I00 = T11 / det;
  const oreal I01 = -T01 / det;
  const oreal I11 = T00 / det;
  const oreal a1 = I00 * b0 + I01 * bb1;
  const oreal a2 = I01 * b0 + I11 * bb1;
  const oreal a0 = (oreal)(1.0 - (double)a1 - (double)a2);

  if (a0 < O_EPS) {
    s->nvrtx = 2;
    slot_move(s, 0, 2);
    W1D(b1, b2, s);
  } else if (a1 < O_EPS) {
    s->nvrtx = 2;
    slot_move(s, 1, 2);
    W1D(b1, b2, s);
  } else if (a2 < O_EPS) {
    s->nvrtx = 2;
    W1D(b1, b2, s);
  }

  const oreal* w00 = vtx(b1, s->vrtx_idx[0][0]);
  const oreal* w01 = vtx(b2, s->vrtx_idx[0][1]);
  const oreal* w10 = vtx(b1, s->vrtx_idx[1][0]);
  const oreal* w11 = vtx(b2, s->vrtx_idx[1][1]);
  const oreal* w20 = vtx(b1, s->vrtx_idx[2][0]);
  const oreal* w21 = vtx(b2, s->vrtx_idx[2][1]);
  for (int t = 0; t < 3; t++) {
    s->witnesses[0][t] = w00[t] * a0 + w10[t] * a1 + w20[t] * a2;
    s->witnesses[1][t] = w01[t] * a0 + w11[t] * a1 + w21[t] * a2;
  }
}

/* W3D: openGJK.c:797-929 */
static void W3D(const opoly* b1, const opoly* b2, osimplex* s) {
  oreal pq[3], pr[3], ps[3], po[3];
  const oreal* p = s->vrtx[0];
  const oreal* q = s->vrtx[1];
  const oreal* r = s->vrtx[2];
  const oreal* sv = s->vrtx[3];
  for (int t = 0; t < 3; t++) {
    pq[t] = q[t] - p[t];
    pr[t] = r[t] - p[t];
    ps[t] = sv[t] - p[t];
    po[t] = -p[t];
  }
  const oreal T00 = v_dot(pq, pq);
  const oreal T01 = v_dot(pq, pr);
  const oreal T02 = v_dot(pq, ps);
  const oreal T11 = v_dot(pr, pr);
  const oreal T12 = v_dot(pr, ps);
  const oreal T22 = v_dot(ps, ps);
  const oreal det00 = T11 * T22 - T12 * T12;
  const oreal det01 = T01 * T22 - T02 * T12;
  const oreal det02 = T01 * T12 - T02 * T11;
  const oreal det = T00 * det00 - T01 * det01 + T02 * det02;
  if (det == 0.0) W2D(b1, b2, s);
  const oreal b0 = v_dot(pq, po);
  const oreal bb1 = v_dot(pr, po);
  const oreal bb2 = v_dot(ps, po);
  const oreal det11 = T00 * T22 - T02 * T02;
  const oreal det12 = T00 * T12 - T01 * T02;
  const oreal det22 = T00 * T11 - T01 * T01;
  const oreal I00 = det00 / det;
  const oreal I01 = -det01 / det;
  const oreal I02 = det02 / det;
  const oreal I11 = det11 / det;
  const oreal I12 = -det12 / det;
  const oreal I22 = det22 / det;
  const oreal a1 = I00 * b0 + I01 * bb1 + I02 * bb2;
  const oreal a2 = I01 * b0 + I11 * bb1 + I12 * bb2;
  const oreal a3 = I02 * b0 + I12 * bb1 + I22 * bb2;
  const oreal a0 = (oreal)(1.0 - (double)a1 - (double)a2 - (double)a3);

  if (a0 < O_EPS) {
    s->nvrtx = 3;
    slot_move(s, 0, 3);
    W2D(b1, b2, s);
  } else if (a1 < O_EPS) {
    s->nvrtx = 3;
    slot_move(s, 1, 3);
    W2D(b1, b2, s);
  } else if (a2 < O_EPS) {
    s->nvrtx = 3;
    slot_move(s, 2, 3);
    W2D(b1, b2, s);
  } else if (a3 < O_EPS) {
    s->nvrtx = 3;
    W2D(b1, b2, s);
  }

  const oreal* w00 = vtx(b1, s->vrtx_idx[0][0]);
  const oreal* w01 = vtx(b2, s->vrtx_idx[0][1]);
  const oreal* w10 = vtx(b1, s->vrtx_idx[1][0]);
  const oreal* w11 = vtx(b2, s->vrtx_idx[1][1]);
  const oreal* w20 = vtx(b1, s->vrtx_idx[2][0]);
  const oreal* w21 = vtx(b2, s->vrtx_idx[2][1]);
  const oreal* w30 = vtx(b1, s->vrtx_idx[3][0]);
  const oreal* w31 = vtx(b2, s->vrtx_idx[3][1]);
  for (int t = 0; t < 3; t++) {
    s->witnesses[0][t] = w00[t] * a0 + w10[t] * a1 + w20[t] * a2 + w30[t] * a3;
    s->witnesses[1][t] = w01[t] * a0 + w11[t] * a1 + w21[t] * a2 + w31[t] * a3;
  }
}

/* compute_witnesses: openGJK.c:931-950 */
static void witnesses(const opoly* b1, const opoly* b2, osimplex* s) {
  switch (s->nvrtx) {
    case 4: W3D(b1, b2, s); break;
    case 3: W2D(b1, b2, s); break;
    case 2: W1D(b1, b2, s); break;
    case 1: W0D(b1, b2, s); break;
    default: break;
  }
}

/* ----------------------------------------------- GJK driver (a6) */

/* compute_minimum_distance: openGJK.c:952-1056 */
oreal oracle_gjk(opoly bd1, opoly bd2, osimplex* s, ostats* st) {
  unsigned int k = 0;
  const unsigned int mk = 25;
  const oreal eps_rel = EPS_REL;
  const oreal eps_tot = EPS_TOT;
  const oreal eps_rel2 = eps_rel * eps_rel;
  oreal w[3], v[3], vminus[3];
  int w_idx[2];
  oreal norm2Wmax = 0;

  v[0] = bd1.coord[0] - bd2.coord[0];
  v[1] = bd1.coord[1] - bd2.coord[1];
  v[2] = bd1.coord[2] - bd2.coord[2];

  s->nvrtx = 1;
  v_copy(s->vrtx[0], v);
  s->vrtx_idx[0][0] = 0;
  s->vrtx_idx[0][1] = 0;
  v_copy(bd1.s, bd1.coord);
  bd1.s_idx = 0;
  v_copy(bd2.s, bd2.coord);
  bd2.s_idx = 0;

  do {
    k++;
    for (int t = 0; t < 3; ++t) vminus[t] = -v[t];
    support(&bd1, vminus);
    support(&bd2, v);
    v_sub(w, bd1.s, bd2.s);
    w_idx[0] = bd1.s_idx;
    w_idx[1] = bd2.s_idx;

    const oreal gap = (v_norm2(v) - v_dot(v, w));
    if (gap <= (eps_rel * v_norm2(v)) || gap < eps_tot) break;
    if (v_norm2(v) < eps_rel2) break;

    const int i = s->nvrtx;
    v_copy(s->vrtx[i], w);
    s->vrtx_idx[i][0] = w_idx[0];
    s->vrtx_idx[i][1] = w_idx[1];
    s->nvrtx++;

    subalgorithm(s, v);

    for (int jj = 0; jj < s->nvrtx; jj++) {
      const oreal tn = v_norm2(s->vrtx[jj]);
      if (tn > norm2Wmax) norm2Wmax = tn;
    }
    if ((v_norm2(v) <= (eps_tot * eps_tot * norm2Wmax))) break;
  } while ((s->nvrtx != 4) && (k != mk));

  if (st) {
    st->gjk_iters = (int)k;
    st->sweeps_a += (int)k;
    st->sweeps_b += (int)k;
  }
  witnesses(&bd1, &bd2, s);
  return O_SQRT(v_norm2(v));
}

/* ----------------------------------------------- EPA (a14 - a22) */

/* EPAFace / EPAPolytope / EPAEdge: EPA.c:56-79 */
typedef struct {
  int v[3];
  int v_idx[3][2];
  oreal normal[3];
  oreal distance;
  bool valid;
} oface;

typedef struct {
  oreal vertices[EPA_MAX_VERTS][3];
  int vertex_indices[EPA_MAX_VERTS][2];
  int num_vertices;
  oface faces[EPA_MAX_FACES];
  int max_face_index;
} opolytope;

typedef struct {
  int v1, v2;
  int v_idx1[2], v_idx2[2];
  bool valid;
} oedge;

static inline oreal bcoord(const opoly* b, int i, int c) { return b->coord[i * 3 + c]; }

/* compute_face_normal_distance: EPA.c:92-129 */
static void face_normal(opolytope* P, int f) {
  oface* F = &P->faces[f];
  const oreal* v0 = P->vertices[F->v[0]];
  const oreal* v1 = P->vertices[F->v[1]];
  const oreal* v2 = P->vertices[F->v[2]];
  oreal e0[3], e1[3];
  v_sub(e0, v1, v0);
  v_sub(e1, v2, v0);
  v_cross(e0, e1, F->normal);
  const oreal nsq = v_norm2(F->normal);
  if (nsq > O_EPS * O_EPS) {
    const oreal nn = O_SQRT(nsq);
    for (int i = 0; i < 3; i++) F->normal[i] /= nn;
    F->distance = v_dot(F->normal, v0);
    if (F->distance < 0) {
      for (int i = 0; i < 3; i++) F->normal[i] = -F->normal[i];
      F->distance = -F->distance;
    }
  } else {
    F->valid = false;
    F->distance = (oreal)1e10;
  }
}

/* is_face_visible: EPA.c:132-140 */
static bool face_visible(const opolytope* P, int f, const oreal* p) {
  const oface* F = &P->faces[f];
  if (!F->valid) return false;
  oreal d[3];
  v_sub(d, p, P->vertices[F->v[0]]);
  return v_dot(F->normal, d) > O_EPS;
}

/* Flip winding if the normal points toward the centroid (EPA.c:204-231,
 * EPA.c:791-819). */
static void fix_winding(opolytope* P, int f, const oreal* centroid) {
  oface* F = &P->faces[f];
  const oreal* v0 = P->vertices[F->v[0]];
  const oreal* v1 = P->vertices[F->v[1]];
  const oreal* v2 = P->vertices[F->v[2]];
  oreal e0[3], e1[3], n[3], tc[3];
  v_sub(e0, v1, v0);
  v_sub(e1, v2, v0);
  v_cross(e0, e1, n);
  v_sub(tc, centroid, v0);
  if (v_dot(n, tc) > 0) {
    int t = F->v[1];
    F->v[1] = F->v[2];
    F->v[2] = t;
    int t0 = F->v_idx[1][0], t1 = F->v_idx[1][1];
    F->v_idx[1][0] = F->v_idx[2][0];
    F->v_idx[1][1] = F->v_idx[2][1];
    F->v_idx[2][0] = t0;
    F->v_idx[2][1] = t1;
  }
}

/* init_epa_polytope: EPA.c:144-235 */
static void epa_init(opolytope* P, const osimplex* s, oreal* centroid) {
  static const int tet[4][3] = {{0, 1, 2}, {0, 3, 1}, {0, 2, 3}, {1, 3, 2}};
  memset(P->faces, 0, sizeof(P->faces));
  P->num_vertices = 4;
  for (int i = 0; i < 4; i++) {
    v_copy(P->vertices[i], s->vrtx[i]);
    P->vertex_indices[i][0] = s->vrtx_idx[i][0];
    P->vertex_indices[i][1] = s->vrtx_idx[i][1];
  }
  centroid[0] = centroid[1] = centroid[2] = 0.0f;
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 3; j++) centroid[j] += P->vertices[i][j] * 0.25f;
  for (int f = 0; f < 4; f++) {
    for (int c = 0; c < 3; c++) {
      P->faces[f].v[c] = tet[f][c];
      P->faces[f].v_idx[c][0] = s->vrtx_idx[tet[f][c]][0];
      P->faces[f].v_idx[c][1] = s->vrtx_idx[tet[f][c]][1];
    }
    P->faces[f].valid = true;
  }
  for (int f = 0; f < 4; f++) fix_winding(P, f, centroid);
  P->max_face_index = 4;
}

/* compute_barycentric_origin: EPA.c:238-304 */
static void bary_origin(const oreal* v0, const oreal* v1, const oreal* v2, oreal* a0,
                        oreal* a1, oreal* a2) {
  oreal e0[3], e1[3];
  v_sub(e0, v1, v0);
  v_sub(e1, v2, v0);
  const oreal d00 = v_dot(e0, e0);
  const oreal d01 = v_dot(e0, e1);
  const oreal d11 = v_dot(e1, e1);
  const oreal d20 = -v_dot(v0, e0);
  const oreal d21 = -v_dot(v0, e1);
  const oreal denom = d00 * d11 - d01 * d01;
  if (O_FABS(denom) < O_EPS) {
    *a0 = *a1 = *a2 = (oreal)1.0 / (oreal)3.0;
    return;
  }
  const oreal inv = (oreal)1.0 / denom;
  const oreal u = (d11 * d20 - d01 * d21) * inv;
  const oreal v = (d00 * d21 - d01 * d20) * inv;
  const oreal w = (oreal)1.0 - u - v;
  if (w < 0) {
    oreal e12[3];
    v_sub(e12, v2, v1);
    oreal t = -v_dot(v1, e12) / v_dot(e12, e12);
    t = O_FMAX((oreal)0.0, O_FMIN((oreal)1.0, t));
    *a0 = 0;
    *a1 = (oreal)1.0 - t;
    *a2 = t;
  } else if (u < 0) {
    oreal t = -v_dot(v0, e1) / v_dot(e1, e1);
    t = O_FMAX((oreal)0.0, O_FMIN((oreal)1.0, t));
    *a0 = (oreal)1.0 - t;
    *a1 = 0;
    *a2 = t;
  } else if (v < 0) {
    oreal t = -v_dot(v0, e0) / v_dot(e0, e0);
    t = O_FMAX((oreal)0.0, O_FMIN((oreal)1.0, t));
    *a0 = (oreal)1.0 - t;
    *a1 = t;
    *a2 = 0;
  } else {
    *a0 = w;
    *a1 = u;
    *a2 = v;
  }
}

/* support_epa: EPA.c:307-344.  Starts from -1e10 with strict '>', so the
 * lowest index wins ties on both bodies. */
static void support_epa(const opoly* b1, const opoly* b2, const oreal* d, oreal* out,
                        int* out_idx, ostats* st) {
  oreal m1 = -1e10f, m2 = -1e10f;
  int i1 = -1, i2 = -1;
  for (int i = 0; i < b1->numpoints; i++) {
    const oreal s = bcoord(b1, i, 0) * d[0] + bcoord(b1, i, 1) * d[1] + bcoord(b1, i, 2) * d[2];
    if (s > m1) {
      m1 = s;
      i1 = i;
    }
  }
  for (int i = 0; i < b2->numpoints; i++) {
    const oreal s = bcoord(b2, i, 0) * d[0] + bcoord(b2, i, 1) * d[1] + bcoord(b2, i, 2) * d[2];
    if (-s > m2) {
      m2 = -s;
      i2 = i;
    }
  }
  if (i1 >= 0 && i2 >= 0) {
    for (int i = 0; i < 3; i++) out[i] = bcoord(b1, i1, i) - bcoord(b2, i2, i);
    out_idx[0] = i1;
    out_idx[1] = i2;
  }
  if (st) {
    st->sweeps_a++;
    st->sweeps_b++;
  }
}

/* set_contact_normal: EPA.c:350-360 */
static void contact_normal_from(const oreal* w1, const oreal* w2, oreal* n) {
  oreal d[3] = {w2[0] - w1[0], w2[1] - w1[1], w2[2] - w1[2]};
  const oreal nn = O_SQRT(v_norm2(d));
  if (nn > O_EPS) {
    n[0] = d[0] / nn;
    n[1] = d[1] / nn;
    n[2] = d[2] / nn;
  } else {
    n[0] = 1.0f;
    n[1] = 0.0f;
    n[2] = 0.0f;
  }
}

/* Witnesses at a support pair, then a normal from them (EPA.c:411-417). */
static void zero_depth_exit(const opoly* b1, const opoly* b2, osimplex* s, oreal* distance,
                            oreal* normal, const int* idx) {
  *distance = 0.0f;
  for (int c = 0; c < 3; ++c) {
    s->witnesses[0][c] = bcoord(b1, idx[0], c);
    s->witnesses[1][c] = bcoord(b2, idx[1], c);
  }
  contact_normal_from(s->witnesses[0], s->witnesses[1], normal);
}

/* Is p farther than eps from every vertex in s[0..n)?  (EPA.c:453-463) */
static bool is_new_point(const osimplex* s, const oreal* p) {
  const oreal eps_sq = O_EPS * O_EPS;
  for (int t = 0; t < s->nvrtx; ++t) {
    const oreal dx = p[0] - s->vrtx[t][0];
    const oreal dy = p[1] - s->vrtx[t][1];
    const oreal dz = p[2] - s->vrtx[t][2];
    const oreal d2 = dx * dx + dy * dy + dz * dz;
    if (d2 < eps_sq) return false;
  }
  return true;
}

static void append_vertex(osimplex* s, const oreal* p, const int* idx) {
  const int t = s->nvrtx;
  v_copy(s->vrtx[t], p);
  s->vrtx_idx[t][0] = idx[0];
  s->vrtx_idx[t][1] = idx[1];
  s->nvrtx = t + 1;
}

/* Blend the closest face's original vertices (EPA.c:637-650). */
static void face_result(const opoly* b1, const opoly* b2, const opolytope* P, const oface* F,
                        osimplex* s, oreal* distance, oreal* normal, oreal dist) {
  oreal a0, a1, a2;
  bary_origin(P->vertices[F->v[0]], P->vertices[F->v[1]], P->vertices[F->v[2]], &a0, &a1, &a2);
  for (int i = 0; i < 3; i++) {
    s->witnesses[0][i] = bcoord(b1, F->v_idx[0][0], i) * a0 + bcoord(b1, F->v_idx[1][0], i) * a1 +
                         bcoord(b1, F->v_idx[2][0], i) * a2;
    s->witnesses[1][i] = bcoord(b2, F->v_idx[0][1], i) * a0 + bcoord(b2, F->v_idx[1][1], i) * a1 +
                         bcoord(b2, F->v_idx[2][1], i) * a2;
    normal[i] = F->normal[i];
  }
  *distance = -dist;
}

/* Closest valid face with distance in [0, 1e10) (EPA.c:608-617). */
static int closest_face(const opolytope* P, oreal* dist) {
  int best = -1;
  oreal bd = 1e10f;
  for (int i = 0; i < P->max_face_index; ++i) {
    if (!P->faces[i].valid) continue;
    if (P->faces[i].distance >= 0.0f && P->faces[i].distance < bd) {
      bd = P->faces[i].distance;
      best = i;
    }
  }
  *dist = bd;
  return best;
}

/* computeCollisionInformation: EPA.c:362-864 */
void oracle_epa(const opoly* bd1, const opoly* bd2, osimplex* simplex, oreal* distance,
                oreal contact_normal[3], ostats* st) {
  if (*distance > O_EPS) {
    contact_normal_from(simplex->witnesses[0], simplex->witnesses[1], contact_normal);
    return;
  }

  if (simplex->nvrtx != 4) {
    oreal nv[3];
    int nidx[2];
    if (simplex->nvrtx == 1) { /* EPA.c:378-419 */
      support_epa(bd1, bd2, simplex->vrtx[0], nv, nidx, st);
      if (is_new_point(simplex, nv)) {
        append_vertex(simplex, nv, nidx);
      } else {
        zero_depth_exit(bd1, bd2, simplex, distance, contact_normal, nidx);
        return;
      }
    }
    if (simplex->nvrtx == 2) { /* EPA.c:420-484 */
      oreal dir[3], edge[3];
      v_sub(edge, simplex->vrtx[1], simplex->vrtx[0]);
      oreal axis[3] = {1.0f, 0.0f, 0.0f};
      const oreal en = O_SQRT(v_norm2(edge));
      if (en > O_EPS && O_FABS(edge[0]) > 0.9f * en) {
        axis[0] = 0.0f;
        axis[1] = 1.0f;
        axis[2] = 0.0f;
      }
      v_cross(edge, axis, dir);
      if (v_norm2(dir) < O_EPS) {
        axis[0] = 0.0f;
        axis[1] = 0.0f;
        axis[2] = 1.0f;
        v_cross(edge, axis, dir);
      }
      support_epa(bd1, bd2, dir, nv, nidx, st);
      if (is_new_point(simplex, nv)) {
        append_vertex(simplex, nv, nidx);
      } else {
        zero_depth_exit(bd1, bd2, simplex, distance, contact_normal, nidx);
        return;
      }
    }
    if (simplex->nvrtx == 3) { /* EPA.c:485-569 */
      oreal dir[3], e0[3], e1[3];
      v_sub(e0, simplex->vrtx[1], simplex->vrtx[0]);
      v_sub(e1, simplex->vrtx[2], simplex->vrtx[0]);
      v_cross(e0, e1, dir);
      support_epa(bd1, bd2, dir, nv, nidx, st);
      if (is_new_point(simplex, nv)) {
        append_vertex(simplex, nv, nidx);
      } else {
        dir[0] = -dir[0];
        dir[1] = -dir[1];
        dir[2] = -dir[2];
      }
      if (simplex->nvrtx == 3) {
        support_epa(bd1, bd2, dir, nv, nidx, st);
        if (is_new_point(simplex, nv)) {
          append_vertex(simplex, nv, nidx);
        } else {
          zero_depth_exit(bd1, bd2, simplex, distance, contact_normal, nidx);
          return;
        }
      }
    }
    if (simplex->nvrtx != 4) { /* EPA.c:572-582 */
      const int best = simplex->nvrtx > 0 ? simplex->nvrtx - 1 : 0;
      zero_depth_exit(bd1, bd2, simplex, distance, contact_normal, simplex->vrtx_idx[best]);
      return;
    }
  }

  opolytope P;
  oreal centroid[3];
  epa_init(&P, simplex, centroid);

  const int max_iterations = 64;
  const oreal tolerance = EPS_TOT;
  int iteration = 0;

  while (iteration < max_iterations && P.num_vertices < EPA_MAX_VERTS - 1) {
    iteration++;
    for (int i = 0; i < P.max_face_index; ++i)
      if (P.faces[i].valid) face_normal(&P, i);

    oreal cdist;
    const int cf = closest_face(&P, &cdist);
    if (cf < 0) break;
    const oface* closest = &P.faces[cf];

    oreal nv[3];
    int nidx[2];
    support_epa(bd1, bd2, P.faces[cf].normal, nv, nidx, st);

    const oreal dist_to_new = v_dot(P.faces[cf].normal, nv);
    const oreal improvement = dist_to_new - cdist;
    if (improvement < tolerance) { /* converged, EPA.c:635-652 */
      face_result(bd1, bd2, &P, closest, simplex, distance, contact_normal, cdist);
      break;
    }

    bool dup = false; /* EPA.c:654-684 */
    {
      const oreal eps_sq = O_EPS * O_EPS;
      for (int i = 0; i < P.num_vertices; i++) {
        const oreal dx = nv[0] - P.vertices[i][0];
        const oreal dy = nv[1] - P.vertices[i][1];
        const oreal dz = nv[2] - P.vertices[i][2];
        if (dx * dx + dy * dy + dz * dz < eps_sq) {
          dup = true;
          break;
        }
      }
    }
    if (dup) {
      face_result(bd1, bd2, &P, closest, simplex, distance, contact_normal, cdist);
      break;
    }

    /* add the vertex and update the running centroid (EPA.c:686-699) */
    const int nid = P.num_vertices;
    v_copy(P.vertices[nid], nv);
    P.vertex_indices[nid][0] = nidx[0];
    P.vertex_indices[nid][1] = nidx[1];
    P.num_vertices++;
    const oreal inv_n = (oreal)1.0 / (oreal)P.num_vertices;
    for (int i = 0; i < 3; i++) centroid[i] += (nv[i] - centroid[i]) * inv_n;

    /* horizon: edges of faces removed now; shared edges cancel (EPA.c:700-759) */
    oedge edges[EPA_MAX_FACES * 3];
    int ne = 0;
    for (int f = 0; f < P.max_face_index; f++) {
      if (!P.faces[f].valid) continue;
      if (!face_visible(&P, f, nv)) continue;
      for (int c = 0; c < 3; c++) {
        if (ne < EPA_MAX_FACES * 3) {
          const int a = c, b = (c + 1) % 3;
          edges[ne].v1 = P.faces[f].v[a];
          edges[ne].v2 = P.faces[f].v[b];
          edges[ne].v_idx1[0] = P.faces[f].v_idx[a][0];
          edges[ne].v_idx1[1] = P.faces[f].v_idx[a][1];
          edges[ne].v_idx2[0] = P.faces[f].v_idx[b][0];
          edges[ne].v_idx2[1] = P.faces[f].v_idx[b][1];
          edges[ne].valid = true;
          ne++;
        }
      }
      P.faces[f].valid = false;
    }
    for (int i = 0; i < ne; i++) {
      if (!edges[i].valid) continue;
      for (int j = i + 1; j < ne; j++) {
        if (!edges[j].valid) continue;
        if ((edges[i].v1 == edges[j].v1 && edges[i].v2 == edges[j].v2) ||
            (edges[i].v1 == edges[j].v2 && edges[i].v2 == edges[j].v1)) {
          edges[i].valid = false;
          edges[j].valid = false;
        }
      }
    }

    /* fan new faces into the first free slots (EPA.c:761-825) */
    for (int i = 0; i < ne; i++) {
      if (!edges[i].valid) continue;
      int slot = -1;
      for (int j = 0; j < EPA_MAX_FACES; j++) {
        if (!P.faces[j].valid) {
          slot = j;
          break;
        }
      }
      if (slot < 0 || slot >= EPA_MAX_FACES) {
        if (st) st->face_overflow++; /* horizon edges left without a face */
        break;
      }
      oface* F = &P.faces[slot];
      F->v[0] = edges[i].v1;
      F->v[1] = edges[i].v2;
      F->v[2] = nid;
      F->v_idx[0][0] = edges[i].v_idx1[0];
      F->v_idx[0][1] = edges[i].v_idx1[1];
      F->v_idx[1][0] = edges[i].v_idx2[0];
      F->v_idx[1][1] = edges[i].v_idx2[1];
      F->v_idx[2][0] = nidx[0];
      F->v_idx[2][1] = nidx[1];
      F->valid = true;
      fix_winding(&P, slot, centroid);
      if (slot >= P.max_face_index) P.max_face_index = slot + 1;
    }
  }

  if (st) st->epa_iters = iteration;

  if (iteration >= max_iterations) { /* cap fallback, EPA.c:828-863 */
    for (int i = 0; i < P.max_face_index; ++i)
      if (P.faces[i].valid) face_normal(&P, i);
    oreal cdist;
    const int cf = closest_face(&P, &cdist);
    if (cf >= 0)
      face_result(bd1, bd2, &P, &P.faces[cf], simplex, distance, contact_normal, cdist);
  }
}

/* ----------------------------------------------- batch driver */

typedef struct {
  int mode, n;
  const oreal* coords;
  const int64_t* off;
  const int32_t* cnt;
  const int32_t* pa;
  const int32_t* pb;
  osimplex* simplices;
  oreal* distances;
  oreal* normals;
  ostats* stats;
  int next; /* next unclaimed pair (atomic) */
  const float* coords32; /* float pool promoted per pair (oracle_batch_f32), or NULL */
} batch_job;

/* Per-worker buffer of the two bodies of one pair, promoted from float. */
typedef struct {
  oreal* v;
  int64_t cap;
} pair_buf;

static const oreal* promote(pair_buf* B, const float* src, int64_t n, int64_t at) {
  for (int64_t k = 0; k < 3 * n; k++) B->v[3 * at + k] = (oreal)src[k];
  return B->v + 3 * at;
}

static void run_pair_buf(batch_job* J, int i, pair_buf* B) {
  opoly b1, b2;
  memset(&b1, 0, sizeof(b1));
  memset(&b2, 0, sizeof(b2));
  b1.numpoints = J->cnt[J->pa[i]];
  b2.numpoints = J->cnt[J->pb[i]];
  if (J->coords32) {
    const int64_t n1 = b1.numpoints > 0 ? b1.numpoints : 0, n2 = b2.numpoints > 0 ? b2.numpoints : 0;
    if (n1 + n2 > B->cap) {
      free(B->v);
      B->cap = 2 * (n1 + n2);
      B->v = (oreal*)malloc((size_t)(3 * B->cap) * sizeof(oreal));
    }
    b1.coord = promote(B, J->coords32 + 3 * J->off[J->pa[i]], n1, 0);
    b2.coord = promote(B, J->coords32 + 3 * J->off[J->pb[i]], n2, n1);
  } else {
    b1.coord = J->coords + 3 * J->off[J->pa[i]];
    b2.coord = J->coords + 3 * J->off[J->pb[i]];
  }
  ostats* st = J->stats ? &J->stats[i] : NULL;
  if (st) memset(st, 0, sizeof(*st));
  if (J->mode & 1) {
    memset(&J->simplices[i], 0, sizeof(osimplex));
    J->distances[i] = oracle_gjk(b1, b2, &J->simplices[i], st);
  }
  if (J->mode & 2) {
    oracle_epa(&b1, &b2, &J->simplices[i], &J->distances[i], &J->normals[3 * i], st);
  }
}

static void* batch_worker(void* arg) {
  batch_job* J = (batch_job*)arg;
  pair_buf B = {NULL, 0};
  const int chunk = 32;
  for (;;) {
    const int start = __atomic_fetch_add(&J->next, chunk, __ATOMIC_RELAXED);
    if (start >= J->n) break;
    const int end = start + chunk < J->n ? start + chunk : J->n;
    for (int i = start; i < end; i++) run_pair_buf(J, i, &B);
  }
  free(B.v);
  return NULL;
}

static void run_batch(batch_job* J, int nthreads) {
  if (nthreads <= 1) {
    batch_worker(J);
    return;
  }
  if (nthreads > 256) nthreads = 256;
  pthread_t th[256];
  int started = 0;
  for (int t = 0; t < nthreads; t++) {
    if (pthread_create(&th[t], NULL, batch_worker, J) != 0) break;
    started++;
  }
  if (started == 0) {
    batch_worker(J);
    return;
  }
  for (int t = 0; t < started; t++) pthread_join(th[t], NULL);
}

void oracle_batch(int mode, int n, const oreal* coords, const int64_t* off, const int32_t* cnt,
                  const int32_t* pair_a, const int32_t* pair_b, osimplex* simplices,
                  oreal* distances, oreal* normals, ostats* stats, int nthreads) {
  if (n <= 0) return;
  batch_job J = {mode, n, coords, off, cnt, pair_a, pair_b, simplices, distances, normals, stats,
                 0, NULL};
  run_batch(&J, nthreads);
}

void oracle_batch_f32(int mode, int n, const float* coords, const int64_t* off,
                      const int32_t* cnt, const int32_t* pair_a, const int32_t* pair_b,
                      osimplex* simplices, oreal* distances, oreal* normals, ostats* stats,
                      int nthreads) {
  if (n <= 0) return;
  batch_job J = {mode, n, NULL, off, cnt, pair_a, pair_b, simplices, distances, normals, stats,
                 0, coords};
  run_batch(&J, nthreads);
}

int oracle_real_size(void) { return (int)sizeof(oreal); }
