/*
 * opengjk-amd: data layouts shared by the C ABI, the C++ drop-in API and the
 * HIP kernels.
 *
 * These are the caller-visible layouts of the reference library
 * (GJK/common.h:46-89).  They are kept byte-for-byte identical so that a
 * program built against the reference headers links and runs against this
 * library unchanged:
 *
 *   precision  sizeof(gkPolytope)  offsets {numpoints,s,s_idx,coord}
 *   FP32       32                  {0, 4, 16, 24}
 *   FP64       48                  {0, 8, 32, 40}
 *
 *   precision  sizeof(gkSimplex)   offsets {nvrtx,vrtx,vrtx_idx,witnesses}
 *   FP32       108                 {0, 4, 52, 84}
 *   FP64       184                 {0, 8, 104, 136}
 *
 * The reference selects FP32 by an unconditional `#define USE_32BITS`
 * (GJK/common.h:44).  Here FP32 is also the default; define
 * OGJK_USE_64BITS before including this header to get the FP64 layouts.
 */
#ifndef OPENGJK_AMD_COMMON_H
#define OPENGJK_AMD_COMMON_H

#include <float.h>
#include <stddef.h>

#ifdef OGJK_USE_64BITS
#define gkFloat double
#define gkEpsilon DBL_EPSILON
#else
#ifndef USE_32BITS
#define USE_32BITS
#endif
#define gkFloat float
#define gkEpsilon FLT_EPSILON
#endif

/* A convex polytope given by its vertex cloud (GJK/common.h:68-78).
 * `coord` is AoS [x0,y0,z0,x1,...] and stays owned by the caller.
 * `s` / `s_idx` are support-point scratch; this library never reads them. */
typedef struct gkPolytope {
  int numpoints;
  gkFloat s[3];
  int s_idx;
  gkFloat* coord;
} gkPolytope;

/* GJK simplex and result record (GJK/common.h:84-89).
 * vrtx[k] = A[vrtx_idx[k][0]] - B[vrtx_idx[k][1]];
 * witnesses[0] lies on body 1, witnesses[1] on body 2. */
typedef struct gkSimplex {
  int nvrtx;
  gkFloat vrtx[4][3];
  int vrtx_idx[4][2];
  gkFloat witnesses[2][3];
} gkSimplex;

/* Pair of polytope indices into a pool (GJK/gpu/openGJK.h:309-312). */
#ifdef __cplusplus
struct gkCollisionPair {
  int idx1;
  int idx2;
};
#else
typedef struct gkCollisionPair {
  int idx1;
  int idx2;
} gkCollisionPair;
#endif

#ifdef __cplusplus
#define OGJK_STATIC_ASSERT(c, m) static_assert(c, m)
#else
#define OGJK_STATIC_ASSERT(c, m) _Static_assert(c, m)
#endif

#ifdef OGJK_USE_64BITS
OGJK_STATIC_ASSERT(sizeof(gkPolytope) == 48, "gkPolytope FP64 layout");
OGJK_STATIC_ASSERT(offsetof(gkPolytope, s) == 8, "gkPolytope.s");
OGJK_STATIC_ASSERT(offsetof(gkPolytope, s_idx) == 32, "gkPolytope.s_idx");
OGJK_STATIC_ASSERT(offsetof(gkPolytope, coord) == 40, "gkPolytope.coord");
OGJK_STATIC_ASSERT(sizeof(gkSimplex) == 184, "gkSimplex FP64 layout");
OGJK_STATIC_ASSERT(offsetof(gkSimplex, vrtx) == 8, "gkSimplex.vrtx");
OGJK_STATIC_ASSERT(offsetof(gkSimplex, vrtx_idx) == 104, "gkSimplex.vrtx_idx");
OGJK_STATIC_ASSERT(offsetof(gkSimplex, witnesses) == 136, "gkSimplex.witnesses");
#else
OGJK_STATIC_ASSERT(sizeof(gkPolytope) == 32, "gkPolytope FP32 layout");
OGJK_STATIC_ASSERT(offsetof(gkPolytope, s) == 4, "gkPolytope.s");
OGJK_STATIC_ASSERT(offsetof(gkPolytope, s_idx) == 16, "gkPolytope.s_idx");
OGJK_STATIC_ASSERT(offsetof(gkPolytope, coord) == 24, "gkPolytope.coord");
OGJK_STATIC_ASSERT(sizeof(gkSimplex) == 108, "gkSimplex FP32 layout");
OGJK_STATIC_ASSERT(offsetof(gkSimplex, vrtx) == 4, "gkSimplex.vrtx");
OGJK_STATIC_ASSERT(offsetof(gkSimplex, vrtx_idx) == 52, "gkSimplex.vrtx_idx");
OGJK_STATIC_ASSERT(offsetof(gkSimplex, witnesses) == 84, "gkSimplex.witnesses");
#endif
OGJK_STATIC_ASSERT(sizeof(gkCollisionPair) == 8, "gkCollisionPair layout");

#endif /* OPENGJK_AMD_COMMON_H */
