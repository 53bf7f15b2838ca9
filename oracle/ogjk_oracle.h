/*
 * Oracle for opengjk-amd: a CPU restatement of the reference GJK (openGJK) and
 * EPA, in plain C, used ONLY as a test checker and as the CPU baseline timed
 * by the benchmark.  Nothing in the shipped library links or calls it.
 *
 * It follows, operation for operation:
 *   GJK/cpu/openGJK.c  (GJK driver, support, S1D/S2D/S3D, W0D..W3D)
 *   GJK/cpu/EPA.c      (EPA, degenerate-simplex growth, barycentric witnesses)
 *   examples/cpu/example.cpp:13-40 (batch semantics: sequential per pair)
 *
 * Precision: compiled twice, with OGJK_REAL = double (the FP64 oracle the GPU
 * path is checked against) and OGJK_REAL = float (bit-compared with the real
 * reference sources compiled as shipped, whose default is FP32).
 *
 * The reference is GPL-3.0-only (GJK/common.h:14); this restatement is test
 * infrastructure kept apart from the product.
 */
#ifndef OGJK_ORACLE_H
#define OGJK_ORACLE_H

#include <stdint.h>

#ifndef OGJK_REAL
#define OGJK_REAL double
#endif

typedef OGJK_REAL oreal;

/* Same field order and meaning as gkPolytope / gkSimplex (GJK/common.h:68-89). */
typedef struct {
  int numpoints;
  oreal s[3];
  int s_idx;
  const oreal* coord;
} opoly;

typedef struct {
  int nvrtx;
  oreal vrtx[4][3];
  int vrtx_idx[4][2];
  oreal witnesses[2][3];
} osimplex;

/* Per-pair instrumentation (not part of the reference): number of GJK
 * iterations, EPA main-loop iterations, body sweeps done by the support
 * functions for body 1 and body 2, and EPA iterations that ran out of face
 * slots with horizon edges left (GJK/cpu/EPA.c:766-775). */
typedef struct {
  int gjk_iters;
  int epa_iters;
  int sweeps_a;
  int sweeps_b;
  int face_overflow;
} ostats;

#ifdef __cplusplus
extern "C" {
#endif

/* compute_minimum_distance (GJK/cpu/openGJK.c:952-1056). */
oreal oracle_gjk(opoly bd1, opoly bd2, osimplex* s, ostats* st);

/* computeCollisionInformation (GJK/cpu/EPA.c:362-864). */
void oracle_epa(const opoly* bd1, const opoly* bd2, osimplex* s, oreal* distance,
                oreal contact_normal[3], ostats* st);

/* Batch driver over a pooled AoS coordinate array.
 *   coords     : AoS pool, body b occupies coords[3*off[b] .. 3*(off[b]+cnt[b]))
 *   pair_a/b   : body index of side 1 / side 2 for each pair
 *   mode       : 1 = GJK only, 2 = EPA only (reads simplices/distances),
 *                3 = GJK then EPA (GJK::GPU::computeGJKAndEPA semantics)
 *   normals    : n*3, written for modes 2 and 3 (may be NULL for mode 1)
 *   stats      : n records or NULL
 *   nthreads   : worker threads (dynamic chunking); <=1 runs inline
 */
void oracle_batch(int mode, int n, const oreal* coords, const int64_t* off,
                  const int32_t* cnt, const int32_t* pair_a, const int32_t* pair_b,
                  osimplex* simplices, oreal* distances, oreal* normals,
                  ostats* stats, int nthreads);

/* As oracle_batch over a float32 AoS pool: each worker promotes the two
 * bodies of a pair exactly to oreal before running it, so a large pool is
 * never converted as a whole. */
void oracle_batch_f32(int mode, int n, const float* coords, const int64_t* off,
                      const int32_t* cnt, const int32_t* pair_a, const int32_t* pair_b,
                      osimplex* simplices, oreal* distances, oreal* normals,
                      ostats* stats, int nthreads);

int oracle_real_size(void);

#ifdef __cplusplus
}
#endif

#endif /* OGJK_ORACLE_H */
