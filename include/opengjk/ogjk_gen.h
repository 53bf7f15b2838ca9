/*
 * opengjk-amd synthetic workload generator (host only, libogjk_gen.so).
 *
 * Counter-based: body b's vertex stream depends only on (seed, family, b), so
 * bodies are generated in parallel and the GPU path and the CPU checker see
 * bit-identical float32 inputs.  Vertices are computed in double and rounded
 * once to float.  Pair i of a batch is (body 2i, body 2i+1).
 *
 * Families (workload shapes of the benchmark configurations):
 *   OGJK_GEN_SPHERE      n = nverts points uniform on a sphere of radius
 *                        U[0.5,1.5], centre U[-box,box]^3
 *   OGJK_GEN_ELLIPSOID   n log-uniform in [nmin, nmax]; points on an
 *                        ellipsoid, axes U[0.3,2], rotated by a uniform random
 *                        quaternion, centre U[-box,box]^3
 *   OGJK_GEN_NOISY_SPHERE n = nverts points u + N(0, sigma)^3 with u uniform
 *                        on the unit sphere, centre U[-box,box]^3
 */
#ifndef OPENGJK_AMD_GEN_H
#define OPENGJK_AMD_GEN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  OGJK_GEN_SPHERE = 0,
  OGJK_GEN_ELLIPSOID = 2,
  OGJK_GEN_NOISY_SPHERE = 3,
};

typedef struct ogjk_gen_spec {
  int family;     /* OGJK_GEN_* */
  uint64_t seed;
  int n_bodies;   /* bodies 0..n_bodies-1 */
  int nverts;     /* fixed vertex count (SPHERE, NOISY_SPHERE) */
  int nmin, nmax; /* log-uniform vertex count range (ELLIPSOID) */
  double box;     /* centres drawn from U[-box, box]^3 */
  double sigma;   /* NOISY_SPHERE perturbation */
} ogjk_gen_spec;

/* Vertex count of every body; returns the total. */
int64_t ogjk_gen_counts(const ogjk_gen_spec* spec, int32_t* counts);

/* Fill AoS float32 coordinates; body b starts at vertex offsets[b]. */
void ogjk_gen_fill(const ogjk_gen_spec* spec, const int32_t* counts, const int64_t* offsets,
                   float* coords, int nthreads);

#ifdef __cplusplus
}
#endif

#endif /* OPENGJK_AMD_GEN_H */
