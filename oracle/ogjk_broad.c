/* Broad-phase oracle: uniform-grid sphere overlap -> pair lists, and the
 * body-pair -> sub-mesh-pair expansion.  TEST INFRASTRUCTURE ONLY: the
 * checker for the GPU broad phase; nothing in the product links it.
 *
 * Restates, single-threaded and in plain C:
 *   insert_objects_kernel   visualization/integrate_final_gjk.cu:467-491
 *   count_pairs_kernel      visualization/integrate_final_gjk.cu:493-528
 *   generate_pairs_kernel   visualization/integrate_final_gjk.cu:530-570
 *   sim_broad_phase         visualization/integrate_final_gjk.cu:916-1002
 *                           (prefix scan of counts, max_pairs clamp)
 *   count_submesh_pairs_kernel / expand_submesh_pairs_kernel
 *                           visualization/integrate_final_gjk.cu:265-302
 *
 * Order of the pair list: by first object; within one object by neighbour
 * cell, z outermost, then y, then x, each from -1 to +1; within a cell by the
 * cell list's order.  The reference fills cell lists with atomicAdd, so its
 * list order (and, when a cell holds more than max_per_cell objects, which
 * ones it keeps) depends on thread timing; this restatement fixes it to
 * ascending object index, one of the orders the reference can produce.
 *
 * Arithmetic is FP32 with every operation rounded (no contraction):
 *   cell   = clamp(floor((p + boundary) / cell_size), 0, grid_size - 1)
 *   overlap: ((dx*dx + dy*dy) + dz*dz) < (r_i + r_j) * (r_i + r_j)
 * A NaN coordinate maps to cell 0 (the reference's float->int conversion of
 * NaN gives 0) and never overlaps anything.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

static int cell_axis(float p, float boundary, float cell_size, int grid_size) {
  const float f = floorf((p + boundary) / cell_size);
  if (!(f >= 0.0f)) return 0; /* negative or NaN */
  if (f > (float)(grid_size - 1)) return grid_size - 1;
  return (int)f;
}

static int overlap(const float* a, const float* b) {
  const float dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
  const float d2 = (dx * dx + dy * dy) + dz * dz;
  const float rs = a[3] + b[3];
  return d2 < rs * rs;
}

/* spheres: n x (x, y, z, radius).  pairs: up to max_pairs x (i, j).
 * Returns the number of overlapping pairs found (before the clamp), or -1 if
 * scratch memory could not be allocated. */
int64_t oracle_broad_phase(const float* spheres, int32_t n, float cell_size, float boundary,
                           int32_t grid_size, int32_t max_per_cell, int32_t* pairs,
                           int64_t max_pairs) {
  if (n <= 0) return 0;
  const int64_t g = grid_size, cells = g * g * g;
  int32_t* cell = malloc(sizeof(int32_t) * (size_t)n);
  int64_t* start = calloc((size_t)cells + 1, sizeof(int64_t));
  int32_t* list = malloc(sizeof(int32_t) * (size_t)n);
  int64_t* fill = malloc(sizeof(int64_t) * (size_t)cells);
  if (!cell || !start || !list || !fill) {
    free(cell); free(start); free(list); free(fill);
    return -1;
  }
  /* insert: cell lists in ascending object order (counting sort) */
  for (int32_t i = 0; i < n; i++) {
    const float* s = spheres + 4 * (int64_t)i;
    const int cx = cell_axis(s[0], boundary, cell_size, grid_size);
    const int cy = cell_axis(s[1], boundary, cell_size, grid_size);
    const int cz = cell_axis(s[2], boundary, cell_size, grid_size);
    cell[i] = (int32_t)(cx + cy * g + cz * g * g);
    start[cell[i] + 1]++;
  }
  for (int64_t c = 0; c < cells; c++) start[c + 1] += start[c];
  memcpy(fill, start, sizeof(int64_t) * (size_t)cells);
  for (int32_t i = 0; i < n; i++) list[fill[cell[i]]++] = i;

  /* count and generate in one pass: the running total is the exclusive scan */
  int64_t total = 0;
  for (int32_t i = 0; i < n; i++) {
    const float* s = spheres + 4 * (int64_t)i;
    const int64_t c = cell[i];
    const int cx = (int)(c % g), cy = (int)((c / g) % g), cz = (int)(c / (g * g));
    for (int dz = -1; dz <= 1; dz++)
      for (int dy = -1; dy <= 1; dy++)
        for (int dx = -1; dx <= 1; dx++) {
          const int nx = cx + dx, ny = cy + dy, nz = cz + dz;
          if (nx < 0 || nx >= grid_size || ny < 0 || ny >= grid_size || nz < 0 ||
              nz >= grid_size)
            continue;
          const int64_t nc = nx + ny * g + nz * g * g;
          int64_t lo = start[nc], hi = start[nc + 1];
          if (max_per_cell > 0 && hi - lo > max_per_cell) hi = lo + max_per_cell;
          for (int64_t k = lo; k < hi; k++) {
            const int32_t j = list[k];
            if (j <= i) continue;
            if (!overlap(s, spheres + 4 * (int64_t)j)) continue;
            if (total < max_pairs) {
              pairs[2 * total] = i;
              pairs[2 * total + 1] = j;
            }
            total++;
          }
        }
  }
  free(cell); free(start); free(list); free(fill);
  return total;
}

/* Body pairs -> sub-mesh pairs: pair (a, b) expands to
 * (sub_off[a] + u, sub_off[b] + v), u-major, u < sub_cnt[a], v < sub_cnt[b].
 * A pair naming a body outside [0, n_bodies) expands to nothing (the
 * reference reads out of bounds there).  Returns the number of sub-mesh pairs
 * (before the clamp to max_sub). */
int64_t oracle_expand_pairs(const int32_t* body_pairs, int32_t n_body_pairs,
                            const int32_t* sub_off, const int32_t* sub_cnt, int32_t n_bodies,
                            int32_t* sub_pairs, int64_t max_sub) {
  int64_t w = 0;
  for (int32_t p = 0; p < n_body_pairs; p++) {
    const int32_t a = body_pairs[2 * p], b = body_pairs[2 * p + 1];
    if (a < 0 || a >= n_bodies || b < 0 || b >= n_bodies) continue;
    for (int32_t u = 0; u < sub_cnt[a]; u++)
      for (int32_t v = 0; v < sub_cnt[b]; v++) {
        if (w < max_sub) {
          sub_pairs[2 * w] = sub_off[a] + u;
          sub_pairs[2 * w + 1] = sub_off[b] + v;
        }
        w++;
      }
  }
  return w;
}
