/* Convex-hull oracle: the incremental hull and its volume moments.  TEST
 * INFRASTRUCTURE ONLY: the checker for the GPU hull builder; nothing in the
 * product links it.
 *
 * Restates, in plain C:
 *   hull_signed_vol6, build_hull  visualization/rendering/mesh_builder.cpp:268-360
 *   hull_volume_moments           visualization/rendering/mesh_builder.cpp:433-462
 *   center_hull_verts             visualization/rendering/mesh_builder.cpp:464-469
 *   compute_hull_inertia_k        visualization/rendering/mesh_builder.cpp:471-481
 *
 * Arithmetic as the reference's host build has it (g++ -O3, no FMA): FP32
 * signed volumes with every operation rounded, FP64 moments summed in face
 * order.  Faces are vertex-index triples; the list order is the reference's
 * (visible faces removed in place, new faces appended in horizon order).
 * Returns the number of faces (0 when n < 4 or the points are collinear or
 * coplanar), or -1 when `cap` faces or horizon edges are exceeded.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

static float vol6(const float* p, const int32_t* f, int32_t q) {
  const float *a = p + 3 * f[0], *b = p + 3 * f[1], *c = p + 3 * f[2], *d = p + 3 * q;
  const float bx = b[0] - a[0], by = b[1] - a[1], bz = b[2] - a[2];
  const float cx = c[0] - a[0], cy = c[1] - a[1], cz = c[2] - a[2];
  const float dx = d[0] - a[0], dy = d[1] - a[1], dz = d[2] - a[2];
  const float nx = by * cz - bz * cy, ny = bz * cx - bx * cz, nz = bx * cy - by * cx;
  return nx * dx + ny * dy + nz * dz;
}

int32_t oracle_build_hull(const float* p, int32_t n, int32_t* faces, int32_t cap) {
  if (n < 4) return 0;
  int32_t i0 = 0, i1 = 1, i2 = -1, i3 = -1;
  {
    int32_t mn[3] = {0, 0, 0}, mx[3] = {0, 0, 0};
    for (int32_t i = 1; i < n; i++)
      for (int k = 0; k < 3; k++) {
        if (p[3 * i + k] < p[3 * mn[k] + k]) mn[k] = i;
        if (p[3 * i + k] > p[3 * mx[k] + k]) mx[k] = i;
      }
    const int32_t c[6] = {mn[0], mx[0], mn[1], mx[1], mn[2], mx[2]};
    float bd = 0;
    for (int a = 0; a < 6; a++)
      for (int b = a + 1; b < 6; b++) {
        const float dx = p[3 * c[b]] - p[3 * c[a]], dy = p[3 * c[b] + 1] - p[3 * c[a] + 1],
                    dz = p[3 * c[b] + 2] - p[3 * c[a] + 2];
        const float d = dx * dx + dy * dy + dz * dz;
        if (d > bd) {
          bd = d;
          i0 = c[a];
          i1 = c[b];
        }
      }
  }
  {
    const float e0 = p[3 * i1] - p[3 * i0], e1 = p[3 * i1 + 1] - p[3 * i0 + 1],
                e2 = p[3 * i1 + 2] - p[3 * i0 + 2];
    float bd = 0;
    for (int32_t i = 0; i < n; i++) {
      if (i == i0 || i == i1) continue;
      const float vx = p[3 * i] - p[3 * i0], vy = p[3 * i + 1] - p[3 * i0 + 1],
                  vz = p[3 * i + 2] - p[3 * i0 + 2];
      const float cx = vy * e2 - vz * e1, cy = vz * e0 - vx * e2, cz = vx * e1 - vy * e0;
      const float d = cx * cx + cy * cy + cz * cz;
      if (d > bd) {
        bd = d;
        i2 = i;
      }
    }
  }
  if (i2 < 0) return 0;
  {
    float bd = 0;
    const int32_t f[3] = {i0, i1, i2};
    for (int32_t i = 0; i < n; i++) {
      if (i == i0 || i == i1 || i == i2) continue;
      float d = vol6(p, f, i);
      if (d < 0) d = -d;
      if (d > bd) {
        bd = d;
        i3 = i;
      }
    }
  }
  if (i3 < 0) return 0;
  {
    const int32_t f[3] = {i0, i1, i2};
    if (vol6(p, f, i3) > 0) {
      const int32_t t = i1;
      i1 = i2;
      i2 = t;
    }
  }
  if (cap < 4) return -1;
  int32_t nf = 4;
  const int32_t init[12] = {i0, i1, i2, i0, i2, i3, i0, i3, i1, i1, i3, i2};
  memcpy(faces, init, sizeof(init));

  int32_t* vis = malloc(sizeof(int32_t) * (size_t)cap);
  int32_t* hor = malloc(sizeof(int32_t) * 2 * (size_t)cap);
  char* alive = malloc((size_t)cap);
  if (!vis || !hor || !alive) {
    free(vis); free(hor); free(alive);
    return -1;
  }
  int32_t result = 0;
  for (int32_t q = 0; q < n; q++) {
    if (q == i0 || q == i1 || q == i2 || q == i3) continue;
    int32_t nv = 0;
    for (int32_t f = 0; f < nf; f++)
      if (vol6(p, faces + 3 * f, q) > 1e-7f) vis[nv++] = f;
    if (nv == 0) continue;
    /* horizon: directed edges of visible faces whose reverse is on no visible face */
    int32_t nh = 0;
    for (int32_t k = 0; k < nv; k++)
      for (int e = 0; e < 3; e++) {
        const int32_t a = faces[3 * vis[k] + e], b = faces[3 * vis[k] + (e + 1) % 3];
        int found = 0;
        for (int32_t k2 = 0; k2 < nv && !found; k2++)
          for (int e2 = 0; e2 < 3; e2++)
            if (faces[3 * vis[k2] + e2] == b && faces[3 * vis[k2] + (e2 + 1) % 3] == a) found = 1;
        if (!found) {
          if (nh >= cap) { result = -1; goto done; }
          hor[2 * nh] = a;
          hor[2 * nh + 1] = b;
          nh++;
        }
      }
    memset(alive, 1, (size_t)nf);
    for (int32_t k = 0; k < nv; k++) alive[vis[k]] = 0;
    int32_t w = 0;
    for (int32_t f = 0; f < nf; f++)
      if (alive[f]) {
        memmove(faces + 3 * w, faces + 3 * f, 3 * sizeof(int32_t));
        w++;
      }
    if (w + nh > cap) { result = -1; goto done; }
    for (int32_t h = 0; h < nh; h++) {
      faces[3 * (w + h)] = hor[2 * h];
      faces[3 * (w + h) + 1] = hor[2 * h + 1];
      faces[3 * (w + h) + 2] = q;
    }
    nf = w + nh;
  }
  result = nf;
done:
  free(vis); free(hor); free(alive);
  return result;
}

/* m[0..6] = volume, first moments x, y, z (times 4 volume), second moments
 * x^2, y^2, z^2, summed over the faces in order. */
void oracle_hull_moments(const float* p, const int32_t* faces, int32_t nf, double* m) {
  double vol = 0, cx = 0, cy = 0, cz = 0, ix2 = 0, iy2 = 0, iz2 = 0;
  for (int32_t f = 0; f < nf; f++) {
    const float *a = p + 3 * faces[3 * f], *b = p + 3 * faces[3 * f + 1],
                *c = p + 3 * faces[3 * f + 2];
    const double dv = ((double)a[0] * ((double)b[1] * (double)c[2] - (double)b[2] * (double)c[1]) -
                       (double)a[1] * ((double)b[0] * (double)c[2] - (double)b[2] * (double)c[0]) +
                       (double)a[2] * ((double)b[0] * (double)c[1] - (double)b[1] * (double)c[0])) /
                      6.0;
    vol += dv;
    cx += dv * (a[0] + b[0] + c[0]);
    cy += dv * (a[1] + b[1] + c[1]);
    cz += dv * (a[2] + b[2] + c[2]);
#define OGJK_I2(ak, bk, ck)                                                                  \
  (dv / 10.0 * ((double)(ak) * (ak) + (double)(ak) * (bk) + (double)(bk) * (bk) +           \
                (double)(ak) * (ck) + (double)(bk) * (ck) + (double)(ck) * (ck)))
    ix2 += OGJK_I2(a[0], b[0], c[0]);
    iy2 += OGJK_I2(a[1], b[1], c[1]);
    iz2 += OGJK_I2(a[2], b[2], c[2]);
#undef OGJK_I2
  }
  m[0] = vol; m[1] = cx; m[2] = cy; m[3] = cz; m[4] = ix2; m[5] = iy2; m[6] = iz2;
}

/* Centre of mass as the reference shifts by it (0 when volume < 1e-10), and
 * the unit inertia factors kx, ky, kz of the moments m. */
void oracle_hull_com(const double* m, float* com) {
  if (m[0] < 1e-10) {
    com[0] = com[1] = com[2] = 0.0f;
    return;
  }
  com[0] = (float)(m[1] / (4.0 * m[0]));
  com[1] = (float)(m[2] / (4.0 * m[0]));
  com[2] = (float)(m[3] / (4.0 * m[0]));
}

void oracle_hull_inertia_k(const double* m, float* k) {
  if (m[0] < 1e-10) {
    k[0] = k[1] = k[2] = 1.0f / 10.0f;
    return;
  }
  const double c0 = m[1] / (4.0 * m[0]), c1 = m[2] / (4.0 * m[0]), c2 = m[3] / (4.0 * m[0]);
  k[0] = (float)((m[4] - m[0] * c0 * c0) / m[0]);
  k[1] = (float)((m[5] - m[0] * c1 * c1) / m[0]);
  k[2] = (float)((m[6] - m[0] * c2 * c2) / m[0]);
}

/* ---------------------------------------------------------------- batches */
#include <pthread.h>

typedef struct {
  const float* pts;
  const int64_t* off;
  const int32_t* cnt;
  int32_t n;
  int32_t* faces;
  const int64_t* face_off;
  int32_t* nfaces;
  float* centred;
  float* k;
  int32_t next;
  pthread_mutex_t mu;
} IngestJob;

static void ingest_one(IngestJob* j, int32_t m) {
  const float* p = j->pts + 3 * j->off[m];
  float* pc = j->centred + 3 * j->off[m];
  int32_t* f = j->faces + 3 * j->face_off[m];
  const int32_t cap = (int32_t)(j->face_off[m + 1] - j->face_off[m]);
  const int32_t n = j->cnt[m];
  double mo[7];
  float com[3];
  int32_t nf = oracle_build_hull(p, n, f, cap);
  oracle_hull_moments(p, f, nf > 0 ? nf : 0, mo);
  oracle_hull_com(mo, com);
  for (int32_t i = 0; i < n; i++)
    for (int c = 0; c < 3; c++) pc[3 * i + c] = p[3 * i + c] - com[c];
  nf = oracle_build_hull(pc, n, f, cap);
  j->nfaces[m] = nf;
  oracle_hull_moments(pc, f, nf > 0 ? nf : 0, mo);
  oracle_hull_inertia_k(mo, j->k + 3 * m);
}

static void* ingest_worker(void* arg) {
  IngestJob* j = (IngestJob*)arg;
  for (;;) {
    pthread_mutex_lock(&j->mu);
    const int32_t m0 = j->next;
    j->next += 64;
    pthread_mutex_unlock(&j->mu);
    if (m0 >= j->n) return NULL;
    const int32_t m1 = m0 + 64 < j->n ? m0 + 64 : j->n;
    for (int32_t m = m0; m < m1; m++) ingest_one(j, m);
  }
}

/* The reference's ingestion of each cloud (hull, centre of mass, centring,
 * hull of the centred points, inertia factors; main_opengl.cpp:114-121, 313)
 * over n clouds on nthreads threads.  faces/nfaces hold the second hull. */
void oracle_ingest_batch(const float* pts, const int64_t* off, const int32_t* cnt, int32_t n,
                         int32_t* faces, const int64_t* face_off, int32_t* nfaces,
                         float* centred, float* k, int32_t nthreads) {
  IngestJob j = {pts, off, cnt, n, faces, face_off, nfaces, centred, k, 0,
                 PTHREAD_MUTEX_INITIALIZER};
  if (nthreads < 1) nthreads = 1;
  if (nthreads > 256) nthreads = 256;
  pthread_t th[256];
  for (int t = 0; t < nthreads; t++) pthread_create(&th[t], NULL, ingest_worker, &j);
  for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
}
