// Mesh ingestion on the device: the incremental convex hull of many point
// clouds at once, and the FP64 volume moments, centre of mass and inertia
// factors of each hull.
//
// The face list is the reference's (visualization/rendering/mesh_builder.cpp:
// 268-360): the same initial tetrahedron, points inserted in index order,
// faces with a 6x signed volume above 1e-7 (FP32) visible, visible faces
// removed in place, one new face per horizon edge appended in the order the
// visible faces and their edges are listed.  The reference runs that loop on
// one CPU thread per cloud; here one wave owns a cloud and each step is
// data-parallel over the wave: visibility over faces, the horizon test over
// the visible faces' edges, the order-preserving compaction by ballots and
// prefix counts.  Moments are summed in face order by one lane, as the
// reference sums them (:433-481), so they are bit-identical too.
#include <hip/hip_runtime.h>

#include <stdint.h>

#include "ogjk_internal.hpp"
#include "opengjk/ogjk.h"

ogjk_status ogjk_fail(ogjk_status code, const char* fmt, ...);

namespace {

constexpr int kWave = 64;
// A cloud whose face capacity is at most this works in LDS (faces, visible
// list and horizon: 24 B per face, 6 KiB per wave) and copies its faces out
// once; larger clouds work in the output buffer and the workspace.
constexpr int kLdsFaces = 256;

__device__ __forceinline__ float vol6(const float* __restrict__ p, int a, int b, int c,
                                      float qx, float qy, float qz) {
  const float ax = p[3 * a], ay = p[3 * a + 1], az = p[3 * a + 2];
  const float bx = p[3 * b] - ax, by = p[3 * b + 1] - ay, bz = p[3 * b + 2] - az;
  const float cx = p[3 * c] - ax, cy = p[3 * c + 1] - ay, cz = p[3 * c + 2] - az;
  const float dx = qx - ax, dy = qy - ay, dz = qz - az;
  const float nx = by * cz - bz * cy, ny = bz * cx - bx * cz, nz = bx * cy - by * cx;
  return nx * dx + ny * dy + nz * dz;
}

__device__ __forceinline__ uint64_t ballot(bool v) { return __ballot(v); }

__device__ __forceinline__ int rank_below(uint64_t mask, int lane) {
  return __popcll(mask & ((uint64_t(1) << lane) - 1));
}

// Wave argmax over (value, index): larger value, then lower index.  Lanes
// with index < 0 hold nothing.
__device__ __forceinline__ void wave_argmax(float& v, int& i) {
#pragma unroll
  for (int m = 1; m < kWave; m <<= 1) {
    const float ov = __shfl_xor(v, m);
    const int oi = __shfl_xor(i, m);
    const bool take = oi >= 0 && (i < 0 || ov > v || (ov == v && oi < i));
    v = take ? ov : v;
    i = take ? oi : i;
  }
}

__device__ __forceinline__ void wave_argmin(float& v, int& i) {
#pragma unroll
  for (int m = 1; m < kWave; m <<= 1) {
    const float ov = __shfl_xor(v, m);
    const int oi = __shfl_xor(i, m);
    const bool take = oi >= 0 && (i < 0 || ov < v || (ov == v && oi < i));
    v = take ? ov : v;
    i = take ? oi : i;
  }
}

// One wave per cloud.  faces/scratch of cloud m start at 3 * face_off[m] and
// hold face_off[m + 1] - face_off[m] faces; nfaces[m] = face count, 0 for no
// hull (fewer than 4 points, collinear or coplanar), -1 when the capacity is
// exceeded, -2 for a non-finite coordinate, -3 when the cloud's face range is
// not inside [0, total_faces] (nothing is written for it).
__global__ __launch_bounds__(kWave) void k_hull(const float* __restrict__ pts,
                                                const int64_t* __restrict__ off,
                                                const int32_t* __restrict__ cnt,
                                                const int64_t* __restrict__ face_off,
                                                int64_t total_faces,
                                                int32_t* __restrict__ faces,
                                                int32_t* __restrict__ nfaces,
                                                int32_t* __restrict__ scratch) {
  const int m = blockIdx.x;
  const int lane = threadIdx.x;
  const int n = cnt[m];
  const float* P = pts + 3 * off[m];
  const int64_t f0 = face_off[m];
  const int64_t cap64 = face_off[m + 1] - f0;
  if (f0 < 0 || cap64 < 0 || face_off[m + 1] > total_faces) {
    if (lane == 0) nfaces[m] = -3;
    return;
  }
  const int cap = cap64 < 0x7fffffff ? (int)cap64 : 0x7fffffff;
  __shared__ int32_t lds[6 * kLdsFaces];
  const bool in_lds = cap64 <= kLdsFaces;
  int32_t* F = in_lds ? lds : faces + 3 * f0;                          // 3 * cap entries
  int32_t* vis = in_lds ? lds + 3 * kLdsFaces : scratch + 3 * f0;     // cap entries
  int32_t* hor = in_lds ? lds + 4 * kLdsFaces : scratch + 3 * f0 + cap64;  // 2 * cap
  if (n < 4) {
    if (lane == 0) nfaces[m] = 0;
    return;
  }
  // axis extremes (first index on ties), rejecting non-finite input
  float mnv[3], mxv[3];
  int mni[3], mxi[3];
  bool bad = false;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    mnv[k] = INFINITY;
    mxv[k] = -INFINITY;
    mni[k] = mxi[k] = -1;
  }
  for (int i = lane; i < n; i += kWave) {
#pragma unroll
    for (int k = 0; k < 3; k++) {
      const float v = P[3 * i + k];
      bad |= !isfinite(v);
      if (mni[k] < 0 || v < mnv[k]) { mnv[k] = v; mni[k] = i; }
      if (mxi[k] < 0 || v > mxv[k]) { mxv[k] = v; mxi[k] = i; }
    }
  }
  if (ballot(bad)) {
    if (lane == 0) nfaces[m] = -2;
    return;
  }
#pragma unroll
  for (int k = 0; k < 3; k++) {
    wave_argmin(mnv[k], mni[k]);
    wave_argmax(mxv[k], mxi[k]);
  }
  // most separated pair among the six
  int i0 = 0, i1 = 1;
  {
    const int c[6] = {mni[0], mxi[0], mni[1], mxi[1], mni[2], mxi[2]};
    float bd = 0;
    for (int a = 0; a < 6; a++)
      for (int b = a + 1; b < 6; b++) {
        const float dx = P[3 * c[b]] - P[3 * c[a]], dy = P[3 * c[b] + 1] - P[3 * c[a] + 1],
                    dz = P[3 * c[b] + 2] - P[3 * c[a] + 2];
        const float d = dx * dx + dy * dy + dz * dz;
        if (d > bd) {
          bd = d;
          i0 = c[a];
          i1 = c[b];
        }
      }
  }
  // farthest from the line i0-i1
  int i2;
  {
    const float e0 = P[3 * i1] - P[3 * i0], e1 = P[3 * i1 + 1] - P[3 * i0 + 1],
                e2 = P[3 * i1 + 2] - P[3 * i0 + 2];
    float bv = 0;
    int bi = -1;
    for (int i = lane; i < n; i += kWave) {
      if (i == i0 || i == i1) continue;
      const float vx = P[3 * i] - P[3 * i0], vy = P[3 * i + 1] - P[3 * i0 + 1],
                  vz = P[3 * i + 2] - P[3 * i0 + 2];
      const float cx = vy * e2 - vz * e1, cy = vz * e0 - vx * e2, cz = vx * e1 - vy * e0;
      const float d = cx * cx + cy * cy + cz * cz;
      if (d > bv) { bv = d; bi = i; }
    }
    wave_argmax(bv, bi);
    i2 = bi;
  }
  if (i2 < 0) {
    if (lane == 0) nfaces[m] = 0;
    return;
  }
  // farthest from the plane i0-i1-i2
  int i3;
  {
    float bv = 0;
    int bi = -1;
    for (int i = lane; i < n; i += kWave) {
      if (i == i0 || i == i1 || i == i2) continue;
      float d = vol6(P, i0, i1, i2, P[3 * i], P[3 * i + 1], P[3 * i + 2]);
      if (d < 0) d = -d;
      if (d > bv) { bv = d; bi = i; }
    }
    wave_argmax(bv, bi);
    i3 = bi;
  }
  if (i3 < 0) {
    if (lane == 0) nfaces[m] = 0;
    return;
  }
  if (vol6(P, i0, i1, i2, P[3 * i3], P[3 * i3 + 1], P[3 * i3 + 2]) > 0) {
    const int t = i1;
    i1 = i2;
    i2 = t;
  }
  if (cap < 4) {
    if (lane == 0) nfaces[m] = -1;
    return;
  }
  if (lane < 12) {
    const int init[12] = {i0, i1, i2, i0, i2, i3, i0, i3, i1, i1, i3, i2};
    F[lane] = init[lane];
  }
  __syncthreads();
  int nf = 4;
  for (int q = 0; q < n; q++) {
    if (q == i0 || q == i1 || q == i2 || q == i3) continue;
    const float qx = P[3 * q], qy = P[3 * q + 1], qz = P[3 * q + 2];
    // visible faces, listed in ascending order
    int nv = 0;
    for (int base = 0; base < nf; base += kWave) {
      const int f = base + lane;
      bool v = false;
      if (f < nf) v = vol6(P, F[3 * f], F[3 * f + 1], F[3 * f + 2], qx, qy, qz) > 1e-7f;
      const uint64_t b = ballot(v);
      if (v) vis[nv + rank_below(b, lane)] = f;
      nv += __popcll(b);
    }
    if (nv == 0) continue;
    __syncthreads();
    // horizon: directed edges of visible faces whose reverse no visible face
    // has, in (visible face, edge) order
    int nh = 0;
    const int ne = 3 * nv;
    for (int base = 0; base < ne; base += kWave) {
      const int h = base + lane;
      bool hz = false;
      int a = 0, b = 0;
      if (h < ne) {
        const int k = h / 3, e = h - 3 * k;
        const int g = vis[k];
        a = F[3 * g + e];
        b = F[3 * g + (e == 2 ? 0 : e + 1)];
        bool found = false;
        for (int k2 = 0; k2 < nv && !found; k2++) {
          const int g2 = vis[k2];
          const int x0 = F[3 * g2], x1 = F[3 * g2 + 1], x2 = F[3 * g2 + 2];
          found = (x0 == b && x1 == a) || (x1 == b && x2 == a) || (x2 == b && x0 == a);
        }
        hz = !found;
      }
      const uint64_t bl = ballot(hz);
      const int r = nh + rank_below(bl, lane);
      if (hz && r < cap) {
        hor[2 * r] = a;
        hor[2 * r + 1] = b;
      }
      nh += __popcll(bl);
    }
    __syncthreads();
    // drop the visible faces, keeping the order of the others
    int w = 0;
    for (int base = 0; base < nf; base += kWave) {
      const int f = base + lane;
      bool alive = false;
      int x0 = 0, x1 = 0, x2 = 0;
      if (f < nf) {
        x0 = F[3 * f];
        x1 = F[3 * f + 1];
        x2 = F[3 * f + 2];
        alive = !(vol6(P, x0, x1, x2, qx, qy, qz) > 1e-7f);
      }
      const uint64_t b = ballot(alive);
      if (alive) {
        const int d = w + rank_below(b, lane);
        F[3 * d] = x0;
        F[3 * d + 1] = x1;
        F[3 * d + 2] = x2;
      }
      w += __popcll(b);
      __syncthreads();
    }
    if (nh > cap || w + nh > cap) {
      if (lane == 0) nfaces[m] = -1;
      return;
    }
    for (int r = lane; r < nh; r += kWave) {
      F[3 * (w + r)] = hor[2 * r];
      F[3 * (w + r) + 1] = hor[2 * r + 1];
      F[3 * (w + r) + 2] = q;
    }
    nf = w + nh;
    __syncthreads();
  }
  if (in_lds) {
    int32_t* out = faces + 3 * f0;
    for (int i = lane; i < 3 * nf; i += kWave) out[i] = F[i];
  }
  if (lane == 0) nfaces[m] = nf;
}

// One thread per cloud: moments summed in face order (mesh_builder.cpp:433-462),
// centre of mass as center_hull_verts shifts by it (0 below volume 1e-10) and
// the inertia factors of compute_hull_inertia_k (:464-481).
__global__ __launch_bounds__(64) void k_hull_moments(const float* __restrict__ pts,
                                                     const int64_t* __restrict__ off,
                                                     const int32_t* __restrict__ faces,
                                                     const int64_t* __restrict__ face_off,
                                                     const int32_t* __restrict__ nfaces, int nm,
                                                     double* __restrict__ moments,
                                                     float* __restrict__ com,
                                                     float* __restrict__ inertia_k) {
  const int m = blockIdx.x * 64 + threadIdx.x;
  if (m >= nm) return;
  const float* P = pts + 3 * off[m];
  const int32_t* F = faces + 3 * face_off[m];
  const int nf = nfaces[m] > 0 ? nfaces[m] : 0;
  double vol = 0, cx = 0, cy = 0, cz = 0, ix2 = 0, iy2 = 0, iz2 = 0;
  for (int f = 0; f < nf; f++) {
    const float* a = P + 3 * F[3 * f];
    const float* b = P + 3 * F[3 * f + 1];
    const float* c = P + 3 * F[3 * f + 2];
    const double dv =
        ((double)a[0] * ((double)b[1] * (double)c[2] - (double)b[2] * (double)c[1]) -
         (double)a[1] * ((double)b[0] * (double)c[2] - (double)b[2] * (double)c[0]) +
         (double)a[2] * ((double)b[0] * (double)c[1] - (double)b[1] * (double)c[0])) /
        6.0;
    vol += dv;
    cx += dv * (a[0] + b[0] + c[0]);
    cy += dv * (a[1] + b[1] + c[1]);
    cz += dv * (a[2] + b[2] + c[2]);
    ix2 += dv / 10.0 *
           ((double)a[0] * a[0] + (double)a[0] * b[0] + (double)b[0] * b[0] +
            (double)a[0] * c[0] + (double)b[0] * c[0] + (double)c[0] * c[0]);
    iy2 += dv / 10.0 *
           ((double)a[1] * a[1] + (double)a[1] * b[1] + (double)b[1] * b[1] +
            (double)a[1] * c[1] + (double)b[1] * c[1] + (double)c[1] * c[1]);
    iz2 += dv / 10.0 *
           ((double)a[2] * a[2] + (double)a[2] * b[2] + (double)b[2] * b[2] +
            (double)a[2] * c[2] + (double)b[2] * c[2] + (double)c[2] * c[2]);
  }
  double* mo = moments + 7 * (int64_t)m;
  mo[0] = vol; mo[1] = cx; mo[2] = cy; mo[3] = cz; mo[4] = ix2; mo[5] = iy2; mo[6] = iz2;
  float* co = com + 3 * (int64_t)m;
  float* ko = inertia_k + 3 * (int64_t)m;
  if (vol < 1e-10) {
    co[0] = co[1] = co[2] = 0.0f;
    ko[0] = ko[1] = ko[2] = 1.0f / 10.0f;
    return;
  }
  const double c0 = cx / (4.0 * vol), c1 = cy / (4.0 * vol), c2 = cz / (4.0 * vol);
  co[0] = (float)c0;
  co[1] = (float)c1;
  co[2] = (float)c2;
  ko[0] = (float)((ix2 - vol * c0 * c0) / vol);
  ko[1] = (float)((iy2 - vol * c1 * c1) / vol);
  ko[2] = (float)((iz2 - vol * c2 * c2) / vol);
}

// points of cloud m minus com[m], one thread per point
__global__ __launch_bounds__(256) void k_center(float* __restrict__ pts,
                                                const int32_t* __restrict__ cloud,
                                                int64_t n_points,
                                                const float* __restrict__ com) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n_points;
       i += (int64_t)gridDim.x * 256) {
    const float* c = com + 3 * (int64_t)cloud[i];
    pts[3 * i] -= c[0];
    pts[3 * i + 1] -= c[1];
    pts[3 * i + 2] -= c[2];
  }
}

ogjk_status hip_err(hipError_t e, const char* what) {
  return ogjk_fail(e == hipErrorOutOfMemory ? OGJK_ERR_NOMEM : OGJK_ERR_HIP, "%s: %s", what,
                   hipGetErrorString(e));
}

}  // namespace

extern "C" {

ogjk_status ogjk_build_hulls(const float* d_points, const int64_t* d_point_off,
                             const int32_t* d_point_cnt, int32_t n_clouds,
                             const int64_t* d_face_off, int64_t total_faces, int32_t* d_faces,
                             int32_t* d_nfaces, void* d_workspace, size_t workspace_bytes,
                             void* stream) {
  if (n_clouds < 0 || total_faces < 0)
    return ogjk_fail(OGJK_ERR_INVALID, "n_clouds and total_faces must be >= 0");
  if (n_clouds == 0) return OGJK_OK;
  if (!d_points || !d_point_off || !d_point_cnt || !d_face_off || !d_faces || !d_nfaces)
    return ogjk_fail(OGJK_ERR_INVALID, "NULL argument");
  const size_t need = (size_t)total_faces * 3 * sizeof(int32_t);
  if (!d_workspace || workspace_bytes < need)
    return ogjk_fail(OGJK_ERR_INVALID, "workspace of %zu bytes needed", need);
  hipLaunchKernelGGL(k_hull, dim3((unsigned)n_clouds), dim3(kWave), 0, (hipStream_t)stream,
                     d_points, d_point_off, d_point_cnt, d_face_off, total_faces, d_faces,
                     d_nfaces, (int32_t*)d_workspace);
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? OGJK_OK : hip_err(e, "hull launch");
}

ogjk_status ogjk_hull_moments(const float* d_points, const int64_t* d_point_off,
                              const int32_t* d_faces, const int64_t* d_face_off,
                              const int32_t* d_nfaces, int32_t n_clouds, double* d_moments,
                              float* d_com, float* d_inertia_k, void* stream) {
  if (n_clouds < 0) return ogjk_fail(OGJK_ERR_INVALID, "n_clouds < 0");
  if (n_clouds == 0) return OGJK_OK;
  if (!d_points || !d_point_off || !d_faces || !d_face_off || !d_nfaces || !d_moments ||
      !d_com || !d_inertia_k)
    return ogjk_fail(OGJK_ERR_INVALID, "NULL argument");
  hipLaunchKernelGGL(k_hull_moments, dim3((unsigned)((n_clouds + 63) / 64)), dim3(64), 0,
                     (hipStream_t)stream, d_points, d_point_off, d_faces, d_face_off, d_nfaces,
                     n_clouds, d_moments, d_com, d_inertia_k);
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? OGJK_OK : hip_err(e, "moments launch");
}

ogjk_status ogjk_center_points(float* d_points, const int32_t* d_point_cloud, int64_t n_points,
                               const float* d_com, void* stream) {
  if (n_points < 0) return ogjk_fail(OGJK_ERR_INVALID, "n_points < 0");
  if (n_points == 0) return OGJK_OK;
  if (!d_points || !d_point_cloud || !d_com) return ogjk_fail(OGJK_ERR_INVALID, "NULL argument");
  const int64_t cb = (n_points + 255) / 256;
  hipLaunchKernelGGL(k_center, dim3((unsigned)(cb < 65536 ? cb : 65536)), dim3(256), 0,
                     (hipStream_t)stream, d_points, d_point_cloud, n_points, d_com);
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? OGJK_OK : hip_err(e, "centre launch");
}

}  // extern "C"
