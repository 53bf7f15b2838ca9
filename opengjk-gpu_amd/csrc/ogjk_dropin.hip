// opengjk-amd C++ drop-in layer: the reference's host API names over the C
// ABI.  Each function cites the reference definition it replaces.
#include <hip/hip_runtime.h>

#include <cstring>
#include <vector>

#include "opengjk/gpu.h"

namespace {

thread_local ogjk_status g_last = OGJK_OK;

void note(ogjk_status s) { g_last = s; }

template <class T>
void dmalloc(T** p, size_t count) {
  *p = nullptr;
  if (count == 0) return;
  if (hipMalloc((void**)p, count * sizeof(T)) != hipSuccess) {
    *p = nullptr;
    note(OGJK_ERR_NOMEM);
  }
}

// Concatenate the bodies' AoS coordinates into one device buffer and build
// device gkPolytope descriptors that point into it (openGJK.cu:2889-2954).
void upload_polytopes(int n, const gkPolytope* host, gkPolytope** d_desc, gkFloat** d_coords) {
  size_t total = 0;
  for (int i = 0; i < n; i++) total += (size_t)(host[i].numpoints > 0 ? host[i].numpoints : 0) * 3;
  dmalloc(d_desc, (size_t)n);
  dmalloc(d_coords, total > 0 ? total : 1);
  if (!*d_desc || !*d_coords) return;
  std::vector<gkPolytope> tmp(host, host + n);
  std::vector<gkFloat> staging(total);
  size_t o = 0;
  for (int i = 0; i < n; i++) {
    const size_t k = (size_t)(host[i].numpoints > 0 ? host[i].numpoints : 0) * 3;
    if (k) std::memcpy(staging.data() + o, host[i].coord, k * sizeof(gkFloat));
    tmp[i].coord = *d_coords + o;
    o += k;
  }
  if (total) (void)hipMemcpy(*d_coords, staging.data(), total * sizeof(gkFloat), hipMemcpyHostToDevice);
  (void)hipMemcpy(*d_desc, tmp.data(), (size_t)n * sizeof(gkPolytope), hipMemcpyHostToDevice);
}

}  // namespace

namespace GJK {
namespace GPU {

// examples/gpu/example.cu:17-21
GJK::Common::PerformanceTimer& timer() {
  static GJK::Common::PerformanceTimer t;
  return t;
}

ogjk_status lastStatus() { return g_last; }

// examples/gpu/example.cu:23-52 (kernel-only timing, as the reference)
void computeDistances(const int n, const gkPolytope* bd1, const gkPolytope* bd2,
                      gkSimplex* simplices, gkFloat* distances) {
  if (n <= 0) return;
  gkPolytope *d1 = nullptr, *d2 = nullptr;
  gkFloat *c1 = nullptr, *c2 = nullptr, *dd = nullptr;
  gkSimplex* ds = nullptr;
  allocate_and_copy_device_arrays(n, bd1, bd2, &d1, &d2, &c1, &c2, &ds, &dd);
  timer().startGpuTimer();
  compute_minimum_distance_device(n, d1, d2, ds, dd);
  timer().endGpuTimer();
  copy_results_from_device(n, ds, dd, simplices, distances);
  free_device_arrays(d1, d2, c1, c2, ds, dd);
}

// examples/gpu/example.cu:54-84
void computeGJKAndEPA(const int n, const gkPolytope* bd1, const gkPolytope* bd2,
                      gkSimplex* simplices, gkFloat* distances, gkFloat* contact_normals) {
  if (n <= 0) return;
  gkPolytope *d1 = nullptr, *d2 = nullptr;
  gkFloat *c1 = nullptr, *c2 = nullptr, *dd = nullptr, *dn = nullptr;
  gkSimplex* ds = nullptr;
  allocate_and_copy_device_arrays(n, bd1, bd2, &d1, &d2, &c1, &c2, &ds, &dd);
  dmalloc(&dn, (size_t)n * 3);
  timer().startGpuTimer();
  // One fused kernel instead of the reference's two back-to-back launches.
  note(ogjk_collide_polytopes(OGJK_MODE_GJK_EPA, d1, d2, n, ds, dd, dn, nullptr));
  (void)hipDeviceSynchronize();
  timer().endGpuTimer();
  copy_results_from_device(n, ds, dd, simplices, distances);
  if (contact_normals && dn)
    (void)hipMemcpy(contact_normals, dn, (size_t)n * 3 * sizeof(gkFloat), hipMemcpyDeviceToHost);
  free_device_arrays(d1, d2, c1, c2, ds, dd);
  if (dn) (void)hipFree(dn);
}

// examples/gpu/example.cu:86-119
void computeEPA(const int n, const gkPolytope* bd1, const gkPolytope* bd2, gkSimplex* simplices,
                gkFloat* distances, gkFloat* contact_normals) {
  if (n <= 0) return;
  gkPolytope *d1 = nullptr, *d2 = nullptr;
  gkFloat *c1 = nullptr, *c2 = nullptr, *dd = nullptr, *dn = nullptr;
  gkSimplex* ds = nullptr;
  allocate_and_copy_device_arrays(n, bd1, bd2, &d1, &d2, &c1, &c2, &ds, &dd);
  if (ds) (void)hipMemcpy(ds, simplices, (size_t)n * sizeof(gkSimplex), hipMemcpyHostToDevice);
  if (dd) (void)hipMemcpy(dd, distances, (size_t)n * sizeof(gkFloat), hipMemcpyHostToDevice);
  dmalloc(&dn, (size_t)n * 3);
  timer().startGpuTimer();
  compute_epa_device(n, d1, d2, ds, dd, dn);
  timer().endGpuTimer();
  copy_results_from_device(n, ds, dd, simplices, distances);
  if (contact_normals && dn)
    (void)hipMemcpy(contact_normals, dn, (size_t)n * 3 * sizeof(gkFloat), hipMemcpyDeviceToHost);
  free_device_arrays(d1, d2, c1, c2, ds, dd);
  if (dn) (void)hipFree(dn);
}

// README.md:39-46 (code form: openGJK.h:112-119 semantics)
void computeCollisionInformation(const int n, const gkPolytope* bd1, const gkPolytope* bd2,
                                 gkSimplex* simplices, gkFloat* distances,
                                 gkFloat* contact_normals) {
  computeEPA(n, bd1, bd2, simplices, distances, contact_normals);
}

// README.md:39-46 (README form with separate witness arrays)
void computeCollisionInformation(const int n, const gkPolytope* bd1, const gkPolytope* bd2,
                                 gkSimplex* simplices, gkFloat* distances, gkFloat* witness1,
                                 gkFloat* witness2, gkFloat* contact_normals) {
  if (n <= 0) return;
  std::vector<gkFloat> nrm;
  gkFloat* nptr = contact_normals;
  if (!nptr) {
    nrm.resize((size_t)n * 3);
    nptr = nrm.data();
  }
  computeEPA(n, bd1, bd2, simplices, distances, nptr);
  for (int i = 0; i < n; i++) {
    for (int c = 0; c < 3; c++) {
      if (witness1) witness1[3 * i + c] = simplices[i].witnesses[0][c];
      if (witness2) witness2[3 * i + c] = simplices[i].witnesses[1][c];
    }
  }
}

}  // namespace GPU
}  // namespace GJK

// ------------------------------------------------------------- flat API

// GJK/gpu/openGJK.cu:2791-2819
void compute_minimum_distance(const int n, const gkPolytope* bd1, const gkPolytope* bd2,
                              gkSimplex* simplices, gkFloat* distances) {
  if (n <= 0) return;
  note(ogjk_collide_host(OGJK_MODE_GJK, n, bd1, bd2, simplices, distances, nullptr));
}

// GJK/gpu/openGJK.cu:2821-2852.  The distances array is uploaded as input:
// EPA decides collision from it (the reference omits that copy here, while
// its namespaced computeEPA performs it, examples/gpu/example.cu:104-105).
void computeCollisionInformation(const int n, const gkPolytope* bd1, const gkPolytope* bd2,
                                 gkSimplex* simplices, gkFloat* distances,
                                 gkFloat* contact_normals) {
  if (n <= 0) return;
  note(ogjk_collide_host(OGJK_MODE_EPA, n, bd1, bd2, simplices, distances, contact_normals));
}

// GJK/gpu/openGJK.cu:2854-2883
void compute_gjk_epa(const int n, const gkPolytope* bd1, const gkPolytope* bd2,
                     gkSimplex* simplices, gkFloat* distances, gkFloat* contact_normals) {
  if (n <= 0) return;
  note(ogjk_collide_host(OGJK_MODE_GJK_EPA, n, bd1, bd2, simplices, distances, contact_normals));
}

// GJK/gpu/openGJK.cu:2956-2966
void allocate_epa_device_arrays(const int n, gkFloat** d_witness1, gkFloat** d_witness2,
                                gkFloat** d_contact_normals) {
  dmalloc(d_witness1, (size_t)n * 3);
  dmalloc(d_witness2, (size_t)n * 3);
  if (d_contact_normals) dmalloc(d_contact_normals, (size_t)n * 3);
}

// GJK/gpu/openGJK.cu:3017-3032
void copy_epa_results_from_device(const int n, const gkFloat* d_witness1,
                                  const gkFloat* d_witness2, const gkFloat* d_contact_normals,
                                  gkFloat* witness1, gkFloat* witness2, gkFloat* contact_normals) {
  const size_t b = (size_t)n * 3 * sizeof(gkFloat);
  (void)hipMemcpy(witness1, d_witness1, b, hipMemcpyDeviceToHost);
  (void)hipMemcpy(witness2, d_witness2, b, hipMemcpyDeviceToHost);
  if (contact_normals && d_contact_normals)
    (void)hipMemcpy(contact_normals, d_contact_normals, b, hipMemcpyDeviceToHost);
}

// GJK/gpu/openGJK.cu:3050-3060
void free_epa_device_arrays(gkFloat* d_witness1, gkFloat* d_witness2, gkFloat* d_contact_normals) {
  if (d_witness1) (void)hipFree(d_witness1);
  if (d_witness2) (void)hipFree(d_witness2);
  if (d_contact_normals) (void)hipFree(d_contact_normals);
}

// GJK/gpu/openGJK.cu:2889-2954
void allocate_and_copy_device_arrays(const int n, const gkPolytope* bd1, const gkPolytope* bd2,
                                     gkPolytope** d_bd1, gkPolytope** d_bd2, gkFloat** d_coord1,
                                     gkFloat** d_coord2, gkSimplex** d_simplices,
                                     gkFloat** d_distances) {
  upload_polytopes(n, bd1, d_bd1, d_coord1);
  upload_polytopes(n, bd2, d_bd2, d_coord2);
  dmalloc(d_simplices, (size_t)n);
  dmalloc(d_distances, (size_t)n);
  if (*d_simplices) (void)hipMemset(*d_simplices, 0, (size_t)n * sizeof(gkSimplex));
}

// GJK/gpu/openGJK.cu:2968-2984
void compute_minimum_distance_device(const int n, const gkPolytope* d_bd1,
                                     const gkPolytope* d_bd2, gkSimplex* d_simplices,
                                     gkFloat* d_distances) {
  if (n <= 0) return;
  note(ogjk_collide_polytopes(OGJK_MODE_GJK, d_bd1, d_bd2, n, d_simplices, d_distances, nullptr,
                              nullptr));
  (void)hipDeviceSynchronize();
}

// GJK/gpu/openGJK.cu:2986-3004
void compute_epa_device(const int n, const gkPolytope* d_bd1, const gkPolytope* d_bd2,
                        gkSimplex* d_simplices, gkFloat* d_distances, gkFloat* d_contact_normals) {
  if (n <= 0) return;
  note(ogjk_collide_polytopes(OGJK_MODE_EPA, d_bd1, d_bd2, n, d_simplices, d_distances,
                              d_contact_normals, nullptr));
  (void)hipDeviceSynchronize();
}

// GJK/gpu/openGJK.cu:3006-3015
void copy_results_from_device(const int n, const gkSimplex* d_simplices,
                              const gkFloat* d_distances, gkSimplex* simplices,
                              gkFloat* distances) {
  if (n <= 0) return;
  (void)hipMemcpy(distances, d_distances, (size_t)n * sizeof(gkFloat), hipMemcpyDeviceToHost);
  (void)hipMemcpy(simplices, d_simplices, (size_t)n * sizeof(gkSimplex), hipMemcpyDeviceToHost);
}

// GJK/gpu/openGJK.cu:3034-3048
void free_device_arrays(gkPolytope* d_bd1, gkPolytope* d_bd2, gkFloat* d_coord1,
                        gkFloat* d_coord2, gkSimplex* d_simplices, gkFloat* d_distances) {
  if (d_bd1) (void)hipFree(d_bd1);
  if (d_bd2) (void)hipFree(d_bd2);
  if (d_coord1) (void)hipFree(d_coord1);
  if (d_coord2) (void)hipFree(d_coord2);
  if (d_simplices) (void)hipFree(d_simplices);
  if (d_distances) (void)hipFree(d_distances);
}

// GJK/gpu/openGJK.cu:3165-3233
void allocate_indexed_device(const int num_polytopes, const int max_pairs,
                             const gkPolytope* polytopes, gkPolytope** d_polytopes,
                             gkFloat** d_coords, gkCollisionPair** d_pairs,
                             gkSimplex** d_simplices, gkFloat** d_distances,
                             gkFloat** d_contact_normals) {
  upload_polytopes(num_polytopes, polytopes, d_polytopes, d_coords);
  dmalloc(d_pairs, (size_t)max_pairs);
  dmalloc(d_simplices, (size_t)max_pairs);
  dmalloc(d_distances, (size_t)max_pairs);
  if (d_contact_normals) dmalloc(d_contact_normals, (size_t)max_pairs * 3);
}

void free_indexed_device(gkPolytope* d_polytopes, gkFloat* d_coords, gkCollisionPair* d_pairs,
                         gkSimplex* d_simplices, gkFloat* d_distances,
                         gkFloat* d_contact_normals) {
  if (d_polytopes) (void)hipFree(d_polytopes);
  if (d_coords) (void)hipFree(d_coords);
  if (d_pairs) (void)hipFree(d_pairs);
  if (d_simplices) (void)hipFree(d_simplices);
  if (d_distances) (void)hipFree(d_distances);
  if (d_contact_normals) (void)hipFree(d_contact_normals);
}

void upload_pairs_device(const int num_pairs, const gkCollisionPair* pairs,
                         gkCollisionPair* d_pairs) {
  if (num_pairs <= 0) return;
  (void)hipMemcpy(d_pairs, pairs, (size_t)num_pairs * sizeof(gkCollisionPair),
                  hipMemcpyHostToDevice);
}

// GJK/gpu/openGJK.cu:3064-3125
void compute_minimum_distance_indexed(const int num_polytopes, const int num_pairs,
                                      const gkPolytope* polytopes, const gkCollisionPair* pairs,
                                      gkSimplex* simplices, gkFloat* distances) {
  if (num_pairs <= 0 || num_polytopes <= 0) return;
  note(ogjk_collide_host_indexed(OGJK_MODE_GJK, num_polytopes, num_pairs, polytopes, pairs,
                                 simplices, distances, nullptr));
}

// GJK/gpu/openGJK.cu:3127-3142
void compute_minimum_distance_indexed_device(const int num_pairs, const gkPolytope* d_polytopes,
                                             const gkCollisionPair* d_pairs,
                                             gkSimplex* d_simplices, gkFloat* d_distances) {
  if (num_pairs <= 0) return;
  // The reference kernel trusts the indices; the pool size is unknown here,
  // so the range check is disabled with the largest pool.
  note(ogjk_collide_polytopes_indexed(OGJK_MODE_GJK, d_polytopes, 0x7fffffff, d_pairs, num_pairs,
                                      d_simplices, d_distances, nullptr, nullptr));
  (void)hipDeviceSynchronize();
}

// GJK/gpu/openGJK.cu:3144-3162
void compute_epa_indexed_device(const int num_pairs, const gkPolytope* d_polytopes,
                                const gkCollisionPair* d_pairs, gkSimplex* d_simplices,
                                gkFloat* d_distances, gkFloat* d_contact_normals) {
  if (num_pairs <= 0) return;
  note(ogjk_collide_polytopes_indexed(OGJK_MODE_EPA, d_polytopes, 0x7fffffff, d_pairs, num_pairs,
                                      d_simplices, d_distances, d_contact_normals, nullptr));
  (void)hipDeviceSynchronize();
}

// GJK/gpu/openGJK.cu:3235-3272
void compute_epa_indexed(const int num_polytopes, const int num_pairs,
                         const gkPolytope* polytopes, const gkCollisionPair* pairs,
                         gkSimplex* simplices, gkFloat* distances, gkFloat* contact_normals) {
  if (num_pairs <= 0) return;
  note(ogjk_collide_host_indexed(OGJK_MODE_EPA, num_polytopes, num_pairs, polytopes, pairs,
                                 simplices, distances, contact_normals));
}

// GJK/gpu/openGJK.cu:3274-3311
void compute_gjk_epa_indexed(const int num_polytopes, const int num_pairs,
                             const gkPolytope* polytopes, const gkCollisionPair* pairs,
                             gkSimplex* simplices, gkFloat* distances, gkFloat* contact_normals) {
  if (num_pairs <= 0) return;
  note(ogjk_collide_host_indexed(OGJK_MODE_GJK_EPA, num_polytopes, num_pairs, polytopes, pairs,
                                 simplices, distances, contact_normals));
}
