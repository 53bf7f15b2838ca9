/*
 * opengjk-amd C++ drop-in API: the reference library's host entry points with
 * identical names, signatures and (Itanium) mangled symbols, implemented over
 * the C ABI in ogjk.h.  A program written against the reference headers
 * (GJK/gpu/openGJK.h, examples/gpu/example.h) links against libogjk.so.
 *
 * Semantics kept from the reference: synchronous calls, results valid on
 * return, n <= 0 is a no-op, `void` returns.  Errors that the reference never
 * reported are recorded instead: ogjk_last_error() / GJK::GPU::lastStatus().
 */
#ifndef OPENGJK_AMD_GPU_H
#define OPENGJK_AMD_GPU_H

#include <chrono>
#include <stdexcept>

#include <hip/hip_runtime.h>

#include "opengjk/common.h"
#include "opengjk/ogjk.h"

namespace GJK {
namespace Common {

/* examples/common/timer.h:15-99: hipEvent GPU timer + steady CPU timer. */
class PerformanceTimer {
 public:
  PerformanceTimer() {
    (void)hipEventCreate(&ev_start_);
    (void)hipEventCreate(&ev_end_);
  }
  ~PerformanceTimer() {
    (void)hipEventDestroy(ev_start_);
    (void)hipEventDestroy(ev_end_);
  }
  void startCpuTimer() {
    if (cpu_started_) throw std::runtime_error("CPU timer already started");
    cpu_started_ = true;
    t0_ = std::chrono::steady_clock::now();
  }
  void endCpuTimer() {
    const auto t1 = std::chrono::steady_clock::now();
    if (!cpu_started_) throw std::runtime_error("CPU timer not started");
    cpu_ms_ = (float)std::chrono::duration<double, std::milli>(t1 - t0_).count();
    cpu_started_ = false;
  }
  void startGpuTimer() {
    if (gpu_started_) throw std::runtime_error("GPU timer already started");
    gpu_started_ = true;
    (void)hipEventRecord(ev_start_, nullptr);
  }
  void endGpuTimer() {
    (void)hipEventRecord(ev_end_, nullptr);
    (void)hipEventSynchronize(ev_end_);
    if (!gpu_started_) throw std::runtime_error("GPU timer not started");
    (void)hipEventElapsedTime(&gpu_ms_, ev_start_, ev_end_);
    gpu_started_ = false;
  }
  float getCpuElapsedTimeForPreviousOperation() { return cpu_ms_; }
  float getGpuElapsedTimeForPreviousOperation() { return gpu_ms_; }
  PerformanceTimer(const PerformanceTimer&) = delete;
  PerformanceTimer(PerformanceTimer&&) = delete;
  PerformanceTimer& operator=(const PerformanceTimer&) = delete;
  PerformanceTimer& operator=(PerformanceTimer&&) = delete;

 private:
  hipEvent_t ev_start_ = nullptr, ev_end_ = nullptr;
  std::chrono::steady_clock::time_point t0_;
  bool cpu_started_ = false, gpu_started_ = false;
  float cpu_ms_ = 0.f, gpu_ms_ = 0.f;
};

}  // namespace Common

namespace GPU {

/* examples/gpu/example.h:9 */
GJK::Common::PerformanceTimer& timer();
/* Status of the calling thread's last drop-in call (no reference counterpart). */
ogjk_status lastStatus();

/* examples/gpu/example.h:21-25 */
void computeDistances(const int n, const gkPolytope* bd1, const gkPolytope* bd2,
                      gkSimplex* simplices, gkFloat* distances);
/* examples/gpu/example.h:41-46 */
void computeEPA(const int n, const gkPolytope* bd1, const gkPolytope* bd2, gkSimplex* simplices,
                gkFloat* distances, gkFloat* contact_normals);
/* examples/gpu/example.h:62-67 */
void computeGJKAndEPA(const int n, const gkPolytope* bd1, const gkPolytope* bd2,
                      gkSimplex* simplices, gkFloat* distances, gkFloat* contact_normals);
/* The README's EPA entry point (README.md:39-46), code form: same as computeEPA. */
void computeCollisionInformation(const int n, const gkPolytope* bd1, const gkPolytope* bd2,
                                 gkSimplex* simplices, gkFloat* distances,
                                 gkFloat* contact_normals);
/* README form (README.md:39-46): also returns the witnesses as n*3 arrays. */
void computeCollisionInformation(const int n, const gkPolytope* bd1, const gkPolytope* bd2,
                                 gkSimplex* simplices, gkFloat* distances, gkFloat* witness1,
                                 gkFloat* witness2, gkFloat* contact_normals = nullptr);

}  // namespace GPU
}  // namespace GJK

/* ---------------- flat API (GJK/gpu/openGJK.h:91-505) ---------------- */

void compute_minimum_distance(const int n, const gkPolytope* bd1, const gkPolytope* bd2,
                              gkSimplex* simplices, gkFloat* distances);
void computeCollisionInformation(const int n, const gkPolytope* bd1, const gkPolytope* bd2,
                                 gkSimplex* simplices, gkFloat* distances,
                                 gkFloat* contact_normals);
void compute_gjk_epa(const int n, const gkPolytope* bd1, const gkPolytope* bd2,
                     gkSimplex* simplices, gkFloat* distances, gkFloat* contact_normals);

void allocate_epa_device_arrays(const int n, gkFloat** d_witness1, gkFloat** d_witness2,
                                gkFloat** d_contact_normals);
void copy_epa_results_from_device(const int n, const gkFloat* d_witness1,
                                  const gkFloat* d_witness2, const gkFloat* d_contact_normals,
                                  gkFloat* witness1, gkFloat* witness2, gkFloat* contact_normals);
void free_epa_device_arrays(gkFloat* d_witness1, gkFloat* d_witness2, gkFloat* d_contact_normals);

void allocate_and_copy_device_arrays(const int n, const gkPolytope* bd1, const gkPolytope* bd2,
                                     gkPolytope** d_bd1, gkPolytope** d_bd2, gkFloat** d_coord1,
                                     gkFloat** d_coord2, gkSimplex** d_simplices,
                                     gkFloat** d_distances);
void compute_minimum_distance_device(const int n, const gkPolytope* d_bd1,
                                     const gkPolytope* d_bd2, gkSimplex* d_simplices,
                                     gkFloat* d_distances);
void compute_epa_device(const int n, const gkPolytope* d_bd1, const gkPolytope* d_bd2,
                        gkSimplex* d_simplices, gkFloat* d_distances, gkFloat* d_contact_normals);
void copy_results_from_device(const int n, const gkSimplex* d_simplices,
                              const gkFloat* d_distances, gkSimplex* simplices,
                              gkFloat* distances);
void free_device_arrays(gkPolytope* d_bd1, gkPolytope* d_bd2, gkFloat* d_coord1,
                        gkFloat* d_coord2, gkSimplex* d_simplices, gkFloat* d_distances);

void allocate_indexed_device(const int num_polytopes, const int max_pairs,
                             const gkPolytope* polytopes, gkPolytope** d_polytopes,
                             gkFloat** d_coords, gkCollisionPair** d_pairs,
                             gkSimplex** d_simplices, gkFloat** d_distances,
                             gkFloat** d_contact_normals);
void free_indexed_device(gkPolytope* d_polytopes, gkFloat* d_coords, gkCollisionPair* d_pairs,
                         gkSimplex* d_simplices, gkFloat* d_distances,
                         gkFloat* d_contact_normals);
void upload_pairs_device(const int num_pairs, const gkCollisionPair* pairs,
                         gkCollisionPair* d_pairs);

void compute_minimum_distance_indexed(const int num_polytopes, const int num_pairs,
                                      const gkPolytope* polytopes, const gkCollisionPair* pairs,
                                      gkSimplex* simplices, gkFloat* distances);
void compute_minimum_distance_indexed_device(const int num_pairs, const gkPolytope* d_polytopes,
                                             const gkCollisionPair* d_pairs,
                                             gkSimplex* d_simplices, gkFloat* d_distances);
void compute_epa_indexed_device(const int num_pairs, const gkPolytope* d_polytopes,
                                const gkCollisionPair* d_pairs, gkSimplex* d_simplices,
                                gkFloat* d_distances, gkFloat* d_contact_normals);
void compute_epa_indexed(const int num_polytopes, const int num_pairs,
                         const gkPolytope* polytopes, const gkCollisionPair* pairs,
                         gkSimplex* simplices, gkFloat* distances, gkFloat* contact_normals);
void compute_gjk_epa_indexed(const int num_polytopes, const int num_pairs,
                             const gkPolytope* polytopes, const gkCollisionPair* pairs,
                             gkSimplex* simplices, gkFloat* distances, gkFloat* contact_normals);

#endif /* OPENGJK_AMD_GPU_H */
