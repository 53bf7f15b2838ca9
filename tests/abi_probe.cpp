// Declares the reference library's host API exactly as its headers do
// (GJK/gpu/openGJK.h:91-505, examples/gpu/example.h:9-67, README.md:39-46) and
// takes each function's address.  Compiled with g++ -c, the object's undefined
// symbols are the mangled names a program built against the reference needs;
// tests/test_abi.py checks that libogjk.so defines every one of them.
#include "opengjk/common.h"

namespace GJK {
namespace Common { class PerformanceTimer; }
namespace GPU {
GJK::Common::PerformanceTimer& timer();
void computeDistances(const int n, const gkPolytope* bd1, const gkPolytope* bd2,
                      gkSimplex* simplices, gkFloat* distances);
void computeEPA(const int n, const gkPolytope* bd1, const gkPolytope* bd2, gkSimplex* simplices,
                gkFloat* distances, gkFloat* contact_normals);
void computeGJKAndEPA(const int n, const gkPolytope* bd1, const gkPolytope* bd2,
                      gkSimplex* simplices, gkFloat* distances, gkFloat* contact_normals);
void computeCollisionInformation(const int n, const gkPolytope* bd1, const gkPolytope* bd2,
                                 gkSimplex* simplices, gkFloat* distances, gkFloat* witness1,
                                 gkFloat* witness2, gkFloat* contact_normals);
}  // namespace GPU
}  // namespace GJK

void compute_minimum_distance(const int n, const gkPolytope* bd1, const gkPolytope* bd2,
                              gkSimplex* simplices, gkFloat* distances);
void computeCollisionInformation(const int n, const gkPolytope* bd1, const gkPolytope* bd2,
                                 gkSimplex* simplices, gkFloat* distances, gkFloat* contact_normals);
void compute_gjk_epa(const int n, const gkPolytope* bd1, const gkPolytope* bd2,
                     gkSimplex* simplices, gkFloat* distances, gkFloat* contact_normals);
void allocate_epa_device_arrays(const int n, gkFloat** d_witness1, gkFloat** d_witness2,
                                gkFloat** d_contact_normals);
void copy_epa_results_from_device(const int n, const gkFloat* d_witness1, const gkFloat* d_witness2,
                                  const gkFloat* d_contact_normals, gkFloat* witness1,
                                  gkFloat* witness2, gkFloat* contact_normals);
void free_epa_device_arrays(gkFloat* d_witness1, gkFloat* d_witness2, gkFloat* d_contact_normals);
void allocate_and_copy_device_arrays(const int n, const gkPolytope* bd1, const gkPolytope* bd2,
                                     gkPolytope** d_bd1, gkPolytope** d_bd2, gkFloat** d_coord1,
                                     gkFloat** d_coord2, gkSimplex** d_simplices,
                                     gkFloat** d_distances);
void compute_minimum_distance_device(const int n, const gkPolytope* d_bd1, const gkPolytope* d_bd2,
                                     gkSimplex* d_simplices, gkFloat* d_distances);
void compute_epa_device(const int n, const gkPolytope* d_bd1, const gkPolytope* d_bd2,
                        gkSimplex* d_simplices, gkFloat* d_distances, gkFloat* d_contact_normals);
void copy_results_from_device(const int n, const gkSimplex* d_simplices, const gkFloat* d_distances,
                              gkSimplex* simplices, gkFloat* distances);
void free_device_arrays(gkPolytope* d_bd1, gkPolytope* d_bd2, gkFloat* d_coord1, gkFloat* d_coord2,
                        gkSimplex* d_simplices, gkFloat* d_distances);
void allocate_indexed_device(const int num_polytopes, const int max_pairs,
                             const gkPolytope* polytopes, gkPolytope** d_polytopes,
                             gkFloat** d_coords, gkCollisionPair** d_pairs, gkSimplex** d_simplices,
                             gkFloat** d_distances, gkFloat** d_contact_normals);
void free_indexed_device(gkPolytope* d_polytopes, gkFloat* d_coords, gkCollisionPair* d_pairs,
                         gkSimplex* d_simplices, gkFloat* d_distances, gkFloat* d_contact_normals);
void upload_pairs_device(const int num_pairs, const gkCollisionPair* pairs,
                         gkCollisionPair* d_pairs);
void compute_minimum_distance_indexed(const int num_polytopes, const int num_pairs,
                                      const gkPolytope* polytopes, const gkCollisionPair* pairs,
                                      gkSimplex* simplices, gkFloat* distances);
void compute_minimum_distance_indexed_device(const int num_pairs, const gkPolytope* d_polytopes,
                                             const gkCollisionPair* d_pairs, gkSimplex* d_simplices,
                                             gkFloat* d_distances);
void compute_epa_indexed_device(const int num_pairs, const gkPolytope* d_polytopes,
                                const gkCollisionPair* d_pairs, gkSimplex* d_simplices,
                                gkFloat* d_distances, gkFloat* d_contact_normals);
void compute_epa_indexed(const int num_polytopes, const int num_pairs, const gkPolytope* polytopes,
                         const gkCollisionPair* pairs, gkSimplex* simplices, gkFloat* distances,
                         gkFloat* contact_normals);
void compute_gjk_epa_indexed(const int num_polytopes, const int num_pairs,
                             const gkPolytope* polytopes, const gkCollisionPair* pairs,
                             gkSimplex* simplices, gkFloat* distances, gkFloat* contact_normals);

void* const kRefs[] = {
    (void*)&GJK::GPU::timer, (void*)&GJK::GPU::computeDistances, (void*)&GJK::GPU::computeEPA,
    (void*)&GJK::GPU::computeGJKAndEPA,
    (void*)static_cast<void (*)(int, const gkPolytope*, const gkPolytope*, gkSimplex*, gkFloat*,
                                gkFloat*, gkFloat*, gkFloat*)>(
        &GJK::GPU::computeCollisionInformation),
    (void*)&compute_minimum_distance, (void*)&computeCollisionInformation, (void*)&compute_gjk_epa,
    (void*)&allocate_epa_device_arrays, (void*)&copy_epa_results_from_device,
    (void*)&free_epa_device_arrays, (void*)&allocate_and_copy_device_arrays,
    (void*)&compute_minimum_distance_device, (void*)&compute_epa_device,
    (void*)&copy_results_from_device, (void*)&free_device_arrays, (void*)&allocate_indexed_device,
    (void*)&free_indexed_device, (void*)&upload_pairs_device,
    (void*)&compute_minimum_distance_indexed, (void*)&compute_minimum_distance_indexed_device,
    (void*)&compute_epa_indexed_device, (void*)&compute_epa_indexed,
    (void*)&compute_gjk_epa_indexed,
};
