/*
 * opengjk-amd C ABI (libogjk.so): batched GJK distance and EPA penetration
 * for pairs of convex polytopes on AMD Instinct MI355X (gfx950).
 *
 * Plain pointers and sizes only.  Every entry point returns an ogjk_status;
 * ogjk_last_error() gives the calling thread's last message.  Results are
 * those of the FP64 reference CPU algorithm (GJK/cpu/openGJK.c, GJK/cpu/EPA.c)
 * rounded once to the gkFloat ABI.
 *
 * Mode (what a call computes), mirroring the reference entry points:
 *   OGJK_MODE_GJK      distance + simplex + witnesses
 *                      (compute_minimum_distance_kernel, GJK/gpu/openGJK.cu:1427)
 *   OGJK_MODE_EPA      EPA only; reads simplices/distances written by GJK
 *                      (compute_epa_kernel, GJK/gpu/openGJK.cu:2729)
 *   OGJK_MODE_GJK_EPA  both, fused in one kernel (GJK::GPU::computeGJKAndEPA,
 *                      examples/gpu/example.cu:54-84)
 *
 * Output conventions (GJK/gpu/openGJK.h:64-119): distance >= 0 for separated
 * pairs; EPA writes distance = -depth for penetrating pairs, overwrites
 * simplices[i].witnesses (witness on body 1, on body 2) and writes the unit
 * contact normal (3 floats per pair, AoS).  A pair with a body of fewer than
 * one vertex or an out-of-range pair index gets distance NaN, nvrtx 0.
 */
#ifndef OPENGJK_AMD_OGJK_H
#define OPENGJK_AMD_OGJK_H

#include <stddef.h>
#include <stdint.h>

#include "opengjk/common.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef enum ogjk_status {
  OGJK_OK = 0,
  OGJK_ERR_INVALID = 1, /* bad argument */
  OGJK_ERR_HIP = 2,     /* HIP runtime error (message in ogjk_last_error) */
  OGJK_ERR_NOMEM = 3,   /* device or host allocation failed */
  OGJK_ERR_UNSUPPORTED = 4
} ogjk_status;

enum { OGJK_MODE_GJK = 1, OGJK_MODE_EPA = 2, OGJK_MODE_GJK_EPA = 3 };

/* Device-resident structure-of-arrays scene.  Body b owns vertices
 * [body_off[b], body_off[b] + body_cnt[b]) of x/y/z; body_r1[b] >=
 * max |x|+|y|+|z| over its vertices (written by ogjk_pack_scene). */
/* Vertex clusters of a scene (optional; see ogjk_build_clusters).  Each body's
 * vertices are copied in order of direction from the body's centre and cut
 * into clusters of 32 slots with a bounding sphere and box each; support sweeps over bodies of at least min_verts
 * vertices then skip every cluster that cannot hold the maximum.  Results are
 * identical with and without clusters.  All arrays are device memory. */
typedef struct ogjk_clusters {
  float* x;            /* n_slots: vertices in cluster order */
  float* y;
  float* z;
  int32_t* orig;       /* n_slots: body-local input index of each slot (-1: padding) */
  int64_t* body_slot;  /* n_bodies: first slot of each body (a multiple of 32), or
                          -1: not built (fewer than min_verts vertices at the build) */
  float* bounds;       /* 8 * n_slots / 32: per cluster centre x, y, z, radius,
                          half extents x, y, z, unused */
  int64_t n_slots;     /* capacity of x, y, z, orig: >= ogjk_cluster_slots() */
  int32_t min_verts;   /* smaller bodies are swept without culling; the build
                          clusters only bodies of at least min_verts vertices, so
                          lowering it afterwards culls no further bodies */
  float* super_bounds; /* 8 * n_supers: bounds of each run of 32 clusters of the
                          bodies of more than 4,096 vertices, as bounds */
  int64_t* body_super; /* n_bodies: first supercluster of each body, -1: none */
  int64_t n_supers;    /* capacity of super_bounds: >= ogjk_cluster_supers(n_slots) */
  int64_t n_supers_used; /* set by the build: superclusters written */
} ogjk_clusters;

typedef struct ogjk_scene {
  const float* x;
  const float* y;
  const float* z;
  const int64_t* body_off;
  const int32_t* body_cnt;
  const float* body_r1;
  int32_t n_bodies;
  const ogjk_clusters* clusters; /* host pointer, or NULL: no culling */
} ogjk_scene;

/* Plain-FP64 scene (double coordinates) for the comparison path: SoA x/y/z,
 * or, when xyz is not NULL, one AoS pool [x0, y0, z0, x1, ...] (the
 * reference's gkPolytope coordinate layout at FP64) and x/y/z are ignored. */
typedef struct ogjk_scene64 {
  const double* x;
  const double* y;
  const double* z;
  const int64_t* body_off;
  const int32_t* body_cnt;
  int32_t n_bodies;
  const double* xyz;
} ogjk_scene64;

/* Per-pair FP64 record for verification (optional, may be NULL):
 * [distance, w0.xyz, w1.xyz, normal.xyz, vrtx[0..3].xyz, 2 unused] */
enum { OGJK_F64_RECORD = 24 };

/* ------------------------------------------------------------ library */

int ogjk_version(void);
const char* ogjk_last_error(void);
/* Number of visible HIP devices (0 without a GPU; never fails). */
int ogjk_device_count(void);
ogjk_status ogjk_sync(void* stream);
/* Launch shape, process-wide, read at launch time: lanes of a wave64 per
 * pair in the GJK kernel (8, 16, 32, 64; default 8) and in the EPA kernel
 * (16, 32, 64; default 32).  fused 0: GJK -> device worklist -> persistent
 * EPA; 1: GJK+EPA as one kernel, one group per pair; 2: the reference's
 * launch shape for the plain-FP64 comparison path (ogjk_collide_scene_f64):
 * a GJK kernel over every pair, then an EPA kernel over every pair, one group
 * per pair each, non-persistent (GJK/gpu/openGJK.cu:2968-3004), where GJK
 * lanes may also be 1 (one thread per pair); other calls run as fused 0.
 * Not synchronised with launches issued concurrently from other threads. */
ogjk_status ogjk_set_config(int gjk_lanes, int epa_lanes, int fused);

/* Size-bucketed launch shape: pairs with max(nA, nB) <= max_verts[b] (and
 * above the previous bound) form bucket b; the last bucket is open-ended
 * (its max_verts entry is ignored).  Each bucket runs GJK on gjk_lanes[b]
 * (8, 16, 32, 64) and EPA on epa_lanes[b] (16, 32, 64) lanes per pair, all
 * scheduled on the device (classify, scatter, persistent per-bucket kernels).
 * Same process-wide scope as ogjk_set_config, which it replaces. */
ogjk_status ogjk_set_buckets(int n_buckets, const int* max_verts, const int* gjk_lanes,
                             const int* epa_lanes);

/* Pairs per workgroup of the size buckets set by ogjk_set_config or
 * ogjk_set_buckets (which reset them to GJK 1 and EPA 4 waves per block):
 * waves per block of the GJK kernel and of the persistent EPA kernel (1, 2
 * or 4) for each of the n_buckets buckets; pairs per block = 64 * waves /
 * lanes per pair.  EPA runs one launch per lane class, shaped by the first
 * bucket of the class.  Same process-wide scope as ogjk_set_config. */
ogjk_status ogjk_set_block_waves(int n_buckets, const int* gjk_waves, const int* epa_waves);

/* Device workspace bytes a GJK or GJK+EPA call over n_pairs needs. */
size_t ogjk_workspace_bytes(int32_t n_pairs);

/* ------------------------------------------------------------ scenes */

/* Pack an AoS float pool (device) into the SoA scene arrays (device) and
 * compute body_r1.  n_verts = total vertices. */
ogjk_status ogjk_pack_scene(const float* d_coords_aos, const int64_t* d_body_off,
                            const int32_t* d_body_cnt, int32_t n_bodies, int64_t n_verts,
                            float* d_x, float* d_y, float* d_z, float* d_r1, void* stream);

/* Pairs over a resident scene; d_pairs holds 2 int32 per pair (idx1, idx2)
 * or is NULL for the implicit pairs (2i, 2i+1).  Asynchronous on `stream`.
 * Replaces compute_minimum_distance_indexed_device / compute_epa_indexed_device
 * (GJK/gpu/openGJK.h:419-457) for SoA scenes. */
ogjk_status ogjk_collide_scene(int mode, const ogjk_scene* scene, const int32_t* d_pairs,
                               int32_t n_pairs, gkSimplex* d_simplices, float* d_distances,
                               float* d_normals, double* d_f64, void* stream);

/* Same, with a caller-provided device workspace (>= ogjk_workspace_bytes;
 * used by OGJK_MODE_GJK and OGJK_MODE_GJK_EPA).  Calls without it share one library-owned
 * workspace per device and must not overlap on that device. */
ogjk_status ogjk_collide_scene_ws(int mode, const ogjk_scene* scene, const int32_t* d_pairs,
                                  int32_t n_pairs, gkSimplex* d_simplices, float* d_distances,
                                  float* d_normals, double* d_f64, void* d_workspace,
                                  size_t workspace_bytes, void* stream);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): as ogjk_collide_scene, but the GJK -> EPA workspace
 * (counters, bucketed pair list, EPA jobs) is a static device array of the
 * library sized for OGJK_NOALLOC_MAX_PAIRS pairs (12.3 MiB, one per device,
 * held until the process ends).  Calls no hipMalloc/hipFree.  More than
 * OGJK_NOALLOC_MAX_PAIRS pairs is OGJK_ERR_INVALID.  Every call on one device
 * shares that array (its counters are cleared by each launch), so calls that
 * overlap on one device -- other streams, other threads -- race. */
#define OGJK_NOALLOC_MAX_PAIRS 65536
ogjk_status ogjk_collide_scene_noalloc(int mode, const ogjk_scene* scene, const int32_t* d_pairs,
                                       int32_t n_pairs, gkSimplex* d_simplices,
                                       float* d_distances, float* d_normals, double* d_f64,
                                       void* stream);

/* A GJK / GJK+EPA launch over a scene captured as one HIP graph: the whole
 * sequence (counter reset, size buckets, GJK, worklist, persistent EPA) is
 * replayed by a single ogjk_graph_launch, with no host work per batch.
 * Pointers, the pair count and the launch configuration are fixed at
 * capture; the data behind the pointers (coordinates, pair lists) may change
 * between launches.  The graph owns its workspace, so launches of different
 * graphs may overlap; launches of one graph must not.  No counterpart in
 * the reference (its launches are synchronous, GJK/gpu/openGJK.cu:2968-3003). */
typedef struct ogjk_graph_s* ogjk_graph;
ogjk_status ogjk_graph_capture_scene(int mode, const ogjk_scene* scene, const int32_t* d_pairs,
                                     int32_t n_pairs, gkSimplex* d_simplices, float* d_distances,
                                     float* d_normals, double* d_f64, ogjk_graph* graph);
ogjk_status ogjk_graph_launch(ogjk_graph graph, void* stream);
ogjk_status ogjk_graph_destroy(ogjk_graph graph);

/* Scene bodies under per-pair rigid poses: pair i is (d_pairs[2i],
 * d_pairs[2i+1]) placed by d_poses[14i..14i+6] and d_poses[14i+7..14i+13],
 * each pose = unit quaternion (w, x, y, z) then translation (x, y, z); world
 * vertex = R(q) x + t in FP64.  Results are in world coordinates and equal a
 * run of the reference on the FP64 world vertices.  Asynchronous on stream;
 * d_workspace as for ogjk_collide_scene_ws.  Clusters are not used.
 * Replaces a transform pass (visualization/integrate_final_gjk.cu:102-118,
 * 304-332) followed by GJK::GPU::computeGJKAndEPA (examples/gpu/example.h:38-43). */
ogjk_status ogjk_collide_posed(int mode, const ogjk_scene* scene, const int32_t* d_pairs,
                               const float* d_poses, int32_t n_pairs, gkSimplex* d_simplices,
                               float* d_distances, float* d_normals, double* d_f64,
                               void* d_workspace, size_t workspace_bytes, void* stream);

/* As ogjk_collide_posed with one pose per scene body: body b is placed by
 * d_body_poses[7b..7b+6], and pair i is d_pairs[i] (a broad-phase list, see
 * ogjk_broad_phase).  Replaces sim_step's transform_to_world_kernel followed
 * by compute_minimum_distance_indexed_device and compute_epa_indexed_device
 * (visualization/integrate_final_gjk.cu:304-332, 1016-1035). */
ogjk_status ogjk_collide_posed_bodies(int mode, const ogjk_scene* scene,
                                      const gkCollisionPair* d_pairs, const float* d_body_poses,
                                      int32_t n_pairs, gkSimplex* d_simplices,
                                      float* d_distances, float* d_normals, double* d_f64,
                                      void* d_workspace, size_t workspace_bytes, void* stream);

/* Compact contact record, 40 B per pair: what a contact consumer reads.  The
 * reference's collision_response_kernel reads distances[p], epa_normals[3p..]
 * and simplices[p].witnesses (visualization/integrate_final_gjk.cu:572-689),
 * 40 of the 124 bytes the gkSimplex outputs hold.  Values are those of the
 * gkSimplex/distance/normal outputs of the same launch, bit for bit; the
 * normal is zero in GJK-only mode, and a pair with an invalid body has every
 * field NaN. */
typedef struct ogjk_contact {
  float distance;      /* > 0: separation; <= 0: -penetration depth */
  float normal[3];
  float witness[2][3]; /* on A, on B */
} ogjk_contact;
OGJK_STATIC_ASSERT(sizeof(ogjk_contact) == 40, "ogjk_contact layout");
OGJK_STATIC_ASSERT(offsetof(ogjk_contact, witness) == 16, "ogjk_contact.witness");

/* As ogjk_collide_scene_ws and ogjk_collide_posed_bodies, writing contact
 * records instead of gkSimplex, distance and normal arrays (mode 1 or 3;
 * d_f64 optional).  No counterpart in the reference: its consumer reads the
 * three arrays above. */
ogjk_status ogjk_collide_scene_contacts(int mode, const ogjk_scene* scene, const int32_t* d_pairs,
                                        int32_t n_pairs, ogjk_contact* d_contacts, double* d_f64,
                                        void* d_workspace, size_t workspace_bytes, void* stream);
ogjk_status ogjk_collide_posed_bodies_contacts(int mode, const ogjk_scene* scene,
                                               const gkCollisionPair* d_pairs,
                                               const float* d_body_poses, int32_t n_pairs,
                                               ogjk_contact* d_contacts, double* d_f64,
                                               void* d_workspace, size_t workspace_bytes,
                                               void* stream);

/* Serving loop over a resident scene (no counterpart in the reference, whose
 * calls are synchronous and allocate per call, GJK/gpu/openGJK.cu:2791-2883).
 * Requests of posed pairs (layout of ogjk_collide_posed, host memory) are
 * admitted into up to `slots` in-flight launches of at most max_batch_pairs
 * pairs each (fewer slots if memory_budget bytes of device memory cannot hold
 * them), batched in arrival order as slots free up, and answered into the
 * caller's host buffers (simplices, normals and f64 may be NULL).  Buffers
 * must stay valid until ogjk_server_wait returns for the ticket.  Latency is
 * submit to results in the caller's buffers. */
typedef struct ogjk_server_config {
  int32_t max_batch_pairs; /* pairs per launch and per request (default 65536) */
  int32_t slots;           /* launches in flight (default 4) */
  int64_t memory_budget;   /* device bytes for slots, 0: no limit */
} ogjk_server_config;
typedef struct ogjk_server_s* ogjk_server;
ogjk_status ogjk_server_create(const ogjk_scene* scene, const ogjk_server_config* config,
                               ogjk_server* server);
ogjk_status ogjk_server_submit(ogjk_server server, const int32_t* pairs, const float* poses,
                               int32_t n_pairs, gkSimplex* simplices, float* distances,
                               float* normals, double* f64, int64_t* ticket);
/* Blocks until the request is answered; latency_s and batch (the launch that
 * carried it) may be NULL.  Each ticket is waited for once. */
ogjk_status ogjk_server_wait(ogjk_server server, int64_t ticket, double* latency_s,
                             int64_t* batch);
ogjk_status ogjk_server_poll(ogjk_server server, int64_t ticket, int* done);
ogjk_status ogjk_server_info(ogjk_server server, int32_t* slots, int32_t* capacity,
                             int64_t* launches, int64_t* pairs_done);
/* Shuts the server down and frees it: submissions are refused from then on,
 * requests not yet launched are cancelled (their ogjk_server_wait returns
 * OGJK_ERR_INVALID), launched batches complete and their requests are
 * answered, and the call returns once every thread blocked in
 * ogjk_server_wait has returned.  No call may start on the server after
 * destroy has been called.  Returns OGJK_ERR_HIP if a launch had failed. */
ogjk_status ogjk_server_destroy(ogjk_server server);

/* ------------------------------------------------------------- broad phase */

/* Uniform grid: grid_size^3 cubic cells of edge cell_size from -boundary on
 * each axis; an object's cell is clamp(floor((p + boundary) / cell_size), 0,
 * grid_size - 1) per axis (NaN -> 0).  A cell lists at most max_per_cell
 * objects, the lowest indices (0: no limit; the reference's grid keeps 512,
 * visualization/sim_config.h:16, chosen by arrival order).  Cell lists, the
 * start table and the sort are in the caller's workspace, sized by
 * ogjk_broad_phase_workspace (needs a device: the sort's size query does). */
typedef struct ogjk_grid {
  float cell_size;
  float boundary;
  int32_t grid_size;
  int32_t max_per_cell;
} ogjk_grid;
ogjk_status ogjk_broad_phase_workspace(int32_t n_objects, const ogjk_grid* grid, size_t* bytes);

/* Spheres d_spheres[4i..4i+3] = centre x, y, z and bounding radius of object
 * i.  Writes every pair (i, j), i < j, in the same or a neighbouring cell with
 * ((dx*dx + dy*dy) + dz*dz) < (r_i + r_j)^2 in FP32: ordered by i, then by
 * neighbour cell (z, then y, then x, each from -1 to +1), then by j.  The
 * first max_pairs are written; *n_pairs = how many, *n_found = all found.
 * Synchronous on stream (the total is read back).  Replaces sim_broad_phase's
 * grid pass: insert_objects_kernel, count_pairs_kernel, the prefix scan and
 * generate_pairs_kernel (visualization/integrate_final_gjk.cu:467-570,
 * 916-956). */
ogjk_status ogjk_broad_phase(const float* d_spheres, int32_t n_objects, const ogjk_grid* grid,
                             gkCollisionPair* d_pairs, int32_t max_pairs, int32_t* n_pairs,
                             int64_t* n_found, void* d_workspace, size_t workspace_bytes,
                             void* stream);

/* Body pairs -> sub-mesh pairs: body pair p = (a, b) becomes
 * (d_sub_off[a] + u, d_sub_off[b] + v) for u < d_sub_cnt[a], v < d_sub_cnt[b],
 * u-major, in body-pair order; a pair naming a body outside [0, n_bodies)
 * expands to nothing.  The first max_sub_pairs are written; synchronous.
 * Replaces count_submesh_pairs_kernel, its scan and
 * expand_submesh_pairs_kernel (visualization/integrate_final_gjk.cu:265-302,
 * 958-1000). */
ogjk_status ogjk_expand_pairs_workspace(int32_t n_body_pairs, size_t* bytes);
ogjk_status ogjk_expand_pairs(const gkCollisionPair* d_body_pairs, int32_t n_body_pairs,
                              const int32_t* d_sub_off, const int32_t* d_sub_cnt,
                              int32_t n_bodies, gkCollisionPair* d_sub_pairs,
                              int32_t max_sub_pairs, int32_t* n_sub_pairs, int64_t* n_found,
                              void* d_workspace, size_t workspace_bytes, void* stream);

/* ---------------------------------------------------------- mesh ingestion */

/* Convex hulls of n_clouds point clouds at once (cloud m: d_point_cnt[m]
 * points at d_points + 3 * d_point_off[m], AoS float32).  Faces (vertex-index
 * triples, local to the cloud) go to d_faces + 3 * d_face_off[m], at most
 * d_face_off[m + 1] - d_face_off[m] of them; d_face_off has n_clouds + 1
 * entries and d_face_off[n_clouds] == total_faces.  d_nfaces[m] = face count,
 * 0 when there is no hull (fewer than 4 points, collinear or coplanar), -1
 * when the capacity was too small, -2 for a non-finite coordinate, -3 when
 * the cloud's face range is not inside [0, total_faces].  The face
 * list is the reference's, in its order, including its behaviour on dense
 * near-coplanar input (more than 2n - 4 faces can result: give about 4n).
 * Workspace: 12 * total_faces bytes.  Asynchronous.  Replaces build_hull
 * (visualization/rendering/mesh_builder.cpp:279-360), run once per cloud on
 * the host there. */
ogjk_status ogjk_build_hulls(const float* d_points, const int64_t* d_point_off,
                             const int32_t* d_point_cnt, int32_t n_clouds,
                             const int64_t* d_face_off, int64_t total_faces, int32_t* d_faces,
                             int32_t* d_nfaces, void* d_workspace, size_t workspace_bytes,
                             void* stream);

/* Per cloud, from its hull: d_moments[7m..7m+6] = volume, first moments
 * (x, y, z, times 4 volume), second moments (x^2, y^2, z^2) in FP64, summed in
 * face order; d_com[3m..] = the centre of mass center_hull_verts shifts by (0
 * below volume 1e-10); d_inertia_k[3m..] = compute_hull_inertia_k's factors.
 * Replaces hull_volume_moments, center_hull_verts' shift and
 * compute_hull_inertia_k (visualization/rendering/mesh_builder.cpp:433-481). */
ogjk_status ogjk_hull_moments(const float* d_points, const int64_t* d_point_off,
                              const int32_t* d_faces, const int64_t* d_face_off,
                              const int32_t* d_nfaces, int32_t n_clouds, double* d_moments,
                              float* d_com, float* d_inertia_k, void* stream);

/* d_points[i] -= d_com[d_point_cloud[i]] for n_points points (FP32), the
 * reference's centring of a cloud (mesh_builder.cpp:464-469). */
ogjk_status ogjk_center_points(float* d_points, const int32_t* d_point_cloud, int64_t n_points,
                               const float* d_com, void* stream);

/* Slots a cluster build needs for a scene of n_verts vertices in n_bodies
 * bodies (padding to 32 per body included). */
int64_t ogjk_cluster_slots(int64_t n_verts, int32_t n_bodies);
/* Superclusters (runs of 32 clusters) a build over n_slots slots can need. */
int64_t ogjk_cluster_supers(int64_t n_slots);

/* Builds the clusters of every body of a packed scene (scene->x/y/z,
 * body_off, body_cnt) with at least clusters->min_verts vertices into
 * caller-allocated arrays.  Synchronous on stream;
 * allocates scratch for the per-body sort.  No counterpart in the reference:
 * an acceleration structure for the support function
 * (GJK/cpu/openGJK.c:615-639, GJK/cpu/EPA.c:307-344) on large hulls. */
ogjk_status ogjk_build_clusters(const ogjk_scene* scene, ogjk_clusters* clusters, void* stream);

/* Plain-FP64 comparison path: double storage, FP64 sweep, no candidate
 * re-rank; same selection rules and results (the denominator of the
 * mixed-precision speed-up). */
ogjk_status ogjk_pack_scene_f64(const float* d_coords_aos, int64_t n_verts, double* d_x,
                                double* d_y, double* d_z, void* stream);
/* The same pool as an FP64 AoS array of 3 * n_verts doubles (ogjk_scene64.xyz). */
ogjk_status ogjk_pack_scene_f64_aos(const float* d_coords_aos, int64_t n_verts, double* d_xyz,
                                    void* stream);
ogjk_status ogjk_collide_scene_f64(int mode, const ogjk_scene64* scene, const int32_t* d_pairs,
                                   int32_t n_pairs, gkSimplex* d_simplices, float* d_distances,
                                   float* d_normals, double* d_f64, void* stream);

/* ------------------------------------------------------------ gkPolytope input */

/* Device gkPolytope arrays whose coord are device AoS pointers; pair i is
 * (d_bd1[i], d_bd2[i]).  Replaces compute_minimum_distance_device and
 * compute_epa_device (GJK/gpu/openGJK.h:237-262).  Asynchronous. */
ogjk_status ogjk_collide_polytopes(int mode, const gkPolytope* d_bd1, const gkPolytope* d_bd2,
                                   int32_t n, gkSimplex* d_simplices, float* d_distances,
                                   float* d_normals, void* stream);

/* Indexed form over a device polytope pool (GJK/gpu/openGJK.h:398-505). */
ogjk_status ogjk_collide_polytopes_indexed(int mode, const gkPolytope* d_pool,
                                           int32_t n_polytopes, const gkCollisionPair* d_pairs,
                                           int32_t n_pairs, gkSimplex* d_simplices,
                                           float* d_distances, float* d_normals, void* stream);

/* Host arrays in, host arrays out, synchronous; device memory is staged
 * internally.  Replaces compute_minimum_distance / computeCollisionInformation
 * / compute_gjk_epa (GJK/gpu/openGJK.h:91-141).  normals may be NULL for GJK.
 * The staging buffers (device and pinned host, up to 256 MiB each per device)
 * are kept between calls until the process ends; larger calls allocate and
 * free their own.  Calls on one device are serialised. */
ogjk_status ogjk_collide_host(int mode, int32_t n, const gkPolytope* bd1, const gkPolytope* bd2,
                              gkSimplex* simplices, float* distances, float* normals);

/* Host indexed form: compute_minimum_distance_indexed / compute_epa_indexed /
 * compute_gjk_epa_indexed (GJK/gpu/openGJK.h:398-505). */
ogjk_status ogjk_collide_host_indexed(int mode, int32_t n_polytopes, int32_t n_pairs,
                                      const gkPolytope* polytopes, const gkCollisionPair* pairs,
                                      gkSimplex* simplices, float* distances, float* normals);

#ifdef __cplusplus
}
#endif

#endif /* OPENGJK_AMD_OGJK_H */
