// opengjk-amd C ABI: argument checking, host staging and launches.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <vector>

#include "ogjk_internal.hpp"
#include "opengjk/ogjk.h"

namespace {

thread_local char g_err[512] = "";

ogjk_status fail(ogjk_status st, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return st;
}

}  // namespace

// Records the thread's error message; shared with the serving runtime.
ogjk_status ogjk_fail(ogjk_status st, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return st;
}

namespace {

ogjk_status hip_fail(hipError_t e, const char* what) {
  return fail(e == hipErrorOutOfMemory ? OGJK_ERR_NOMEM : OGJK_ERR_HIP, "%s: %s", what,
              hipGetErrorString(e));
}

#define OGJK_HIP(call, what)                     \
  do {                                           \
    const hipError_t e_ = (call);                \
    if (e_ != hipSuccess) return hip_fail(e_, what); \
  } while (0)

bool valid_mode(int m) { return m == OGJK_MODE_GJK || m == OGJK_MODE_EPA || m == OGJK_MODE_GJK_EPA; }

// Library-owned GJK+EPA workspace, one per device, grown on demand.  Calls
// that use it must not overlap on one device (pass an explicit workspace to
// ogjk_collide_scene_ws for concurrent streams).
struct CachedWorkspace {
  std::mutex mu;
  void* ptr[64] = {};
  size_t bytes[64] = {};
};
CachedWorkspace g_ws;

ogjk_status cached_workspace(size_t need, void** out) {
  int dev = 0;
  OGJK_HIP(hipGetDevice(&dev), "hipGetDevice");
  if (dev < 0 || dev >= 64) return fail(OGJK_ERR_UNSUPPORTED, "device id %d", dev);
  std::lock_guard<std::mutex> lock(g_ws.mu);
  if (g_ws.bytes[dev] < need) {
    if (g_ws.ptr[dev]) OGJK_HIP(hipFree(g_ws.ptr[dev]), "hipFree(workspace)");
    g_ws.ptr[dev] = nullptr;
    g_ws.bytes[dev] = 0;
    const size_t grow = need + need / 4;
    OGJK_HIP(hipMalloc(&g_ws.ptr[dev], grow), "hipMalloc(workspace)");
    g_ws.bytes[dev] = grow;
  }
  *out = g_ws.ptr[dev];
  return OGJK_OK;
}

// One device allocation, freed on scope exit.
struct DeviceArena {
  char* base = nullptr;
  ~DeviceArena() {
    if (base) (void)hipFree(base);
  }
  static size_t up(size_t b) { return (b + 255) & ~size_t(255); }
  hipError_t reserve(size_t bytes) { return bytes ? hipMalloc((void**)&base, bytes) : hipSuccess; }
};

// Host staging for the host entry points.  Inputs are laid out in one
// buffer exactly as they will sit on the device, so a call makes one upload
// and one download:
//   x | y | z | off | cnt | r1 | pairs | simplices | distances | normals | workspace
// (each piece 256-byte aligned).  Upload: everything before the simplices, or
// through the distances for EPA, which reads them.  Download: simplices
// through distances, or through normals when they are asked for.  The
// workspace lives on the device only.
struct StageLayout {
  size_t x, y, z, off, cnt, r1, pairs, simp, dist, nrm, ws, host_end, total;
  StageLayout(size_t nv, size_t nb, size_t n, bool indexed) {
    size_t o = 0;
    auto put = [&](size_t bytes) {
      const size_t at = o;
      o += DeviceArena::up(bytes);
      return at;
    };
    x = put(nv * 4);
    y = put(nv * 4);
    z = put(nv * 4);
    off = put(nb * 8);
    cnt = put(nb * 4);
    r1 = put(nb * 4);
    pairs = put(indexed ? n * 8 : 0);
    simp = put(n * sizeof(gkSimplex));
    dist = put(n * 4);
    nrm = put(n * 12);
    host_end = o;
    ws = put(ogjk::workspace_bytes((int)n));
    total = o;
  }
};

// Per-device staging kept between calls: a device buffer and a pinned host
// buffer, grown on demand up to kStageCacheMax bytes each.  A call that needs
// more allocates its own device buffer and a pageable host buffer and frees
// them on return.  The mutex serialises the host entry points on a device.
constexpr size_t kStageCacheMax = size_t(256) << 20;
struct StageCache {
  std::mutex mu;
  char* dev = nullptr;
  char* host = nullptr;
  size_t dev_bytes = 0, host_bytes = 0;
};
StageCache g_stage[64];

ogjk_status stage_grow(StageCache& c, const StageLayout& L) {
  const auto sized = [](size_t need) {
    return std::min(std::max(need + need / 4, size_t(1) << 20), kStageCacheMax);
  };
  if (c.dev_bytes < L.total) {
    if (c.dev) OGJK_HIP(hipFree(c.dev), "hipFree(staging)");
    c.dev = nullptr;
    c.dev_bytes = 0;
    const size_t b = sized(L.total);
    OGJK_HIP(hipMalloc((void**)&c.dev, b), "hipMalloc(staging)");
    c.dev_bytes = b;
  }
  if (c.host_bytes < L.host_end) {
    if (c.host) OGJK_HIP(hipHostFree(c.host), "hipHostFree(staging)");
    c.host = nullptr;
    c.host_bytes = 0;
    const size_t b = sized(L.host_end);
    OGJK_HIP(hipHostMalloc((void**)&c.host, b, hipHostMallocDefault), "hipHostMalloc(staging)");
    c.host_bytes = b;
  }
  return OGJK_OK;
}

// Stage `lists` (interleave: body 2i = lists[0][i], 2i+1 = lists[1][i], the
// implicit pairs; otherwise one pool addressed by `pairs`), run `mode` over
// `n` pairs and copy the results back.  Synchronous.
ogjk_status run_host(int mode, const gkPolytope* const* lists, const int* counts, int nlists,
                     bool interleave, const gkCollisionPair* pairs, int n, gkSimplex* simplices,
                     float* distances, float* normals) {
  int64_t nb = 0, nv = 0;
  for (int l = 0; l < nlists; l++) nb += counts[l];
  auto body_at = [&](int64_t b) -> const gkPolytope& {
    if (interleave) return lists[b & 1][b >> 1];
    return lists[0][b];
  };
  for (int64_t b = 0; b < nb; b++) {
    const gkPolytope& p = body_at(b);
    if (p.numpoints < 0) return fail(OGJK_ERR_INVALID, "body %lld: numpoints %d < 0", (long long)b, p.numpoints);
    if (p.numpoints > 0 && p.coord == nullptr)
      return fail(OGJK_ERR_INVALID, "body %lld: coord is NULL", (long long)b);
    nv += p.numpoints;
  }
  const StageLayout L((size_t)nv, (size_t)nb, (size_t)n, pairs != nullptr);
  int dev = 0;
  OGJK_HIP(hipGetDevice(&dev), "hipGetDevice");
  if (dev < 0 || dev >= 64) return fail(OGJK_ERR_UNSUPPORTED, "device id %d", dev);
  StageCache& c = g_stage[dev];
  std::lock_guard<std::mutex> lock(c.mu);
  char *hb, *db;
  std::vector<char> hvec;
  DeviceArena big;
  if (L.total <= kStageCacheMax) {
    const ogjk_status s = stage_grow(c, L);
    if (s != OGJK_OK) return s;
    hb = c.host;
    db = c.dev;
  } else {
    hvec.resize(L.host_end);
    OGJK_HIP(big.reserve(L.total), "hipMalloc(staging)");
    hb = hvec.data();
    db = big.base;
  }

  float* hx = (float*)(hb + L.x);
  float* hy = (float*)(hb + L.y);
  float* hz = (float*)(hb + L.z);
  int64_t* hoff = (int64_t*)(hb + L.off);
  int32_t* hcnt = (int32_t*)(hb + L.cnt);
  float* hr1 = (float*)(hb + L.r1);
  int64_t o = 0;
  for (int64_t b = 0; b < nb; b++) {
    const gkPolytope& p = body_at(b);
    float m = 0.0f;
    for (int i = 0; i < p.numpoints; i++) {
      const float vx = p.coord[3 * i], vy = p.coord[3 * i + 1], vz = p.coord[3 * i + 2];
      hx[o + i] = vx;
      hy[o + i] = vy;
      hz[o + i] = vz;
      m = std::fmax(m, std::fabs(vx) + std::fabs(vy) + std::fabs(vz));
    }
    hoff[b] = o;
    hcnt[b] = p.numpoints;
    hr1[b] = m * (1.0f + 0x1.0p-22f);  // per-body L1 bound, rounded up
    o += p.numpoints;
  }
  if (pairs) std::memcpy(hb + L.pairs, pairs, (size_t)n * 8);
  size_t up = L.simp;
  if (mode == OGJK_MODE_EPA) {
    std::memcpy(hb + L.simp, simplices, (size_t)n * sizeof(gkSimplex));
    std::memcpy(hb + L.dist, distances, (size_t)n * 4);
    up = L.dist + (size_t)n * 4;
  }
  const bool want_n = (mode & OGJK_MODE_EPA) && normals;
  const size_t down = (want_n ? L.nrm + (size_t)n * 12 : L.dist + (size_t)n * 4) - L.simp;

  OGJK_HIP(hipMemcpyAsync(db, hb, up, hipMemcpyHostToDevice, nullptr), "hipMemcpy(upload)");
  const ogjk::SceneArgs sa{(float*)(db + L.x), (float*)(db + L.y), (float*)(db + L.z),
                           (int64_t*)(db + L.off), (int32_t*)(db + L.cnt), (float*)(db + L.r1),
                           (int32_t)nb, pairs ? (int32_t*)(db + L.pairs) : nullptr};
  const ogjk::OutArgs oa{(gkSimplex*)(db + L.simp), (float*)(db + L.dist), (float*)(db + L.nrm),
                         nullptr};
  OGJK_HIP(ogjk::launch_scene(mode, sa, oa, n, db + L.ws, nullptr), "kernel launch");
  OGJK_HIP(hipMemcpyAsync(hb + L.simp, db + L.simp, down, hipMemcpyDeviceToHost, nullptr),
           "hipMemcpy(download)");
  OGJK_HIP(hipStreamSynchronize(nullptr), "kernel");
  std::memcpy(simplices, hb + L.simp, (size_t)n * sizeof(gkSimplex));
  std::memcpy(distances, hb + L.dist, (size_t)n * 4);
  if (want_n) std::memcpy(normals, hb + L.nrm, (size_t)n * 12);
  return OGJK_OK;
}

}  // namespace

// ------------------------------------------------- allocation-free variant
//
// The GJK -> EPA workspace of ogjk_collide_scene_noalloc: a static device
// array sized for OGJK_NOALLOC_MAX_PAIRS pairs, one copy per device, held
// until the process ends.  Host code never names `scratch` in a kernel
// argument or copy (on the host it is only a stand-in for the symbol): its
// device address comes from hipGetSymbolAddress, once per device, on that
// device's first call.
constexpr size_t kNoallocBytes = ogjk::workspace_bytes_for(OGJK_NOALLOC_MAX_PAIRS);
static __device__ __align__(256) unsigned char scratch[kNoallocBytes];

namespace {
struct NoallocScratch {
  std::mutex mu;
  void* ptr[64] = {};
};
NoallocScratch g_noalloc;

ogjk_status noalloc_scratch(void** out) {
  int dev = 0;
  OGJK_HIP(hipGetDevice(&dev), "hipGetDevice");
  if (dev < 0 || dev >= 64) return fail(OGJK_ERR_UNSUPPORTED, "device id %d", dev);
  std::lock_guard<std::mutex> lock(g_noalloc.mu);
  if (!g_noalloc.ptr[dev]) {
    void* p = nullptr;
    OGJK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    if (!p) return fail(OGJK_ERR_HIP, "hipGetSymbolAddress(scratch) returned NULL");
    g_noalloc.ptr[dev] = p;
  }
  *out = g_noalloc.ptr[dev];
  return OGJK_OK;
}
}  // namespace

extern "C" {

int ogjk_version(void) { return 100; }

const char* ogjk_last_error(void) { return g_err; }

int ogjk_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

ogjk_status ogjk_set_config(int gjk_lanes, int epa_lanes, int fused) {
  ogjk::LaunchConfig c;
  c.nb = 1;
  c.gjk_lanes[0] = gjk_lanes;
  c.epa_lanes[0] = epa_lanes;
  c.fused = fused;
  if (ogjk::set_config(c) != 0)
    return fail(OGJK_ERR_INVALID,
                "lanes per pair: GJK 8/16/32/64 (1 with fused 2), EPA 16/32/64; fused 0, 1 or 2 "
                "(got %d, %d, %d)", gjk_lanes, epa_lanes, fused);
  return OGJK_OK;
}

ogjk_status ogjk_set_buckets(int n_buckets, const int* max_verts, const int* gjk_lanes,
                             const int* epa_lanes) {
  if (n_buckets < 1 || n_buckets > ogjk::kMaxBuckets || !max_verts || !gjk_lanes || !epa_lanes)
    return fail(OGJK_ERR_INVALID, "1..%d buckets with non-NULL tables", ogjk::kMaxBuckets);
  ogjk::LaunchConfig c;
  c.nb = n_buckets;
  for (int b = 0; b < n_buckets; b++) {
    c.max_verts[b] = b < n_buckets - 1 ? max_verts[b] : 0x7fffffff;
    c.gjk_lanes[b] = gjk_lanes[b];
    c.epa_lanes[b] = epa_lanes[b];
  }
  if (ogjk::set_config(c) != 0)
    return fail(OGJK_ERR_INVALID, "bucket bounds must increase; lanes GJK 8/16/32/64, EPA 16/32/64");
  return OGJK_OK;
}

ogjk_status ogjk_set_block_waves(int n_buckets, const int* gjk_waves, const int* epa_waves) {
  ogjk::LaunchConfig c = ogjk::get_config();
  if (n_buckets != c.nb || !gjk_waves || !epa_waves)
    return fail(OGJK_ERR_INVALID, "%d bucket(s) configured, got %d", c.nb, n_buckets);
  for (int b = 0; b < n_buckets; b++) {
    c.gjk_waves[b] = gjk_waves[b];
    c.epa_waves[b] = epa_waves[b];
  }
  if (ogjk::set_config(c) != 0)
    return fail(OGJK_ERR_INVALID, "waves per block must be 1, 2 or 4");
  return OGJK_OK;
}

size_t ogjk_workspace_bytes(int32_t n_pairs) { return ogjk::workspace_bytes(n_pairs); }

ogjk_status ogjk_sync(void* stream) {
  OGJK_HIP(hipStreamSynchronize((hipStream_t)stream), "hipStreamSynchronize");
  return OGJK_OK;
}

ogjk_status ogjk_pack_scene(const float* d_coords_aos, const int64_t* d_body_off,
                            const int32_t* d_body_cnt, int32_t n_bodies, int64_t n_verts,
                            float* d_x, float* d_y, float* d_z, float* d_r1, void* stream) {
  if (n_bodies < 0 || n_verts < 0) return fail(OGJK_ERR_INVALID, "negative size");
  if (n_verts > 0 && (!d_coords_aos || !d_x || !d_y || !d_z))
    return fail(OGJK_ERR_INVALID, "NULL coordinate pointer");
  if (n_bodies > 0 && (!d_body_off || !d_body_cnt || !d_r1))
    return fail(OGJK_ERR_INVALID, "NULL body table pointer");
  OGJK_HIP(ogjk::launch_pack(d_coords_aos, d_body_off, d_body_cnt, n_bodies, n_verts, d_x, d_y,
                             d_z, d_r1, (hipStream_t)stream),
           "pack launch");
  return OGJK_OK;
}

ogjk_status ogjk_pack_scene_f64_aos(const float* d_coords_aos, int64_t n_verts, double* d_xyz,
                                    void* stream) {
  if (n_verts < 0) return fail(OGJK_ERR_INVALID, "negative size");
  if (n_verts > 0 && (!d_coords_aos || !d_xyz)) return fail(OGJK_ERR_INVALID, "NULL pointer");
  OGJK_HIP(ogjk::launch_pack_f64_aos(d_coords_aos, n_verts, d_xyz, (hipStream_t)stream),
           "pack launch");
  return OGJK_OK;
}

ogjk_status ogjk_pack_scene_f64(const float* d_coords_aos, int64_t n_verts, double* d_x,
                                double* d_y, double* d_z, void* stream) {
  if (n_verts < 0) return fail(OGJK_ERR_INVALID, "negative size");
  if (n_verts > 0 && (!d_coords_aos || !d_x || !d_y || !d_z))
    return fail(OGJK_ERR_INVALID, "NULL coordinate pointer");
  OGJK_HIP(ogjk::launch_pack_f64(d_coords_aos, n_verts, d_x, d_y, d_z, (hipStream_t)stream),
           "pack launch");
  return OGJK_OK;
}

// Validates a float scene and its optional clusters into launch arguments.
static ogjk_status scene_args(const ogjk_scene* scene, const int32_t* d_pairs, int32_t n_pairs,
                              ogjk::SceneArgs* sa) {
  if (!scene || !scene->x || !scene->y || !scene->z || !scene->body_off || !scene->body_cnt ||
      !scene->body_r1)
    return fail(OGJK_ERR_INVALID, "incomplete scene");
  if (!d_pairs && (int64_t)scene->n_bodies < 2 * (int64_t)n_pairs)
    return fail(OGJK_ERR_INVALID, "implicit pairs need %lld bodies, scene has %d",
                2 * (long long)n_pairs, scene->n_bodies);
  *sa = ogjk::SceneArgs{scene->x, scene->y, scene->z, scene->body_off, scene->body_cnt,
                        scene->body_r1, scene->n_bodies, d_pairs};
  if (const ogjk_clusters* cl = scene->clusters) {
    if (!cl->x || !cl->y || !cl->z || !cl->orig || !cl->body_slot || !cl->bounds ||
        !cl->super_bounds || !cl->body_super)
      return fail(OGJK_ERR_INVALID, "incomplete clusters");
    sa->csbounds = cl->n_supers_used > 0 ? cl->super_bounds : nullptr;
    sa->cx = cl->x;
    sa->cy = cl->y;
    sa->cz = cl->z;
    sa->corig = cl->orig;
    sa->cslot = cl->body_slot;
    sa->cbounds = cl->bounds;
    sa->cluster_min = cl->min_verts > 1 ? cl->min_verts : 1;
  }
  return OGJK_OK;
}

static ogjk_status check_out(int mode, int32_t n, gkSimplex* s, float* d, float* nrm) {
  if (!valid_mode(mode)) return fail(OGJK_ERR_INVALID, "mode %d is not 1, 2 or 3", mode);
  if (n < 0) return fail(OGJK_ERR_INVALID, "n = %d < 0", n);
  if (n > 0 && (!s || !d)) return fail(OGJK_ERR_INVALID, "NULL simplices/distances");
  (void)nrm;
  return OGJK_OK;
}

ogjk_status ogjk_collide_scene_ws(int mode, const ogjk_scene* scene, const int32_t* d_pairs,
                                  int32_t n_pairs, gkSimplex* d_simplices, float* d_distances,
                                  float* d_normals, double* d_f64, void* d_workspace,
                                  size_t workspace_bytes, void* stream) {
  const ogjk_status st = check_out(mode, n_pairs, d_simplices, d_distances, d_normals);
  if (st != OGJK_OK) return st;
  if (n_pairs == 0) return OGJK_OK;
  if (mode != OGJK_MODE_EPA && (!d_workspace || workspace_bytes < ogjk::workspace_bytes(n_pairs)))
    return fail(OGJK_ERR_INVALID, "workspace of %zu bytes needed", ogjk::workspace_bytes(n_pairs));
  ogjk::SceneArgs sa;
  const ogjk_status s2 = scene_args(scene, d_pairs, n_pairs, &sa);
  if (s2 != OGJK_OK) return s2;
  const ogjk::OutArgs oa{d_simplices, d_distances, d_normals, d_f64};
  OGJK_HIP(ogjk::launch_scene(mode, sa, oa, n_pairs, d_workspace, (hipStream_t)stream),
           "kernel launch");
  return OGJK_OK;
}

static ogjk_status check_contacts(int mode, int32_t n, ogjk_contact* contacts, void* ws,
                                  size_t ws_bytes) {
  if (mode != OGJK_MODE_GJK && mode != OGJK_MODE_GJK_EPA)
    return fail(OGJK_ERR_INVALID, "contact records need mode 1 or 3, got %d", mode);
  if (n < 0) return fail(OGJK_ERR_INVALID, "n = %d < 0", n);
  if (n > 0 && !contacts) return fail(OGJK_ERR_INVALID, "NULL contacts");
  if (n > 0 && (!ws || ws_bytes < ogjk::workspace_bytes(n)))
    return fail(OGJK_ERR_INVALID, "workspace of %zu bytes needed", ogjk::workspace_bytes(n));
  return OGJK_OK;
}

ogjk_status ogjk_collide_scene_contacts(int mode, const ogjk_scene* scene, const int32_t* d_pairs,
                                        int32_t n_pairs, ogjk_contact* d_contacts, double* d_f64,
                                        void* d_workspace, size_t workspace_bytes, void* stream) {
  const ogjk_status st = check_contacts(mode, n_pairs, d_contacts, d_workspace, workspace_bytes);
  if (st != OGJK_OK || n_pairs == 0) return st;
  ogjk::SceneArgs sa;
  const ogjk_status s2 = scene_args(scene, d_pairs, n_pairs, &sa);
  if (s2 != OGJK_OK) return s2;
  ogjk::OutArgs oa{nullptr, nullptr, nullptr, d_f64};
  oa.contacts = (float*)d_contacts;
  OGJK_HIP(ogjk::launch_scene(mode, sa, oa, n_pairs, d_workspace, (hipStream_t)stream),
           "kernel launch");
  return OGJK_OK;
}

ogjk_status ogjk_collide_posed_bodies_contacts(int mode, const ogjk_scene* scene,
                                               const gkCollisionPair* d_pairs,
                                               const float* d_body_poses, int32_t n_pairs,
                                               ogjk_contact* d_contacts, double* d_f64,
                                               void* d_workspace, size_t workspace_bytes,
                                               void* stream) {
  const ogjk_status st = check_contacts(mode, n_pairs, d_contacts, d_workspace, workspace_bytes);
  if (st != OGJK_OK || n_pairs == 0) return st;
  if (!d_pairs || !d_body_poses) return fail(OGJK_ERR_INVALID, "posed pairs need pairs and poses");
  ogjk::SceneArgs sa;
  const ogjk_status s2 = scene_args(scene, (const int32_t*)d_pairs, n_pairs, &sa);
  if (s2 != OGJK_OK) return s2;
  ogjk::OutArgs oa{nullptr, nullptr, nullptr, d_f64};
  oa.contacts = (float*)d_contacts;
  OGJK_HIP(ogjk::launch_posed(mode, sa, d_body_poses, true, oa, n_pairs, d_workspace,
                              (hipStream_t)stream),
           "kernel launch");
  return OGJK_OK;
}

ogjk_status ogjk_collide_posed(int mode, const ogjk_scene* scene, const int32_t* d_pairs,
                               const float* d_poses, int32_t n_pairs, gkSimplex* d_simplices,
                               float* d_distances, float* d_normals, double* d_f64,
                               void* d_workspace, size_t workspace_bytes, void* stream) {
  const ogjk_status st = check_out(mode, n_pairs, d_simplices, d_distances, d_normals);
  if (st != OGJK_OK) return st;
  if (n_pairs == 0) return OGJK_OK;
  if (!d_pairs || !d_poses) return fail(OGJK_ERR_INVALID, "posed pairs need pairs and poses");
  if (mode != OGJK_MODE_EPA && (!d_workspace || workspace_bytes < ogjk::workspace_bytes(n_pairs)))
    return fail(OGJK_ERR_INVALID, "workspace of %zu bytes needed", ogjk::workspace_bytes(n_pairs));
  ogjk::SceneArgs sa;
  const ogjk_status s2 = scene_args(scene, d_pairs, n_pairs, &sa);
  if (s2 != OGJK_OK) return s2;
  const ogjk::OutArgs oa{d_simplices, d_distances, d_normals, d_f64};
  OGJK_HIP(ogjk::launch_posed(mode, sa, d_poses, false, oa, n_pairs, d_workspace,
                              (hipStream_t)stream),
           "kernel launch");
  return OGJK_OK;
}

ogjk_status ogjk_collide_posed_bodies(int mode, const ogjk_scene* scene,
                                      const gkCollisionPair* d_pairs, const float* d_body_poses,
                                      int32_t n_pairs, gkSimplex* d_simplices,
                                      float* d_distances, float* d_normals, double* d_f64,
                                      void* d_workspace, size_t workspace_bytes, void* stream) {
  const ogjk_status st = check_out(mode, n_pairs, d_simplices, d_distances, d_normals);
  if (st != OGJK_OK) return st;
  if (n_pairs == 0) return OGJK_OK;
  if (!d_pairs || !d_body_poses) return fail(OGJK_ERR_INVALID, "posed pairs need pairs and poses");
  if (mode != OGJK_MODE_EPA && (!d_workspace || workspace_bytes < ogjk::workspace_bytes(n_pairs)))
    return fail(OGJK_ERR_INVALID, "workspace of %zu bytes needed", ogjk::workspace_bytes(n_pairs));
  ogjk::SceneArgs sa;
  const ogjk_status s2 = scene_args(scene, (const int32_t*)d_pairs, n_pairs, &sa);
  if (s2 != OGJK_OK) return s2;
  const ogjk::OutArgs oa{d_simplices, d_distances, d_normals, d_f64};
  OGJK_HIP(ogjk::launch_posed(mode, sa, d_body_poses, true, oa, n_pairs, d_workspace,
                              (hipStream_t)stream),
           "kernel launch");
  return OGJK_OK;
}

// ------------------------------------------------------------ HIP graphs

struct ogjk_graph_s {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  void* workspace = nullptr;
  ~ogjk_graph_s() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (workspace) (void)hipFree(workspace);
  }
};

ogjk_status ogjk_graph_capture_scene(int mode, const ogjk_scene* scene, const int32_t* d_pairs,
                                     int32_t n_pairs, gkSimplex* d_simplices, float* d_distances,
                                     float* d_normals, double* d_f64, ogjk_graph* out) {
  if (!out) return fail(OGJK_ERR_INVALID, "NULL graph handle");
  *out = nullptr;
  const ogjk_status st = check_out(mode, n_pairs, d_simplices, d_distances, d_normals);
  if (st != OGJK_OK) return st;
  ogjk::SceneArgs sa;
  const ogjk_status s2 = scene_args(scene, d_pairs, n_pairs, &sa);
  if (s2 != OGJK_OK) return s2;
  auto g = std::make_unique<ogjk_graph_s>();
  const size_t ws_bytes = ogjk::workspace_bytes(n_pairs);
  OGJK_HIP(hipMalloc(&g->workspace, ws_bytes), "hipMalloc(graph workspace)");
  hipStream_t cs = nullptr;
  OGJK_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking), "hipStreamCreate");
  hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
  if (e == hipSuccess) {
    const ogjk::OutArgs oa{d_simplices, d_distances, d_normals, d_f64};
    const hipError_t le = ogjk::launch_scene(mode, sa, oa, n_pairs, g->workspace, cs);
    e = hipStreamEndCapture(cs, &g->graph);
    if (le != hipSuccess) e = le;
  }
  if (e == hipSuccess) e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
  (void)hipStreamDestroy(cs);
  OGJK_HIP(e, "graph capture");
  *out = g.release();
  return OGJK_OK;
}

ogjk_status ogjk_graph_launch(ogjk_graph g, void* stream) {
  if (!g || !g->exec) return fail(OGJK_ERR_INVALID, "NULL graph");
  OGJK_HIP(hipGraphLaunch(g->exec, (hipStream_t)stream), "hipGraphLaunch");
  return OGJK_OK;
}

ogjk_status ogjk_graph_destroy(ogjk_graph g) {
  delete g;
  return OGJK_OK;
}

int64_t ogjk_cluster_slots(int64_t n_verts, int32_t n_bodies) {
  return ogjk::cluster_slots(n_verts, n_bodies);
}

int64_t ogjk_cluster_supers(int64_t n_slots) { return ogjk::cluster_supers(n_slots); }

ogjk_status ogjk_build_clusters(const ogjk_scene* scene, ogjk_clusters* cl, void* stream) {
  if (!scene || !cl) return fail(OGJK_ERR_INVALID, "NULL scene or clusters");
  if (scene->n_bodies < 0) return fail(OGJK_ERR_INVALID, "negative body count");
  if (scene->n_bodies == 0) return OGJK_OK;
  if (!scene->x || !scene->y || !scene->z || !scene->body_off || !scene->body_cnt)
    return fail(OGJK_ERR_INVALID, "incomplete scene");
  if (!cl->x || !cl->y || !cl->z || !cl->orig || !cl->body_slot || !cl->bounds ||
      !cl->super_bounds || !cl->body_super)
    return fail(OGJK_ERR_INVALID, "incomplete clusters");
  const hipError_t e = ogjk::build_clusters(
      scene->x, scene->y, scene->z, scene->body_off, scene->body_cnt, scene->n_bodies, cl->x,
      cl->y, cl->z, cl->orig, cl->body_slot, cl->bounds, cl->n_slots, cl->super_bounds,
      cl->body_super, cl->n_supers, &cl->n_supers_used, cl->min_verts, (hipStream_t)stream);
  if (e == hipErrorInvalidValue)
    return fail(OGJK_ERR_INVALID, "clusters need more than %lld slots or %lld superclusters, or "
                "a body exceeds 2^27 vertices", (long long)cl->n_slots, (long long)cl->n_supers);
  if (e == hipErrorOutOfMemory) return fail(OGJK_ERR_NOMEM, "cluster build scratch");
  OGJK_HIP(e, "cluster build");
  return OGJK_OK;
}

ogjk_status ogjk_collide_scene(int mode, const ogjk_scene* scene, const int32_t* d_pairs,
                               int32_t n_pairs, gkSimplex* d_simplices, float* d_distances,
                               float* d_normals, double* d_f64, void* stream) {
  void* ws = nullptr;
  const size_t need = ogjk::workspace_bytes(n_pairs);
  if (mode != OGJK_MODE_EPA && n_pairs > 0) {
    const ogjk_status s2 = cached_workspace(need, &ws);
    if (s2 != OGJK_OK) return s2;
  }
  return ogjk_collide_scene_ws(mode, scene, d_pairs, n_pairs, d_simplices, d_distances, d_normals,
                               d_f64, ws, need, stream);
}

ogjk_status ogjk_collide_scene_noalloc(int mode, const ogjk_scene* scene, const int32_t* d_pairs,
                                       int32_t n_pairs, gkSimplex* d_simplices,
                                       float* d_distances, float* d_normals, double* d_f64,
                                       void* stream) {
  if (n_pairs > OGJK_NOALLOC_MAX_PAIRS)
    return fail(OGJK_ERR_INVALID, "ogjk_collide_scene_noalloc: %d pairs exceed the maximum of %d",
                n_pairs, (int)OGJK_NOALLOC_MAX_PAIRS);
  const ogjk_status st = check_out(mode, n_pairs, d_simplices, d_distances, d_normals);
  if (st != OGJK_OK || n_pairs == 0) return st;
  void* ws = nullptr;
  const ogjk_status s2 = noalloc_scratch(&ws);
  if (s2 != OGJK_OK) return s2;
  return ogjk_collide_scene_ws(mode, scene, d_pairs, n_pairs, d_simplices, d_distances, d_normals,
                               d_f64, ws, kNoallocBytes, stream);
}

ogjk_status ogjk_collide_scene_f64(int mode, const ogjk_scene64* scene, const int32_t* d_pairs,
                                   int32_t n_pairs, gkSimplex* d_simplices, float* d_distances,
                                   float* d_normals, double* d_f64, void* stream) {
  const ogjk_status st = check_out(mode, n_pairs, d_simplices, d_distances, d_normals);
  if (st != OGJK_OK) return st;
  if (n_pairs == 0) return OGJK_OK;
  if (!scene || (!scene->xyz && (!scene->x || !scene->y || !scene->z)) || !scene->body_off ||
      !scene->body_cnt)
    return fail(OGJK_ERR_INVALID, "incomplete scene");
  if (!d_pairs && (int64_t)scene->n_bodies < 2 * (int64_t)n_pairs)
    return fail(OGJK_ERR_INVALID, "implicit pairs need %lld bodies, scene has %d",
                2 * (long long)n_pairs, scene->n_bodies);
  ogjk::SceneArgs sa{scene->x, scene->y, scene->z, scene->body_off, scene->body_cnt,
                     nullptr, scene->n_bodies, d_pairs};
  sa.xyz = scene->xyz;
  const ogjk::OutArgs oa{d_simplices, d_distances, d_normals, d_f64};
  void* ws = nullptr;
  if (mode != OGJK_MODE_EPA) {
    const ogjk_status s2 = cached_workspace(ogjk::workspace_bytes(n_pairs), &ws);
    if (s2 != OGJK_OK) return s2;
  }
  OGJK_HIP(ogjk::launch_scene_f64(mode, sa, oa, n_pairs, ws, (hipStream_t)stream), "kernel launch");
  return OGJK_OK;
}

ogjk_status ogjk_collide_polytopes(int mode, const gkPolytope* d_bd1, const gkPolytope* d_bd2,
                                   int32_t n, gkSimplex* d_simplices, float* d_distances,
                                   float* d_normals, void* stream) {
  const ogjk_status st = check_out(mode, n, d_simplices, d_distances, d_normals);
  if (st != OGJK_OK) return st;
  if (n == 0) return OGJK_OK;
  if (!d_bd1 || !d_bd2) return fail(OGJK_ERR_INVALID, "NULL polytope array");
  const ogjk::PolyArgs pa{d_bd1, d_bd2, nullptr, 0};
  const ogjk::OutArgs oa{d_simplices, d_distances, d_normals, nullptr};
  void* ws = nullptr;
  if (mode != OGJK_MODE_EPA) {
    const ogjk_status s2 = cached_workspace(ogjk::workspace_bytes(n), &ws);
    if (s2 != OGJK_OK) return s2;
  }
  OGJK_HIP(ogjk::launch_poly(mode, pa, oa, n, ws, (hipStream_t)stream), "kernel launch");
  return OGJK_OK;
}

ogjk_status ogjk_collide_polytopes_indexed(int mode, const gkPolytope* d_pool,
                                           int32_t n_polytopes, const gkCollisionPair* d_pairs,
                                           int32_t n_pairs, gkSimplex* d_simplices,
                                           float* d_distances, float* d_normals, void* stream) {
  const ogjk_status st = check_out(mode, n_pairs, d_simplices, d_distances, d_normals);
  if (st != OGJK_OK) return st;
  if (n_pairs == 0) return OGJK_OK;
  if (!d_pool || !d_pairs || n_polytopes <= 0)
    return fail(OGJK_ERR_INVALID, "NULL pool/pairs or empty pool");
  const ogjk::PolyArgs pa{d_pool, nullptr, d_pairs, n_polytopes};
  const ogjk::OutArgs oa{d_simplices, d_distances, d_normals, nullptr};
  void* ws = nullptr;
  if (mode != OGJK_MODE_EPA) {
    const ogjk_status s2 = cached_workspace(ogjk::workspace_bytes(n_pairs), &ws);
    if (s2 != OGJK_OK) return s2;
  }
  OGJK_HIP(ogjk::launch_poly(mode, pa, oa, n_pairs, ws, (hipStream_t)stream), "kernel launch");
  return OGJK_OK;
}

ogjk_status ogjk_collide_host(int mode, int32_t n, const gkPolytope* bd1, const gkPolytope* bd2,
                              gkSimplex* simplices, float* distances, float* normals) {
  const ogjk_status st = check_out(mode, n, simplices, distances, normals);
  if (st != OGJK_OK) return st;
  if (n == 0) return OGJK_OK;
  if (!bd1 || !bd2) return fail(OGJK_ERR_INVALID, "NULL polytope array");
  const gkPolytope* lists[2] = {bd1, bd2};
  const int counts[2] = {n, n};
  return run_host(mode, lists, counts, 2, true, nullptr, n, simplices, distances, normals);
}

ogjk_status ogjk_collide_host_indexed(int mode, int32_t n_polytopes, int32_t n_pairs,
                                      const gkPolytope* polytopes, const gkCollisionPair* pairs,
                                      gkSimplex* simplices, float* distances, float* normals) {
  const ogjk_status st = check_out(mode, n_pairs, simplices, distances, normals);
  if (st != OGJK_OK) return st;
  if (n_pairs == 0) return OGJK_OK;
  if (!polytopes || !pairs || n_polytopes <= 0)
    return fail(OGJK_ERR_INVALID, "NULL polytopes/pairs or empty pool");
  const gkPolytope* lists[1] = {polytopes};
  const int counts[1] = {n_polytopes};
  return run_host(mode, lists, counts, 1, false, pairs, n_pairs, simplices, distances, normals);
}

}  // extern "C"
