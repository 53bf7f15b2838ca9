// Broad phase: uniform-grid sphere overlap -> gkCollisionPair lists, and the
// body-pair -> sub-mesh-pair expansion, on the device.
//
// Same pair set and order as the reference's grid
// (visualization/integrate_final_gjk.cu:467-570, 916-1002, 265-302), built the
// other way round: instead of fixed-size per-cell slot arrays filled by
// atomics, objects are counting-sorted by cell into one array (a histogram by
// atomics, one scan that is also the cell start table, placement, then each
// object's rank among the lower indices of its cell, so every cell lists its
// objects in ascending index, one of the orders the reference's atomics can
// produce; no cell overflows).  Threads walk objects in cell order,
// so the lanes of a wave query the same neighbour cells and their loads
// coalesce or hit in cache; the three x-neighbours of a row are one contiguous
// range when cells are not capped.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <stdint.h>

#include <cmath>

#include "ogjk_internal.hpp"
#include "opengjk/ogjk.h"

ogjk_status ogjk_fail(ogjk_status code, const char* fmt, ...);

namespace {

constexpr int kBlock = 256;
constexpr int64_t kMaxCells = (int64_t(1) << 30) - 1;  // cells + 1 scanned as int

__device__ __forceinline__ int cell_axis(float p, float boundary, float cell, int g) {
  const float f = floorf((p + boundary) / cell);
  if (!(f >= 0.0f)) return 0;  // negative or NaN
  if (f > (float)(g - 1)) return g - 1;
  return (int)f;
}

__device__ __forceinline__ bool overlap(const float4 a, const float4 b) {
  const float dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
  const float d2 = (dx * dx + dy * dy) + dz * dz;
  const float rs = a.w + b.w;
  return d2 < rs * rs;
}

__global__ __launch_bounds__(kBlock) void k_bp_zero(int32_t* __restrict__ a, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock)
    a[i] = 0;
}

// keys[i] = cell of object i; slot[i] = its arrival slot in the cell's count.
__global__ __launch_bounds__(kBlock) void k_bp_hist(const float4* __restrict__ sph, int n,
                                                    float cell, float boundary, int g,
                                                    uint32_t* __restrict__ keys,
                                                    int32_t* __restrict__ slot,
                                                    int32_t* __restrict__ count) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const float4 s = sph[i];
  const uint32_t cx = cell_axis(s.x, boundary, cell, g);
  const uint32_t cy = cell_axis(s.y, boundary, cell, g);
  const uint32_t cz = cell_axis(s.z, boundary, cell, g);
  const uint32_t k = cx + (uint32_t)g * (cy + (uint32_t)g * cz);
  keys[i] = k;
  slot[i] = atomicAdd(count + k, 1);
}

// Crowded cells: objects sorted by (cell, index) by a stable radix sort
// instead of ranked by a scan of their cell.  order[t] = the object at
// sorted position t, okeys[t] its cell.
__global__ __launch_bounds__(kBlock) void k_bp_iota(int32_t* __restrict__ v, int n) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) v[i] = i;
}


__global__ __launch_bounds__(kBlock) void k_bp_place(const uint32_t* __restrict__ keys,
                                                     const int32_t* __restrict__ slot,
                                                     const int32_t* __restrict__ start, int n,
                                                     int32_t* __restrict__ tmp) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) tmp[start[keys[i]] + slot[i]] = i;
}

// An object in cell order: its sphere, index and cell, one 32-byte line.
struct alignas(32) BpRec {
  float4 s;
  int32_t id;
  uint32_t key;
  int32_t pad[2];
};

// A cell holding more objects than this is ordered by the radix sort (the
// per-object rank scan costs its cell's occupancy).
constexpr int kRankScanMax = 64;

// The count and emit kernels of the rank-scan pass run unless *gate is set
// (some cell was too crowded to rank), and always when gate is null (the
// radix-sorted pass).
__device__ __forceinline__ bool gated_off(const int32_t* gate) {
  return gate != nullptr && *gate != 0;
}

// Object i goes to start[cell] + (number of lower indices in its cell): cells
// list their objects in ascending index.  pos[i] = that position.
// An object of a cell holding more than kRankScanMax objects ranks nothing
// and sets *crowded (one atomic per wave); the host then sorts instead.
__global__ __launch_bounds__(kBlock) void k_bp_rank(const float4* __restrict__ sph,
                                                    const uint32_t* __restrict__ keys,
                                                    const int32_t* __restrict__ start,
                                                    const int32_t* __restrict__ tmp, int n,
                                                    BpRec* __restrict__ rec,
                                                    int32_t* __restrict__ pos,
                                                    int32_t* __restrict__ crowded) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  const uint32_t k = i < n ? keys[i] : 0u;
  const int lo = i < n ? start[k] : 0, hi = i < n ? start[k + 1] : 0;
  const bool big = hi - lo > kRankScanMax;
  const uint64_t any = __ballot(big);
  if (any != 0 && (threadIdx.x & 63) == (unsigned)__builtin_ctzll(any)) atomicOr(crowded, 1);
  if (i >= n || big) return;
  int r = 0;
  for (int t = lo; t < hi; t++) r += tmp[t] < i;
  BpRec e;
  e.s = sph[i];
  e.id = i;
  e.key = k;
  e.pad[0] = e.pad[1] = 0;
  rec[lo + r] = e;
  pos[i] = lo + r;
}

__global__ __launch_bounds__(kBlock) void k_bp_from_sorted(const float4* __restrict__ sph,
                                                           const uint32_t* __restrict__ okeys,
                                                           const int32_t* __restrict__ order,
                                                           int n, BpRec* __restrict__ rec,
                                                           int32_t* __restrict__ pos) {
  const int t = blockIdx.x * kBlock + threadIdx.x;
  if (t >= n) return;
  const int i = order[t];
  BpRec e;
  e.s = sph[i];
  e.id = i;
  e.key = okeys[t];
  e.pad[0] = e.pad[1] = 0;
  rec[t] = e;
  pos[i] = t;
}

// Partners of an object the count pass keeps for the emit pass.
constexpr int kKeep = 8;

// Calls f(j) for every j > o overlapping the object at sorted position s, in
// neighbour-cell order (z, y, x from -1 to +1), ascending j within a cell.
template <typename F>
__device__ __forceinline__ void bp_walk(const BpRec* __restrict__ rec,
                                        const int32_t* __restrict__ start, int s, int g, int cap,
                                        F&& f) {
  const BpRec me = rec[s];
  const int o = me.id;
  const uint32_t k = me.key;
  const int cx = (int)(k % (uint32_t)g);
  const int cy = (int)((k / (uint32_t)g) % (uint32_t)g);
  const int cz = (int)(k / ((uint32_t)g * (uint32_t)g));
  const int x0 = cx > 0 ? cx - 1 : 0, x1 = cx < g - 1 ? cx + 1 : g - 1;
  if (cap == 0) {
    // uncapped: the x-neighbours of each of the 9 rows are one contiguous
    // range; all 18 bounds are loaded before any candidate (indices clamped
    // into the grid, rows outside it emptied)
    int lo[9], hi[9];
#pragma unroll
    for (int r = 0; r < 9; r++) {
      const int nz = cz + r / 3 - 1, ny = cy + r % 3 - 1;
      const bool in = nz >= 0 && nz < g && ny >= 0 && ny < g;
      const int64_t row =
          (int64_t)g * ((in ? ny : cy) + (int64_t)g * (in ? nz : cz));
      lo[r] = start[row + x0];
      hi[r] = in ? start[row + x1 + 1] : lo[r];
    }
#pragma unroll
    for (int r = 0; r < 9; r++)
      for (int t = lo[r]; t < hi[r]; t += 4) {
        // four candidates in flight; indices past the row repeat its last
        // entry and are skipped
        float4 q[4];
        int j[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
          const int tu = t + u < hi[r] ? t + u : hi[r] - 1;
          q[u] = rec[tu].s;
          j[u] = rec[tu].id;
        }
#pragma unroll
        for (int u = 0; u < 4; u++)
          if (t + u < hi[r] && j[u] > o && overlap(me.s, q[u])) f(j[u]);
      }
    return;
  }
  for (int dz = -1; dz <= 1; dz++) {
    const int nz = cz + dz;
    if (nz < 0 || nz >= g) continue;
    for (int dy = -1; dy <= 1; dy++) {
      const int ny = cy + dy;
      if (ny < 0 || ny >= g) continue;
      const int64_t row = (int64_t)g * (ny + (int64_t)g * nz);
      for (int nx = x0; nx <= x1; nx++) {
        const int lo = start[row + nx];
        int hi = start[row + nx + 1];
        if (hi - lo > cap) hi = lo + cap;
        for (int t = lo; t < hi; t++) {
          const float4 q = rec[t].s;
          const int j = rec[t].id;
          if (j > o && overlap(me.s, q)) f(j);
        }
      }
    }
  }
}

// Workgroups go to the 8 XCDs round-robin; a grid of a multiple of 8 blocks
// is remapped so that each XCD takes one contiguous eighth (a slab of cells
// whose neighbour rows mostly stay in that XCD's L2).
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_block() {
  const int per = gridDim.x / kXcds;
  return (blockIdx.x % kXcds) * per + blockIdx.x / kXcds;
}
unsigned xcd_blocks(int64_t threads) {
  const int64_t b = (threads + kBlock - 1) / kBlock;
  return (unsigned)((b + kXcds - 1) / kXcds * kXcds);
}

// Count pass, one thread per object in cell order: counts[o] = number of
// pairs (o, j), j > o (counts[n] = 0, so the exclusive scan's last entry is
// the total), and the first kKeep partners in keep[kKeep o ..].  Launched on
// xcd_blocks(n + 1) blocks.
__global__ __launch_bounds__(kBlock) void k_bp_count(const BpRec* __restrict__ rec,
                                                     const int32_t* __restrict__ start, int n,
                                                     int g, int cap,
                                                     int64_t* __restrict__ counts,
                                                     int32_t* __restrict__ keep,
                                                     const int32_t* __restrict__ gate) {
  const int s = xcd_block() * kBlock + threadIdx.x;
  if (gated_off(gate)) return;
  if (s >= n) {
    if (s == n) counts[n] = 0;
    return;
  }
  const int o = rec[s].id;
  int64_t cnt = 0;
  int32_t* kp = keep + (int64_t)kKeep * o;
  bp_walk(rec, start, s, g, cap, [&](int j) {
    if (cnt < kKeep) kp[cnt] = j;
    cnt++;
  });
  counts[o] = cnt;
}

// Emit pass, one thread per object in index order (the output is in index
// order): copies the kept partners to offsets[o] or, for an object with more
// than kKeep, walks its neighbours again; positions at or past max_pairs are
// dropped.
__global__ __launch_bounds__(kBlock) void k_bp_emit(const BpRec* __restrict__ rec,
                                                    const int32_t* __restrict__ start,
                                                    const int32_t* __restrict__ pos, int n, int g,
                                                    int cap, const int64_t* __restrict__ counts,
                                                    const int32_t* __restrict__ keep,
                                                    const int64_t* __restrict__ offsets,
                                                    gkCollisionPair* __restrict__ pairs,
                                                    int64_t max_pairs,
                                                    const int32_t* __restrict__ gate) {
  const int o = blockIdx.x * kBlock + threadIdx.x;
  if (o >= n || gated_off(gate)) return;
  int64_t w = offsets[o];
  const int64_t c = counts[o];
  if (c <= kKeep) {
    const int32_t* kp = keep + (int64_t)kKeep * o;
    for (int q = 0; q < c && w < max_pairs; q++, w++) pairs[w] = gkCollisionPair{o, kp[q]};
    return;
  }
  bp_walk(rec, start, pos[o], g, cap, [&](int j) {
    if (w < max_pairs) pairs[w] = gkCollisionPair{o, j};
    w++;
  });
}

// Sub-mesh expansion: counts[p] = sub_cnt[a] * sub_cnt[b] (0 for an index out
// of [0, n_bodies)), counts[n] = 0.
__global__ __launch_bounds__(kBlock) void k_sub_counts(const gkCollisionPair* __restrict__ bp,
                                                       int n, const int32_t* __restrict__ sub_cnt,
                                                       int nb, int64_t* __restrict__ counts) {
  const int p = blockIdx.x * kBlock + threadIdx.x;
  if (p > n) return;
  if (p == n) {
    counts[n] = 0;
    return;
  }
  const gkCollisionPair q = bp[p];
  int64_t c = 0;
  if (q.idx1 >= 0 && q.idx1 < nb && q.idx2 >= 0 && q.idx2 < nb) {
    const int64_t ca = sub_cnt[q.idx1], cb = sub_cnt[q.idx2];
    if (ca > 0 && cb > 0) c = ca * cb;
  }
  counts[p] = c;
}

__global__ __launch_bounds__(kBlock) void k_sub_expand(
    const gkCollisionPair* __restrict__ bp, int n, const int32_t* __restrict__ sub_off,
    const int32_t* __restrict__ sub_cnt, const int64_t* __restrict__ counts,
    const int64_t* __restrict__ offsets, gkCollisionPair* __restrict__ out, int64_t max_out) {
  const int p = blockIdx.x * kBlock + threadIdx.x;
  if (p >= n) return;
  const int64_t c = counts[p];
  if (c == 0) return;
  const gkCollisionPair q = bp[p];
  const int oa = sub_off[q.idx1], ob = sub_off[q.idx2];
  const int ca = sub_cnt[q.idx1], cb = sub_cnt[q.idx2];
  int64_t w = offsets[p];
  for (int u = 0; u < ca && w < max_out; u++)
    for (int v = 0; v < cb && w < max_out; v++) out[w++] = gkCollisionPair{oa + u, ob + v};
}

inline size_t up256(size_t b) { return (b + 255) & ~size_t(255); }

// Workspace of the broad phase for n objects over `cells` cells.
struct BpLayout {
  size_t keys, slot, count, start, tmp, rec, pos, keep, counts, offsets, okeys, temp, total;
  size_t temp_bytes;
  int key_bits;  // bits of a cell key (the radix sort's range)
};

hipError_t bp_layout_query(int n, int64_t cells, BpLayout* L);

// The library's temp-storage queries cost host time on every call, so the
// layout of the last (n, cells) is kept per thread.
hipError_t bp_layout(int n, int64_t cells, BpLayout* L) {
  thread_local int last_n = -1;
  thread_local int64_t last_cells = -1;
  thread_local BpLayout last;
  if (n != last_n || cells != last_cells) {
    const hipError_t e = bp_layout_query(n, cells, &last);
    if (e != hipSuccess) return e;
    last_n = n;
    last_cells = cells;
  }
  *L = last;
  return hipSuccess;
}

hipError_t bp_layout_query(int n, int64_t cells, BpLayout* L) {
  size_t cell_b = 0, scan_b = 0, sort_b = 0;
  hipError_t e = hipcub::DeviceScan::ExclusiveSum(nullptr, cell_b, (const int32_t*)nullptr,
                                                  (int32_t*)nullptr, (int)(cells + 1));
  if (e != hipSuccess) return e;
  e = hipcub::DeviceScan::ExclusiveSum(nullptr, scan_b, (const int64_t*)nullptr,
                                       (int64_t*)nullptr, n + 1);
  if (e != hipSuccess) return e;
  L->key_bits = 1;
  while (L->key_bits < 31 && ((int64_t)1 << L->key_bits) < cells) L->key_bits++;
  e = hipcub::DeviceRadixSort::SortPairs(nullptr, sort_b, (const uint32_t*)nullptr,
                                         (uint32_t*)nullptr, (const int32_t*)nullptr,
                                         (int32_t*)nullptr, n, 0, L->key_bits);
  if (e != hipSuccess) return e;
  L->temp_bytes = cell_b > scan_b ? cell_b : scan_b;
  if (sort_b > L->temp_bytes) L->temp_bytes = sort_b;
  size_t o = 0;
  L->keys = o;    o += up256(sizeof(uint32_t) * (size_t)n);
  L->slot = o;    o += up256(sizeof(int32_t) * (size_t)n);
  // count[cells] stays 0 (the scan's last input), count[cells + 1] = the
  // crowded-cell flag
  L->count = o;   o += up256(sizeof(int32_t) * ((size_t)cells + 2));
  L->start = o;   o += up256(sizeof(int32_t) * ((size_t)cells + 1));
  L->tmp = o;     o += up256(sizeof(int32_t) * (size_t)n);
  L->rec = o;     o += up256(sizeof(BpRec) * (size_t)n);
  L->pos = o;     o += up256(sizeof(int32_t) * (size_t)n);
  L->keep = o;    o += up256(sizeof(int32_t) * kKeep * (size_t)n);
  L->counts = o;  o += up256(sizeof(int64_t) * ((size_t)n + 1));
  L->offsets = o; o += up256(sizeof(int64_t) * ((size_t)n + 1));
  L->okeys = o;   o += up256(sizeof(uint32_t) * (size_t)n);
  L->temp = o;    o += up256(L->temp_bytes);
  L->total = o;
  return hipSuccess;
}

struct ExLayout {
  size_t counts, offsets, temp, total, temp_bytes;
};

hipError_t ex_layout(int n, ExLayout* L) {
  size_t scan_b = 0;
  const hipError_t e = hipcub::DeviceScan::ExclusiveSum(nullptr, scan_b, (const int64_t*)nullptr,
                                                        (int64_t*)nullptr, n + 1);
  if (e != hipSuccess) return e;
  L->temp_bytes = scan_b;
  size_t o = 0;
  L->counts = o;  o += up256(sizeof(int64_t) * ((size_t)n + 1));
  L->offsets = o; o += up256(sizeof(int64_t) * ((size_t)n + 1));
  L->temp = o;    o += up256(scan_b);
  L->total = o;
  return hipSuccess;
}

ogjk_status hip_err(hipError_t e, const char* what) {
  return ogjk_fail(e == hipErrorOutOfMemory ? OGJK_ERR_NOMEM : OGJK_ERR_HIP, "%s: %s", what,
                   hipGetErrorString(e));
}

#define BP_HIP(call, what)                              \
  do {                                                  \
    const hipError_t e_ = (call);                       \
    if (e_ != hipSuccess) return hip_err(e_, what);     \
  } while (0)

ogjk_status check_grid(const ogjk_grid* g, int64_t* cells) {
  if (!g) return ogjk_fail(OGJK_ERR_INVALID, "NULL grid");
  if (!(g->cell_size > 0.0f) || !std::isfinite(g->cell_size) || !std::isfinite(g->boundary))
    return ogjk_fail(OGJK_ERR_INVALID, "cell_size must be finite and > 0, boundary finite");
  if (g->grid_size < 1) return ogjk_fail(OGJK_ERR_INVALID, "grid_size %d < 1", g->grid_size);
  if (g->max_per_cell < 0) return ogjk_fail(OGJK_ERR_INVALID, "max_per_cell < 0");
  const int64_t c = (int64_t)g->grid_size * g->grid_size * g->grid_size;
  if (c > kMaxCells)
    return ogjk_fail(OGJK_ERR_UNSUPPORTED, "grid of %lld cells exceeds 2^30 - 1", (long long)c);
  *cells = c;
  return OGJK_OK;
}

unsigned blocks_for(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

}  // namespace

extern "C" {

ogjk_status ogjk_broad_phase_workspace(int32_t n_objects, const ogjk_grid* grid, size_t* bytes) {
  if (!bytes) return ogjk_fail(OGJK_ERR_INVALID, "NULL bytes");
  if (n_objects < 0) return ogjk_fail(OGJK_ERR_INVALID, "n_objects < 0");
  int64_t cells = 0;
  const ogjk_status st = check_grid(grid, &cells);
  if (st != OGJK_OK) return st;
  BpLayout L;
  BP_HIP(bp_layout(n_objects, cells, &L), "workspace query");
  *bytes = L.total;
  return OGJK_OK;
}

ogjk_status ogjk_broad_phase(const float* d_spheres, int32_t n_objects, const ogjk_grid* grid,
                             gkCollisionPair* d_pairs, int32_t max_pairs, int32_t* n_pairs,
                             int64_t* n_found, void* d_workspace, size_t workspace_bytes,
                             void* stream) {
  if (n_pairs) *n_pairs = 0;
  if (n_found) *n_found = 0;
  if (n_objects < 0 || max_pairs < 0)
    return ogjk_fail(OGJK_ERR_INVALID, "n_objects and max_pairs must be >= 0");
  int64_t cells = 0;
  const ogjk_status st = check_grid(grid, &cells);
  if (st != OGJK_OK) return st;
  if (n_objects == 0) return OGJK_OK;
  if (!d_spheres || (max_pairs > 0 && !d_pairs))
    return ogjk_fail(OGJK_ERR_INVALID, "NULL spheres or pairs");
  BpLayout L;
  BP_HIP(bp_layout(n_objects, cells, &L), "workspace query");
  if (!d_workspace || workspace_bytes < L.total)
    return ogjk_fail(OGJK_ERR_INVALID, "workspace of %zu bytes needed", L.total);
  char* ws = (char*)d_workspace;
  auto* keys = (uint32_t*)(ws + L.keys);
  auto* slot = (int32_t*)(ws + L.slot);
  auto* count = (int32_t*)(ws + L.count);
  auto* start = (int32_t*)(ws + L.start);
  auto* tmp = (int32_t*)(ws + L.tmp);
  auto* rec = (BpRec*)(ws + L.rec);
  auto* pos = (int32_t*)(ws + L.pos);
  auto* keep = (int32_t*)(ws + L.keep);
  auto* counts = (int64_t*)(ws + L.counts);
  auto* offsets = (int64_t*)(ws + L.offsets);
  void* temp = ws + L.temp;
  const hipStream_t st_ = (hipStream_t)stream;
  const int n = n_objects, g = grid->grid_size;
  const float4* sph = (const float4*)d_spheres;

  // counting sort by cell; start[] is the exclusive scan of the histogram
  const int64_t zb = blocks_for(cells + 2);
  int32_t* crowded = count + cells + 1;
  hipLaunchKernelGGL(k_bp_zero, dim3((unsigned)(zb < 4096 ? zb : 4096)), dim3(kBlock), 0, st_,
                     count, cells + 2);
  hipLaunchKernelGGL(k_bp_hist, dim3(blocks_for(n)), dim3(kBlock), 0, st_, sph, n,
                     grid->cell_size, grid->boundary, g, keys, slot, count);
  BP_HIP(hipGetLastError(), "cell histogram");
  size_t tb = L.temp_bytes;
  BP_HIP(hipcub::DeviceScan::ExclusiveSum(temp, tb, count, start, (int)(cells + 1), st_),
         "cell scan");
  // Cells of at most kRankScanMax objects: each object ranks itself within its
  // cell.  When a cell holds more, k_bp_rank sets *crowded, the count and
  // emit kernels of this pass return at once, and the host then orders the
  // objects with the radix sort below and counts and emits again.
  hipLaunchKernelGGL(k_bp_place, dim3(blocks_for(n)), dim3(kBlock), 0, st_, keys, slot, start,
                     n, tmp);
  hipLaunchKernelGGL(k_bp_rank, dim3(blocks_for(n)), dim3(kBlock), 0, st_, sph, keys, start, tmp,
                     n, rec, pos, crowded);
  BP_HIP(hipGetLastError(), "cell sort");
  int64_t total = 0;
  int32_t is_crowded = 0;
  for (int pass = 0; pass < 2; pass++) {
    const int32_t* gate = pass == 0 ? crowded : nullptr;
    hipLaunchKernelGGL(k_bp_count, dim3(xcd_blocks((int64_t)n + 1)), dim3(kBlock), 0, st_, rec,
                       start, n, g, grid->max_per_cell, counts, keep, gate);
    BP_HIP(hipGetLastError(), "pair count");
    tb = L.temp_bytes;
    BP_HIP(hipcub::DeviceScan::ExclusiveSum(temp, tb, counts, offsets, n + 1, st_), "pair scan");
    if (max_pairs > 0) {
      hipLaunchKernelGGL(k_bp_emit, dim3(blocks_for(n)), dim3(kBlock), 0, st_, rec, start, pos,
                         n, g, grid->max_per_cell, counts, keep, offsets, d_pairs,
                         (int64_t)max_pairs, gate);
      BP_HIP(hipGetLastError(), "pair emit");
    }
    BP_HIP(hipMemcpyAsync(&total, offsets + n, sizeof(int64_t), hipMemcpyDeviceToHost, st_),
           "pair total");
    if (pass == 0)
      BP_HIP(hipMemcpyAsync(&is_crowded, crowded, sizeof(int32_t), hipMemcpyDeviceToHost, st_),
             "cell occupancy");
    BP_HIP(hipStreamSynchronize(st_), "broad phase");
    if (!is_crowded) break;
    if (pass == 0) {
      // crowded cells: a stable radix sort by cell lists each cell's objects
      // in ascending index
      auto* okeys = (uint32_t*)(ws + L.okeys);
      hipLaunchKernelGGL(k_bp_iota, dim3(blocks_for(n)), dim3(kBlock), 0, st_, slot, n);
      tb = L.temp_bytes;
      BP_HIP(hipcub::DeviceRadixSort::SortPairs(temp, tb, keys, okeys, slot, tmp, n, 0,
                                                L.key_bits, st_),
             "cell radix sort");
      hipLaunchKernelGGL(k_bp_from_sorted, dim3(blocks_for(n)), dim3(kBlock), 0, st_, sph, okeys,
                         tmp, n, rec, pos);
      BP_HIP(hipGetLastError(), "cell sort");
    }
  }
  if (n_found) *n_found = total;
  if (n_pairs) *n_pairs = (int32_t)(total < max_pairs ? total : max_pairs);
  return OGJK_OK;
}

ogjk_status ogjk_expand_pairs_workspace(int32_t n_body_pairs, size_t* bytes) {
  if (!bytes) return ogjk_fail(OGJK_ERR_INVALID, "NULL bytes");
  if (n_body_pairs < 0) return ogjk_fail(OGJK_ERR_INVALID, "n_body_pairs < 0");
  ExLayout L;
  BP_HIP(ex_layout(n_body_pairs, &L), "workspace query");
  *bytes = L.total;
  return OGJK_OK;
}

ogjk_status ogjk_expand_pairs(const gkCollisionPair* d_body_pairs, int32_t n_body_pairs,
                              const int32_t* d_sub_off, const int32_t* d_sub_cnt,
                              int32_t n_bodies, gkCollisionPair* d_sub_pairs,
                              int32_t max_sub_pairs, int32_t* n_sub_pairs, int64_t* n_found,
                              void* d_workspace, size_t workspace_bytes, void* stream) {
  if (n_sub_pairs) *n_sub_pairs = 0;
  if (n_found) *n_found = 0;
  if (n_body_pairs < 0 || max_sub_pairs < 0 || n_bodies < 0)
    return ogjk_fail(OGJK_ERR_INVALID, "counts must be >= 0");
  if (n_body_pairs == 0) return OGJK_OK;
  if (!d_body_pairs || !d_sub_off || !d_sub_cnt || (max_sub_pairs > 0 && !d_sub_pairs))
    return ogjk_fail(OGJK_ERR_INVALID, "NULL input or output");
  ExLayout L;
  BP_HIP(ex_layout(n_body_pairs, &L), "workspace query");
  if (!d_workspace || workspace_bytes < L.total)
    return ogjk_fail(OGJK_ERR_INVALID, "workspace of %zu bytes needed", L.total);
  char* ws = (char*)d_workspace;
  auto* counts = (int64_t*)(ws + L.counts);
  auto* offsets = (int64_t*)(ws + L.offsets);
  const hipStream_t st_ = (hipStream_t)stream;
  const int n = n_body_pairs;
  hipLaunchKernelGGL(k_sub_counts, dim3(blocks_for((int64_t)n + 1)), dim3(kBlock), 0, st_,
                     d_body_pairs, n, d_sub_cnt, n_bodies, counts);
  BP_HIP(hipGetLastError(), "sub-mesh counts");
  size_t tb = L.temp_bytes;
  BP_HIP(hipcub::DeviceScan::ExclusiveSum(ws + L.temp, tb, counts, offsets, n + 1, st_),
         "sub-mesh scan");
  if (max_sub_pairs > 0) {
    hipLaunchKernelGGL(k_sub_expand, dim3(blocks_for(n)), dim3(kBlock), 0, st_, d_body_pairs, n,
                       d_sub_off, d_sub_cnt, counts, offsets, d_sub_pairs,
                       (int64_t)max_sub_pairs);
    BP_HIP(hipGetLastError(), "sub-mesh expand");
  }
  int64_t total = 0;
  BP_HIP(hipMemcpyAsync(&total, offsets + n, sizeof(int64_t), hipMemcpyDeviceToHost, st_),
         "sub-mesh total");
  BP_HIP(hipStreamSynchronize(st_), "sub-mesh expansion");
  if (n_found) *n_found = total;
  if (n_sub_pairs) *n_sub_pairs = (int32_t)(total < max_sub_pairs ? total : max_sub_pairs);
  return OGJK_OK;
}

}  // extern "C"
