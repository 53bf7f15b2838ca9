// Test driver for the reference-named C++ API of libogjk.so, written the way
// a program built against the reference headers calls it: the API
// consistency section of examples/main.cpp:378-644 (computeDistances twice,
// the indexed host and device forms, computeGJKAndEPA, computeEPA /
// computeCollisionInformation after GJK, compute_gjk_epa_indexed,
// compute_epa_indexed), plus the flat *_device path
// (GJK/gpu/openGJK.h:212-299) and device-built gkPolytope arrays
// (visualization/integrate_final_gjk.cu:691-704).
//
//   dropin_api_check <input.bin> <out_dir>
//
// input.bin: int32 n_pairs, int32 cnt[2 n_pairs], float32 coords[3 * sum cnt]
// (bodies in order A0, B0, A1, B1, ...; pair i is (A_i, B_i)).  For every
// call it writes <out_dir>/<name>.bin: n gkSimplex (108 B), n distances,
// 3n normals (zero when the call computes none).  The comparison with the
// FP64 oracle is done by the Python test that runs this program.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "opengjk/gpu.h"

namespace {

int g_fail = 0;

void expect(bool ok, const char* what) {
  if (!ok) {
    std::fprintf(stderr, "FAIL: %s (%s)\n", what, ogjk_last_error());
    g_fail = 1;
  }
}

void check_status(const char* what) {
  expect(GJK::GPU::lastStatus() == OGJK_OK, what);
}

struct Out {
  std::vector<gkSimplex> s;
  std::vector<gkFloat> d, nrm;
  explicit Out(int n) : s(n), d(n), nrm(3 * (size_t)n, 0.0f) {
    for (auto& x : s) std::memset(&x, 0, sizeof(x));
  }
};

void write_out(const std::string& dir, const char* name, const Out& o) {
  const std::string path = dir + "/" + name + ".bin";
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) {
    std::fprintf(stderr, "FAIL: cannot write %s\n", path.c_str());
    g_fail = 1;
    return;
  }
  const size_t n = o.d.size();
  std::fwrite(o.s.data(), sizeof(gkSimplex), n, f);
  std::fwrite(o.d.data(), sizeof(gkFloat), n, f);
  std::fwrite(o.nrm.data(), sizeof(gkFloat), 3 * n, f);
  std::fclose(f);
}

// Device-built polytope descriptors, as a simulation builds them from its
// own device buffers (integrate_final_gjk.cu:691-704).
__global__ void build_polytopes(gkPolytope* out, gkFloat* coords, const long long* off,
                                const int* cnt, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  gkPolytope p;
  p.numpoints = cnt[i];
  p.s[0] = p.s[1] = p.s[2] = 0.0f;
  p.s_idx = 0;
  p.coord = coords + 3 * off[i];
  out[i] = p;
}

#define HIPCHECK(x) expect((x) == hipSuccess, #x)

}  // namespace

int main(int argc, char** argv) {
  if (argc != 3) {
    std::fprintf(stderr, "usage: %s input.bin out_dir\n", argv[0]);
    return 2;
  }
  const std::string dir = argv[2];
  FILE* f = std::fopen(argv[1], "rb");
  if (!f) return 2;
  int n = 0;
  if (std::fread(&n, 4, 1, f) != 1 || n <= 0) return 2;
  std::vector<int> cnt(2 * (size_t)n);
  if (std::fread(cnt.data(), 4, cnt.size(), f) != cnt.size()) return 2;
  std::vector<long long> off(cnt.size());
  long long total = 0;
  for (size_t b = 0; b < cnt.size(); b++) {
    off[b] = total;
    total += cnt[b];
  }
  std::vector<gkFloat> coords(3 * (size_t)total);
  if (std::fread(coords.data(), 4, coords.size(), f) != coords.size()) return 2;
  std::fclose(f);

  // host polytopes as main.cpp:144-160 builds them
  std::vector<gkPolytope> p1(n), p2(n);
  for (int i = 0; i < n; i++) {
    for (int side = 0; side < 2; side++) {
      gkPolytope& p = side ? p2[i] : p1[i];
      const size_t b = 2 * (size_t)i + side;
      p.numpoints = cnt[b];
      p.s[0] = p.s[1] = p.s[2] = 0.0f;
      p.s_idx = 0;
      p.coord = coords.data() + 3 * off[b];
    }
  }
  // indexed pool in reverse body order: pair i = (2n-1-2i, 2n-2-2i)
  const int npool = 2 * n;
  std::vector<gkPolytope> pool(npool);
  std::vector<gkCollisionPair> pairs(n);
  for (int i = 0; i < n; i++) {
    pool[npool - 1 - 2 * i] = p1[i];
    pool[npool - 2 - 2 * i] = p2[i];
    pairs[i].idx1 = npool - 1 - 2 * i;
    pairs[i].idx2 = npool - 2 - 2 * i;
  }

  // ---- GJK: GJK::GPU::computeDistances, twice (main.cpp:166-170)
  for (int rep = 0; rep < 2; rep++) {
    Out o(n);
    GJK::GPU::computeDistances(n, p1.data(), p2.data(), o.s.data(), o.d.data());
    check_status("GJK::GPU::computeDistances");
    expect(GJK::GPU::timer().getGpuElapsedTimeForPreviousOperation() >= 0.0f, "timer");
    write_out(dir, rep ? "gjk_ns_2" : "gjk_ns_1", o);
  }
  Out gjk(n);
  // ---- GJK: flat compute_minimum_distance (openGJK.h:91-97)
  compute_minimum_distance(n, p1.data(), p2.data(), gjk.s.data(), gjk.d.data());
  check_status("compute_minimum_distance");
  write_out(dir, "gjk_flat", gjk);
  // ---- GJK: allocate_and_copy + compute_minimum_distance_device + copy + free
  {
    Out o(n);
    gkPolytope *d1 = nullptr, *d2 = nullptr;
    gkFloat *c1 = nullptr, *c2 = nullptr, *dd = nullptr;
    gkSimplex* ds = nullptr;
    allocate_and_copy_device_arrays(n, p1.data(), p2.data(), &d1, &d2, &c1, &c2, &ds, &dd);
    expect(d1 && d2 && c1 && c2 && ds && dd, "allocate_and_copy_device_arrays");
    compute_minimum_distance_device(n, d1, d2, ds, dd);
    check_status("compute_minimum_distance_device");
    copy_results_from_device(n, ds, dd, o.s.data(), o.d.data());
    write_out(dir, "gjk_device", o);
    // ---- EPA on the same device arrays: compute_epa_device (openGJK.h:255-262)
    gkFloat* dn = nullptr;
    HIPCHECK(hipMalloc(&dn, 3 * (size_t)n * sizeof(gkFloat)));
    compute_epa_device(n, d1, d2, ds, dd, dn);
    check_status("compute_epa_device");
    Out e(n);
    copy_results_from_device(n, ds, dd, e.s.data(), e.d.data());
    HIPCHECK(hipMemcpy(e.nrm.data(), dn, 3 * (size_t)n * sizeof(gkFloat), hipMemcpyDeviceToHost));
    write_out(dir, "epa_device", e);
    HIPCHECK(hipFree(dn));
    free_device_arrays(d1, d2, c1, c2, ds, dd);
  }
  // ---- GJK: compute_minimum_distance_indexed (main.cpp:405-430)
  Out gidx(n);
  compute_minimum_distance_indexed(npool, n, pool.data(), pairs.data(), gidx.s.data(),
                                   gidx.d.data());
  check_status("compute_minimum_distance_indexed");
  write_out(dir, "gjk_indexed", gidx);
  // ---- GJK: compute_minimum_distance_indexed_device, descriptors rewritten
  // on the host to point into one device staging buffer (main.cpp:436-506)
  {
    long long pool_total = 0;
    for (int k = 0; k < npool; k++) pool_total += pool[k].numpoints;
    gkFloat* d_coords = nullptr;
    gkPolytope* d_pool = nullptr;
    gkCollisionPair* d_pairs = nullptr;
    gkSimplex* d_s = nullptr;
    gkFloat* d_d = nullptr;
    HIPCHECK(hipMalloc(&d_coords, 3 * (size_t)pool_total * sizeof(gkFloat)));
    HIPCHECK(hipMalloc(&d_pool, npool * sizeof(gkPolytope)));
    HIPCHECK(hipMalloc(&d_pairs, n * sizeof(gkCollisionPair)));
    HIPCHECK(hipMalloc(&d_s, n * sizeof(gkSimplex)));
    HIPCHECK(hipMalloc(&d_d, n * sizeof(gkFloat)));
    std::vector<gkPolytope> tmp(pool);
    std::vector<gkFloat> staging(3 * (size_t)pool_total);
    size_t o3 = 0;
    for (int k = 0; k < npool; k++) {
      const size_t m = 3 * (size_t)pool[k].numpoints;
      std::memcpy(staging.data() + o3, pool[k].coord, m * sizeof(gkFloat));
      tmp[k].coord = d_coords + o3;
      o3 += m;
    }
    HIPCHECK(hipMemcpy(d_coords, staging.data(), staging.size() * sizeof(gkFloat),
                       hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(d_pool, tmp.data(), npool * sizeof(gkPolytope), hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(d_pairs, pairs.data(), n * sizeof(gkCollisionPair), hipMemcpyHostToDevice));
    HIPCHECK(hipMemset(d_s, 0, n * sizeof(gkSimplex)));
    GJK::GPU::timer().startGpuTimer();
    compute_minimum_distance_indexed_device(n, d_pool, d_pairs, d_s, d_d);
    GJK::GPU::timer().endGpuTimer();
    check_status("compute_minimum_distance_indexed_device (host-built)");
    Out o(n);
    HIPCHECK(hipMemcpy(o.s.data(), d_s, n * sizeof(gkSimplex), hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(o.d.data(), d_d, n * sizeof(gkFloat), hipMemcpyDeviceToHost));
    write_out(dir, "gjk_indexed_device_host_built", o);
    HIPCHECK(hipFree(d_coords));
    HIPCHECK(hipFree(d_pool));
    HIPCHECK(hipFree(d_pairs));
    HIPCHECK(hipFree(d_s));
    HIPCHECK(hipFree(d_d));
  }
  // ---- GJK + EPA on descriptors built on the device from a device pool
  // (integrate_final_gjk.cu:691-704, 1016-1035): compute_minimum_distance_
  // indexed_device then compute_epa_indexed_device, pairs uploaded by
  // upload_pairs_device
  {
    const int nb = 2 * n;
    gkFloat* d_coords = nullptr;
    long long* d_off = nullptr;
    int* d_cnt = nullptr;
    gkPolytope* d_polys = nullptr;
    HIPCHECK(hipMalloc(&d_coords, coords.size() * sizeof(gkFloat)));
    HIPCHECK(hipMalloc(&d_off, nb * sizeof(long long)));
    HIPCHECK(hipMalloc(&d_cnt, nb * sizeof(int)));
    HIPCHECK(hipMalloc(&d_polys, nb * sizeof(gkPolytope)));
    HIPCHECK(hipMemcpy(d_coords, coords.data(), coords.size() * sizeof(gkFloat),
                       hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(d_off, off.data(), nb * sizeof(long long), hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(d_cnt, cnt.data(), nb * sizeof(int), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(build_polytopes, dim3((nb + 255) / 256), dim3(256), 0, 0, d_polys,
                       d_coords, (const long long*)d_off, (const int*)d_cnt, nb);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipDeviceSynchronize());
    std::vector<gkCollisionPair> pr(n);
    for (int i = 0; i < n; i++) {
      pr[i].idx1 = 2 * i;
      pr[i].idx2 = 2 * i + 1;
    }
    gkCollisionPair* d_pairs = nullptr;
    gkSimplex* d_s = nullptr;
    gkFloat *d_d = nullptr, *d_n = nullptr;
    HIPCHECK(hipMalloc(&d_pairs, n * sizeof(gkCollisionPair)));
    HIPCHECK(hipMalloc(&d_s, n * sizeof(gkSimplex)));
    HIPCHECK(hipMalloc(&d_d, n * sizeof(gkFloat)));
    HIPCHECK(hipMalloc(&d_n, 3 * (size_t)n * sizeof(gkFloat)));
    upload_pairs_device(n, pr.data(), d_pairs);
    compute_minimum_distance_indexed_device(n, d_polys, d_pairs, d_s, d_d);
    check_status("compute_minimum_distance_indexed_device (device-built)");
    Out o(n);
    HIPCHECK(hipMemcpy(o.s.data(), d_s, n * sizeof(gkSimplex), hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(o.d.data(), d_d, n * sizeof(gkFloat), hipMemcpyDeviceToHost));
    write_out(dir, "gjk_indexed_device_device_built", o);
    compute_epa_indexed_device(n, d_polys, d_pairs, d_s, d_d, d_n);
    check_status("compute_epa_indexed_device (device-built)");
    Out e(n);
    HIPCHECK(hipMemcpy(e.s.data(), d_s, n * sizeof(gkSimplex), hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(e.d.data(), d_d, n * sizeof(gkFloat), hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(e.nrm.data(), d_n, 3 * (size_t)n * sizeof(gkFloat), hipMemcpyDeviceToHost));
    write_out(dir, "epa_indexed_device_device_built", e);
    HIPCHECK(hipFree(d_coords));
    HIPCHECK(hipFree(d_off));
    HIPCHECK(hipFree(d_cnt));
    HIPCHECK(hipFree(d_polys));
    HIPCHECK(hipFree(d_pairs));
    HIPCHECK(hipFree(d_s));
    HIPCHECK(hipFree(d_d));
    HIPCHECK(hipFree(d_n));
  }
  // ---- indexed device arrays through the library's own allocators
  // (allocate_indexed_device / upload_pairs_device / free_indexed_device)
  {
    gkPolytope* d_pool = nullptr;
    gkFloat *d_coords = nullptr, *d_d = nullptr, *d_n = nullptr;
    gkCollisionPair* d_pairs = nullptr;
    gkSimplex* d_s = nullptr;
    allocate_indexed_device(npool, n, pool.data(), &d_pool, &d_coords, &d_pairs, &d_s, &d_d, &d_n);
    expect(d_pool && d_coords && d_pairs && d_s && d_d && d_n, "allocate_indexed_device");
    upload_pairs_device(n, pairs.data(), d_pairs);
    compute_minimum_distance_indexed_device(n, d_pool, d_pairs, d_s, d_d);
    compute_epa_indexed_device(n, d_pool, d_pairs, d_s, d_d, d_n);
    check_status("allocate_indexed_device path");
    Out e(n);
    HIPCHECK(hipMemcpy(e.s.data(), d_s, n * sizeof(gkSimplex), hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(e.d.data(), d_d, n * sizeof(gkFloat), hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(e.nrm.data(), d_n, 3 * (size_t)n * sizeof(gkFloat), hipMemcpyDeviceToHost));
    write_out(dir, "epa_indexed_device_allocated", e);
    free_indexed_device(d_pool, d_coords, d_pairs, d_s, d_d, d_n);
  }
  // ---- GJK + EPA: GJK::GPU::computeGJKAndEPA (main.cpp:509-529)
  {
    Out o(n);
    GJK::GPU::computeGJKAndEPA(n, p1.data(), p2.data(), o.s.data(), o.d.data(), o.nrm.data());
    check_status("GJK::GPU::computeGJKAndEPA");
    write_out(dir, "gjkepa_ns", o);
  }
  // ---- GJK + EPA: flat compute_gjk_epa (openGJK.h:134-141)
  {
    Out o(n);
    compute_gjk_epa(n, p1.data(), p2.data(), o.s.data(), o.d.data(), o.nrm.data());
    check_status("compute_gjk_epa");
    write_out(dir, "gjkepa_flat", o);
  }
  // ---- GJK + EPA: compute_gjk_epa_indexed (main.cpp:558-595)
  {
    Out o(n);
    compute_gjk_epa_indexed(npool, n, pool.data(), pairs.data(), o.s.data(), o.d.data(),
                            o.nrm.data());
    check_status("compute_gjk_epa_indexed");
    write_out(dir, "gjkepa_indexed", o);
  }
  // ---- EPA after GJK: GJK::GPU::computeDistances then computeEPA (main.cpp:531-556)
  {
    Out o(n);
    GJK::GPU::computeDistances(n, p1.data(), p2.data(), o.s.data(), o.d.data());
    GJK::GPU::computeEPA(n, p1.data(), p2.data(), o.s.data(), o.d.data(), o.nrm.data());
    check_status("GJK::GPU::computeEPA");
    write_out(dir, "epa_ns", o);
  }
  // ---- EPA after GJK: GJK::GPU::computeCollisionInformation, both forms
  {
    Out o(n);
    GJK::GPU::computeDistances(n, p1.data(), p2.data(), o.s.data(), o.d.data());
    GJK::GPU::computeCollisionInformation(n, p1.data(), p2.data(), o.s.data(), o.d.data(),
                                          o.nrm.data());
    check_status("GJK::GPU::computeCollisionInformation");
    write_out(dir, "epa_ci_ns", o);
    Out r(n);
    std::vector<gkFloat> w1(3 * (size_t)n), w2(3 * (size_t)n);
    GJK::GPU::computeDistances(n, p1.data(), p2.data(), r.s.data(), r.d.data());
    GJK::GPU::computeCollisionInformation(n, p1.data(), p2.data(), r.s.data(), r.d.data(),
                                          w1.data(), w2.data(), r.nrm.data());
    check_status("GJK::GPU::computeCollisionInformation (README form)");
    bool same = true;
    for (int i = 0; i < n; i++)
      for (int c = 0; c < 3; c++)
        same &= w1[3 * i + c] == r.s[i].witnesses[0][c] && w2[3 * i + c] == r.s[i].witnesses[1][c];
    expect(same, "README form: witness arrays equal simplices[i].witnesses");
    write_out(dir, "epa_ci_readme", r);
  }
  // ---- EPA after GJK: flat computeCollisionInformation (openGJK.h:112-119)
  {
    Out o(n);
    o.s = gjk.s;
    o.d = gjk.d;
    computeCollisionInformation(n, p1.data(), p2.data(), o.s.data(), o.d.data(), o.nrm.data());
    check_status("computeCollisionInformation");
    write_out(dir, "epa_flat", o);
  }
  // ---- EPA after GJK: compute_epa_indexed on the indexed GJK outputs (main.cpp:597-637)
  {
    Out o(n);
    o.s = gidx.s;
    o.d = gidx.d;
    compute_epa_indexed(npool, n, pool.data(), pairs.data(), o.s.data(), o.d.data(), o.nrm.data());
    check_status("compute_epa_indexed");
    write_out(dir, "epa_indexed", o);
  }
  // ---- the EPA result allocators and copies (openGJK.h:155-194)
  {
    gkFloat *dw1 = nullptr, *dw2 = nullptr, *dn = nullptr;
    allocate_epa_device_arrays(n, &dw1, &dw2, &dn);
    expect(dw1 && dw2 && dn, "allocate_epa_device_arrays");
    std::vector<gkFloat> a(3 * (size_t)n), b(3 * (size_t)n), c(3 * (size_t)n);
    for (size_t k = 0; k < a.size(); k++) {
      a[k] = (gkFloat)k;
      b[k] = -(gkFloat)k;
      c[k] = 0.5f * (gkFloat)k;
    }
    HIPCHECK(hipMemcpy(dw1, a.data(), a.size() * sizeof(gkFloat), hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(dw2, b.data(), b.size() * sizeof(gkFloat), hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(dn, c.data(), c.size() * sizeof(gkFloat), hipMemcpyHostToDevice));
    std::vector<gkFloat> a2(a.size()), b2(b.size()), c2(c.size());
    copy_epa_results_from_device(n, dw1, dw2, dn, a2.data(), b2.data(), c2.data());
    expect(a2 == a && b2 == b && c2 == c, "copy_epa_results_from_device round trip");
    free_epa_device_arrays(dw1, dw2, dn);
  }
  std::printf("%s: %d pairs, %s\n", argv[0], n, g_fail ? "FAILED" : "all calls returned");
  return g_fail;
}
