// Synthetic workload generator: see include/opengjk/ogjk_gen.h.
// Only correctly rounded operations (+ - * / sqrt) and glibc's exp/log/cos
// are used, so a given image produces the same floats on every host.
#include "opengjk/ogjk_gen.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <thread>
#include <vector>

namespace {

inline uint64_t mix64(uint64_t z) {
  z += 0x9e3779b97f4a7c15ULL;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

// splitmix64 stream keyed by (seed, family, body, purpose).
struct Stream {
  uint64_t s;
  Stream(uint64_t seed, int family, int64_t body, uint64_t purpose) {
    s = mix64(mix64(seed ^ 0x6a09e667f3bcc909ULL) ^
              mix64((uint64_t)body * 0xd1b54a32d192ed03ULL + (uint64_t)family * 0x9e37ULL +
                    purpose * 0xa0761d6478bd642fULL));
  }
  uint64_t next() {
    s += 0x9e3779b97f4a7c15ULL;
    uint64_t z = s;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
  }
  double u01() { return (double)(next() >> 11) * 0x1.0p-53; }  // [0,1)
  double sym() { return 2.0 * u01() - 1.0; }                    // [-1,1)
  // Uniform direction on S^2 by rejection from the cube.
  void dir3(double* d) {
    for (;;) {
      const double x = sym(), y = sym(), z = sym();
      const double s = x * x + y * y + z * z;
      if (s <= 1.0 && s > 1e-6) {
        const double r = std::sqrt(s);
        d[0] = x / r;
        d[1] = y / r;
        d[2] = z / r;
        return;
      }
    }
  }
  // Uniform unit quaternion by rejection from the 4-cube.
  void quat(double* q) {
    for (;;) {
      const double a = sym(), b = sym(), c = sym(), d = sym();
      const double s = a * a + b * b + c * c + d * d;
      if (s <= 1.0 && s > 1e-6) {
        const double r = std::sqrt(s);
        q[0] = a / r;
        q[1] = b / r;
        q[2] = c / r;
        q[3] = d / r;
        return;
      }
    }
  }
  // Standard normal (Box-Muller, cosine branch).
  double normal() {
    double u1 = u01();
    if (u1 < 1e-300) u1 = 1e-300;
    const double u2 = u01();
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  }
};

enum : uint64_t { kCount = 1, kShape = 2, kVerts = 3 };

int body_count(const ogjk_gen_spec& sp, int64_t b) {
  if (sp.family == OGJK_GEN_ELLIPSOID) {
    Stream st(sp.seed, sp.family, b, kCount);
    const double lo = std::log((double)sp.nmin), hi = std::log((double)sp.nmax + 1.0);
    int n = (int)std::floor(std::exp(lo + (hi - lo) * st.u01()));
    return std::min(std::max(n, sp.nmin), sp.nmax);
  }
  return sp.nverts;
}

void fill_body(const ogjk_gen_spec& sp, int64_t b, int n, float* out) {
  Stream shape(sp.seed, sp.family, b, kShape);
  Stream vs(sp.seed, sp.family, b, kVerts);
  double c[3];
  if (sp.family == OGJK_GEN_SPHERE) {
    const double r = 0.5 + shape.u01();
    for (double& ci : c) ci = sp.box * shape.sym();
    double d[3];
    for (int k = 0; k < n; k++) {
      vs.dir3(d);
      for (int t = 0; t < 3; t++) out[3 * k + t] = (float)(c[t] + r * d[t]);
    }
  } else if (sp.family == OGJK_GEN_ELLIPSOID) {
    double ax[3], q[4];
    for (double& a : ax) a = 0.3 + 1.7 * shape.u01();
    shape.quat(q);
    for (double& ci : c) ci = sp.box * shape.sym();
    const double w = q[0], x = q[1], y = q[2], z = q[3];
    const double R[3][3] = {
        {1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)},
        {2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)},
        {2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)}};
    double d[3];
    for (int k = 0; k < n; k++) {
      vs.dir3(d);
      const double e[3] = {ax[0] * d[0], ax[1] * d[1], ax[2] * d[2]};
      for (int t = 0; t < 3; t++)
        out[3 * k + t] = (float)(c[t] + (R[t][0] * e[0] + R[t][1] * e[1] + R[t][2] * e[2]));
    }
  } else {  // OGJK_GEN_NOISY_SPHERE
    for (double& ci : c) ci = sp.box * shape.sym();
    double d[3];
    for (int k = 0; k < n; k++) {
      vs.dir3(d);
      for (int t = 0; t < 3; t++) out[3 * k + t] = (float)(c[t] + d[t] + sp.sigma * vs.normal());
    }
  }
}

}  // namespace

extern "C" int64_t ogjk_gen_counts(const ogjk_gen_spec* spec, int32_t* counts) {
  int64_t total = 0;
  for (int b = 0; b < spec->n_bodies; b++) {
    counts[b] = body_count(*spec, b);
    total += counts[b];
  }
  return total;
}

extern "C" void ogjk_gen_fill(const ogjk_gen_spec* spec, const int32_t* counts,
                              const int64_t* offsets, float* coords, int nthreads) {
  const int nb = spec->n_bodies;
  if (nthreads <= 1 || nb < 64) {
    for (int b = 0; b < nb; b++) fill_body(*spec, b, counts[b], coords + 3 * offsets[b]);
    return;
  }
  std::atomic<int> next{0};
  auto work = [&]() {
    for (;;) {
      const int start = next.fetch_add(64);
      if (start >= nb) return;
      const int end = std::min(nb, start + 64);
      for (int b = start; b < end; b++) fill_body(*spec, b, counts[b], coords + 3 * offsets[b]);
    }
  };
  std::vector<std::thread> th;
  for (int t = 0; t < nthreads; t++) th.emplace_back(work);
  for (auto& t : th) t.join();
}
