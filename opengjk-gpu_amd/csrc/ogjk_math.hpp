// opengjk-amd device math: FP64 vector helpers and wave64 group collectives.
//
// Everything here is compiled with FMA contraction off (the build passes
// -ffp-contract=off and the pragma below), so each a*b+c rounds twice, exactly
// like the CPU checker.  FMAs are only ever written explicitly.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#pragma clang fp contract(off)


namespace ogjk {

#define OGJK_DEV __device__ __forceinline__

// DBL_EPSILON / FLT_EPSILON as the reference uses them (GJK/common.h:48,55).
constexpr double kEps64 = 2.220446049250313080847e-16;
constexpr float kEps32 = 1.1920928955078125e-07f;

struct D3 {
  double x, y, z;
};

OGJK_DEV D3 d3(double x, double y, double z) { return D3{x, y, z}; }
OGJK_DEV D3 sub(D3 a, D3 b) { return D3{a.x - b.x, a.y - b.y, a.z - b.z}; }
OGJK_DEV D3 neg(D3 a) { return D3{-a.x, -a.y, -a.z}; }
// (a0*b0 + a1*b1) + a2*b2, the association of the reference's dotProduct macro.
OGJK_DEV double dot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
OGJK_DEV double norm2(D3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
OGJK_DEV D3 cross(D3 a, D3 b) {
  return D3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
OGJK_DEV D3 sel(bool c, D3 a, D3 b) { return D3{c ? a.x : b.x, c ? a.y : b.y, c ? a.z : b.z}; }

// ------------------------------------------------------------ collectives
// A "group" is LPP consecutive lanes of one wave64 (LPP in {8,16,32,64}).
// All lanes of a group are kept convergent around every collective.

OGJK_DEV int lane_id() { return __lane_id(); }

template <int LPP>
OGJK_DEV uint64_t group_ballot(bool p) {
  const uint64_t b = __ballot(p);
  if constexpr (LPP == 64) {
    return b;
  } else {
    const int base = lane_id() & ~(LPP - 1);
    return (b >> base) & ((1ull << LPP) - 1);
  }
}

template <int LPP>
OGJK_DEV bool group_any(bool p) {
  return group_ballot<LPP>(p) != 0;
}

// Number of set predicates in lower lanes of the group.
template <int LPP>
OGJK_DEV int group_prefix(uint64_t gb) {
  const int l = lane_id() & (LPP - 1);
  return __popcll(gb & ((1ull << l) - 1ull));
}

// Exchange step m (1, 2, 4, ...) of an all-reduce over LPP lanes.  Steps
// 1 and 2 are quad permutes, 4 and 8 are the row half-mirror and mirror
// (lane i <-> 7-i, 15-i): applied in increasing m after the lower steps have
// made each block of m lanes uniform they pair every block with its partner
// block, like a butterfly, without leaving the VALU.  Larger steps cross DPP
// rows through ds_bpermute.
template <int M>
OGJK_DEV int xchg_i32(int v) {
  if constexpr (M == 1) return __builtin_amdgcn_update_dpp(0, v, 0xB1, 0xF, 0xF, false);
  else if constexpr (M == 2) return __builtin_amdgcn_update_dpp(0, v, 0x4E, 0xF, 0xF, false);
  else if constexpr (M == 4) return __builtin_amdgcn_update_dpp(0, v, 0x141, 0xF, 0xF, false);
  else if constexpr (M == 8) return __builtin_amdgcn_update_dpp(0, v, 0x140, 0xF, 0xF, false);
  else if constexpr (M == 16) {
    // v_permlane16_swap with both operands = v: result 0 holds the even
    // 16-lane row's value, result 1 the odd row's, in both rows of a pair
    const auto r = __builtin_amdgcn_permlane16_swap((unsigned)v, (unsigned)v, false, false);
    return (int)((lane_id() & 16) ? r[0] : r[1]);
  } else {
    static_assert(M == 32, "exchange step");
    const auto r = __builtin_amdgcn_permlane32_swap((unsigned)v, (unsigned)v, false, false);
    return (int)((lane_id() & 32) ? r[0] : r[1]);
  }
}
template <int M>
OGJK_DEV float xchg(float v) {
  return __builtin_bit_cast(float, xchg_i32<M>(__builtin_bit_cast(int, v)));
}
template <int M>
OGJK_DEV int xchg(int v) {
  return xchg_i32<M>(v);
}
template <int M>
OGJK_DEV double xchg(double v) {
  const uint64_t u = __builtin_bit_cast(uint64_t, v);
  const int lo = xchg_i32<M>((int)(uint32_t)u), hi = xchg_i32<M>((int)(uint32_t)(u >> 32));
  return __builtin_bit_cast(double, ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}

// The two values a symmetric reduction step m combines, in either order:
// (own, partner).  Steps 16 and 32 use gfx950's v_permlane16_swap /
// v_permlane32_swap with both operands = v: the two results hold v's even and
// odd rows (lower and upper halves), i.e. each lane's own value and its
// xor-16 (xor-32) partner's, without an LDS round trip.
template <int M>
OGJK_DEV void xpair_i32(int v, int& a, int& b) {
  if constexpr (M == 16) {
    const auto r = __builtin_amdgcn_permlane16_swap((unsigned)v, (unsigned)v, false, false);
    a = (int)r[0];
    b = (int)r[1];
  } else if constexpr (M == 32) {
    const auto r = __builtin_amdgcn_permlane32_swap((unsigned)v, (unsigned)v, false, false);
    a = (int)r[0];
    b = (int)r[1];
  } else {
    a = v;
    b = xchg_i32<M>(v);
  }
}
template <int M>
OGJK_DEV void xpair(float v, float& a, float& b) {
  int x, y;
  xpair_i32<M>(__builtin_bit_cast(int, v), x, y);
  a = __builtin_bit_cast(float, x);
  b = __builtin_bit_cast(float, y);
}
template <int M>
OGJK_DEV void xpair(int v, int& a, int& b) {
  xpair_i32<M>(v, a, b);
}
template <int M>
OGJK_DEV void xpair(double v, double& a, double& b) {
  const uint64_t u = __builtin_bit_cast(uint64_t, v);
  int lo0, lo1, hi0, hi1;
  xpair_i32<M>((int)(uint32_t)u, lo0, lo1);
  xpair_i32<M>((int)(uint32_t)(u >> 32), hi0, hi1);
  a = __builtin_bit_cast(double, ((uint64_t)(uint32_t)hi0 << 32) | (uint32_t)lo0);
  b = __builtin_bit_cast(double, ((uint64_t)(uint32_t)hi1 << 32) | (uint32_t)lo1);
}

template <int LPP>
OGJK_DEV double shfl_d(double v, int src) {
  return __shfl(v, src, LPP);
}
template <int LPP>
OGJK_DEV int shfl_i(int v, int src) {
  return __shfl(v, src, LPP);
}
template <int LPP>
OGJK_DEV float shfl_f(float v, int src) {
  return __shfl(v, src, LPP);
}

template <int LPP, int M = 1>
OGJK_DEV float group_max_f(float v) {
  if constexpr (M < LPP) {
    float a, b;
    xpair<M>(v, a, b);
    v = fmaxf(a, b);
    return group_max_f<LPP, 2 * M>(v);
  } else {
    return v;
  }
}

// Max of (value, index) with the lowest index winning ties.  NaN values must
// already be mapped to -inf by the caller, so the result is group-uniform.
template <int LPP, int M = 1>
OGJK_DEV void group_argmax_d(double& v, int& i) {
  if constexpr (M < LPP) {
    double va, vb;
    int ia, ib;
    xpair<M>(v, va, vb);
    xpair<M>(i, ia, ib);
    const bool tb = vb > va || (vb == va && ib < ia);  // symmetric in (a, b)
    v = tb ? vb : va;
    i = tb ? ib : ia;
    group_argmax_d<LPP, 2 * M>(v, i);
  }
}

// Min of (value, index), lowest index on ties; NaN must be mapped to +inf.
template <int LPP, int M = 1>
OGJK_DEV void group_argmin_d(double& v, int& i) {
  if constexpr (M < LPP) {
    double va, vb;
    int ia, ib;
    xpair<M>(v, va, vb);
    xpair<M>(i, ia, ib);
    const bool tb = vb < va || (vb == va && ib < ia);  // symmetric in (a, b)
    v = tb ? vb : va;
    i = tb ? ib : ia;
    group_argmin_d<LPP, 2 * M>(v, i);
  }
}

}  // namespace ogjk
