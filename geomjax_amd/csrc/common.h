// Device helpers shared by the integrator kernels (gfx950, wave64).
//
// Chains are mapped to aligned groups of LPC lanes ("lanes per chain",
// a power of two <= 64).  Every per-chain vector is spread over its group in
// a blocked layout: lane l of the group holds elements [l*EPL, (l+1)*EPL).
// Group-wide sums and maxima use DPP inside a 16-lane row and
// ds_bpermute-based shuffles across rows; every lane of the group ends with a
// bit-identical result (each butterfly step adds a pair in both orders, and
// floating-point addition is commutative), so branch decisions taken on a
// reduced value are uniform across the group.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gjx {

// ---------------------------------------------------------------------------
// DPP moves.  Control codes (GCN/CDNA DPP encoding):
//   quad_perm [1,0,3,2] = 0xB1 (xor 1), quad_perm [2,3,0,1] = 0x4E (xor 2),
//   row_half_mirror = 0x141 (lane i <-> 7-i in each half row),
//   row_mirror = 0x140 (lane i <-> 15-i in each row).
// After the two quad steps all lanes of a quad hold the quad total, so the
// mirror steps pair whole quads / half rows, as an xor butterfly would.
// ---------------------------------------------------------------------------
template <int CTRL>
__device__ __forceinline__ float dpp_f32(float v) {
  return __builtin_bit_cast(
      float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, 0xF, 0xF, false));
}

template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
  uint64_t b = __builtin_bit_cast(uint64_t, v);
  int lo = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)b, CTRL, 0xF, 0xF, false);
  int hi = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)(b >> 32), CTRL, 0xF, 0xF, false);
  return __builtin_bit_cast(double, ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}

template <int CTRL>
__device__ __forceinline__ float dpp(float v) { return dpp_f32<CTRL>(v); }
template <int CTRL>
__device__ __forceinline__ double dpp(double v) { return dpp_f64<CTRL>(v); }

struct OpSum {
  template <class T> __device__ __forceinline__ static T f(T a, T b) { return a + b; }
};
struct OpMax {
  template <class T> __device__ __forceinline__ static T f(T a, T b) { return a > b ? a : b; }
};

// Reduce over an aligned group of W lanes (W power of two, 1..64).
template <int W, class Op, class T>
__device__ __forceinline__ T group_reduce(T v) {
  if constexpr (W >= 2) v = Op::f(v, dpp<0xB1>(v));
  if constexpr (W >= 4) v = Op::f(v, dpp<0x4E>(v));
  if constexpr (W >= 8) v = Op::f(v, dpp<0x141>(v));
  if constexpr (W >= 16) v = Op::f(v, dpp<0x140>(v));
  if constexpr (W >= 32) v = Op::f(v, (T)__shfl_xor(v, 16, 64));
  if constexpr (W >= 64) v = Op::f(v, (T)__shfl_xor(v, 32, 64));
  return v;
}

template <int W, class T>
__device__ __forceinline__ T group_sum(T v) { return group_reduce<W, OpSum>(v); }
template <int W, class T>
__device__ __forceinline__ T group_max(T v) { return group_reduce<W, OpMax>(v); }

// Exclusive prefix sum over the W lanes of a group (Hillis-Steele with
// shuffles; used once per transition, not in the inner loop).
template <int W, class T>
__device__ __forceinline__ T group_exclusive_scan(T v, int lane_in_group) {
  T inc = v;
#pragma unroll
  for (int off = 1; off < W; off <<= 1) {
    T up = (T)__shfl_up(inc, off, W);
    if (lane_in_group >= off) inc += up;
  }
  return inc - v;
}

// Number of integration steps of chain c in transition tr: the per-chain
// value of the dynamic kernel ([T, C], clamped to [0, L]), or L for all.
__device__ __forceinline__ int chain_steps(const int32_t* Lc, int L, int C, int tr, int c) {
  if (Lc == nullptr) return L;
  const int v = Lc[(size_t)tr * C + c];
  return v < 0 ? 0 : (v > L ? L : v);
}

// ---------------------------------------------------------------------------
// Counter-based PRNG: Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel
// random numbers: as easy as 1, 2, 3", SC'11), written from the published
// round function: two 32x32->64 multiplies per round, Weyl key schedule.
// ---------------------------------------------------------------------------
struct u32x4 { uint32_t x, y, z, w; };

__device__ __forceinline__ u32x4 philox4x32_10(u32x4 ctr, uint32_t k0, uint32_t k1) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
  const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    uint64_t p0 = (uint64_t)M0 * ctr.x;
    uint64_t p1 = (uint64_t)M1 * ctr.z;
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    ctr = u32x4{hi1 ^ ctr.y ^ k0, lo1, hi0 ^ ctr.w ^ k1, lo0};
    k0 += W0;
    k1 += W1;
  }
  return ctr;
}

// Uniform in (0, 1): never 0 so that log() below is finite.
__device__ __forceinline__ double u01_f64(uint32_t a, uint32_t b) {
  uint64_t m = ((uint64_t)a << 21) ^ (uint64_t)(b >> 11);  // 53 random bits
  m &= ((1ull << 53) - 1);
  return ((double)m + 0.5) * (1.0 / 9007199254740992.0);
}
__device__ __forceinline__ float u01_f32(uint32_t a) {
  return ((float)(a >> 8) + 0.5f) * (1.0f / 16777216.0f);
}

// Two independent standard normals by the Box-Muller transform.
__device__ __forceinline__ void normal2(u32x4 r, double& n0, double& n1) {
  double u1 = u01_f64(r.x, r.y), u2 = u01_f64(r.z, r.w);
  double rad = sqrt(-2.0 * log(u1));
  double s, c;
  sincospi(2.0 * u2, &s, &c);
  n0 = rad * c;
  n1 = rad * s;
}
__device__ __forceinline__ void normal2(u32x4 r, float& n0, float& n1) {
  // float path: 24-bit uniforms, math in float
  float u1 = u01_f32(r.x), u2 = u01_f32(r.y);
  float rad = sqrtf(-2.0f * logf(u1));
  float s, c;
  sincospif(2.0f * u2, &s, &c);
  n0 = rad * c;
  n1 = rad * s;
}

// Purpose tags keep the streams of different draws disjoint.
enum : uint32_t {
  kStreamMomentum = 0x10000000u,   // + element index
  kStreamMomentumV = 0x20000000u,  // replicated last coordinate
  kStreamAccept = 0x30000000u,
  kStreamNutsDirection = 0x40000000u,  // + expansion index
  kStreamNutsMerge = 0x50000000u,      // + expansion index
  kStreamNutsStep = 0x60000000u        // + integration step index within the transition
};

template <class T> struct is_f64 { static constexpr bool value = false; };
template <> struct is_f64<double> { static constexpr bool value = true; };

constexpr double kLog2Pi = 1.8378770664093454835606594728112;

// max that propagates NaN (jnp.max does; a plain a > b ? a : b would drop it)
template <class R>
__device__ __forceinline__ R nan_max(R a, R b) {
  return (a > b || a != a) ? a : b;
}

struct NanMax {
  template <class T> __device__ __forceinline__ static T f(T a, T b) { return nan_max(a, b); }
};

template <class T>
__device__ __forceinline__ T ld(const void* p, size_t i) { return static_cast<const T*>(p)[i]; }
template <class T>
__device__ __forceinline__ void st(void* p, size_t i, T v) { static_cast<T*>(p)[i] = v; }

}  // namespace gjx
