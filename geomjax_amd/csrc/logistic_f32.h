// Float32 building blocks shared by the tuned logistic-regression kernels:
// fragment loads, the bf16 hi/lo split and the three-product MFMA that
// carries ~16 significant bits per term with float32 accumulation, fast
// float32 transcendentals, MFMA/VALU interleaving hints, and the feature
// permutation that aligns V^T = G^-1 X^T accumulator slots with the row layout.
#pragma once

#include <hip/hip_runtime.h>

namespace gjx {
namespace lf32 {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef float f32x32 __attribute__((ext_vector_type(32)));

__device__ __forceinline__ void load8(const float* p, float (&v)[8]) {
  const float4 a = reinterpret_cast<const float4*>(p)[0];
  const float4 b = reinterpret_cast<const float4*>(p)[1];
  v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
  v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
}

typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));

// v = hi + lo + O(2^-17 |v|), hi and lo in bf16 (round to nearest even).
// Two values per conversion (v_cvt_pk_bf16_f32); hi is widened back to
// float32 by placing its bits in the upper half of a word.
__device__ __forceinline__ void split8(const float (&v)[8], bf16x8& hi, bf16x8& lo) {
#pragma unroll
  for (int i = 0; i < 8; i += 2) {
    const bf16x2 h = __builtin_convertvector((f32x2{v[i], v[i + 1]}), bf16x2);
    const uint32_t hb = __builtin_bit_cast(uint32_t, h);
    const float h0 = __builtin_bit_cast(float, hb << 16);
    const float h1 = __builtin_bit_cast(float, hb & 0xffff0000u);
    const bf16x2 r = __builtin_convertvector((f32x2{v[i] - h0, v[i + 1] - h1}), bf16x2);
    hi[i] = h[0];
    hi[i + 1] = h[1];
    lo[i] = r[0];
    lo[i + 1] = r[1];
  }
}

// acc += (ah + al)(bh + bl) without the lo*lo term
__device__ __forceinline__ f32x16 mfma3(const bf16x8& ah, const bf16x8& al, const bf16x8& bh, const bf16x8& bl,
                                        f32x16 acc) {
  acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, bh, acc, 0, 0, 0);
  acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bl, acc, 0, 0, 0);
  acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bh, acc, 0, 0, 0);
  return acc;
}

// acc += (ah + al)(bh + bl) with all four products: the lo*lo term is about
// 2^-16 of each product and, for the diagonal of a Gram, always of one sign
__device__ __forceinline__ f32x16 mfma4(const bf16x8& ah, const bf16x8& al, const bf16x8& bh, const bf16x8& bl,
                                        f32x16 acc) {
  acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, bl, acc, 0, 0, 0);
  return mfma3(ah, al, bh, bl, acc);
}

// sigmoid with the hardware exp2 and reciprocal (a few ulp: the state is float32)
__device__ __forceinline__ float sigmoidf(float x) { return __builtin_amdgcn_rcpf(1.0f + __expf(-x)); }
__device__ __forceinline__ float sqrt_fast(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ double softplus(double x) { return fmax(x, 0.0) + log1p(exp(-fabs(x))); }

// Ask the scheduler to issue the region as NM groups of (1 MFMA, NV VALU), so
// that independent VALU work runs while each MFMA executes.
template <int NM, int NV>
__device__ __forceinline__ void interleave_mfma_valu() {
#pragma unroll
  for (int i = 0; i < NM; ++i) {
    __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
    __builtin_amdgcn_sched_group_barrier(0x002, NV, 0);
  }
}

// phi: accumulator feature slot m = 32t + 8a + 4h + c  ->  feature 16(j/8) + 8h + j%8, j = 16t + 4a + c
__device__ __forceinline__ int phi(int m) {
  const int t = m >> 5, a = (m >> 3) & 3, h = (m >> 2) & 1, c = m & 3;
  const int j = 16 * t + 4 * a + c;
  return (j >> 3) * 16 + h * 8 + (j & 7);
}

}  // namespace lf32
}  // namespace gjx
