// mlp.h — MFMA building blocks of the learned light proposal (gfx950, bf16 in, fp32 accumulate).
//
// v_mfma_f32_32x32x16_bf16 multiplies a 32x16 tile A by a 16x32 tile B into a 32x32 fp32 tile.
// Operand layout for wave64 lane l (checked on the device by tests/test_gpu_mlp.py):
//   A: 8 bf16 = A[row = l % 32][k = 8 * (l / 32) + j], j = 0..7
//   B: 8 bf16 = B[k = 8 * (l / 32) + j][col = l % 32]
//   D: 16 fp32, acc[i] = D[row = 8 * (i / 4) + 4 * (l / 32) + i % 4][col = l % 32]
#pragma once

#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rm {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kIn = 128;      // feature width (114 used, zero padded)
constexpr int kHidden = 128;  // hidden width
constexpr int kFeatures = 114;
// cluster counts: 64, 128 or 256 (k_propose holds a quad's numerators in LDS), or a multiple of
// kSuper up to kMaxClusters (k_propose_wide streams them in column blocks of kSuper), with at most
// kWideCand candidates per pixel
constexpr int kSuper = 128;
constexpr int kMaxClusters = 4096;
constexpr int kWideCand = 8;

__device__ __forceinline__ f32x16 mfma32(bf16x8 a, bf16x8 b, f32x16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}

// bf16 <-> fp32 with round-to-nearest-even (NaN kept quiet)
// float -> bf16, round to nearest even: one v_cvt_pk_bf16_f32 on gfx950
__device__ __forceinline__ uint16_t f2bf(float x) { return __builtin_bit_cast(uint16_t, static_cast<__bf16>(x)); }
// two floats -> packed bf16 pair (a in the low half), one instruction
__device__ __forceinline__ uint32_t f2bf2(float a, float b) {
    typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
    typedef float f32x2_t __attribute__((ext_vector_type(2)));
    const f32x2_t v = {a, b};
    return __builtin_bit_cast(uint32_t, __builtin_convertvector(v, bf16x2_t));
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float(static_cast<uint32_t>(h) << 16); }

// Offsets (elements) of W1..W4 ([out][in]) and b1..b4 in the flat parameter arrays.
__host__ __device__ constexpr size_t w_off(int layer) { return static_cast<size_t>(layer) * kHidden * kIn; }
__host__ __device__ constexpr size_t b_off(int layer) { return static_cast<size_t>(layer) * kHidden; }
__host__ __device__ constexpr size_t n_weights(int K) { return 3 * static_cast<size_t>(kHidden) * kIn + static_cast<size_t>(K) * kHidden; }
__host__ __device__ constexpr size_t n_bias(int K) { return 3 * static_cast<size_t>(kHidden) + K; }

}  // namespace rm
