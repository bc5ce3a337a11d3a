// MFMA operand/result layouts used by the gfx950 kernels of this backend.
//
// Every layout below is checked against a host GEMM by the probe entry point
// bdh_probe_mfma (tests/test_gpu_kernels.py) before a kernel relies on it.
//
//  f32_32x32x2 : A 32x2, B 2x32, D 32x32 (16 floats/lane)
//     A: lane l holds A[l%32][l/32]     B: lane l holds B[l/32][l%32]
//     D: acc r of lane l is D[8*(r/4) + 4*(l/32) + r%4][l%32]
//  f32_16x16x4 : A 16x4, B 4x16, D 16x16 (4 floats/lane)
//     A: A[l%16][l/16]   B: B[l/16][l%16]   D: acc r -> D[4*(l/16) + r][l%16]
//  f64_16x16x4 : A, B as f32_16x16x4 (doubles); D: acc r -> D[4*r + l/16][l%16]
//     (rows interleaved, unlike the f32 form; measured with the probe)
//  bf16 32x32x16 (gfx950): A 32x16, B 16x32, 8 bf16 per lane
//     A: lane l holds A[l%32][8*(l/32) + 0..7]   B: B[8*(l/32) + 0..7][l%32]
//     D: as f32_32x32x2
//  bf16 16x16x32 (gfx950): A: A[l%16][8*(l/16) + 0..7]  B: B[8*(l/16)+0..7][l%16]
//     D: as f32_16x16x4
#pragma once

#include <hip/hip_runtime.h>

namespace bdh {

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef double doublex4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ floatx16 mfma_f32_32x32x2(float a, float b, floatx16 c)
{
    return __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ floatx4 mfma_f32_16x16x4(float a, float b, floatx4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ doublex4 mfma_f64_16x16x4(double a, double b, doublex4 c)
{
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ floatx16 mfma_bf16_32x32x16(bf16x8 a, bf16x8 b, floatx16 c)
{
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ floatx4 mfma_bf16_16x16x32(bf16x8 a, bf16x8 b, floatx4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

}  // namespace bdh
