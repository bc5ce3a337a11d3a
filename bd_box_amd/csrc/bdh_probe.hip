// In-bounds probes of the MFMA operand/result layouts assumed in bdh_mfma.hpp.
// One wavefront loads A (M x K, row-major) and B (K x N, row-major) with the assumed
// operand layout, issues a single MFMA on a zero accumulator and scatters the result with
// the assumed output layout into C (M x N, row-major); the host compares C with A.B.
#include "bdh_common.hpp"
#include "bdh_kernels.hpp"
#include "bdh_mfma.hpp"

namespace bdh {

__global__ void probe_f32_32x32x2(const float* A, const float* B, float* C)
{
    const int l = threadIdx.x;
    float a = A[(l % 32) * 2 + l / 32];
    float b = B[(l / 32) * 32 + l % 32];
    floatx16 acc;
    for (int r = 0; r < 16; ++r) acc[r] = 0.f;
    acc = mfma_f32_32x32x2(a, b, acc);
    for (int r = 0; r < 16; ++r) C[(8 * (r / 4) + 4 * (l / 32) + r % 4) * 32 + l % 32] = acc[r];
}

__global__ void probe_f32_16x16x4(const float* A, const float* B, float* C)
{
    const int l = threadIdx.x;
    float a = A[(l % 16) * 4 + l / 16];
    float b = B[(l / 16) * 16 + l % 16];
    floatx4 acc = {0.f, 0.f, 0.f, 0.f};
    acc = mfma_f32_16x16x4(a, b, acc);
    for (int r = 0; r < 4; ++r) C[(4 * (l / 16) + r) * 16 + l % 16] = acc[r];
}

__global__ void probe_f64_16x16x4(const double* A, const double* B, double* C)
{
    const int l = threadIdx.x;
    double a = A[(l % 16) * 4 + l / 16];
    double b = B[(l / 16) * 16 + l % 16];
    doublex4 acc = {0.0, 0.0, 0.0, 0.0};
    acc = mfma_f64_16x16x4(a, b, acc);
    for (int r = 0; r < 4; ++r) C[(4 * r + l / 16) * 16 + l % 16] = acc[r];
}

__global__ void probe_bf16_32x32x16(const float* A, const float* B, float* C)
{
    const int l = threadIdx.x;
    bf16x8 a, b;
    for (int q = 0; q < 8; ++q) {
        a[q] = (__bf16)A[(l % 32) * 16 + 8 * (l / 32) + q];
        b[q] = (__bf16)B[(8 * (l / 32) + q) * 32 + l % 32];
    }
    floatx16 acc;
    for (int r = 0; r < 16; ++r) acc[r] = 0.f;
    acc = mfma_bf16_32x32x16(a, b, acc);
    for (int r = 0; r < 16; ++r) C[(8 * (r / 4) + 4 * (l / 32) + r % 4) * 32 + l % 32] = acc[r];
}

__global__ void probe_bf16_16x16x32(const float* A, const float* B, float* C)
{
    const int l = threadIdx.x;
    bf16x8 a, b;
    for (int q = 0; q < 8; ++q) {
        a[q] = (__bf16)A[(l % 16) * 32 + 8 * (l / 16) + q];
        b[q] = (__bf16)B[(8 * (l / 16) + q) * 16 + l % 16];
    }
    floatx4 acc = {0.f, 0.f, 0.f, 0.f};
    acc = mfma_bf16_16x16x32(a, b, acc);
    for (int r = 0; r < 4; ++r) C[(4 * (l / 16) + r) * 16 + l % 16] = acc[r];
}

// kind: 0 f32 32x32x2, 1 f32 16x16x4, 2 f64 16x16x4, 3 bf16 32x32x16, 4 bf16 16x16x32.
// Device buffers sized for the kind (A: MxK, B: KxN, C: MxN).
int probe_mfma(int kind, const void* A, const void* B, void* C, hipStream_t st)
{
    switch (kind) {
    case 0: hipLaunchKernelGGL(probe_f32_32x32x2, dim3(1), dim3(64), 0, st, (const float*)A, (const float*)B, (float*)C); break;
    case 1: hipLaunchKernelGGL(probe_f32_16x16x4, dim3(1), dim3(64), 0, st, (const float*)A, (const float*)B, (float*)C); break;
    case 2: hipLaunchKernelGGL(probe_f64_16x16x4, dim3(1), dim3(64), 0, st, (const double*)A, (const double*)B, (double*)C); break;
    case 3: hipLaunchKernelGGL(probe_bf16_32x32x16, dim3(1), dim3(64), 0, st, (const float*)A, (const float*)B, (float*)C); break;
    case 4: hipLaunchKernelGGL(probe_bf16_16x16x32, dim3(1), dim3(64), 0, st, (const float*)A, (const float*)B, (float*)C); break;
    default: set_error("probe_mfma: unknown kind"); return -2;
    }
    BDH_LAUNCH_CHECK();
    return 0;
}

}  // namespace bdh
