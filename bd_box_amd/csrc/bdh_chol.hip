// Blocked right-looking Cholesky D = L L^T and the triangular product L.x on gfx950.
//
// Replaces the reference's my_dpotrf / my_dtrmv (src/myblas.c:45-118: LAPACK
// dpotrf("U") + dtrmv("U","T"), or MAGMA spotrf_gpu('L') + cublasStrmv in the CUDA
// build).  The factor is stored lower, column-major, in place of D; n is padded to a
// multiple of PAD_TO with an identity block so every tile is whole.
//
// Per panel of CHOL_NB = 64 columns:
//   panel kernel   one workgroup per 32 rows at and below the diagonal block: each loads
//                  the 64x64 diagonal block, factors it in FP64 in LDS (redundantly in every
//                  workgroup, 64^3/3 flops), then solves its own 32 rows X L_kk^T = B in
//                  FP64 and stores them in T.  No separate POTRF launch, no inverse.
//   trailing SYRK  A22 -= L21 L21^T over 128x128 lower tiles on MFMA:
//                  T = float : v_mfma_f32_32x32x2_f32, FP32 accumulation
//                  T = double: v_mfma_f64_16x16x4_f64
// FP64 is used for the diagonal-block factorization and the panel solve in both
// precisions; in the FP32 variant the factor is rounded to FP32 when stored.
#include "bdh_common.hpp"
#include "bdh_kernels.hpp"
#include "bdh_mfma.hpp"

namespace bdh {

constexpr int PANEL_ROWS = 32;

template <typename T>
__global__ void __launch_bounds__(256)
chol_panel_kernel(T* __restrict__ A, long long ld, int c0, int* __restrict__ status)
{
    __shared__ double sL[CHOL_NB][CHOL_NB + 1];
    __shared__ double sX[PANEL_ROWS][CHOL_NB + 1];
    const int tid = threadIdx.x;
    const int rb = blockIdx.x;
    const int r0 = c0 + rb * PANEL_ROWS;
    const T* Ad = A + (long long)c0 * ld + c0;

    for (int idx = tid; idx < CHOL_NB * CHOL_NB; idx += 256) {
        int r = idx % CHOL_NB, c = idx / CHOL_NB;
        sL[r][c] = (r >= c) ? (double)Ad[(long long)c * ld + r] : 0.0;
    }
    const bool solve = rb >= CHOL_NB / PANEL_ROWS;
    if (solve) {
        const T* Ax = A + (long long)c0 * ld + r0;
        for (int idx = tid; idx < PANEL_ROWS * CHOL_NB; idx += 256) {
            int r = idx % PANEL_ROWS, c = idx / PANEL_ROWS;
            sX[r][c] = (double)Ax[(long long)c * ld + r];
        }
    }
    __syncthreads();
    // right-looking factorization of the diagonal block
    for (int j = 0; j < CHOL_NB; ++j) {
        double d = sL[j][j];
        if (!(d > 0.0)) {
            if (tid == 0) atomicCAS(status, 0, c0 + j + 1);
            d = __builtin_nan("");
        }
        const double s = sqrt(d);
        const double inv = 1.0 / s;
        __syncthreads();
        if (tid < CHOL_NB) {
            if (tid > j) sL[tid][j] *= inv;
            else if (tid == j) sL[j][j] = s;
        }
        __syncthreads();
        const int r = tid & (CHOL_NB - 1);
        const double lrj = sL[r][j];
        for (int c = j + 1 + (tid >> 6); c <= r; c += 4) sL[r][c] -= lrj * sL[c][j];
        __syncthreads();
    }
    if (solve) {
        // X L^T = B, column by column
        for (int c = 0; c < CHOL_NB; ++c) {
            if (tid < PANEL_ROWS) sX[tid][c] /= sL[c][c];
            __syncthreads();
            const int r = tid & (PANEL_ROWS - 1);
            const double xc = sX[r][c];
            for (int c2 = c + 1 + (tid >> 5); c2 < CHOL_NB; c2 += 8) sX[r][c2] -= xc * sL[c2][c];
            __syncthreads();
        }
        T* Ax = A + (long long)c0 * ld + r0;
        for (int idx = tid; idx < PANEL_ROWS * CHOL_NB; idx += 256) {
            int r = idx % PANEL_ROWS, c = idx / PANEL_ROWS;
            Ax[(long long)c * ld + r] = (T)sX[r][c];
        }
    } else {
        // rows of the diagonal block itself: store the lower part of L_kk
        T* Aw = A + (long long)c0 * ld + r0;
        const int rbase = rb * PANEL_ROWS;
        for (int idx = tid; idx < PANEL_ROWS * CHOL_NB; idx += 256) {
            int r = idx % PANEL_ROWS, c = idx / PANEL_ROWS;
            if (rbase + r >= c) Aw[(long long)c * ld + r] = (T)sL[rbase + r][c];
        }
    }
}

// Lower-triangular tile index (ti >= tj) from a linear index x.
__device__ __forceinline__ void tri_tile(int x, int& ti, int& tj)
{
    int t = (int)((sqrt(8.0 * (double)x + 1.0) - 1.0) * 0.5);
    while ((t + 1) * (t + 2) / 2 <= x) ++t;
    while (t * (t + 1) / 2 > x) --t;
    ti = t;
    tj = x - t * (t + 1) / 2;
}

// ------------------------------------------------------------------ FP32 SYRK (MFMA)
// Workgroup 256 = 4 waves in 2x2; tile 128x128; each wave 64x64 = 2x2 blocks of 32x32.
constexpr int SY_TILE = 128;
constexpr int SY_KC32 = 32;

__global__ void __launch_bounds__(256)
syrk_f32_kernel(float* __restrict__ A, long long ld, int c0, int t0, int n_pad)
{
    __shared__ float sA[SY_KC32][SY_TILE + 4];
    __shared__ float sB[SY_KC32][SY_TILE + 4];
    int ti, tj;
    tri_tile(blockIdx.x, ti, tj);
    const int ri0 = t0 + ti * SY_TILE;
    const int rj0 = t0 + tj * SY_TILE;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wr = wave >> 1, wc = wave & 1;
    const int l32 = lane & 31, lh = lane >> 5;

    floatx16 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[a][b][r] = 0.f;

    for (int k0 = 0; k0 < CHOL_NB; k0 += SY_KC32) {
        __syncthreads();
        // 32 columns x 128 rows, float4 per thread-iteration
        for (int idx = tid; idx < SY_KC32 * SY_TILE / 4; idx += 256) {
            const int k = idx / (SY_TILE / 4);
            const int r4 = (idx % (SY_TILE / 4)) * 4;
            const float* col = A + (long long)(c0 + k0 + k) * ld;
            float4 va = make_float4(0.f, 0.f, 0.f, 0.f), vb = va;
            if (ri0 + r4 < n_pad) va = *reinterpret_cast<const float4*>(col + ri0 + r4);
            if (rj0 + r4 < n_pad) vb = *reinterpret_cast<const float4*>(col + rj0 + r4);
            sA[k][r4 + 0] = va.x; sA[k][r4 + 1] = va.y; sA[k][r4 + 2] = va.z; sA[k][r4 + 3] = va.w;
            sB[k][r4 + 0] = vb.x; sB[k][r4 + 1] = vb.y; sB[k][r4 + 2] = vb.z; sB[k][r4 + 3] = vb.w;
        }
        __syncthreads();
#pragma unroll 4
        for (int k = 0; k < SY_KC32; k += 2) {
            float av[2], bv[2];
#pragma unroll
            for (int a = 0; a < 2; ++a) av[a] = sA[k + lh][wr * 64 + a * 32 + l32];
#pragma unroll
            for (int b = 0; b < 2; ++b) bv[b] = sB[k + lh][wc * 64 + b * 32 + l32];
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b) acc[a][b] = mfma_f32_32x32x2(av[a], bv[b], acc[a][b]);
        }
    }
    // C -= acc ; lane owns column l32 and rows 8*(r/4) + 4*lh + (0..3) of each 32x32 block
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            const int col = rj0 + wc * 64 + b * 32 + l32;
            if (col >= n_pad) continue;
            float* cp = A + (long long)col * ld;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int row = ri0 + wr * 64 + a * 32 + 8 * q + 4 * lh;
                if (row >= n_pad) continue;
                float4 c = *reinterpret_cast<float4*>(cp + row);
                c.x -= acc[a][b][4 * q + 0];
                c.y -= acc[a][b][4 * q + 1];
                c.z -= acc[a][b][4 * q + 2];
                c.w -= acc[a][b][4 * q + 3];
                *reinterpret_cast<float4*>(cp + row) = c;
            }
        }
}

// ------------------------------------------------------------------ FP64 SYRK (MFMA)
// Same tiling; each wave 64x64 = 4x4 blocks of 16x16 (v_mfma_f64_16x16x4_f64).  The f64
// MFMA returns row 4*r + l/16 in accumulator r of lane l; feeding the A operand rows in the
// order perm(m) = 4*(m%4) + m/4 makes accumulator r hold row 4*(l/16) + r instead, so each
// lane owns 4 consecutive rows of one column (two 16-byte stores).
constexpr int SY_KC64 = 16;

__global__ void __launch_bounds__(256)
syrk_f64_kernel(double* __restrict__ A, long long ld, int c0, int t0, int n_pad)
{
    __shared__ double sA[SY_KC64][SY_TILE + 2];
    __shared__ double sB[SY_KC64][SY_TILE + 2];
    int ti, tj;
    tri_tile(blockIdx.x, ti, tj);
    const int ri0 = t0 + ti * SY_TILE;
    const int rj0 = t0 + tj * SY_TILE;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wr = wave >> 1, wc = wave & 1;
    const int l16 = lane & 15, lq = lane >> 4;
    const int prow = 4 * (l16 & 3) + (l16 >> 2);

    doublex4 acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) acc[a][b][r] = 0.0;

    for (int k0 = 0; k0 < CHOL_NB; k0 += SY_KC64) {
        __syncthreads();
        for (int idx = tid; idx < SY_KC64 * SY_TILE / 2; idx += 256) {
            const int k = idx / (SY_TILE / 2);
            const int r2 = (idx % (SY_TILE / 2)) * 2;
            const double* col = A + (long long)(c0 + k0 + k) * ld;
            double2 va = make_double2(0.0, 0.0), vb = va;
            if (ri0 + r2 < n_pad) va = *reinterpret_cast<const double2*>(col + ri0 + r2);
            if (rj0 + r2 < n_pad) vb = *reinterpret_cast<const double2*>(col + rj0 + r2);
            sA[k][r2] = va.x; sA[k][r2 + 1] = va.y;
            sB[k][r2] = vb.x; sB[k][r2 + 1] = vb.y;
        }
        __syncthreads();
#pragma unroll 2
        for (int k = 0; k < SY_KC64; k += 4) {
            double av[4], bv[4];
#pragma unroll
            for (int a = 0; a < 4; ++a) av[a] = sA[k + lq][wr * 64 + a * 16 + prow];
#pragma unroll
            for (int b = 0; b < 4; ++b) bv[b] = sB[k + lq][wc * 64 + b * 16 + l16];
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int b = 0; b < 4; ++b) acc[a][b] = mfma_f64_16x16x4(av[a], bv[b], acc[a][b]);
        }
    }
    // lane owns column l16 and rows 4*lq + (0..3) of each 16x16 block
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int col = rj0 + wc * 64 + b * 16 + l16;
            const int row = ri0 + wr * 64 + a * 16 + 4 * lq;
            if (col >= n_pad || row >= n_pad) continue;
            double* cp = A + (long long)col * ld + row;
            double2 c0v = *reinterpret_cast<double2*>(cp);
            double2 c1v = *reinterpret_cast<double2*>(cp + 2);
            c0v.x -= acc[a][b][0]; c0v.y -= acc[a][b][1];
            c1v.x -= acc[a][b][2]; c1v.y -= acc[a][b][3];
            *reinterpret_cast<double2*>(cp) = c0v;
            *reinterpret_cast<double2*>(cp + 2) = c1v;
        }
}

template <typename T>
int chol_factor(T* d_A, long long ld, int n_pad, int* d_status, int syrk_mode, hipStream_t st)
{
    BDH_CHECK(n_pad % PAD_TO == 0, "chol_factor: n_pad must be a multiple of 128");
    BDH_CHECK(ld >= n_pad && ld % 4 == 0, "chol_factor: bad leading dimension");
    for (int c0 = 0; c0 < n_pad; c0 += CHOL_NB) {
        const int rows = n_pad - c0;
        hipLaunchKernelGGL(chol_panel_kernel<T>, dim3(rows / PANEL_ROWS), dim3(256), 0, st,
                           d_A, ld, c0, d_status);
        BDH_LAUNCH_CHECK();
        const int t0 = c0 + CHOL_NB;
        if (t0 >= n_pad) break;
        const int nt = ceil_div(n_pad - t0, SY_TILE);
        const int ntiles = nt * (nt + 1) / 2;
        if constexpr (sizeof(T) == 4) {
            (void)syrk_mode;
            hipLaunchKernelGGL(syrk_f32_kernel, dim3(ntiles), dim3(256), 0, st,
                               reinterpret_cast<float*>(d_A), ld, c0, t0, n_pad);
        } else {
            hipLaunchKernelGGL(syrk_f64_kernel, dim3(ntiles), dim3(256), 0, st,
                               reinterpret_cast<double*>(d_A), ld, c0, t0, n_pad);
        }
        BDH_LAUNCH_CHECK();
    }
    return 0;
}

// ------------------------------------------------------------------ y = L x (partials)
// Tile: (64 lanes x V rows) x TRMV_COLS columns, V = 16 / sizeof(T) so each lane issues one
// 16-byte load per column; the 4 waves split the tile's columns and meet in LDS.
// P[cb][r] receives row r's partial over column block cb (FP64 accumulation).
constexpr int TRMV_COLS = 256;

template <typename T>
__global__ void __launch_bounds__(256)
trmv_partial_kernel(const T* __restrict__ L, long long ld, int n, const double* __restrict__ x,
                    double* __restrict__ P)
{
    constexpr int V = 16 / sizeof(T);
    constexpr int ROWS = 64 * V;
    const int rb = blockIdx.x, cb = blockIdx.y;
    const int r0 = rb * ROWS, cbeg = cb * TRMV_COLS;
    if (cbeg > r0 + ROWS - 1 || cbeg >= n) return;  // tile entirely above the diagonal
    __shared__ double s_red[4][ROWS];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int rbase = r0 + lane * V;
    double acc[V];
#pragma unroll
    for (int q = 0; q < V; ++q) acc[q] = 0.0;
    const int cend = min(n, cbeg + TRMV_COLS);
    const int wcols = TRMV_COLS / 4;
    const int c_lo = cbeg + wave * wcols, c_hi = min(cend, c_lo + wcols);
    if (rbase < n) {
        const bool full = (c_hi - 1) <= rbase;  // every column of this wave is left of the lane's rows
        if (full) {
#pragma unroll 4
            for (int c = c_lo; c < c_hi; ++c) {
                const double xc = x[c];
                if constexpr (V == 4) {
                    float4 v = *reinterpret_cast<const float4*>(L + (long long)c * ld + rbase);
                    acc[0] += (double)v.x * xc; acc[1] += (double)v.y * xc;
                    acc[2] += (double)v.z * xc; acc[3] += (double)v.w * xc;
                } else {
                    double2 v = *reinterpret_cast<const double2*>(L + (long long)c * ld + rbase);
                    acc[0] += v.x * xc; acc[1] += v.y * xc;
                }
            }
        } else {
            for (int c = c_lo; c < c_hi && c <= rbase + V - 1; ++c) {
                const double xc = x[c];
#pragma unroll
                for (int q = 0; q < V; ++q)
                    if (rbase + q >= c) acc[q] += (double)L[(long long)c * ld + rbase + q] * xc;
            }
        }
    }
#pragma unroll
    for (int q = 0; q < V; ++q) s_red[wave][lane * V + q] = acc[q];
    __syncthreads();
    for (int r = tid; r < ROWS; r += 256) {
        const int row = r0 + r;
        if (row < n)
            P[(long long)cb * n + row] = s_red[0][r] + s_red[1][r] + s_red[2][r] + s_red[3][r];
    }
}

int trmv_col_blocks(int n) { return ceil_div(n, TRMV_COLS); }

template <typename T>
int launch_trmv_partials(const T* d_L, long long ld, int n, const double* d_x, double* d_P,
                         hipStream_t st)
{
    constexpr int ROWS = 64 * (16 / sizeof(T));
    BDH_CHECK(ld % (16 / sizeof(T)) == 0, "trmv: leading dimension not 16-byte aligned");
    dim3 grid(ceil_div(n, ROWS), trmv_col_blocks(n));
    hipLaunchKernelGGL(trmv_partial_kernel<T>, grid, dim3(256), 0, st, d_L, ld, n, d_x, d_P);
    BDH_LAUNCH_CHECK();
    return 0;
}

// y[r] = scale * sum_{cb <= r / blk_cols} P[cb][r]
__global__ void sum_partials_kernel(const double* __restrict__ P, int n, int nblk, int blk_cols,
                                    double scale, double* __restrict__ y)
{
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    int last = blk_cols > 0 ? min(nblk - 1, r / blk_cols) : nblk - 1;
    double s = 0.0;
    for (int cb = 0; cb <= last; ++cb) s += P[(long long)cb * n + r];
    y[r] = scale * s;
}

int launch_sum_partials(const double* d_P, int n, int nblk, int blk_cols, double scale,
                        double* d_y, hipStream_t st)
{
    hipLaunchKernelGGL(sum_partials_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, st, d_P, n,
                       nblk, blk_cols, scale, d_y);
    BDH_LAUNCH_CHECK();
    return 0;
}

template int chol_factor<float>(float*, long long, int, int*, int, hipStream_t);
template int chol_factor<double>(double*, long long, int, int*, int, hipStream_t);
template int launch_trmv_partials<float>(const float*, long long, int, const double*, double*, hipStream_t);
template int launch_trmv_partials<double>(const double*, long long, int, const double*, double*, hipStream_t);

}  // namespace bdh
