// Blocked right-looking Cholesky D = L L^T and the triangular product L.x on gfx950.
//
// Replaces the reference's my_dpotrf / my_dtrmv (src/myblas.c:45-118: LAPACK
// dpotrf("U") + dtrmv("U","T"), or MAGMA spotrf_gpu('L') + cublasStrmv in the CUDA
// build).  The factor is stored lower, column-major, in place of D; n is padded to a
// multiple of PAD_TO with an identity block so every tile is whole.
//
// Per panel of CHOL_NB = 64 columns:
//   panel kernel   one workgroup per 256 rows below the diagonal block (plus one that stores
//                  L_kk): each factors the 64x64 diagonal block in FP64 in registers of one
//                  wavefront (redundantly, 64^3/3 flops), then every thread solves one row
//                  x L_kk^T = b in FP64 and stores it in T.  No separate POTRF launch, no
//                  inverse.
//   trailing SYRK  A22 -= L21 L21^T over 128x128 lower tiles on MFMA:
//                  T = float : v_mfma_f32_32x32x2_f32, FP32 accumulation
//                  T = double: v_mfma_f64_16x16x4_f64
// FP64 is used for the diagonal-block factorization and the panel solve in both
// precisions; in the FP32 variant the factor is rounded to FP32 when stored.
#include "bdh_common.hpp"
#include "bdh_kernels.hpp"
#include "bdh_mfma.hpp"

namespace bdh {

constexpr int PANEL_ROWS = 256;  // rows solved per workgroup, one per thread

// Value of v held by lane `src` (uniform), through two v_readlane_b32.
__device__ __forceinline__ double readlane_f64(double v, int src)
{
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)u, src);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u >> 32), src);
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// Panel kernel.  Workgroup 0 stores L_kk; workgroup b >= 1 solves rows
// c0 + 64 + 256 (b - 1) + tid.  In every workgroup, wave 0 first factors the 64x64
// diagonal block with each lane holding one row in registers (FP64): for column j the pivot
// and the entries of the scaled column come from the lanes holding them (v_readlane: scalar
// operands of the FMAs), and every lane applies the rank-1 update to its own row -- no LDS
// traffic and no barriers inside the 64 steps.  After one
// barrier each thread loads and solves its row
// x L_kk^T = b column by column (right-looking: x_c is final, then x_k -= x_c L[k][c] for
// k > c) against L_kk stored transposed in LDS, so every read is a broadcast.
template <typename T>
__global__ void __launch_bounds__(256)
chol_panel_kernel(T* __restrict__ A, long long ld, int c0, int n_pad, int* __restrict__ status)
{
    __shared__ __align__(16) double sLT[CHOL_NB][CHOL_NB + 2];  // sLT[c][k] = L_kk[k][c]
    __shared__ double s_inv[CHOL_NB];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int b = blockIdx.x;
    const int r = c0 + CHOL_NB + (b - 1) * PANEL_ROWS + tid;  // row solved by this thread
    const bool solves = b > 0 && r < n_pad;
    if (wave == 0) {
        const T* Ad = A + (long long)c0 * ld + c0 + lane;
        double a[CHOL_NB];
#pragma unroll
        for (int c = 0; c < CHOL_NB; ++c) a[c] = (c <= lane) ? (double)Ad[(long long)c * ld] : 0.0;
#pragma unroll
        for (int j = 0; j < CHOL_NB; ++j) {
            double d = readlane_f64(a[j], j);
            if (!(d > 0.0)) {
                if (lane == 0) atomicCAS(status, 0, c0 + j + 1);
                d = __builtin_nan("");
            }
            const double sd = sqrt(d);
            const double inv = 1.0 / sd;
            a[j] = (lane == j) ? sd : a[j] * inv;
#pragma unroll
            for (int k = j + 1; k < CHOL_NB; ++k) a[k] -= a[j] * readlane_f64(a[j], k);
        }
#pragma unroll
        for (int c = 0; c < CHOL_NB; ++c) sLT[c][lane] = (c <= lane) ? a[c] : 0.0;
        __builtin_amdgcn_wave_barrier();
        // (a[lane] would index the register array at run time and push it to scratch)
        s_inv[lane] = 1.0 / sLT[lane][lane];
    }
    __syncthreads();
    if (b == 0) {
        // store the lower part of L_kk
        T* Aw = A + (long long)c0 * ld + c0;
        for (int idx = tid; idx < CHOL_NB * CHOL_NB; idx += 256) {
            const int rr = idx % CHOL_NB, cc = idx / CHOL_NB;
            if (rr >= cc) Aw[(long long)cc * ld + rr] = (T)sLT[cc][rr];
        }
        return;
    }
    if (!solves) return;
    // the row is loaded only now: keeping it live across the factorization would double
    // the register footprint of wave 0
    double x[CHOL_NB];
    const T* Axl = A + (long long)c0 * ld + r;
#pragma unroll
    for (int c = 0; c < CHOL_NB; ++c) x[c] = (double)Axl[(long long)c * ld];
#pragma unroll
    for (int c = 0; c < CHOL_NB; ++c) {
        x[c] *= s_inv[c];
        const double xc = x[c];
#pragma unroll
        for (int k = c + 1; k < CHOL_NB; ++k) x[k] -= xc * sLT[c][k];
    }
    T* Ax = A + (long long)c0 * ld + r;
#pragma unroll
    for (int c = 0; c < CHOL_NB; ++c) Ax[(long long)c * ld] = (T)x[c];
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
int chol_factor(T* d_A, long long ld, int n_pad, int* d_status, int syrk_mode, hipStream_t st,
                PhaseRecorder* rec)
{
    BDH_CHECK(n_pad % PAD_TO == 0, "chol_factor: n_pad must be a multiple of 128");
    BDH_CHECK(ld >= n_pad && ld % 4 == 0, "chol_factor: bad leading dimension");
    for (int c0 = 0; c0 < n_pad; c0 += CHOL_NB) {
        const int below = n_pad - c0 - CHOL_NB;
        if (rec) rec->begin(PH_PANEL, st);
        hipLaunchKernelGGL(chol_panel_kernel<T>, dim3(1 + ceil_div(below, PANEL_ROWS)), dim3(256),
                           0, st, d_A, ld, c0, n_pad, d_status);
        BDH_LAUNCH_CHECK();
        if (rec) rec->end(PH_PANEL, st);
        const int t0 = c0 + CHOL_NB;
        if (t0 >= n_pad) break;
        const int nt = ceil_div(n_pad - t0, SY_TILE);
        const int ntiles = nt * (nt + 1) / 2;
        if (rec) rec->begin(PH_SYRK, st);
        if constexpr (sizeof(T) == 4) {
            (void)syrk_mode;
            hipLaunchKernelGGL(syrk_f32_kernel, dim3(ntiles), dim3(256), 0, st,
                               reinterpret_cast<float*>(d_A), ld, c0, t0, n_pad);
        } else {
            hipLaunchKernelGGL(syrk_f64_kernel, dim3(ntiles), dim3(256), 0, st,
                               reinterpret_cast<double*>(d_A), ld, c0, t0, n_pad);
        }
        BDH_LAUNCH_CHECK();
        if (rec) rec->end(PH_SYRK, st);
    }
    return 0;
}

// ------------------------------------------------------------------ y = L x (partials)
// Tile: (64 lanes x V rows) x TRMV_COLS columns, V = 16 / sizeof(T) so each lane issues one
// 16-byte load per column; the 4 waves split the tile's columns (unrolled by 8 for loads in
// flight) and meet in LDS.  Entries above the diagonal are masked with a select after the
// load (the storage there is finite but meaningless), so diagonal tiles take the same
// vector path.  P[cb][r] receives row r's partial over column block cb (FP64 accumulation).
constexpr int TRMV_COLS = 256;

template <typename T>
__device__ __forceinline__ void trmv_accum(const T* __restrict__ col, int rbase, int c, double xc,
                                           bool mask, double* acc)
{
    constexpr int V = 16 / sizeof(T);
    if constexpr (V == 4) {
        const float4 v = *reinterpret_cast<const float4*>(col + rbase);
        acc[0] += (double)((!mask || rbase + 0 >= c) ? v.x : 0.f) * xc;
        acc[1] += (double)((!mask || rbase + 1 >= c) ? v.y : 0.f) * xc;
        acc[2] += (double)((!mask || rbase + 2 >= c) ? v.z : 0.f) * xc;
        acc[3] += (double)((!mask || rbase + 3 >= c) ? v.w : 0.f) * xc;
    } else {
        const double2 v = *reinterpret_cast<const double2*>(col + rbase);
        acc[0] += ((!mask || rbase + 0 >= c) ? v.x : 0.0) * xc;
        acc[1] += ((!mask || rbase + 1 >= c) ? v.y : 0.0) * xc;
    }
}

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
    // the whole tile is below the diagonal when its last column precedes its first row
    const bool mask = (cbeg + TRMV_COLS - 1) > r0;
    if (rbase < n) {
        int c = c_lo;
        for (; c + 8 <= c_hi; c += 8) {
#pragma unroll
            for (int u = 0; u < 8; ++u)
                trmv_accum<T>(L + (long long)(c + u) * ld, rbase, c + u, x[c + u], mask, acc);
        }
        for (; c < c_hi; ++c) trmv_accum<T>(L + (long long)c * ld, rbase, c, x[c], mask, acc);
    }
#pragma unroll
    for (int q = 0; q < V; ++q) s_red[wave][lane * V + q] = acc[q];
    __syncthreads();
    for (int rr = tid; rr < ROWS; rr += 256) {
        const int row = r0 + rr;
        if (row < n)
            P[(long long)cb * n + row] = (s_red[0][rr] + s_red[1][rr]) + (s_red[2][rr] + s_red[3][rr]);
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

int trmv_blk_cols() { return TRMV_COLS; }

template int chol_factor<float>(float*, long long, int, int*, int, hipStream_t, PhaseRecorder*);
template int chol_factor<double>(double*, long long, int, int*, int, hipStream_t, PhaseRecorder*);
template int launch_trmv_partials<float>(const float*, long long, int, const double*, double*, hipStream_t);
template int launch_trmv_partials<double>(const double*, long long, int, const double*, double*, hipStream_t);

}  // namespace bdh
