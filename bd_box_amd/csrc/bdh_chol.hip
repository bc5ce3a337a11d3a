// Blocked right-looking Cholesky D = L L^T and the triangular product L.x on gfx950.
//
// Replaces the reference's my_dpotrf / my_dtrmv (src/myblas.c:45-118: LAPACK
// dpotrf("U") + dtrmv("U","T"), or MAGMA spotrf_gpu('L') + cublasStrmv in the CUDA
// build).  The factor is stored lower, column-major, in place of D; n is padded to a
// multiple of PAD_TO with an identity block so every tile is whole.
//
// Per panel of CHOL_NB = 64 columns c0..c0+63:
//   diagonal block  L_kk of the 64x64 block in FP64 (factor_diag64): the four wavefronts
//                   own 16 columns each and run as a pipeline, each applying the earlier
//                   columns' updates as they are published in LDS and then factoring its
//                   own.  Stored in T into A and in FP64 into Lbuf.
//   TRSM            L_ik = A_ik L_kk^{-T} for every row below, FP64 forward substitution,
//                   four lanes per row (a quad exchanges each finished entry through DPP).
//   trailing SYRK   A22 -= L21 L21^T over 128x128 lower tiles on MFMA:
//                   T = float : v_mfma_f32_32x32x2_f32, FP32 accumulation
//                   T = double: v_mfma_f64_16x16x4_f64
//                   Workgroup 0 owns the tile holding the next panel's diagonal block; it
//                   keeps that block in LDS and factors it right after its update, so the
//                   next diagonal factorization overlaps the rest of the SYRK (lookahead
//                   without a second stream).  Only panel 0 needs its own launch.
// FP64 is used for the diagonal blocks and the panel solve in both precisions; in the FP32
// variant the factor is rounded to FP32 when stored.
#include <algorithm>

#include "bdh_common.hpp"
#include "bdh_kernels.hpp"
#include "bdh_mfma.hpp"

namespace bdh {

constexpr int DG_LD = CHOL_NB + 1;   // row pitch (doubles) of the diagonal block in LDS
constexpr int SY_TILE_C = 128;       // SYRK tile edge
constexpr int TRSM_ROWS = 64;        // rows per TRSM workgroup (4 lanes per row)

// Value of v held by lane `src` (uniform), through two v_readlane_b32.
__device__ __forceinline__ double readlane_f64(double v, int src)
{
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)u, src);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u >> 32), src);
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// 1/sqrt(x): hardware estimate refined by two Newton steps y <- y (3 - x y^2) / 2.
__device__ __forceinline__ double rsqrt_f64(double x)
{
    double y = __builtin_amdgcn_rsq(x);
    y = y * (1.5 - 0.5 * x * y * y);
    y = y * (1.5 - 0.5 * x * y * y);
    return y;
}

// Per diagonal block, the workspace Lb holds (FP64):
//   Lb[0 .. 4095]      L_kk column-major, zeros above the diagonal
//   Lb[4096 + 256 J]   inverse of its J-th 16x16 diagonal sub-block, column-major, J = 0..3
constexpr int LB_STRIDE = CHOL_NB * CHOL_NB + 4 * 256;

// Bounded wait (inside one workgroup) until the LDS counter reaches `need`; returns the
// counter's value then, or -1 if it never did.  Every wavefront that publishes to the counter
// runs concurrently with the waiting one, so the wait ends unless a wavefront faulted; the
// bound keeps a defect from hanging the GPU.
constexpr int LDS_SPIN = 1 << 22;
__device__ __forceinline__ int wait_lds_count(int* cnt, int need)
{
    for (int it = 0; it < LDS_SPIN; ++it) {
        const int v = __hip_atomic_load(cnt, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
        if (v >= need) return v;
        __builtin_amdgcn_s_sleep(1);
    }
    return -1;
}

// In-place FP64 Cholesky of the 64x64 block held in LDS, called by all 256 threads.
//   in : sD[r * DG_LD + c], lower part (r >= c) read
//   out: the same memory as L[c * 64 + r] (column-major, pitch 64, zeros above the diagonal),
//        the inverses of its four 16x16 diagonal sub-blocks after it (sD + 4096 + 256 J, the
//        layout load_lkk leaves: sD is then sLT / sInv of the solve), and the same in Lb.
//        sD must hold CHOL_NB * CHOL_NB + 4 * 256 doubles.
// The four wavefronts form a pipeline over column groups: wavefront w owns columns
// 16w .. 16w+15 (lane = row, one register per column).  It applies the rank-1 updates of the
// 16w earlier columns as the wavefronts before it publish them (column in LDS, then an LDS
// counter), then factors its own 16 columns one by one, publishing each.  So the pivot chain
// runs through the wavefronts back to back and every update off it runs beside it.
// Within a wavefront's own columns, pivot t+1 = a_{t+1,t+1} - (a_{t+1,t} / sqrt(d_t))^2 is
// formed from two broadcasts (v_readlane) of the pre-update entries, issued before the
// pivot's reciprocal square root is known; the rank-1 update of the later own columns reads
// the scaled column back from LDS as broadcasts, and the same entries build the inverse of
// the own 16x16 sub-block step by step, so it is complete with the last own column.
// A non-positive (or NaN) pivot records row c0 + j + 1 in *status (the first one wins) and the
// factorization continues with NaN, as the reference's dpotrf reports INFO.
// head(), tail(): work of the caller's for the wavefronts' idle stretches (wavefront w's own
// columns begin only after the 16 w columns before them, and wavefront 0 waits from its last
// column to the end): each wavefront runs head() once after its block entries are in
// registers, before it takes the earlier columns, and tail() once after its own columns.
template <typename Head, typename Tail>
__device__ __forceinline__ void factor_diag64(double* __restrict__ sD, int c0, int* __restrict__ status,
                                              double* __restrict__ Lb, Head&& head, Tail&& tail)
{
    __shared__ int s_done;   // columns published so far
    __shared__ int s_bad;    // first non-positive pivot (64: none); -1: a wait timed out
    const int tid = threadIdx.x, lane = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int cw = 16 * w;   // first own column
    __syncthreads();   // the block in sD may have been written by one wavefront alone
    double p[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) p[u] = (lane >= cw + u) ? sD[lane * DG_LD + cw + u] : 0.0;
    if (tid == 0) {
        s_done = 0;
        s_bad = CHOL_NB;
    }
    __syncthreads();   // every block entry is in registers: sD is reused for L
    head();
    double* L = sD;
    bool ok = true;
    // updates by the earlier column groups, two columns at a time as they are published
    // (cw is a multiple of 16, so a pair never straddles two groups)
    int avail = 0;
    for (int c = 0; c < cw; c += 2) {
        if (avail < c + 2) {
            avail = wait_lds_count(&s_done, c + 2);
            if (avail < 0) {
                ok = false;
                break;
            }
        }
        // every LDS read of the pair is issued before its first use
        const double x0 = L[c * CHOL_NB + lane], x1 = L[(c + 1) * CHOL_NB + lane];
        const double2* b0 = reinterpret_cast<const double2*>(L + c * CHOL_NB + cw);
        const double2* b1 = reinterpret_cast<const double2*>(L + (c + 1) * CHOL_NB + cw);
        double2 v0[8], v1[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            v0[u] = b0[u];
            v1[u] = b1[u];
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            p[2 * u] = __builtin_fma(-x1, v1[u].x, __builtin_fma(-x0, v0[u].x, p[2 * u]));
            p[2 * u + 1] = __builtin_fma(-x1, v1[u].y, __builtin_fma(-x0, v0[u].y, p[2 * u + 1]));
        }
    }
    int bad = CHOL_NB;
    if (ok) {
        // inverse of the own 16x16 diagonal sub-block, built beside the factorization: lane
        // q < 16 solves L_ww y = e_q right-looking, step t taking column t's 1 / l_tt and the
        // entries l_{cw+u, cw+t} the column update broadcasts anyway
        double y[16];
#pragma unroll
        for (int r = 0; r < 16; ++r) y[r] = (r == lane) ? 1.0 : 0.0;
        double piv = readlane_f64(p[0], cw);
#pragma unroll
        for (int t = 0; t < 16; ++t) {
            const int c = cw + t;
            if (!(piv > 0.0) && bad == CHOL_NB) bad = c;
            // the next pivot a_{c+1,c+1} - l_{c+1,c}^2 from the entries before this column's
            // update (two broadcasts issued before the reciprocal square root is known)
            double a1 = 0.0, a2 = 0.0;
            if (t < 15) {
                a1 = readlane_f64(p[t], c + 1);
                a2 = readlane_f64(p[t + 1], c + 1);
            }
            const double rs = rsqrt_f64(piv);
            double nxt = 0.0;
            if (t < 15) {
                const double le = a1 * rs;   // l_{c+1, c}
                nxt = __builtin_fma(-le, le, a2);
            }
            const double lp = p[t] * rs, ld = piv * rs;
            double l = (lane == c) ? ld : lp;
            l = (lane < c) ? 0.0 : l;
            p[t] = l;
            L[c * CHOL_NB + lane] = l;
            y[t] *= rs;
            if (t < 15) {
                // l_{cw+u, c} of the later own columns: LDS broadcasts of the column this
                // wavefront has just stored (its own stores, in order before these loads)
                __builtin_amdgcn_wave_barrier();
                double lc[16];
#pragma unroll
                for (int u = t + 1; u < 16; ++u) lc[u] = L[c * CHOL_NB + cw + u];
#pragma unroll
                for (int u = t + 1; u < 16; ++u) p[u] = __builtin_fma(-l, lc[u], p[u]);
#pragma unroll
                for (int u = t + 1; u < 16; ++u) {
                    y[u] = __builtin_fma(-lc[u], y[t], y[u]);
                    // keeps the step here: left to the scheduler, the steps sink to the stores
                    // at the end and every broadcast stays live until then
                    asm volatile("" : "+v"(y[u]));
                }
            }
            // columns are published in pairs (the consumers take them in pairs)
            if (t & 1) __hip_atomic_store(&s_done, c + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
            piv = nxt;
        }
        // own columns into Lb (zeros above the diagonal)
#pragma unroll
        for (int u = 0; u < 16; ++u) Lb[(cw + u) * CHOL_NB + lane] = (lane >= cw + u) ? p[u] : 0.0;
        if (lane < 16) {
            double* inv = Lb + CHOL_NB * CHOL_NB + 256 * w;
            double* sinv = sD + CHOL_NB * CHOL_NB + 256 * w;   // the solve's sInv layout
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                inv[lane * 16 + r] = y[r];
                sinv[lane * 16 + r] = y[r];
            }
        }
    }
    tail();
    if (lane == 0) {
        if (!ok) atomicExch(&s_bad, -1);
        else if (bad < CHOL_NB) atomicMin(&s_bad, bad);
    }
    __syncthreads();   // L complete in LDS, every wavefront's Lb stores issued
    if (tid == 0) {
        if (s_bad < 0) atomicCAS(status, 0, -1);
        else if (s_bad < CHOL_NB) atomicCAS(status, 0, c0 + s_bad + 1);
    }
}

__device__ __forceinline__ void factor_diag64(double* __restrict__ sD, int c0, int* __restrict__ status,
                                              double* __restrict__ Lb)
{
    factor_diag64(sD, c0, status, Lb, [] {}, [] {});
}

// L (lower part) of the factored block, L[c * 64 + r] in LDS, into A, in T.
template <typename T>
__device__ __forceinline__ void store_diag64(const double* __restrict__ L, T* __restrict__ A,
                                             long long ld, int c0)
{
    for (int e = threadIdx.x; e < CHOL_NB * CHOL_NB; e += 256) {
        const int r = e % CHOL_NB, c = e / CHOL_NB;
        if (r >= c) A[(long long)(c0 + c) * ld + c0 + r] = (T)L[c * CHOL_NB + r];
    }
}

// Panel 0's diagonal block (later ones are factored inside the SYRK kernel).
template <typename T>
__device__ __forceinline__ void chol_diag_body(T* __restrict__ A, long long ld, int c0,
                                               double* __restrict__ Lb, int* __restrict__ flags,
                                               int nflags, int* __restrict__ status)
{
    __shared__ double sD[CHOL_NB * CHOL_NB + 4 * 256];
    static_assert(CHOL_NB * CHOL_NB + 4 * 256 >= CHOL_NB * DG_LD, "diagonal block LDS");
    for (int k = threadIdx.x; k < nflags; k += 256) flags[k] = 0;
    double v[CHOL_NB * CHOL_NB / 256];
#pragma unroll
    for (int i = 0; i < CHOL_NB * CHOL_NB / 256; ++i) {
        const int e = threadIdx.x + 256 * i, r = e % CHOL_NB, c = e / CHOL_NB;
        v[i] = (r >= c) ? (double)A[(long long)(c0 + c) * ld + c0 + r] : 0.0;
    }
#pragma unroll
    for (int i = 0; i < CHOL_NB * CHOL_NB / 256; ++i) {
        const int e = threadIdx.x + 256 * i, r = e % CHOL_NB, c = e / CHOL_NB;
        sD[r * DG_LD + c] = v[i];
    }
    factor_diag64(sD, c0, status, Lb);   // begins with a barrier
    store_diag64<T>(sD, A, ld, c0);
}

template <typename T>
__global__ void __launch_bounds__(256)
chol_diag_kernel(T* __restrict__ A, long long ld, int c0, double* __restrict__ Lb,
                 int* __restrict__ flags, int nflags, int* __restrict__ status)
{
    chol_diag_body<T>(A, ld, c0, Lb, flags, nflags, status);
}

// Broadcast within each quad of lanes from quad lane Q (DPP quad_perm).
template <int Q>
__device__ __forceinline__ double quad_bcast_t(double v)
{
    constexpr int ctrl = Q | (Q << 2) | (Q << 4) | (Q << 6);
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const int lo = __builtin_amdgcn_mov_dpp((int)(unsigned)u, ctrl, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp((int)(unsigned)(u >> 32), ctrl, 0xF, 0xF, false);
    return __builtin_bit_cast(double, ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ double quad_bcast(double v, int q)
{
    switch (q) {
        case 0: return quad_bcast_t<0>(v);
        case 1: return quad_bcast_t<1>(v);
        case 2: return quad_bcast_t<2>(v);
        default: return quad_bcast_t<3>(v);
    }
}

// TRSM x L_kk^T = b for 16 rows (one wavefront), blocked by the four 16-column sub-blocks J
// of the panel, on v_mfma_f64_16x16x4_f64:
//   Y_J = B_J - sum_{K<J} X_K L_JK^T,   X_J = Y_J Linv_JJ^T
// Y and X live in the MFMA result layout (accumulator r of lane l = row 4 r + l/16, column
// l%16 of the 16x16 block); a finished block goes through LDS (pitch 17) to become an A
// operand (lane l: row l%16, k = 4 s + l/16).  sLT[k * 64 + c] = L_kk[c][k];
// sInv[256 J + c' * 16 + c] = Linv_JJ[c][c'].  All FP64.
constexpr int SC_PITCH = 17;

struct Trsm16 {
    doublex4 y[4];   // B_J in, X_J out, result layout
};

template <typename T>
__device__ __forceinline__ void trsm16_load(const T* __restrict__ A, long long ld, int t0, int g0,
                                            int re, int lane, Trsm16& g)
{
    const int lr = lane & 15, lq = lane >> 4;
#pragma unroll
    for (int J = 0; J < 4; ++J) {
        const T* col = A + (long long)(t0 + 16 * J + lr) * ld;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int row = g0 + 4 * r + lq;
            g.y[J][r] = (row < re) ? (double)col[row] : 0.0;
        }
    }
}

template <typename T>
__device__ __forceinline__ void trsm16_solve_store(T* __restrict__ A, long long ld, int t0, int g0,
                                                   int re, int lane, Trsm16& g,
                                                   const double* __restrict__ sLT,
                                                   const double* __restrict__ sInv,
                                                   double* __restrict__ sc)
{
    const int lr = lane & 15, lq = lane >> 4;
    double xa[4][4];   // A-operand form of the finished X_K: xa[K][s] = X_K[lr][4 s + lq]
#pragma unroll
    for (int J = 0; J < 4; ++J) {
        doublex4 yj = g.y[J];
#pragma unroll
        for (int K = 0; K < J; ++K)
#pragma unroll
            for (int s4 = 0; s4 < 4; ++s4) {
                const double b = sLT[(16 * K + 4 * s4 + lq) * CHOL_NB + 16 * J + lr];
                yj = mfma_f64_16x16x4(-xa[K][s4], b, yj);
            }
        // Y_J into A-operand form
#pragma unroll
        for (int r = 0; r < 4; ++r) sc[(4 * r + lq) * SC_PITCH + lr] = yj[r];
        __builtin_amdgcn_wave_barrier();
        double ya[4];
#pragma unroll
        for (int s4 = 0; s4 < 4; ++s4) ya[s4] = sc[lr * SC_PITCH + 4 * s4 + lq];
        __builtin_amdgcn_wave_barrier();
        doublex4 xj = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int s4 = 0; s4 < 4; ++s4)
            xj = mfma_f64_16x16x4(ya[s4], sInv[256 * J + (4 * s4 + lq) * 16 + lr], xj);
        T* col = A + (long long)(t0 + 16 * J + lr) * ld;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int row = g0 + 4 * r + lq;
            if (row < re) col[row] = (T)xj[r];
        }
        if (J < 3) {
#pragma unroll
            for (int r = 0; r < 4; ++r) sc[(4 * r + lq) * SC_PITCH + lr] = xj[r];
            __builtin_amdgcn_wave_barrier();
#pragma unroll
            for (int s4 = 0; s4 < 4; ++s4) xa[J][s4] = sc[lr * SC_PITCH + 4 * s4 + lq];
            __builtin_amdgcn_wave_barrier();
        }
    }
}

// Two 16-row groups a and b of one wavefront solved together: the same L operands, two
// independent MFMA chains interleaved (each hides the other's latency).  Rows of b at or
// beyond re are neither loaded nor stored.
template <typename T>
__device__ __forceinline__ void trsm16x2_solve_store(T* __restrict__ A, long long ld, int t0, int ga0,
                                                     int gb0, int re, int lane, Trsm16& ga, Trsm16& gb,
                                                     const double* __restrict__ sLT,
                                                     const double* __restrict__ sInv,
                                                     double* __restrict__ sc)
{
    const int lr = lane & 15, lq = lane >> 4;
    double xa[4][4], xb[4][4];
#pragma unroll
    for (int J = 0; J < 4; ++J) {
        doublex4 ya = ga.y[J], yb = gb.y[J];
#pragma unroll
        for (int K = 0; K < J; ++K)
#pragma unroll
            for (int s4 = 0; s4 < 4; ++s4) {
                const double b = sLT[(16 * K + 4 * s4 + lq) * CHOL_NB + 16 * J + lr];
                ya = mfma_f64_16x16x4(-xa[K][s4], b, ya);
                yb = mfma_f64_16x16x4(-xb[K][s4], b, yb);
            }
        double yaa[4], yba[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) sc[(4 * r + lq) * SC_PITCH + lr] = ya[r];
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int s4 = 0; s4 < 4; ++s4) yaa[s4] = sc[lr * SC_PITCH + 4 * s4 + lq];
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int r = 0; r < 4; ++r) sc[(4 * r + lq) * SC_PITCH + lr] = yb[r];
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int s4 = 0; s4 < 4; ++s4) yba[s4] = sc[lr * SC_PITCH + 4 * s4 + lq];
        __builtin_amdgcn_wave_barrier();
        doublex4 xja = {0.0, 0.0, 0.0, 0.0}, xjb = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int s4 = 0; s4 < 4; ++s4) {
            const double iv = sInv[256 * J + (4 * s4 + lq) * 16 + lr];
            xja = mfma_f64_16x16x4(yaa[s4], iv, xja);
            xjb = mfma_f64_16x16x4(yba[s4], iv, xjb);
        }
        T* col = A + (long long)(t0 + 16 * J + lr) * ld;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int rowa = ga0 + 4 * r + lq, rowb = gb0 + 4 * r + lq;
            if (rowa < re) col[rowa] = (T)xja[r];
            if (rowb < re) col[rowb] = (T)xjb[r];
        }
        if (J < 3) {
#pragma unroll
            for (int r = 0; r < 4; ++r) sc[(4 * r + lq) * SC_PITCH + lr] = xja[r];
            __builtin_amdgcn_wave_barrier();
#pragma unroll
            for (int s4 = 0; s4 < 4; ++s4) xa[J][s4] = sc[lr * SC_PITCH + 4 * s4 + lq];
            __builtin_amdgcn_wave_barrier();
#pragma unroll
            for (int r = 0; r < 4; ++r) sc[(4 * r + lq) * SC_PITCH + lr] = xjb[r];
            __builtin_amdgcn_wave_barrier();
#pragma unroll
            for (int s4 = 0; s4 < 4; ++s4) xb[J][s4] = sc[lr * SC_PITCH + 4 * s4 + lq];
            __builtin_amdgcn_wave_barrier();
        }
    }
}

// L_kk and its diagonal sub-block inverses (Lb, FP64) into LDS.
__device__ __forceinline__ void load_lkk(const double* __restrict__ Lb, double* __restrict__ sLT,
                                         double* __restrict__ sInv)
{
    // Lb is column-major (Lb[k * CHOL_NB + c] = L_kk[c][k]), which is the sLT layout.  All
    // loads are issued before the LDS stores (which the compiler would otherwise order
    // between them).
    constexpr int NL = CHOL_NB * CHOL_NB / 256, NI = 4;
    double v[NL + NI];
#pragma unroll
    for (int i = 0; i < NL + NI; ++i) v[i] = Lb[threadIdx.x + 256 * i];
#pragma unroll
    for (int i = 0; i < NL; ++i) sLT[threadIdx.x + 256 * i] = v[i];
#pragma unroll
    for (int i = 0; i < NI; ++i) sInv[threadIdx.x + 256 * i] = v[NL + i];
    __syncthreads();
}

// LDS of a TRSM workgroup: sLT, sInv, and 16 x SC_PITCH doubles per wavefront.
constexpr int TRSM_LDS_DOUBLES = CHOL_NB * CHOL_NB + 4 * 256 + 4 * 16 * SC_PITCH;

// Panel c0's TRSM as its own launch (panel 0 only; the others run inside the SYRK kernel):
// 64 rows per workgroup, 16 per wavefront.
template <typename T>
__device__ __forceinline__ void chol_trsm_body(T* __restrict__ A, long long ld, int c0, int n_pad,
                                               const double* __restrict__ Lb, int bid)
{
    __shared__ double smem[TRSM_LDS_DOUBLES];
    double* sLT = smem;
    double* sInv = sLT + CHOL_NB * CHOL_NB;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    double* sc = sInv + 4 * 256 + wave * 16 * SC_PITCH;
    const int g0 = c0 + CHOL_NB + bid * TRSM_ROWS + 16 * wave;
    Trsm16 g;
    trsm16_load<T>(A, ld, c0, g0, n_pad, lane, g);
    load_lkk(Lb, sLT, sInv);
    if (g0 < n_pad) trsm16_solve_store<T>(A, ld, c0, g0, n_pad, lane, g, sLT, sInv, sc);
}

template <typename T>
__global__ void __launch_bounds__(256)
chol_trsm_kernel(T* __restrict__ A, long long ld, int c0, int n_pad, const double* __restrict__ Lb)
{
    chol_trsm_body<T>(A, ld, c0, n_pad, Lb, blockIdx.x);
}

// Flags of the diagonal blocks, after the FP64 blocks in the workspace: flag k is set once
// block k's factor is in Lb (cleared by the panel-0 diagonal kernel of each factorization).
__device__ __forceinline__ int* chol_flags(double* Lbuf, int n_pad)
{
    return reinterpret_cast<int*>(Lbuf + (long long)(n_pad / CHOL_NB) * LB_STRIDE);
}

// Bounded wait (about a second) for a flag set by another workgroup of the same grid.  The
// polls are relaxed; one agent-scope acquire follows the flag (an acquiring poll would
// invalidate this XCD's L2 at every iteration, under the feet of the other workgroups).
constexpr int CHOL_SPIN = 1 << 20;
// SLEEP: s_sleep units (64 cycles each) between polls; the bound on the polls scales with it.
template <int SLEEP = 2>
__device__ __forceinline__ bool wait_flag(int* flag)
{
    for (int it = 0; it < CHOL_SPIN * 2 / SLEEP; ++it) {
        if (__hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            return true;
        }
        __builtin_amdgcn_s_sleep(SLEEP);
    }
    return false;
}

// Tail of the SYRK kernels (tile column 0): the next panel's diagonal block was factored by
// workgroup 0 (diag); every tile of column 0 then solves its own rows of that panel, 16 rows
// per wavefront at a time.  The rows are loaded before the wait (this workgroup wrote
// them), so the wait hides the loads.
template <typename T>
__device__ __forceinline__ void fused_trsm(T* __restrict__ A, long long ld, int t0, int ri0, int n_pad,
                                           double* __restrict__ Lb_next, int* __restrict__ flag,
                                           int* __restrict__ status, bool diag, unsigned char* smem,
                                           const double* __restrict__ sD)
{
    double* sLT = reinterpret_cast<double*>(smem);
    double* sInv = sLT + CHOL_NB * CHOL_NB;
    __shared__ int s_ok;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    double* sc = sInv + 4 * 256 + wave * 16 * SC_PITCH;
    const int rb = max(ri0, t0 + CHOL_NB), re = min(n_pad, ri0 + SY_TILE_C);
    const int g0 = rb + 16 * wave, g1 = g0 + 64;   // this wavefront's two 16-row groups
    if (diag) {
        // the FP64 block is in Lb (factor_diag64); the T copy in A is not read by this kernel
        // each wavefront's Lb stores written back (release only: no L2 invalidation), then
        // the flag
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __syncthreads();
        if (tid == 0) __hip_atomic_store(flag, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        store_diag64<T>(sD, A, ld, t0);
    }
    // this workgroup's own C stores are read back below: workgroup scope suffices (no L2
    // write-back / invalidation, which an agent-scope fence would cost every such workgroup)
    __syncthreads();
    Trsm16 x0, x1;
    if (g0 < re) trsm16_load<T>(A, ld, t0, g0, re, lane, x0);
    if (g1 < re) trsm16_load<T>(A, ld, t0, g1, re, lane, x1);
    else x1 = Trsm16{};
    if (tid == 0) {
        if (diag) {
            s_ok = 1;
        } else {
            s_ok = wait_flag(flag) ? 1 : 0;
            if (!s_ok) atomicCAS(status, 0, -1);
        }
    }
    __syncthreads();
    if (!s_ok) return;
    load_lkk(Lb_next, sLT, sInv);
    if (g1 < re) trsm16x2_solve_store<T>(A, ld, t0, g0, g1, re, lane, x0, x1, sLT, sInv, sc);
    else if (g0 < re) trsm16_solve_store<T>(A, ld, t0, g0, re, lane, x0, sLT, sInv, sc);
}

// Tile (ti, tj), ti >= tj, of a lower trapezoid with ni tile rows, enumerated column by
// column (tile (0, 0) first).
__device__ __forceinline__ void trap_tile(int x, int ni, int& ti, int& tj)
{
    int j = 0;
    while (x >= ni - j) {
        x -= ni - j;
        ++j;
    }
    ti = j + x;
    tj = j;
}

// ------------------------------------------------------------------ FP32 SYRK (MFMA)
// Workgroup 256 = 4 waves in 2x2; tile 128x128; each wave 64x64 = 2x2 blocks of 32x32.
//   SPLIT = false: v_mfma_f32_32x32x2_f32 on the FP32 operands.
//   SPLIT = true : each FP32 operand x = hi + lo with hi = bf16(x), lo = bf16(x - hi)
//                  (x - hi is exact in FP32), and a b = hi_a hi_b + hi_a lo_b + lo_a hi_b
//                  on v_mfma_f32_32x32x16_bf16 with FP32 accumulation; the dropped lo_a lo_b
//                  and the split residuals are below 2^-17 |a b|.
constexpr int SY_TILE = SY_TILE_C;
__device__ __forceinline__ int ceil_div_dev(int a, int b) { return (a + b - 1) / b; }
constexpr int SY_KC32 = 32;
constexpr int SY_KCH = 32;                 // k per stage of the split kernel
constexpr int SY_HP = SY_KCH / 2 + 4;      // row pitch of a split stage, in 32-bit words

// LDS of the SYRK kernels: the operand stages, reused for the next diagonal block.
constexpr int SY_SMEM_F32 = 2 * SY_KC32 * (SY_TILE + 4) * 4;
constexpr int SY_SMEM_SPLIT = 4 * SY_TILE * SY_HP * 4;
constexpr int SY_SMEM_MAX2 = SY_SMEM_SPLIT > SY_SMEM_F32 ? SY_SMEM_SPLIT : SY_SMEM_F32;
constexpr int SY_SMEM = SY_SMEM_MAX2 > TRSM_LDS_DOUBLES * 8 ? SY_SMEM_MAX2 : TRSM_LDS_DOUBLES * 8;
static_assert(SY_SMEM >= CHOL_NB * DG_LD * 8, "diagonal block does not fit the SYRK LDS");
static_assert(SY_SMEM >= TRSM_LDS_DOUBLES * 8, "TRSM operands do not fit the SYRK LDS");

// bf16 bits of x, round to nearest even (finite inputs)
__device__ __forceinline__ unsigned bf16_bits(float x)
{
    const unsigned u = __builtin_bit_cast(unsigned, x);
    return (u + 0x7FFFu + ((u >> 16) & 1u)) >> 16;
}
__device__ __forceinline__ void split_bf16(float x, unsigned& hi, unsigned& lo)
{
    hi = bf16_bits(x);
    lo = bf16_bits(x - __builtin_bit_cast(float, hi << 16));
}

// acc += A[ri0 .. ri0+127, kb .. kb+K) A[rj0 .. rj0+127, kb .. kb+K)^T for one 128x128 tile on
// MFMA (rows at or beyond n_pad read as zero).  Workgroup 256 = 4 waves in 2x2, each 64x64 =
// 2x2 blocks of 32x32; acc[a][b] is the 32x32 block (a, b) of the wave in the MFMA result
// layout.  The operand stages use smem (SY_SMEM bytes).
template <bool SPLIT>
__device__ __forceinline__ void tile_acc_f32(const float* __restrict__ A, long long ld, int kb, int K, int ri0,
                                             int rj0, int n_pad, unsigned char* smem, floatx16 (&acc)[2][2])
{
    float (*sA)[SY_TILE + 4] = reinterpret_cast<float (*)[SY_TILE + 4]>(smem);
    float (*sB)[SY_TILE + 4] = reinterpret_cast<float (*)[SY_TILE + 4]>(smem + SY_KC32 * (SY_TILE + 4) * 4);
    // split stages: [0] A hi, [1] A lo, [2] B hi, [3] B lo; row-major (row, k pairs)
    unsigned (*sH)[SY_TILE][SY_HP] = reinterpret_cast<unsigned (*)[SY_TILE][SY_HP]>(smem);
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wr = wave >> 1, wc = wave & 1;
    const int l32 = lane & 31, lh = lane >> 5;
    if constexpr (SPLIT) {
        // staging: thread t loads row t of the A rows (t < 128) or of the B rows, SY_KCH
        // consecutive k (one coalesced load per k across the wave), splits, and writes
        // k-contiguous hi / lo words
        const bool isB = tid >= SY_TILE;
        const int srow = isB ? rj0 + tid - SY_TILE : ri0 + tid;
        const int lrow = tid & (SY_TILE - 1);
        const bool live = srow < n_pad;
        // the next chunk's loads are issued before this chunk's MFMAs (register double
        // buffering: v is converted into LDS before it is refilled)
        float v[SY_KCH];
        {
            const float* src = A + (long long)kb * ld + srow;
#pragma unroll
            for (int k = 0; k < SY_KCH; ++k) v[k] = live ? src[(long long)k * ld] : 0.f;
        }
        for (int k0 = 0; k0 < K; k0 += SY_KCH) {
            __syncthreads();
#pragma unroll
            for (int k = 0; k < SY_KCH; k += 8) {
                unsigned h[8], l[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) split_bf16(v[k + u], h[u], l[u]);
                const uint4 hw = make_uint4(h[0] | (h[1] << 16), h[2] | (h[3] << 16),
                                            h[4] | (h[5] << 16), h[6] | (h[7] << 16));
                const uint4 lw = make_uint4(l[0] | (l[1] << 16), l[2] | (l[3] << 16),
                                            l[4] | (l[5] << 16), l[6] | (l[7] << 16));
                *reinterpret_cast<uint4*>(&sH[isB ? 2 : 0][lrow][k / 2]) = hw;
                *reinterpret_cast<uint4*>(&sH[isB ? 3 : 1][lrow][k / 2]) = lw;
            }
            __syncthreads();
            if (k0 + SY_KCH < K) {
                const float* src = A + (long long)(kb + k0 + SY_KCH) * ld + srow;
#pragma unroll
                for (int k = 0; k < SY_KCH; ++k) v[k] = live ? src[(long long)k * ld] : 0.f;
            }
#pragma unroll
            for (int t = 0; t < SY_KCH / 16; ++t) {
                const int w = 8 * t + 4 * lh;   // lane holds k = 16 t + 8 lh + 0..7
                bf16x8 ah[2], al[2], bh[2], bl[2];
#pragma unroll
                for (int a = 0; a < 2; ++a) {
                    const int r = wr * 64 + a * 32 + l32;
                    ah[a] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(&sH[0][r][w]));
                    al[a] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(&sH[1][r][w]));
                }
#pragma unroll
                for (int b = 0; b < 2; ++b) {
                    const int r = wc * 64 + b * 32 + l32;
                    bh[b] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(&sH[2][r][w]));
                    bl[b] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(&sH[3][r][w]));
                }
#pragma unroll
                for (int a = 0; a < 2; ++a)
#pragma unroll
                    for (int b = 0; b < 2; ++b) {
                        acc[a][b] = mfma_bf16_32x32x16(al[a], bh[b], acc[a][b]);
                        acc[a][b] = mfma_bf16_32x32x16(ah[a], bl[b], acc[a][b]);
                        acc[a][b] = mfma_bf16_32x32x16(ah[a], bh[b], acc[a][b]);
                    }
            }
        }
    } else {
        // register double buffering: the next chunk's global loads are issued before this
        // chunk's MFMAs (each thread: 4 float4 of the A rows and 4 of the B rows per chunk)
        constexpr int NV = SY_KC32 * SY_TILE / 4 / 256;
        float4 va[NV], vb[NV];
        auto load_chunk = [&](int k0) {
#pragma unroll
            for (int i = 0; i < NV; ++i) {
                const int idx = tid + 256 * i;
                const int k = idx / (SY_TILE / 4);
                const int r4 = (idx % (SY_TILE / 4)) * 4;
                const float* col = A + (long long)(kb + k0 + k) * ld;
                va[i] = (ri0 + r4 < n_pad) ? *reinterpret_cast<const float4*>(col + ri0 + r4) : make_float4(0.f, 0.f, 0.f, 0.f);
                vb[i] = (rj0 + r4 < n_pad) ? *reinterpret_cast<const float4*>(col + rj0 + r4) : make_float4(0.f, 0.f, 0.f, 0.f);
            }
        };
        load_chunk(0);
        for (int k0 = 0; k0 < K; k0 += SY_KC32) {
            __syncthreads();
#pragma unroll
            for (int i = 0; i < NV; ++i) {
                const int idx = tid + 256 * i;
                const int k = idx / (SY_TILE / 4);
                const int r4 = (idx % (SY_TILE / 4)) * 4;
                *reinterpret_cast<float4*>(&sA[k][r4]) = va[i];
                *reinterpret_cast<float4*>(&sB[k][r4]) = vb[i];
            }
            __syncthreads();
            if (k0 + SY_KC32 < K) load_chunk(k0 + SY_KC32);
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
    }
}

// C -= acc for the tile (ri0, rj0): columns at or beyond t1 and rows at or beyond n_pad are
// left alone.  With to_lds (wavefront 0 of the workgroup whose tile holds the next diagonal
// block) the updated block goes to sD (row pitch DG_LD, FP64) instead of A.
__device__ __forceinline__ void tile_epilogue_f32(float* __restrict__ A, long long ld, int ri0, int rj0, int t1,
                                                  int n_pad, const floatx16 (&acc)[2][2], bool to_lds,
                                                  double* __restrict__ sD)
{
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wr = wave >> 1, wc = wave & 1;
    const int l32 = lane & 31, lh = lane >> 5;
    // C -= acc ; lane owns column l32 and rows 8*(r/4) + 4*lh + (0..3) of each 32x32 block.
    // All 16 loads of C are issued before any store (the stores could alias them as far as
    // the compiler knows, which would serialise every load behind the previous store).
    float4 cv[2][2][4];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            const int col = rj0 + wc * 64 + b * 32 + l32;
            const float* cp = A + (long long)col * ld;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int row = ri0 + wr * 64 + a * 32 + 8 * q + 4 * lh;
                cv[a][b][q] = (col < t1 && row < n_pad) ? *reinterpret_cast<const float4*>(cp + row)
                                                         : make_float4(0.f, 0.f, 0.f, 0.f);
            }
        }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            const int col = rj0 + wc * 64 + b * 32 + l32;
            if (col >= t1) continue;
            float* cp = A + (long long)col * ld;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int row = ri0 + wr * 64 + a * 32 + 8 * q + 4 * lh;
                if (row >= n_pad) continue;
                float4 c = cv[a][b][q];
                c.x -= acc[a][b][4 * q + 0];
                c.y -= acc[a][b][4 * q + 1];
                c.z -= acc[a][b][4 * q + 2];
                c.w -= acc[a][b][4 * q + 3];
                if (to_lds) {
                    const int lr = row - ri0, lc = col - rj0;
                    sD[(lr + 0) * DG_LD + lc] = c.x;
                    sD[(lr + 1) * DG_LD + lc] = c.y;
                    sD[(lr + 2) * DG_LD + lc] = c.z;
                    sD[(lr + 3) * DG_LD + lc] = c.w;
                } else {
                    *reinterpret_cast<float4*>(cp + row) = c;
                }
            }
        }
}

// Three waves per SIMD: the accumulators live in VGPRs next to the fused TRSM's registers.
template <bool SPLIT>
__device__ __forceinline__ void syrk_f32_body(float* __restrict__ A, long long ld, int kb, int K, int t0,
                                              int t1, int n_pad, double* __restrict__ Lb_next,
                                              int* __restrict__ flag_next, int* __restrict__ status,
                                              int bid)
{
    __shared__ __align__(16) unsigned char smem[SY_SMEM];
    double* sD = reinterpret_cast<double*>(smem);
    // workgroup 0 holds tile (0, 0), whose top-left 64x64 block (wave 0) is the next
    // panel's diagonal block
    const bool diag = bid == 0;
    int ti, tj;
    trap_tile(bid, ceil_div_dev(n_pad - t0, SY_TILE), ti, tj);
    const int ri0 = t0 + ti * SY_TILE;
    const int rj0 = t0 + tj * SY_TILE;
    floatx16 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[a][b][r] = 0.f;
    tile_acc_f32<SPLIT>(A, ld, kb, K, ri0, rj0, n_pad, smem, acc);
    if (diag) __syncthreads();   // operand stages are reused for the diagonal block
    tile_epilogue_f32(A, ld, ri0, rj0, t1, n_pad, acc, diag && (threadIdx.x >> 6) == 0, sD);
    // wavefront 0 wrote the whole diagonal block to LDS itself and starts factoring it while
    // the other wavefronts finish their stores (the factor's first barrier joins them)
    if (diag) factor_diag64(sD, t0, status, Lb_next);
    if (tj == 0) fused_trsm<float>(A, ld, t0, ri0, n_pad, Lb_next, flag_next, status, diag, smem, sD);
}

template <bool SPLIT>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3, 8)))
syrk_f32_kernel(float* __restrict__ A, long long ld, int kb, int K, int t0, int t1, int n_pad,
                double* __restrict__ Lb_next, int* __restrict__ flag_next, int* __restrict__ status)
{
    syrk_f32_body<SPLIT>(A, ld, kb, K, t0, t1, n_pad, Lb_next, flag_next, status, blockIdx.x);
}

// ------------------------------------------------------------------ FP64 SYRK (MFMA)
// Same tiling; each wave 64x64 = 4x4 blocks of 16x16 (v_mfma_f64_16x16x4_f64).  The f64
// MFMA returns row 4*r + l/16 in accumulator r of lane l; feeding the A operand rows in the
// order perm(m) = 4*(m%4) + m/4 makes accumulator r hold row 4*(l/16) + r instead, so each
// lane owns 4 consecutive rows of one column (two 16-byte stores).
constexpr int SY_KC64 = 16;

// acc += A[ri0.., kb .. kb+K) A[rj0.., kb .. kb+K)^T for one 128x128 tile on FP64 MFMA:
// each wave 64x64 = 4x4 blocks of 16x16, A-operand rows fed in the order perm(m) (above).
__device__ __forceinline__ void tile_acc_f64(const double* __restrict__ A, long long ld, int kb, int K, int ri0,
                                             int rj0, int n_pad, unsigned char* smem, doublex4 (&acc)[4][4])
{
    static_assert(2 * SY_KC64 * (SY_TILE + 2) * 8 <= SY_SMEM, "SYRK LDS");
    double (*sA)[SY_TILE + 2] = reinterpret_cast<double (*)[SY_TILE + 2]>(smem);
    double (*sB)[SY_TILE + 2] = reinterpret_cast<double (*)[SY_TILE + 2]>(smem + SY_KC64 * (SY_TILE + 2) * 8);
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wr = wave >> 1, wc = wave & 1;
    const int l16 = lane & 15, lq = lane >> 4;
    const int prow = 4 * (l16 & 3) + (l16 >> 2);
    for (int k0 = 0; k0 < K; k0 += SY_KC64) {
        __syncthreads();
        for (int idx = tid; idx < SY_KC64 * SY_TILE / 2; idx += 256) {
            const int k = idx / (SY_TILE / 2);
            const int r2 = (idx % (SY_TILE / 2)) * 2;
            const double* col = A + (long long)(kb + k0 + k) * ld;
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
}

// C -= acc for the tile (ri0, rj0) (columns < t1, rows < n_pad), or into sD (to_lds).
__device__ __forceinline__ void tile_epilogue_f64(double* __restrict__ A, long long ld, int ri0, int rj0, int t1,
                                                  int n_pad, const doublex4 (&acc)[4][4], bool to_lds,
                                                  double* __restrict__ sD)
{
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wr = wave >> 1, wc = wave & 1;
    const int l16 = lane & 15, lq = lane >> 4;
    // lane owns column l16 and rows 4*lq + (0..3) of each 16x16 block; every load of C is
    // issued before the stores
    double2 cv[4][4][2];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int col = rj0 + wc * 64 + b * 16 + l16;
            const int row = ri0 + wr * 64 + a * 16 + 4 * lq;
            const double* cp = A + (long long)col * ld + row;
            const bool in = col < t1 && row < n_pad;
            cv[a][b][0] = in ? *reinterpret_cast<const double2*>(cp) : make_double2(0.0, 0.0);
            cv[a][b][1] = in ? *reinterpret_cast<const double2*>(cp + 2) : make_double2(0.0, 0.0);
        }
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int col = rj0 + wc * 64 + b * 16 + l16;
            const int row = ri0 + wr * 64 + a * 16 + 4 * lq;
            if (col >= t1 || row >= n_pad) continue;
            double* cp = A + (long long)col * ld + row;
            double2 c0v = cv[a][b][0];
            double2 c1v = cv[a][b][1];
            c0v.x -= acc[a][b][0]; c0v.y -= acc[a][b][1];
            c1v.x -= acc[a][b][2]; c1v.y -= acc[a][b][3];
            if (to_lds) {
                const int lr = row - ri0, lc = col - rj0;
                sD[(lr + 0) * DG_LD + lc] = c0v.x;
                sD[(lr + 1) * DG_LD + lc] = c0v.y;
                sD[(lr + 2) * DG_LD + lc] = c1v.x;
                sD[(lr + 3) * DG_LD + lc] = c1v.y;
            } else {
                *reinterpret_cast<double2*>(cp) = c0v;
                *reinterpret_cast<double2*>(cp + 2) = c1v;
            }
        }
}

__device__ __forceinline__ void syrk_f64_body(double* __restrict__ A, long long ld, int kb, int K, int t0,
                                              int t1, int n_pad, double* __restrict__ Lb_next,
                                              int* __restrict__ flag_next, int* __restrict__ status,
                                              int bid)
{
    __shared__ __align__(16) unsigned char smem[SY_SMEM];
    double* sD = reinterpret_cast<double*>(smem);
    const bool diag = bid == 0;
    int ti, tj;
    trap_tile(bid, ceil_div_dev(n_pad - t0, SY_TILE), ti, tj);
    const int ri0 = t0 + ti * SY_TILE;
    const int rj0 = t0 + tj * SY_TILE;
    doublex4 acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) acc[a][b][r] = 0.0;
    tile_acc_f64(A, ld, kb, K, ri0, rj0, n_pad, smem, acc);
    if (diag) __syncthreads();   // operand stages are reused for the diagonal block
    tile_epilogue_f64(A, ld, ri0, rj0, t1, n_pad, acc, diag && (threadIdx.x >> 6) == 0, sD);
    // wavefront 0 wrote the whole diagonal block to LDS itself and starts factoring it while
    // the other wavefronts finish their stores (the factor's first barrier joins them)
    if (diag) factor_diag64(sD, t0, status, Lb_next);
    if (tj == 0) fused_trsm<double>(A, ld, t0, ri0, n_pad, Lb_next, flag_next, status, diag, smem, sD);
}

__global__ void __launch_bounds__(256)
syrk_f64_kernel(double* __restrict__ A, long long ld, int kb, int K, int t0, int t1, int n_pad,
                double* __restrict__ Lb_next, int* __restrict__ flag_next, int* __restrict__ status)
{
    syrk_f64_body(A, ld, kb, K, t0, t1, n_pad, Lb_next, flag_next, status, blockIdx.x);
}

// Two-level blocking: outer panels of CHOL_OUTER columns are factored panel by panel
// (64 columns) with updates restricted to the outer panel; the trailing matrix beyond it is
// updated once per outer panel with K = CHOL_OUTER, which divides its HBM traffic (read +
// write of the trailing lower triangle) by CHOL_OUTER / 64.
#ifndef CHOL_OUTER_COLS
#define CHOL_OUTER_COLS 256
#endif
constexpr int CHOL_OUTER = CHOL_OUTER_COLS;
static_assert(CHOL_OUTER % CHOL_NB == 0 && CHOL_OUTER >= CHOL_NB, "outer panel width");

// Update launch k (k = 0, 1, ...) of an n_pad matrix, after panel k = columns 64 k ..:
// inside an outer panel of `outer` columns, the update of the outer panel's later columns; at
// the end of an outer panel, the whole outer panel's update of the trailing matrix.  Inside
// the outer panel the schedule is either right-looking (panel k updates all of the outer
// panel's later columns, K = 64) or left-looking (all of the outer panel's panels so far
// update only the next panel's 64 columns, K = 64, 128, ...).  Every schedule and width
// leaves the same factor up to rounding.  False once no columns are left.
#ifndef CHOL_LEFT_MAX
#define CHOL_LEFT_MAX 4096
#endif
__host__ __device__ inline bool chol_step(int n_pad, bool left, int outer, int k, int& kb, int& K, int& t0,
                                          int& t1)
{
    const int c0 = k * CHOL_NB;
    t0 = c0 + CHOL_NB;
    if (t0 >= n_pad) return false;
    const int P0 = c0 / outer * outer;
    const int P1 = n_pad < P0 + outer ? n_pad : P0 + outer;
    if (t0 < P1 && left) {
        kb = P0;
        K = c0 + CHOL_NB - P0;
        t1 = t0 + CHOL_NB;
    } else if (t0 < P1) {
        kb = c0;
        K = CHOL_NB;
        t1 = P1;
    } else {
        kb = P0;
        K = P1 - P0;
        t1 = n_pad;
    }
    return true;
}

// The single-system schedule: left-looking for matrices up to CHOL_LEFT_MAX, whose launches
// are bound by the HBM traffic of the updates, right-looking beyond, where the chain of
// diagonal blocks bounds them and the longer left-looking update of the first tile would
// lengthen it.
bool chol_left_looking(int n_pad) { return n_pad <= CHOL_LEFT_MAX; }
int chol_outer_default() { return CHOL_OUTER; }

// Workgroups of an update launch: the lower-trapezoid tiles of rows t0..n_pad, columns
// t0..t1.
__host__ __device__ inline int update_tiles(int n_pad, int t0, int t1)
{
    const int ni = (n_pad - t0 + SY_TILE - 1) / SY_TILE;
    const int nj = (t1 - t0 + SY_TILE - 1) / SY_TILE;
    return nj * ni - nj * (nj - 1) / 2;
}

template <typename T>
int launch_update(T* d_A, long long ld, int kb, int K, int t0, int t1, int n_pad, double* d_Lb_next,
                  int* d_flag_next, int* d_status, int syrk_mode, hipStream_t st)
{
    const int ntiles = update_tiles(n_pad, t0, t1);
    if constexpr (sizeof(T) == 4) {
        if (syrk_mode == SYRK_BF16X3)
            hipLaunchKernelGGL(syrk_f32_kernel<true>, dim3(ntiles), dim3(256), 0, st,
                               reinterpret_cast<float*>(d_A), ld, kb, K, t0, t1, n_pad, d_Lb_next,
                               d_flag_next, d_status);
        else
            hipLaunchKernelGGL(syrk_f32_kernel<false>, dim3(ntiles), dim3(256), 0, st,
                               reinterpret_cast<float*>(d_A), ld, kb, K, t0, t1, n_pad, d_Lb_next,
                               d_flag_next, d_status);
    } else {
        hipLaunchKernelGGL(syrk_f64_kernel, dim3(ntiles), dim3(256), 0, st,
                           reinterpret_cast<double*>(d_A), ld, kb, K, t0, t1, n_pad, d_Lb_next,
                           d_flag_next, d_status);
    }
    BDH_LAUNCH_CHECK();
    return 0;
}

// ------------------------------------------------------------------ factorization by epochs
// One launch per outer panel P (columns P0 .. P1, CHOL_OUTER wide) for the large single-system
// matrices.  Each launch runs three things side by side, which the per-panel launches above
// run one after another:
//   the panel factorization of P (its 64-column panels k: diagonal block, triangular solve of
//   the rows below, right-looking update of P's later columns), on the row strips of P;
//   the lookahead: the next outer panel's columns (P1 .. Q1) receive the previous outer
//   panel's update and then each panel k's as soon as its rows are solved, so they are final
//   when the launch ends and the next launch starts its diagonal block at once;
//   the previous outer panel's update of every column beyond Q1.
// Invariant at the start of epoch P: columns P0 .. P1 carry the updates of every column below
// P0; columns beyond P1 those of every column below P0 - CHOL_OUTER.
// Roles by workgroup index (the chain's workgroups have the lowest indices, so they are
// dispatched first and all resident; every wait is for a lower index):
//   b < R            row strip s = b (rows P0 + 128 s ..); for each panel k of P, in order:
//                    the diagonal block if it lies in the strip (flag D_k) or a wait for D_k;
//                    the solve of the strip's rows below block k (flag T_k,strip); panel k's
//                    update of the strip's rows of P's later columns (right-looking, the next
//                    diagonal block first; see epoch_strip), after the flags T_k of the strips
//                    that hold those columns' rows.
//   R <= b < R + NL  a 128 x 128 tile of the lookahead columns: K = CHOL_OUTER from the previous
//                    outer panel, then K = 64 from each panel k after T_k of its two strips,
//                    accumulated in registers and applied to C once.
//   later b          a tile beyond Q1: the previous outer panel's update (K = CHOL_OUTER).
struct EpochArgs {
    int P0, P1, Q1;   // this outer panel, the end of the lookahead columns
    int prevP0;       // the previous outer panel's first column (< 0 in the first epoch)
    int R, NL, NB;    // row strips, lookahead tiles, bulk tiles
};


// Flags of the epoch schedule, after the diagonal-block flags in the workspace: T_k,g for
// panel k and global row strip g (rows 128 g ..), n_pad / 64 x n_pad / 128 of them.
__device__ __forceinline__ int* epoch_tflags(double* Lbuf, int n_pad)
{
    return chol_flags(Lbuf, n_pad) + n_pad / CHOL_NB;
}

// The epoch schedule's solve flags sit one per 128-byte line, so the tiles that poll one of
// them do not share a line with the flags the diagonal chain sets and reads.
constexpr int FLAG_STRIDE = 32;
__device__ __forceinline__ int* tflag(int* tflags, int k, int g, int gstrips)
{
    return tflags + ((long long)k * gstrips + g) * FLAG_STRIDE;
}

// The workgroup's stores so far made visible at agent scope, then *flag set.
__device__ __forceinline__ void publish_flag(int* flag)
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __syncthreads();
    if (threadIdx.x == 0) __hip_atomic_store(flag, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

// All threads: wait until *flag is set (thread 0 polls, bounded); false after a timeout, which
// is reported through status.
template <int SLEEP = 2>
__device__ __forceinline__ bool wg_wait_flag(int* flag, int* status)
{
    __shared__ int s_flag_ok;
    if (threadIdx.x == 0) {
        s_flag_ok = wait_flag<SLEEP>(flag) ? 1 : 0;
        if (!s_flag_ok) atomicCAS(status, 0, -1);
    }
    __syncthreads();
    return s_flag_ok != 0;
}

// All threads: wait until *f1 and *f2 are set (thread 0 polls both relaxed, then acquires once:
// each agent-scope acquire invalidates this XCD's L2, which the tiles around share).
template <int SLEEP = 2>
__device__ __forceinline__ bool wg_wait_flags2(int* f1, int* f2, int* status)
{
    __shared__ int s_flags_ok;
    if (threadIdx.x == 0) {
        int ok = 0;
        for (int it = 0; it < CHOL_SPIN * 2 / SLEEP; ++it) {
            if (__hip_atomic_load(f1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0 &&
                __hip_atomic_load(f2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
                ok = 1;
                break;
            }
            __builtin_amdgcn_s_sleep(SLEEP);
        }
        s_flags_ok = ok;
        if (!ok) atomicCAS(status, 0, -1);
    }
    __syncthreads();
    return s_flags_ok != 0;
}

// C[ri0.., rj0..] -= A[ri0.., kb .. kb+K) A[rj0.., kb .. kb+K)^T for one 128 x 128 tile
// (columns < t1, rows < n_pad).
template <typename T, bool SPLIT>
__device__ __forceinline__ void epoch_tile_update(T* __restrict__ A, long long ld, int kb, int K, int ri0,
                                                  int rj0, int t1, int n_pad, unsigned char* smem)
{
    if constexpr (sizeof(T) == 4) {
        floatx16 acc[2][2];
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int r = 0; r < 16; ++r) acc[a][b][r] = 0.f;
        tile_acc_f32<SPLIT>(A, ld, kb, K, ri0, rj0, n_pad, smem, acc);
        tile_epilogue_f32(A, ld, ri0, rj0, t1, n_pad, acc, false, nullptr);
    } else {
        doublex4 acc[4][4];
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b)
#pragma unroll
                for (int r = 0; r < 4; ++r) acc[a][b][r] = 0.0;
        tile_acc_f64(A, ld, kb, K, ri0, rj0, n_pad, smem, acc);
        tile_epilogue_f64(A, ld, ri0, rj0, t1, n_pad, acc, false, nullptr);
    }
}

// acc[q] = sum_k A[ra + i][kb + k] A[ca + 32 q + j][kb + k] over NQ 32 x 32 quadrants side by side
// (rows ra .., columns ca + 32 q ..), K a multiple of 32, on one wavefront's MFMA with operands
// loaded straight from A (no LDS), KC = 32 / NQ k at a time with the next KC in flight.
// v_mfma_f32_32x32x2_f32: lane l feeds row l % 32, k = l / 32; acc[q][r] holds row
// 8 (r / 4) + 4 (l / 32) + r % 4, column l % 32 of quadrant q.
template <int NQ>
__device__ __forceinline__ void quad_acc(const float* __restrict__ A, long long ld, int ra, int ca, int kb, int K,
                                         floatx16 (&acc)[NQ])
{
    constexpr int KC = 32 / NQ, NJ = KC / 2;
    const int lane = threadIdx.x & 63, l32 = lane & 31, lh = lane >> 5;
#pragma unroll
    for (int q = 0; q < NQ; ++q)
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[q][r] = 0.f;
    float av[NJ], bv[NQ][NJ];
    auto load = [&](int k0, float (&a)[NJ], float (&bq)[NQ][NJ]) {
#pragma unroll
        for (int j = 0; j < NJ; ++j) {
            const float* col = A + (long long)(kb + k0 + 2 * j + lh) * ld;
            a[j] = col[ra + l32];
#pragma unroll
            for (int q = 0; q < NQ; ++q) bq[q][j] = col[ca + 32 * q + l32];
        }
    };
    load(0, av, bv);
    for (int k0 = 0; k0 < K; k0 += KC) {
        float an[NJ], bn[NQ][NJ];
        if (k0 + KC < K) load(k0 + KC, an, bn);
#pragma unroll
        for (int j = 0; j < NJ; ++j)
#pragma unroll
            for (int q = 0; q < NQ; ++q) acc[q] = mfma_f32_32x32x2(av[j], bv[q][j], acc[q]);
        if (k0 + KC < K) {
#pragma unroll
            for (int j = 0; j < NJ; ++j) {
                av[j] = an[j];
#pragma unroll
                for (int q = 0; q < NQ; ++q) bv[q][j] = bn[q][j];
            }
        }
    }
}

// The same in FP64, K a multiple of 16: v_mfma_f64_16x16x4_f64 on the 2 x 2 tiles of each
// quadrant, lane l feeds row l % 16, k = l / 16; acc[q][i][j][r] holds row 16 i + 4 r + l / 16,
// column 16 j + l % 16 of quadrant q.  16 k at a time with the next 16 in flight.
template <int NQ>
__device__ __forceinline__ void quad_acc(const double* __restrict__ A, long long ld, int ra, int ca, int kb, int K,
                                         doublex4 (&acc)[NQ][2][2])
{
    const int lane = threadIdx.x & 63, l16 = lane & 15, lq = lane >> 4;
#pragma unroll
    for (int q = 0; q < NQ; ++q)
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) acc[q][i][j] = doublex4{0.0, 0.0, 0.0, 0.0};
    double av[4][2], bv[NQ][4][2];
    auto load = [&](int k0, double (&a)[4][2], double (&bq)[NQ][4][2]) {
#pragma unroll
        for (int s4 = 0; s4 < 4; ++s4) {
            const double* col = A + (long long)(kb + k0 + 4 * s4 + lq) * ld;
            a[s4][0] = col[ra + l16];
            a[s4][1] = col[ra + 16 + l16];
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                bq[q][s4][0] = col[ca + 32 * q + l16];
                bq[q][s4][1] = col[ca + 32 * q + 16 + l16];
            }
        }
    };
    load(0, av, bv);
    for (int k0 = 0; k0 < K; k0 += 16) {
        double an[4][2], bn[NQ][4][2];
        if (k0 + 16 < K) load(k0 + 16, an, bn);
#pragma unroll
        for (int s4 = 0; s4 < 4; ++s4)
#pragma unroll
            for (int q = 0; q < NQ; ++q)
#pragma unroll
                for (int i = 0; i < 2; ++i)
#pragma unroll
                    for (int j = 0; j < 2; ++j)
                        acc[q][i][j] = mfma_f64_16x16x4(av[s4][i], bv[q][s4][j], acc[q][i][j]);
        if (k0 + 16 < K) {
#pragma unroll
            for (int s4 = 0; s4 < 4; ++s4)
#pragma unroll
                for (int i = 0; i < 2; ++i) {
                    av[s4][i] = an[s4][i];
#pragma unroll
                    for (int q = 0; q < NQ; ++q) bv[q][s4][i] = bn[q][s4][i];
                }
        }
    }
}

// A[ra + i][ca + j] - acc of a quadrant, into dst[i * pitch + j] (LDS, FP64) or back into A
// (dst null).
__device__ __forceinline__ void quad_sub(float* __restrict__ A, long long ld, int ra, int ca, const floatx16& acc,
                                         double* __restrict__ dst, int pitch)
{
    const int lane = threadIdx.x & 63, l32 = lane & 31, lh = lane >> 5;
    float v[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int row = 8 * (r >> 2) + 4 * lh + (r & 3);
        v[r] = A[(long long)(ca + l32) * ld + ra + row] - acc[r];
    }
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int row = 8 * (r >> 2) + 4 * lh + (r & 3);
        if (dst) dst[row * pitch + l32] = (double)v[r];
        else A[(long long)(ca + l32) * ld + ra + row] = v[r];
    }
}

__device__ __forceinline__ void quad_sub(double* __restrict__ A, long long ld, int ra, int ca,
                                         const doublex4 (&acc)[2][2], double* __restrict__ dst, int pitch)
{
    const int lane = threadIdx.x & 63, l16 = lane & 15, lq = lane >> 4;
    double v[2][2][4];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r)
                v[i][j][r] = A[(long long)(ca + 16 * j + l16) * ld + ra + 16 * i + 4 * r + lq] - acc[i][j][r];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = 16 * i + 4 * r + lq, col = 16 * j + l16;
                if (dst) dst[row * pitch + col] = v[i][j][r];
                else A[(long long)(ca + col) * ld + ra + row] = v[i][j][r];
            }
}

template <typename T, int NQ>
struct QuadAcc;
template <int NQ>
struct QuadAcc<float, NQ> {
    floatx16 v[NQ];
};
template <int NQ>
struct QuadAcc<double, NQ> {
    doublex4 v[NQ][2][2];
};

// sD[r * DG_LD + c] = A[c1 + r][c1 + c] - sum_k A[c1 + r][kb + k] A[c1 + c][kb + k] for the 64 x 64
// diagonal block at c1 (the strict upper 32 x 32 quadrant is not written; the factor reads only
// r >= c), on MFMA in T: three wavefronts take the quadrants (0, 0), (1, 0), (1, 1).
template <typename T>
__device__ __forceinline__ void diag_block_update(T* __restrict__ A, long long ld, int c1, int kb, int K,
                                                  double* __restrict__ sD)
{
    const int wave = threadIdx.x >> 6;
    if (wave == 3) return;
    const int qr = wave == 0 ? 0 : 1, qc = wave == 2 ? 1 : 0;
    QuadAcc<T, 1> acc;
    quad_acc<1>(A, ld, c1 + 32 * qr, c1 + 32 * qc, kb, K, acc.v);
    quad_sub(A, ld, c1 + 32 * qr, c1 + 32 * qc, acc.v[0], sD + 32 * qr * DG_LD + 32 * qc, DG_LD);
}

// One wavefront: A[ra + i][ca + j] -= sum_k A[ra + i][kb + k] A[ca + j][kb + k] over the
// 32 x 32 NQ quadrants at rows ra .., columns ca, ca + 32, ...
template <typename T, int NQ>
__device__ __forceinline__ void wave_quads_update(T* __restrict__ A, long long ld, int ra, int ca, int kb, int K)
{
    QuadAcc<T, NQ> acc;
    quad_acc<NQ>(A, ld, ra, ca, kb, K, acc.v);
#pragma unroll
    for (int q = 0; q < NQ; ++q) quad_sub(A, ld, ra, ca + 32 * q, acc.v[q], nullptr, 0);
}

// A[rb + i][cb + j] -= sum_k A[rb + i][kb + k] A[cb + j][kb + k] for the rows rb .. rb + 32 W
// of the 64 columns cb .. (the block lies below the diagonal): wavefront W0 + w, w < W, takes
// rows rb + 32 w .. with both column quadrants.
template <typename T, int W, int W0 = 0>
__device__ __forceinline__ void block_update64(T* __restrict__ A, long long ld, int rb, int cb, int kb, int K)
{
    const int wave = (threadIdx.x >> 6) - W0;
    if (wave < 0 || wave >= W) return;
    const int ra = rb + 32 * wave;
    QuadAcc<T, 2> acc;
    quad_acc<2>(A, ld, ra, cb, kb, K, acc.v);
    quad_sub(A, ld, ra, cb, acc.v[0], nullptr, 0);
    quad_sub(A, ld, ra, cb + 32, acc.v[1], nullptr, 0);
}

// Row strip s of outer panel P (see above), right-looking: after solving its rows of panel k
// the strip applies panel k to its rows of the later columns of P.  The diagonal chain comes
// first and stays in LDS.  A strip holding the next two diagonal blocks kA (rows r0 ..) and
// kB = kA + 1:
//   panel kA - 1: only block kA's update (64 x 64, K = 64) is made, into LDS; the strip's lower
//                 block row (rows r0 + 64 ..) keeps panel kA - 1 pending;
//   block kA:     factor, publish D_kA, then the pending panel's update of block (kB, kA)
//                 (64 x 64, into A), the solve of rows r0 + 64 .. with L_kA, and block kB's
//                 update with panels kA - 1 and kA (K = 128) into LDS;
//   block kB:     factor, publish D_kB.
// Only the epoch's first block is read from A.
template <typename T, bool SPLIT>
__device__ __forceinline__ void epoch_strip(T* __restrict__ A, long long ld, int n_pad, double* __restrict__ Lbuf,
                                            int* __restrict__ status, const EpochArgs& e, int s,
                                            unsigned char* smem)
{
    int* dflags = chol_flags(Lbuf, n_pad);
    int* tflags = epoch_tflags(Lbuf, n_pad);
    const int gstrips = n_pad / SY_TILE;
    const int r0 = e.P0 + SY_TILE * s, r1 = min(r0 + SY_TILE, n_pad);
    const int g = r0 / SY_TILE;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    double* sD = reinterpret_cast<double*>(smem);
    double* sLT = sD;
    double* sInv = sLT + CHOL_NB * CHOL_NB;
    double* sc = sInv + 4 * 256 + wave * 16 * SC_PITCH;
    bool block_in_lds = false;   // sD holds the next diagonal block, fully updated
    int pend = -1;   // first column of panel kA - 1 while it is pending on the lower block row
    int kdef = -1;   // first column of panel kA - 2 while it is deferred on block kB
    for (int c0 = e.P0; c0 < e.P1 && c0 < r1; c0 += CHOL_NB) {
        const int k = c0 / CHOL_NB;
        double* Lb = Lbuf + (long long)k * LB_STRIDE;
        const bool diag = c0 >= r0;   // block k lies in this strip
        int rs = r0;                  // first row of the strip solved with panel k
        if (diag) {
            if (!block_in_lds) {
                // the epoch's first block: it carries every update (earlier launches'); all 16
                // loads of a thread are issued before its LDS stores
                __syncthreads();
                T v[CHOL_NB * CHOL_NB / 256];
#pragma unroll
                for (int i = 0; i < CHOL_NB * CHOL_NB / 256; ++i) {
                    const int e2 = tid + 256 * i, r = e2 % CHOL_NB, c = e2 / CHOL_NB;
                    v[i] = A[(long long)(c0 + c) * ld + c0 + r];
                }
#pragma unroll
                for (int i = 0; i < CHOL_NB * CHOL_NB / 256; ++i) {
                    const int e2 = tid + 256 * i, r = e2 % CHOL_NB, c = e2 / CHOL_NB;
                    sD[r * DG_LD + c] = (r >= c) ? (double)v[i] : 0.0;
                }
            }
            rs = c0 + CHOL_NB;
            // Made in the factor's idle stretches (the operand rows were solved by this
            // workgroup, behind a barrier): the pending panel on block (kB, kA), which the solve
            // below reads (wavefront 3, before its columns), and the deferred and pending panels
            // on block kB, so that its update after the solve takes panel kA alone (quadrants
            // (1, 0), (1, 1) by wavefront 2 before its columns, (0, 0) by wavefront 0 after).
            // In FP64 (slower MFMA, shorter idle stretches to fill) only block (kB, kA) is made
            // there, by wavefronts 2 and 3, and block kB takes its panels after the solve.
            const bool lower = pend >= 0 && rs < r1;
            const int klo = kdef >= 0 ? kdef : pend, kK = lower ? pend + CHOL_NB - klo : 0;
            factor_diag64(sD, c0, status, Lb,
                [&] {
                    if (!lower) return;
                    if constexpr (sizeof(T) == 4) {
                        if (wave == 3) {
                            wave_quads_update<T, 2>(A, ld, rs, c0, pend, CHOL_NB);
                            wave_quads_update<T, 2>(A, ld, rs + 32, c0, pend, CHOL_NB);
                        } else if (wave == 2) {
                            wave_quads_update<T, 2>(A, ld, rs + 32, rs, klo, kK);
                        }
                    } else {
                        block_update64<T, 2, 2>(A, ld, rs, c0, pend, CHOL_NB);
                    }
                },
                [&] {
                    if constexpr (sizeof(T) == 4)
                        if (lower && wave == 0) wave_quads_update<T, 1>(A, ld, rs, rs, klo, kK);
                });                              // begins with a barrier; leaves sLT / sInv
            publish_flag(dflags + k);            // Lb (the stores of store_diag64 are not read here)
            store_diag64<T>(sD, A, ld, c0);
        }
        block_in_lds = false;
        if (rs < r1) {
            // rows rs .. r1 of panel k, two 16-row groups per wavefront; the rows carry every
            // update once this workgroup's stores to them are done (a barrier), and are loaded
            // before the wait for D_k
            __syncthreads();
            const int g0 = rs + 16 * wave, g1 = g0 + 64;
            Trsm16 x0, x1;
            if (g0 < r1) trsm16_load<T>(A, ld, c0, g0, r1, lane, x0);
            if (g1 < r1) trsm16_load<T>(A, ld, c0, g1, r1, lane, x1);
            else x1 = Trsm16{};
            if (!diag) {
                if (!wg_wait_flag(dflags + k, status)) return;   // a barrier: smem is free
                load_lkk(Lb, sLT, sInv);    // ends with a barrier
            }
            if (g1 < r1) trsm16x2_solve_store<T>(A, ld, c0, g0, g1, r1, lane, x0, x1, sLT, sInv, sc);
            else if (g0 < r1) trsm16_solve_store<T>(A, ld, c0, g0, r1, lane, x0, sLT, sInv, sc);
        }
        publish_flag(tflag(tflags, k, g, gstrips));
        // panel k on the strip's rows of columns cb .. ce of P
        const int cb = c0 + CHOL_NB, ce = min(e.P1, r1);
        if (cb >= ce) continue;
        if (diag) {
            // block kA was factored: below the strip's diagonal the later columns hold block kB
            // alone, which takes panel kA into LDS (FP32: the earlier panels came in the
            // factor; FP64: with the pending panel, contiguous)
            __syncthreads();   // the solve's stores done; smem is free
            const int kb = sizeof(T) == 4 || pend < 0 ? c0 : pend;
            diag_block_update<T>(A, ld, cb, kb, cb - kb, sD);
            block_in_lds = true;
            pend = -1;
            kdef = -1;
        } else if (cb == r0) {
            // the next panel is block kA: its update (the strip's own rows of panel k) into LDS,
            // the lower block row's left pending
            __syncthreads();
            diag_block_update<T>(A, ld, cb, c0, CHOL_NB, sD);
            block_in_lds = true;
            pend = c0;
        } else {
            // on a strip holding diagonal blocks, FP32 leaves block kB's columns to its
            // factor's idle stretches (kdef)
            const int cu = sizeof(T) == 4 && r0 < e.P1 ? min(ce, r0 + CHOL_NB) : ce;
            if (cu < ce) kdef = c0;
            if constexpr (sizeof(T) == 4) {
                // tiles of up to 128 columns (LDS-staged, faster in FP32 than the pieces below),
                // after the flags T_k of the strips holding rows cb .. cu
                for (int gg = cb / SY_TILE; gg <= (cu - 1) / SY_TILE; ++gg)
                    if (gg != g && !wg_wait_flag(tflag(tflags, k, gg, gstrips), status)) return;
                for (int rj0 = cb; rj0 < cu; rj0 += SY_TILE) {
                    __syncthreads();   // smem is free (the solve's operands, the previous tile's stages)
                    floatx16 acc[2][2];
#pragma unroll
                    for (int a2 = 0; a2 < 2; ++a2)
#pragma unroll
                        for (int b2 = 0; b2 < 2; ++b2)
#pragma unroll
                            for (int r = 0; r < 16; ++r) acc[a2][b2][r] = 0.f;
                    tile_acc_f32<SPLIT>(A, ld, c0, CHOL_NB, r0, rj0, n_pad, smem, acc);
                    tile_epilogue_f32(A, ld, r0, rj0, cu, n_pad, acc, false, nullptr);
                }
            } else {
                // 64-column pieces in column order (the next panel's first), each after the
                // flag T_k of the strip holding its rows; a piece on the strip's own diagonal
                // blocks takes only the rows from its diagonal down
                __syncthreads();   // the solve's stores done
                for (int j = cb; j < cu; j += CHOL_NB) {
                    const int gg = j / SY_TILE;
                    if (gg != g && !wg_wait_flag(tflag(tflags, k, gg, gstrips), status)) return;
                    if (j < r0) block_update64<T, 4>(A, ld, r0, j, c0, CHOL_NB);
                    else if (j == r0) block_update64<T, 4>(A, ld, j, j, c0, CHOL_NB);
                    else block_update64<T, 2>(A, ld, j, j, c0, CHOL_NB);
                }
            }
        }
    }
}

// Lookahead tile x (rows from P1, columns P1 .. Q1).
template <typename T, bool SPLIT>
__device__ __forceinline__ void epoch_lookahead(T* __restrict__ A, long long ld, int n_pad,
                                                double* __restrict__ Lbuf, int* __restrict__ status,
                                                const EpochArgs& e, int x, unsigned char* smem)
{
    int* tflags = epoch_tflags(Lbuf, n_pad);
    const int gstrips = n_pad / SY_TILE;
    int ti, tj;
    trap_tile(x, ceil_div_dev(n_pad - e.P1, SY_TILE), ti, tj);
    const int ri0 = e.P1 + SY_TILE * ti, rj0 = e.P1 + SY_TILE * tj;
    const int gi = ri0 / SY_TILE, gj = rj0 / SY_TILE;
    if constexpr (sizeof(T) == 4) {
        floatx16 acc[2][2];
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int r = 0; r < 16; ++r) acc[a][b][r] = 0.f;
        if (e.prevP0 >= 0) tile_acc_f32<SPLIT>(A, ld, e.prevP0, e.P0 - e.prevP0, ri0, rj0, n_pad, smem, acc);
        for (int c0 = e.P0; c0 < e.P1; c0 += CHOL_NB) {
            const int k = c0 / CHOL_NB;
            if (!wg_wait_flags2<16>(tflag(tflags, k, gi, gstrips), tflag(tflags, k, gj, gstrips), status)) return;
            tile_acc_f32<SPLIT>(A, ld, c0, CHOL_NB, ri0, rj0, n_pad, smem, acc);
        }
        tile_epilogue_f32(A, ld, ri0, rj0, e.Q1, n_pad, acc, false, nullptr);
    } else {
        doublex4 acc[4][4];
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b)
#pragma unroll
                for (int r = 0; r < 4; ++r) acc[a][b][r] = 0.0;
        if (e.prevP0 >= 0) tile_acc_f64(A, ld, e.prevP0, e.P0 - e.prevP0, ri0, rj0, n_pad, smem, acc);
        for (int c0 = e.P0; c0 < e.P1; c0 += CHOL_NB) {
            const int k = c0 / CHOL_NB;
            if (!wg_wait_flags2<16>(tflag(tflags, k, gi, gstrips), tflag(tflags, k, gj, gstrips), status)) return;
            tile_acc_f64(A, ld, c0, CHOL_NB, ri0, rj0, n_pad, smem, acc);
        }
        tile_epilogue_f64(A, ld, ri0, rj0, e.Q1, n_pad, acc, false, nullptr);
    }
}

// Tile (ti, tj), ti >= tj, of the lower triangle of an ni x ni tile matrix at position x of an
// order that walks 8 x 8 super-tiles (column-major over the super-tiles, column-major inside
// an off-diagonal one, trap_tile's order inside a diagonal one).
__device__ __forceinline__ void super_tile(int x, int ni, int& ti, int& tj)
{
    constexpr int S = 8;
    const int ns = (ni + S - 1) / S;
    for (int SJ = 0; SJ < ns; ++SJ) {
        const int c0 = S * SJ, mc = min(S, ni - c0);
        const int dc = mc * (mc + 1) / 2;
        if (x < dc) {
            trap_tile(x, mc, ti, tj);
            ti += c0;
            tj += c0;
            return;
        }
        x -= dc;
        const int below = (ni - c0 - mc) * mc;   // the off-diagonal super-tiles of this column
        if (x < below) {
            const int blk = x / (S * mc), rem = x - blk * S * mc;
            const int r0 = c0 + S * (blk + 1), mr = min(S, ni - r0);
            ti = r0 + rem % mr;
            tj = c0 + rem / mr;
            return;
        }
        x -= below;
    }
    ti = tj = 0;
}

template <typename T, bool SPLIT>
__device__ __forceinline__ void chol_epoch_body(T* __restrict__ A, long long ld, int n_pad,
                                                double* __restrict__ Lbuf, int* __restrict__ status,
                                                EpochArgs e)
{
    __shared__ __align__(16) unsigned char smem[SY_SMEM];
    const int b = blockIdx.x;
    if (b < e.R) {
        epoch_strip<T, SPLIT>(A, ld, n_pad, Lbuf, status, e, b, smem);
    } else if (b < e.R + e.NL) {
        epoch_lookahead<T, SPLIT>(A, ld, n_pad, Lbuf, status, e, b - e.R, smem);
    } else {
        // workgroups b, b + 8, ... run on one XCD: give each XCD a contiguous run of the
        // super-tile order, so the tiles it runs at a time share their operand row blocks in
        // its L2
        const int xb = b - e.R - e.NL;
        const int c = xb & 7, q = e.NB >> 3, r = e.NB & 7;
        int ti, tj;
        super_tile(c * q + min(c, r) + (xb >> 3), ceil_div_dev(n_pad - e.Q1, SY_TILE), ti, tj);
        epoch_tile_update<T, SPLIT>(A, ld, e.prevP0, e.P0 - e.prevP0, e.Q1 + SY_TILE * ti, e.Q1 + SY_TILE * tj,
                                    n_pad, n_pad, smem);
    }
}

// Two waves per SIMD: the row strips hold the diagonal factor's, the solve's and the tile
// update's registers in one kernel.
template <bool SPLIT>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 8)))
chol_epoch_f32_kernel(float* __restrict__ A, long long ld, int n_pad, double* __restrict__ Lbuf,
                      int* __restrict__ status, EpochArgs e)
{
    chol_epoch_body<float, SPLIT>(A, ld, n_pad, Lbuf, status, e);
}

__global__ void __launch_bounds__(256)
chol_epoch_f64_kernel(double* __restrict__ A, long long ld, int n_pad, double* __restrict__ Lbuf,
                      int* __restrict__ status, EpochArgs e)
{
    chol_epoch_body<double, false>(A, ld, n_pad, Lbuf, status, e);
}

// Clears the diagonal-block and solve flags of one factorization (a kernel rather than a
// memset node, so a captured factorization replays it correctly).
__global__ void chol_epoch_clear_kernel(int* __restrict__ flags, int count)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) flags[i] = 0;
}

// Matrices at least this large factor by epochs; smaller ones (one or two outer panels, where
// nothing is left to overlap) by the per-panel launches.
#ifndef CHOL_EPOCH_MIN
#define CHOL_EPOCH_MIN 1024
#endif

template <typename T>
int chol_factor_epochs(T* d_A, long long ld, int n_pad, double* d_Lbuf, int* d_status, int syrk_mode,
                       hipStream_t st, PhaseRecorder* rec)
{
    const int nblk = n_pad / CHOL_NB, gstrips = n_pad / SY_TILE;
    int* d_flags = reinterpret_cast<int*>(d_Lbuf + (long long)nblk * LB_STRIDE);
    const int nflags = nblk + nblk * gstrips * FLAG_STRIDE;
    if (rec) rec->begin(PH_PANEL, st);
    hipLaunchKernelGGL(chol_epoch_clear_kernel, dim3(ceil_div(nflags, 256)), dim3(256), 0, st, d_flags, nflags);
    BDH_LAUNCH_CHECK();
    if (rec) rec->end(PH_PANEL, st);
    for (int P0 = 0; P0 < n_pad; P0 += CHOL_OUTER) {
        EpochArgs e;
        e.P0 = P0;
        e.P1 = std::min(P0 + CHOL_OUTER, n_pad);
        e.Q1 = std::min(e.P1 + CHOL_OUTER, n_pad);
        e.prevP0 = P0 - CHOL_OUTER;
        e.R = (n_pad - P0) / SY_TILE;
        e.NL = e.Q1 > e.P1 ? update_tiles(n_pad, e.P1, e.Q1) : 0;
        e.NB = (e.prevP0 >= 0 && e.Q1 < n_pad) ? update_tiles(n_pad, e.Q1, n_pad) : 0;
        const dim3 grid(e.R + e.NL + e.NB);
        if (rec) rec->begin(PH_SYRK, st);
        if constexpr (sizeof(T) == 4) {
            if (syrk_mode == SYRK_BF16X3)
                hipLaunchKernelGGL(chol_epoch_f32_kernel<true>, grid, dim3(256), 0, st, reinterpret_cast<float*>(d_A),
                                   ld, n_pad, d_Lbuf, d_status, e);
            else
                hipLaunchKernelGGL(chol_epoch_f32_kernel<false>, grid, dim3(256), 0, st, reinterpret_cast<float*>(d_A),
                                   ld, n_pad, d_Lbuf, d_status, e);
        } else {
            hipLaunchKernelGGL(chol_epoch_f64_kernel, grid, dim3(256), 0, st, reinterpret_cast<double*>(d_A), ld,
                               n_pad, d_Lbuf, d_status, e);
        }
        BDH_LAUNCH_CHECK();
        if (rec) rec->end(PH_SYRK, st);
    }
    return 0;
}

// Matrices of CHOL_EPOCH_MIN and more factor by epochs (above).  Smaller ones: one launch per
// panel, the update kernel of panel k also factoring diagonal block k + 1 (workgroup 0) and
// solving panel k + 1 (its tile column 0, after workgroup 0's flag).
template <typename T>
int chol_factor(T* d_A, long long ld, int n_pad, double* d_Lbuf, int* d_status, int syrk_mode,
                hipStream_t st, PhaseRecorder* rec)
{
    BDH_CHECK(n_pad % PAD_TO == 0, "chol_factor: n_pad must be a multiple of 128");
    BDH_CHECK(ld >= n_pad && ld % 4 == 0, "chol_factor: bad leading dimension");
    BDH_CHECK(syrk_mode != SYRK_BF16X3 || sizeof(T) == 4, "chol_factor: split SYRK needs FP32 storage");
    if (n_pad >= CHOL_EPOCH_MIN) return chol_factor_epochs<T>(d_A, ld, n_pad, d_Lbuf, d_status, syrk_mode, st, rec);
    constexpr long long LB = LB_STRIDE;
    const int nblk = n_pad / CHOL_NB;
    int* d_flags = reinterpret_cast<int*>(d_Lbuf + nblk * LB);
    if (rec) rec->begin(PH_PANEL, st);
    hipLaunchKernelGGL(chol_diag_kernel<T>, dim3(1), dim3(256), 0, st, d_A, ld, 0, d_Lbuf, d_flags,
                       nblk, d_status);
    BDH_LAUNCH_CHECK();
    if (n_pad > CHOL_NB) {
        hipLaunchKernelGGL(chol_trsm_kernel<T>, dim3(ceil_div(n_pad - CHOL_NB, TRSM_ROWS)), dim3(256),
                           0, st, d_A, ld, 0, n_pad, (const double*)d_Lbuf);
        BDH_LAUNCH_CHECK();
    }
    if (rec) rec->end(PH_PANEL, st);
    int kb, K, t0, t1;
    const bool left = chol_left_looking(n_pad);
    for (int k = 0; chol_step(n_pad, left, CHOL_OUTER, k, kb, K, t0, t1); ++k) {
        if (rec) rec->begin(PH_SYRK, st);
        const int rc = launch_update<T>(d_A, ld, kb, K, t0, t1, n_pad, d_Lbuf + (k + 1) * LB,
                                        d_flags + k + 1, d_status, syrk_mode, st);
        if (rc) return rc;
        if (rec) rec->end(PH_SYRK, st);
    }
    return 0;
}

size_t chol_lbuf_doubles(int n_pad)
{
    const size_t nblk = n_pad / CHOL_NB, gstrips = n_pad / SY_TILE;
    // FP64 blocks with inverses, then the int flags: one per diagonal block, then (epochs
    // only) the solve flags, one per panel and row strip, each on its own 128-byte line
    if (n_pad < CHOL_EPOCH_MIN) return nblk * LB_STRIDE + (nblk + 1) / 2;
    return nblk * LB_STRIDE + (nblk + nblk * gstrips * FLAG_STRIDE + 1) / 2;
}

// ------------------------------------------------------------------ Y = L X (partials)
// X holds m right-hand sides (m <= M, X[k * ldx + c]): with L fixed between refactorizations,
// the noise of a whole refactor cycle is one pass over L instead of one per step.
// Tile: (64 lanes x V rows) x TRMV_COLS columns, V = 16 / sizeof(T) so each lane issues one
// 16-byte load per column; the 4 waves split the tile's columns (unrolled for loads in
// flight) and meet in LDS.  Entries above the diagonal are masked with a select after the
// load (the storage there is finite but meaningless), so diagonal tiles take the same
// vector path.  P[cb][k][r] receives row r's partial of right-hand side k over column
// block cb (FP64 products and accumulation).
constexpr int TRMV_COLS = 256;

template <typename T, int M>
__device__ __forceinline__ void trmv_accum(const T* __restrict__ col, int rbase, int c,
                                           const double* __restrict__ xc, bool mask,
                                           double (*acc)[M])
{
    constexpr int V = 16 / sizeof(T);
    double v[V];
    if constexpr (V == 4) {
        const float4 w = *reinterpret_cast<const float4*>(col + rbase);
        v[0] = (!mask || rbase + 0 >= c) ? (double)w.x : 0.0;
        v[1] = (!mask || rbase + 1 >= c) ? (double)w.y : 0.0;
        v[2] = (!mask || rbase + 2 >= c) ? (double)w.z : 0.0;
        v[3] = (!mask || rbase + 3 >= c) ? (double)w.w : 0.0;
    } else {
        const double2 w = *reinterpret_cast<const double2*>(col + rbase);
        v[0] = (!mask || rbase + 0 >= c) ? w.x : 0.0;
        v[1] = (!mask || rbase + 1 >= c) ? w.y : 0.0;
    }
#pragma unroll
    for (int q = 0; q < V; ++q)
#pragma unroll
        for (int k = 0; k < M; ++k) acc[q][k] += v[q] * xc[k];
}

template <typename T, int M>
__device__ __forceinline__ void trmv_partial_body(const T* __restrict__ L, long long ld, int n,
                                                  const double* __restrict__ X, long long ldx, int m,
                                                  double* __restrict__ P, int rb, int cb)
{
    constexpr int V = 16 / sizeof(T);
    constexpr int ROWS = 64 * V;
    constexpr int UNR = M >= 8 ? 4 : (M >= 4 ? 4 : 8);
    const int r0 = rb * ROWS, cbeg = cb * TRMV_COLS;
    if (cbeg > r0 + ROWS - 1 || cbeg >= n) return;  // tile entirely above the diagonal
    __shared__ double s_red[4][ROWS];
    // the wavefront index is uniform: its columns, and so the X values, come through the
    // scalar cache into SGPRs instead of one vector load per lane
    const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int rbase = r0 + lane * V;
    double acc[V][M];
#pragma unroll
    for (int q = 0; q < V; ++q)
#pragma unroll
        for (int k = 0; k < M; ++k) acc[q][k] = 0.0;
    const int cend = min(n, cbeg + TRMV_COLS);
    const int wcols = TRMV_COLS / 4;
    const int c_lo = cbeg + wave * wcols, c_hi = min(cend, c_lo + wcols);
    // the whole tile is below the diagonal when its last column precedes its first row
    const bool mask = (cbeg + TRMV_COLS - 1) > r0;
    if (rbase < n) {
        int c = c_lo;
        for (; c + UNR <= c_hi; c += UNR) {
            double xc[UNR][M];
#pragma unroll
            for (int u = 0; u < UNR; ++u)
#pragma unroll
                for (int k = 0; k < M; ++k) xc[u][k] = (k < m) ? X[k * ldx + c + u] : 0.0;
#pragma unroll
            for (int u = 0; u < UNR; ++u)
                trmv_accum<T, M>(L + (long long)(c + u) * ld, rbase, c + u, xc[u], mask, acc);
        }
        for (; c < c_hi; ++c) {
            double xc[M];
#pragma unroll
            for (int k = 0; k < M; ++k) xc[k] = (k < m) ? X[k * ldx + c] : 0.0;
            trmv_accum<T, M>(L + (long long)c * ld, rbase, c, xc, mask, acc);
        }
    }
    for (int k = 0; k < m; ++k) {
        __syncthreads();
#pragma unroll
        for (int q = 0; q < V; ++q) {
            double a = 0.0;
#pragma unroll
            for (int kk = 0; kk < M; ++kk) a = (kk == k) ? acc[q][kk] : a;
            s_red[wave][lane * V + q] = a;
        }
        __syncthreads();
        for (int rr = tid; rr < ROWS; rr += 256) {
            const int row = r0 + rr;
            if (row < n)
                P[((long long)cb * m + k) * n + row] =
                    (s_red[0][rr] + s_red[1][rr]) + (s_red[2][rr] + s_red[3][rr]);
        }
    }
}

template <typename T, int M>
__global__ void __launch_bounds__(256)
trmv_partial_kernel(const T* __restrict__ L, long long ld, int n, const double* __restrict__ X,
                    long long ldx, int m, double* __restrict__ P)
{
    trmv_partial_body<T, M>(L, ld, n, X, ldx, m, P, blockIdx.x, blockIdx.y);
}

int trmv_col_blocks(int n) { return ceil_div(n, TRMV_COLS); }

template <typename T>
int launch_trmv_partials(const T* d_L, long long ld, int n, const double* d_X, long long ldx, int m,
                         double* d_P, hipStream_t st)
{
    constexpr int ROWS = 64 * (16 / sizeof(T));
    BDH_CHECK(ld % (16 / sizeof(T)) == 0, "trmv: leading dimension not 16-byte aligned");
    BDH_CHECK(m >= 1 && m <= TRMV_MAX_RHS, "trmv: 1..16 right-hand sides");
    dim3 grid(ceil_div(n, ROWS), trmv_col_blocks(n));
    if (m == 1)
        hipLaunchKernelGGL((trmv_partial_kernel<T, 1>), grid, dim3(256), 0, st, d_L, ld, n, d_X, ldx, m, d_P);
    else if (m == 2)
        hipLaunchKernelGGL((trmv_partial_kernel<T, 2>), grid, dim3(256), 0, st, d_L, ld, n, d_X, ldx, m, d_P);
    else if (m <= 4)
        hipLaunchKernelGGL((trmv_partial_kernel<T, 4>), grid, dim3(256), 0, st, d_L, ld, n, d_X, ldx, m, d_P);
    else if (m <= 8)
        hipLaunchKernelGGL((trmv_partial_kernel<T, 8>), grid, dim3(256), 0, st, d_L, ld, n, d_X, ldx, m, d_P);
    else if (m <= 10)
        hipLaunchKernelGGL((trmv_partial_kernel<T, 10>), grid, dim3(256), 0, st, d_L, ld, n, d_X, ldx, m, d_P);
    else
        hipLaunchKernelGGL((trmv_partial_kernel<T, 16>), grid, dim3(256), 0, st, d_L, ld, n, d_X, ldx, m, d_P);
    BDH_LAUNCH_CHECK();
    return 0;
}

int trmv_blk_cols() { return TRMV_COLS; }

// ------------------------------------------------------------------ batches of matrices
// Launch k of every chain's factorization runs as one grid: workgroup x maps to its chain
// through the prefix table of launch k and then does exactly what workgroup
// x - prefix[c] of the chain's own launch k does (chol_factor), on the chain's matrix,
// workspace, flags and status.  A chain's diagonal workgroup (local index 0) has the lowest
// index of the chain's workgroups, so it is dispatched before any workgroup that waits for
// its flag, as in the single-matrix launch.  Chains whose factorization has fewer launches
// drop out of the later grids.
int chol_update_steps(int n_pad)
{
    int kb, K, t0, t1, k = 0;
    while (chol_step(n_pad, true, CHOL_OUTER, k, kb, K, t0, t1)) ++k;
    return k;
}

int chol_update_blocks(int n_pad, bool left, int outer, int k)
{
    int kb, K, t0, t1;
    return chol_step(n_pad, left, outer, k, kb, K, t0, t1) ? update_tiles(n_pad, t0, t1) : 0;
}

int chol_trsm0_blocks(int n_pad) { return n_pad > CHOL_NB ? ceil_div(n_pad - CHOL_NB, TRSM_ROWS) : 0; }

template <typename T>
int trmv_blocks(int n)
{
    constexpr int ROWS = 64 * (16 / sizeof(T));
    return ceil_div(n, ROWS) * trmv_col_blocks(n);
}

template <typename T>
__global__ void __launch_bounds__(256)
chol_diag_batch_kernel(T* __restrict__ A, double* __restrict__ Lbuf, const BatchChain* __restrict__ ch,
                       int* __restrict__ status)
{
    const int c = blockIdx.x;
    const BatchChain b = ch[c];
    double* Lb = Lbuf + b.lb;
    chol_diag_body<T>(A + b.mat, b.n_pad, 0, Lb, chol_flags(Lb, b.n_pad), b.n_pad / CHOL_NB, status + c);
}

template <typename T>
__global__ void __launch_bounds__(256)
chol_trsm_batch_kernel(T* __restrict__ A, const double* __restrict__ Lbuf,
                       const BatchChain* __restrict__ ch, const int* __restrict__ prefix, int nch)
{
    const int c = batch_find(prefix, nch, blockIdx.x);
    const BatchChain b = ch[c];
    chol_trsm_body<T>(A + b.mat, b.n_pad, 0, b.n_pad, Lbuf + b.lb, blockIdx.x - prefix[c]);
}

template <bool SPLIT>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3, 8)))
syrk_f32_batch_kernel(float* __restrict__ A, double* __restrict__ Lbuf, const BatchChain* __restrict__ ch,
                      const int* __restrict__ prefix, int nch, int k, int* __restrict__ status)
{
    const int c = batch_find(prefix, nch, blockIdx.x);
    const BatchChain b = ch[c];
    int kb, K, t0, t1;
    chol_step(b.n_pad, b.left != 0, b.outer, k, kb, K, t0, t1);
    double* Lb = Lbuf + b.lb;
    syrk_f32_body<SPLIT>(A + b.mat, b.n_pad, kb, K, t0, t1, b.n_pad, Lb + (long long)(k + 1) * LB_STRIDE,
                         chol_flags(Lb, b.n_pad) + k + 1, status + c, blockIdx.x - prefix[c]);
}

__global__ void __launch_bounds__(256)
syrk_f64_batch_kernel(double* __restrict__ A, double* __restrict__ Lbuf, const BatchChain* __restrict__ ch,
                      const int* __restrict__ prefix, int nch, int k, int* __restrict__ status)
{
    const int c = batch_find(prefix, nch, blockIdx.x);
    const BatchChain b = ch[c];
    int kb, K, t0, t1;
    chol_step(b.n_pad, b.left != 0, b.outer, k, kb, K, t0, t1);
    double* Lb = Lbuf + b.lb;
    syrk_f64_body(A + b.mat, b.n_pad, kb, K, t0, t1, b.n_pad, Lb + (long long)(k + 1) * LB_STRIDE,
                  chol_flags(Lb, b.n_pad) + k + 1, status + c, blockIdx.x - prefix[c]);
}

template <typename T>
int chol_factor_batch(T* d_A, double* d_Lbuf, const BatchPlan& plan, int* d_status, int syrk_mode,
                      hipStream_t st, PhaseRecorder* rec)
{
    BDH_CHECK(syrk_mode != SYRK_BF16X3 || sizeof(T) == 4, "chol_factor_batch: split SYRK needs FP32 storage");
    BDH_CHECK(sizeof(T) == 8 || syrk_mode != SYRK_F64, "chol_factor_batch: FP32 storage needs an FP32 SYRK");
    if (rec) rec->begin(PH_PANEL, st);
    hipLaunchKernelGGL(chol_diag_batch_kernel<T>, dim3(plan.nch), dim3(256), 0, st, d_A, d_Lbuf, plan.ch,
                       d_status);
    BDH_LAUNCH_CHECK();
    if (plan.trsm_blocks > 0) {
        hipLaunchKernelGGL(chol_trsm_batch_kernel<T>, dim3(plan.trsm_blocks), dim3(256), 0, st, d_A,
                           (const double*)d_Lbuf, plan.ch, plan.trsm_pre, plan.nch);
        BDH_LAUNCH_CHECK();
    }
    if (rec) rec->end(PH_PANEL, st);
    for (int k = 0; k < plan.chol_steps; ++k) {
        const int blocks = plan.chol_blocks[k];
        if (blocks == 0) continue;
        const int* pre = plan.chol_pre + (long long)k * (plan.nch + 1);
        if (rec) rec->begin(PH_SYRK, st);
        if constexpr (sizeof(T) == 4) {
            if (syrk_mode == SYRK_BF16X3)
                hipLaunchKernelGGL(syrk_f32_batch_kernel<true>, dim3(blocks), dim3(256), 0, st,
                                   reinterpret_cast<float*>(d_A), d_Lbuf, plan.ch, pre, plan.nch, k, d_status);
            else
                hipLaunchKernelGGL(syrk_f32_batch_kernel<false>, dim3(blocks), dim3(256), 0, st,
                                   reinterpret_cast<float*>(d_A), d_Lbuf, plan.ch, pre, plan.nch, k, d_status);
        } else {
            hipLaunchKernelGGL(syrk_f64_batch_kernel, dim3(blocks), dim3(256), 0, st,
                               reinterpret_cast<double*>(d_A), d_Lbuf, plan.ch, pre, plan.nch, k, d_status);
        }
        BDH_LAUNCH_CHECK();
        if (rec) rec->end(PH_SYRK, st);
    }
    return 0;
}

template <typename T, int M>
__global__ void __launch_bounds__(256)
trmv_batch_kernel(const T* __restrict__ A, const double* __restrict__ X, long long x_stride, int m,
                  const BatchChain* __restrict__ ch, const int* __restrict__ prefix, int nch,
                  double* __restrict__ Pt)
{
    constexpr int ROWS = 64 * (16 / sizeof(T));
    const int c = batch_find(prefix, nch, blockIdx.x);
    const BatchChain b = ch[c];
    const int x = blockIdx.x - prefix[c];
    const int gx = (b.n + ROWS - 1) / ROWS;
    const double* Xc = x_stride > 0 ? X + 3LL * b.bead0 : X + b.xi_off;
    const long long ldx = x_stride > 0 ? x_stride : b.ldx;
    trmv_partial_body<T, M>(A + b.mat, b.n_pad, b.n, Xc, ldx, m, Pt + b.pt, x % gx, x / gx);
}

template <typename T>
int launch_trmv_batch(const T* d_A, const double* d_X, long long x_stride, int m, const BatchPlan& plan,
                      double* d_Pt, hipStream_t st)
{
    BDH_CHECK(m >= 1 && m <= TRMV_MAX_RHS, "trmv: 1..16 right-hand sides");
    const dim3 grid(plan.trmv_blocks);
#define TB_LAUNCH(MM)                                                                              \
    hipLaunchKernelGGL((trmv_batch_kernel<T, MM>), grid, dim3(256), 0, st, d_A, d_X, x_stride, m, \
                       plan.ch, plan.trmv_pre, plan.nch, d_Pt)
    if (m == 1) TB_LAUNCH(1);
    else if (m == 2) TB_LAUNCH(2);
    else if (m <= 4) TB_LAUNCH(4);
    else if (m <= 8) TB_LAUNCH(8);
    else if (m <= 10) TB_LAUNCH(10);
    else TB_LAUNCH(16);
#undef TB_LAUNCH
    BDH_LAUNCH_CHECK();
    return 0;
}

template int chol_factor_batch<float>(float*, double*, const BatchPlan&, int*, int, hipStream_t, PhaseRecorder*);
template int chol_factor_batch<double>(double*, double*, const BatchPlan&, int*, int, hipStream_t, PhaseRecorder*);
template int launch_trmv_batch<float>(const float*, const double*, long long, int, const BatchPlan&, double*, hipStream_t);
template int launch_trmv_batch<double>(const double*, const double*, long long, int, const BatchPlan&, double*, hipStream_t);
template int trmv_blocks<float>(int);
template int trmv_blocks<double>(int);

template int chol_factor<float>(float*, long long, int, double*, int*, int, hipStream_t, PhaseRecorder*);
template int chol_factor<double>(double*, long long, int, double*, int*, int, hipStream_t, PhaseRecorder*);
template int launch_trmv_partials<float>(const float*, long long, int, const double*, long long, int, double*, hipStream_t);
template int launch_trmv_partials<double>(const double*, long long, int, const double*, long long, int, double*, hipStream_t);

}  // namespace bdh
