// Free-space Rotne-Prager-Yamakawa tensor on gfx950: tile-wise assembly of the lower
// triangle of D (the on-the-fly products over the same blocks are in bdh_pairs.hip).
//
// Block formula (src/diff_alg/diff_tensor.c:45-92), r = |r_i - r_j|, e = (r_i - r_j)/r:
//   r >= a_i + a_j : (kT/8 pi eta r) [ (1 + (a_i^2+a_j^2)/(3 r^2)) I + (1 - (a_i^2+a_j^2)/r^2) e e^T ]
//   r <  a_i + a_j : (kT/6 pi eta abar) [ (1 - 9r/(32 abar)) I + (3r/(32 abar)) e e^T ],
//                    abar = (a_i + a_j)/2
//   diagonal       : kT/(6 pi eta a_i) I                       (diff_decl.c:94-98)
// Bead separations are formed in FP64 from the FP64 positions and then rounded to the
// arithmetic type T, so FP32 assembly keeps relative accuracy for distant beads.
//
// Storage: column-major, leading dimension ld; block (i, j), i > j, at rows 3i.., columns
// 3j.. (the transpose of the reference's upper-triangle placement, diff_tensor.c:425-435,
// which is the same matrix since D is symmetric).
#include "bdh_common.hpp"
#include "bdh_kernels.hpp"

namespace bdh {

template <typename T>
struct Rpy6 {
    T xx, yy, zz, xy, xz, yz;
};

template <typename T>
__device__ __forceinline__ T rsqrt_t(T x);
template <>
__device__ __forceinline__ float rsqrt_t<float>(float x) { return __frsqrt_rn(x); }
template <>
__device__ __forceinline__ double rsqrt_t<double>(double x) { return 1.0 / sqrt(x); }

// Off-diagonal block for separation d (in FP64) and radii ai, aj.
template <typename T>
__device__ __forceinline__ Rpy6<T> rpy_pair(double dx, double dy, double dz, T ai, T aj,
                                            T preDii, T preDij)
{
    T x = (T)dx, y = (T)dy, z = (T)dz;
    T r2 = x * x + y * y + z * z;
    T inv_r = rsqrt_t<T>(r2);
    T r = r2 * inv_r;
    T ex = x * inv_r, ey = y * inv_r, ez = z * inv_r;
    T d, o;
    if (r < ai + aj) {
        T sig = (T)0.5 * (ai + aj);
        T pre = preDii / sig;
        T c = (T)(3.0 / 32.0) * r / sig;
        d = pre * ((T)1 - (T)3 * c);  // 1 - 9r/(32 sig)
        o = pre * c;                   // 3r/(32 sig)
    } else {
        T pre = preDij * inv_r;
        T s = (ai * ai + aj * aj) * inv_r * inv_r;
        d = pre * ((T)1 + s * (T)(1.0 / 3.0));
        o = pre * ((T)1 - s);
    }
    Rpy6<T> b;
    b.xx = d + o * ex * ex;
    b.yy = d + o * ey * ey;
    b.zz = d + o * ez * ez;
    b.xy = o * ex * ey;
    b.xz = o * ex * ez;
    b.yz = o * ey * ez;
    return b;
}

// ----------------------------------------------------------------------------- assembly
// Workgroup (64, 4): tx -> row bead i (64 per tile), ty -> 4 column beads each (16 per
// tile).  Tiles strictly above the diagonal exit at once.  A wavefront's stores to one
// column cover 192 consecutive rows (3 per lane).
constexpr int ASM_ROWS = 64;
constexpr int ASM_COLS = 16;

// With pos2, the blocks are the mean of those at pos and pos2 (the IG-T tensor
// (D(r0) + D(r*)) / 2 of compute_hydro_IGT, diff_equation.c:327-352; the diagonal blocks do
// not depend on the positions).
template <typename T>
__device__ __forceinline__ void rpy_assemble_body(const double4* __restrict__ pos,
                                                  const double4* __restrict__ pos2, int N, T preDii,
                                                  T preDij, T* __restrict__ A, long long ld, int bx,
                                                  int by)
{
    const int i0 = bx * ASM_ROWS;
    const int j0 = by * ASM_COLS;
    if (j0 > i0 + ASM_ROWS - 1) return;  // tile above the diagonal
    __shared__ double4 s_col[ASM_COLS], s_col2[ASM_COLS];
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int tid = ty * ASM_ROWS + tx;
    if (tid < ASM_COLS) {
        int j = j0 + tid;
        s_col[tid] = (j < N) ? pos[j] : make_double4(0, 0, 0, 1);
        if (pos2) s_col2[tid] = (j < N) ? pos2[j] : make_double4(0, 0, 0, 1);
    }
    __syncthreads();
    const int i = i0 + tx;
    if (i >= N) return;
    const double4 pi = pos[i];
    const double4 pi2 = pos2 ? pos2[i] : pi;
    const T ai = (T)pi.w;
#pragma unroll
    for (int q = 0; q < ASM_COLS / 4; ++q) {
        const int jl = ty * (ASM_COLS / 4) + q;
        const int j = j0 + jl;
        if (j >= N || j > i) continue;
        T* c0 = A + (long long)(3 * j) * ld + 3 * i;
        if (j == i) {
            T dii = preDii / ai;
            c0[0] = dii; c0[1] = 0; c0[2] = 0;
            c0[ld + 1] = dii; c0[ld + 2] = 0;
            c0[2 * ld + 2] = dii;
            continue;
        }
        const double4 pj = s_col[jl];
        Rpy6<T> b = rpy_pair<T>(pi.x - pj.x, pi.y - pj.y, pi.z - pj.z, ai, (T)pj.w, preDii, preDij);
        if (pos2) {
            const double4 pj2 = s_col2[jl];
            const Rpy6<T> b2 = rpy_pair<T>(pi2.x - pj2.x, pi2.y - pj2.y, pi2.z - pj2.z, ai, (T)pj2.w,
                                           preDii, preDij);
            b.xx = (b.xx + b2.xx) / (T)2; b.yy = (b.yy + b2.yy) / (T)2; b.zz = (b.zz + b2.zz) / (T)2;
            b.xy = (b.xy + b2.xy) / (T)2; b.xz = (b.xz + b2.xz) / (T)2; b.yz = (b.yz + b2.yz) / (T)2;
        }
        c0[0] = b.xx; c0[1] = b.xy; c0[2] = b.xz;
        c0[ld + 0] = b.xy; c0[ld + 1] = b.yy; c0[ld + 2] = b.yz;
        c0[2 * ld + 0] = b.xz; c0[2 * ld + 1] = b.yz; c0[2 * ld + 2] = b.zz;
    }
}

template <typename T>
__global__ void __launch_bounds__(256)
rpy_assemble_kernel(const double4* __restrict__ pos, const double4* __restrict__ pos2, int N,
                    T preDii, T preDij, T* __restrict__ A, long long ld)
{
    rpy_assemble_body<T>(pos, pos2, N, preDii, preDij, A, ld, blockIdx.x, blockIdx.y);
}

// Every chain of a batch: workgroup x of chain c is tile (x % gx, x / gx) of its own grid.
template <typename T>
__global__ void __launch_bounds__(256)
rpy_assemble_batch_kernel(const double4* __restrict__ pos, T preDii, T preDij, T* __restrict__ A,
                          const BatchChain* __restrict__ ch, const int* __restrict__ prefix, int nch)
{
    const int c = batch_find(prefix, nch, blockIdx.x);
    const BatchChain b = ch[c];
    const int x = blockIdx.x - prefix[c];
    const int gx = (b.N + ASM_ROWS - 1) / ASM_ROWS;
    rpy_assemble_body<T>(pos + b.bead0, nullptr, b.N, preDii, preDij, A + b.mat, b.n_pad, x % gx, x / gx);
}

int rpy_blocks(int N) { return ceil_div(N, ASM_ROWS) * ceil_div(N, ASM_COLS); }

template <typename T>
int launch_rpy_assemble_batch(const double4* d_pos, double preDii, double preDij, T* d_A,
                              const BatchPlan& plan, hipStream_t st)
{
    hipLaunchKernelGGL(rpy_assemble_batch_kernel<T>, dim3(plan.asm_blocks), dim3(ASM_ROWS, 4), 0, st,
                       d_pos, (T)preDii, (T)preDij, d_A, plan.ch, plan.asm_pre, plan.nch);
    BDH_LAUNCH_CHECK();
    return 0;
}

// Padding rows/columns n..n_pad-1 form an identity block so the blocked factorization
// runs on whole tiles; everything else of the matrix starts at zero.
template <typename T>
__global__ void set_identity_pad_kernel(T* __restrict__ A, long long ld, int n, int n_pad)
{
    int r = n + blockIdx.x * blockDim.x + threadIdx.x;
    if (r < n_pad) A[(long long)r * ld + r] = (T)1;
}

template <typename T>
int launch_rpy_assemble(const double4* d_pos, int N, double preDii, double preDij, T* d_A,
                        long long ld, hipStream_t st, const double4* d_pos2)
{
    dim3 grid(ceil_div(N, ASM_ROWS), ceil_div(N, ASM_COLS));
    hipLaunchKernelGGL(rpy_assemble_kernel<T>, grid, dim3(ASM_ROWS, 4), 0, st, d_pos, d_pos2, N,
                       (T)preDii, (T)preDij, d_A, ld);
    BDH_LAUNCH_CHECK();
    return 0;
}

template <typename T>
int launch_set_identity_pad(T* d_A, long long ld, int n, int n_pad, hipStream_t st)
{
    if (n_pad <= n) return 0;
    hipLaunchKernelGGL(set_identity_pad_kernel<T>, dim3(ceil_div(n_pad - n, 256)), dim3(256), 0,
                       st, d_A, ld, n, n_pad);
    BDH_LAUNCH_CHECK();
    return 0;
}

template int launch_rpy_assemble<float>(const double4*, int, double, double, float*, long long, hipStream_t,
                                        const double4*);
template int launch_rpy_assemble<double>(const double4*, int, double, double, double*, long long, hipStream_t,
                                         const double4*);
template int launch_rpy_assemble_batch<float>(const double4*, double, double, float*, const BatchPlan&, hipStream_t);
template int launch_rpy_assemble_batch<double>(const double4*, double, double, double*, const BatchPlan&, hipStream_t);
template int launch_set_identity_pad<float>(float*, long long, int, int, hipStream_t);
template int launch_set_identity_pad<double>(double*, long long, int, int, hipStream_t);

}  // namespace bdh
