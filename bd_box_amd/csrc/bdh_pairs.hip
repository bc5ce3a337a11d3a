// Pair sums over hydrodynamic blocks evaluated on the fly, and the Geyer-Winter TEA noise.
//
// pair_sums_kernel<T, EW, TEA>: for every row bead i, over the column beads j of one split,
//   f_i += D_ij F_j                        (D.F, diff_equation.c:55-117 / 119-214)
// and with TEA (diff_rand.c:86-135, the sums of compute_r's DIFF_GEYER branch)
//   x_i += D_ij X_j,  X_j = xi_j sqrt(2 dt / D_jj)          (j != i)
//   c_i += sum_l D_ij[., l]^2 / D_jj,l                       (j != i)
//   e_i += sum of the 9 entries of D_ij                      (j != i)
// D_ij is the free-space RPY block (EW = false) or the Beenakker-Ewald block (EW = true);
// D_jj in the TEA sums is the free-space diagonal kT/(6 pi eta a_j) (diag_D, as the
// reference uses it).  Each pair is evaluated once per row it belongs to (full rows, no
// atomics); accumulation is FP64.
//
// TEA closing (diff_rand.c:185-224): eps = 2 (sum_{i<j} D_ij entries) / (n (n - 1)),
// clamped to 1; a = (n - 1) eps^2 - (n - 2) eps; beta = (1 - sqrt(1 - a)) / a (0.5 when
// |a| < 1e-7); Rs_r = (1 + beta^2 C_r / D_rr)^(-1/2) (beta Feff_r + D_rr X_r).
#include <vector>

#include "bdh_common.hpp"
#include "bdh_ewald.hpp"
#include "bdh_kernels.hpp"

namespace bdh {

// ------------------------------------------------------------------ Ewald tables
__constant__ EwaldTables c_ew;
__constant__ short4 c_hfg[EW_MAXV];
__constant__ double4 c_recip[EW_MAXV];
__constant__ double c_recip_z[EW_MAXV];

int ewald_upload(double box, int R)
{
    BDH_CHECK(R >= 0 && R <= 5, "Ewald lattice radius must be 0..5");
    EwaldTables t{};
    std::vector<short4> hfg;
    std::vector<double4> rec;
    std::vector<double> recz;
    t.box = box;
    t.inv_L = 1.0 / box;
    t.alpha = SQRTPI * t.inv_L;            // init_tensor, diff_tensor.c:551-554
    t.alpha2 = t.alpha * t.alpha;
    t.alpha3 = t.alpha * t.alpha2;
    t.inv_alpha2 = 1.0 / t.alpha2;
    t.preQ = t.inv_L * t.inv_L * t.inv_L;
    t.self_a = -6.0 * t.alpha / SQRTPI;
    t.self_b = 40.0 * t.alpha3 / (3.0 * SQRTPI);
    const int rr = R * R;
    for (int h = -R; h <= R; ++h) {
        const int nf = (int)std::sqrt((double)(rr - h * h));
        for (int f = -nf; f <= nf; ++f) {
            const int ng = (int)std::sqrt((double)(rr - h * h - f * f));
            for (int g = -ng; g <= ng; ++g) {
                hfg.push_back(make_short4((short)h, (short)f, (short)g, 0));
                const double s = (double)(h * h + f * f + g * g);
                if (s == 0.0) {
                    rec.push_back(make_double4(0, 0, 0, 0));
                    recz.push_back(0.0);
                    continue;
                }
                const double nrm = std::sqrt(s);
                const double k2 = s * 4.0 * PI_F * PI_F;
                const double G = (1.0 + (t.inv_alpha2 * k2 / 4.0) * (1.0 + t.inv_alpha2 * k2 / 2.0)) *
                                 (6.0 * PI_F * std::exp(-t.inv_alpha2 * k2 / 4.0) / k2);
                rec.push_back(make_double4(G, k2 / 3.0, h / nrm, f / nrm));
                recz.push_back(g / nrm);
            }
        }
    }
    t.nv = (int)hfg.size();
    BDH_CHECK(t.nv <= EW_MAXV, "Ewald table overflow");
    BDH_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_ew), &t, sizeof(t)));
    BDH_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_hfg), hfg.data(), sizeof(short4) * hfg.size()));
    BDH_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_recip), rec.data(), sizeof(double4) * rec.size()));
    BDH_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_recip_z), recz.data(), sizeof(double) * recz.size()));
    return t.nv;
}

// ------------------------------------------------------------------ blocks
template <typename T>
__device__ __forceinline__ T rsqrt_p(T x);
template <>
__device__ __forceinline__ float rsqrt_p<float>(float x) { return __frsqrt_rn(x); }
template <>
__device__ __forceinline__ double rsqrt_p<double>(double x) { return 1.0 / sqrt(x); }

// Free-space RPY block for separation d = r_i - r_j (diff_tensor.c:45-92), or the diagonal.
template <typename T>
__device__ __forceinline__ Blk6 free_block(double dx, double dy, double dz, double ai, double aj,
                                           bool self, double preDii, double preDij)
{
    if (self) {
        const double dii = preDii / ai;
        return Blk6{dii, dii, dii, 0.0, 0.0, 0.0};
    }
    const T x = (T)dx, y = (T)dy, z = (T)dz, ta = (T)ai, tb = (T)aj;
    const T r2 = x * x + y * y + z * z;
    const T inv_r = rsqrt_p<T>(r2);
    const T r = r2 * inv_r;
    const T ex = x * inv_r, ey = y * inv_r, ez = z * inv_r;
    T d, o;
    if (r < ta + tb) {
        const T sig = (T)0.5 * (ta + tb);
        const T pre = (T)preDii / sig;
        const T c = (T)(3.0 / 32.0) * r / sig;
        d = pre * ((T)1 - (T)3 * c);
        o = pre * c;
    } else {
        const T pre = (T)preDij * inv_r;
        const T s = (ta * ta + tb * tb) * inv_r * inv_r;
        d = pre * ((T)1 + s * (T)(1.0 / 3.0));
        o = pre * ((T)1 - s);
    }
    return Blk6{(double)(d + o * ex * ex), (double)(d + o * ey * ey), (double)(d + o * ez * ez),
                (double)(o * ex * ey), (double)(o * ex * ez), (double)(o * ey * ez)};
}

template <typename T, bool EW>
__device__ __forceinline__ Blk6 pair_block(const double4& pi, const double4& pj, bool self,
                                           double preDii, double preDij)
{
    if constexpr (EW)
        return beenakker_block<T>(pj.x - pi.x, pj.y - pi.y, pj.z - pi.z, pi.w, pj.w, self, preDii,
                                  preDij);
    else
        return free_block<T>(pi.x - pj.x, pi.y - pj.y, pi.z - pj.z, pi.w, pj.w, self, preDii,
                             preDij);
}

// ------------------------------------------------------------------ pair sums
constexpr int PS_ROWS = 64;
constexpr int PS_CHUNK = 128;

template <typename T, bool EW, bool TEA>
__global__ void __launch_bounds__(256)
pair_sums_kernel(const double4* __restrict__ pos, const double* __restrict__ F,
                 const double* __restrict__ xi, int N, int cols_per_split, double preDii,
                 double preDij, double dt, double* __restrict__ Pf, double* __restrict__ Px,
                 double* __restrict__ Pc, double* __restrict__ Pe)
{
    __shared__ double4 s_pos[PS_CHUNK];
    __shared__ double s_f[PS_CHUNK * 3];
    __shared__ double s_x[TEA ? PS_CHUNK * 3 : 1];
    __shared__ double s_red[TEA ? 10 : 3][4][PS_ROWS];
    const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * PS_ROWS + tx;
    const int i = blockIdx.x * PS_ROWS + tx;
    const int jb = blockIdx.y * cols_per_split;
    const int je = min(N, jb + cols_per_split);
    const double4 pi = (i < N) ? pos[i] : make_double4(0, 0, 0, 1);
    double f0 = 0, f1 = 0, f2 = 0, x0 = 0, x1 = 0, x2 = 0, c0 = 0, c1 = 0, c2 = 0, e = 0;
    for (int b0 = jb; b0 < je; b0 += PS_CHUNK) {
        const int cn = min(PS_CHUNK, je - b0);
        __syncthreads();
        for (int q = tid; q < cn; q += 256) {
            const double4 p = pos[b0 + q];
            s_pos[q] = p;
            s_f[3 * q] = F[3 * (b0 + q)];
            s_f[3 * q + 1] = F[3 * (b0 + q) + 1];
            s_f[3 * q + 2] = F[3 * (b0 + q) + 2];
            if constexpr (TEA) {
                const double sx = sqrt(2.0 * dt / (preDii / p.w));
                s_x[3 * q] = xi[3 * (b0 + q)] * sx;
                s_x[3 * q + 1] = xi[3 * (b0 + q) + 1] * sx;
                s_x[3 * q + 2] = xi[3 * (b0 + q) + 2] * sx;
            }
        }
        __syncthreads();
        if (i >= N) continue;
        for (int q = ty; q < cn; q += 4) {
            const int j = b0 + q;
            const bool self = (j == i);
            const double4 pj = s_pos[q];
            const Blk6 B = pair_block<T, EW>(pi, pj, self, preDii, preDij);
            const double fx = s_f[3 * q], fy = s_f[3 * q + 1], fz = s_f[3 * q + 2];
            f0 += B.xx * fx + B.xy * fy + B.xz * fz;
            f1 += B.xy * fx + B.yy * fy + B.yz * fz;
            f2 += B.xz * fx + B.yz * fy + B.zz * fz;
            if constexpr (TEA) {
                if (!self) {
                    const double X0 = s_x[3 * q], X1 = s_x[3 * q + 1], X2 = s_x[3 * q + 2];
                    x0 += B.xx * X0 + B.xy * X1 + B.xz * X2;
                    x1 += B.xy * X0 + B.yy * X1 + B.yz * X2;
                    x2 += B.xz * X0 + B.yz * X1 + B.zz * X2;
                    const double inv_d = pj.w / preDii;  // 1 / D_jj, isotropic
                    c0 += (B.xx * B.xx + B.xy * B.xy + B.xz * B.xz) * inv_d;
                    c1 += (B.xy * B.xy + B.yy * B.yy + B.yz * B.yz) * inv_d;
                    c2 += (B.xz * B.xz + B.yz * B.yz + B.zz * B.zz) * inv_d;
                    e += B.xx + B.yy + B.zz + 2.0 * (B.xy + B.xz + B.yz);
                }
            }
        }
    }
    s_red[0][ty][tx] = f0; s_red[1][ty][tx] = f1; s_red[2][ty][tx] = f2;
    if constexpr (TEA) {
        s_red[3][ty][tx] = x0; s_red[4][ty][tx] = x1; s_red[5][ty][tx] = x2;
        s_red[6][ty][tx] = c0; s_red[7][ty][tx] = c1; s_red[8][ty][tx] = c2;
        s_red[9][ty][tx] = e;
    }
    __syncthreads();
    if (ty == 0 && i < N) {
        const long long o = (long long)blockIdx.y * 3 * N + 3 * i;
#pragma unroll
        for (int k = 0; k < 3; ++k)
            Pf[o + k] = (s_red[k][0][tx] + s_red[k][1][tx]) + (s_red[k][2][tx] + s_red[k][3][tx]);
        if constexpr (TEA) {
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                Px[o + k] = (s_red[3 + k][0][tx] + s_red[3 + k][1][tx]) + (s_red[3 + k][2][tx] + s_red[3 + k][3][tx]);
                Pc[o + k] = (s_red[6 + k][0][tx] + s_red[6 + k][1][tx]) + (s_red[6 + k][2][tx] + s_red[6 + k][3][tx]);
            }
            Pe[(long long)blockIdx.y * N + i] = (s_red[9][0][tx] + s_red[9][1][tx]) + (s_red[9][2][tx] + s_red[9][3][tx]);
        }
    }
}

int pair_splits(int N, bool ewald)
{
    const int rows = ceil_div(N, PS_ROWS);
    int s = ceil_div(ewald ? 4096 : 2048, rows);
    const int max_s = ceil_div(N, 32);
    if (s > max_s) s = max_s;
    return s < 1 ? 1 : s;
}

template <typename T>
int launch_pair_sums(const double4* d_pos, const double* d_F, const double* d_xi, int N,
                     double preDii, double preDij, double dt, bool ewald, bool tea, int splits,
                     double* d_Pf, double* d_Px, double* d_Pc, double* d_Pe, hipStream_t st)
{
    dim3 grid(ceil_div(N, PS_ROWS), splits);
    dim3 blk(PS_ROWS, 4);
    const int cps = ceil_div(N, splits);
    if (ewald && tea)
        hipLaunchKernelGGL((pair_sums_kernel<T, true, true>), grid, blk, 0, st, d_pos, d_F, d_xi, N, cps,
                           preDii, preDij, dt, d_Pf, d_Px, d_Pc, d_Pe);
    else if (ewald)
        hipLaunchKernelGGL((pair_sums_kernel<T, true, false>), grid, blk, 0, st, d_pos, d_F, d_xi, N, cps,
                           preDii, preDij, dt, d_Pf, d_Px, d_Pc, d_Pe);
    else if (tea)
        hipLaunchKernelGGL((pair_sums_kernel<T, false, true>), grid, blk, 0, st, d_pos, d_F, d_xi, N, cps,
                           preDii, preDij, dt, d_Pf, d_Px, d_Pc, d_Pe);
    else
        hipLaunchKernelGGL((pair_sums_kernel<T, false, false>), grid, blk, 0, st, d_pos, d_F, d_xi, N, cps,
                           preDii, preDij, dt, d_Pf, d_Px, d_Pc, d_Pe);
    BDH_LAUNCH_CHECK();
    return 0;
}

// ------------------------------------------------------------------ Ewald assembly
// Lower triangle of the Beenakker-Ewald D (rows of i, columns of j, i >= j), the same
// tiling as the free-space assembly: (64 row beads) x (16 column beads) per workgroup.
template <typename T>
__global__ void __launch_bounds__(256)
ewald_assemble_kernel(const double4* __restrict__ pos, int N, double preDii, double preDij,
                      T* __restrict__ A, long long ld)
{
    const int i0 = blockIdx.x * 64, j0 = blockIdx.y * 16;
    if (j0 > i0 + 63) return;
    const int i = i0 + threadIdx.x;
    if (i >= N) return;
    const double4 pi = pos[i];
    for (int q = 0; q < 4; ++q) {
        const int j = j0 + threadIdx.y * 4 + q;
        if (j >= N || j > i) continue;
        const Blk6 B = beenakker_block<T>(pos[j].x - pi.x, pos[j].y - pi.y, pos[j].z - pi.z, pi.w,
                                          pos[j].w, i == j, preDii, preDij);
        T* c0 = A + (long long)(3 * j) * ld + 3 * i;
        c0[0] = (T)B.xx; c0[1] = (T)B.xy; c0[2] = (T)B.xz;
        c0[ld] = (T)B.xy; c0[ld + 1] = (T)B.yy; c0[ld + 2] = (T)B.yz;
        c0[2 * ld] = (T)B.xz; c0[2 * ld + 1] = (T)B.yz; c0[2 * ld + 2] = (T)B.zz;
    }
}

template <typename T>
int launch_ewald_assemble(const double4* d_pos, int N, double preDii, double preDij, T* d_A,
                          long long ld, hipStream_t st)
{
    dim3 grid(ceil_div(N, 64), ceil_div(N, 16));
    hipLaunchKernelGGL(ewald_assemble_kernel<T>, grid, dim3(64, 4), 0, st, d_pos, N, preDii, preDij,
                       d_A, ld);
    BDH_LAUNCH_CHECK();
    return 0;
}

// ------------------------------------------------------------------ TEA closing
// One block: eps from the row partials (each unordered pair counted from both rows), then
// beta; out[0] = beta, out[1] = eps.
__global__ void __launch_bounds__(256)
tea_beta_kernel(const double* __restrict__ Pe, int S, int N, double* __restrict__ out)
{
    __shared__ double s[256];
    double a = 0.0;
    for (long long k = threadIdx.x; k < (long long)S * N; k += 256) a += Pe[k];
    s[threadIdx.x] = a;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double n = 3.0 * N;
        double eps = 0.5 * s[0];
        eps /= n;
        eps /= (n - 1.0);
        eps *= 2.0;
        if (eps > 1.0) eps = 1.0;
        const double aa = (3.0 * N - 1.0) * eps * eps - (3.0 * N - 2.0) * eps;
        const double beta = (fabs(aa) < (double)0.0000001f) ? 0.5 : (1.0 - sqrt(1.0 - aa)) / aa;
        out[0] = beta;
        out[1] = eps;
    }
}

__global__ void tea_rs_kernel(const double4* __restrict__ pos, const double* __restrict__ xi,
                              const double* __restrict__ Feff, const double* __restrict__ C,
                              const double* __restrict__ bt, int N, double preDii, double dt,
                              double* __restrict__ Rs)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= 3 * N) return;
    const double beta = bt[0];
    const double dii = preDii / pos[r / 3].w;
    const double X = xi[r] * sqrt(2.0 * dt / dii);
    double c = C[r] * (beta * beta * (1.0 / dii));
    c = sqrt(1.0 / (1.0 + c));
    Rs[r] = c * (beta * Feff[r] + dii * X);
}

int launch_tea_close(const double4* d_pos, const double* d_xi, const double* d_Feff,
                     const double* d_C, const double* d_Pe, int S, int N, double preDii, double dt,
                     double* d_bt, double* d_Rs, hipStream_t st)
{
    hipLaunchKernelGGL(tea_beta_kernel, dim3(1), dim3(256), 0, st, d_Pe, S, N, d_bt);
    BDH_LAUNCH_CHECK();
    hipLaunchKernelGGL(tea_rs_kernel, dim3(ceil_div(3 * N, 256)), dim3(256), 0, st, d_pos, d_xi, d_Feff,
                       d_C, d_bt, N, preDii, dt, d_Rs);
    BDH_LAUNCH_CHECK();
    return 0;
}

template int launch_pair_sums<float>(const double4*, const double*, const double*, int, double, double, double, bool, bool, int, double*, double*, double*, double*, hipStream_t);
template int launch_pair_sums<double>(const double4*, const double*, const double*, int, double, double, double, bool, bool, int, double*, double*, double*, double*, hipStream_t);
template int launch_ewald_assemble<float>(const double4*, int, double, double, float*, long long, hipStream_t);
template int launch_ewald_assemble<double>(const double4*, int, double, double, double*, long long, hipStream_t);

}  // namespace bdh
