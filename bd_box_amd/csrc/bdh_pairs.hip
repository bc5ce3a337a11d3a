// Pair sums over hydrodynamic blocks evaluated on the fly, and the Geyer-Winter TEA noise.
//
// pair_sums_kernel<T, EW, TEA>: for every row bead i, over the column beads j of a tile,
//   f_i += D_ij F_j                        (D.F, diff_equation.c:55-117 / 119-214)
// and with TEA (diff_rand.c:86-135, the sums of compute_r's DIFF_GEYER branch)
//   x_i += D_ij X_j,  X_j = xi_j sqrt(2 dt / D_jj)          (j != i)
//   c_i += sum_l D_ij[., l]^2 / D_jj,l                       (j != i)
//   e_i += sum of the 9 entries of D_ij                      (j != i)
// D_ij is the free-space RPY block (EW = false) or the Beenakker-Ewald block (EW = true);
// with MEAN it is the IG-T mean (D_ij(r*) + D_ij(r0)) / 2 of the blocks at two position sets
// (pos2 = r*, diff_equation.c:341-366), whose TEA sums the IG-T corrector draws from;
// D_jj in the TEA sums is the free-space diagonal kT/(6 pi eta a_j) (diag_D, as the
// reference uses it).  Each unordered pair is evaluated once and serves both of its rows
// (no atomics); accumulation is FP64.
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
__constant__ double4 c_recip[EW_MAXV];
__constant__ double c_recip_z[EW_MAXV];
__constant__ double c_recip_w[EW_MAXV];

// Method of the tables last uploaded (the launchers pick the matching instantiation; every
// launch is preceded by an upload check for the calling system).
static int g_uploaded_method = EW_BEENAKKER;

int ewald_upload(double box, int R, int method, double alpha_smith, double alpha_scale, int R_real,
                 int R_recip, double r_cut)
{
    BDH_CHECK(R >= 0 && R <= 5, "Ewald lattice radius must be 0..5");
    BDH_CHECK(method == EW_SMITH || method == EW_BEENAKKER, "Ewald method");
    const bool ext = alpha_scale > 0.0;
    if (!ext) R_real = R_recip = R;
    BDH_CHECK(method == EW_BEENAKKER || !ext, "Ewald: a split of its own only for the Beenakker sums");
    BDH_CHECK(R_real >= 0 && R_real <= 5 && R_recip >= 0 && R_recip <= 5, "Ewald split: radii 0..5");
    R = R_real > R_recip ? R_real : R_recip;
    EwaldTables t{};
    int nv = 0;
    std::vector<double4> rec;
    std::vector<double> recz, recw;
    t.method = method;
    t.box = box;
    t.inv_L = 1.0 / box;
    if (method == EW_BEENAKKER) {
        t.alpha = SQRTPI * t.inv_L;            // init_tensor, diff_tensor.c:551-554
        if (ext) t.alpha *= alpha_scale;
    } else {
        t.alpha = alpha_smith;                  // reduced units (the .prm ewald_alpha)
    }
    t.alpha2 = t.alpha * t.alpha;
    t.alpha3 = t.alpha * t.alpha2;
    t.inv_alpha2 = 1.0 / t.alpha2;
    t.preQ = t.inv_L * t.inv_L * t.inv_L;
    t.preO = 0.75 * t.inv_L;
    t.self_a = -6.0 * t.alpha / SQRTPI;
    t.self_b = 40.0 * t.alpha3 / (3.0 * SQRTPI);
    // M_PI_ALPHA_SQ = -M_PIF * M_PIF / alpha2 with the product of the float pis in float
    // (diff_tensor.c:557)
    const float pif = 3.14159265358979323846f;
    const float mpp = -pif * pif;
    const double m_pi_alpha_sq = (double)mpp / t.alpha2;
    if (method == EW_BEENAKKER) {
        // the half sphere (h, f, g) > 0 lexicographically, |.|^2 <= R_recip^2, h then f then g
        // ascending (beenakker_block's loop order), each vector standing for itself and its
        // negative (2 G)
        const int rk = R_recip * R_recip;
        bool any = false;
        for (int h = 0; h * h <= rk; ++h) {
            const int nf = (int)std::sqrt((double)(rk - h * h));
            for (int f = (h == 0 ? 0 : -nf); f <= nf; ++f) {
                const int ng = (int)std::sqrt((double)(rk - h * h - f * f));
                for (int g = (h == 0 && f == 0 ? 1 : -ng); g <= ng; ++g) {
                    ++nv;
                    const double s = (double)(h * h + f * f + g * g);
                    const double nrm = std::sqrt(s);
                    // The reference forms k = 2 pi (h, f, g) in box units (diff_tensor.c:372-376)
                    // against alpha and sigma in A, so its reciprocal terms underflow to zero for
                    // any box of more than a few A; the extension takes k = 2 pi (h, f, g) / L,
                    // the wave vectors of the box, as Beenakker's sum has them.
                    const double k2 = ext ? s * 4.0 * M_PI * M_PI * t.inv_L * t.inv_L : s * 4.0 * PI_F * PI_F;
                    const double six_pi = ext ? 6.0 * M_PI : 6.0 * PI_F;
                    const double G = (1.0 + (t.inv_alpha2 * k2 / 4.0) * (1.0 + t.inv_alpha2 * k2 / 2.0)) *
                                     (six_pi * std::exp(-t.inv_alpha2 * k2 / 4.0) / k2);
                    any = any || G != 0.0;
                    rec.push_back(make_double4(2.0 * G, k2 / 3.0, h / nrm, f / nrm));
                    recz.push_back(g / nrm);
                    recw.push_back(0.0);
                }
            }
        }
        t.RR_recip = rk;
        t.nrecip = any ? nv : 0;
    } else {
        const int rr = R * R;
        for (int h = -R; h <= R; ++h) {
            const int nf = (int)std::sqrt((double)(rr - h * h));
            for (int f = -nf; f <= nf; ++f) {
                const int ng = (int)std::sqrt((double)(rr - h * h - f * f));
                for (int g = -ng; g <= ng; ++g) {
                    ++nv;
                    const double s = (double)(h * h + f * f + g * g);
                    if (s == 0.0) {
                        rec.push_back(make_double4(0, 0, 0, 0));
                        recz.push_back(0.0);
                        recw.push_back(0.0);
                        continue;
                    }
                    // diff_tensor.c:199-225; e = (h, f, g) / |.| as get_ev2_norms forms it
                    const double nrm = std::sqrt(s);
                    const double inv = 1.0 / nrm;
                    const double ex = h * inv, ey = f * inv, ez = g * inv;
                    const double G1 = 2.0 * std::exp(m_pi_alpha_sq * s) / ((double)pif * s);
                    rec.push_back(make_double4(G1, -1.0 + m_pi_alpha_sq * s, ex, ey));
                    recz.push_back(ez);
                    recw.push_back(4.0 * (double)pif * std::exp(m_pi_alpha_sq * s));
                }
            }
        }
    }
    t.nv = nv;
    t.R = R;
    t.RR_real = R_real * R_real;
    t.min_image = ext ? 1 : 0;
    BDH_CHECK(r_cut <= 0.0 || ext, "Ewald: a real-space cutoff only for the extension split");
    t.rc2 = r_cut > 0.0 ? r_cut * r_cut : 1e300;
    BDH_CHECK(t.nv <= EW_MAXV, "Ewald table overflow");
    BDH_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_ew), &t, sizeof(t)));
    if (!rec.empty()) {
        BDH_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_recip), rec.data(), sizeof(double4) * rec.size()));
        BDH_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_recip_z), recz.data(), sizeof(double) * recz.size()));
        BDH_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_recip_w), recw.data(), sizeof(double) * recw.size()));
    }
    g_uploaded_method = method;
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

// EW: 0 free space, 1 Beenakker-Ewald, 2 Smith-Ewald
template <typename T, int EW>
__device__ __forceinline__ Blk6 pair_block(const double4& pi, const double4& pj, bool self,
                                           double preDii, double preDij)
{
    if constexpr (EW == 1)
        return beenakker_block<T>(pj.x - pi.x, pj.y - pi.y, pj.z - pi.z, pi.w, pj.w, self, preDii,
                                  preDij);
    else if constexpr (EW == 2)
        return smith_block<T>(pj.x - pi.x, pj.y - pi.y, pj.z - pi.z, pi.w, pj.w, self, preDii,
                              preDij);
    else
        return free_block<T>(pi.x - pj.x, pi.y - pj.y, pi.z - pj.z, pi.w, pj.w, self, preDii,
                             preDij);
}

// ------------------------------------------------------------------ pair sums
// Tiles of 64 beads.  Workgroup b takes the tile pair (I, J), I <= J, of the upper triangle
// (b = J (J + 1) / 2 + I).  Lane t owns row bead I*64 + t; wave w takes the shifts
// s = w, w + 4, ... and evaluates the pair (t, (t + s) mod 64) of tile J, so the 64 column
// beads of one shift are distinct within a wave and its column sums go to the wave's own
// LDS rows without conflicts.
//   I < J : every unordered pair once; the block serves both rows, D_ji = D_ij (a block is
//           an even function of the separation, and symmetric).
//   I == J: every ordered pair, rows only.
// Row sums of tile I land in split J, column sums of tile J in split I: every split of
// every row is written exactly once, so the partials need no clearing.  The eps sums of
// the workgroup (each unordered pair i != j counted twice, as from both rows) are reduced
// into Pe[b].
constexpr int PS_TILE = 64;

int pair_splits(int N, bool) { return ceil_div(N, PS_TILE); }
int pair_eparts(int N) { const int t = ceil_div(N, PS_TILE); return t * (t + 1) / 2; }

template <typename T, int EW, bool TEA, bool MEAN = false>
__device__ __forceinline__ void pair_sums_body(const double4* __restrict__ pos, const double* __restrict__ F,
                                               const double* __restrict__ xi, int N, double preDii,
                                               double preDij, double dt, double* __restrict__ Pf,
                                               double* __restrict__ Px, double* __restrict__ Pc,
                                               double* __restrict__ Pe, int b,
                                               const double4* __restrict__ pos2 = nullptr)
{
    constexpr int NV = TEA ? 10 : 3;
    __shared__ double4 s_pos[PS_TILE];
    __shared__ double4 s_pos2[MEAN ? PS_TILE : 1];
    __shared__ double s_f[3 * PS_TILE];
    __shared__ double s_x[TEA ? 3 * PS_TILE : 1];
    __shared__ double s_col[4][NV][PS_TILE];
    __shared__ double s_row[4][NV][PS_TILE];
    const int t = threadIdx.x, w = threadIdx.y, tid = w * PS_TILE + t;
    int J = (int)((sqrt(8.0 * b + 1.0) - 1.0) * 0.5);
    while (J * (J + 1) / 2 > b) --J;
    while ((J + 1) * (J + 2) / 2 <= b) ++J;
    const int I = b - J * (J + 1) / 2;
    const bool diag = I == J;
    const int i = I * PS_TILE + t, j0 = J * PS_TILE;
    if (tid < PS_TILE) {
        const int j = j0 + tid;
        const bool ok = j < N;
        const double4 p = ok ? pos[j] : make_double4(0, 0, 0, 1);
        s_pos[tid] = p;
        if constexpr (MEAN) s_pos2[tid] = ok ? pos2[j] : make_double4(0, 0, 0, 1);
#pragma unroll
        for (int k = 0; k < 3; ++k) s_f[3 * tid + k] = ok ? F[3 * j + k] : 0.0;
        if constexpr (TEA) {
            const double sx = sqrt(2.0 * dt / (preDii / p.w));
#pragma unroll
            for (int k = 0; k < 3; ++k) s_x[3 * tid + k] = ok ? xi[3 * j + k] * sx : 0.0;
        }
    }
#pragma unroll
    for (int v = 0; v < NV; ++v) s_col[w][v][t] = 0.0;
    const bool iv = i < N;
    const double4 pi = iv ? pos[i] : make_double4(0, 0, 0, 1);
    double4 pi2 = pi;
    if constexpr (MEAN) pi2 = iv ? pos2[i] : make_double4(0, 0, 0, 1);
    double Fi[3] = {0, 0, 0}, Xi[3] = {0, 0, 0};
    if (iv && !diag) {
#pragma unroll
        for (int k = 0; k < 3; ++k) Fi[k] = F[3 * i + k];
        if constexpr (TEA) {
            const double sx = sqrt(2.0 * dt / (preDii / pi.w));
#pragma unroll
            for (int k = 0; k < 3; ++k) Xi[k] = xi[3 * i + k] * sx;
        }
    }
    const double inv_di = pi.w / preDii;  // 1 / D_ii, isotropic (diag_D)
    __syncthreads();
    double acc[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) acc[v] = 0.0;
    for (int sh = w; sh < PS_TILE; sh += 4) {
        const int jl = (t + sh) & (PS_TILE - 1);
        if (!iv || j0 + jl >= N) continue;
        const bool self = diag && sh == 0;
        const double4 pj = s_pos[jl];
        Blk6 B = pair_block<T, EW>(pi, pj, self, preDii, preDij);
        if constexpr (MEAN) {
            const Blk6 B2 = pair_block<T, EW>(pi2, s_pos2[jl], self, preDii, preDij);
            B.xx = (B2.xx + B.xx) / 2;
            B.yy = (B2.yy + B.yy) / 2;
            B.zz = (B2.zz + B.zz) / 2;
            B.xy = (B2.xy + B.xy) / 2;
            B.xz = (B2.xz + B.xz) / 2;
            B.yz = (B2.yz + B.yz) / 2;
        }
        const double fx = s_f[3 * jl], fy = s_f[3 * jl + 1], fz = s_f[3 * jl + 2];
        acc[0] += B.xx * fx + B.xy * fy + B.xz * fz;
        acc[1] += B.xy * fx + B.yy * fy + B.yz * fz;
        acc[2] += B.xz * fx + B.yz * fy + B.zz * fz;
        double q0 = 0, q1 = 0, q2 = 0, e = 0;
        if constexpr (TEA) {
            if (!self) {
                const double X0 = s_x[3 * jl], X1 = s_x[3 * jl + 1], X2 = s_x[3 * jl + 2];
                acc[3] += B.xx * X0 + B.xy * X1 + B.xz * X2;
                acc[4] += B.xy * X0 + B.yy * X1 + B.yz * X2;
                acc[5] += B.xz * X0 + B.yz * X1 + B.zz * X2;
                q0 = B.xx * B.xx + B.xy * B.xy + B.xz * B.xz;
                q1 = B.xy * B.xy + B.yy * B.yy + B.yz * B.yz;
                q2 = B.xz * B.xz + B.yz * B.yz + B.zz * B.zz;
                const double inv_dj = pj.w / preDii;
                acc[6] += q0 * inv_dj;
                acc[7] += q1 * inv_dj;
                acc[8] += q2 * inv_dj;
                e = B.xx + B.yy + B.zz + 2.0 * (B.xy + B.xz + B.yz);
                acc[9] += e;
            }
        }
        if (!diag) {
            double* c = &s_col[w][0][jl];
            c[0] += B.xx * Fi[0] + B.xy * Fi[1] + B.xz * Fi[2];
            c[PS_TILE] += B.xy * Fi[0] + B.yy * Fi[1] + B.yz * Fi[2];
            c[2 * PS_TILE] += B.xz * Fi[0] + B.yz * Fi[1] + B.zz * Fi[2];
            if constexpr (TEA) {
                c[3 * PS_TILE] += B.xx * Xi[0] + B.xy * Xi[1] + B.xz * Xi[2];
                c[4 * PS_TILE] += B.xy * Xi[0] + B.yy * Xi[1] + B.yz * Xi[2];
                c[5 * PS_TILE] += B.xz * Xi[0] + B.yz * Xi[1] + B.zz * Xi[2];
                c[6 * PS_TILE] += q0 * inv_di;
                c[7 * PS_TILE] += q1 * inv_di;
                c[8 * PS_TILE] += q2 * inv_di;
                c[9 * PS_TILE] += e;
            }
        }
    }
#pragma unroll
    for (int v = 0; v < NV; ++v) s_row[w][v][t] = acc[v];
    __syncthreads();
#define sum4(a, v) (((a)[0][v][t] + (a)[1][v][t]) + ((a)[2][v][t] + (a)[3][v][t]))
    if (w == 0 && iv) {
        const long long o = (long long)J * 3 * N + 3 * i;
#pragma unroll
        for (int k = 0; k < 3; ++k) Pf[o + k] = sum4(s_row, k);
        if constexpr (TEA) {
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                Px[o + k] = sum4(s_row, 3 + k);
                Pc[o + k] = sum4(s_row, 6 + k);
            }
        }
    }
    if (w == 1 && !diag && j0 + t < N) {
        const long long o = (long long)I * 3 * N + 3 * (j0 + t);
#pragma unroll
        for (int k = 0; k < 3; ++k) Pf[o + k] = sum4(s_col, k);
        if constexpr (TEA) {
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                Px[o + k] = sum4(s_col, 3 + k);
                Pc[o + k] = sum4(s_col, 6 + k);
            }
        }
    }
    if constexpr (TEA) {
        if (w == 2) {
            double e = sum4(s_row, 9);
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) e += __shfl_xor(e, off);
            if (t == 0) Pe[b] = diag ? e : 2.0 * e;
        }
    }
#undef sum4
}

template <typename T, int EW, bool TEA, bool MEAN = false>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 8)))
pair_sums_kernel(const double4* __restrict__ pos, const double* __restrict__ F,
                 const double* __restrict__ xi, int N, double preDii, double preDij, double dt,
                 double* __restrict__ Pf, double* __restrict__ Px, double* __restrict__ Pc,
                 double* __restrict__ Pe, const double4* __restrict__ pos2)
{
    pair_sums_body<T, EW, TEA, MEAN>(pos, F, xi, N, preDii, preDij, dt, Pf, Px, Pc, Pe, blockIdx.x, pos2);
}

// Free-space D.F of every chain of a batch: workgroup x of chain c is the chain's tile pair x.
template <typename T>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 8)))
pair_sums_batch_kernel(const double4* __restrict__ pos, const double* __restrict__ F, double preDii,
                       double preDij, const BatchChain* __restrict__ ch, const int* __restrict__ prefix,
                       int nch, double* __restrict__ Pp)
{
    const int c = batch_find(prefix, nch, blockIdx.x);
    const BatchChain bc = ch[c];
    pair_sums_body<T, 0, false>(pos + bc.bead0, F + 3LL * bc.bead0, nullptr, bc.N, preDii, preDij, 0.0,
                                Pp + bc.pp, nullptr, nullptr, nullptr, blockIdx.x - prefix[c]);
}

template <typename T>
int launch_pair_sums_batch(const double4* d_pos, const double* d_F, double preDii, double preDij,
                           const BatchPlan& plan, double* d_Pp, hipStream_t st)
{
    hipLaunchKernelGGL(pair_sums_batch_kernel<T>, dim3(plan.pair_blocks), dim3(PS_TILE, 4), 0, st, d_pos,
                       d_F, preDii, preDij, plan.ch, plan.pair_pre, plan.nch, d_Pp);
    BDH_LAUNCH_CHECK();
    return 0;
}

template <typename T>
int launch_pair_sums(const double4* d_pos, const double* d_F, const double* d_xi, int N,
                     double preDii, double preDij, double dt, bool ewald, bool tea, int splits,
                     double* d_Pf, double* d_Px, double* d_Pc, double* d_Pe, hipStream_t st,
                     const double4* d_pos2)
{
    BDH_CHECK(splits == pair_splits(N, ewald), "launch_pair_sums: split count");
    BDH_CHECK(!d_pos2 || tea, "launch_pair_sums: the mean tensor only for TEA sums");
    const dim3 grid(pair_eparts(N)), blk(PS_TILE, 4);
    const int ew = ewald ? (g_uploaded_method == EW_SMITH ? 2 : 1) : 0;
#define PS_LAUNCH(E, TE)                                                                          \
    hipLaunchKernelGGL((pair_sums_kernel<T, E, TE>), grid, blk, 0, st, d_pos, d_F, d_xi, N, preDii, \
                       preDij, dt, d_Pf, d_Px, d_Pc, d_Pe, (const double4*)nullptr)
#define PS_MEAN(E)                                                                                 \
    hipLaunchKernelGGL((pair_sums_kernel<T, E, true, true>), grid, blk, 0, st, d_pos, d_F, d_xi, N, \
                       preDii, preDij, dt, d_Pf, d_Px, d_Pc, d_Pe, d_pos2)
    if (d_pos2) {
        if (ew == 1) PS_MEAN(1);
        else if (ew == 2) PS_MEAN(2);
        else PS_MEAN(0);
    } else if (ew == 1) {
        if (tea) PS_LAUNCH(1, true); else PS_LAUNCH(1, false);
    } else if (ew == 2) {
        if (tea) PS_LAUNCH(2, true); else PS_LAUNCH(2, false);
    } else {
        if (tea) PS_LAUNCH(0, true); else PS_LAUNCH(0, false);
    }
#undef PS_LAUNCH
#undef PS_MEAN
    BDH_LAUNCH_CHECK();
    return 0;
}

// ------------------------------------------------------------------ single block
// One 3x3 block of D in FP64 (compute_Dij, diff_tensor.c:437-470, and the block functions it
// dispatches to: diff_tensor_calc_pos :45-92, the Smith and Beenakker blocks :102-422), EW as
// pair_block, as nine column-major values (the block is symmetric).
template <int EW>
__global__ void single_block_kernel(const double4* __restrict__ pos, int i, int j, double preDii,
                                    double preDij, double* __restrict__ out)
{
    if (threadIdx.x != 0) return;
    const Blk6 B = pair_block<double, EW>(pos[i], pos[j], i == j, preDii, preDij);
    out[0] = B.xx; out[1] = B.xy; out[2] = B.xz;
    out[3] = B.xy; out[4] = B.yy; out[5] = B.yz;
    out[6] = B.xz; out[7] = B.yz; out[8] = B.zz;
}

int launch_single_block(const double4* d_pos, int i, int j, int ew, double preDii, double preDij,
                        double* d_out, hipStream_t st)
{
    if (ew == 1)
        hipLaunchKernelGGL(single_block_kernel<1>, dim3(1), dim3(64), 0, st, d_pos, i, j, preDii, preDij, d_out);
    else if (ew == 2)
        hipLaunchKernelGGL(single_block_kernel<2>, dim3(1), dim3(64), 0, st, d_pos, i, j, preDii, preDij, d_out);
    else
        hipLaunchKernelGGL(single_block_kernel<0>, dim3(1), dim3(64), 0, st, d_pos, i, j, preDii, preDij, d_out);
    BDH_LAUNCH_CHECK();
    return 0;
}

// ------------------------------------------------------------------ Ewald assembly
// Lower triangle of the Beenakker-Ewald D (rows of i, columns of j, i >= j), the same
// tiling as the free-space assembly: (64 row beads) x (16 column beads) per workgroup.
// With pos2, the mean of the blocks at pos and pos2 (the IG-T tensor).  EW as pair_block.
template <typename T, int EW>
__global__ void __launch_bounds__(256)
ewald_assemble_kernel(const double4* __restrict__ pos, const double4* __restrict__ pos2, int N,
                      double preDii, double preDij, T* __restrict__ A, long long ld)
{
    const int i0 = blockIdx.x * 64, j0 = blockIdx.y * 16;
    if (j0 > i0 + 63) return;
    const int i = i0 + threadIdx.x;
    if (i >= N) return;
    const double4 pi = pos[i];
    for (int q = 0; q < 4; ++q) {
        const int j = j0 + threadIdx.y * 4 + q;
        if (j >= N || j > i) continue;
        Blk6 B = pair_block<T, EW>(pi, pos[j], i == j, preDii, preDij);
        if (pos2 && i != j) {
            const double4 qi = pos2[i], qj = pos2[j];
            const Blk6 B2 = pair_block<T, EW>(qi, qj, false, preDii, preDij);
            B.xx = (B.xx + B2.xx) / 2; B.yy = (B.yy + B2.yy) / 2; B.zz = (B.zz + B2.zz) / 2;
            B.xy = (B.xy + B2.xy) / 2; B.xz = (B.xz + B2.xz) / 2; B.yz = (B.yz + B2.yz) / 2;
        }
        T* c0 = A + (long long)(3 * j) * ld + 3 * i;
        c0[0] = (T)B.xx; c0[1] = (T)B.xy; c0[2] = (T)B.xz;
        c0[ld] = (T)B.xy; c0[ld + 1] = (T)B.yy; c0[ld + 2] = (T)B.yz;
        c0[2 * ld] = (T)B.xz; c0[2 * ld + 1] = (T)B.yz; c0[2 * ld + 2] = (T)B.zz;
    }
}

template <typename T>
int launch_ewald_assemble(const double4* d_pos, int N, double preDii, double preDij, T* d_A,
                          long long ld, hipStream_t st, const double4* d_pos2)
{
    dim3 grid(ceil_div(N, 64), ceil_div(N, 16));
    if (g_uploaded_method == EW_SMITH)
        hipLaunchKernelGGL((ewald_assemble_kernel<T, 2>), grid, dim3(64, 4), 0, st, d_pos, d_pos2, N,
                           preDii, preDij, d_A, ld);
    else
        hipLaunchKernelGGL((ewald_assemble_kernel<T, 1>), grid, dim3(64, 4), 0, st, d_pos, d_pos2, N,
                           preDii, preDij, d_A, ld);
    BDH_LAUNCH_CHECK();
    return 0;
}

// ------------------------------------------------------------------ TEA closing
// One block: eps from the per-workgroup partials of the pair sums (each unordered pair
// counted twice, as from both rows), then beta; out[0] = beta, out[1] = eps.
__global__ void __launch_bounds__(256)
tea_beta_kernel(const double* __restrict__ Pe, int S, int N, double* __restrict__ out)
{
    __shared__ double s[256];
    double a = 0.0;
    for (int k = threadIdx.x; k < S; k += 256) a += Pe[k];
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

template int launch_pair_sums<float>(const double4*, const double*, const double*, int, double, double, double, bool, bool, int, double*, double*, double*, double*, hipStream_t, const double4*);
template int launch_pair_sums<double>(const double4*, const double*, const double*, int, double, double, double, bool, bool, int, double*, double*, double*, double*, hipStream_t, const double4*);
template int launch_pair_sums_batch<float>(const double4*, const double*, double, double, const BatchPlan&, double*, hipStream_t);
template int launch_pair_sums_batch<double>(const double4*, const double*, double, double, const BatchPlan&, double*, hipStream_t);
template int launch_ewald_assemble<float>(const double4*, int, double, double, float*, long long, hipStream_t,
                                          const double4*);
template int launch_ewald_assemble<double>(const double4*, int, double, double, double*, long long, hipStream_t,
                                           const double4*);

}  // namespace bdh
