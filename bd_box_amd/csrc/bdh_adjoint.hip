// Reverse-mode derivative (vector-Jacobian product) of one FP64 Ermak-McCammon step with
// Cholesky noise refactored at the step's positions (the step of comp_step_ermak /
// compute_hydro_ermak, src/diff_alg/diff_comp.c:152-179, diff_equation.c:216-286,
// diff_rand.c:64-85):
//
//   r' = r + c D(r) F(r) + s L(r) xi,        c = dt / kT,  s = sqrt(2 dt),  D = L L^T
//
// For a cotangent g of r' the cotangent of r is
//
//   rbar = g + grad_r <U, D(r)> - c H(r) (D g)
//   U    = c sym(g F^T) + sym(S),  S = L^-T Phi(L^T Lbar) L^-1,  Lbar = s tril(g xi^T)
//
// H is the Hessian of the potential energy (dF/dr = -H), Phi the lower triangle with the
// diagonal halved (derivative of the Cholesky factor, Murray 2016).  With a = L^T g,
// Phi(L^T Lbar) = s Phi(a xi^T), so S comes from two triangular solves with n right-hand
// sides: Y1 = L^-T (s Phi(a xi^T)), Y2 = L^-T Y1^T = S^T.
//
// Kernels (FP64; the two solves are the O(n^3) part):
//   adj_lt_g_kernel         a = L^T g
//   adj_diag_inv_kernel     inverses of L's 64 x 64 diagonal blocks
//   adj_rhs_kernel, adj_trsm_super_kernel, adj_trsm_update_kernel
//                           L^T X = B blocked by 64 rows, four block rows per step (their
//                           triangle, then one K = 256 update of the rows above), B =
//                           s Phi(a xi^T) or Y1^T
//   adj_symmetrize_kernel   Z = Y2 + Y2^T
//   adj_contract_kernel     grad_r <U, D(r)> over the RPY blocks (diff_tensor.c:45-92), column splits
//   adj_hvp_pairs_kernel,   g + contraction - c H y  (pair and bond Hessians, the forward
//   adj_hvp_finish_kernel   pass's cutoffs and bonded-partner scaling; bond.c:43-85)
#include <algorithm>

#include "bdh_common.hpp"
#include "bdh_kernels.hpp"
#include "bdh_mfma.hpp"
#include "bdh_pairterm.hpp"

namespace bdh {

__global__ void __launch_bounds__(256)
adj_lt_g_kernel(const double* __restrict__ L, long long ld, int n, const double* __restrict__ g,
                double* __restrict__ a)
{
    __shared__ double s_red[256];
    const int i = blockIdx.x;
    const double* col = L + (long long)i * ld;
    double acc = 0.0;
    for (int k = i + threadIdx.x; k < n; k += 256) acc += col[k] * g[k];
    s_red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s_red[threadIdx.x] += s_red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) a[i] = s_red[0];
}

// ---- L^T X = B for np right-hand sides (np = the padded order: L's padding is an identity
// block, so the padded rows and columns of X stay zero), blocked by 64 rows from the last
// block up: X_J <- L_JJ^-T X_J, then X_I -= L_JI^T X_J for every I < J.  X is column-major
// (ld np), solved in place after the right-hand sides are written into it.
constexpr int ADJ_NB = 64;
constexpr int ADJ_SUPER = 4;   // block rows of L^T X = B per triangular launch / update K

// mode 0: X(i, q) = s a_i xi_q (i > q), s a_q xi_q / 2 (i = q), 0 (i < q or beyond n)
// mode 1: X = Y^T (64 x 64 tiles through LDS)
__global__ void __launch_bounds__(256)
adj_rhs_kernel(int mode, int np, int n, const double* __restrict__ a, const double* __restrict__ xi,
               double s, const double* __restrict__ Y, double* __restrict__ X)
{
    __shared__ double t[ADJ_NB][ADJ_NB + 1];
    const int i0 = blockIdx.x * ADJ_NB, q0 = blockIdx.y * ADJ_NB;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    if (mode == 0) {
        for (int qq = ty; qq < ADJ_NB; qq += 4) {
            const int i = i0 + tx, q = q0 + qq;
            double v = 0.0;
            if (i < n && q < n) v = i > q ? s * a[i] * xi[q] : (i == q ? 0.5 * s * a[i] * xi[q] : 0.0);
            X[(long long)q * np + i] = v;
        }
        return;
    }
    // X(i, q) = Y(q, i): read the tile of Y with rows q0.., columns i0.. (column-major)
#pragma unroll
    for (int u = 0; u < ADJ_NB / 4; ++u) t[ty + 4 * u][tx] = Y[(long long)(i0 + ty + 4 * u) * np + q0 + tx];
    __syncthreads();
#pragma unroll
    for (int u = 0; u < ADJ_NB / 4; ++u) X[(long long)(q0 + ty + 4 * u) * np + i0 + tx] = t[tx][ty + 4 * u];
}

// P = sA^T sB for 64 x 64 LDS tiles (sA[k][i], sB[k][q], k < 64) on v_mfma_f64_16x16x4_f64:
// wavefront w owns the 32 x 32 quarter (rows 32 (w / 2), columns 32 (w % 2)) as 2 x 2 blocks;
// accumulator r of lane l holds row 4 r + l / 16, column l % 16 of its block.
__device__ __forceinline__ void tile_tn_mfma_acc(const double (*sA)[ADJ_NB + 1],
                                                 const double (*sB)[ADJ_NB + 1], doublex4 acc[2][2])
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r0 = 32 * (wave >> 1), c0 = 32 * (wave & 1);
    const int lr = lane & 15, lk = lane >> 4;
#pragma unroll 4
    for (int k = 0; k < ADJ_NB; k += 4) {
        double av[2], bv[2];
#pragma unroll
        for (int a = 0; a < 2; ++a) av[a] = sA[k + lk][r0 + 16 * a + lr];
#pragma unroll
        for (int b = 0; b < 2; ++b) bv[b] = sB[k + lk][c0 + 16 * b + lr];
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b) acc[a][b] = mfma_f64_16x16x4(av[a], bv[b], acc[a][b]);
    }
}

__device__ __forceinline__ void tile_tn_mfma(const double (*sA)[ADJ_NB + 1], const double (*sB)[ADJ_NB + 1],
                                             doublex4 acc[2][2])
{
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = doublex4{0.0, 0.0, 0.0, 0.0};
    tile_tn_mfma_acc(sA, sB, acc);
}

// A 64 x 64 tile of a column-major matrix (column c at P + c ld, rows contiguous) through
// registers: every thread's 16 loads are issued before any of them is stored to LDS, as
// sT[k][c] = P(k, c).
struct TileRegs {
    double v[16];
};
__device__ __forceinline__ void tile_load(const double* __restrict__ P, long long ld, TileRegs& t)
{
    const int k = threadIdx.x & 63, c = threadIdx.x >> 6;
#pragma unroll
    for (int u = 0; u < 16; ++u) t.v[u] = P[(long long)(c + 4 * u) * ld + k];
}
__device__ __forceinline__ void tile_to_lds(const TileRegs& t, double (*sT)[ADJ_NB + 1])
{
    const int k = threadIdx.x & 63, c = threadIdx.x >> 6;
#pragma unroll
    for (int u = 0; u < 16; ++u) sT[k][c + 4 * u] = t.v[u];
}

// sT[i][q] = P(i, q) from tile_tn_mfma's accumulators.
__device__ __forceinline__ void tile_store_lds(double (*sT)[ADJ_NB + 1], const doublex4 acc[2][2])
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r0 = 32 * (wave >> 1), c0 = 32 * (wave & 1);
    const int lr = lane & 15, lk = lane >> 4;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) sT[r0 + 16 * a + lk + 4 * r][c0 + 16 * b + lr] = acc[a][b][r];
}

// X(i0 + i, q0 + q) = P(i, q) (sub = false) or -= P(i, q) (sub = true) from tile_tn_mfma's
// accumulators.
__device__ __forceinline__ void tile_store(double* __restrict__ X, int np, int i0, int q0,
                                           const doublex4 acc[2][2], bool sub)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r0 = 32 * (wave >> 1), c0 = 32 * (wave & 1);
    const int lr = lane & 15, lk = lane >> 4;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            double* xc = X + (long long)(q0 + c0 + 16 * b + lr) * np + i0 + r0 + 16 * a + lk;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                if (sub) xc[4 * r] -= acc[a][b][r];
                else xc[4 * r] = acc[a][b][r];
            }
        }
}

// Inverses of the 64 x 64 diagonal blocks of L (workgroup J, thread c: column c of
// L_JJ^-1 by forward substitution), column-major into Minv + 4096 J.
__global__ void __launch_bounds__(64)
adj_diag_inv_kernel(const double* __restrict__ L, long long ld, double* __restrict__ Minv)
{
    __shared__ double sL[ADJ_NB][ADJ_NB + 1];   // sL[i][k] = L(64 J + i, 64 J + k)
    const int j0 = blockIdx.x * ADJ_NB, c = threadIdx.x;
#pragma unroll 16
    for (int k = 0; k < ADJ_NB; ++k) sL[c][k] = L[(long long)(j0 + k) * ld + j0 + c];
    __syncthreads();
    double m[ADJ_NB];
#pragma unroll
    for (int i = 0; i < ADJ_NB; ++i) {
        double b = (i == c) ? 1.0 : 0.0;
#pragma unroll
        for (int k = 0; k < i; ++k) b -= sL[i][k] * m[k];
        m[i] = (i >= c) ? b / sL[i][i] : 0.0;
    }
    double* out = Minv + (long long)blockIdx.x * ADJ_NB * ADJ_NB + (long long)c * ADJ_NB;
#pragma unroll
    for (int i = 0; i < ADJ_NB; ++i) out[i] = m[i];
}

// Block rows Jlo..Jhi (at most ADJ_SUPER) of L^T X = B for one ADJ_SC-column tile of X
// (workgroup = column tile), the rows below Jhi already applied.  The tile's rows of these
// block rows stay in LDS: for J = Jhi down to Jlo, X_J <- M_J^T X_J (M_J = L_JJ^-1), then
// X_I -= L_JI^T X_J for Jlo <= I < J.  The A operands come in that fixed order (M_J, then the
// L_JI), each loaded into registers while the previous one is multiplied.  Wavefront w owns
// rows 16 w.. of every 64 x ADJ_SC product (two 16 x 16 blocks on v_mfma_f64_16x16x4_f64).
constexpr int ADJ_SC = 32;

__global__ void __launch_bounds__(256)
adj_trsm_super_kernel(const double* __restrict__ L, long long ld, const double* __restrict__ Minv, int np,
                      int Jlo, int Jhi, double* __restrict__ X)
{
    __shared__ double sX[ADJ_SUPER][ADJ_NB][ADJ_SC + 1];   // sX[j][k][q] = X(64 (Jlo + j) + k, q0 + q)
    __shared__ double sA[2][ADJ_NB][ADJ_NB + 1];          // sA[.][k][i] = A(k, i)
    const int q0 = blockIdx.x * ADJ_SC;
    const int nbr = Jhi - Jlo + 1;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int lr = lane & 15, lk = lane >> 4;
    const int tk = threadIdx.x & 63, tc = threadIdx.x >> 6;   // loads: row tk, columns tc + 4 u
    TileRegs ta;
    tile_load(Minv + (long long)Jhi * ADJ_NB * ADJ_NB, ADJ_NB, ta);
#pragma unroll
    for (int j = 0; j < ADJ_SUPER; ++j) {
        if (j < nbr) {
            const double* src = X + (long long)q0 * np + (long long)(Jlo + j) * ADJ_NB + tk;
            double v[ADJ_SC / 4];
#pragma unroll
            for (int u = 0; u < ADJ_SC / 4; ++u) v[u] = src[(long long)(tc + 4 * u) * np];
#pragma unroll
            for (int u = 0; u < ADJ_SC / 4; ++u) sX[j][tk][tc + 4 * u] = v[u];
        }
    }
    int step = 0;
    for (int J = Jhi; J >= Jlo; --J) {
        const int j = J - Jlo;
        for (int I = J; I >= Jlo; --I, ++step) {
            const int buf = step & 1;
            tile_to_lds(ta, sA[buf]);
            __syncthreads();   // sA[buf] and the last writes of sX are visible
            // the next A operand: L_J,I-1, else M_J-1
            const int nJ = I > Jlo ? J : J - 1, nI = I > Jlo ? I - 1 : J - 1;
            if (nJ >= Jlo) {
                if (nI == nJ) tile_load(Minv + (long long)nJ * ADJ_NB * ADJ_NB, ADJ_NB, ta);
                else tile_load(L + (long long)nI * ADJ_NB * ld + (long long)nJ * ADJ_NB, ld, ta);
            }
            doublex4 acc[2] = {doublex4{0.0, 0.0, 0.0, 0.0}, doublex4{0.0, 0.0, 0.0, 0.0}};
#pragma unroll 4
            for (int k = 0; k < ADJ_NB; k += 4) {
                const double av = sA[buf][k + lk][16 * wave + lr];
#pragma unroll
                for (int b = 0; b < 2; ++b) acc[b] = mfma_f64_16x16x4(av, sX[j][k + lk][16 * b + lr], acc[b]);
            }
            if (I == J) {
                __syncthreads();   // every wavefront is done reading X_J
#pragma unroll
                for (int b = 0; b < 2; ++b)
#pragma unroll
                    for (int r = 0; r < 4; ++r) sX[j][16 * wave + lk + 4 * r][16 * b + lr] = acc[b][r];
            } else {
                const int i = I - Jlo;
#pragma unroll
                for (int b = 0; b < 2; ++b)
#pragma unroll
                    for (int r = 0; r < 4; ++r) sX[i][16 * wave + lk + 4 * r][16 * b + lr] -= acc[b][r];
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < ADJ_SUPER; ++j) {
        if (j < nbr) {
            double* dst = X + (long long)q0 * np + (long long)(Jlo + j) * ADJ_NB + tk;
#pragma unroll
            for (int u = 0; u < ADJ_SC / 4; ++u) dst[(long long)(tc + 4 * u) * np] = sX[j][tk][tc + 4 * u];
        }
    }
}

// X_I -= sum_{kk < kb} L_{J+kk,I}^T X_{J+kk}: workgroup (I, column tile), K = 64 kb on MFMA,
// the next 64-row slice of both operands loaded while the current one is multiplied.
__global__ void __launch_bounds__(256)
adj_trsm_update_kernel(const double* __restrict__ L, long long ld, int np, int J, int kb,
                       double* __restrict__ X)
{
    __shared__ double sA[ADJ_NB][ADJ_NB + 1];   // sA[k][i] = L(64 (J + kk) + k, 64 I + i)
    __shared__ double sB[ADJ_NB][ADJ_NB + 1];   // sB[k][q] = X(64 (J + kk) + k, q0 + q)
    const int I = blockIdx.x, q0 = blockIdx.y * ADJ_NB;
    const int i0 = I * ADJ_NB;
    const double* La = L + (long long)i0 * ld;
    const double* Xq = X + (long long)q0 * np;
    doublex4 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = doublex4{0.0, 0.0, 0.0, 0.0};
    TileRegs ta, tb;
    tile_load(La + J * ADJ_NB, ld, ta);
    tile_load(Xq + J * ADJ_NB, np, tb);
    for (int kk = 0; kk < kb; ++kk) {
        if (kk) __syncthreads();   // the previous slice's products are done with sA and sB
        tile_to_lds(ta, sA);
        tile_to_lds(tb, sB);
        __syncthreads();
        if (kk + 1 < kb) {
            const int j1 = (J + kk + 1) * ADJ_NB;
            tile_load(La + j1, ld, ta);
            tile_load(Xq + j1, np, tb);
        }
        tile_tn_mfma_acc(sA, sB, acc);
    }
    tile_store(X, np, i0, q0, acc, true);
}

// Z = Y + Y^T (np x np, column-major; 64 x 64 tiles through LDS), so that the contraction
// reads one coalesced matrix.
__global__ void __launch_bounds__(256)
adj_symmetrize_kernel(int np, const double* __restrict__ Y, double* __restrict__ Z)
{
    __shared__ double t[ADJ_NB][ADJ_NB + 1];
    const int i0 = blockIdx.x * ADJ_NB, q0 = blockIdx.y * ADJ_NB;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    // t[ii][tx] = Y(q0 + tx, i0 + ii) = Y^T(i0 + ii, q0 + tx)
#pragma unroll
    for (int u = 0; u < ADJ_NB / 4; ++u) t[ty + 4 * u][tx] = Y[(long long)(i0 + ty + 4 * u) * np + q0 + tx];
    __syncthreads();
#pragma unroll
    for (int u = 0; u < ADJ_NB / 4; ++u) {
        const int qq = ty + 4 * u;
        const long long e = (long long)(q0 + qq) * np + i0 + tx;
        Z[e] = Y[e] + t[tx][qq];
    }
}

// f, f', g, g' of the RPY block B = f I + g e e^T at separation r (radii ai, aj).
__device__ __forceinline__ void rpy_fg(double r, double ai, double aj, double preDii, double preDij,
                                       double& fp, double& gv, double& gp)
{
    if (r < ai + aj) {
        const double abar = 0.5 * (ai + aj);
        fp = -9.0 * preDii / (32.0 * abar * abar);
        gv = preDii / abar * (3.0 * r / (32.0 * abar));
        gp = 3.0 * preDii / (32.0 * abar * abar);
    } else {
        const double sig2 = ai * ai + aj * aj, r2 = r * r;
        fp = preDij * (-1.0 / r2 - sig2 / (r2 * r2));
        gv = preDij * (1.0 / r - sig2 / (r2 * r));
        gp = preDij * (-1.0 / r2 + 3.0 * sig2 / (r2 * r2));
    }
}

// Term of pair (k, j) of out_k = sum_{j != k} grad_d <W_kj, B(d)> at d = r_k - r_j, with
// W_kj = c (g_k F_j^T + F_k g_j^T) + Z(3k.., 3j..),  Z = Y2 + Y2^T  (W = 2 U_kj).
__device__ __forceinline__ void contract_pair(const double4& pk, const double4& pj, int k, int j,
                                              double preDii, double preDij, const double* __restrict__ Z,
                                              long long ldz, const double* gk, const double* Fk,
                                              const double* __restrict__ g, const double* __restrict__ F,
                                              double c, double* acc)
{
    const double d[3] = {pk.x - pj.x, pk.y - pj.y, pk.z - pj.z};
    const double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    const double e[3] = {d[0] / r, d[1] / r, d[2] / r};
    double fp, gv, gp;
    rpy_fg(r, pk.w, pj.w, preDii, preDij, fp, gv, gp);
    double W[3][3];
#pragma unroll
    for (int u = 0; u < 3; ++u)
#pragma unroll
        for (int v = 0; v < 3; ++v) {
            const long long row = 3LL * k + u, col = 3LL * j + v;
            W[u][v] = c * (gk[u] * F[3 * j + v] + Fk[u] * g[3 * j + v]) + Z[col * ldz + row];
        }
    const double trW = W[0][0] + W[1][1] + W[2][2];
    double We[3], WTe[3];
#pragma unroll
    for (int u = 0; u < 3; ++u) {
        We[u] = W[u][0] * e[0] + W[u][1] * e[1] + W[u][2] * e[2];
        WTe[u] = W[0][u] * e[0] + W[1][u] * e[1] + W[2][u] * e[2];
    }
    const double eWe = e[0] * We[0] + e[1] * We[1] + e[2] * We[2];
    double v[3];
#pragma unroll
    for (int u = 0; u < 3; ++u) v[u] = We[u] + WTe[u];
    const double ev = e[0] * v[0] + e[1] * v[1] + e[2] * v[2];
#pragma unroll
    for (int u = 0; u < 3; ++u)
        acc[u] += (fp * trW + gp * eWe) * e[u] + (gv / r) * (v[u] - ev * e[u]);
}

// Workgroup (x, s): row beads k = 64 x + lane, column beads j of split s
// (js beads from s js); wave w takes j = s js + w, + 4, ...  The four waves' sums meet in
// LDS in wave order and land in P[s][3k + u]; the splits are summed afterwards in split
// order, so the result does not depend on scheduling.
__global__ void __launch_bounds__(256)
adj_contract_kernel(const double4* __restrict__ pos, int N, double preDii, double preDij,
                    const double* __restrict__ Z, long long ldz, const double* __restrict__ g,
                    const double* __restrict__ F, double c, int js, double* __restrict__ P)
{
    __shared__ double red[3][3][64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int k = blockIdx.x * 64 + lane;
    const int j0 = blockIdx.y * js, j1 = min(N, j0 + js);
    double acc[3] = {0.0, 0.0, 0.0};
    if (k < N) {
        const double4 pk = pos[k];
        const double gk[3] = {g[3 * k], g[3 * k + 1], g[3 * k + 2]};
        const double Fk[3] = {F[3 * k], F[3 * k + 1], F[3 * k + 2]};
        for (int j = j0 + wave; j < j1; j += 4) {
            if (j == k) continue;
            contract_pair(pk, pos[j], k, j, preDii, preDij, Z, ldz, gk, Fk, g, F, c, acc);
        }
    }
    if (wave > 0) {
#pragma unroll
        for (int u = 0; u < 3; ++u) red[wave - 1][u][lane] = acc[u];
    }
    __syncthreads();
    if (wave == 0 && k < N) {
        double* out = P + (long long)blockIdx.y * 3 * N;
#pragma unroll
        for (int u = 0; u < 3; ++u)
            out[3 * k + u] = ((acc[u] + red[0][u][lane]) + red[1][u][lane]) + red[2][u][lane];
    }
}

// Pair part of (H y)_k over the column beads of split s (the forward pass's cutoffs and
// bonded-partner scaling); workgroups, waves and the wave-order LDS sum as in
// adj_contract_kernel, partials in P[s][3k + u].
__global__ void __launch_bounds__(256)
adj_hvp_pairs_kernel(const double4* __restrict__ pos, const double* __restrict__ Q,
                     const double2* __restrict__ LJ, int N, Topology topo, ForceParams fp,
                     const double* __restrict__ y, int js, double* __restrict__ P)
{
    __shared__ double red[3][3][64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int k = blockIdx.x * 64 + lane;
    const int j0 = blockIdx.y * js, j1 = min(N, j0 + js);
    double acc[3] = {0.0, 0.0, 0.0};
    if (k < N) {
        const double4 pk = pos[k];
        const double2 ljk = LJ[k];
        const double qk = Q[k];
        const double yk[3] = {y[3 * k], y[3 * k + 1], y[3 * k + 2]};
        const int nb0 = fp.use_conn ? topo.conn_off[k] : 0;
        const int nb1 = fp.use_conn ? topo.conn_off[k + 1] : 0;
        const double lj6 = fp.lj_6_term;
        for (int j = j0 + wave; j < j1; j += 4) {
            if (j == k) continue;
            const double4 pj = pos[j];
            const double d[3] = {pk.x - pj.x, pk.y - pj.y, pk.z - pj.z};
            const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
            if (r2 > fp.far_r2) continue;
            const double r = sqrt(r2);
            const bool conn = bonded(topo, nb0, nb1, j);
            double d1 = 0.0, d2 = 0.0;
            if (fp.alpha_lj != 0.0) {
                const double2 ljj = LJ[j];
                const double sig = ljk.x + ljj.x;
                const double cut = fp.cutoff_lj > 0.0 ? fp.cutoff_lj : (double)1.122462f * sig;
                const double scale = conn ? fp.bond_lj_scale : 1.0;
                if (r <= cut && scale != 0.0) {
                    const double eps = sqrt(ljk.y * ljj.y);
                    double s6 = sig / r;
                    s6 *= s6;
                    s6 *= s6 * s6;
                    d1 += scale * fp.alpha_lj * eps * (-12.0 * s6 * s6 + 6.0 * lj6 * s6) / r;
                    d2 += scale * fp.alpha_lj * eps * (156.0 * s6 * s6 - 42.0 * lj6 * s6) / r2;
                }
            }
            if (fp.elec) {
                const double scale = conn ? fp.bond_c_scale : 1.0;
                if (r <= fp.cutoff_c && scale != 0.0) {
                    const double qq = fp.gam * qk * Q[j];
                    const double ex = exp(-fp.kappa_c * r);
                    const double kap = fp.kappa_c;
                    d1 += -scale * qq * ex * (1.0 / r2 + kap / r);
                    d2 += scale * qq * ex * (2.0 / (r2 * r) + 2.0 * kap / r2 + kap * kap / r);
                }
            }
            if (d1 == 0.0 && d2 == 0.0) continue;
            const double e[3] = {d[0] / r, d[1] / r, d[2] / r};
            const double dy[3] = {yk[0] - y[3 * j], yk[1] - y[3 * j + 1], yk[2] - y[3 * j + 2]};
            const double ed = e[0] * dy[0] + e[1] * dy[1] + e[2] * dy[2];
    #pragma unroll
            for (int u = 0; u < 3; ++u) acc[u] += d2 * ed * e[u] + (d1 / r) * (dy[u] - ed * e[u]);
        }
    }
    if (wave > 0) {
#pragma unroll
        for (int u = 0; u < 3; ++u) red[wave - 1][u][lane] = acc[u];
    }
    __syncthreads();
    if (wave == 0 && k < N) {
        double* out = P + (long long)blockIdx.y * 3 * N;
#pragma unroll
        for (int u = 0; u < 3; ++u)
            out[3 * k + u] = ((acc[u] + red[0][u][lane]) + red[1][u][lane]) + red[2][u][lane];
    }
}

// out_k = g_k + contraction_k - c (H y)_k: the pair partials summed in split order, then the
// bond Hessians (bond.c:43-85).
__global__ void __launch_bounds__(256)
adj_hvp_finish_kernel(const double4* __restrict__ pos, int N, Topology topo, const double* __restrict__ y,
                      const double* __restrict__ P, int S, const double* __restrict__ contract,
                      const double* __restrict__ g, double c, double* __restrict__ out)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= N) return;
    const double4 pk = pos[k];
    const double yk[3] = {y[3 * k], y[3 * k + 1], y[3 * k + 2]};
    double acc[3] = {0.0, 0.0, 0.0};
    for (int sp = 0; sp < S; ++sp) {
#pragma unroll
        for (int u = 0; u < 3; ++u) acc[u] += P[(long long)sp * 3 * N + 3 * k + u];
    }
    for (int b = topo.bond_off[k]; b < topo.bond_off[k + 1]; ++b) {
        const int p = topo.bond_p[b];
        const double r0 = topo.bond_parm[3 * b], rmax = topo.bond_parm[3 * b + 1];
        const double H = topo.bond_parm[3 * b + 2];
        const double4 pp = pos[p];
        const double d[3] = {pk.x - pp.x, pk.y - pp.y, pk.z - pp.z};
        const double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        const double e[3] = {d[0] / r, d[1] / r, d[2] / r};
        const double den = 1.0 - r * r / (rmax * rmax);
        const double b1 = H * (r - r0) / den;
        const double b2 = H * (den + (r - r0) * 2.0 * r / (rmax * rmax)) / (den * den);
        const double dy[3] = {yk[0] - y[3 * p], yk[1] - y[3 * p + 1], yk[2] - y[3 * p + 2]};
        const double ed = e[0] * dy[0] + e[1] * dy[1] + e[2] * dy[2];
#pragma unroll
        for (int u = 0; u < 3; ++u) acc[u] += b2 * ed * e[u] + (b1 / r) * (dy[u] - ed * e[u]);
    }
#pragma unroll
    for (int u = 0; u < 3; ++u) out[3 * k + u] = g[3 * k + u] + contract[3 * k + u] - c * acc[u];
}

__global__ void __launch_bounds__(256)
adj_promote_kernel(const float* __restrict__ A, long long count, double* __restrict__ Ld)
{
    const long long e = (long long)blockIdx.x * 1024 + 4 * threadIdx.x;
    if (e + 3 < count) {
        const float4 v = *reinterpret_cast<const float4*>(A + e);
        Ld[e] = v.x;
        Ld[e + 1] = v.y;
        Ld[e + 2] = v.z;
        Ld[e + 3] = v.w;
    } else {
        for (long long q = e; q < count; ++q) Ld[q] = A[q];
    }
}

int launch_adj_promote(const float* d_A, long long count, double* d_Ld, hipStream_t st)
{
    BDH_CHECK(count > 0 && count % 4 == 0, "adjoint: promote count");
    hipLaunchKernelGGL(adj_promote_kernel, dim3((unsigned)ceil_div_ll(count, 1024)), dim3(256), 0, st, d_A, count,
                       d_Ld);
    BDH_LAUNCH_CHECK();
    return 0;
}

int launch_adj_lt_g(const double* d_L, long long ld, int n, const double* d_g, double* d_a, hipStream_t st)
{
    hipLaunchKernelGGL(adj_lt_g_kernel, dim3(n), dim3(256), 0, st, d_L, ld, n, d_g, d_a);
    BDH_LAUNCH_CHECK();
    return 0;
}

int launch_adj_diag_inverses(const double* d_L, long long ld, int np, double* d_Minv, hipStream_t st)
{
    BDH_CHECK(np % ADJ_NB == 0 && ld >= np, "adjoint solve: padded order");
    hipLaunchKernelGGL(adj_diag_inv_kernel, dim3(np / ADJ_NB), dim3(ADJ_NB), 0, st, d_L, ld, d_Minv);
    BDH_LAUNCH_CHECK();
    return 0;
}

int launch_adj_upper_solve(const double* d_L, long long ld, int np, int n, int mode, const double* d_a,
                           const double* d_xi, double s, const double* d_B, const double* d_Minv,
                           double* d_X, hipStream_t st)
{
    BDH_CHECK(np % ADJ_NB == 0 && ld >= np, "adjoint solve: padded order");
    const int nb = np / ADJ_NB;
    hipLaunchKernelGGL(adj_rhs_kernel, dim3(nb, nb), dim3(256), 0, st, mode, np, n, d_a, d_xi, s, d_B, d_X);
    BDH_LAUNCH_CHECK();
    // ADJ_SUPER block rows at a time: their own triangle in one launch (a workgroup per column
    // tile), then every block row above them in one update launch with K = 64 ADJ_SUPER
    for (int Jhi = nb - 1; Jhi >= 0; Jhi -= ADJ_SUPER) {
        const int Jlo = std::max(0, Jhi - ADJ_SUPER + 1);
        hipLaunchKernelGGL(adj_trsm_super_kernel, dim3(np / ADJ_SC), dim3(256), 0, st, d_L, ld, d_Minv, np, Jlo,
                           Jhi, d_X);
        BDH_LAUNCH_CHECK();
        if (Jlo > 0) {
            hipLaunchKernelGGL(adj_trsm_update_kernel, dim3(Jlo, nb), dim3(256), 0, st, d_L, ld, np, Jlo,
                               Jhi - Jlo + 1, d_X);
            BDH_LAUNCH_CHECK();
        }
    }
    return 0;
}

int launch_adj_symmetrize(int np, const double* d_Y, double* d_Z, hipStream_t st)
{
    BDH_CHECK(np % ADJ_NB == 0, "adjoint: padded order");
    hipLaunchKernelGGL(adj_symmetrize_kernel, dim3(np / ADJ_NB, np / ADJ_NB), dim3(256), 0, st, np, d_Y, d_Z);
    BDH_LAUNCH_CHECK();
    return 0;
}

int adj_contract_splits(int N)
{
    const int rows = ceil_div(N, 64);
    return std::max(1, std::min({ADJ_CONTRACT_MAX_SPLITS, ceil_div(1024, rows), ceil_div(N, 4)}));
}

int launch_adj_contract(const double4* d_pos, int N, double preDii, double preDij, const double* d_Y2,
                        long long ldy, const double* d_g, const double* d_F, double c, double* d_part,
                        double* d_out, hipStream_t st)
{
    const int S = adj_contract_splits(N);
    const int js = ceil_div(N, S);
    hipLaunchKernelGGL(adj_contract_kernel, dim3(ceil_div(N, 64), S), dim3(256), 0, st, d_pos, N, preDii,
                       preDij, d_Y2, ldy, d_g, d_F, c, js, d_part);
    BDH_LAUNCH_CHECK();
    ReduceJobs jobs{};
    jobs.job[0] = ReduceJob{d_part, d_out, 3 * N, S, 0, 1.0};
    return launch_reduce_partials(jobs, 1, 3 * N, st);
}

int launch_adj_hvp(const double4* d_pos, const double* d_Q, const double2* d_LJ, int N,
                   const Topology& topo, const ForceParams& fp, const double* d_y,
                   const double* d_contract, const double* d_g, double c, double* d_part, double* d_out,
                   hipStream_t st)
{
    const int S = adj_contract_splits(N);
    const int js = ceil_div(N, S);
    hipLaunchKernelGGL(adj_hvp_pairs_kernel, dim3(ceil_div(N, 64), S), dim3(256), 0, st, d_pos, d_Q, d_LJ,
                       N, topo, fp, d_y, js, d_part);
    BDH_LAUNCH_CHECK();
    hipLaunchKernelGGL(adj_hvp_finish_kernel, dim3(ceil_div(N, 256)), dim3(256), 0, st, d_pos, N, topo, d_y,
                       (const double*)d_part, S, d_contract, d_g, c, d_out);
    BDH_LAUNCH_CHECK();
    return 0;
}


// ---- parameter cotangents of one step: theta_bar = (c D g)^T dF/dtheta for the per-bead LJ
// epsilon_k and radius R_k, the charges Q_k and the per-bond constants H_b.  With
// F_k = -sum_j phi'(r_kj) e_kj, pair (k, j) contributes -(y_k - y_j).e_kj dphi'/dtheta:
//   LJ  (LJ.c:37-64, sigma_kj = R_k + R_j, eps_kj = sqrt(eps_k eps_j), data.c:97-102)
//        phi' = a eps_kj (-12 s6^2 + 6 lj6 s6) / r                 -> d/deps_k = phi' / (2 eps_k)
//        d phi' / d sigma = a eps_kj (-144 s6^2 + 36 lj6 s6) / (sigma r)  (= d / dR_k)
//   DH  (electro.c:36-56)  phi' = -(gamma/eps_r) Q_k Q_j e^{-kappa r} (1/r^2 + kappa/r)
//        d phi' / dQ_k = -(gamma/eps_r) Q_j e^{-kappa r} (1/r^2 + kappa/r)
//   bond (bond.c:43-85)   phi' = H (r - r0) / (1 - r^2/rmax^2)     -> d/dH = (r - r0) / den
// Row k sums its pairs over one column split (four wavefronts interleave the columns and
// meet in LDS in wave order); P[split][3][N] holds (eps, R, Q) sums, combined in split order.
__global__ void __launch_bounds__(256)
adj_param_pairs_kernel(const double4* __restrict__ pos, const double* __restrict__ Q,
                       const double2* __restrict__ LJ, int N, Topology topo, ForceParams fp,
                       const double* __restrict__ y, int js, double* __restrict__ P)
{
    __shared__ double red[3][3][64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int k = blockIdx.x * 64 + lane;
    const int j0 = blockIdx.y * js, j1 = min(N, j0 + js);
    double acc[3] = {0.0, 0.0, 0.0};
    if (k < N) {
        const double4 pk = pos[k];
        const double2 ljk = LJ[k];
        const double yk[3] = {y[3 * k], y[3 * k + 1], y[3 * k + 2]};
        const int nb0 = fp.use_conn ? topo.conn_off[k] : 0;
        const int nb1 = fp.use_conn ? topo.conn_off[k + 1] : 0;
        const double lj6 = fp.lj_6_term;
        for (int j = j0 + wave; j < j1; j += 4) {
            if (j == k) continue;
            const double4 pj = pos[j];
            const double d[3] = {pk.x - pj.x, pk.y - pj.y, pk.z - pj.z};
            const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
            if (r2 > fp.far_r2) continue;
            const double r = sqrt(r2);
            const bool conn = bonded(topo, nb0, nb1, j);
            const double ed = (d[0] * (yk[0] - y[3 * j]) + d[1] * (yk[1] - y[3 * j + 1]) +
                               d[2] * (yk[2] - y[3 * j + 2])) / r;
            if (fp.alpha_lj != 0.0) {
                const double2 ljj = LJ[j];
                const double sig = ljk.x + ljj.x;
                const double cut = fp.cutoff_lj > 0.0 ? fp.cutoff_lj : (double)1.122462f * sig;
                const double scale = conn ? fp.bond_lj_scale : 1.0;
                if (r <= cut && scale != 0.0) {
                    const double eps = sqrt(ljk.y * ljj.y);
                    double s6 = sig / r;
                    s6 *= s6;
                    s6 *= s6 * s6;
                    const double a = scale * fp.alpha_lj * eps;
                    acc[0] += ed * (a * (-12.0 * s6 * s6 + 6.0 * lj6 * s6) / r);
                    acc[1] += ed * (a * (-144.0 * s6 * s6 + 36.0 * lj6 * s6) / (sig * r));
                }
            }
            if (fp.elec) {
                const double scale = conn ? fp.bond_c_scale : 1.0;
                if (r <= fp.cutoff_c && scale != 0.0) {
                    const double ex = exp(-fp.kappa_c * r);
                    acc[2] += ed * (-scale * fp.gam * Q[j] * ex * (1.0 / r2 + fp.kappa_c / r));
                }
            }
        }
    }
    if (wave > 0) {
#pragma unroll
        for (int u = 0; u < 3; ++u) red[wave - 1][u][lane] = acc[u];
    }
    __syncthreads();
    if (wave == 0 && k < N) {
        double* out = P + (long long)blockIdx.y * 3 * N;
#pragma unroll
        for (int u = 0; u < 3; ++u)
            out[u * N + k] = ((acc[u] + red[0][u][lane]) + red[1][u][lane]) + red[2][u][lane];
    }
}

// Per bead: the split partials summed in split order, scaled by -c (and 1 / (2 eps_k) for
// epsilon); per bond, at its lower-index end, -c (y_k - y_p).e (r - r0) / den into
// H_bar[bond_id] (each bond written once).
__global__ void __launch_bounds__(256)
adj_param_finish_kernel(const double4* __restrict__ pos, const double2* __restrict__ LJ, int N,
                        Topology topo, const int* __restrict__ bond_id, const double* __restrict__ y,
                        const double* __restrict__ P, int S, double c, double* __restrict__ eps_bar,
                        double* __restrict__ R_bar, double* __restrict__ Q_bar, double* __restrict__ H_bar)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= N) return;
    double acc[3] = {0.0, 0.0, 0.0};
    for (int sp = 0; sp < S; ++sp) {
#pragma unroll
        for (int u = 0; u < 3; ++u) acc[u] += P[(long long)sp * 3 * N + u * N + k];
    }
    eps_bar[k] = -c * acc[0] / (2.0 * LJ[k].y);
    R_bar[k] = -c * acc[1];
    Q_bar[k] = -c * acc[2];
    const double4 pk = pos[k];
    for (int b = topo.bond_off[k]; b < topo.bond_off[k + 1]; ++b) {
        const int p = topo.bond_p[b];
        if (p < k) continue;
        const double r0 = topo.bond_parm[3 * b], rmax = topo.bond_parm[3 * b + 1];
        const double4 pp = pos[p];
        const double d[3] = {pk.x - pp.x, pk.y - pp.y, pk.z - pp.z};
        const double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        const double den = 1.0 - r * r / (rmax * rmax);
        const double ed = (d[0] * (y[3 * k] - y[3 * p]) + d[1] * (y[3 * k + 1] - y[3 * p + 1]) +
                           d[2] * (y[3 * k + 2] - y[3 * p + 2])) / r;
        H_bar[bond_id[b]] = -c * ed * (r - r0) / den;
    }
}

int launch_adj_params(const double4* d_pos, const double* d_Q, const double2* d_LJ, int N,
                      const Topology& topo, const int* d_bond_id, const ForceParams& fp,
                      const double* d_y, double c, double* d_part, double* d_eps_bar,
                      double* d_R_bar, double* d_Q_bar, double* d_H_bar, hipStream_t st)
{
    const int S = adj_contract_splits(N);
    const int js = ceil_div(N, S);
    hipLaunchKernelGGL(adj_param_pairs_kernel, dim3(ceil_div(N, 64), S), dim3(256), 0, st, d_pos, d_Q, d_LJ,
                       N, topo, fp, d_y, js, d_part);
    BDH_LAUNCH_CHECK();
    hipLaunchKernelGGL(adj_param_finish_kernel, dim3(ceil_div(N, 256)), dim3(256), 0, st, d_pos, d_LJ, N, topo,
                       d_bond_id, d_y, (const double*)d_part, S, c, d_eps_bar, d_R_bar, d_Q_bar, d_H_bar);
    BDH_LAUNCH_CHECK();
    return 0;
}

}  // namespace bdh
