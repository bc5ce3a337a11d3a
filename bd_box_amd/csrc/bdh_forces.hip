// Bonded and nonbonded forces, and the Ermak-McCammon position update, on gfx950.
//
// Forces restate the reference's compute_forces with brute-force pair lists
// (src/potentials/calc_func.c:106-228, force_single :370-409); the pair and bond terms are
// in bdh_pairterm.hpp.  Bonded partners use bond_lj_scale / bond_c_scale (0 removes the
// term) whenever a scale differs from 1, exactly when the reference allocates its partner
// table (bond.c:113-130); otherwise bonded pairs interact like any other pair.
// Up to FT_MAX_N beads (forces_tiles_body): tiles of 64 beads, one workgroup per tile pair
// (I <= J) of the upper triangle, each unordered pair evaluated once, as the reference's
// loop over j > i does (calc_func.c:188-228): the pair's force goes to the row bead i and its
// negative to the column bead j.  Row sums of tile I land in split J, column sums of tile J
// in split I, so every split of every bead is written exactly once (no clearing, no
// atomics, deterministic sums).  Beyond FT_MAX_N (forces_body): 64 row beads and one split
// of the column beads per workgroup, full row sums with every pair evaluated from both
// sides.  Pair arithmetic runs in T (float in the mixed-precision step), separations and
// accumulation in FP64.
#include "bdh_common.hpp"
#include "bdh_kernels.hpp"
#include "bdh_pairterm.hpp"

namespace bdh {

constexpr int FR_ROWS = 64;
constexpr int FR_CHUNK = 256;

// Workgroup (bx, by): row beads 64 bx .., column split by; PE (when given) receives the
// workgroup's energies at 3 pe_idx.
template <typename T>
__device__ __forceinline__ void forces_body(const double4* __restrict__ pos, const double* __restrict__ Q,
                                            const double2* __restrict__ LJ, int N, const Topology& topo,
                                            const ForceParams& fp, int cols_per_split,
                                            double* __restrict__ PF, double* __restrict__ PE,
                                            int* __restrict__ flags, int bx, int by, int pe_idx)
{
    __shared__ double4 s_pos[FR_CHUNK];
    __shared__ double s_q[FR_CHUNK];
    __shared__ double2 s_lj[FR_CHUNK];
    __shared__ double s_red[6][4][FR_ROWS];
    const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * FR_ROWS + tx;
    const int i = bx * FR_ROWS + tx;
    const int jb = by * cols_per_split;
    const int je = min(N, jb + cols_per_split);
    const double4 pi = (i < N) ? pos[i] : make_double4(0, 0, 0, 1);
    const double qi = (i < N) ? Q[i] : 0.0;
    const double2 lji = (i < N) ? LJ[i] : make_double2(0, 0);
    const int nb0 = (i < N && fp.use_conn) ? topo.conn_off[i] : 0;
    const int nb1 = (i < N && fp.use_conn) ? topo.conn_off[i + 1] : 0;
    PairAcc a{0, 0, 0, 0, 0, 0};
    for (int c0 = jb; c0 < je; c0 += FR_CHUNK) {
        const int cn = min(FR_CHUNK, je - c0);
        __syncthreads();
        for (int q = tid; q < cn; q += 256) {
            s_pos[q] = pos[c0 + q];
            s_q[q] = Q[c0 + q];
            s_lj[q] = LJ[c0 + q];
        }
        __syncthreads();
        if (i >= N) continue;
        for (int q = ty; q < cn; q += 4) {
            const int j = c0 + q;
            if (j == i) continue;
            double dx = s_pos[q].x - pi.x, dy = s_pos[q].y - pi.y, dz = s_pos[q].z - pi.z;
            const double r2 = pair_r2(dx, dy, dz, fp.pbc, fp.box);
            if (r2 <= fp.far_r2)
                pair_term<T>(dx, dy, dz, r2, qi, s_q[q], lji, s_lj[q], bonded(topo, nb0, nb1, j), fp, a);
        }
    }
    // bonds (only in the first column split, so each bond end is added once)
    double e_b = 0.0;
    if (i < N && by == 0 && ty == 0) e_b = bond_terms(pos, i, pi, topo, fp, a);
    s_red[0][ty][tx] = a.fx;
    s_red[1][ty][tx] = a.fy;
    s_red[2][ty][tx] = a.fz;
    s_red[3][ty][tx] = a.e_lj;
    s_red[4][ty][tx] = a.e_c;
    s_red[5][ty][tx] = e_b;
    __syncthreads();
    if (ty == 0 && i < N) {
        double* out = PF + (long long)by * 3 * N + 3 * i;
        out[0] = s_red[0][0][tx] + s_red[0][1][tx] + s_red[0][2][tx] + s_red[0][3][tx];
        out[1] = s_red[1][0][tx] + s_red[1][1][tx] + s_red[1][2][tx] + s_red[1][3][tx];
        out[2] = s_red[2][0][tx] + s_red[2][1][tx] + s_red[2][2][tx] + s_red[2][3][tx];
    }
    // energies: pair terms are seen from both ends (x 1/2), bonds likewise
    __syncthreads();
    if (PE && tid < 3) {
        double e = 0;
        for (int y = 0; y < 4; ++y)
            for (int x = 0; x < FR_ROWS; ++x) e += s_red[3 + tid][y][x];
        PE[3 * pe_idx + tid] = 0.5 * e;
    }
    if (a.rep && fp.check_overlap) atomicOr(flags, 1);
    if (a.rep & 2) atomicOr(flags, 2);
}

constexpr int FT_TILE = 64;
constexpr int FT_MAX_N = 16384;   // tile-pair partials: 3N x ceil(N / 64) doubles

// Tile pair b = J (J + 1) / 2 + I of the upper triangle of 64-bead tiles.  Lane t owns row
// bead I*64 + t; wave w takes the shifts s = w, w + 4, ... and pairs row t with column bead
// (t + s) mod 64 of tile J, so the column beads of one shift are distinct within a wave and
// its column sums go to the wave's own LDS rows.  Diagonal tiles (I = J) visit every ordered
// pair and keep row sums only; their workgroup also adds every row bead's bond terms.
template <typename T>
__device__ __forceinline__ void forces_tiles_body(const double4* __restrict__ pos, const double* __restrict__ Q,
                                                  const double2* __restrict__ LJ, int N, const Topology& topo,
                                                  const ForceParams& fp, double* __restrict__ PF,
                                                  double* __restrict__ PE, int* __restrict__ flags, int b,
                                                  int pe_idx)
{
    __shared__ double4 s_pos[FT_TILE];
    __shared__ double s_q[FT_TILE];
    __shared__ double2 s_lj[FT_TILE];
    __shared__ double s_col[4][3][FT_TILE];
    __shared__ double s_row[4][6][FT_TILE];
    const int t = threadIdx.x, w = threadIdx.y, tid = w * FT_TILE + t;
    int J = (int)((sqrt(8.0 * b + 1.0) - 1.0) * 0.5);
    while (J * (J + 1) / 2 > b) --J;
    while ((J + 1) * (J + 2) / 2 <= b) ++J;
    const int I = b - J * (J + 1) / 2;
    const bool diag = I == J;
    const int i = I * FT_TILE + t, j0 = J * FT_TILE;
    if (tid < FT_TILE) {
        const int j = j0 + tid;
        const bool ok = j < N;
        s_pos[tid] = ok ? pos[j] : make_double4(0, 0, 0, 1);
        s_q[tid] = ok ? Q[j] : 0.0;
        s_lj[tid] = ok ? LJ[j] : make_double2(0, 0);
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) s_col[w][k][t] = 0.0;
    const bool iv = i < N;
    const double4 pi = iv ? pos[i] : make_double4(0, 0, 0, 1);
    const double qi = iv ? Q[i] : 0.0;
    const double2 lji = iv ? LJ[i] : make_double2(0, 0);
    const int nb0 = (iv && fp.use_conn) ? topo.conn_off[i] : 0;
    const int nb1 = (iv && fp.use_conn) ? topo.conn_off[i + 1] : 0;
    __syncthreads();
    PairAcc a{0, 0, 0, 0, 0, 0};
    for (int sh = w; sh < FT_TILE; sh += 4) {
        const int jl = (t + sh) & (FT_TILE - 1);
        const int j = j0 + jl;
        if (!iv || j >= N || (diag && sh == 0)) continue;
        double dx = s_pos[jl].x - pi.x, dy = s_pos[jl].y - pi.y, dz = s_pos[jl].z - pi.z;
        const double r2 = pair_r2(dx, dy, dz, fp.pbc, fp.box);
        if (r2 > fp.far_r2) continue;
        PairAcc p{0, 0, 0, 0, 0, 0};
        pair_term<T>(dx, dy, dz, r2, qi, s_q[jl], lji, s_lj[jl], bonded(topo, nb0, nb1, j), fp, p);
        a.fx += p.fx;
        a.fy += p.fy;
        a.fz += p.fz;
        a.e_lj += p.e_lj;
        a.e_c += p.e_c;
        a.rep |= p.rep;
        if (!diag) {
            s_col[w][0][jl] -= p.fx;
            s_col[w][1][jl] -= p.fy;
            s_col[w][2][jl] -= p.fz;
        }
    }
    double e_b = 0.0;
    if (iv && diag && w == 0) e_b = bond_terms(pos, i, pi, topo, fp, a);
    s_row[w][0][t] = a.fx;
    s_row[w][1][t] = a.fy;
    s_row[w][2][t] = a.fz;
    s_row[w][3][t] = a.e_lj;
    s_row[w][4][t] = a.e_c;
    s_row[w][5][t] = e_b;
    __syncthreads();
#define FT_SUM4(arr, v) ((((arr)[0][v][t] + (arr)[1][v][t]) + (arr)[2][v][t]) + (arr)[3][v][t])
    if (w == 0 && iv) {
        double* out = PF + (long long)J * 3 * N + 3 * i;
#pragma unroll
        for (int k = 0; k < 3; ++k) out[k] = FT_SUM4(s_row, k);
    }
    if (w == 1 && !diag && j0 + t < N) {
        double* out = PF + (long long)I * 3 * N + 3 * (j0 + t);
#pragma unroll
        for (int k = 0; k < 3; ++k) out[k] = FT_SUM4(s_col, k);
    }
    if (PE && w == 2) {
        // pairs of a diagonal tile are seen from both ends; bonds from both of theirs
        double e[3] = {FT_SUM4(s_row, 3), FT_SUM4(s_row, 4), FT_SUM4(s_row, 5)};
#pragma unroll
        for (int k = 0; k < 3; ++k)
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) e[k] += __shfl_xor(e[k], off);
        if (t == 0) {
            PE[3 * pe_idx] = diag ? 0.5 * e[0] : e[0];
            PE[3 * pe_idx + 1] = diag ? 0.5 * e[1] : e[1];
            PE[3 * pe_idx + 2] = 0.5 * e[2];
        }
    }
#undef FT_SUM4
    if (a.rep && fp.check_overlap) atomicOr(flags, 1);
    if (a.rep & 2) atomicOr(flags, 2);
}

template <typename T>
__global__ void __launch_bounds__(256)
forces_tiles_kernel(const double4* __restrict__ pos, const double* __restrict__ Q,
                    const double2* __restrict__ LJ, int N, Topology topo, ForceParams fp,
                    double* __restrict__ PF, double* __restrict__ PE, int* __restrict__ flags)
{
    forces_tiles_body<T>(pos, Q, LJ, N, topo, fp, PF, PE, flags, blockIdx.x, blockIdx.x);
}

template <typename T>
__global__ void __launch_bounds__(256)
forces_kernel(const double4* __restrict__ pos, const double* __restrict__ Q,
              const double2* __restrict__ LJ, int N, Topology topo, ForceParams fp,
              int cols_per_split, double* __restrict__ PF, double* __restrict__ PE,
              int* __restrict__ flags)
{
    forces_body<T>(pos, Q, LJ, N, topo, fp, cols_per_split, PF, PE, flags, blockIdx.x, blockIdx.y,
                   blockIdx.y * gridDim.x + blockIdx.x);
}

// Every chain of a batch: workgroup x of chain c is (x % gx, x / gx) of the chain's own grid
// (gx row blocks, fsplits column splits).  The bond tables of chain c start at entry
// bead0 + c of the concatenated CSR offsets (N + 1 per chain) and name chain-local beads.
template <typename T>
__global__ void __launch_bounds__(256)
forces_batch_kernel(const double4* __restrict__ pos, const double* __restrict__ Q,
                    const double2* __restrict__ LJ, Topology topo, ForceParams fp,
                    const BatchChain* __restrict__ ch, const int* __restrict__ prefix, int nch,
                    double* __restrict__ Pf, int* __restrict__ flags)
{
    const int c = batch_find(prefix, nch, blockIdx.x);
    const BatchChain b = ch[c];
    const int x = blockIdx.x - prefix[c];
    Topology t = topo;
    t.conn_off += b.bead0 + c;
    t.bond_off += b.bead0 + c;
    if (b.N <= FT_MAX_N) {
        forces_tiles_body<T>(pos + b.bead0, Q + b.bead0, LJ + b.bead0, b.N, t, fp, Pf + b.pf, nullptr, flags,
                             x, 0);
    } else {
        const int gx = (b.N + FR_ROWS - 1) / FR_ROWS;
        forces_body<T>(pos + b.bead0, Q + b.bead0, LJ + b.bead0, b.N, t, fp, (b.N + b.fsplits - 1) / b.fsplits,
                       Pf + b.pf, nullptr, flags, x % gx, x / gx, 0);
    }
}

template <typename T>
int launch_forces_batch(const double4* d_pos, const double* d_Q, const double2* d_LJ,
                        const Topology& topo, const ForceParams& fp, const BatchPlan& plan,
                        double* d_Pf, int* d_flags, hipStream_t st)
{
    hipLaunchKernelGGL(forces_batch_kernel<T>, dim3(plan.force_blocks), dim3(FR_ROWS, 4), 0, st, d_pos,
                       d_Q, d_LJ, topo, fp, plan.ch, plan.force_pre, plan.nch, d_Pf, d_flags);
    BDH_LAUNCH_CHECK();
    return 0;
}

int force_splits(int N)
{
    if (N <= FT_MAX_N) return ceil_div(N, FT_TILE);
    int rows = ceil_div(N, FR_ROWS);
    int s = ceil_div(1024, rows);
    int max_s = ceil_div(N, 64);
    if (s > max_s) s = max_s;
    return s < 1 ? 1 : s;
}

int force_blocks(int N)
{
    if (N <= FT_MAX_N) {
        const int nt = ceil_div(N, FT_TILE);
        return nt * (nt + 1) / 2;
    }
    return ceil_div(N, FR_ROWS) * force_splits(N);
}

template <typename T>
int launch_forces(const double4* d_pos, const double* d_Q, const double2* d_LJ, int N,
                  const Topology& topo, const ForceParams& fp, double* d_PF, double* d_PE,
                  int* d_flags, int splits, hipStream_t st)
{
    if (N <= FT_MAX_N) {
        BDH_CHECK(splits == ceil_div(N, FT_TILE), "forces: tile-pair partials expect one split per tile");
        hipLaunchKernelGGL(forces_tiles_kernel<T>, dim3(force_blocks(N)), dim3(FT_TILE, 4), 0, st, d_pos, d_Q,
                           d_LJ, N, topo, fp, d_PF, d_PE, d_flags);
        BDH_LAUNCH_CHECK();
        return 0;
    }
    dim3 grid(ceil_div(N, FR_ROWS), splits);
    hipLaunchKernelGGL(forces_kernel<T>, grid, dim3(FR_ROWS, 4), 0, st, d_pos, d_Q, d_LJ, N,
                       topo, fp, ceil_div(N, splits), d_PF, d_PE, d_flags);
    BDH_LAUNCH_CHECK();
    return 0;
}

// E = (E_LJ, E_coulomb, E_bond): sum of the per-workgroup partials (one block, tree sum)
__global__ void __launch_bounds__(256)
energy_reduce_kernel(const double* __restrict__ PE, int pe_blocks, double* __restrict__ E)
{
    __shared__ double s[3][256];
    const int t = threadIdx.x;
    double a0 = 0, a1 = 0, a2 = 0;
    for (int k = t; k < pe_blocks; k += 256) {
        a0 += PE[3 * k];
        a1 += PE[3 * k + 1];
        a2 += PE[3 * k + 2];
    }
    s[0][t] = a0; s[1][t] = a1; s[2][t] = a2;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) { s[0][t] += s[0][t + w]; s[1][t] += s[1][t + w]; s[2][t] += s[2][t + w]; }
        __syncthreads();
    }
    if (t < 3) E[t] = s[t][0];
}

int launch_energy_reduce(const double* d_PE, int pe_blocks, double* d_E, hipStream_t st)
{
    hipLaunchKernelGGL(energy_reduce_kernel, dim3(1), dim3(256), 0, st, d_PE, pe_blocks, d_E);
    BDH_LAUNCH_CHECK();
    return 0;
}

// out_j[r] = scale_j * sum_{s <= last_j(r)} P_j[s][r] for up to three jobs (blockIdx.y),
// last_j(r) = S_j - 1, or min(S_j - 1, r / blk_cols_j) for the triangular partials of L.x.
// 64 rows per block; the 4 waves take every 4th partial and meet in LDS.
__global__ void __launch_bounds__(256)
reduce_partials_kernel(ReduceJobs jobs)
{
    __shared__ double s_red[4][64];
    const ReduceJob& J = jobs.job[blockIdx.y];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r = blockIdx.x * 64 + lane;
    double acc = 0.0;
    if (r < J.n) {
        const int rr = J.period > 0 ? r % J.period : r;
        const int last = J.blk_cols > 0 ? min(J.S - 1, rr / J.blk_cols) : J.S - 1;
        const double* p = J.P + r;
        int sidx = wave;
        for (; sidx + 12 <= last; sidx += 16) {
            const double v0 = p[(long long)sidx * J.n];
            const double v1 = p[(long long)(sidx + 4) * J.n];
            const double v2 = p[(long long)(sidx + 8) * J.n];
            const double v3 = p[(long long)(sidx + 12) * J.n];
            acc += (v0 + v1) + (v2 + v3);
        }
        for (; sidx <= last; sidx += 4) acc += p[(long long)sidx * J.n];
    }
    s_red[wave][lane] = acc;
    __syncthreads();
    if (wave == 0 && r < J.n)
        J.out[r] = J.scale * ((s_red[0][lane] + s_red[1][lane]) + (s_red[2][lane] + s_red[3][lane]));
}

// The sums of reduce_partials_kernel, in the same order, per chain of a batch: workgroup x
// of chain c covers rows 64 x' .. of the chain's 3N, right-hand side blockIdx.y.
__global__ void __launch_bounds__(256)
reduce_batch_kernel(const double* __restrict__ P, int kind, int m, double* __restrict__ out,
                    long long out_stride, double scale, const BatchChain* __restrict__ ch,
                    const int* __restrict__ prefix, int nch, int blk_cols)
{
    __shared__ double s_red[4][64];
    const int c = batch_find(prefix, nch, blockIdx.x);
    const BatchChain b = ch[c];
    const int k = blockIdx.y;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r = (blockIdx.x - prefix[c]) * 64 + lane;
    const int n = b.n;
    const int S = kind == 0 ? b.psplits : (kind == 1 ? b.fsplits : b.tcols);
    // split s of right-hand side k: row (s m + k) of length n (m = 1 but for L.x)
    const double* base = P + (kind == 0 ? b.pp : (kind == 1 ? b.pf : b.pt)) + (long long)k * n;
    const long long stride = (long long)m * n;
    double acc = 0.0;
    if (r < n) {
        const int last = kind == 2 ? min(S - 1, r / blk_cols) : S - 1;
        const double* p = base + r;
        int sidx = wave;
        for (; sidx + 12 <= last; sidx += 16) {
            const double v0 = p[sidx * stride];
            const double v1 = p[(sidx + 4) * stride];
            const double v2 = p[(sidx + 8) * stride];
            const double v3 = p[(sidx + 12) * stride];
            acc += (v0 + v1) + (v2 + v3);
        }
        for (; sidx <= last; sidx += 4) acc += p[sidx * stride];
    }
    s_red[wave][lane] = acc;
    __syncthreads();
    if (wave == 0 && r < n)
        out[k * out_stride + 3LL * b.bead0 + r] =
            scale * ((s_red[0][lane] + s_red[1][lane]) + (s_red[2][lane] + s_red[3][lane]));
}

int launch_reduce_batch(const double* d_P, int kind, int m, double* d_out, long long out_stride,
                        double scale, const BatchPlan& plan, hipStream_t st)
{
    BDH_CHECK(kind >= 0 && kind <= 2 && m >= 1 && (kind == 2 || m == 1), "launch_reduce_batch: kind");
    hipLaunchKernelGGL(reduce_batch_kernel, dim3(plan.rows_blocks, m), dim3(256), 0, st, d_P, kind, m, d_out,
                       out_stride, scale, plan.ch, plan.rows_pre, plan.nch, trmv_blk_cols());
    BDH_LAUNCH_CHECK();
    return 0;
}

int launch_reduce_partials(const ReduceJobs& jobs, int njobs, int max_n, hipStream_t st)
{
    if (njobs <= 0) return 0;
    hipLaunchKernelGGL(reduce_partials_kernel, dim3(ceil_div(max_n, 64), njobs), dim3(256), 0, st,
                       jobs);
    BDH_LAUNCH_CHECK();
    return 0;
}

// ------------------------------------------------------------------ E-M position update
// r_i <- r_i + Rf_i + Rs_i, then the periodic wrap of trans_to_box (diff_equation.c:254-264,
// trans.c:101-107).  MODE adds a flow field's term (FlowTerm); its products are rounded one
// by one, as the reference's host code forms them.
template <int MODE>
__global__ void em_update_kernel(double4* __restrict__ pos, int N, const double* __restrict__ Rf,
                                 const double* __restrict__ Rs, int pbc, double box, FlowTerm fl)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    double4 p = pos[i];
    if constexpr (MODE == FLOW_NONE) {
        p.x = (Rf[3 * i] + Rs[3 * i]) + p.x;
        p.y = (Rf[3 * i + 1] + Rs[3 * i + 1]) + p.y;
        p.z = (Rf[3 * i + 2] + Rs[3 * i + 2]) + p.z;
    } else if constexpr (MODE == FLOW_MOVE) {
#pragma clang fp contract(off)
        const double* G = fl.G;
        const double t0 = fl.dt * (G[0] * p.x + G[1] * p.y + G[2] * p.z);
        const double t1 = fl.dt * (G[3] * p.x + G[4] * p.y + G[5] * p.z);
        const double t2 = fl.dt * (G[6] * p.x + G[7] * p.y + G[8] * p.z);
        p.x = Rf[3 * i] + Rs[3 * i] + p.x + t0;
        p.y = Rf[3 * i + 1] + Rs[3 * i + 1] + p.y + t1;
        p.z = Rf[3 * i + 2] + Rs[3 * i + 2] + p.z + t2;
        if (fl.FS) {
            fl.FS[3 * i] = t0 + fl.dt * (G[0] * p.x + G[1] * p.y + G[2] * p.z);
            fl.FS[3 * i + 1] = t1 + fl.dt * (G[3] * p.x + G[4] * p.y + G[5] * p.z);
            fl.FS[3 * i + 2] = t2 + fl.dt * (G[6] * p.x + G[7] * p.y + G[8] * p.z);
        }
    } else {
        p.x = p.x + (Rf[3 * i] + Rs[3 * i] + 0.5 * fl.FS[3 * i]);
        p.y = p.y + (Rf[3 * i + 1] + Rs[3 * i + 1] + 0.5 * fl.FS[3 * i + 1]);
        p.z = p.z + (Rf[3 * i + 2] + Rs[3 * i + 2] + 0.5 * fl.FS[3 * i + 2]);
    }
    if (pbc) {
        const double inv = 1.0 / box;
        p.x -= box * floor(p.x * inv + 0.5);
        p.y -= box * floor(p.y * inv + 0.5);
        p.z -= box * floor(p.z * inv + 0.5);
    }
    pos[i] = p;
}

// reduce_partials_kernel's sum of S partial rows (same order, same scaling) into Rf, and
// em_update_kernel's move of the coordinate each entry belongs to, in one launch: entry r of
// the 3N-vector updates component r % 3 of bead r / 3 with the same expression, so positions
// and Rf are bitwise those of the two launches.
__global__ void __launch_bounds__(256)
reduce_move_kernel(const double* __restrict__ P, int S, int n, double scale, double* __restrict__ Rf,
                   double* __restrict__ pos, const double* __restrict__ Rs, int pbc, double box)
{
    __shared__ double s_red[4][64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r = blockIdx.x * 64 + lane;
    double acc = 0.0;
    if (r < n) {
        const double* p = P + r;
        int sidx = wave;
        for (; sidx + 12 <= S - 1; sidx += 16) {
            const double v0 = p[(long long)sidx * n];
            const double v1 = p[(long long)(sidx + 4) * n];
            const double v2 = p[(long long)(sidx + 8) * n];
            const double v3 = p[(long long)(sidx + 12) * n];
            acc += (v0 + v1) + (v2 + v3);
        }
        for (; sidx <= S - 1; sidx += 4) acc += p[(long long)sidx * n];
    }
    s_red[wave][lane] = acc;
    __syncthreads();
    if (wave == 0 && r < n) {
        double rf;
        {
            // rounded on its own, as the stored Rf that em_update_kernel reads (no FMA with
            // the add below)
#pragma clang fp contract(off)
            rf = scale * ((s_red[0][lane] + s_red[1][lane]) + (s_red[2][lane] + s_red[3][lane]));
        }
        Rf[r] = rf;
        double* q = pos + 4LL * (r / 3) + r % 3;
        double x = (rf + Rs[r]) + *q;
        if (pbc) {
            const double inv = 1.0 / box;
            x -= box * floor(x * inv + 0.5);
        }
        *q = x;
    }
}

int launch_reduce_move(const double* d_P, int S, int n, double scale, double* d_Rf, double4* d_pos,
                       const double* d_Rs, int pbc, double box, hipStream_t st)
{
    BDH_CHECK(S > 0 && n % 3 == 0, "reduce_move: bad shape");
    hipLaunchKernelGGL(reduce_move_kernel, dim3(ceil_div(n, 64)), dim3(256), 0, st, d_P, S, n, scale, d_Rf,
                       reinterpret_cast<double*>(d_pos), d_Rs, pbc, box);
    BDH_LAUNCH_CHECK();
    return 0;
}

int launch_em_update(double4* d_pos, int N, const double* d_Rf, const double* d_Rs, int pbc,
                     double box, hipStream_t st, const FlowTerm* flow)
{
    const FlowTerm fl = flow ? *flow : FlowTerm{};
    const dim3 grid(ceil_div(N, 128)), block(128);
    if (fl.mode == FLOW_MOVE) {
        hipLaunchKernelGGL(em_update_kernel<FLOW_MOVE>, grid, block, 0, st, d_pos, N, d_Rf, d_Rs, pbc, box, fl);
    } else if (fl.mode == FLOW_CORRECTOR) {
        BDH_CHECK(fl.FS != nullptr, "em_update: the corrector's flow term needs FS");
        hipLaunchKernelGGL(em_update_kernel<FLOW_CORRECTOR>, grid, block, 0, st, d_pos, N, d_Rf, d_Rs, pbc, box,
                           fl);
    } else {
        hipLaunchKernelGGL(em_update_kernel<FLOW_NONE>, grid, block, 0, st, d_pos, N, d_Rf, d_Rs, pbc, box, fl);
    }
    BDH_LAUNCH_CHECK();
    return 0;
}

template int launch_forces_batch<float>(const double4*, const double*, const double2*, const Topology&, const ForceParams&, const BatchPlan&, double*, int*, hipStream_t);
template int launch_forces_batch<double>(const double4*, const double*, const double2*, const Topology&, const ForceParams&, const BatchPlan&, double*, int*, hipStream_t);
template int launch_forces<float>(const double4*, const double*, const double2*, int, const Topology&, const ForceParams&, double*, double*, int*, int, hipStream_t);
template int launch_forces<double>(const double4*, const double*, const double2*, int, const Topology&, const ForceParams&, double*, double*, int*, int, hipStream_t);

}  // namespace bdh
