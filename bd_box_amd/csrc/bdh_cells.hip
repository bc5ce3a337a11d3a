// Cell lists built on the GPU and the nonbonded forces evaluated over them.
//
// The reference builds its neighbour structures on the host at every force call
// (calc_func.c:114-115): hashed spatial buckets (nb_list bucket_space, bucket.c:471-509)
// or sorted corner keys (bucket_list, bucket.c:173-379), both of cell edge
// max_force_cutoff (pre_func.c:197-231: the largest potential cutoff, or twice the
// largest bead radius, plus 0.01).  Pairs found there go through force_single, whose
// per-potential cutoffs select the interacting pairs -- the same set the brute-force
// loop selects (calc_func.c:188-228).  Here, all on the device and rebuilt every call:
//   cell_setup    one block: bounding box (bc none) or the periodic cube; cubic cells of
//                 edge >= the cell cutoff, at most CELL_MAX_DIM per axis
//   cell_count    cell of every bead and per-cell counts (integer atomics)
//   cell_scan     exclusive scan of the counts (one block)
//   cell_scatter  bead indices into their cell's slots
//   cell_order    each cell's beads sorted by index, so the lists -- and the order of every
//                 force sum -- do not depend on the order the atomics ran in
// Up to 65536 beads the five steps run in one workgroup with the counts in LDS (one launch,
// at most 24^3 cells); larger systems take the multi-kernel path above.
// A bead's neighbours are the beads of the 27 surrounding cells (each distinct cell once
// when a periodic axis has fewer than 3 cells).  The pair and bond terms are the
// brute-force kernel's (bdh_pairterm.hpp); with the cell edge at least the largest
// cutoff, every pair the brute-force kernel finds interacting is visited, so the set of
// interacting pairs is identical.
#include "bdh_common.hpp"
#include "bdh_kernels.hpp"
#include "bdh_pairterm.hpp"

namespace bdh {

struct CellGrid {
    double ox, oy, oz;  // origin
    double h, inv_h;    // cell edge
    int nx, ny, nz;
    int ncell;
};

// Cubic cells of edge >= cutoff, at most maxdim per axis: tiling the periodic cube, or
// covering the bounding box lo..hi.
__device__ CellGrid make_grid(double lx, double ly, double lz, double hx, double hy, double hz,
                              double cutoff, int pbc, double box, int maxdim)
{
    CellGrid G;
    if (pbc) {
        int n = (int)floor(box / cutoff);
        if (n < 1) n = 1;
        if (n > maxdim) n = maxdim;
        G.nx = G.ny = G.nz = n;
        G.h = box / n;
        G.ox = G.oy = G.oz = -0.5 * box;
    } else {
        // extents of the finite coordinates (none finite, e.g. after a failed step: one cell)
        double ex = hx - lx, ey = hy - ly, ez = hz - lz;
        const bool ok = ex >= 0.0 && ey >= 0.0 && ez >= 0.0 && ex < 1e30 && ey < 1e30 && ez < 1e30;
        if (!ok) { ex = ey = ez = 0.0; lx = ly = lz = 0.0; }
        const double h = fmax(cutoff, fmax(fmax(ex, ey), ez) / (maxdim - 1));
        G.h = h;
        G.ox = lx; G.oy = ly; G.oz = lz;
        G.nx = max(1, min(maxdim, (int)(ex / h) + 1));
        G.ny = max(1, min(maxdim, (int)(ey / h) + 1));
        G.nz = max(1, min(maxdim, (int)(ez / h) + 1));
    }
    G.inv_h = 1.0 / G.h;
    G.ncell = G.nx * G.ny * G.nz;
    return G;
}

__global__ void __launch_bounds__(256)
cell_setup_kernel(const double4* __restrict__ pos, int N, double cutoff, int pbc, double box,
                  CellGrid* __restrict__ g)
{
    __shared__ double smin[3][256], smax[3][256];
    const int t = threadIdx.x;
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    if (!pbc) {
        for (int i = t; i < N; i += 256) {
            const double4 p = pos[i];
            lo[0] = fmin(lo[0], p.x); hi[0] = fmax(hi[0], p.x);
            lo[1] = fmin(lo[1], p.y); hi[1] = fmax(hi[1], p.y);
            lo[2] = fmin(lo[2], p.z); hi[2] = fmax(hi[2], p.z);
        }
    }
    for (int k = 0; k < 3; ++k) { smin[k][t] = lo[k]; smax[k][t] = hi[k]; }
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w)
            for (int k = 0; k < 3; ++k) {
                smin[k][t] = fmin(smin[k][t], smin[k][t + w]);
                smax[k][t] = fmax(smax[k][t], smax[k][t + w]);
            }
        __syncthreads();
    }
    if (t == 0) *g = make_grid(smin[0][0], smin[1][0], smin[2][0], smax[0][0], smax[1][0], smax[2][0],
                               cutoff, pbc, box, CELL_MAX_DIM);
}

// Cell index along one axis, always in [0, n): non-finite coordinates go to cell 0.
__device__ __forceinline__ int cell_coord(double x, double o, double inv_h, int n, int pbc)
{
    double f = floor((x - o) * inv_h);
    if (pbc) f -= (double)n * floor(f / (double)n);
    if (!(f >= 0.0)) return 0;
    return f >= (double)n ? n - 1 : (int)f;
}

__global__ void cell_zero_kernel(const CellGrid* __restrict__ g, int* __restrict__ count,
                                 int* __restrict__ fill)
{
    const int ncell = g->ncell;
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < ncell; c += gridDim.x * blockDim.x) {
        count[c] = 0;
        fill[c] = 0;
    }
}

__global__ void cell_count_kernel(const double4* __restrict__ pos, int N, int pbc,
                                  const CellGrid* __restrict__ g, int* __restrict__ key,
                                  int* __restrict__ count)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const CellGrid G = *g;
    const double4 p = pos[i];
    const int cx = cell_coord(p.x, G.ox, G.inv_h, G.nx, pbc);
    const int cy = cell_coord(p.y, G.oy, G.inv_h, G.ny, pbc);
    const int cz = cell_coord(p.z, G.oz, G.inv_h, G.nz, pbc);
    const int c = (cz * G.ny + cy) * G.nx + cx;
    key[i] = c;
    atomicAdd(&count[c], 1);
}

// exclusive scan of count[0..ncell) into start[0..ncell]; one block of 1024 threads
__global__ void __launch_bounds__(1024)
cell_scan_kernel(const int* __restrict__ count, const CellGrid* __restrict__ g,
                 int* __restrict__ start)
{
    __shared__ int s[1024];
    const int ncell = g->ncell;
    const int t = threadIdx.x;
    const int per = (ncell + 1023) / 1024;
    const int b = min(ncell, t * per), e = min(ncell, b + per);
    int sum = 0;
    for (int c = b; c < e; ++c) sum += count[c];
    s[t] = sum;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const int v = (t >= off) ? s[t - off] : 0;
        __syncthreads();
        s[t] += v;
        __syncthreads();
    }
    int run = s[t] - sum;
    for (int c = b; c < e; ++c) { start[c] = run; run += count[c]; }
    if (t == 1023) start[ncell] = s[1023];
}

__global__ void cell_scatter_kernel(int N, const int* __restrict__ key, const int* __restrict__ start,
                                    int* __restrict__ fill, int* __restrict__ sorted)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const int c = key[i];
    sorted[start[c] + atomicAdd(&fill[c], 1)] = i;
}

__global__ void cell_order_kernel(const CellGrid* __restrict__ g, const int* __restrict__ start,
                                  int* __restrict__ sorted)
{
    const int ncell = g->ncell;
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < ncell; c += gridDim.x * blockDim.x) {
        const int b = start[c], e = start[c + 1];
        for (int k = b + 1; k < e; ++k) {
            const int v = sorted[k];
            int m = k - 1;
            while (m >= b && sorted[m] > v) { sorted[m + 1] = sorted[m]; --m; }
            sorted[m + 1] = v;
        }
    }
}

// The whole build in one workgroup with the cell counts in LDS (grids of at most
// CELL_ONE_DIM^3 cells): one launch instead of six for the usual system sizes.
constexpr int CELL_ONE_DIM = 24;
constexpr int CELL_ONE_MAX = CELL_ONE_DIM * CELL_ONE_DIM * CELL_ONE_DIM;
constexpr int CELL_ONE_THREADS = 1024;
constexpr int CELL_ONE_MAX_N = 1 << 16;

__global__ void __launch_bounds__(CELL_ONE_THREADS)
cell_build_one_kernel(const double4* __restrict__ pos, int N, double cutoff, int pbc, double box,
                      CellGrid* __restrict__ g, int* __restrict__ key, int* __restrict__ start,
                      int* __restrict__ sorted)
{
    __shared__ int s_cnt[CELL_ONE_MAX];
    __shared__ int s_part[CELL_ONE_THREADS];
    __shared__ double s_red[6][CELL_ONE_THREADS / 64];
    __shared__ CellGrid s_g;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    if (!pbc) {
        for (int i = t; i < N; i += CELL_ONE_THREADS) {
            const double4 p = pos[i];
            lo[0] = fmin(lo[0], p.x); hi[0] = fmax(hi[0], p.x);
            lo[1] = fmin(lo[1], p.y); hi[1] = fmax(hi[1], p.y);
            lo[2] = fmin(lo[2], p.z); hi[2] = fmax(hi[2], p.z);
        }
#pragma unroll
        for (int w = 32; w > 0; w >>= 1)
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                lo[k] = fmin(lo[k], __shfl_xor(lo[k], w));
                hi[k] = fmax(hi[k], __shfl_xor(hi[k], w));
            }
        if (lane == 0)
            for (int k = 0; k < 3; ++k) { s_red[k][wave] = lo[k]; s_red[3 + k][wave] = hi[k]; }
    }
    __syncthreads();
    if (t == 0) {
        if (!pbc)
            for (int w = 1; w < CELL_ONE_THREADS / 64; ++w)
                for (int k = 0; k < 3; ++k) {
                    lo[k] = fmin(s_red[k][0], s_red[k][w]);
                    hi[k] = fmax(s_red[3 + k][0], s_red[3 + k][w]);
                    s_red[k][0] = lo[k];
                    s_red[3 + k][0] = hi[k];
                }
        s_g = make_grid(s_red[0][0], s_red[1][0], s_red[2][0], s_red[3][0], s_red[4][0], s_red[5][0],
                        cutoff, pbc, box, CELL_ONE_DIM);
        *g = s_g;
    }
    __syncthreads();
    const CellGrid G = s_g;
    for (int c = t; c < G.ncell; c += CELL_ONE_THREADS) s_cnt[c] = 0;
    __syncthreads();
    for (int i = t; i < N; i += CELL_ONE_THREADS) {
        const double4 p = pos[i];
        const int c = (cell_coord(p.z, G.oz, G.inv_h, G.nz, pbc) * G.ny +
                       cell_coord(p.y, G.oy, G.inv_h, G.ny, pbc)) * G.nx +
                      cell_coord(p.x, G.ox, G.inv_h, G.nx, pbc);
        key[i] = c;
        atomicAdd(&s_cnt[c], 1);
    }
    __syncthreads();
    // exclusive scan: each thread sums a contiguous run of cells, then a scan of the sums
    const int per = (G.ncell + CELL_ONE_THREADS - 1) / CELL_ONE_THREADS;
    const int b = min(G.ncell, t * per), e = min(G.ncell, b + per);
    int sum = 0;
    for (int c = b; c < e; ++c) sum += s_cnt[c];
    s_part[t] = sum;
    __syncthreads();
    for (int off = 1; off < CELL_ONE_THREADS; off <<= 1) {
        const int v = (t >= off) ? s_part[t - off] : 0;
        __syncthreads();
        s_part[t] += v;
        __syncthreads();
    }
    int run = s_part[t] - sum;
    for (int c = b; c < e; ++c) {
        const int n = s_cnt[c];
        s_cnt[c] = run;
        start[c] = run;
        run += n;
    }
    if (t == 0) start[G.ncell] = N;
    __syncthreads();
    for (int i = t; i < N; i += CELL_ONE_THREADS) sorted[atomicAdd(&s_cnt[key[i]], 1)] = i;
    __syncthreads();
    // s_cnt[c] is now the end of cell c; order each cell's beads by index
    for (int c = t; c < G.ncell; c += CELL_ONE_THREADS) {
        const int cb = c ? s_cnt[c - 1] : 0, ce = s_cnt[c];
        for (int k = cb + 1; k < ce; ++k) {
            const int v = sorted[k];
            int m = k - 1;
            while (m >= cb && sorted[m] > v) { sorted[m + 1] = sorted[m]; --m; }
            sorted[m + 1] = v;
        }
    }
}

int launch_cell_build(const double4* d_pos, int N, double cutoff, int pbc, double box,
                      const CellList& cl, hipStream_t st)
{
    BDH_CHECK(cutoff > 0.0, "cell list: cutoff must be positive");
    CellGrid* g = (CellGrid*)cl.grid;
    if (N <= CELL_ONE_MAX_N) {
        hipLaunchKernelGGL(cell_build_one_kernel, dim3(1), dim3(CELL_ONE_THREADS), 0, st, d_pos, N,
                           cutoff, pbc, box, g, cl.key, cl.start, cl.sorted);
        BDH_LAUNCH_CHECK();
        return 0;
    }
    hipLaunchKernelGGL(cell_setup_kernel, dim3(1), dim3(256), 0, st, d_pos, N, cutoff, pbc, box, g);
    BDH_LAUNCH_CHECK();
    hipLaunchKernelGGL(cell_zero_kernel, dim3(256), dim3(256), 0, st, (const CellGrid*)g, cl.count,
                       cl.fill);
    BDH_LAUNCH_CHECK();
    hipLaunchKernelGGL(cell_count_kernel, dim3(ceil_div(N, 256)), dim3(256), 0, st, d_pos, N, pbc,
                       (const CellGrid*)g, cl.key, cl.count);
    BDH_LAUNCH_CHECK();
    hipLaunchKernelGGL(cell_scan_kernel, dim3(1), dim3(1024), 0, st, cl.count, (const CellGrid*)g,
                       cl.start);
    BDH_LAUNCH_CHECK();
    hipLaunchKernelGGL(cell_scatter_kernel, dim3(ceil_div(N, 256)), dim3(256), 0, st, N, cl.key,
                       cl.start, cl.fill, cl.sorted);
    BDH_LAUNCH_CHECK();
    hipLaunchKernelGGL(cell_order_kernel, dim3(256), dim3(256), 0, st, (const CellGrid*)g, cl.start,
                       cl.sorted);
    BDH_LAUNCH_CHECK();
    return 0;
}

// The distinct neighbour cells of the cell holding a bead, enumerated by candidate index
// q = (kz * my + ky) * mx + kx.
struct NbrCells {
    int cx, cy, cz;
    int mx, my, mz;  // candidates per axis: 3, or n for a periodic axis of n < 3 cells
};

__device__ __forceinline__ NbrCells nbr_cells(const CellGrid& G, double4 p, int pbc)
{
    NbrCells c;
    c.cx = cell_coord(p.x, G.ox, G.inv_h, G.nx, pbc);
    c.cy = cell_coord(p.y, G.oy, G.inv_h, G.ny, pbc);
    c.cz = cell_coord(p.z, G.oz, G.inv_h, G.nz, pbc);
    c.mx = (pbc && G.nx < 3) ? G.nx : 3;
    c.my = (pbc && G.ny < 3) ? G.ny : 3;
    c.mz = (pbc && G.nz < 3) ? G.nz : 3;
    return c;
}

// axis coordinate of candidate k, or -1 if it lies outside a non-periodic grid
__device__ __forceinline__ int nbr_axis(int c, int k, int m, int n, int pbc)
{
    if (pbc) {
        if (m < 3) return k;
        const int v = c + k - 1;
        return v < 0 ? v + n : (v >= n ? v - n : v);
    }
    const int v = c + k - 1;
    return (v < 0 || v >= n) ? -1 : v;
}

__device__ __forceinline__ int nbr_cell(const CellGrid& G, const NbrCells& c, int q, int pbc)
{
    const int kx = q % c.mx, ky = (q / c.mx) % c.my, kz = q / (c.mx * c.my);
    const int x = nbr_axis(c.cx, kx, c.mx, G.nx, pbc);
    const int y = nbr_axis(c.cy, ky, c.my, G.ny, pbc);
    const int z = nbr_axis(c.cz, kz, c.mz, G.nz, pbc);
    if (x < 0 || y < 0 || z < 0) return -1;
    return (z * G.ny + y) * G.nx + x;
}

// ------------------------------------------------------------------------ forces
// CELL_LANES lanes per row bead; lane l takes neighbour cells l, l + CELL_LANES, ... and
// the lanes' sums meet in a fixed shuffle tree (deterministic).  Full rows: no atomics.
constexpr int CELL_LANES = 8;
constexpr int CF_THREADS = 128;

template <typename T>
__global__ void __launch_bounds__(CF_THREADS)
forces_cells_kernel(const double4* __restrict__ pos, const double* __restrict__ Q,
                    const double2* __restrict__ LJ, int N, Topology topo, ForceParams fp,
                    const CellGrid* __restrict__ g, const int* __restrict__ start,
                    const int* __restrict__ sorted, double* __restrict__ F,
                    double* __restrict__ PE, int* __restrict__ flags)
{
    __shared__ double s_e[3][CF_THREADS / CELL_LANES];
    const int sub = threadIdx.x % CELL_LANES;
    const int slot = threadIdx.x / CELL_LANES;
    const int i = blockIdx.x * (CF_THREADS / CELL_LANES) + slot;
    PairAcc a{0, 0, 0, 0, 0, 0};
    double e_b = 0.0;
    if (i < N) {
        const CellGrid G = *g;
        const double4 pi = pos[i];
        const double qi = Q[i];
        const double2 lji = LJ[i];
        const int nb0 = fp.use_conn ? topo.conn_off[i] : 0;
        const int nb1 = fp.use_conn ? topo.conn_off[i + 1] : 0;
        const NbrCells nc = nbr_cells(G, pi, fp.pbc);
        const int ncand = nc.mx * nc.my * nc.mz;
        for (int q = sub; q < ncand; q += CELL_LANES) {
            const int c = nbr_cell(G, nc, q, fp.pbc);
            if (c < 0) continue;
            const int e = start[c + 1];
            for (int k = start[c]; k < e; ++k) {
                const int j = sorted[k];
                if (j == i) continue;
                const double4 pj = pos[j];
                double dx = pj.x - pi.x, dy = pj.y - pi.y, dz = pj.z - pi.z;
                const double r2 = pair_r2(dx, dy, dz, fp.pbc, fp.box);
                if (r2 <= fp.far_r2)
                    pair_term<T>(dx, dy, dz, r2, qi, Q[j], lji, LJ[j], bonded(topo, nb0, nb1, j), fp, a);
            }
        }
        if (sub == 0) e_b = bond_terms(pos, i, pi, topo, fp, a);
    }
#pragma unroll
    for (int w = CELL_LANES / 2; w > 0; w >>= 1) {
        a.fx += __shfl_xor(a.fx, w, CELL_LANES);
        a.fy += __shfl_xor(a.fy, w, CELL_LANES);
        a.fz += __shfl_xor(a.fz, w, CELL_LANES);
        a.e_lj += __shfl_xor(a.e_lj, w, CELL_LANES);
        a.e_c += __shfl_xor(a.e_c, w, CELL_LANES);
        a.rep |= __shfl_xor(a.rep, w, CELL_LANES);
    }
    if (sub == 0) {
        if (i < N) {
            F[3 * i] = a.fx;
            F[3 * i + 1] = a.fy;
            F[3 * i + 2] = a.fz;
        }
        s_e[0][slot] = a.e_lj;
        s_e[1][slot] = a.e_c;
        s_e[2][slot] = e_b;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        double s = 0;
        for (int k = 0; k < CF_THREADS / CELL_LANES; ++k) s += s_e[threadIdx.x][k];
        PE[3 * blockIdx.x + threadIdx.x] = 0.5 * s;
    }
    if (a.rep && fp.check_overlap) atomicOr(flags, 1);
    if (a.rep & 2) atomicOr(flags, 2);
}

int cell_force_blocks(int N) { return ceil_div(N, CF_THREADS / CELL_LANES); }

template <typename T>
int launch_forces_cells(const double4* d_pos, const double* d_Q, const double2* d_LJ, int N,
                        const Topology& topo, const ForceParams& fp, const CellList& cl,
                        double* d_F, double* d_PE, int* d_flags, hipStream_t st)
{
    hipLaunchKernelGGL(forces_cells_kernel<T>, dim3(cell_force_blocks(N)), dim3(CF_THREADS), 0, st,
                       d_pos, d_Q, d_LJ, N, topo, fp, (const CellGrid*)cl.grid, cl.start, cl.sorted,
                       d_F, d_PE, d_flags);
    BDH_LAUNCH_CHECK();
    return 0;
}

// ------------------------------------------------------------------- pair export
// Pairs i < j with |r_j - r_i| <= cutoff (FP64, pair_r2) found through the cell list,
// written in whatever order the atomics give (the caller sorts).  count may exceed cap:
// then only cap pairs were stored.
__global__ void cell_pairs_kernel(const double4* __restrict__ pos, int N,
                                  const CellGrid* __restrict__ g, const int* __restrict__ start,
                                  const int* __restrict__ sorted, double cutoff, int pbc,
                                  double box, int* __restrict__ pairs, long long cap,
                                  unsigned long long* __restrict__ count)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const CellGrid G = *g;
    const double4 pi = pos[i];
    const NbrCells nc = nbr_cells(G, pi, pbc);
    const int ncand = nc.mx * nc.my * nc.mz;
    for (int q = 0; q < ncand; ++q) {
        const int c = nbr_cell(G, nc, q, pbc);
        if (c < 0) continue;
        for (int k = start[c]; k < start[c + 1]; ++k) {
            const int j = sorted[k];
            if (j <= i) continue;
            const double4 pj = pos[j];
            double dx = pj.x - pi.x, dy = pj.y - pi.y, dz = pj.z - pi.z;
            const double r = sqrt(pair_r2(dx, dy, dz, pbc, box));
            if (r <= cutoff) {
                const unsigned long long s = atomicAdd(count, 1ull);
                if ((long long)s < cap) {
                    pairs[2 * s] = i;
                    pairs[2 * s + 1] = j;
                }
            }
        }
    }
}

__global__ void clear_u64_kernel(unsigned long long* p) { *p = 0ull; }

int launch_cell_pairs(const double4* d_pos, int N, const CellList& cl, double cutoff, int pbc,
                      double box, int* d_pairs, long long cap, unsigned long long* d_count,
                      hipStream_t st)
{
    hipLaunchKernelGGL(clear_u64_kernel, dim3(1), dim3(1), 0, st, d_count);
    BDH_LAUNCH_CHECK();
    hipLaunchKernelGGL(cell_pairs_kernel, dim3(ceil_div(N, 128)), dim3(128), 0, st, d_pos, N,
                       (const CellGrid*)cl.grid, cl.start, cl.sorted, cutoff, pbc, box, d_pairs, cap,
                       d_count);
    BDH_LAUNCH_CHECK();
    return 0;
}

size_t cell_grid_bytes() { return sizeof(CellGrid); }

template int launch_forces_cells<float>(const double4*, const double*, const double2*, int, const Topology&, const ForceParams&, const CellList&, double*, double*, int*, hipStream_t);
template int launch_forces_cells<double>(const double4*, const double*, const double2*, int, const Topology&, const ForceParams&, const CellList&, double*, double*, int*, hipStream_t);

}  // namespace bdh
