// Kernel launchers of the HIP backend (all asynchronous on the given stream).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace bdh {

// Measurement hook: when a recorder is passed, launches are bracketed by begin/end on the
// launch stream (HIP events in the C ABI's implementation).
enum Phase { PH_NOISE = 0, PH_ASSEMBLE, PH_PANEL, PH_SYRK, PH_MATVEC, PH_TRMV, PH_MOVE,
             PH_FORCES, PH_COUNT };
struct PhaseRecorder {
    virtual ~PhaseRecorder() = default;
    virtual void begin(int phase, hipStream_t st) = 0;
    virtual void end(int phase, hipStream_t st) = 0;
};

// ---- batches of independent systems
// Chain c of a batch owns beads bead0 .. bead0 + N - 1 of the concatenated bead arrays,
// entries 3 bead0 .. of the concatenated 3N-vectors, an n_pad x n_pad column-major matrix
// (ld = n_pad) at A + mat, FP64 factor workspace at Lbuf + lb, its Gaussians at xi + xi_off
// (rows of ldx), and partial-sum rows at pp (pair sums: psplits rows), pf (forces: fsplits
// rows) and pt (L.x: tcols x 16 rows).  Every per-chain launch shape is the one the
// single-system launchers use, so a chain's arithmetic is the same as on its own, as long as
// `left` (the update schedule of its factorization, chol_left_looking) and `outer` (its outer
// panel width) are the ones the single system picks for its n_pad.
struct BatchChain {
    int N, n, n_pad, bead0;
    int psplits, fsplits, tcols, ldx;
    int left, outer;
    long long mat, lb, xi_off, pp, pf, pt;
};

// Workgroup x of a batched launch belongs to chain c with prefix[c] <= x < prefix[c + 1]
// (chains without work have equal bounds and are never chosen).  x is uniform, so the
// search runs on scalar loads.
__device__ __forceinline__ int batch_find(const int* __restrict__ prefix, int nch, int x)
{
    int lo = 0, hi = nch;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (prefix[mid] <= x) lo = mid; else hi = mid;
    }
    return lo;
}

// Prefix tables (nch + 1 ints on the device) and workgroup totals of one batch's launches.
struct BatchPlan {
    int nch = 0;
    const BatchChain* ch = nullptr;
    const int *asm_pre = nullptr, *pair_pre = nullptr, *force_pre = nullptr, *trmv_pre = nullptr;
    const int *rows_pre = nullptr, *diag_pre = nullptr, *trsm_pre = nullptr;
    int asm_blocks = 0, pair_blocks = 0, force_blocks = 0, trmv_blocks = 0, rows_blocks = 0,
        trsm_blocks = 0;
    const int* chol_pre = nullptr;   // steps x (nch + 1): launch k of every chain's factorization
    int chol_steps = 0;
    const int* chol_blocks = nullptr;   // host: workgroups of launch k
};

// ---- random stream (bdh_rng.hip)
uint64_t lcg_seed(uint32_t seed);
int launch_gauss_streams(uint64_t* d_states, double* d_out, const long long* d_offsets,
                         const int* d_counts, int nstreams, int* d_status, hipStream_t st);
// `count` Gaussians of one stream (*d_state, updated) into d_out, the values and final state
// of launch_gauss_streams for that stream, over many workgroups; d_scratch holds at least
// gauss_wide_scratch_bytes(count) bytes.
size_t gauss_wide_scratch_bytes(long long max_count);
int launch_gauss_wide(uint64_t* d_state, double* d_out, int count, void* d_scratch, size_t scratch_bytes,
                      int* d_status, hipStream_t st);
// Uploads the LCG jump tables to the current device (once); a synchronous copy, so it must
// run before any stream capture that records a Gaussian launch.
int gauss_prepare();

// ---- RPY tensor (bdh_rpy.hip)
template <typename T>
int launch_rpy_assemble(const double4* d_pos, int N, double preDii, double preDij, T* d_A,
                        long long ld, hipStream_t st, const double4* d_pos2 = nullptr);
template <typename T>
int launch_set_identity_pad(T* d_A, long long ld, int n, int n_pad, hipStream_t st);
int rpy_blocks(int N);   // workgroups of one assembly
template <typename T>
int launch_rpy_assemble_batch(const double4* d_pos, double preDii, double preDij, T* d_A,
                              const BatchPlan& plan, hipStream_t st);

// ---- pair sums, Ewald, TEA (bdh_pairs.hip)
int ewald_upload(double box, int R, int method, double alpha_smith, double alpha_scale, int R_real,
                 int R_recip, double r_cut);
int pair_splits(int N, bool ewald);   // row partials per bead: one per 64-bead tile
int pair_eparts(int N);                // eps partials of the TEA sums: one per tile pair
// d_pos2 (TEA only): the sums of the mean tensor (D(pos2) + D(pos)) / 2 (the IG-T corrector).
template <typename T>
int launch_pair_sums(const double4* d_pos, const double* d_F, const double* d_xi, int N,
                     double preDii, double preDij, double dt, bool ewald, bool tea, int splits,
                     double* d_Pf, double* d_Px, double* d_Pc, double* d_Pe, hipStream_t st,
                     const double4* d_pos2 = nullptr);
// Free-space D.F partials of every chain (pair_sums_kernel's arithmetic, no TEA).
template <typename T>
int launch_pair_sums_batch(const double4* d_pos, const double* d_F, double preDii, double preDij,
                           const BatchPlan& plan, double* d_Pp, hipStream_t st);
template <typename T>
int launch_ewald_assemble(const double4* d_pos, int N, double preDii, double preDij, T* d_A,
                          long long ld, hipStream_t st, const double4* d_pos2 = nullptr);
// S = pair_eparts(N) eps partials in d_Pe
int launch_tea_close(const double4* d_pos, const double* d_xi, const double* d_Feff,
                     const double* d_C, const double* d_Pe, int S, int N, double preDii, double dt,
                     double* d_bt, double* d_Rs, hipStream_t st);

// ---- Cholesky and triangular products (bdh_chol.hip)
constexpr int CHOL_NB = 64;   // panel width
constexpr int PAD_TO = 128;   // n is padded to a multiple of this with an identity block
enum SyrkMode { SYRK_F32 = 0, SYRK_F64 = 1, SYRK_BF16X3 = 2 };
// d_Lbuf: chol_lbuf_doubles(n_pad) doubles of workspace (FP64 diagonal blocks).
size_t chol_lbuf_doubles(int n_pad);
template <typename T>
int chol_factor(T* d_A, long long ld, int n_pad, double* d_Lbuf, int* d_status, int syrk_mode,
                hipStream_t st, PhaseRecorder* rec = nullptr);
int trmv_col_blocks(int n);
// Y = L X for m <= TRMV_MAX_RHS right-hand sides X[k * ldx + c]; partials P[cb][k][n].
constexpr int TRMV_MAX_RHS = 16;
template <typename T>
int launch_trmv_partials(const T* d_L, long long ld, int n, const double* d_X, long long ldx, int m,
                         double* d_P, hipStream_t st);
int trmv_blk_cols();   // columns per partial of the L.x kernel
// Per-chain workgroup counts of the batched launches (host).
int chol_update_steps(int n_pad);                 // update launches of one factorization
bool chol_left_looking(int n_pad);                // the single-system update schedule
int chol_outer_default();                         // the single-system outer panel width
int chol_update_blocks(int n_pad, bool left, int outer, int k);   // workgroups of launch k (0 when none)
int chol_trsm0_blocks(int n_pad);                 // panel-0 solve workgroups
template <typename T>
int trmv_blocks(int n);                           // row x column blocks of the L.x partials
// Factorization of every chain's matrix, launch k of all chains together.
template <typename T>
int chol_factor_batch(T* d_A, double* d_Lbuf, const BatchPlan& plan, int* d_status, int syrk_mode,
                      hipStream_t st, PhaseRecorder* rec = nullptr);
// L.x partials of every chain: X of chain c at X + xi_off (rows of ldx), or, with
// x_stride > 0, at X + 3 bead0 (rows of x_stride).
template <typename T>
int launch_trmv_batch(const T* d_A, const double* d_X, long long x_stride, int m, const BatchPlan& plan,
                      double* d_Pt, hipStream_t st);

// ---- forces (bdh_forces.hip)
struct ForceParams {
    double alpha_lj, cutoff_lj;   // cutoff_lj <= 0 -> 1.122462 (R_i + R_j)
    double lj_6_term;
    int elec;
    double gam;                   // gamma_c / epsilon_c
    double kappa_c, cutoff_c;
    double bond_lj_scale, bond_c_scale;
    int use_conn;                 // bonded-partner table active (scales != 1)
    int check_overlap;
    int pbc;
    double box;
    double far_r2;                // beyond this squared separation a pair adds nothing (no
                                  // force, no overlap): pair_term is skipped
};
// Per-bead adjacency: conn_* lists bonded partners (pair exclusion / scaling), bond_* lists
// each bead's bonds (partner, r0, rmax, H) so both ends are accumulated without atomics.
struct Topology {
    const int* conn_off;
    const int* conn;
    const int* bond_off;
    const int* bond_p;
    const double* bond_parm;
};
int force_splits(int N);   // partial rows of the forces (one per 64-bead tile up to 16384 beads)
int force_blocks(int N);   // workgroups (and energy partials) of one force launch
template <typename T>
int launch_forces(const double4* d_pos, const double* d_Q, const double2* d_LJ, int N,
                  const Topology& topo, const ForceParams& fp, double* d_PF, double* d_PE,
                  int* d_flags, int splits, hipStream_t st);
int launch_energy_reduce(const double* d_PE, int pe_blocks, double* d_E, hipStream_t st);

// ---- cell lists (bdh_cells.hip)
constexpr int CELL_MAX_DIM = 48;                                   // cells per axis
constexpr int CELL_MAX = CELL_MAX_DIM * CELL_MAX_DIM * CELL_MAX_DIM;
struct CellList {
    void* grid;     // device CellGrid (cell_grid_bytes())
    int* key;       // N: cell of each bead
    int* count;     // CELL_MAX
    int* fill;      // CELL_MAX
    int* start;     // CELL_MAX + 1
    int* sorted;    // N: bead indices grouped by cell, ascending within a cell
};
size_t cell_grid_bytes();
int launch_cell_build(const double4* d_pos, int N, double cutoff, int pbc, double box,
                      const CellList& cl, hipStream_t st);
int cell_force_blocks(int N);
template <typename T>
int launch_forces_cells(const double4* d_pos, const double* d_Q, const double2* d_LJ, int N,
                        const Topology& topo, const ForceParams& fp, const CellList& cl,
                        double* d_F, double* d_PE, int* d_flags, hipStream_t st);
int launch_cell_pairs(const double4* d_pos, int N, const CellList& cl, double cutoff, int pbc,
                      double box, int* d_pairs, long long cap, unsigned long long* d_count,
                      hipStream_t st);

// Sums of partial rows written by the split kernels (D.F, L.xi, forces).
struct ReduceJob {
    const double* P;   // S partial rows of length n
    double* out;
    int n, S, blk_cols;  // blk_cols > 0: row r sums partials 0..r/blk_cols only
    double scale;
    int period;          // > 0: the triangular bound uses r % period (stacked vectors)
};
struct ReduceJobs {
    ReduceJob job[3];
};
int launch_reduce_partials(const ReduceJobs& jobs, int njobs, int max_n, hipStream_t st);
// The same sums per chain of a batch: kind 0 pair-sum partials (pp, psplits), 1 force
// partials (pf, fsplits), 2 L.x partials of m right-hand sides (pt, tcols, triangular);
// out[k * out_stride + 3 bead0 + r] = scale * sum.
int launch_reduce_batch(const double* d_P, int kind, int m, double* d_out, long long out_stride,
                        double scale, const BatchPlan& plan, hipStream_t st);
template <typename T>
int launch_forces_batch(const double4* d_pos, const double* d_Q, const double2* d_LJ,
                        const Topology& topo, const ForceParams& fp, const BatchPlan& plan,
                        double* d_Pf, int* d_flags, hipStream_t st);

// ---- integrator (bdh_forces.hip)
// Rf = scale * (sum of S partial rows of length n, reduce_partials order), then the move of
// launch_em_update with that Rf; one launch, bitwise the same as the two.
int launch_reduce_move(const double* d_P, int S, int n, double scale, double* d_Rf, double4* d_pos,
                       const double* d_Rs, int pbc, double box, hipStream_t st);
// A flow field's share of the move, vel_grad_tensor (diff_equation.c:265-285, :304-312):
//   FLOW_MOVE       r' = ((Rf + Rs) + r) + t with t = dt G r, and, when FS is set,
//                   FS = t + dt G r' (r' before the box wrap);
//   FLOW_CORRECTOR  r' = r + ((Rf + Rs) + FS / 2), the IG-T corrector.
enum { FLOW_NONE = 0, FLOW_MOVE = 1, FLOW_CORRECTOR = 2 };
struct FlowTerm {
    double G[9] = {};        // velocity gradient, row-major (row, column)
    double dt = 0.0;
    double* FS = nullptr;    // 3N: written by FLOW_MOVE (if set), read by FLOW_CORRECTOR
    int mode = FLOW_NONE;
};
int launch_em_update(double4* d_pos, int N, const double* d_Rf, const double* d_Rs, int pbc,
                     double box, hipStream_t st, const FlowTerm* flow = nullptr);

// ---- reverse-mode derivative of an FP64 step (bdh_adjoint.hip)
int launch_adj_lt_g(const double* d_L, long long ld, int n, const double* d_g, double* d_a, hipStream_t st);
// L^T X = B for n right-hand sides: mode 0 B = s Phi(a xi^T), mode 1 B = Y^T (d_B = Y)
// (np = padded order, a multiple of 64; X and B are np x np column-major)
// inverses of the 64 x 64 diagonal blocks (np / 64 x 4096 doubles)
int launch_adj_diag_inverses(const double* d_L, long long ld, int np, double* d_Minv, hipStream_t st);
// Ld[e] = A[e] for e < count (an FP32 factor promoted to FP64; count a multiple of 4)
int launch_adj_promote(const float* d_A, long long count, double* d_Ld, hipStream_t st);
int launch_adj_upper_solve(const double* d_L, long long ld, int np, int n, int mode, const double* d_a,
                           const double* d_xi, double s, const double* d_B, const double* d_Minv,
                           double* d_X, hipStream_t st);
int launch_adj_symmetrize(int np, const double* d_Y, double* d_Z, hipStream_t st);
// d_Y2 = Z = Y2 + Y2^T (ld ldy).  The column splits' partial sums go to d_part
// (adj_contract_splits(N) rows of 3N doubles) and are summed in split order into d_out.
constexpr int ADJ_CONTRACT_MAX_SPLITS = 64;
int adj_contract_splits(int N);
int launch_adj_contract(const double4* d_pos, int N, double preDii, double preDij, const double* d_Y2,
                        long long ldy, const double* d_g, const double* d_F, double c, double* d_part,
                        double* d_out, hipStream_t st);
// d_part: adj_contract_splits(N) rows of 3N doubles for the pair partials.
int launch_adj_hvp(const double4* d_pos, const double* d_Q, const double2* d_LJ, int N,
                   const Topology& topo, const ForceParams& fp, const double* d_y,
                   const double* d_contract, const double* d_g, double c, double* d_part, double* d_out,
                   hipStream_t st);
// Parameter cotangents of one step for y = D g (bdh_adjoint.hip): eps_bar, R_bar, Q_bar (N)
// and H_bar (one per bond; bond_id maps each per-bead bond entry to its bond).  d_part:
// adj_contract_splits(N) rows of 3N doubles.
// One FP64 block of D (bdh_pairs.hip): ew 0 free space, 1 Beenakker-Ewald, 2 Smith-Ewald;
// nine column-major values into d_out.
int launch_single_block(const double4* d_pos, int i, int j, int ew, double preDii, double preDij,
                        double* d_out, hipStream_t st);
int launch_adj_params(const double4* d_pos, const double* d_Q, const double2* d_LJ, int N,
                      const Topology& topo, const int* d_bond_id, const ForceParams& fp,
                      const double* d_y, double c, double* d_part, double* d_eps_bar,
                      double* d_R_bar, double* d_Q_bar, double* d_H_bar, hipStream_t st);

// ---- probes (bdh_probe.hip)
int probe_mfma(int kind, const void* A, const void* B, void* C, hipStream_t st);

}  // namespace bdh
