// C ABI of the backend (include/bdbox_hip.h): device state of one simulated system and
// the orchestration of one Ermak-McCammon step on a single HIP stream.
//
// Step (reference: comp_step_ermak / compute_hydro_ermak, src/diff_alg/diff_comp.c:152-179,
// diff_equation.c:216-264, diff_rand.c:47-85):
//   1. xi        Gaussians: device drand48 + polar Box-Muller stream, or caller-supplied
//   2. refactor  every `refactor_every` steps: assemble D (lower, T) and factor it in place
//   3. Rf        (dt/kT) D F with RPY blocks recomputed on the fly (D is not re-read)
//   4. Rs        sqrt(2 dt) L xi (triangular product over the stored factor)
//   5. move      r += Rf + Rs in FP64 (periodic wrap when bc is periodic)
//   6. forces    at the new positions, for the next step
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/bdbox_hip.h"
#include "bdh_common.hpp"
#include "bdh_ewald.hpp"
#include "bdh_kernels.hpp"

namespace bdh {

static thread_local std::string g_err;
void set_error(const std::string& msg) { g_err = msg; }
const char* last_error() { return g_err.c_str(); }

// Rf/Rs without hydrodynamic interactions (diff_equation.c:138-146, diff_rand.c:54-61):
// Rf[r] = (dt/kT) D_ii F_r ; Rs[r] = xi_r sqrt(D_ii 2 dt)
__global__ void diag_mobility_kernel(const double4* __restrict__ pos, const double* __restrict__ F,
                                     const double* __restrict__ xi, int N, double preDii, double dt,
                                     double rf_scale, double* __restrict__ Rf, double* __restrict__ Rs)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= 3 * N) return;
    const double dii = preDii / pos[r / 3].w;
    if (Rf) Rf[r] = rf_scale * dii * F[r];
    if (Rs) Rs[r] = xi[r] * sqrt(dii * 2.0 * dt);
}

// The TEA pair sums of a large periodic system over beads in a space-filling-curve order
// perm: pos_s[i] = pos[perm[i]] (pos2 alike), F_s and X_s the 3-vectors alike.  A 64-bead tile
// is then a compact region, and a wavefront evaluates few real-space images beyond those its
// pairs need.
__global__ void pair_order_gather_kernel(const int* __restrict__ perm, int N, const double4* __restrict__ pos,
                                         const double4* __restrict__ pos2, const double* __restrict__ F,
                                         const double* __restrict__ X, double4* __restrict__ pos_s,
                                         double4* __restrict__ pos2_s, double* __restrict__ F_s,
                                         double* __restrict__ X_s)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const int j = perm[i];
    pos_s[i] = pos[j];
    if (pos2) pos2_s[i] = pos2[j];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        F_s[3 * i + k] = F[3 * j + k];
        X_s[3 * i + k] = X[3 * j + k];
    }
}

// a[perm[i]] = a_s[i] and, with b, b[perm[i]] = b_s[i] (3-vectors)
__global__ void pair_order_scatter_kernel(const int* __restrict__ perm, int N, const double* __restrict__ a_s,
                                          double* __restrict__ a, const double* __restrict__ b_s,
                                          double* __restrict__ b)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const int j = perm[i];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        a[3 * j + k] = a_s[3 * i + k];
        if (b) b[3 * j + k] = b_s[3 * i + k];
    }
}

// out = (a + b) / 2, the IG-T deterministic move Rf0 (diff_equation.c:322-325)
__global__ void mean2_kernel(const double* __restrict__ a, const double* __restrict__ b, int n,
                             double* __restrict__ out)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < n) out[r] = 0.5 * (a[r] + b[r]);
}

// Copy the lower triangle of a column-major T matrix to a full symmetric FP64 matrix.
template <typename T>
__global__ void lower_to_full_kernel(const T* __restrict__ A, long long lda, int n,
                                     double* __restrict__ D, long long ldd, int symmetric)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    if (r >= n) return;
    double v;
    if (r >= c) v = (double)A[(long long)c * lda + r];
    else v = symmetric ? (double)A[(long long)r * lda + c] : 0.0;
    D[(long long)c * ldd + r] = v;
}

template <typename T>
__global__ void full_to_lower_kernel(const double* __restrict__ D, long long ldd, int n,
                                     T* __restrict__ A, long long lda, int from_upper)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    if (r >= n || r < c) return;
    const double v = from_upper ? D[(long long)r * ldd + c] : D[(long long)c * ldd + r];
    A[(long long)c * lda + r] = (T)v;
}

// HIP-event recorder for the measurement hook (one event pair per bracketed launch).
struct EventRecorder : PhaseRecorder {
    struct Rec { int phase; hipEvent_t e0, e1; };
    std::vector<Rec> recs;
    hipEvent_t open_ev[PH_COUNT] = {};
    void begin(int phase, hipStream_t st) override
    {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return;
        (void)hipEventRecord(e, st);
        open_ev[phase] = e;
    }
    void end(int phase, hipStream_t st) override
    {
        hipEvent_t e;
        if (!open_ev[phase] || hipEventCreate(&e) != hipSuccess) return;
        (void)hipEventRecord(e, st);
        recs.push_back({phase, open_ev[phase], e});
        open_ev[phase] = nullptr;
    }
    void clear()
    {
        for (auto& r : recs) { (void)hipEventDestroy(r.e0); (void)hipEventDestroy(r.e1); }
        recs.clear();
    }
    ~EventRecorder() override { clear(); }
};

}  // namespace bdh

using namespace bdh;

struct bdh_system {
    bdh_config cfg{};
    int N = 0, n = 0, n_pad = 0;
    long long ld = 0;
    bool mixed = false;
    int syrk_mode = SYRK_F64;
    int device = 0;
    hipStream_t st = nullptr;
    double preDii = 0, preDij = 0;
    ForceParams fp{};
    Topology topo{};
    // device buffers
    double4* pos = nullptr;
    double* Q = nullptr;
    double2* LJ = nullptr;
    int *conn_off = nullptr, *conn = nullptr, *bond_off = nullptr, *bond_p = nullptr;
    double* bond_parm = nullptr;
    double *F = nullptr, *E = nullptr, *PF = nullptr, *PE = nullptr;
    int fsplits = 0, fblocks = 0;
    bool cells = false;             // nonbonded pairs over a device cell list
    double cell_cut = 0;            // cell edge lower bound (largest cutoff + 0.01)
    CellList cl{};
    void* A = nullptr;
    double* Lbuf = nullptr;         // FP64 diagonal blocks of the factorization
    double* xi = nullptr;
    size_t xi_cap = 0;              // doubles in xi
    double *Pf = nullptr, *Pr = nullptr, *Rf = nullptr, *Rs = nullptr, *noise = nullptr;
    double* RsB = nullptr;          // sqrt(2 dt) L xi for the steps of one refactor cycle
    const double* last_Rs = nullptr;  // Rs of the last step taken
    double last_dt = 0;             // dt of the last step taken (halved by ermak_var retries)
    double4* pos_prev = nullptr;    // positions and forces before a step that may be rejected
    double* F_prev = nullptr;
    void* A2 = nullptr;             // IG-T: the averaged tensor and its factor
    double *Rf0 = nullptr, *Rfs = nullptr;   // IG-T: the mean deterministic move, Rf at r*
    long long ldx = 0;              // stride of drawn noise vectors: n, or n + 1 if n is odd
    int msplits = 0, tblocks = 0;
    bool ewald = false, tea = false;
    double *Px = nullptr, *Pc = nullptr, *Pe = nullptr, *Feff = nullptr, *Cs = nullptr, *bt = nullptr;
    uint64_t* rng = nullptr;
    long long* rng_off = nullptr;
    int* rng_cnt = nullptr;
    void* gwork = nullptr;          // the many-workgroup Gaussian launches' plan and counts
    size_t gwork_bytes = 0;
    int* flags = nullptr;   // [0] force flags, [1] Cholesky status, [2] rng status
    long long steps_done = 0;
    bool have_L = false;
    EventRecorder* rec = nullptr;   // non-null while phase timing is enabled
    hipGraphExec_t factor_exec = nullptr;   // assembly + factorization, captured once
    double* adj = nullptr;          // reverse-mode step: Y1, Y2 (n x n), then vectors
    unsigned char* fcache = nullptr;  // saved factors (fcache_slots x n_pad^2, the storage type of A)
    int fcache_slots = 0;
    int nbonds = 0;
    int* bond_id = nullptr;         // bond index of each per-bead bond entry
    std::vector<int> csr_bond;      // the same on the host
    double* pgrad = nullptr;        // parameter cotangents: eps, R, Q (N each), H (nbonds)
    double ew_scale = 0.0;          // Beenakker split (bdh_set_ewald_split): alpha / (sqrt(pi) / L),
                                    // 0 for the reference's split
    int ew_Rr = -1, ew_Rk = -1;     // real / reciprocal sphere radii (-1: ewald_real)
    double ew_rcut = 0.0;           // real-space image cutoff in A (bdh_set_ewald_cutoff; 0: none)
    bool flow = false;              // a flow field (bdh_set_flow): velocity gradient G
    double G[9] = {};
    double* FS = nullptr;           // the flow terms of the last predictor (IG-T corrector)
    double4* pos_star = nullptr;    // IG-T with TEA noise: the predictor's positions r*
    int* perm = nullptr;            // periodic TEA: the pair sums' bead order (set_pair_order)
    double4 *pos_s = nullptr, *pos2_s = nullptr;
    double *F_s = nullptr, *X_s = nullptr, *Rf_s = nullptr, *Rs_s = nullptr;
    long long order_step = 0;       // steps_done when perm was last computed
};

namespace {

template <typename P>
int dmalloc(P** p, size_t bytes)
{
    if (bytes == 0) bytes = 16;
    BDH_HIP(hipMalloc((void**)p, bytes));
    return 0;
}

template <typename P>
void dfree(P*& p)
{
    if (p) (void)hipFree((void*)p);
    p = nullptr;
}

int check_flags(bdh_system* s)
{
    int h[3];
    BDH_HIP(hipMemcpyAsync(h, s->flags, sizeof(h), hipMemcpyDeviceToHost, s->st));
    BDH_HIP(hipStreamSynchronize(s->st));
    if (h[1] < 0) {
        set_error("Cholesky decomposition error: a panel did not receive its diagonal block");
        return -3;
    }
    if (h[1]) {
        set_error("Cholesky decomposition error: non-positive pivot at row " + std::to_string(h[1]));
        return -3;
    }
    if (h[0] & 2) {
        set_error("Overstretched bond");
        return -4;
    }
    if (h[2]) {
        set_error("Gaussian stream did not complete");
        return -5;
    }
    return 0;
}

#define REC_BEGIN(ph) do { if (s->rec) s->rec->begin(ph, s->st); } while (0)
#define REC_END(ph) do { if (s->rec) s->rec->end(ph, s->st); } while (0)

int alloc_cells(bdh_system* s)
{
    if (s->cl.grid) return 0;
    if (dmalloc(&s->cl.grid, cell_grid_bytes())) return -1;
    if (dmalloc(&s->cl.key, sizeof(int) * s->N)) return -1;
    if (dmalloc(&s->cl.count, sizeof(int) * CELL_MAX)) return -1;
    if (dmalloc(&s->cl.fill, sizeof(int) * CELL_MAX)) return -1;
    if (dmalloc(&s->cl.start, sizeof(int) * (CELL_MAX + 1))) return -1;
    if (dmalloc(&s->cl.sorted, sizeof(int) * s->N)) return -1;
    return 0;
}

int compute_forces_dev(bdh_system* s, bool energies = false)
{
    REC_BEGIN(PH_FORCES);
    int rc = 0;
    if (s->cells) {
        if (launch_cell_build(s->pos, s->N, s->cell_cut, s->fp.pbc, s->fp.box, s->cl, s->st))
            return -1;
        rc = s->mixed ? launch_forces_cells<float>(s->pos, s->Q, s->LJ, s->N, s->topo, s->fp, s->cl,
                                                   s->F, s->PE, s->flags, s->st)
                      : launch_forces_cells<double>(s->pos, s->Q, s->LJ, s->N, s->topo, s->fp, s->cl,
                                                    s->F, s->PE, s->flags, s->st);
        if (!rc && energies) rc = launch_energy_reduce(s->PE, s->fblocks, s->E, s->st);
        REC_END(PH_FORCES);
        return rc;
    }
    if (s->mixed) {
        if (launch_forces<float>(s->pos, s->Q, s->LJ, s->N, s->topo, s->fp, s->PF, s->PE,
                                 s->flags, s->fsplits, s->st)) return -1;
    } else {
        if (launch_forces<double>(s->pos, s->Q, s->LJ, s->N, s->topo, s->fp, s->PF, s->PE,
                                  s->flags, s->fsplits, s->st)) return -1;
    }
    ReduceJobs jobs{};
    jobs.job[0] = ReduceJob{s->PF, s->F, s->n, s->fsplits, 0, 1.0};
    rc = launch_reduce_partials(jobs, 1, s->n, s->st);
    if (!rc && energies) rc = launch_energy_reduce(s->PE, s->fblocks, s->E, s->st);
    REC_END(PH_FORCES);
    return rc;
}

// The Ewald tables live in constant memory of the module: reload them when the system
// asks for a different box or lattice radius than the last one loaded.
double g_ew_box = -1.0, g_ew_alpha = -1.0, g_ew_scale = -1.0;
int g_ew_R = -1, g_ew_method = -1, g_ew_Rr = -2, g_ew_Rk = -2;
double g_ew_rcut = -1.0;
int g_ew_dev = -1;

int ensure_ewald(bdh_system* s)
{
    if (!s->ewald) return 0;
    const bdh_config& c = s->cfg;
    if (g_ew_box == c.box && g_ew_R == c.ewald_real && g_ew_method == c.ewald_method &&
        (c.ewald_method != EW_SMITH || g_ew_alpha == c.ewald_alpha) && g_ew_dev == s->device &&
        g_ew_scale == s->ew_scale && g_ew_Rr == s->ew_Rr && g_ew_Rk == s->ew_Rk && g_ew_rcut == s->ew_rcut)
        return 0;
    if (ewald_upload(c.box, c.ewald_real, c.ewald_method, c.ewald_alpha, s->ew_scale, s->ew_Rr, s->ew_Rk,
                     s->ew_rcut) < 0)
        return -1;
    g_ew_box = c.box;
    g_ew_R = c.ewald_real;
    g_ew_method = c.ewald_method;
    g_ew_alpha = c.ewald_alpha;
    g_ew_scale = s->ew_scale;
    g_ew_Rr = s->ew_Rr;
    g_ew_Rk = s->ew_Rk;
    g_ew_rcut = s->ew_rcut;
    g_ew_dev = s->device;
    return 0;
}

#ifndef BDH_FACTOR_GRAPH
#define BDH_FACTOR_GRAPH 1
#endif

int assemble_and_factor_launches(bdh_system* s, void* Abuf, const double4* pos2);

// D at the current positions (with pos2: the mean of the tensors at both sets of positions)
// into Abuf (default: the system's matrix), factored in place.  The refactor of the system's
// own matrix (the same ~150 launches with the same arguments every time) is captured into a
// HIP graph on first use and replayed afterwards; phase timing runs the launches directly.
int assemble_and_factor(bdh_system* s, void* Abuf = nullptr, const double4* pos2 = nullptr)
{
    if (ensure_ewald(s)) return -1;
    const bool graph = BDH_FACTOR_GRAPH && !Abuf && !pos2 && !s->rec;
    if (!graph) return assemble_and_factor_launches(s, Abuf, pos2);
    if (!s->factor_exec) {
        BDH_HIP(hipStreamBeginCapture(s->st, hipStreamCaptureModeThreadLocal));
        const int rc = assemble_and_factor_launches(s, nullptr, nullptr);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(s->st, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        BDH_HIP(e);
        const hipError_t ei = hipGraphInstantiate(&s->factor_exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ei != hipSuccess) s->factor_exec = nullptr;
        BDH_HIP(ei);
    }
    BDH_HIP(hipGraphLaunch(s->factor_exec, s->st));
    s->have_L = true;
    return 0;
}

int assemble_and_factor_launches(bdh_system* s, void* Abuf, const double4* pos2)
{
    void* M = Abuf ? Abuf : s->A;
    if (s->mixed) {
        float* A = (float*)M;
        REC_BEGIN(PH_ASSEMBLE);
        if (s->ewald ? launch_ewald_assemble<float>(s->pos, s->N, s->preDii, s->preDij, A, s->ld, s->st, pos2)
                     : launch_rpy_assemble<float>(s->pos, s->N, s->preDii, s->preDij, A, s->ld, s->st, pos2))
            return -1;
        REC_END(PH_ASSEMBLE);
        if (chol_factor<float>(A, s->ld, s->n_pad, s->Lbuf, s->flags + 1, s->syrk_mode, s->st, s->rec)) return -1;
    } else {
        double* A = (double*)M;
        REC_BEGIN(PH_ASSEMBLE);
        if (s->ewald ? launch_ewald_assemble<double>(s->pos, s->N, s->preDii, s->preDij, A, s->ld, s->st, pos2)
                     : launch_rpy_assemble<double>(s->pos, s->N, s->preDii, s->preDij, A, s->ld, s->st, pos2))
            return -1;
        REC_END(PH_ASSEMBLE);
        if (chol_factor<double>(A, s->ld, s->n_pad, s->Lbuf, s->flags + 1, SYRK_F64, s->st, s->rec)) return -1;
    }
    if (M == s->A) s->have_L = true;
    return 0;
}

// Pair sums over blocks evaluated on the fly: D.F, and for TEA also D.X, C and eps (of the
// mean tensor of the positions and pos2 when pos2 is set).
int matvec_dev(bdh_system* s, const double* xi = nullptr, bool tea = false, double dt = 0.0,
               const double4* pos2 = nullptr)
{
    if (ensure_ewald(s)) return -1;
    if (dt == 0.0) dt = s->cfg.dt;
    if (s->mixed)
        return launch_pair_sums<float>(s->pos, s->F, xi, s->N, s->preDii, s->preDij, dt, s->ewald,
                                       tea, s->msplits, s->Pf, s->Px, s->Pc, s->Pe, s->st, pos2);
    return launch_pair_sums<double>(s->pos, s->F, xi, s->N, s->preDii, s->preDij, dt, s->ewald,
                                    tea, s->msplits, s->Pf, s->Px, s->Pc, s->Pe, s->st, pos2);
}

// The Geyer-Winter TEA displacement Rs (into s->Rs) from the Gaussians X, one pass over the
// pairs (diff_rand.c:86-230), and with Rf_out also Rf = (dt / kT) D F from the same pass.
// pos2: the sums of the mean tensor of the positions and pos2 (the IG-T corrector).
int tea_parts(bdh_system* s, const double* X, double dt, double* Rf_out, const double4* pos2 = nullptr)
{
    if (s->perm) {
        // the same sums over the beads in the pair order, the results scattered back
        const int N = s->N;
        const dim3 g(ceil_div(N, 256)), b(256);
        hipLaunchKernelGGL(pair_order_gather_kernel, g, b, 0, s->st, (const int*)s->perm, N, (const double4*)s->pos,
                           pos2, (const double*)s->F, X, s->pos_s, s->pos2_s, s->F_s, s->X_s);
        BDH_LAUNCH_CHECK();
        if (ensure_ewald(s)) return -1;
        const double4* p2 = pos2 ? s->pos2_s : nullptr;
        const int rc = s->mixed
            ? launch_pair_sums<float>(s->pos_s, s->F_s, s->X_s, N, s->preDii, s->preDij, dt, s->ewald, true,
                                      s->msplits, s->Pf, s->Px, s->Pc, s->Pe, s->st, p2)
            : launch_pair_sums<double>(s->pos_s, s->F_s, s->X_s, N, s->preDii, s->preDij, dt, s->ewald, true,
                                       s->msplits, s->Pf, s->Px, s->Pc, s->Pe, s->st, p2);
        if (rc) return -1;
        ReduceJobs jobs{};
        int nj = 0;
        if (Rf_out) jobs.job[nj++] = ReduceJob{s->Pf, s->Rf_s, s->n, s->msplits, 0, dt / (kB * s->cfg.T)};
        jobs.job[nj++] = ReduceJob{s->Px, s->Feff, s->n, s->msplits, 0, 1.0};
        jobs.job[nj++] = ReduceJob{s->Pc, s->Cs, s->n, s->msplits, 0, 1.0};
        if (launch_reduce_partials(jobs, nj, s->n, s->st)) return -1;
        if (launch_tea_close(s->pos_s, s->X_s, s->Feff, s->Cs, s->Pe, pair_eparts(N), N, s->preDii, dt, s->bt,
                             s->Rs_s, s->st))
            return -1;
        hipLaunchKernelGGL(pair_order_scatter_kernel, g, b, 0, s->st, (const int*)s->perm, N,
                           (const double*)s->Rs_s, s->Rs, (const double*)(Rf_out ? s->Rf_s : nullptr), Rf_out);
        BDH_LAUNCH_CHECK();
        return 0;
    }
    if (matvec_dev(s, X, true, dt, pos2)) return -1;
    ReduceJobs jobs{};
    int nj = 0;
    if (Rf_out) jobs.job[nj++] = ReduceJob{s->Pf, Rf_out, s->n, s->msplits, 0, dt / (kB * s->cfg.T)};
    jobs.job[nj++] = ReduceJob{s->Px, s->Feff, s->n, s->msplits, 0, 1.0};
    jobs.job[nj++] = ReduceJob{s->Pc, s->Cs, s->n, s->msplits, 0, 1.0};
    if (launch_reduce_partials(jobs, nj, s->n, s->st)) return -1;
    return launch_tea_close(s->pos, X, s->Feff, s->Cs, s->Pe, pair_eparts(s->N), s->N, s->preDii, dt, s->bt,
                            s->Rs, s->st);
}

int trmv_dev(bdh_system* s, const double* X, long long ldx, int m, const void* L = nullptr)
{
    const void* M = L ? L : s->A;
    if (s->mixed)
        return launch_trmv_partials<float>((const float*)M, s->ld, s->n, X, ldx, m, s->Pr, s->st);
    return launch_trmv_partials<double>((const double*)M, s->ld, s->n, X, ldx, m, s->Pr, s->st);
}

// Rs of m consecutive steps, sqrt(2 dt) L [xi_1 .. xi_m], in one pass over L (the factor
// is fixed within a refactor cycle); stacked into RsB[j n + r].
int cycle_noise(bdh_system* s, const double* X, long long ldx, int m, double dt, const void* L = nullptr)
{
    REC_BEGIN(PH_TRMV);
    if (trmv_dev(s, X, ldx, m, L)) return -1;
    ReduceJobs jobs{};
    jobs.job[0] = ReduceJob{s->Pr, s->RsB, m * s->n, s->tblocks, trmv_blk_cols(),
                            std::sqrt(2.0 * dt), s->n};
    if (launch_reduce_partials(jobs, 1, m * s->n, s->st)) return -1;
    REC_END(PH_TRMV);
    return 0;
}

// The flow-field term of a move of length dt (FLOW_NONE without a flow field).
FlowTerm flow_term(const bdh_system* s, int mode, double dt, double* FS = nullptr)
{
    FlowTerm f;
    if (!s->flow) return f;
    for (int k = 0; k < 9; ++k) f.G[k] = s->G[k];
    f.dt = dt;
    f.FS = FS;
    f.mode = mode;
    return f;
}

// One E-M step.  Cholesky: Rs_dev = this step's precomputed sqrt(2 dt) L xi.
int one_step(bdh_system* s, const double* xi_dev, const double* Rs_dev, double dt)
{
    const bdh_config& c = s->cfg;
    const double rf_scale = dt / (kB * c.T);
    const FlowTerm fl = flow_term(s, FLOW_MOVE, dt);
    if (c.hydro == BDH_HYDRO_CHOLESKY) {
        REC_BEGIN(PH_MATVEC);
        if (matvec_dev(s)) return -1;
        REC_END(PH_MATVEC);
        REC_BEGIN(PH_MOVE);
        if (s->flow) {
            // the flow term needs all three components of a bead: Rf first, then the move
            ReduceJobs jobs{};
            jobs.job[0] = ReduceJob{s->Pf, s->Rf, s->n, s->msplits, 0, rf_scale};
            if (launch_reduce_partials(jobs, 1, s->n, s->st)) return -1;
            if (launch_em_update(s->pos, s->N, s->Rf, Rs_dev, c.bc == 1, c.box, s->st, &fl)) return -1;
        } else if (launch_reduce_move(s->Pf, s->msplits, s->n, rf_scale, s->Rf, s->pos, Rs_dev, c.bc == 1,
                                      c.box, s->st)) {
            return -1;
        }
        s->last_Rs = Rs_dev;
        REC_END(PH_MOVE);
    } else if (c.hydro == BDH_HYDRO_GEYER) {
        // TEA-HI: one pass over the pairs gives D.F, D.X, C and eps (diff_rand.c:86-230)
        REC_BEGIN(PH_MATVEC);
        if (tea_parts(s, xi_dev, dt, s->Rf)) return -1;
        REC_END(PH_MATVEC);
        REC_BEGIN(PH_MOVE);
        if (launch_em_update(s->pos, s->N, s->Rf, s->Rs, c.bc == 1, c.box, s->st, &fl)) return -1;
        REC_END(PH_MOVE);
    } else {
        REC_BEGIN(PH_MOVE);
        hipLaunchKernelGGL(diag_mobility_kernel, dim3(ceil_div(s->n, 256)), dim3(256), 0, s->st,
                           s->pos, s->F, xi_dev, s->N, s->preDii, dt, rf_scale, s->Rf, s->Rs);
        BDH_LAUNCH_CHECK();
        if (launch_em_update(s->pos, s->N, s->Rf, s->Rs, c.bc == 1, c.box, s->st, &fl)) return -1;
        REC_END(PH_MOVE);
    }
    if (compute_forces_dev(s)) return -1;
    return 0;
}

// Gaussians of m consecutive steps in one launch: m calls of gauss_rand for n values equal
// one call for m * ldx values with ldx = n + (n odd) -- every call starts a fresh pair and
// drops the last value of its last pair when n is odd -- so step j reads xi + j * ldx.
int draw_noise(bdh_system* s, int m)
{
    REC_BEGIN(PH_NOISE);
    const int count = m == 1 ? s->n : (int)(m * s->ldx);   // as rng_cnt[m]
    int rc = launch_gauss_wide(s->rng, s->xi, count, s->gwork, s->gwork_bytes, s->flags + 2, s->st);
    REC_END(PH_NOISE);
    return rc;
}

// One step with the reference's rejection loop (comp_step_ermak, diff_comp.c:152-179): when
// the forces at the new positions report an overlap, the positions and forces are restored
// and the step is redrawn with new Gaussians (ermak_var also halves dt, at most 16 times);
// more than move_attempts rejections are fatal.  The factor of the step is kept (the
// reference refactors only when curr_iter changes).
int step_with_retries(bdh_system* s, const double* xi_given, bool refactor)
{
    const bool chol = s->cfg.hydro == BDH_HYDRO_CHOLESKY;
    if (chol && refactor && assemble_and_factor(s)) return -1;
    BDH_HIP(hipMemcpyAsync(s->pos_prev, s->pos, sizeof(double4) * s->N, hipMemcpyDeviceToDevice, s->st));
    BDH_HIP(hipMemcpyAsync(s->F_prev, s->F, sizeof(double) * s->n, hipMemcpyDeviceToDevice, s->st));
    double dt = s->cfg.dt;
    for (int count = 0;; ) {
        // supplied Gaussians serve the first attempt; redraws come from the device stream
        const double* X = count == 0 ? xi_given : nullptr;
        if (!X) {
            if (draw_noise(s, 1)) return -1;
            X = s->xi;
        }
        if (chol && cycle_noise(s, X, s->n, 1, dt)) return -1;
        BDH_HIP(hipMemsetAsync(s->flags, 0, sizeof(int), s->st));
        if (one_step(s, X, chol ? s->RsB : nullptr, dt)) return -1;
        int fl = 0;
        BDH_HIP(hipMemcpyAsync(&fl, s->flags, sizeof(int), hipMemcpyDeviceToHost, s->st));
        BDH_HIP(hipStreamSynchronize(s->st));
        BDH_CHECK(!(fl & 2), "Overstretched bond");
        if (!(fl & 1)) break;
        if (s->cfg.algorithm == BDH_ERMAK_VAR && count < 16) dt /= 2;
        BDH_HIP(hipMemcpyAsync(s->pos, s->pos_prev, sizeof(double4) * s->N, hipMemcpyDeviceToDevice, s->st));
        BDH_HIP(hipMemcpyAsync(s->F, s->F_prev, sizeof(double) * s->n, hipMemcpyDeviceToDevice, s->st));
        ++count;
        BDH_CHECK(count <= s->cfg.move_attempts, "Wrong conformation, can't move");
    }
    s->last_dt = dt;
    s->steps_done++;
    return 0;
}

// Rf = (dt / kT) D F at the current positions and forces into out, for every hydro mode
// (compute_Rf, diff_equation.c:119-214: D.F on the fly, or diag_D without hydrodynamics).
int rf_dev(bdh_system* s, double* out, double dt)
{
    const double scale = dt / (kB * s->cfg.T);
    if (s->cfg.hydro != BDH_HYDRO_NONE) {
        if (matvec_dev(s)) return -1;
        ReduceJobs jobs{};
        jobs.job[0] = ReduceJob{s->Pf, out, s->n, s->msplits, 0, scale};
        return launch_reduce_partials(jobs, 1, s->n, s->st);
    }
    hipLaunchKernelGGL(diag_mobility_kernel, dim3(ceil_div(s->n, 256)), dim3(256), 0, s->st, s->pos,
                       s->F, (const double*)nullptr, s->N, s->preDii, dt, scale, out, (double*)nullptr);
    BDH_LAUNCH_CHECK();
    return 0;
}

// Rs from the next Gaussians of the device stream (compute_r, diff_rand.c:47-236):
// sqrt(2 dt) L xi (L the factor in Lm), the TEA displacement (of the mean tensor of the
// positions and pos2 when pos2 is set), or xi sqrt(2 dt diag_D) without hydrodynamics.
const double* rs_dev(bdh_system* s, double dt, const void* Lm, const double4* pos2 = nullptr)
{
    if (draw_noise(s, 1)) return nullptr;
    if (s->cfg.hydro == BDH_HYDRO_CHOLESKY)
        return cycle_noise(s, s->xi, s->n, 1, dt, Lm) ? nullptr : s->RsB;
    if (s->cfg.hydro == BDH_HYDRO_GEYER) return tea_parts(s, s->xi, dt, nullptr, pos2) ? nullptr : s->Rs;
    hipLaunchKernelGGL(diag_mobility_kernel, dim3(ceil_div(s->n, 256)), dim3(256), 0, s->st, s->pos,
                       s->F, (const double*)s->xi, s->N, s->preDii, dt, 0.0, (double*)nullptr, s->Rs);
    return hipGetLastError() == hipSuccess ? s->Rs : nullptr;
}

// Forces at the current positions; returns the force flags (1 overlap, 2 overstretched bond).
int forces_flags(bdh_system* s, int& fl)
{
    BDH_HIP(hipMemsetAsync(s->flags, 0, sizeof(int), s->st));
    if (compute_forces_dev(s)) return -1;
    BDH_HIP(hipMemcpyAsync(&fl, s->flags, sizeof(int), hipMemcpyDeviceToHost, s->st));
    BDH_HIP(hipStreamSynchronize(s->st));
    BDH_CHECK(!(fl & 2), "Overstretched bond");
    return 0;
}

// One IG-T predictor-corrector step, comp_step_IGT_const / _var (diff_comp.c:76-144) over
// compute_hydro_IGT and its second phase (diff_equation.c:288-374):
//   predictor  r* = r0 + Rf(r0) + sqrt(2 dt) L(r0) xi1, forces at r*;
//   corrector  r  = r0 + (Rf(r0) + Rf(r*)) / 2 + sqrt(2 dt) L_avg xi2, forces at r,
//              L_avg the factor of (D(r0) + D(r*)) / 2.
// With TEA noise the two random moves are the TEA displacements of D(r0) and of
// (D(r0) + D(r*)) / 2, whose blocks are re-evaluated at r0 and r* (input.c:497-501 makes the
// reference read them from its stored tensors).
// A rejected predictor is redrawn with L(r0) kept (igt_var also halves dt).  A rejected
// corrector is redrawn alone with L_avg kept (igt_const; its first rejection is not
// counted), or undoes the whole step (igt_var: dt halved, L(r0) kept).
int step_igt(bdh_system* s)
{
    const bool chol = s->cfg.hydro == BDH_HYDRO_CHOLESKY;
    const bool var = s->cfg.algorithm == BDH_IGT_VAR;
    const size_t pb = sizeof(double4) * s->N, fb = sizeof(double) * s->n;
    double dt = s->cfg.dt;
    int count = 0, fl = 0;
    BDH_HIP(hipMemcpyAsync(s->pos_prev, s->pos, pb, hipMemcpyDeviceToDevice, s->st));
    BDH_HIP(hipMemcpyAsync(s->F_prev, s->F, fb, hipMemcpyDeviceToDevice, s->st));
    if (chol && assemble_and_factor(s)) return -1;
    const double* Rs = nullptr;
    for (;;) {
        // predictor
        if (rf_dev(s, s->Rf, dt)) return -1;
        if (!(Rs = rs_dev(s, dt, s->A))) return -1;
        const FlowTerm pf = flow_term(s, FLOW_MOVE, dt, s->FS);
        if (launch_em_update(s->pos, s->N, s->Rf, Rs, s->cfg.bc == 1, s->cfg.box, s->st, &pf)) return -1;
        if (forces_flags(s, fl)) return -1;
        if (fl & 1) {
            BDH_HIP(hipMemcpyAsync(s->pos, s->pos_prev, pb, hipMemcpyDeviceToDevice, s->st));
            BDH_HIP(hipMemcpyAsync(s->F, s->F_prev, fb, hipMemcpyDeviceToDevice, s->st));
            if (var && count < 16) dt /= 2;
            BDH_CHECK(++count <= s->cfg.move_attempts, "Wrong conformation, can't move");
            continue;
        }
        // mean deterministic move and mean tensor of (r0, r*)
        if (rf_dev(s, s->Rfs, dt)) return -1;
        hipLaunchKernelGGL(mean2_kernel, dim3(ceil_div(s->n, 256)), dim3(256), 0, s->st, s->Rf, s->Rfs,
                           s->n, s->Rf0);
        BDH_LAUNCH_CHECK();
        if (chol && assemble_and_factor(s, s->A2, s->pos_prev)) return -1;
        if (s->tea) BDH_HIP(hipMemcpyAsync(s->pos_star, s->pos, pb, hipMemcpyDeviceToDevice, s->st));
        BDH_HIP(hipMemcpyAsync(s->pos, s->pos_prev, pb, hipMemcpyDeviceToDevice, s->st));
        // corrector
        bool ok = false;
        for (int fails = 0;; ++fails) {
            if (!(Rs = rs_dev(s, dt, s->A2, s->pos_star))) return -1;
            const FlowTerm cf = flow_term(s, FLOW_CORRECTOR, dt, s->FS);
            if (launch_em_update(s->pos, s->N, s->Rf0, Rs, s->cfg.bc == 1, s->cfg.box, s->st, &cf)) return -1;
            if (forces_flags(s, fl)) return -1;
            if (!(fl & 1)) {
                ok = true;
                break;
            }
            BDH_HIP(hipMemcpyAsync(s->pos, s->pos_prev, pb, hipMemcpyDeviceToDevice, s->st));
            if (var) break;
            if (fails > 0) BDH_CHECK(++count <= s->cfg.move_attempts, "Wrong conformation, can't move");
        }
        if (ok) break;
        BDH_HIP(hipMemcpyAsync(s->F, s->F_prev, fb, hipMemcpyDeviceToDevice, s->st));
        if (count < 16) dt /= 2;
        BDH_CHECK(++count <= s->cfg.move_attempts, "Wrong conformation, can't move");
    }
    BDH_HIP(hipMemcpyAsync(s->Rf, s->Rf0, fb, hipMemcpyDeviceToDevice, s->st));
    s->last_Rs = Rs;
    s->last_dt = dt;
    s->steps_done++;
    return 0;
}

// steps steps; have_xi: the caller's Gaussians are in s->xi ([step][n]).  E-M steps without
// rejections are taken in chunks of up to TRMV_MAX_RHS that never cross a refactorization.
int run_steps(bdh_system* s, int steps, bool have_xi)
{
    if (s->cfg.algorithm == BDH_IGT_CONST || s->cfg.algorithm == BDH_IGT_VAR) {
        BDH_CHECK(!have_xi, "IG-T steps draw their Gaussians from the device stream");
        for (int k = 0; k < steps; ++k)
            if (step_igt(s)) return -1;
        return 0;
    }
    const int n = s->n;
    const int R = s->cfg.refactor_every;
    const bool chol = s->cfg.hydro == BDH_HYDRO_CHOLESKY;
    if (s->cfg.check_overlap) {
        // rejections possible: step by step, each draw in the reference's order
        for (int k = 0; k < steps; ++k) {
            const bool refactor = chol && (!s->have_L || s->steps_done % R == 0);
            if (step_with_retries(s, have_xi ? s->xi + (size_t)k * n : nullptr, refactor)) return -1;
        }
        return 0;
    }
    for (int k = 0; k < steps;) {
        int m = std::min(steps - k, TRMV_MAX_RHS);
        bool refactor = false;
        if (chol) {
            const int ph = (int)(s->steps_done % R);
            refactor = !s->have_L || ph == 0;
            m = std::min(m, refactor ? R : R - ph);
        }
        const double* X;
        long long ldx;
        if (have_xi) {
            X = s->xi + (size_t)k * n;
            ldx = n;
        } else {
            if (draw_noise(s, m)) return -1;
            X = s->xi;
            ldx = s->ldx;
        }
        if (chol) {
            // D and L at the positions of the cycle's first step (diff_comp.c:156)
            if (refactor && assemble_and_factor(s)) return -1;
            if (cycle_noise(s, X, ldx, m, s->cfg.dt)) return -1;
        }
        for (int j = 0; j < m; ++j) {
            if (one_step(s, X + (size_t)j * ldx, chol ? s->RsB + (size_t)j * n : nullptr, s->cfg.dt))
                return -1;
            s->steps_done++;
        }
        s->last_dt = s->cfg.dt;
        k += m;
    }
    return 0;
}

}  // namespace

namespace {
template <typename T>
int dpotrf_impl(char uplo, double* A, int n, int ld, int* info, int syrk_mode)
{
    const int n_pad = ceil_div(n, PAD_TO) * PAD_TO;
    const long long lda = n_pad;
    hipStream_t st;
    BDH_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    T* d_A = nullptr;
    double* d_full = nullptr;
    double* d_Lbuf = nullptr;
    int* d_status = nullptr;
    int rc = 0;
    if (dmalloc(&d_A, sizeof(T) * lda * n_pad) || dmalloc(&d_full, sizeof(double) * (size_t)n * n) ||
        dmalloc(&d_Lbuf, sizeof(double) * chol_lbuf_doubles(n_pad)) || dmalloc(&d_status, sizeof(int))) {
        rc = -1;
    } else {
        const bool up = (uplo == 'U' || uplo == 'u');
        dim3 grid(ceil_div(n, 256), n);
        (void)hipMemsetAsync(d_A, 0, sizeof(T) * lda * n_pad, st);
        (void)hipMemsetAsync(d_status, 0, sizeof(int), st);
        (void)hipMemcpy2DAsync(d_full, sizeof(double) * n, A, sizeof(double) * ld, sizeof(double) * n, n,
                               hipMemcpyHostToDevice, st);
        hipLaunchKernelGGL(full_to_lower_kernel<T>, grid, dim3(256), 0, st, d_full, (long long)n, n,
                           d_A, lda, up ? 1 : 0);
        rc = launch_set_identity_pad<T>(d_A, lda, n, n_pad, st);
        if (!rc) rc = chol_factor<T>(d_A, lda, n_pad, d_Lbuf, d_status, syrk_mode, st);
        if (!rc) {
            hipLaunchKernelGGL(lower_to_full_kernel<T>, grid, dim3(256), 0, st, (const T*)d_A, lda, n,
                               d_full, (long long)n, 0);
            std::vector<double> Lh((size_t)n * n);
            (void)hipMemcpyAsync(Lh.data(), d_full, sizeof(double) * (size_t)n * n, hipMemcpyDeviceToHost, st);
            (void)hipMemcpyAsync(info, d_status, sizeof(int), hipMemcpyDeviceToHost, st);
            if (hipStreamSynchronize(st) != hipSuccess) { set_error("bdh_dpotrf: sync failed"); rc = -1; }
            if (!rc) {
                for (int c = 0; c < n; ++c)
                    for (int r = c; r < n; ++r) {
                        const double v = Lh[(size_t)c * n + r];
                        if (up) A[(size_t)r * ld + c] = v;   // U = L^T in the upper triangle
                        else A[(size_t)c * ld + r] = v;
                    }
            }
        }
    }
    dfree(d_A); dfree(d_full); dfree(d_Lbuf); dfree(d_status);
    (void)hipStreamDestroy(st);
    return rc;
}

// ------------------------------------------------------------------ allocation-free Gaussians
// bdh_gauss_noalloc stages through this file-scope device array instead of buffers from
// hipMalloc: a header (stream state at 0, output offset at 8, count at 16, status at 20) and
// room for BDH_GAUSS_NOALLOC_MAX doubles after it.  Every device the library runs on holds its
// own copy until the process ends.  In host code `scratch` is only a stand-in for the symbol:
// its device address is looked up once per device (hipGetSymbolAddress, with that device
// current) and only that address is handed to copies and kernels.
constexpr size_t GAUSS_SCRATCH_HDR = 256;
__device__ __align__(256) unsigned char scratch[GAUSS_SCRATCH_HDR + sizeof(double) * BDH_GAUSS_NOALLOC_MAX];

constexpr int SCRATCH_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch_addr[SCRATCH_MAX_DEVICES] = {};

// Device address of `scratch` on the current device (looked up on the first call per device).
int scratch_address(unsigned char** addr)
{
    int dev = 0;
    BDH_HIP(hipGetDevice(&dev));
    BDH_CHECK(dev >= 0 && dev < SCRATCH_MAX_DEVICES, "scratch: device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_addr[dev]) {
        void* p = nullptr;
        BDH_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        BDH_CHECK(p != nullptr, "scratch: no device address");
        g_scratch_addr[dev] = static_cast<unsigned char*>(p);
    }
    *addr = g_scratch_addr[dev];
    return 0;
}

}  // namespace

extern "C" {

const char* bdh_last_error(void) { return bdh::last_error(); }

int bdh_init(int device)
{
    int n = 0;
    BDH_HIP(hipGetDeviceCount(&n));
    BDH_CHECK(device >= 0 && device < n, "bdh_init: no such device");
    BDH_HIP(hipSetDevice(device));
    if (gauss_prepare()) return -1;
    unsigned char* scratch_base = nullptr;   // the allocation-free entry points' array
    return scratch_address(&scratch_base);
}

void bdh_destroy(bdh_system* s)
{
    if (!s) return;
    (void)hipSetDevice(s->device);
    if (s->st) (void)hipStreamSynchronize(s->st);
    delete s->rec;
    s->rec = nullptr;
    dfree(s->pos); dfree(s->Q); dfree(s->LJ);
    dfree(s->conn_off); dfree(s->conn); dfree(s->bond_off); dfree(s->bond_p); dfree(s->bond_parm);
    dfree(s->F); dfree(s->E); dfree(s->PF); dfree(s->PE);
    dfree(s->A); dfree(s->Lbuf); dfree(s->xi); dfree(s->Pf); dfree(s->Pr); dfree(s->Rf); dfree(s->Rs);
    dfree(s->noise); dfree(s->RsB); dfree(s->pos_prev); dfree(s->F_prev); dfree(s->A2); dfree(s->Rf0); dfree(s->Rfs); dfree(s->FS); dfree(s->pos_star); dfree(s->perm); dfree(s->pos_s); dfree(s->pos2_s); dfree(s->F_s); dfree(s->X_s);
    dfree(s->Rf_s); dfree(s->Rs_s); dfree(s->rng); dfree(s->rng_off); dfree(s->rng_cnt); dfree(s->gwork); dfree(s->flags);
    dfree(s->Px); dfree(s->Pc); dfree(s->Pe); dfree(s->Feff); dfree(s->Cs); dfree(s->bt);
    dfree(s->cl.grid); dfree(s->cl.key); dfree(s->cl.count); dfree(s->cl.fill);
    dfree(s->cl.start); dfree(s->cl.sorted);
    if (s->factor_exec) (void)hipGraphExecDestroy(s->factor_exec);
    dfree(s->adj);
    dfree(s->fcache);
    dfree(s->bond_id);
    dfree(s->pgrad);
    if (s->st) (void)hipStreamDestroy(s->st);
    delete s;
}

// Reach of any pair term (the DH and LJ cutoffs and the overlap test R_i + R_j > r, padded
// well above the rounding of the T-precision comparisons in pair_term), and the cell edge of
// init_cutoff (pre_func.c:197-231): every potential's cutoff and twice the largest radius,
// plus 0.01.  Both follow the LJ radii.
static void set_reach(bdh_system* s, const double* LJ)
{
    const bdh_config* cfg = &s->cfg;
    double rmax = 0.0;
    for (int i = 0; i < s->N; ++i) rmax = std::max(rmax, LJ[2 * i]);
    double reach = 2.0 * rmax;
    if (cfg->elec) reach = std::max(reach, cfg->cutoff_c);
    if (cfg->alpha_lj != 0.0)
        reach = std::max(reach, cfg->cutoff_lj > 0 ? cfg->cutoff_lj : 1.122462 * 2.0 * rmax);
    reach = reach * (1.0 + 1e-5) + 1e-5;
    s->fp.far_r2 = reach * reach;
    double cut = 2.0 * rmax;
    if (cfg->elec) cut = std::max(cut, cfg->cutoff_c);
    if (cfg->alpha_lj != 0.0)
        cut = std::max(cut, cfg->cutoff_lj > 0 ? cfg->cutoff_lj : 1.1225 * 2.0 * rmax);
    s->cell_cut = cut + 0.01;
}

static int create_impl(const bdh_config* cfg, int N, const double* coord, const double* Q,
                       const double* LJ, int nbonds, const int* bonds, const double* bond_parms,
                       bdh_system* s)
{
    BDH_CHECK(cfg && coord && Q && LJ, "bdh_create: null argument");
    BDH_CHECK(N > 0, "bdh_create: no beads");
    BDH_CHECK(cfg->hydro == BDH_HYDRO_NONE || cfg->hydro == BDH_HYDRO_CHOLESKY ||
              cfg->hydro == BDH_HYDRO_GEYER, "bdh_create: hydro must be none, cholesky or geyer");
    BDH_CHECK(cfg->bc == 0 || cfg->bc == 1, "bdh_create: bc must be none or periodic");
    // periodic boundaries with hydrodynamics mean an Ewald-summed tensor (input.c:390)
    BDH_CHECK(!(cfg->bc == 1 && cfg->hydro != BDH_HYDRO_NONE &&
                cfg->ewald_method != EW_SMITH && cfg->ewald_method != EW_BEENAKKER),
              "bdh_create: ewald_method smith (0) or beenakker (1)");
    BDH_CHECK(!(cfg->bc == 1 && cfg->hydro != BDH_HYDRO_NONE && cfg->ewald_method == EW_SMITH &&
                !(cfg->ewald_alpha > 0)), "bdh_create: ewald_alpha must be positive");
    BDH_CHECK(!(cfg->bc == 1 && cfg->hydro != BDH_HYDRO_NONE) ||
              (cfg->ewald_real >= 0 && cfg->ewald_real <= 5), "bdh_create: ewald_real 0..5");
    BDH_CHECK(cfg->bc == 0 || cfg->box > 0, "bdh_create: box edge");
    BDH_CHECK(cfg->precision == BDH_FP64 || cfg->precision == BDH_MIXED || cfg->precision == BDH_MIXED_SPLIT,
              "bdh_create: precision");
    BDH_CHECK(cfg->dt > 0 && cfg->T > 0 && cfg->visc != 0 && cfg->vfactor != 0, "bdh_create: dt/T/visc");
    BDH_CHECK(nbonds == 0 || (bonds && bond_parms), "bdh_create: bonds");
    const bool igt = cfg->algorithm == BDH_IGT_CONST || cfg->algorithm == BDH_IGT_VAR;
    BDH_CHECK(cfg->algorithm == BDH_ERMAK_CONST || cfg->algorithm == BDH_ERMAK_VAR || igt,
              "bdh_create: algorithm (ermak_const, ermak_var, igt_const or igt_var)");
    BDH_CHECK(!igt || cfg->refactor_every <= 1, "bdh_create: IG-T factors D at every step (refactor_every 1)");
    BDH_CHECK(cfg->move_attempts >= 0, "bdh_create: move_attempts");
    s->cfg = *cfg;
    if (s->cfg.refactor_every < 1) s->cfg.refactor_every = 1;
    s->last_dt = cfg->dt;
    BDH_HIP(hipGetDevice(&s->device));
    BDH_HIP(hipStreamCreateWithFlags(&s->st, hipStreamNonBlocking));
    s->N = N;
    s->n = 3 * N;
    s->n_pad = ceil_div(s->n, PAD_TO) * PAD_TO;
    s->ld = s->n_pad;
    s->mixed = cfg->precision != BDH_FP64;
    s->syrk_mode = cfg->precision == BDH_MIXED_SPLIT ? SYRK_BF16X3 : (s->mixed ? SYRK_F32 : SYRK_F64);
    s->preDii = kB_1_6PI * cfg->T / (cfg->visc * cfg->vfactor);   // diff_decl.c:93
    s->preDij = kB_1_8PI * cfg->T / (cfg->visc * cfg->vfactor);   // diff_decl.c:100
    for (int i = 0; i < N; ++i)
        BDH_CHECK(coord[4 * i + 3] > 0.0, "bdh_create: bead radius must be positive");
    for (int b = 0; b < nbonds; ++b)
        BDH_CHECK(bonds[2 * b] >= 0 && bonds[2 * b] < N && bonds[2 * b + 1] >= 0 &&
                  bonds[2 * b + 1] < N && bonds[2 * b] != bonds[2 * b + 1],
                  "bdh_create: bond index out of range");

    // force parameters
    ForceParams& fp = s->fp;
    fp.alpha_lj = cfg->alpha_lj;
    fp.cutoff_lj = cfg->cutoff_lj;
    fp.lj_6_term = cfg->lj_6_term ? 1.0 : 0.0;
    fp.elec = cfg->elec;
    fp.gam = cfg->gamma_c / cfg->epsilon_c;
    fp.kappa_c = cfg->kappa_c;
    fp.cutoff_c = cfg->cutoff_c;
    fp.bond_lj_scale = cfg->bond_lj_scale;
    fp.bond_c_scale = cfg->bond_c_scale;
    fp.use_conn = nbonds && ((cfg->elec && cfg->bond_c_scale != 1) ||
                             (cfg->alpha_lj != 0.0 && cfg->bond_lj_scale != 1));
    fp.check_overlap = cfg->check_overlap;
    fp.pbc = cfg->bc == 1;
    fp.box = cfg->box;
    set_reach(s, LJ);

    // per-bead adjacency (CSR), host-built
    std::vector<int> deg(N + 1, 0);
    for (int b = 0; b < nbonds; ++b) { deg[bonds[2 * b]]++; deg[bonds[2 * b + 1]]++; }
    std::vector<int> off(N + 1, 0);
    for (int i = 0; i < N; ++i) off[i + 1] = off[i] + deg[i];
    std::vector<int> fill(off.begin(), off.end() - 1), partner(std::max(1, off[N]));
    std::vector<double> parm(3 * std::max(1, off[N]));
    s->nbonds = nbonds;
    s->csr_bond.assign(std::max(1, off[N]), 0);
    for (int b = 0; b < nbonds; ++b) {
        const int i = bonds[2 * b], j = bonds[2 * b + 1];
        for (int e = 0; e < 2; ++e) {
            const int me = e ? j : i, other = e ? i : j;
            const int k = fill[me]++;
            partner[k] = other;
            s->csr_bond[k] = b;
            parm[3 * k] = bond_parms[3 * b];
            parm[3 * k + 1] = bond_parms[3 * b + 1];
            parm[3 * k + 2] = bond_parms[3 * b + 2];
        }
    }
    if (dmalloc(&s->conn_off, sizeof(int) * (N + 1))) return -1;
    if (dmalloc(&s->bond_off, sizeof(int) * (N + 1))) return -1;
    if (dmalloc(&s->conn, sizeof(int) * partner.size())) return -1;
    if (dmalloc(&s->bond_p, sizeof(int) * partner.size())) return -1;
    if (dmalloc(&s->bond_parm, sizeof(double) * parm.size())) return -1;
    if (dmalloc(&s->bond_id, sizeof(int) * s->csr_bond.size())) return -1;
    BDH_HIP(hipMemcpy(s->bond_id, s->csr_bond.data(), sizeof(int) * s->csr_bond.size(), hipMemcpyHostToDevice));
    BDH_HIP(hipMemcpy(s->conn_off, off.data(), sizeof(int) * (N + 1), hipMemcpyHostToDevice));
    BDH_HIP(hipMemcpy(s->bond_off, off.data(), sizeof(int) * (N + 1), hipMemcpyHostToDevice));
    BDH_HIP(hipMemcpy(s->conn, partner.data(), sizeof(int) * partner.size(), hipMemcpyHostToDevice));
    BDH_HIP(hipMemcpy(s->bond_p, partner.data(), sizeof(int) * partner.size(), hipMemcpyHostToDevice));
    BDH_HIP(hipMemcpy(s->bond_parm, parm.data(), sizeof(double) * parm.size(), hipMemcpyHostToDevice));
    s->topo = Topology{s->conn_off, s->conn, s->bond_off, s->bond_p, s->bond_parm};

    // beads
    if (dmalloc(&s->pos, sizeof(double4) * N)) return -1;
    if (dmalloc(&s->Q, sizeof(double) * N)) return -1;
    if (dmalloc(&s->LJ, sizeof(double2) * N)) return -1;
    BDH_HIP(hipMemcpy(s->pos, coord, sizeof(double) * 4 * N, hipMemcpyHostToDevice));
    BDH_HIP(hipMemcpy(s->Q, Q, sizeof(double) * N, hipMemcpyHostToDevice));
    BDH_HIP(hipMemcpy(s->LJ, LJ, sizeof(double) * 2 * N, hipMemcpyHostToDevice));

    // work buffers
    s->cells = cfg->nb_list != 0;
    s->fsplits = force_splits(N);
    s->fblocks = s->cells ? cell_force_blocks(N) : force_blocks(N);
    if (s->cells) {
        if (alloc_cells(s)) return -1;
    }
    s->ewald = cfg->bc == 1 && cfg->hydro != BDH_HYDRO_NONE;
    s->tea = cfg->hydro == BDH_HYDRO_GEYER;
    s->msplits = pair_splits(N, s->ewald);
    s->tblocks = trmv_col_blocks(s->n);
    if (dmalloc(&s->F, sizeof(double) * s->n)) return -1;
    if (dmalloc(&s->E, sizeof(double) * 4)) return -1;
    if (dmalloc(&s->PF, sizeof(double) * s->n * s->fsplits)) return -1;
    if (dmalloc(&s->PE, sizeof(double) * 3 * s->fblocks)) return -1;
    if (dmalloc(&s->Pf, sizeof(double) * s->n * std::max(1, s->msplits))) return -1;
    if (s->tea) {
        if (dmalloc(&s->Px, sizeof(double) * s->n * s->msplits)) return -1;
        if (dmalloc(&s->Pc, sizeof(double) * s->n * s->msplits)) return -1;
        if (dmalloc(&s->Pe, sizeof(double) * pair_eparts(N))) return -1;
        if (dmalloc(&s->Feff, sizeof(double) * s->n)) return -1;
        if (dmalloc(&s->Cs, sizeof(double) * s->n)) return -1;
        if (dmalloc(&s->bt, sizeof(double) * 2)) return -1;
    }
    if (dmalloc(&s->Pr, sizeof(double) * s->n * s->tblocks * TRMV_MAX_RHS)) return -1;
    if (dmalloc(&s->RsB, sizeof(double) * s->n * TRMV_MAX_RHS)) return -1;
    if (cfg->check_overlap || igt) {
        if (dmalloc(&s->pos_prev, sizeof(double4) * N)) return -1;
        if (dmalloc(&s->F_prev, sizeof(double) * s->n)) return -1;
    }
    if (igt) {
        if (dmalloc(&s->Rf0, sizeof(double) * s->n)) return -1;
        if (dmalloc(&s->Rfs, sizeof(double) * s->n)) return -1;
        if (s->tea && dmalloc(&s->pos_star, sizeof(double4) * N)) return -1;
    }
    if (dmalloc(&s->Rf, sizeof(double) * s->n)) return -1;
    if (dmalloc(&s->Rs, sizeof(double) * s->n)) return -1;
    if (dmalloc(&s->noise, sizeof(double) * s->n)) return -1;
    s->ldx = s->n + (s->n & 1);
    s->xi_cap = (size_t)s->ldx * TRMV_MAX_RHS;
    if (dmalloc(&s->xi, sizeof(double) * s->xi_cap)) return -1;
    if (dmalloc(&s->flags, sizeof(int) * 4)) return -1;
    // every clear goes on the system's own (non-blocking) stream, ordered before the
    // kernels that follow; a null-stream hipMemset would not be
    BDH_HIP(hipMemsetAsync(s->flags, 0, sizeof(int) * 4, s->st));
    BDH_HIP(hipMemsetAsync(s->Rf, 0, sizeof(double) * s->n, s->st));
    BDH_HIP(hipMemsetAsync(s->Rs, 0, sizeof(double) * s->n, s->st));
    if (dmalloc(&s->rng, sizeof(uint64_t))) return -1;
    if (dmalloc(&s->rng_off, sizeof(long long))) return -1;
    if (dmalloc(&s->rng_cnt, sizeof(int) * (TRMV_MAX_RHS + 1))) return -1;
    s->gwork_bytes = gauss_wide_scratch_bytes((long long)s->xi_cap);
    if (dmalloc(&s->gwork, s->gwork_bytes)) return -1;
    {
        uint64_t st0 = lcg_seed(cfg->rand_seed);
        long long o = 0;
        int cnt[TRMV_MAX_RHS + 1];
        for (int m = 0; m <= TRMV_MAX_RHS; ++m) cnt[m] = (int)(m * s->ldx);
        cnt[1] = s->n;   // a single draw keeps the reference's own length
        BDH_HIP(hipMemcpy(s->rng, &st0, sizeof(st0), hipMemcpyHostToDevice));
        BDH_HIP(hipMemcpy(s->rng_off, &o, sizeof(o), hipMemcpyHostToDevice));
        BDH_HIP(hipMemcpy(s->rng_cnt, cnt, sizeof(cnt), hipMemcpyHostToDevice));
    }
    if (cfg->hydro == BDH_HYDRO_CHOLESKY) {
        const size_t es = s->mixed ? sizeof(float) : sizeof(double);
        const size_t bytes = es * (size_t)s->ld * (size_t)s->n_pad;
        if (dmalloc(&s->Lbuf, sizeof(double) * chol_lbuf_doubles(s->n_pad))) return -1;
        for (void** M : {&s->A, &s->A2}) {
            if (M == &s->A2 && !igt) break;
            if (dmalloc(M, bytes)) return -1;
            BDH_HIP(hipMemsetAsync(*M, 0, bytes, s->st));
            if (s->mixed) {
                if (launch_set_identity_pad<float>((float*)*M, s->ld, s->n, s->n_pad, s->st)) return -1;
            } else {
                if (launch_set_identity_pad<double>((double*)*M, s->ld, s->n, s->n_pad, s->st)) return -1;
            }
        }
    }
    // initial forces (main_loop.c:91)
    if (compute_forces_dev(s)) return -1;
    BDH_HIP(hipStreamSynchronize(s->st));
    return check_flags(s);
}

// The pair sums' bead order of a periodic TEA system with at least PAIR_ORDER_MIN beads: the
// Morton order of the positions (wrapped into the box, 1024 cells per axis), ties by index.
// Set at creation and renewed once PAIR_ORDER_EVERY steps have passed since (at the end of a
// bdh_step call, or with positions from bdh_set_coords), as the beads diffuse away from it.  It
// affects only which beads share a tile, not the sums (beyond their rounding).
constexpr int PAIR_ORDER_MIN = 256;
constexpr long long PAIR_ORDER_EVERY = 4096;

static uint32_t spread_bits10(uint32_t v)
{
    v &= 0x3FF;
    v = (v | (v << 16)) & 0x030000FF;
    v = (v | (v << 8)) & 0x0300F00F;
    v = (v | (v << 4)) & 0x030C30C3;
    v = (v | (v << 2)) & 0x09249249;
    return v;
}

static int set_pair_order(bdh_system* s, const double* coord)
{
    if (!(s->tea && s->ewald && s->N >= PAIR_ORDER_MIN)) return 0;
    const int N = s->N, n = s->n;
    if (!s->perm) {
        if (dmalloc(&s->perm, sizeof(int) * N) || dmalloc(&s->pos_s, sizeof(double4) * N) ||
            dmalloc(&s->F_s, sizeof(double) * n) || dmalloc(&s->X_s, sizeof(double) * n) ||
            dmalloc(&s->Rf_s, sizeof(double) * n) || dmalloc(&s->Rs_s, sizeof(double) * n))
            return -1;
        if (dmalloc(&s->pos2_s, sizeof(double4) * N)) return -1;
    }
    const double L = s->cfg.box;
    std::vector<std::pair<uint32_t, int>> key(N);
    for (int i = 0; i < N; ++i) {
        uint32_t q[3];
        for (int k = 0; k < 3; ++k) {
            double u = coord[4 * i + k] / L + 0.5;
            u -= std::floor(u);
            q[k] = (uint32_t)std::min(1023.0, std::max(0.0, std::floor(u * 1024.0)));
        }
        key[i] = {spread_bits10(q[0]) | (spread_bits10(q[1]) << 1) | (spread_bits10(q[2]) << 2), i};
    }
    std::sort(key.begin(), key.end());
    std::vector<int> perm(N);
    for (int i = 0; i < N; ++i) perm[i] = key[i].second;
    BDH_HIP(hipMemcpyAsync(s->perm, perm.data(), sizeof(int) * N, hipMemcpyHostToDevice, s->st));
    BDH_HIP(hipStreamSynchronize(s->st));
    s->order_step = s->steps_done;
    return 0;
}

// Renews the pair order when it is PAIR_ORDER_EVERY steps old (or with force), from the host
// positions coord (N x 4) or, without them, from the device's.
static int refresh_pair_order(bdh_system* s, const double* coord, bool force)
{
    if (!s->perm || (!force && s->steps_done - s->order_step < PAIR_ORDER_EVERY)) return 0;
    std::vector<double> h;
    if (!coord) {
        h.resize(4 * (size_t)s->N);
        BDH_HIP(hipMemcpyAsync(h.data(), s->pos, sizeof(double4) * s->N, hipMemcpyDeviceToHost, s->st));
        BDH_HIP(hipStreamSynchronize(s->st));
        coord = h.data();
    }
    return set_pair_order(s, coord);
}

int bdh_create(const bdh_config* cfg, int N, const double* coord, const double* Q,
               const double* LJ, int nbonds, const int* bonds, const double* bond_parms,
               bdh_system** out)
{
    if (!out) { set_error("bdh_create: null out"); return -2; }
    *out = nullptr;
    bdh_system* s = new bdh_system();
    int rc = create_impl(cfg, N, coord, Q, LJ, nbonds, bonds, bond_parms, s);
    if (!rc) rc = set_pair_order(s, coord);
    if (rc) {
        std::string keep = g_err;
        bdh_destroy(s);
        g_err = keep;
        return rc;
    }
    *out = s;
    return 0;
}

int bdh_step(bdh_system* s, int steps, const double* xi)
{
    BDH_CHECK(s && steps >= 0, "bdh_step: bad arguments");
    BDH_HIP(hipSetDevice(s->device));
    if (xi) {
        if ((size_t)steps * s->n > s->xi_cap) {
            BDH_HIP(hipStreamSynchronize(s->st));
            dfree(s->xi);
            s->xi_cap = (size_t)steps * s->n;
            if (dmalloc(&s->xi, sizeof(double) * s->xi_cap)) return -1;
        }
        BDH_HIP(hipMemcpyAsync(s->xi, xi, sizeof(double) * s->n * (size_t)steps,
                               hipMemcpyHostToDevice, s->st));
    }
    if (run_steps(s, steps, xi != nullptr)) return -1;
    BDH_HIP(hipStreamSynchronize(s->st));
    if (refresh_pair_order(s, nullptr, false)) return -1;
    return check_flags(s);
}

int bdh_compute_D(bdh_system* s, double* D, int ld)
{
    BDH_CHECK(s && D && ld >= s->n, "bdh_compute_D: bad arguments");
    BDH_HIP(hipSetDevice(s->device));
    const size_t es = s->mixed ? sizeof(float) : sizeof(double);
    void* tmp = nullptr;
    double* full = nullptr;
    BDH_HIP(hipMallocAsync(&tmp, es * (size_t)s->ld * s->n_pad, s->st));
    BDH_HIP(hipMallocAsync((void**)&full, sizeof(double) * (size_t)s->n * s->n, s->st));
    BDH_HIP(hipMemsetAsync(tmp, 0, es * (size_t)s->ld * s->n_pad, s->st));
    dim3 grid(ceil_div(s->n, 256), s->n);
    int rc = ensure_ewald(s);
    if (rc) {
    } else if (s->mixed) {
        rc = s->ewald ? launch_ewald_assemble<float>(s->pos, s->N, s->preDii, s->preDij, (float*)tmp, s->ld, s->st)
                      : launch_rpy_assemble<float>(s->pos, s->N, s->preDii, s->preDij, (float*)tmp, s->ld, s->st);
        if (!rc) hipLaunchKernelGGL(lower_to_full_kernel<float>, grid, dim3(256), 0, s->st,
                                    (const float*)tmp, s->ld, s->n, full, (long long)s->n, 1);
    } else {
        rc = s->ewald ? launch_ewald_assemble<double>(s->pos, s->N, s->preDii, s->preDij, (double*)tmp, s->ld, s->st)
                      : launch_rpy_assemble<double>(s->pos, s->N, s->preDii, s->preDij, (double*)tmp, s->ld, s->st);
        if (!rc) hipLaunchKernelGGL(lower_to_full_kernel<double>, grid, dim3(256), 0, s->st,
                                    (const double*)tmp, s->ld, s->n, full, (long long)s->n, 1);
    }
    if (!rc) {
        BDH_HIP(hipMemcpy2DAsync(D, sizeof(double) * ld, full, sizeof(double) * s->n,
                                 sizeof(double) * s->n, s->n, hipMemcpyDeviceToHost, s->st));
    }
    BDH_HIP(hipFreeAsync(tmp, s->st));
    BDH_HIP(hipFreeAsync(full, s->st));
    BDH_HIP(hipStreamSynchronize(s->st));
    return rc;
}

int bdh_compute_forces(bdh_system* s, double* F, double* E, int* repeat)
{
    BDH_CHECK(s, "bdh_compute_forces: null system");
    BDH_HIP(hipSetDevice(s->device));
    BDH_HIP(hipMemsetAsync(s->flags, 0, sizeof(int), s->st));
    if (compute_forces_dev(s, E != nullptr)) return -1;
    int fl = 0;
    if (F) BDH_HIP(hipMemcpyAsync(F, s->F, sizeof(double) * s->n, hipMemcpyDeviceToHost, s->st));
    if (E) BDH_HIP(hipMemcpyAsync(E, s->E, sizeof(double) * 3, hipMemcpyDeviceToHost, s->st));
    BDH_HIP(hipMemcpyAsync(&fl, s->flags, sizeof(int), hipMemcpyDeviceToHost, s->st));
    BDH_HIP(hipStreamSynchronize(s->st));
    if (repeat) *repeat = (fl & 1) ? 1 : 0;
    BDH_CHECK(!(fl & 2), "Overstretched bond");
    return 0;
}

int bdh_neighbour_pairs(bdh_system* s, double cutoff, int* pairs, long long cap, long long* count)
{
    BDH_CHECK(s && count && cutoff > 0.0 && cap >= 0 && (pairs || cap == 0),
              "bdh_neighbour_pairs: bad arguments");
    BDH_HIP(hipSetDevice(s->device));
    if (alloc_cells(s)) return -1;
    int* d_pairs = nullptr;
    unsigned long long* d_count = nullptr;
    const long long pcap = std::max(1LL, cap);
    BDH_HIP(hipMallocAsync((void**)&d_pairs, sizeof(int) * 2 * pcap, s->st));
    BDH_HIP(hipMallocAsync((void**)&d_count, sizeof(unsigned long long), s->st));
    int rc = launch_cell_build(s->pos, s->N, cutoff, s->fp.pbc, s->fp.box, s->cl, s->st);
    if (!rc) rc = launch_cell_pairs(s->pos, s->N, s->cl, cutoff, s->fp.pbc, s->fp.box, d_pairs, cap,
                                    d_count, s->st);
    unsigned long long got = 0;
    if (!rc) {
        BDH_HIP(hipMemcpyAsync(&got, d_count, sizeof(got), hipMemcpyDeviceToHost, s->st));
        BDH_HIP(hipStreamSynchronize(s->st));
        const long long kept = std::min<long long>((long long)got, cap);
        if (kept > 0) {
            BDH_HIP(hipMemcpyAsync(pairs, d_pairs, sizeof(int) * 2 * kept, hipMemcpyDeviceToHost, s->st));
            BDH_HIP(hipStreamSynchronize(s->st));
            std::vector<std::pair<int, int>> v(kept);
            for (long long k = 0; k < kept; ++k) v[k] = {pairs[2 * k], pairs[2 * k + 1]};
            std::sort(v.begin(), v.end());
            for (long long k = 0; k < kept; ++k) { pairs[2 * k] = v[k].first; pairs[2 * k + 1] = v[k].second; }
        }
    }
    (void)hipFreeAsync(d_pairs, s->st);
    (void)hipFreeAsync(d_count, s->st);
    BDH_HIP(hipStreamSynchronize(s->st));
    *count = (long long)got;
    return rc;
}

int bdh_factor(bdh_system* s, double* L, int ld)
{
    BDH_CHECK(s && s->A, "bdh_factor: system has no Cholesky state");
    BDH_HIP(hipSetDevice(s->device));
    if (assemble_and_factor(s)) return -1;
    if (L) {
        BDH_CHECK(ld >= s->n, "bdh_factor: ld < 3N");
        double* full = nullptr;
        BDH_HIP(hipMallocAsync((void**)&full, sizeof(double) * (size_t)s->n * s->n, s->st));
        dim3 grid(ceil_div(s->n, 256), s->n);
        if (s->mixed)
            hipLaunchKernelGGL(lower_to_full_kernel<float>, grid, dim3(256), 0, s->st,
                               (const float*)s->A, s->ld, s->n, full, (long long)s->n, 0);
        else
            hipLaunchKernelGGL(lower_to_full_kernel<double>, grid, dim3(256), 0, s->st,
                               (const double*)s->A, s->ld, s->n, full, (long long)s->n, 0);
        BDH_LAUNCH_CHECK();
        BDH_HIP(hipMemcpy2DAsync(L, sizeof(double) * ld, full, sizeof(double) * s->n,
                                 sizeof(double) * s->n, s->n, hipMemcpyDeviceToHost, s->st));
        BDH_HIP(hipFreeAsync(full, s->st));
    }
    BDH_HIP(hipStreamSynchronize(s->st));
    return check_flags(s);
}

int bdh_noise_product(bdh_system* s, const double* x, double* y)
{
    BDH_CHECK(s && s->A && s->have_L && x && y, "bdh_noise_product: no factor");
    BDH_HIP(hipSetDevice(s->device));
    BDH_HIP(hipMemcpyAsync(s->noise, x, sizeof(double) * s->n, hipMemcpyHostToDevice, s->st));
    if (trmv_dev(s, s->noise, s->n, 1)) return -1;
    ReduceJobs jobs{};
    jobs.job[0] = ReduceJob{s->Pr, s->Rs, s->n, s->tblocks, trmv_blk_cols(), 1.0};
    if (launch_reduce_partials(jobs, 1, s->n, s->st)) return -1;
    BDH_HIP(hipMemcpyAsync(y, s->Rs, sizeof(double) * s->n, hipMemcpyDeviceToHost, s->st));
    BDH_HIP(hipStreamSynchronize(s->st));
    return 0;
}

int bdh_compute_Rf(bdh_system* s, double* Rf)
{
    BDH_CHECK(s && Rf, "bdh_compute_Rf: bad arguments");
    BDH_HIP(hipSetDevice(s->device));
    const double rf_scale = s->cfg.dt / (kB * s->cfg.T);
    if (s->cfg.hydro != BDH_HYDRO_NONE) {
        if (matvec_dev(s)) return -1;
        ReduceJobs jobs{};
        jobs.job[0] = ReduceJob{s->Pf, s->Rf, s->n, s->msplits, 0, rf_scale};
        if (launch_reduce_partials(jobs, 1, s->n, s->st)) return -1;
    } else {
        hipLaunchKernelGGL(diag_mobility_kernel, dim3(ceil_div(s->n, 256)), dim3(256), 0, s->st,
                           s->pos, s->F, s->F, s->N, s->preDii, s->cfg.dt, rf_scale, s->Rf,
                           (double*)nullptr);
        BDH_LAUNCH_CHECK();
    }
    BDH_HIP(hipMemcpyAsync(Rf, s->Rf, sizeof(double) * s->n, hipMemcpyDeviceToHost, s->st));
    BDH_HIP(hipStreamSynchronize(s->st));
    return 0;
}

// Rs of compute_r (diff_rand.c:47-236) at the current positions from the next 3N Gaussians
// of the device stream: sqrt(2 dt) L xi (the factor refreshed when `refactor`), the TEA
// displacement, or sqrt(2 dt diag_D) xi.  Returns the device vector, nullptr on error.
static const double* noise_at(bdh_system* s, double dt, bool refactor)
{
    if (s->cfg.hydro == BDH_HYDRO_CHOLESKY && (refactor || !s->have_L) && assemble_and_factor(s)) return nullptr;
    return rs_dev(s, dt, s->A);
}

static int ensure_prev(bdh_system* s)
{
    if (!s->pos_prev && dmalloc(&s->pos_prev, sizeof(double4) * s->N)) return -1;
    if (!s->F_prev && dmalloc(&s->F_prev, sizeof(double) * s->n)) return -1;
    return 0;
}

int bdh_compute_Dij(bdh_system* s, int i, int j, int kind, double* tD)
{
    BDH_CHECK(s && tD && i >= 0 && i < s->N && j >= 0 && j < s->N, "bdh_compute_Dij: bad arguments");
    BDH_CHECK(kind >= 0 && kind <= 3, "bdh_compute_Dij: kind 0 (the system's tensor), 1 free space, "
                                      "2 Smith-Ewald, 3 Beenakker-Ewald");
    BDH_HIP(hipSetDevice(s->device));
    int ew = 0;
    if (kind == 0) ew = s->ewald ? (s->cfg.ewald_method == EW_SMITH ? 2 : 1) : 0;
    else if (kind >= 2) {
        BDH_CHECK(s->ewald && s->cfg.ewald_method == (kind == 2 ? EW_SMITH : EW_BEENAKKER),
                  "bdh_compute_Dij: the Ewald block of a method the system was not created with");
        ew = kind == 2 ? 2 : 1;
    }
    BDH_CHECK(ew != 0 || kind == 0 || i != j, "bdh_compute_Dij: the free-space block needs i != j");
    if (ensure_ewald(s)) return -1;
    if (launch_single_block(s->pos, i, j, ew, s->preDii, s->preDij, s->noise, s->st)) return -1;
    BDH_HIP(hipMemcpyAsync(tD, s->noise, sizeof(double) * 9, hipMemcpyDeviceToHost, s->st));
    BDH_HIP(hipStreamSynchronize(s->st));
    return 0;
}

int bdh_compute_r(bdh_system* s, double* Rs)
{
    BDH_CHECK(s && Rs, "bdh_compute_r: bad arguments");
    BDH_HIP(hipSetDevice(s->device));
    const double* d = noise_at(s, s->cfg.dt, s->cfg.refactor_every == 1);
    if (!d) return -1;
    BDH_HIP(hipMemcpyAsync(Rs, d, sizeof(double) * s->n, hipMemcpyDeviceToHost, s->st));
    BDH_HIP(hipStreamSynchronize(s->st));
    return check_flags(s);
}

int bdh_ermak_move(bdh_system* s, double dt, double* coord_out)
{
    BDH_CHECK(s && coord_out && dt > 0, "bdh_ermak_move: bad arguments");
    BDH_CHECK(s->cfg.algorithm == BDH_ERMAK_CONST || s->cfg.algorithm == BDH_ERMAK_VAR,
              "bdh_ermak_move: an Ermak-McCammon system");
    BDH_HIP(hipSetDevice(s->device));
    if (ensure_prev(s)) return -1;
    const bool refactor = !s->have_L || s->steps_done % s->cfg.refactor_every == 0;
    BDH_HIP(hipMemcpyAsync(s->pos_prev, s->pos, sizeof(double4) * s->N, hipMemcpyDeviceToDevice, s->st));
    if (rf_dev(s, s->Rf, dt)) return -1;
    const double* Rs = noise_at(s, dt, refactor);
    if (!Rs) return -1;
    const FlowTerm fl = flow_term(s, FLOW_MOVE, dt, s->FS);
    if (launch_em_update(s->pos, s->N, s->Rf, Rs, s->cfg.bc == 1, s->cfg.box, s->st, &fl)) return -1;
    BDH_HIP(hipMemcpyAsync(coord_out, s->pos, sizeof(double4) * s->N, hipMemcpyDeviceToHost, s->st));
    BDH_HIP(hipMemcpyAsync(s->pos, s->pos_prev, sizeof(double4) * s->N, hipMemcpyDeviceToDevice, s->st));
    if (Rs != s->Rs) BDH_HIP(hipMemcpyAsync(s->Rs, Rs, sizeof(double) * s->n, hipMemcpyDeviceToDevice, s->st));
    s->last_Rs = s->Rs;
    s->last_dt = dt;
    s->steps_done++;
    BDH_HIP(hipStreamSynchronize(s->st));
    return check_flags(s);
}

int bdh_igt_phase(bdh_system* s, int phase, double dt, double* coord_out, int* result)
{
    BDH_CHECK(s && coord_out && result && dt > 0 && (phase == 0 || phase == 1), "bdh_igt_phase: bad arguments");
    BDH_CHECK(s->cfg.algorithm == BDH_IGT_CONST || s->cfg.algorithm == BDH_IGT_VAR,
              "bdh_igt_phase: an IG-T system");
    BDH_HIP(hipSetDevice(s->device));
    const bool chol = s->cfg.hydro == BDH_HYDRO_CHOLESKY;
    const size_t pb = sizeof(double4) * s->N, fb = sizeof(double) * s->n;
    int fl = 0;
    *result = 0;
    if (phase == 0) {
        // predictor r* = r0 + Rf(r0) + sqrt(2 dt) L(r0) xi1 and forces at r*
        BDH_HIP(hipMemcpyAsync(s->pos_prev, s->pos, pb, hipMemcpyDeviceToDevice, s->st));
        BDH_HIP(hipMemcpyAsync(s->F_prev, s->F, fb, hipMemcpyDeviceToDevice, s->st));
        if (chol && assemble_and_factor(s)) return -1;
        if (rf_dev(s, s->Rf, dt)) return -1;
        const double* Rs = rs_dev(s, dt, s->A);
        if (!Rs) return -1;
        const FlowTerm pf = flow_term(s, FLOW_MOVE, dt, s->FS);
        if (launch_em_update(s->pos, s->N, s->Rf, Rs, s->cfg.bc == 1, s->cfg.box, s->st, &pf)) return -1;
        if (forces_flags(s, fl)) return -1;
        if (fl & 1) {
            // positions back at r0; the forces stay those of r* (the caller restores them)
            BDH_HIP(hipMemcpyAsync(s->pos, s->pos_prev, pb, hipMemcpyDeviceToDevice, s->st));
            BDH_HIP(hipMemcpyAsync(coord_out, s->pos, pb, hipMemcpyDeviceToHost, s->st));
            BDH_HIP(hipStreamSynchronize(s->st));
            *result = 1;
            return 0;
        }
        // mean deterministic move and the factor of the mean tensor of (r0, r*)
        if (rf_dev(s, s->Rfs, dt)) return -1;
        hipLaunchKernelGGL(mean2_kernel, dim3(ceil_div(s->n, 256)), dim3(256), 0, s->st, s->Rf, s->Rfs, s->n,
                           s->Rf0);
        BDH_LAUNCH_CHECK();
        if (chol && assemble_and_factor(s, s->A2, s->pos_prev)) return -1;
        if (s->tea) BDH_HIP(hipMemcpyAsync(s->pos_star, s->pos, pb, hipMemcpyDeviceToDevice, s->st));
        BDH_HIP(hipMemcpyAsync(s->pos, s->pos_prev, pb, hipMemcpyDeviceToDevice, s->st));
    }
    // corrector r = r0 + Rf0 + sqrt(2 dt) L_avg xi2 (TEA: the sums of the mean tensor) and
    // forces at r
    const double* Rs = rs_dev(s, dt, s->A2, s->pos_star);
    if (!Rs) return -1;
    const FlowTerm cf = flow_term(s, FLOW_CORRECTOR, dt, s->FS);
    if (launch_em_update(s->pos, s->N, s->Rf0, Rs, s->cfg.bc == 1, s->cfg.box, s->st, &cf)) return -1;
    if (forces_flags(s, fl)) return -1;
    BDH_HIP(hipMemcpyAsync(coord_out, s->pos, pb, hipMemcpyDeviceToHost, s->st));
    BDH_HIP(hipStreamSynchronize(s->st));
    s->last_Rs = Rs;
    s->last_dt = dt;
    *result = (fl & 1) ? 2 : 0;
    return 0;
}

int bdh_tea_coefficients(bdh_system* s, double* beta, double* eps)
{
    BDH_CHECK(s && beta && eps && s->tea, "bdh_tea_coefficients: a TEA system");
    double h[2];
    BDH_HIP(hipMemcpyAsync(h, s->bt, sizeof(h), hipMemcpyDeviceToHost, s->st));
    BDH_HIP(hipStreamSynchronize(s->st));
    *beta = h[0];
    *eps = h[1];
    return 0;
}

int bdh_set_ewald_split(bdh_system* s, double alpha_scale, int R_real, int R_recip)
{
    BDH_CHECK(s && s->ewald && s->cfg.ewald_method == EW_BEENAKKER,
              "bdh_set_ewald_split: a periodic system with the Beenakker tensor");
    BDH_CHECK(alpha_scale >= 0.0 && (alpha_scale == 0.0 || (R_real >= 0 && R_real <= 5 && R_recip >= 0 &&
                                                             R_recip <= 5)),
              "bdh_set_ewald_split: alpha scale >= 0, radii 0..5");
    BDH_HIP(hipSetDevice(s->device));
    s->ew_scale = alpha_scale;
    s->ew_Rr = alpha_scale > 0.0 ? R_real : -1;
    s->ew_Rk = alpha_scale > 0.0 ? R_recip : -1;
    if (alpha_scale == 0.0) s->ew_rcut = 0.0;
    // the captured refactor reads the tables of the last upload: rebuild it with the new ones
    if (s->factor_exec) {
        BDH_HIP(hipGraphExecDestroy(s->factor_exec));
        s->factor_exec = nullptr;
    }
    return ensure_ewald(s);
}

int bdh_set_ewald_cutoff(bdh_system* s, double r_cut)
{
    BDH_CHECK(s && s->ewald && s->cfg.ewald_method == EW_BEENAKKER && s->ew_scale > 0.0,
              "bdh_set_ewald_cutoff: a periodic Beenakker system with its own split (bdh_set_ewald_split)");
    BDH_CHECK(r_cut >= 0.0, "bdh_set_ewald_cutoff: r_cut >= 0");
    BDH_HIP(hipSetDevice(s->device));
    s->ew_rcut = r_cut;
    if (s->factor_exec) {
        BDH_HIP(hipGraphExecDestroy(s->factor_exec));
        s->factor_exec = nullptr;
    }
    return ensure_ewald(s);
}

int bdh_set_flow(bdh_system* s, const double* G)
{
    BDH_CHECK(s, "bdh_set_flow: null system");
    BDH_HIP(hipSetDevice(s->device));
    if (!G) {
        s->flow = false;
        return 0;
    }
    for (int k = 0; k < 9; ++k) BDH_CHECK(std::isfinite(G[k]), "bdh_set_flow: G must be finite");
    if (!s->FS && dmalloc(&s->FS, sizeof(double) * s->n)) return -1;
    BDH_HIP(hipMemsetAsync(s->FS, 0, sizeof(double) * s->n, s->st));
    for (int k = 0; k < 9; ++k) s->G[k] = G[k];
    s->flow = true;
    return 0;
}

int bdh_set_forces(bdh_system* s, const double* F)
{
    BDH_CHECK(s && F, "bdh_set_forces: bad arguments");
    BDH_HIP(hipSetDevice(s->device));
    BDH_HIP(hipMemcpyAsync(s->F, F, sizeof(double) * s->n, hipMemcpyHostToDevice, s->st));
    BDH_HIP(hipStreamSynchronize(s->st));
    return 0;
}

int bdh_get_coords(bdh_system* s, double* coord)
{
    BDH_CHECK(s && coord, "bdh_get_coords: bad arguments");
    BDH_HIP(hipMemcpyAsync(coord, s->pos, sizeof(double) * 4 * s->N, hipMemcpyDeviceToHost, s->st));
    BDH_HIP(hipStreamSynchronize(s->st));
    return 0;
}

int bdh_set_coords(bdh_system* s, const double* coord)
{
    BDH_CHECK(s && coord, "bdh_set_coords: bad arguments");
    BDH_HIP(hipMemcpyAsync(s->pos, coord, sizeof(double) * 4 * s->N, hipMemcpyHostToDevice, s->st));
    if (refresh_pair_order(s, coord, s->steps_done == 0)) return -1;
    if (compute_forces_dev(s)) return -1;
    BDH_HIP(hipStreamSynchronize(s->st));
    return 0;
}

// Bytes of the factor in A: n_pad^2 entries of FP64, or of FP32 for the mixed precisions.
static size_t factor_bytes(const bdh_system* s)
{
    return (s->mixed ? sizeof(float) : sizeof(double)) * (size_t)s->n_pad * s->n_pad;
}

int bdh_factor_cache(bdh_system* s, int slots)
{
    BDH_CHECK(s && slots >= 0, "bdh_factor_cache: bad arguments");
    BDH_HIP(hipSetDevice(s->device));
    BDH_HIP(hipStreamSynchronize(s->st));
    dfree(s->fcache);
    s->fcache_slots = 0;
    if (slots == 0) return 0;
    if (dmalloc(&s->fcache, factor_bytes(s) * slots)) return -1;
    s->fcache_slots = slots;
    return 0;
}

int bdh_save_factor(bdh_system* s, int slot)
{
    BDH_CHECK(s && s->have_L, "bdh_save_factor: no factor");
    BDH_CHECK(slot >= 0 && slot < s->fcache_slots, "bdh_save_factor: slot out of range");
    BDH_HIP(hipSetDevice(s->device));
    const size_t fb = factor_bytes(s);
    BDH_HIP(hipMemcpyAsync(s->fcache + fb * slot, s->A, fb, hipMemcpyDeviceToDevice, s->st));
    return 0;
}

struct ParamBars {
    double *eps, *R, *Q, *H;
};

static int step_vjp_impl(bdh_system* s, int slot, const double* coord, const double* xi,
                         const double* gbar, double* rbar, const ParamBars* pb = nullptr);

int bdh_step_vjp(bdh_system* s, const double* coord, const double* xi, const double* gbar, double* rbar)
{
    return step_vjp_impl(s, -1, coord, xi, gbar, rbar);
}

int bdh_step_vjp_cached(bdh_system* s, int slot, const double* coord, const double* xi,
                        const double* gbar, double* rbar)
{
    BDH_CHECK(s && slot >= 0 && slot < s->fcache_slots, "bdh_step_vjp_cached: slot out of range");
    return step_vjp_impl(s, slot, coord, xi, gbar, rbar);
}

int bdh_step_vjp_params(bdh_system* s, int slot, const double* coord, const double* xi,
                        const double* gbar, double* rbar, double* eps_bar, double* R_bar,
                        double* Q_bar, double* H_bar)
{
    BDH_CHECK(s && (slot < 0 || slot < s->fcache_slots), "bdh_step_vjp_params: slot out of range");
    const ParamBars pb{eps_bar, R_bar, Q_bar, H_bar};
    return step_vjp_impl(s, slot, coord, xi, gbar, rbar, &pb);
}

int bdh_set_params(bdh_system* s, const double* Q, const double* LJ, const double* bond_parms)
{
    BDH_CHECK(s, "bdh_set_params: null system");
    BDH_HIP(hipSetDevice(s->device));
    if (Q) BDH_HIP(hipMemcpyAsync(s->Q, Q, sizeof(double) * s->N, hipMemcpyHostToDevice, s->st));
    if (LJ) {
        for (int i = 0; i < s->N; ++i)
            BDH_CHECK(LJ[2 * i] >= 0.0 && LJ[2 * i + 1] >= 0.0, "bdh_set_params: LJ radius and epsilon must be >= 0");
        BDH_HIP(hipMemcpyAsync(s->LJ, LJ, sizeof(double) * 2 * s->N, hipMemcpyHostToDevice, s->st));
        set_reach(s, LJ);
    }
    if (bond_parms && s->nbonds) {
        std::vector<double> parm(3 * s->csr_bond.size());
        for (size_t k = 0; k < s->csr_bond.size(); ++k)
            for (int u = 0; u < 3; ++u) parm[3 * k + u] = bond_parms[3 * s->csr_bond[k] + u];
        BDH_HIP(hipMemcpyAsync(s->bond_parm, parm.data(), sizeof(double) * parm.size(), hipMemcpyHostToDevice,
                               s->st));
        BDH_HIP(hipStreamSynchronize(s->st));   // parm is a host temporary
    }
    // the forces of the current positions under the new parameters (the next step uses them)
    if (compute_forces_dev(s)) return -1;
    BDH_HIP(hipStreamSynchronize(s->st));
    return check_flags(s);
}

static int step_vjp_impl(bdh_system* s, int slot, const double* coord, const double* xi,
                         const double* gbar, double* rbar, const ParamBars* pb)
{
    BDH_CHECK(s && coord && xi && gbar && rbar, "bdh_step_vjp: bad arguments");
    BDH_CHECK(s->cfg.hydro == BDH_HYDRO_CHOLESKY && !s->ewald, "bdh_step_vjp: free-space Cholesky noise only");
    // the derivative is that of one accepted ermak_const step with the given Gaussians: a
    // rejected step (check_overlap) is redrawn from the stream with the same or a halved dt,
    // which the given-xi step does not follow
    BDH_CHECK(!s->cfg.check_overlap && s->cfg.algorithm == BDH_ERMAK_CONST,
              "bdh_step_vjp: ermak_const steps without check_overlap only");
    BDH_CHECK(!s->flow, "bdh_step_vjp: no flow field");
    BDH_HIP(hipSetDevice(s->device));
    const int n = s->n, np = s->n_pad;
    const size_t nn = (size_t)np * np;
    const size_t ninv = (size_t)(np / 64) * 4096;
    // mixed precision: room for the FP32 factor promoted to FP64 after the vectors
    const size_t nL = s->mixed ? nn : 0;
    if (!s->adj && dmalloc(&s->adj, sizeof(double) * (2 * nn + 7 * (size_t)n + ninv + nL))) return -1;
    double* Y1 = s->adj;
    double* Y2 = Y1 + nn;
    double* xi_d = Y2 + nn;
    double* g_d = xi_d + n;
    double* a_d = g_d + n;
    double* Dg = a_d + n;
    double* ctr = Dg + n;
    double* out = ctr + n;
    double* Minv = out + n;
    double* Lp = Minv + ninv;   // mixed: the factor in FP64
    // the forward step's state at coord: forces and the factor of D
    BDH_HIP(hipMemcpyAsync(s->pos, coord, sizeof(double) * 4 * s->N, hipMemcpyHostToDevice, s->st));
    if (compute_forces_dev(s)) return -1;
    if (slot >= 0) {
        // the factor at coord, saved when the forward pass stepped from there
        BDH_HIP(hipMemcpyAsync(s->A, s->fcache + factor_bytes(s) * slot, factor_bytes(s),
                               hipMemcpyDeviceToDevice, s->st));
        s->have_L = true;
    } else if (assemble_and_factor(s)) {
        return -1;
    }
    BDH_HIP(hipMemcpyAsync(xi_d, xi, sizeof(double) * n, hipMemcpyHostToDevice, s->st));
    BDH_HIP(hipMemcpyAsync(g_d, gbar, sizeof(double) * n, hipMemcpyHostToDevice, s->st));
    const double dt = s->cfg.dt;
    const double c = dt / (kB * s->cfg.T);
    const double sq = std::sqrt(2.0 * dt);
    // the derivative of the step that was taken: with a mixed-precision factor, of the FP32
    // factor the noise was drawn with (promoted; the solves below run in FP64)
    const double* L = (const double*)s->A;
    if (s->mixed) {
        if (launch_adj_promote((const float*)s->A, (long long)nn, Lp, s->st)) return -1;
        L = Lp;
    }
    // S^T = L^-T (L^-T s Phi(a xi^T))^T, a = L^T g
    if (launch_adj_lt_g(L, s->ld, n, g_d, a_d, s->st)) return -1;
    if (launch_adj_diag_inverses(L, s->ld, np, Minv, s->st)) return -1;
    if (launch_adj_upper_solve(L, s->ld, np, n, 0, a_d, xi_d, sq, nullptr, Minv, Y1, s->st)) return -1;
    if (launch_adj_upper_solve(L, s->ld, np, n, 1, nullptr, nullptr, 0.0, Y1, Minv, Y2, s->st)) return -1;
    // Z = Y2 + Y2^T into Y1's storage (Y1 is no longer needed); Y2's storage then holds the
    // contraction's partial sums (at most 64 x n doubles <= n_pad^2)
    if (launch_adj_symmetrize(np, Y2, Y1, s->st)) return -1;
    if (launch_adj_contract(s->pos, s->N, s->preDii, s->preDij, Y1, np, g_d, s->F, c, Y2, ctr, s->st))
        return -1;
    // D g from blocks on the fly, then g + contraction - c H (D g)
    if (launch_pair_sums<double>(s->pos, g_d, nullptr, s->N, s->preDii, s->preDij, dt, false, false,
                                 s->msplits, s->Pf, s->Px, s->Pc, s->Pe, s->st))
        return -1;
    ReduceJobs jobs{};
    jobs.job[0] = ReduceJob{s->Pf, Dg, n, s->msplits, 0, 1.0};
    if (launch_reduce_partials(jobs, 1, n, s->st)) return -1;
    if (launch_adj_hvp(s->pos, s->Q, s->LJ, s->N, s->topo, s->fp, Dg, ctr, g_d, c, Y2, out, s->st)) return -1;
    BDH_HIP(hipMemcpyAsync(rbar, out, sizeof(double) * n, hipMemcpyDeviceToHost, s->st));
    if (pb) {
        // (c D g)^T dF/dtheta; the pair partials reuse Y2's storage, free again after the
        // Hessian-vector product (stream order)
        const int N = s->N;
        if (!s->pgrad && dmalloc(&s->pgrad, sizeof(double) * (3 * (size_t)N + std::max(1, s->nbonds)))) return -1;
        double* pe = s->pgrad;
        if (launch_adj_params(s->pos, s->Q, s->LJ, N, s->topo, s->bond_id, s->fp, Dg, c, Y2, pe, pe + N,
                              pe + 2 * N, pe + 3 * N, s->st))
            return -1;
        double* outs[4] = {pb->eps, pb->R, pb->Q, pb->H};
        for (int q = 0; q < 4; ++q) {
            const size_t cnt = q < 3 ? (size_t)N : (size_t)s->nbonds;
            if (outs[q] && cnt)
                BDH_HIP(hipMemcpyAsync(outs[q], pe + (size_t)q * N, sizeof(double) * cnt, hipMemcpyDeviceToHost,
                                       s->st));
        }
    }
    BDH_HIP(hipStreamSynchronize(s->st));
    return check_flags(s);
}

int bdh_get_forces(bdh_system* s, double* F)
{
    BDH_CHECK(s && F, "bdh_get_forces: bad arguments");
    BDH_HIP(hipMemcpyAsync(F, s->F, sizeof(double) * s->n, hipMemcpyDeviceToHost, s->st));
    BDH_HIP(hipStreamSynchronize(s->st));
    return 0;
}

int bdh_last_dt(bdh_system* s, double* dt)
{
    BDH_CHECK(s && dt, "bdh_last_dt: bad arguments");
    *dt = s->last_dt;
    return 0;
}

int bdh_get_last_displacement(bdh_system* s, double* Rf, double* Rs)
{
    BDH_CHECK(s, "bdh_get_last_displacement: null system");
    if (Rf) BDH_HIP(hipMemcpyAsync(Rf, s->Rf, sizeof(double) * s->n, hipMemcpyDeviceToHost, s->st));
    const double* rs = s->last_Rs ? s->last_Rs : s->Rs;
    if (Rs) BDH_HIP(hipMemcpyAsync(Rs, rs, sizeof(double) * s->n, hipMemcpyDeviceToHost, s->st));
    BDH_HIP(hipStreamSynchronize(s->st));
    return 0;
}

int bdh_get_rng_state(bdh_system* s, uint64_t* state)
{
    BDH_CHECK(s && state, "bdh_get_rng_state: bad arguments");
    BDH_HIP(hipMemcpyAsync(state, s->rng, sizeof(uint64_t), hipMemcpyDeviceToHost, s->st));
    BDH_HIP(hipStreamSynchronize(s->st));
    return 0;
}

int bdh_set_rng_state(bdh_system* s, uint64_t state)
{
    BDH_CHECK(s, "bdh_set_rng_state: null system");
    state &= (1ULL << 48) - 1;
    BDH_HIP(hipMemcpyAsync(s->rng, &state, sizeof(uint64_t), hipMemcpyHostToDevice, s->st));
    BDH_HIP(hipStreamSynchronize(s->st));
    return 0;
}

uint64_t bdh_seed_state(uint32_t seed) { return lcg_seed(seed); }

int bdh_gauss(uint64_t* state, double* out, int count)
{
    BDH_CHECK(state && (out || count == 0) && count >= 0, "bdh_gauss: bad arguments");
    if (count == 0) return 0;
    hipStream_t st;
    BDH_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    uint64_t* d_state = nullptr;
    double* d_out = nullptr;
    long long* d_off = nullptr;
    int *d_cnt = nullptr, *d_status = nullptr;
    int rc = 0;
    long long off = 0;
    int h_status = 0;
    if (dmalloc(&d_state, 8) || dmalloc(&d_out, sizeof(double) * count) || dmalloc(&d_off, 8) ||
        dmalloc(&d_cnt, 4) || dmalloc(&d_status, 4)) {
        rc = -1;
    } else {
        (void)hipMemcpyAsync(d_state, state, 8, hipMemcpyHostToDevice, st);
        (void)hipMemcpyAsync(d_off, &off, 8, hipMemcpyHostToDevice, st);
        (void)hipMemcpyAsync(d_cnt, &count, 4, hipMemcpyHostToDevice, st);
        (void)hipMemsetAsync(d_status, 0, 4, st);
        rc = launch_gauss_streams(d_state, d_out, d_off, d_cnt, 1, d_status, st);
        if (!rc) {
            (void)hipMemcpyAsync(out, d_out, sizeof(double) * count, hipMemcpyDeviceToHost, st);
            (void)hipMemcpyAsync(state, d_state, 8, hipMemcpyDeviceToHost, st);
            (void)hipMemcpyAsync(&h_status, d_status, 4, hipMemcpyDeviceToHost, st);
            if (hipStreamSynchronize(st) != hipSuccess) { set_error("bdh_gauss: sync failed"); rc = -1; }
            if (!rc && h_status) { set_error("bdh_gauss: stream did not complete"); rc = -5; }
        }
    }
    dfree(d_state); dfree(d_out); dfree(d_off); dfree(d_cnt); dfree(d_status);
    (void)hipStreamDestroy(st);
    return rc;
}

int bdh_gauss_noalloc(uint64_t* state, double* out, int count)
{
    BDH_CHECK(state && (out || count == 0) && count >= 0, "bdh_gauss_noalloc: bad arguments");
    BDH_CHECK(count <= BDH_GAUSS_NOALLOC_MAX,
              "bdh_gauss_noalloc: count exceeds BDH_GAUSS_NOALLOC_MAX (1048576); use bdh_gauss");
    if (count == 0) return 0;
    unsigned char* base = nullptr;
    if (scratch_address(&base)) return -1;
    // the header is written whole on every call: the array keeps the previous call's contents
    struct Hdr {
        uint64_t state;
        long long off;
        int cnt;
        int status;
    };
    static_assert(sizeof(Hdr) == 24 && sizeof(Hdr) <= GAUSS_SCRATCH_HDR, "scratch header layout");
    const Hdr h{*state, 0, count, 0};
    Hdr back{};
    uint64_t* d_state = reinterpret_cast<uint64_t*>(base);
    long long* d_off = reinterpret_cast<long long*>(base + offsetof(Hdr, off));
    int* d_cnt = reinterpret_cast<int*>(base + offsetof(Hdr, cnt));
    int* d_status = reinterpret_cast<int*>(base + offsetof(Hdr, status));
    double* d_out = reinterpret_cast<double*>(base + GAUSS_SCRATCH_HDR);
    const hipStream_t st = hipStreamPerThread;
    BDH_HIP(hipMemcpyAsync(base, &h, sizeof(Hdr), hipMemcpyHostToDevice, st));
    int rc = launch_gauss_streams(d_state, d_out, d_off, d_cnt, 1, d_status, st);
    if (!rc) {
        BDH_HIP(hipMemcpyAsync(out, d_out, sizeof(double) * count, hipMemcpyDeviceToHost, st));
        BDH_HIP(hipMemcpyAsync(&back, base, sizeof(Hdr), hipMemcpyDeviceToHost, st));
        BDH_HIP(hipStreamSynchronize(st));
        BDH_CHECK(back.status == 0, "bdh_gauss_noalloc: stream did not complete");
        *state = back.state;
    }
    return rc;
}

int bdh_dpotrf(char uplo, double* A, int n, int ld, int precision)
{
    BDH_CHECK(A && n > 0 && ld >= n, "bdh_dpotrf: bad arguments");
    BDH_CHECK(uplo == 'U' || uplo == 'u' || uplo == 'L' || uplo == 'l', "bdh_dpotrf: uplo");
    int info = 0;
    int rc = precision == BDH_FP64 ? dpotrf_impl<double>(uplo, A, n, ld, &info, SYRK_F64)
                                   : dpotrf_impl<float>(uplo, A, n, ld, &info,
                                                        precision == BDH_MIXED_SPLIT ? SYRK_BF16X3 : SYRK_F32);
    if (rc) return rc;
    BDH_CHECK(info >= 0, "bdh_dpotrf: a panel did not receive its diagonal block");
    return info;
}

int bdh_dtrmv(char uplo, const double* A, const double* v, double* o, int n, int ld, int precision)
{
    BDH_CHECK(A && v && o && n > 0 && ld >= n, "bdh_dtrmv: bad arguments");
    const bool up = (uplo == 'U' || uplo == 'u');
    const int n_pad = ceil_div(n, PAD_TO) * PAD_TO;
    hipStream_t st;
    BDH_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    double *d_full = nullptr, *d_x = nullptr, *d_P = nullptr, *d_y = nullptr;
    void* d_A = nullptr;
    const size_t es = precision != BDH_FP64 ? 4 : 8;
    const int nblk = trmv_col_blocks(n);
    int rc = 0;
    if (dmalloc(&d_full, sizeof(double) * (size_t)n * n) || dmalloc(&d_A, es * (size_t)n_pad * n_pad) ||
        dmalloc(&d_x, sizeof(double) * n) || dmalloc(&d_P, sizeof(double) * (size_t)n * nblk) ||
        dmalloc(&d_y, sizeof(double) * n)) {
        rc = -1;
    } else {
        dim3 grid(ceil_div(n, 256), n);
        (void)hipMemsetAsync(d_A, 0, es * (size_t)n_pad * n_pad, st);
        (void)hipMemcpy2DAsync(d_full, sizeof(double) * n, A, sizeof(double) * ld, sizeof(double) * n, n,
                               hipMemcpyHostToDevice, st);
        (void)hipMemcpyAsync(d_x, v, sizeof(double) * n, hipMemcpyHostToDevice, st);
        if (es == 4) {
            hipLaunchKernelGGL(full_to_lower_kernel<float>, grid, dim3(256), 0, st, d_full, (long long)n, n,
                               (float*)d_A, (long long)n_pad, up ? 1 : 0);
            rc = launch_trmv_partials<float>((const float*)d_A, n_pad, n, d_x, n, 1, d_P, st);
        } else {
            hipLaunchKernelGGL(full_to_lower_kernel<double>, grid, dim3(256), 0, st, d_full, (long long)n, n,
                               (double*)d_A, (long long)n_pad, up ? 1 : 0);
            rc = launch_trmv_partials<double>((const double*)d_A, n_pad, n, d_x, n, 1, d_P, st);
        }
        if (!rc) {
            ReduceJobs jobs{};
            jobs.job[0] = ReduceJob{d_P, d_y, n, nblk, trmv_blk_cols(), 1.0};
            rc = launch_reduce_partials(jobs, 1, n, st);
        }
        if (!rc) {
            (void)hipMemcpyAsync(o, d_y, sizeof(double) * n, hipMemcpyDeviceToHost, st);
            if (hipStreamSynchronize(st) != hipSuccess) { set_error("bdh_dtrmv: sync failed"); rc = -1; }
        }
    }
    dfree(d_full); dfree(d_A); dfree(d_x); dfree(d_P); dfree(d_y);
    (void)hipStreamDestroy(st);
    return rc;
}

int bdh_probe_mfma(int kind, const void* A, const void* B, void* C)
{
    static const int dims[5][3] = {{32, 2, 32}, {16, 4, 16}, {16, 4, 16}, {32, 16, 32}, {16, 32, 16}};
    BDH_CHECK(kind >= 0 && kind < 5 && A && B && C, "bdh_probe_mfma: bad arguments");
    const int M = dims[kind][0], K = dims[kind][1], N = dims[kind][2];
    const size_t es = kind == 2 ? 8 : 4;
    void *dA = nullptr, *dB = nullptr, *dC = nullptr;
    int rc = 0;
    if (dmalloc(&dA, es * M * K) || dmalloc(&dB, es * K * N) || dmalloc(&dC, es * M * N)) {
        rc = -1;
    } else {
        (void)hipMemcpy(dA, A, es * M * K, hipMemcpyHostToDevice);
        (void)hipMemcpy(dB, B, es * K * N, hipMemcpyHostToDevice);
        (void)hipMemset(dC, 0, es * M * N);
        rc = probe_mfma(kind, dA, dB, dC, nullptr);
        if (!rc && hipMemcpy(C, dC, es * M * N, hipMemcpyDeviceToHost) != hipSuccess) {
            set_error("bdh_probe_mfma: copy back failed");
            rc = -1;
        }
    }
    dfree(dA); dfree(dB); dfree(dC);
    return rc;
}

int bdh_time_factor(bdh_system* s, int reps, double* ms)
{
    BDH_CHECK(s && s->A && reps > 0 && ms, "bdh_time_factor: bad arguments");
    hipEvent_t e0, e1;
    BDH_HIP(hipEventCreate(&e0));
    BDH_HIP(hipEventCreate(&e1));
    BDH_HIP(hipEventRecord(e0, s->st));
    for (int r = 0; r < reps; ++r)
        if (assemble_and_factor(s)) return -1;
    BDH_HIP(hipEventRecord(e1, s->st));
    BDH_HIP(hipEventSynchronize(e1));
    float t = 0;
    BDH_HIP(hipEventElapsedTime(&t, e0, e1));
    *ms = t / reps;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return check_flags(s);
}

int bdh_time_steps(bdh_system* s, int steps, double* ms)
{
    BDH_CHECK(s && steps > 0 && ms, "bdh_time_steps: bad arguments");
    hipEvent_t e0, e1;
    BDH_HIP(hipEventCreate(&e0));
    BDH_HIP(hipEventCreate(&e1));
    BDH_HIP(hipEventRecord(e0, s->st));
    if (run_steps(s, steps, false)) return -1;
    BDH_HIP(hipEventRecord(e1, s->st));
    BDH_HIP(hipEventSynchronize(e1));
    float t = 0;
    BDH_HIP(hipEventElapsedTime(&t, e0, e1));
    *ms = t;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return check_flags(s);
}

int bdh_phase_timing(bdh_system* s, int enable)
{
    BDH_CHECK(s, "bdh_phase_timing: null system");
    BDH_HIP(hipStreamSynchronize(s->st));
    delete s->rec;
    s->rec = enable ? new EventRecorder() : nullptr;
    return 0;
}

int bdh_phase_times(bdh_system* s, double* ms, long long* counts)
{
    BDH_CHECK(s && s->rec && ms && counts, "bdh_phase_times: timing not enabled");
    BDH_HIP(hipStreamSynchronize(s->st));
    for (int p = 0; p < PH_COUNT; ++p) { ms[p] = 0.0; counts[p] = 0; }
    for (auto& r : s->rec->recs) {
        float t = 0.f;
        BDH_HIP(hipEventElapsedTime(&t, r.e0, r.e1));
        ms[r.phase] += t;
        counts[r.phase] += 1;
    }
    s->rec->clear();
    return 0;
}

int bdh_sync(bdh_system* s)
{
    BDH_CHECK(s, "bdh_sync: null system");
    BDH_HIP(hipStreamSynchronize(s->st));
    return check_flags(s);
}

}  // extern "C"

// ------------------------------------------------------------------ batches of chains
// Independent systems of uneven sizes sharing one configuration, stepped together: each
// phase of the step (assembly, every launch of the factorization, L.xi, D.F, move, forces,
// the Gaussian streams) is one launch over all chains.  Per chain the launch shapes and the
// arithmetic are those of a bdh_system of its own, so chain c's trajectory is the one
// bdh_create / bdh_step give for its beads with rand_seed seeds[c].
struct bdh_batch {
    bdh_config cfg{};
    int B = 0, Ntot = 0, ntot = 0;
    bool mixed = false;
    int syrk_mode = SYRK_F64;
    int device = 0;
    hipStream_t st = nullptr;
    double preDii = 0, preDij = 0;
    ForceParams fp{};
    Topology topo{};
    std::vector<BatchChain> hc;
    std::vector<int> chol_blocks;   // workgroups of update launch k (host)
    std::vector<int> tab;           // host copy of every prefix table
    int left_max = -1;              // update schedule (batch_set_tiles)
    int outer = 0;                  // outer panel width of every chain (batch_set_tiles)
    BatchPlan plan;
    BatchChain* d_ch = nullptr;
    int* d_tab = nullptr;           // every prefix table
    double4* pos = nullptr;
    double* Q = nullptr;
    double2* LJ = nullptr;
    int *conn_off = nullptr, *conn = nullptr, *bond_off = nullptr, *bond_p = nullptr;
    double* bond_parm = nullptr;
    double *F = nullptr, *Rf = nullptr, *RsB = nullptr, *xi = nullptr, *xi_user = nullptr;
    size_t xi_user_cap = 0;
    double *Pp = nullptr, *Pf = nullptr, *Pt = nullptr;
    void* A = nullptr;
    double* Lbuf = nullptr;
    uint64_t* rng = nullptr;
    long long* rng_off = nullptr;
    int* rng_cnt = nullptr;         // [m][B]: m * ldx_c Gaussians for m steps
    int* flags = nullptr;           // [0] force flags, [1] rng status
    int* status = nullptr;          // [B] factorization status per chain
    long long steps_done = 0;
    bool have_L = false;
    EventRecorder* rec = nullptr;
    hipGraphExec_t factor_exec = nullptr;   // assembly + factorization of every chain, captured once
    hipGraphExec_t cycle_exec = nullptr;    // a whole refactor cycle of device-stream steps
};

namespace {

int batch_check(bdh_batch* b)
{
    std::vector<int> st(b->B);
    int h[2];
    BDH_HIP(hipMemcpyAsync(h, b->flags, sizeof(h), hipMemcpyDeviceToHost, b->st));
    BDH_HIP(hipMemcpyAsync(st.data(), b->status, sizeof(int) * b->B, hipMemcpyDeviceToHost, b->st));
    BDH_HIP(hipStreamSynchronize(b->st));
    for (int c = 0; c < b->B; ++c) {
        if (st[c] < 0) {
            set_error("Cholesky decomposition error in chain " + std::to_string(c) +
                      ": a panel did not receive its diagonal block");
            return -3;
        }
        if (st[c]) {
            set_error("Cholesky decomposition error in chain " + std::to_string(c) +
                      ": non-positive pivot at row " + std::to_string(st[c]));
            return -3;
        }
    }
    if (h[0] & 2) {
        set_error("Overstretched bond");
        return -4;
    }
    if (h[1]) {
        set_error("Gaussian stream did not complete");
        return -5;
    }
    return 0;
}

#define BREC_BEGIN(ph) do { if (b->rec) b->rec->begin(ph, b->st); } while (0)
#define BREC_END(ph) do { if (b->rec) b->rec->end(ph, b->st); } while (0)

int batch_forces(bdh_batch* b)
{
    BREC_BEGIN(PH_FORCES);
    const int rc = b->mixed
        ? launch_forces_batch<float>(b->pos, b->Q, b->LJ, b->topo, b->fp, b->plan, b->Pf, b->flags, b->st)
        : launch_forces_batch<double>(b->pos, b->Q, b->LJ, b->topo, b->fp, b->plan, b->Pf, b->flags, b->st);
    if (rc || launch_reduce_batch(b->Pf, 1, 1, b->F, 0, 1.0, b->plan, b->st)) return -1;
    BREC_END(PH_FORCES);
    return 0;
}

int batch_factor_launches(bdh_batch* b);

// The refactor of every chain (the same ~50 launches with the same arguments each time) is
// captured into a HIP graph on first use and replayed; phase timing runs the launches.
int batch_factor(bdh_batch* b)
{
    if (b->rec || !BDH_FACTOR_GRAPH) return batch_factor_launches(b);
    if (!b->factor_exec) {
        BDH_HIP(hipStreamBeginCapture(b->st, hipStreamCaptureModeThreadLocal));
        const int rc = batch_factor_launches(b);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(b->st, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        BDH_HIP(e);
        const hipError_t ei = hipGraphInstantiate(&b->factor_exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ei != hipSuccess) b->factor_exec = nullptr;
        BDH_HIP(ei);
    }
    BDH_HIP(hipGraphLaunch(b->factor_exec, b->st));
    b->have_L = true;
    return 0;
}

// Update schedule and outer panel width of every chain's factorization: chains with n_pad <=
// left_max update left-looking inside each outer panel of `outer` columns, the others
// right-looking (chol_step); left_max < 0 gives each chain the schedule a single system of
// its size uses.  The update launches' prefix tables are rewritten, and the captured graphs,
// whose grids they size, are dropped.
int batch_set_tiles(bdh_batch* b, int left_max, int outer)
{
    BDH_CHECK(outer == 256 || outer == 512 || outer == 1024, "bdh_batch_set_tiles: outer width 256, 512 or 1024");
    const int B = b->B, W = B + 1;
    for (auto& h : b->hc) {
        h.left = left_max < 0 ? chol_left_looking(h.n_pad) : h.n_pad <= left_max;
        h.outer = outer;
    }
    for (int k = 0; k < b->plan.chol_steps; ++k) {
        int* t = b->tab.data() + (size_t)(6 + k) * W;
        long long acc = 0;
        for (int c = 0; c < B; ++c) {
            t[c] = (int)acc;
            acc += chol_update_blocks(b->hc[c].n_pad, b->hc[c].left != 0, b->hc[c].outer, k);
        }
        BDH_CHECK(acc <= (long long)INT32_MAX, "bdh_batch_set_schedule: batch too large");
        t[B] = (int)acc;
        b->chol_blocks[k] = (int)acc;
    }
    BDH_HIP(hipStreamSynchronize(b->st));
    BDH_HIP(hipMemcpy(b->d_tab, b->tab.data(), sizeof(int) * b->tab.size(), hipMemcpyHostToDevice));
    BDH_HIP(hipMemcpy(b->d_ch, b->hc.data(), sizeof(BatchChain) * B, hipMemcpyHostToDevice));
    if (b->cycle_exec) {
        BDH_HIP(hipGraphExecDestroy(b->cycle_exec));
        b->cycle_exec = nullptr;
    }
    if (b->factor_exec) {
        BDH_HIP(hipGraphExecDestroy(b->factor_exec));
        b->factor_exec = nullptr;
    }
    b->left_max = left_max;
    b->outer = outer;
    return 0;
}

int batch_set_schedule(bdh_batch* b, int left_max) { return batch_set_tiles(b, left_max, b->outer); }

// Candidate schedules: all right-looking (0), then left-looking for the chains up to each
// of at most BDH_TUNE_MAX - 1 of the batch's distinct padded sizes (evenly spaced in sorted
// order, the largest always included, which makes every chain left-looking).
std::vector<int> batch_tune_candidates(const bdh_batch* b)
{
    std::vector<int> sizes;
    for (const auto& h : b->hc) sizes.push_back(h.n_pad);
    std::sort(sizes.begin(), sizes.end());
    sizes.erase(std::unique(sizes.begin(), sizes.end()), sizes.end());
    std::vector<int> cand{0};
    const int S = (int)sizes.size(), M = std::min(S, BDH_TUNE_MAX - 1);
    for (int i = 1; i <= M; ++i) cand.push_back(sizes[(size_t)((long long)i * S / M) - 1]);
    return cand;
}

// Each candidate (left_max, outer): one refactor to capture its graph, then `reps` timed
// replays (HIP events on the batch's stream); its time is the fastest replay.  Every replay
// recomputes the same matrices from the unchanged positions, so the candidates see identical
// work.  The fastest candidate is kept.
int batch_tune(bdh_batch* b, int reps, const std::vector<std::pair<int, int>>& cand, int* ib_out,
               std::vector<double>& t)
{
    hipEvent_t e0, e1;
    BDH_HIP(hipEventCreate(&e0));
    BDH_HIP(hipEventCreate(&e1));
    int rc = 0, ib = 0;
    t.assign(cand.size(), 0.0);
    for (size_t i = 0; i < cand.size() && !rc; ++i) {
        rc = batch_set_tiles(b, cand[i].first, cand[i].second) || batch_factor(b) || batch_check(b) ? -1 : 0;
        double tmin = 1e30;
        for (int r = 0; r < reps && !rc; ++r) {
            float ms = 0.f;
            const bool ok = hipEventRecord(e0, b->st) == hipSuccess && batch_factor(b) == 0 &&
                            hipEventRecord(e1, b->st) == hipSuccess && hipEventSynchronize(e1) == hipSuccess &&
                            hipEventElapsedTime(&ms, e0, e1) == hipSuccess;
            if (!ok) {
                set_error("bdh_batch_autotune: a timed refactor failed");
                rc = -1;
            }
            tmin = std::min(tmin, (double)ms);
        }
        if (!rc) rc = batch_check(b);
        t[i] = tmin;
        if (t[i] < t[ib]) ib = (int)i;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    if (batch_set_tiles(b, cand[ib].first, cand[ib].second)) return -1;
    b->have_L = false;
    *ib_out = ib;
    return 0;
}

int batch_autotune(bdh_batch* b, int reps, int* cand_out, double* ms_out, int* ncand, int* best)
{
    std::vector<std::pair<int, int>> cand;
    for (int l : batch_tune_candidates(b)) cand.emplace_back(l, b->outer);
    std::vector<double> t;
    int ib = 0;
    if (batch_tune(b, reps, cand, &ib, t)) return -1;
    *ncand = (int)cand.size();
    *best = cand[ib].first;
    for (size_t i = 0; i < cand.size(); ++i) {
        if (cand_out) cand_out[i] = cand[i].first;
        if (ms_out) ms_out[i] = t[i];
    }
    return 0;
}

// The schedules of batch_tune_candidates under each outer panel width 256, 512 and 1024
// (widths no wider than the largest chain's padded order).
int batch_autotune_tiles(bdh_batch* b, int reps, int* cand_left, int* cand_outer, double* ms_out,
                         int* ncand, int* best)
{
    int nmax = 0;
    for (const auto& h : b->hc) nmax = std::max(nmax, h.n_pad);
    std::vector<std::pair<int, int>> cand;
    for (int outer : {256, 512, 1024}) {
        if (outer > 256 && outer > nmax) break;
        for (int l : batch_tune_candidates(b)) cand.emplace_back(l, outer);
    }
    BDH_CHECK((int)cand.size() <= BDH_TUNE_TILES_MAX, "bdh_batch_autotune_tiles: too many candidates");
    std::vector<double> t;
    int ib = 0;
    if (batch_tune(b, reps, cand, &ib, t)) return -1;
    *ncand = (int)cand.size();
    *best = ib;
    for (size_t i = 0; i < cand.size(); ++i) {
        if (cand_left) cand_left[i] = cand[i].first;
        if (cand_outer) cand_outer[i] = cand[i].second;
        if (ms_out) ms_out[i] = t[i];
    }
    return 0;
}

int batch_factor_launches(bdh_batch* b)
{
    BREC_BEGIN(PH_ASSEMBLE);
    if (b->mixed) {
        if (launch_rpy_assemble_batch<float>(b->pos, b->preDii, b->preDij, (float*)b->A, b->plan, b->st))
            return -1;
        BREC_END(PH_ASSEMBLE);
        if (chol_factor_batch<float>((float*)b->A, b->Lbuf, b->plan, b->status, b->syrk_mode, b->st, b->rec))
            return -1;
    } else {
        if (launch_rpy_assemble_batch<double>(b->pos, b->preDii, b->preDij, (double*)b->A, b->plan, b->st))
            return -1;
        BREC_END(PH_ASSEMBLE);
        if (chol_factor_batch<double>((double*)b->A, b->Lbuf, b->plan, b->status, SYRK_F64, b->st, b->rec))
            return -1;
    }
    b->have_L = true;
    return 0;
}

// The single-system schedule of run_steps (no rejections): chunks of up to TRMV_MAX_RHS
// steps that never cross a refactorization; X == nullptr draws the Gaussians of every chain.
// m steps of one refactor cycle (refactor: the cycle starts at its factorization).  Noise:
// the caller's X (x_stride = Gaussians per step), or the chains' device streams (X null).
// direct: the factorization as launches rather than its captured graph.
int batch_cycle(bdh_batch* b, int m, bool refactor, const double* X, long long x_stride, bool direct)
{
    const double dt = b->cfg.dt;
    const double rf_scale = dt / (kB * b->cfg.T);
    if (!X) {
        BREC_BEGIN(PH_NOISE);
        if (launch_gauss_streams(b->rng, b->xi, b->rng_off, b->rng_cnt + (size_t)m * b->B, b->B,
                                 b->flags + 1, b->st))
            return -1;
        BREC_END(PH_NOISE);
        X = b->xi;
        x_stride = 0;
    }
    if (refactor && (direct ? batch_factor_launches(b) : batch_factor(b))) return -1;
    BREC_BEGIN(PH_TRMV);
    if (b->mixed ? launch_trmv_batch<float>((const float*)b->A, X, x_stride, m, b->plan, b->Pt, b->st)
                 : launch_trmv_batch<double>((const double*)b->A, X, x_stride, m, b->plan, b->Pt, b->st))
        return -1;
    if (launch_reduce_batch(b->Pt, 2, m, b->RsB, b->ntot, std::sqrt(2.0 * dt), b->plan, b->st)) return -1;
    BREC_END(PH_TRMV);
    for (int j = 0; j < m; ++j) {
        BREC_BEGIN(PH_MATVEC);
        if (b->mixed ? launch_pair_sums_batch<float>(b->pos, b->F, b->preDii, b->preDij, b->plan, b->Pp, b->st)
                     : launch_pair_sums_batch<double>(b->pos, b->F, b->preDii, b->preDij, b->plan, b->Pp, b->st))
            return -1;
        BREC_END(PH_MATVEC);
        BREC_BEGIN(PH_MOVE);
        if (launch_reduce_batch(b->Pp, 0, 1, b->Rf, 0, rf_scale, b->plan, b->st)) return -1;
        if (launch_em_update(b->pos, b->Ntot, b->Rf, b->RsB + (size_t)j * b->ntot, 0, 0.0, b->st)) return -1;
        BREC_END(PH_MOVE);
        if (batch_forces(b)) return -1;
    }
    return 0;
}

// Steps of every chain.  A whole refactor cycle on the device streams is one graph launch:
// its ~(50 + 5 R) launches (the chains' Gaussians, assembly and factorization, L.xi of the
// cycle, then per step D.F, move and forces) are captured on first use and replayed; every
// argument (pointers, grid sizes, the stream counters kept in device memory) is the same from
// one cycle to the next.  Partial cycles, given Gaussians and phase timing run the launches.
int batch_run(bdh_batch* b, int steps, const double* X_user)
{
    const int R = b->cfg.refactor_every;
    for (int k = 0; k < steps;) {
        const int ph = (int)(b->steps_done % R);
        const bool refactor = !b->have_L || ph == 0;
        const int m = std::min(std::min(steps - k, TRMV_MAX_RHS), refactor ? R : R - ph);
        if (!X_user && BDH_FACTOR_GRAPH && !b->rec && refactor && m == R) {
            if (!b->cycle_exec) {
                if (gauss_prepare()) return -1;   // its table upload cannot be captured
                BDH_HIP(hipStreamBeginCapture(b->st, hipStreamCaptureModeThreadLocal));
                const int rc = batch_cycle(b, R, true, nullptr, 0, true);
                hipGraph_t g = nullptr;
                const hipError_t e = hipStreamEndCapture(b->st, &g);
                if (rc) {
                    if (g) (void)hipGraphDestroy(g);
                    return rc;
                }
                BDH_HIP(e);
                const hipError_t ei = hipGraphInstantiate(&b->cycle_exec, g, nullptr, nullptr, 0);
                (void)hipGraphDestroy(g);
                if (ei != hipSuccess) b->cycle_exec = nullptr;
                BDH_HIP(ei);
            }
            BDH_HIP(hipGraphLaunch(b->cycle_exec, b->st));
            b->have_L = true;
        } else {
            const double* X = X_user ? X_user + (size_t)k * b->ntot : nullptr;
            if (batch_cycle(b, m, refactor, X, X_user ? b->ntot : 0, false)) return -1;
        }
        b->steps_done += m;
        k += m;
    }
    return 0;
}

int batch_create_impl(const bdh_config* cfg, int B, const int* N, const double* coord, const double* Q,
                      const double* LJ, const int* nbonds, const int* bonds, const double* bond_parms,
                      const uint32_t* seeds, bdh_batch* b)
{
    BDH_CHECK(cfg && N && coord && Q && LJ && nbonds && seeds, "bdh_batch_create: null argument");
    BDH_CHECK(B > 0, "bdh_batch_create: no chains");
    BDH_CHECK(cfg->hydro == BDH_HYDRO_CHOLESKY, "bdh_batch_create: hydro must be cholesky");
    BDH_CHECK(cfg->bc == 0, "bdh_batch_create: free space only (bc none)");
    BDH_CHECK(cfg->algorithm == BDH_ERMAK_CONST || cfg->algorithm == BDH_ERMAK_VAR,
              "bdh_batch_create: algorithm ermak_const");
    BDH_CHECK(!cfg->check_overlap, "bdh_batch_create: check_overlap is not supported in a batch");
    BDH_CHECK(cfg->nb_list == 0, "bdh_batch_create: brute-force pairs only (nb_list brute)");
    BDH_CHECK(cfg->precision == BDH_FP64 || cfg->precision == BDH_MIXED || cfg->precision == BDH_MIXED_SPLIT,
              "bdh_batch_create: precision");
    BDH_CHECK(cfg->dt > 0 && cfg->T > 0 && cfg->visc != 0 && cfg->vfactor != 0, "bdh_batch_create: dt/T/visc");
    b->cfg = *cfg;
    if (b->cfg.refactor_every < 1) b->cfg.refactor_every = 1;
    BDH_HIP(hipGetDevice(&b->device));
    BDH_HIP(hipStreamCreateWithFlags(&b->st, hipStreamNonBlocking));
    b->B = B;
    b->mixed = cfg->precision != BDH_FP64;
    b->syrk_mode = cfg->precision == BDH_MIXED_SPLIT ? SYRK_BF16X3 : (b->mixed ? SYRK_F32 : SYRK_F64);
    b->preDii = kB_1_6PI * cfg->T / (cfg->visc * cfg->vfactor);
    b->preDij = kB_1_8PI * cfg->T / (cfg->visc * cfg->vfactor);

    // layout
    b->hc.resize(B);
    long long bead = 0, mat = 0, lb = 0, xo = 0, pp = 0, pf = 0, pt = 0, bonds_tot = 0;
    int max_steps = 0;
    b->outer = chol_outer_default();
    for (int c = 0; c < B; ++c) {
        BDH_CHECK(N[c] > 0 && N[c] <= (1 << 20), "bdh_batch_create: chain size");
        BDH_CHECK(nbonds[c] >= 0 && (nbonds[c] == 0 || (bonds && bond_parms)), "bdh_batch_create: bonds");
        BatchChain& h = b->hc[c];
        h.N = N[c];
        h.n = 3 * N[c];
        h.n_pad = ceil_div(h.n, PAD_TO) * PAD_TO;
        h.bead0 = (int)bead;
        h.psplits = pair_splits(h.N, false);
        h.fsplits = force_splits(h.N);
        h.tcols = trmv_col_blocks(h.n);
        h.ldx = h.n + (h.n & 1);
        h.left = chol_left_looking(h.n_pad);
        h.outer = chol_outer_default();
        h.mat = mat;
        h.lb = lb;
        h.xi_off = xo;
        h.pp = pp;
        h.pf = pf;
        h.pt = pt;
        bead += h.N;
        mat += (long long)h.n_pad * h.n_pad;
        lb += ((long long)chol_lbuf_doubles(h.n_pad) + 1) / 2 * 2;   // 16-byte aligned workspaces
        xo += (long long)TRMV_MAX_RHS * h.ldx;
        pp += (long long)h.psplits * h.n;
        pf += (long long)h.fsplits * h.n;
        pt += (long long)h.tcols * TRMV_MAX_RHS * h.n;
        bonds_tot += nbonds[c];
        max_steps = std::max(max_steps, chol_update_steps(h.n_pad));
    }
    BDH_CHECK(bead <= (1 << 26), "bdh_batch_create: too many beads");
    b->Ntot = (int)bead;
    b->ntot = 3 * b->Ntot;
    for (long long i = 0; i < bead; ++i)
        BDH_CHECK(coord[4 * i + 3] > 0.0, "bdh_batch_create: bead radius must be positive");

    // prefix tables: asm, pair, force, trmv, rows, trsm, then max_steps chol tables
    const int W = B + 1;
    const int ntab = 6 + max_steps;
    std::vector<int> tab((size_t)ntab * W, 0);
    b->chol_blocks.assign(max_steps, 0);
    auto fill = [&](int t, auto count) -> int {
        long long acc = 0;
        for (int c = 0; c < B; ++c) {
            tab[(size_t)t * W + c] = (int)acc;
            acc += count(b->hc[c]);
        }
        tab[(size_t)t * W + B] = (int)acc;
        return acc <= (long long)INT32_MAX ? (int)acc : -1;
    };
    b->plan.asm_blocks = fill(0, [](const BatchChain& h) { return rpy_blocks(h.N); });
    b->plan.pair_blocks = fill(1, [](const BatchChain& h) { return pair_eparts(h.N); });
    b->plan.force_blocks = fill(2, [](const BatchChain& h) { return force_blocks(h.N); });
    b->plan.trmv_blocks = b->mixed ? fill(3, [](const BatchChain& h) { return trmv_blocks<float>(h.n); })
                                   : fill(3, [](const BatchChain& h) { return trmv_blocks<double>(h.n); });
    b->plan.rows_blocks = fill(4, [](const BatchChain& h) { return ceil_div(h.n, 64); });
    b->plan.trsm_blocks = fill(5, [](const BatchChain& h) { return chol_trsm0_blocks(h.n_pad); });
    for (int k = 0; k < max_steps; ++k)
        b->chol_blocks[k] = fill(6 + k, [k](const BatchChain& h) { return chol_update_blocks(h.n_pad, h.left, h.outer, k); });
    BDH_CHECK(b->plan.asm_blocks >= 0 && b->plan.pair_blocks >= 0 && b->plan.force_blocks >= 0 &&
              b->plan.trmv_blocks >= 0 && b->plan.rows_blocks >= 0 && b->plan.trsm_blocks >= 0,
              "bdh_batch_create: batch too large");
    for (int k = 0; k < max_steps; ++k)
        BDH_CHECK(b->chol_blocks[k] >= 0, "bdh_batch_create: batch too large");
    if (dmalloc(&b->d_tab, sizeof(int) * tab.size())) return -1;
    BDH_HIP(hipMemcpy(b->d_tab, tab.data(), sizeof(int) * tab.size(), hipMemcpyHostToDevice));
    if (dmalloc(&b->d_ch, sizeof(BatchChain) * B)) return -1;
    BDH_HIP(hipMemcpy(b->d_ch, b->hc.data(), sizeof(BatchChain) * B, hipMemcpyHostToDevice));
    b->tab = std::move(tab);
    b->plan.nch = B;
    b->plan.ch = b->d_ch;
    b->plan.asm_pre = b->d_tab;
    b->plan.pair_pre = b->d_tab + W;
    b->plan.force_pre = b->d_tab + 2 * W;
    b->plan.trmv_pre = b->d_tab + 3 * W;
    b->plan.rows_pre = b->d_tab + 4 * W;
    b->plan.trsm_pre = b->d_tab + 5 * W;
    b->plan.chol_pre = b->d_tab + 6 * W;
    b->plan.chol_steps = max_steps;
    b->plan.chol_blocks = b->chol_blocks.data();

    // forces (a single set of parameters; the pair-skip reach is the largest of any chain)
    ForceParams& fp = b->fp;
    fp.alpha_lj = cfg->alpha_lj;
    fp.cutoff_lj = cfg->cutoff_lj;
    fp.lj_6_term = cfg->lj_6_term ? 1.0 : 0.0;
    fp.elec = cfg->elec;
    fp.gam = cfg->gamma_c / cfg->epsilon_c;
    fp.kappa_c = cfg->kappa_c;
    fp.cutoff_c = cfg->cutoff_c;
    fp.bond_lj_scale = cfg->bond_lj_scale;
    fp.bond_c_scale = cfg->bond_c_scale;
    fp.use_conn = bonds_tot && ((cfg->elec && cfg->bond_c_scale != 1) ||
                                (cfg->alpha_lj != 0.0 && cfg->bond_lj_scale != 1));
    fp.check_overlap = 0;
    fp.pbc = 0;
    fp.box = 0.0;
    {
        double rmax = 0.0;
        for (long long i = 0; i < bead; ++i) rmax = std::max(rmax, LJ[2 * i]);
        double reach = 2.0 * rmax;
        if (cfg->elec) reach = std::max(reach, cfg->cutoff_c);
        if (cfg->alpha_lj != 0.0)
            reach = std::max(reach, cfg->cutoff_lj > 0 ? cfg->cutoff_lj : 1.122462 * 2.0 * rmax);
        reach = reach * (1.0 + 1e-5) + 1e-5;
        fp.far_r2 = reach * reach;
    }

    // bonds: per chain a CSR of N + 1 offsets (global positions in the concatenated lists)
    // naming chain-local partners
    std::vector<int> coff(b->Ntot + B), partner(std::max<long long>(1, 2 * bonds_tot));
    std::vector<double> parm(3 * std::max<long long>(1, 2 * bonds_tot));
    {
        long long base = 0, bo = 0;
        for (int c = 0; c < B; ++c) {
            const BatchChain& h = b->hc[c];
            const int* bd = bonds + 2 * bo;
            const double* bp = bond_parms + 3 * bo;
            std::vector<int> deg(h.N, 0);
            for (int q = 0; q < nbonds[c]; ++q) {
                const int i = bd[2 * q], j = bd[2 * q + 1];
                BDH_CHECK(i >= 0 && i < h.N && j >= 0 && j < h.N && i != j,
                          "bdh_batch_create: bond index out of range");
                deg[i]++;
                deg[j]++;
            }
            int* off = coff.data() + h.bead0 + c;
            off[0] = (int)base;
            for (int i = 0; i < h.N; ++i) off[i + 1] = off[i] + deg[i];
            std::vector<int> fillp(off, off + h.N);
            for (int q = 0; q < nbonds[c]; ++q) {
                const int i = bd[2 * q], j = bd[2 * q + 1];
                for (int e = 0; e < 2; ++e) {
                    const int me = e ? j : i, other = e ? i : j;
                    const int at = fillp[me]++;
                    partner[at] = other;
                    parm[3 * at] = bp[3 * q];
                    parm[3 * at + 1] = bp[3 * q + 1];
                    parm[3 * at + 2] = bp[3 * q + 2];
                }
            }
            base = off[h.N];
            bo += nbonds[c];
        }
    }
    if (dmalloc(&b->conn_off, sizeof(int) * coff.size())) return -1;
    if (dmalloc(&b->conn, sizeof(int) * partner.size())) return -1;
    if (dmalloc(&b->bond_parm, sizeof(double) * parm.size())) return -1;
    BDH_HIP(hipMemcpy(b->conn_off, coff.data(), sizeof(int) * coff.size(), hipMemcpyHostToDevice));
    BDH_HIP(hipMemcpy(b->conn, partner.data(), sizeof(int) * partner.size(), hipMemcpyHostToDevice));
    BDH_HIP(hipMemcpy(b->bond_parm, parm.data(), sizeof(double) * parm.size(), hipMemcpyHostToDevice));
    // the bonded-partner table and the bond list are the same CSR
    b->topo = Topology{b->conn_off, b->conn, b->conn_off, b->conn, b->bond_parm};

    // beads and vectors
    std::vector<double2> lj(b->Ntot);
    for (int i = 0; i < b->Ntot; ++i) lj[i] = make_double2(LJ[2 * i], LJ[2 * i + 1]);
    if (dmalloc(&b->pos, sizeof(double4) * b->Ntot)) return -1;
    if (dmalloc(&b->Q, sizeof(double) * b->Ntot)) return -1;
    if (dmalloc(&b->LJ, sizeof(double2) * b->Ntot)) return -1;
    BDH_HIP(hipMemcpy(b->pos, coord, sizeof(double4) * b->Ntot, hipMemcpyHostToDevice));
    BDH_HIP(hipMemcpy(b->Q, Q, sizeof(double) * b->Ntot, hipMemcpyHostToDevice));
    BDH_HIP(hipMemcpy(b->LJ, lj.data(), sizeof(double2) * b->Ntot, hipMemcpyHostToDevice));
    if (dmalloc(&b->F, sizeof(double) * b->ntot)) return -1;
    if (dmalloc(&b->Rf, sizeof(double) * b->ntot)) return -1;
    if (dmalloc(&b->RsB, sizeof(double) * b->ntot * TRMV_MAX_RHS)) return -1;
    if (dmalloc(&b->xi, sizeof(double) * xo)) return -1;
    if (dmalloc(&b->Pp, sizeof(double) * pp)) return -1;
    if (dmalloc(&b->Pf, sizeof(double) * pf)) return -1;
    if (dmalloc(&b->Pt, sizeof(double) * pt)) return -1;
    const size_t esz = b->mixed ? sizeof(float) : sizeof(double);
    if (dmalloc(&b->A, esz * mat)) return -1;
    BDH_HIP(hipMemsetAsync(b->A, 0, esz * mat, b->st));
    for (int c = 0; c < B; ++c) {
        const BatchChain& h = b->hc[c];
        const int rc = b->mixed ? launch_set_identity_pad<float>((float*)b->A + h.mat, h.n_pad, h.n, h.n_pad, b->st)
                                : launch_set_identity_pad<double>((double*)b->A + h.mat, h.n_pad, h.n, h.n_pad, b->st);
        if (rc) return -1;
    }
    if (dmalloc(&b->Lbuf, sizeof(double) * lb)) return -1;
    BDH_HIP(hipMemsetAsync(b->Lbuf, 0, sizeof(double) * lb, b->st));

    // Gaussian streams
    std::vector<uint64_t> rs(B);
    std::vector<long long> ro(B);
    std::vector<int> cnt((size_t)(TRMV_MAX_RHS + 1) * B);
    for (int c = 0; c < B; ++c) {
        rs[c] = lcg_seed(seeds[c]);
        ro[c] = b->hc[c].xi_off;
        for (int m = 0; m <= TRMV_MAX_RHS; ++m) cnt[(size_t)m * B + c] = m * b->hc[c].ldx;
    }
    if (dmalloc(&b->rng, sizeof(uint64_t) * B)) return -1;
    if (dmalloc(&b->rng_off, sizeof(long long) * B)) return -1;
    if (dmalloc(&b->rng_cnt, sizeof(int) * cnt.size())) return -1;
    BDH_HIP(hipMemcpy(b->rng, rs.data(), sizeof(uint64_t) * B, hipMemcpyHostToDevice));
    BDH_HIP(hipMemcpy(b->rng_off, ro.data(), sizeof(long long) * B, hipMemcpyHostToDevice));
    BDH_HIP(hipMemcpy(b->rng_cnt, cnt.data(), sizeof(int) * cnt.size(), hipMemcpyHostToDevice));
    if (dmalloc(&b->flags, sizeof(int) * 4)) return -1;
    if (dmalloc(&b->status, sizeof(int) * B)) return -1;
    BDH_HIP(hipMemsetAsync(b->flags, 0, sizeof(int) * 4, b->st));
    BDH_HIP(hipMemsetAsync(b->status, 0, sizeof(int) * B, b->st));

    // initial forces (main_loop.c:91)
    if (batch_forces(b)) return -1;
    return batch_check(b);
}

}  // namespace

extern "C" {

void bdh_batch_destroy(bdh_batch* b)
{
    if (!b) return;
    if (b->st) (void)hipStreamSynchronize(b->st);
    dfree(b->d_ch); dfree(b->d_tab);
    dfree(b->pos); dfree(b->Q); dfree(b->LJ);
    dfree(b->conn_off); dfree(b->conn); dfree(b->bond_parm);
    dfree(b->F); dfree(b->Rf); dfree(b->RsB); dfree(b->xi); dfree(b->xi_user);
    dfree(b->Pp); dfree(b->Pf); dfree(b->Pt); dfree(b->A); dfree(b->Lbuf);
    dfree(b->rng); dfree(b->rng_off); dfree(b->rng_cnt); dfree(b->flags); dfree(b->status);
    delete b->rec;
    if (b->factor_exec) (void)hipGraphExecDestroy(b->factor_exec);
    if (b->cycle_exec) (void)hipGraphExecDestroy(b->cycle_exec);
    if (b->st) (void)hipStreamDestroy(b->st);
    delete b;
}

int bdh_batch_create(const bdh_config* cfg, int nchains, const int* N, const double* coord,
                     const double* Q, const double* LJ, const int* nbonds, const int* bonds,
                     const double* bond_parms, const uint32_t* seeds, bdh_batch** out)
{
    BDH_CHECK(out, "bdh_batch_create: null output");
    *out = nullptr;
    bdh_batch* b = new bdh_batch();
    const int rc = batch_create_impl(cfg, nchains, N, coord, Q, LJ, nbonds, bonds, bond_parms, seeds, b);
    if (rc) {
        const std::string msg = last_error();
        bdh_batch_destroy(b);
        set_error(msg);
        return rc;
    }
    *out = b;
    return 0;
}

int bdh_batch_step(bdh_batch* b, int steps, const double* xi)
{
    BDH_CHECK(b && steps >= 0, "bdh_batch_step: bad arguments");
    if (steps == 0) return 0;
    const double* X = nullptr;
    if (xi) {
        const size_t need = (size_t)steps * b->ntot;
        if (need > b->xi_user_cap) {
            dfree(b->xi_user);
            if (dmalloc(&b->xi_user, sizeof(double) * need)) return -1;
            b->xi_user_cap = need;
        }
        BDH_HIP(hipMemcpyAsync(b->xi_user, xi, sizeof(double) * need, hipMemcpyHostToDevice, b->st));
        X = b->xi_user;
    }
    if (batch_run(b, steps, X)) return -1;
    return batch_check(b);
}

int bdh_batch_set_schedule(bdh_batch* b, int left_max)
{
    BDH_CHECK(b, "bdh_batch_set_schedule: null batch");
    return batch_set_schedule(b, left_max);
}

int bdh_batch_set_tiles(bdh_batch* b, int left_max, int outer)
{
    BDH_CHECK(b, "bdh_batch_set_tiles: null batch");
    return batch_set_tiles(b, left_max, outer);
}

int bdh_batch_get_tiles(bdh_batch* b, int* left_max, int* outer)
{
    BDH_CHECK(b && left_max && outer, "bdh_batch_get_tiles: bad arguments");
    *left_max = b->left_max;
    *outer = b->outer;
    return 0;
}

int bdh_batch_autotune_tiles(bdh_batch* b, int reps, int* cand_left, int* cand_outer, double* ms,
                             int* ncand, int* best)
{
    BDH_CHECK(b && reps > 0 && ncand && best, "bdh_batch_autotune_tiles: bad arguments");
    BDH_CHECK(b->steps_done == 0, "bdh_batch_autotune_tiles: call before the first step (the tuning "
                                  "refactors at the current positions)");
    return batch_autotune_tiles(b, reps, cand_left, cand_outer, ms, ncand, best);
}

int bdh_batch_get_schedule(bdh_batch* b, int* left_max)
{
    BDH_CHECK(b && left_max, "bdh_batch_get_schedule: bad arguments");
    *left_max = b->left_max;
    return 0;
}

int bdh_batch_autotune(bdh_batch* b, int reps, int* cand, double* ms, int* ncand, int* best)
{
    BDH_CHECK(b && reps > 0 && ncand && best, "bdh_batch_autotune: bad arguments");
    BDH_CHECK(b->steps_done == 0, "bdh_batch_autotune: call before the first step (the tuning "
                                  "refactors at the current positions)");
    return batch_autotune(b, reps, cand, ms, ncand, best);
}

int bdh_batch_get_coords(bdh_batch* b, double* coord)
{
    BDH_CHECK(b && coord, "bdh_batch_get_coords: bad arguments");
    BDH_HIP(hipMemcpyAsync(coord, b->pos, sizeof(double4) * b->Ntot, hipMemcpyDeviceToHost, b->st));
    BDH_HIP(hipStreamSynchronize(b->st));
    return 0;
}

int bdh_batch_get_forces(bdh_batch* b, double* F)
{
    BDH_CHECK(b && F, "bdh_batch_get_forces: bad arguments");
    BDH_HIP(hipMemcpyAsync(F, b->F, sizeof(double) * b->ntot, hipMemcpyDeviceToHost, b->st));
    BDH_HIP(hipStreamSynchronize(b->st));
    return 0;
}

int bdh_batch_time_steps(bdh_batch* b, int steps, double* ms)
{
    BDH_CHECK(b && steps > 0 && ms, "bdh_batch_time_steps: bad arguments");
    hipEvent_t e0, e1;
    BDH_HIP(hipEventCreate(&e0));
    BDH_HIP(hipEventCreate(&e1));
    BDH_HIP(hipEventRecord(e0, b->st));
    if (batch_run(b, steps, nullptr)) return -1;
    BDH_HIP(hipEventRecord(e1, b->st));
    BDH_HIP(hipEventSynchronize(e1));
    float t = 0;
    BDH_HIP(hipEventElapsedTime(&t, e0, e1));
    *ms = t;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return batch_check(b);
}

int bdh_batch_phase_timing(bdh_batch* b, int enable)
{
    BDH_CHECK(b, "bdh_batch_phase_timing: null batch");
    BDH_HIP(hipStreamSynchronize(b->st));
    delete b->rec;
    b->rec = enable ? new EventRecorder() : nullptr;
    return 0;
}

int bdh_batch_phase_times(bdh_batch* b, double* ms, long long* counts)
{
    BDH_CHECK(b && b->rec && ms && counts, "bdh_batch_phase_times: timing not enabled");
    BDH_HIP(hipStreamSynchronize(b->st));
    for (int p = 0; p < PH_COUNT; ++p) { ms[p] = 0.0; counts[p] = 0; }
    for (auto& r : b->rec->recs) {
        float t = 0.f;
        BDH_HIP(hipEventElapsedTime(&t, r.e0, r.e1));
        ms[r.phase] += t;
        counts[r.phase] += 1;
    }
    b->rec->clear();
    return 0;
}

}  // extern "C"
