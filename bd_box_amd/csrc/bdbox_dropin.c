/*
 * Drop-in backend for the unmodified BD_BOX driver.
 *
 * Link-time substitute for the translation units the reference swaps per backend
 * (src/Makefile.am:40-70): diff_alg/diff_decl.c, diff_tensor.c, diff_equation.c,
 * diff_rand.c, diff_comp.c, potentials/calc_func.c and myblas.c.  It defines exactly the
 * symbols the remaining driver TUs (main.c, main_loop.c, input.c, output.c, restart.c,
 * data.c, trans.c, ...) take from them -- the entry points of diff_tensor.h:166-253 and
 * calc_func.h:60-106 and the registry variables of the .prm keywords these TUs own
 * (declared with the reference's own MAKE_STR_IN, diff_decl.c:66-81, calc_func.c:61-63) --
 * and forwards every call to the HIP backend through include/bdbox_hip.h.
 *
 * Compiled against the reference's headers (it is the binding a BD_BOX maintainer adds);
 * see oracle/Makefile target `dropin` for the link line.
 *
 * Supported: algorithm ermak_const / ermak_var (with the reference's rejection of steps that
 * overlap when check_overlap is set, diff_comp.c:152-179) and the IG-T predictor-corrector
 * igt_const / igt_var (diff_comp.c:76-144; hydro none or cholesky), hydro none / cholesky /
 * geyer, bc none or periodic (cube; Smith- or Beenakker-Ewald tensor with hydrodynamics), nb_list
 * brute or either bucket method (both map to the backend's device cell list, which finds
 * the same interacting pairs), FP64 or mixed precision
 * (environment BDBOX_HIP_PRECISION=fp64|mixed|split, default fp64), refactor interval
 * (BDBOX_HIP_REFACTOR, default 1 = the reference's semantics).  Anything else stops with
 * UNERR like the reference does for unimplemented methods (diff_comp.c:230).
 */
#if HAVE_CONFIG_H
#include <config.h>
#endif

#include <stdlib.h>
#include <string.h>

#include "data.h"
#include "err.h"
#include "input.h"
#include "main_loop.h"
#include "myblas.h"
#include "output.h"
#include "rand_move.h"
#include "trans.h"
#include "diff_alg/diff_tensor.h"
#include "potentials/calc_func.h"
#include "potentials/bond.h"
#include "potentials/electro.h"
#include "potentials/LJ.h"
#include "potentials/bucket.h"
#include "diff_alg/newton.h"

#include "../../include/bdbox_hip.h"

/* ---- registry variables and globals owned by the replaced TUs ---- */
DOUBLE* D = NULL;
DOUBLE* Ds = NULL;
DOUBLE* diag_D = NULL;
DOUBLE* inv_diag_D = NULL;
INT ewald = 0;
DOUBLE preDij;
DOUBLE preDii;
INT save_iter = -1;
INT curr_iter = 0;
DOUBLE* E;
/* work vectors of diff_tensor.h:126-145; the backend keeps its own on the device, these hold
   the last move's Rf / Rs (Rf[0], Rs[0]) for callers of compute_hydro_ermak / compute_r */
DOUBLE* X = NULL;
DOUBLE** Rf = NULL;
DOUBLE** Rs = NULL;
DOUBLE* epsilon = NULL;
DOUBLE* FS = NULL;
DOUBLE* Rf0 = NULL;
DOUBLE* S[2] = {NULL, NULL};

MAKE_STR_IN( DOUBLE, visc, 0.01, "viscovity" )
MAKE_STR_IN( DOUBLE, vfactor, 14.4, "factor for viscovity units conversion" )
MAKE_STR_IN( DOUBLE, curr_dt, 0, "inner value" )
MAKE_STR_IN( NO_CHOLS_GEYER, hydro, 0, "method used to evaluate hydrodynamic interactions" )
MAKE_STR_IN( ERMAK_IGTCONST_IGTVAR, algorithm, 2, "integration algorithm" )
MAKE_STR_IN( INT, move_attempts, 10000, "number of attempts to draw a new random vector in case of an overlap" )
MAKE_STR_IN( PTR_DOUBLE9, vel_grad_tensor, NULL, "the velocity gradient tensor matrix (row,column)" )
MAKE_STR_IN( DOUBLE, e_collision, 1, "restitution parameter used in the elastic collision method" )
MAKE_STR_IN( INT, ewald_real, 10, "magnitude of the real lattice vectors" )
MAKE_STR_IN( INT, ewald_recip, 10, "magnitude of the reciprocal lattice vectors" )
MAKE_STR_IN( DOUBLE, ewald_alpha, 1.77245385090551602729, "Ewald summation convergence control" )
MAKE_STR_IN( EWALD_METHOD, ewald_method, 0, "method used to built the Ewald-summed tensor" )
MAKE_STR_IN( YESNO, geyer_on_the_fly, 1, "whether to use 3x3 submatrices of D (yes) or the whole D during evaluation of HI (no) with TEA-HI" )
MAKE_STR_IN( DOUBLE, bond_lj_scale, 1.0, "scaling factor for Lennard-Jones interactions between bonded partners 1-2" )
MAKE_STR_IN( DOUBLE, bond_c_scale, 1.0, "scaling factor for electrostatic interactions between bonded partners 1-2" )
MAKE_STR_IN( YESNO, check_overlap, 1, "whether to detect overlaps after each step" )

static bdh_system* g_sys = NULL;

static void fail(const char* what)
{
    char buff[1024];
    snprintf(buff, sizeof(buff), "HIP backend: %s: %s", what, bdh_last_error());
    UNERR(buff);
}

/* 48-bit drand48 state inside glibc's struct drand48_data (x[0..2], little end first);
   read and written through the reference's own get_state/set_state (rand_move.c:170-198) */
static uint64_t host_rng_state(void)
{
    char* st;
    INT sz;
    unsigned short x[3];
    get_state(&st, &sz);
    memcpy(x, st, sizeof(x));
    return (uint64_t)x[0] | ((uint64_t)x[1] << 16) | ((uint64_t)x[2] << 32);
}

static void set_host_rng_state(uint64_t s)
{
    char* st;
    INT sz;
    unsigned short x[3];
    get_state(&st, &sz);
    x[0] = (unsigned short)(s & 0xFFFF);
    x[1] = (unsigned short)((s >> 16) & 0xFFFF);
    x[2] = (unsigned short)((s >> 32) & 0xFFFF);
    memcpy(st, x, sizeof(x));
}

void init_diff_tensor(void)
{
    bdh_config cfg;
    const char* prec = getenv("BDBOX_HIP_PRECISION");
    const char* refac = getenv("BDBOX_HIP_REFACTOR");
    int i, rc;
    if (algorithm != DIFF_ALG_ERMAK_CONST && algorithm != DIFF_ALG_ERMAK_VAR &&
        algorithm != DIFF_ALG_IGT_CONST && algorithm != DIFF_ALG_IGT_VAR)
        UNERR("HIP backend: algorithm ermak_const / ermak_var / igt_const / igt_var");
    if (hydro != DIFF_NONE && hydro != DIFF_CHOLESKY && hydro != DIFF_GEYER)
        UNERR("HIP backend: hydro none / cholesky / geyer");
    if (bc != BC_NONE && bc != BC_BOX) UNERR("HIP backend: bc none or periodic");
    if (bc == BC_BOX && (box[0] != box[1] || box[0] != box[2]))
        UNERR("HIP backend: periodic box must be a cube");
    /* The reference's on-the-fly TEA path with an Ewald tensor (diff_rand.c:100-130) keeps
       one tD per bead i across all its partners k, and the Ewald block functions add into
       it (diff_tensor.c:265-400), so its C, Feff and eps are not the tensor's sums.  This
       backend always forms the sums of the tensor itself, which is what the reference's
       stored-D path (geyer_on_the_fly no, diff_rand.c:136-164) computes. */
    if (hydro == DIFF_GEYER && bc == BC_BOX && geyer_on_the_fly)
        warning("HIP backend: geyer_on_the_fly with an Ewald tensor gives the TEA sums of the "
                "tensor (those of geyer_on_the_fly no); the reference's on-the-fly Ewald path "
                "accumulates the blocks of a bead into one another, so trajectories differ",
                __FILE__, __LINE__);
    memset(&cfg, 0, sizeof(cfg));
    cfg.dt = dt; cfg.T = T; cfg.visc = visc; cfg.vfactor = vfactor;
    cfg.alpha_lj = alpha_lj; cfg.cutoff_lj = cutoff_lj; cfg.lj_6_term = lj_6_term;
    cfg.elec = elec; cfg.epsilon_c = epsilon_c; cfg.kappa_c = kappa_c; cfg.gamma_c = gamma_c;
    cfg.cutoff_c = cutoff_c; cfg.bond_lj_scale = bond_lj_scale; cfg.bond_c_scale = bond_c_scale;
    cfg.check_overlap = check_overlap; cfg.bc = bc == BC_BOX; cfg.box = box[0];
    cfg.nb_list = nb_list;
    cfg.algorithm = algorithm; cfg.move_attempts = move_attempts;
    cfg.ewald_alpha = ewald_alpha;
    cfg.hydro = hydro; cfg.ewald_method = ewald_method; cfg.ewald_real = ewald_real;
    cfg.precision = !prec || !strcmp(prec, "fp64") ? BDH_FP64
                  : !strcmp(prec, "mixed")            ? BDH_MIXED
                  : !strcmp(prec, "split")            ? BDH_MIXED_SPLIT
                                                      : -1;
    if (cfg.precision < 0) UNERR("BDBOX_HIP_PRECISION: fp64, mixed or split");
    cfg.refactor_every = refac ? atoi(refac) : 1;
    cfg.rand_seed = (uint32_t)rand_seed;
    if (bdh_init(0)) fail("init");
    rc = bdh_create(&cfg, size, coord, Q, LJ, bond_size, bond_pairs, bond_parms, &g_sys);
    if (rc) fail("create");
    /* continue the host stream where the driver left it (init_rand in main.c:89) */
    if (bdh_set_rng_state(g_sys, host_rng_state())) fail("rng state");
    /* a flow field, vel_grad_tensor (diff_equation.c:265-285) */
    if (vel_grad_tensor && bdh_set_flow(g_sys, vel_grad_tensor)) fail("vel_grad_tensor");
    /* diag_D as init_diff_tensor builds it (diff_decl.c:93-98) */
    preDii = kB_1_6PI * T / (visc * vfactor);
    preDij = kB_1_8PI * T / (visc * vfactor);
    Rf = (DOUBLE**)malloc(sizeof(DOUBLE*)); CHMEM(Rf);
    Rs = (DOUBLE**)malloc(sizeof(DOUBLE*)); CHMEM(Rs);
    Rf[0] = (DOUBLE*)calloc(size * DIMS0, sizeof(DOUBLE)); CHMEM(Rf[0]);
    Rs[0] = (DOUBLE*)calloc(size * DIMS0, sizeof(DOUBLE)); CHMEM(Rs[0]);
    Rf0 = (DOUBLE*)calloc(size * DIMS1, sizeof(DOUBLE)); CHMEM(Rf0);
    init_tensor();
    diag_D = (DOUBLE*)malloc(sizeof(DOUBLE) * size * DIMS0); CHMEM(diag_D);
    inv_diag_D = (DOUBLE*)malloc(sizeof(DOUBLE) * size * DIMS0); CHMEM(inv_diag_D);
    for (i = 0; i < size * DIMS0; ++i)
    {
        diag_D[i] = preDii / coord[DIMS1 * (i / DIMS0) + 3];
        inv_diag_D[i] = 1 / diag_D[i];
    }
}

void free_diff_tensor(void)
{
    bdh_destroy(g_sys);
    g_sys = NULL;
    free(diag_D); diag_D = NULL;
    free(inv_diag_D); inv_diag_D = NULL;
    free(Ds); Ds = NULL;
    if (Rf) { free(Rf[0]); free(Rf); Rf = NULL; }
    if (Rs) { free(Rs[0]); free(Rs); Rs = NULL; }
    free(Rf0); Rf0 = NULL;
}

void init_forces(void)
{
    init_cutoff();
    init_Fs();
}

void free_forces(void)
{
    free_Fs();
    free(E);
}

/* compute_forces (calc_func.c:106-186): forces and energies at the current positions into
   F[0] / _E; returns the reference's repeat flag */
INT compute_forces(DOUBLE* _E, DOUBLE t)
{
    double Eb[3];
    int rep = 0;
    (void)t;
    if (bdh_set_coords(g_sys, coord)) fail("set_coords");
    if (bdh_compute_forces(g_sys, F[0], _E ? Eb : NULL, &rep)) fail("compute_forces");
    if (_E)
    {
        memset(_E, 0, sizeof(DOUBLE) * (ENERGY_LAST + 1));
        _E[ENERGY_LJ] = Eb[0];
        _E[ENERGY_COULOMB] = Eb[1];
        _E[ENERGY_BOND] = Eb[2];
    }
    return rep;
}

/* save_curr_coord / save_curr_forces / load_coord / load_forces (diff_decl.c:219-237) on
   the driver's save_coord / save_forces (data.c:57-58); host coord and F[0] are current after
   every compute_hydro, and a load puts the backend back at the restored state */
void save_curr_coord(void)
{
    memcpy(save_coord, coord, sizeof(DOUBLE) * DIMS1 * size);
}

void save_curr_forces(void)
{
    memcpy(save_forces, F[0], sizeof(DOUBLE) * DIMS0 * size);
}

void load_coord(void)
{
    memcpy(coord, save_coord, sizeof(DOUBLE) * DIMS1 * size);
    if (bdh_set_coords(g_sys, coord)) fail("load_coord");
}

void load_forces(void)
{
    memcpy(F[0], save_forces, sizeof(DOUBLE) * DIMS0 * size);
    if (bdh_set_forces(g_sys, F[0])) fail("load_forces");
}

static int output_step(INT s)
{
    return (save_xyz_freq && s % save_xyz_freq == 0) || (save_dcd_freq && s % save_dcd_freq == 0) ||
           (save_rst_freq && s % save_rst_freq == 0) || (save_enr_freq && s % save_enr_freq == 0);
}

/* compute_hydro (diff_comp.c:210-232) for the E-M and IG-T steps: one step on the device,
   rejected attempts redrawn there; positions, forces and the random-stream state come back
   to the host on output steps only */
void compute_hydro(DOUBLE* _E)
{
    uint64_t st;
    double step_dt;
    curr_iter++;
    if (bdh_step(g_sys, 1, NULL)) fail("step");
    if (bdh_last_dt(g_sys, &step_dt)) fail("step dt");
    curr_dt = step_dt;
    curr_time += curr_dt;
    if (bdh_get_coords(g_sys, coord)) fail("get_coords");
    if (bdh_get_forces(g_sys, F[0])) fail("get_forces");
    if (output_step(step) || step + 1 >= bdsteps)
    {
        if (bdh_get_rng_state(g_sys, &st)) fail("rng state");
        set_host_rng_state(st);
        if (_E)
        {
            double Eb[3];
            int rep;
            if (bdh_compute_forces(g_sys, NULL, Eb, &rep)) fail("energies");
            memset(_E, 0, sizeof(DOUBLE) * (ENERGY_LAST + 1));
            _E[ENERGY_LJ] = Eb[0];
            _E[ENERGY_COULOMB] = Eb[1];
            _E[ENERGY_BOND] = Eb[2];
        }
    }
}

/* compute_single_D (diff_tensor.c:561): the full tensor, read by single_point_print */
void compute_single_D(DOUBLE* _D)
{
    if (bdh_compute_D(g_sys, _D, size * DIMS0)) fail("compute_D");
}

/* ---- the rest of the backend interface of diff_tensor.h:166-253 and myblas.h:37-51.  The
   driver's own loop reaches the backend through compute_hydro / compute_forces only; these
   serve other callers of the same interface.  Blocks and tensors are those of the backend's
   current positions: the coordinates of the last compute_forces, compute_hydro, load_coord
   or compute_D. */

static void sync_host_forces(DOUBLE* _E)
{
    if (bdh_get_forces(g_sys, F[0])) fail("get_forces");
    if (_E)
    {
        double Eb[3];
        int rep;
        if (bdh_compute_forces(g_sys, NULL, Eb, &rep)) fail("energies");
        memset(_E, 0, sizeof(DOUBLE) * (ENERGY_LAST + 1));
        _E[ENERGY_LJ] = Eb[0];
        _E[ENERGY_COULOMB] = Eb[1];
        _E[ENERGY_BOND] = Eb[2];
    }
}

/* init_tensor (diff_tensor.c:546-559): the device's Ewald tables are built with the system;
   what the driver can see is the Beenakker splitting parameter, forced to sqrt(pi)/L */
void init_tensor(void)
{
    if (bc == BC_BOX && hydro && ewald_method == EWALD_METHOD_BEENAKKER)
        ewald_alpha = SQRTPI / box[0];
}

/* compute_D (diff_tensor.c:472-544): the whole symmetric tensor at coord (the reference
   fills the upper triangle; the lower one is its mirror) */
void compute_D(DOUBLE* _D)
{
    if (!hydro) return;
    if (bdh_set_coords(g_sys, coord)) fail("set_coords");
    if (bdh_compute_D(g_sys, _D, size * DIMS0)) fail("compute_D");
}

void compute_Dij(INT i, INT j, DOUBLE* tD)
{
    if (bdh_compute_Dij(g_sys, i, j, 0, tD)) fail("compute_Dij");
}

void diff_tensor_calc_pos(INT i, INT j, DOUBLE* tD)
{
    if (bdh_compute_Dij(g_sys, i, j, 1, tD)) fail("diff_tensor_calc_pos");
}

void diff_tensor_calc_ewald_smith(INT i, INT j, DOUBLE* tD)
{
    if (bdh_compute_Dij(g_sys, i, j, 2, tD)) fail("diff_tensor_calc_ewald_smith");
}

void diff_tensor_calc_ewald_smith_diag(INT i, DOUBLE* tD)
{
    if (bdh_compute_Dij(g_sys, i, i, 2, tD)) fail("diff_tensor_calc_ewald_smith_diag");
}

void diff_tensor_calc_ewald_beenakker(INT i, INT j, DOUBLE* tD)
{
    if (bdh_compute_Dij(g_sys, i, j, 3, tD)) fail("diff_tensor_calc_ewald_beenakker");
}

void diff_tensor_calc_ewald_beenakker_diag(INT i, DOUBLE* tD)
{
    if (bdh_compute_Dij(g_sys, i, i, 3, tD)) fail("diff_tensor_calc_ewald_beenakker_diag");
}

/* compute_r (diff_rand.c:47-236): Rs[0] at coord from the next Gaussians of the stream.
   The factor of D stays on the device; _D is not overwritten with it. */
void compute_r(DOUBLE* _D)
{
    (void)_D;
    if (bdh_set_coords(g_sys, coord)) fail("set_coords");
    if (bdh_compute_r(g_sys, Rs[0])) fail("compute_r");
}

/* compute_hydro_ermak (diff_equation.c:216-285): coord_out = coord + Rf + Rs (+ dt G coord
   in a flow field), Rf[0] and Rs[0] the move's parts; the caller evaluates forces at
   coord_out */
void compute_hydro_ermak(DOUBLE* coord_out)
{
    if (bdh_ermak_move(g_sys, curr_dt > 0 ? curr_dt : dt, coord_out)) fail("compute_hydro_ermak");
    if (bdh_get_last_displacement(g_sys, Rf[0], Rs[0])) fail("displacement");
}

/* compute_hydro_IGT (diff_equation.c:317-374): predictor, mean tensor, then the corrector;
   1 if the predictor overlaps (coord unchanged, F[0] the predictor's forces), 2 if the
   corrector does (coord the corrector's positions) */
INT compute_hydro_IGT(DOUBLE* _E)
{
    int res = 0;
    curr_iter++;
    if (bdh_igt_phase(g_sys, 0, curr_dt > 0 ? curr_dt : dt, coord, &res)) fail("compute_hydro_IGT");
    sync_host_forces(res == 1 ? NULL : _E);
    return res;
}

/* compute_hydro_IGT_second_phase (diff_equation.c:288-315): the corrector from coord */
INT compute_hydro_IGT_second_phase(DOUBLE* _E)
{
    int res = 0;
    if (bdh_igt_phase(g_sys, 1, curr_dt > 0 ? curr_dt : dt, coord, &res)) fail("compute_hydro_IGT_second_phase");
    sync_host_forces(_E);
    return res;
}

/* my_dpotrf / my_dtrmv (myblas.c:45-118) on the device, LAPACK's 'U' convention: A = U^T U
   with U in the upper triangle; INFO as LAPACK returns it */
INT my_dpotrf(DOUBLE* A, INT n, INT LD)
{
    const int info = bdh_dpotrf('U', A, n, LD, BDH_FP64);
    if (info < 0) fail("my_dpotrf");
    return info;
}

void my_dtrmv(DOUBLE* A, DOUBLE* v, DOUBLE* o, INT n, INT LD)
{
    if (bdh_dtrmv('U', A, v, o, n, LD, BDH_FP64)) fail("my_dtrmv");
}

/* init_cholesky / free_cholesky (myblas.c:120-132): the factorization's workspaces live on
   the device and belong to each call, so there is no host vector to allocate */
void init_cholesky(void) {}
void free_cholesky(void) {}

/* newton.h:33-37: the elastic-collision moves of ermak_newton are not part of this backend
   (init_diff_tensor accepts no algorithm that reaches them) */
INT newton_moves(INT* count)
{
    (void)count;
    UNERR("HIP backend: elastic collisions (ermak_newton) are not implemented");
    return 1;
}

void newton_init_vel(void) { UNERR("HIP backend: elastic collisions (ermak_newton) are not implemented"); }
void init_newton(void) {}
void free_newton(void) {}
