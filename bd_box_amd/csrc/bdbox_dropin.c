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
    if ((algorithm == DIFF_ALG_IGT_CONST || algorithm == DIFF_ALG_IGT_VAR) && hydro == DIFF_GEYER)
        UNERR("HIP backend: IG-T with hydro geyer is not supported");
    if (hydro != DIFF_NONE && hydro != DIFF_CHOLESKY && hydro != DIFF_GEYER)
        UNERR("HIP backend: hydro none / cholesky / geyer");
    if (vel_grad_tensor) UNERR("HIP backend: no flow field");
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
    /* diag_D as init_diff_tensor builds it (diff_decl.c:93-98) */
    preDii = kB_1_6PI * T / (visc * vfactor);
    preDij = kB_1_8PI * T / (visc * vfactor);
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

void save_curr_coord(void) {}
void save_curr_forces(void) {}

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
    if (output_step(step) || step + 1 >= bdsteps)
    {
        if (bdh_get_coords(g_sys, coord)) fail("get_coords");
        if (bdh_get_forces(g_sys, F[0])) fail("get_forces");
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
