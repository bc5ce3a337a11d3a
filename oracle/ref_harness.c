/*
 * Golden-vector harness for the reference CPU path (test infrastructure only).
 *
 * Linked against the reference's own object files (see oracle/Makefile, target `ref`),
 * it replaces the reference's main() (src/main.c:52-142) with one that performs the same
 * start-up sequence -- parse_args / read_prm_file / apply_options / init_rand /
 * read_str_file / init_diff_tensor / init_forces (src/main.c:87-121) -- and then dumps one
 * quantity of the hot path with 17 significant digits, so the FP64 restatement in
 * oracle/bdbox_oracle.c and the HIP kernels can be pinned against the reference itself.
 *
 * usage: ref_harness MODE [ARGS] -- FILE.prm [--key=value ...]
 *   params        derived constants and the (centred) bead table
 *   D             full symmetric 3N x 3N diffusion tensor   (compute_D, diff_tensor.c:472)
 *   forces        forces F[0] and energies E               (compute_forces, calc_func.c:106)
 *   chol          lower Cholesky factor L, D = L L^T        (my_dpotrf/my_dtrmv, myblas.c:45-118)
 *   gauss M K     K successive gauss_rand(.., M) draws      (rand_move.c:132-168)
 *   step1         X (scaled noise), Rf and Rs of the first step   (diff_equation.c:216-264)
 *   traj K        positions and curr_dt after each of K steps (main_loop.c:91-117;
 *                 curr_dt < dt after ermak_var rejections, diff_comp.c:152-179)
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "data.h"
#include "err.h"
#include "input.h"
#include "main_loop.h"
#include "myblas.h"
#include "rand_move.h"
#include "trans.h"
#include "diff_alg/diff_tensor.h"
#include "potentials/calc_func.h"
#include "potentials/LJ.h"
#include "potentials/electro.h"

static void print_vec(const char* tag, const DOUBLE* v, int n)
{
    int i;
    printf("%s %d\n", tag, n);
    for (i = 0; i < n; ++i) printf("%.17g\n", (double)v[i]);
}

static void ensure_Ds(void)
{
    /* geyer_on_the_fly runs never allocate Ds (diff_decl.c:155-159); the reference's own
       single-point dump allocates it on demand the same way (main_loop.c:133-137) */
    if (Ds == NULL)
    {
        Ds = (DOUBLE*)malloc(sizeof(DOUBLE) * size * size * DIMS0 * DIMS0);
        memset(Ds, 0, sizeof(DOUBLE) * size * size * DIMS0 * DIMS0);
    }
}

static void dump_full_D(void)
{
    int n = size * DIMS0, i, j;
    ensure_Ds();
    compute_D(Ds);
    printf("D %d\n", n);
    for (i = 0; i < n; ++i)
    {
        for (j = 0; j < n; ++j)
        {
            /* upper triangle, column major, LD = n (diff_tensor.c:425-435) */
            double v = (i <= j) ? Ds[i + j * n] : Ds[j + i * n];
            printf("%.17g%c", v, j + 1 == n ? '\n' : ' ');
        }
    }
}

static void dump_chol(void)
{
    int n = size * DIMS0, i, j;
    DOUBLE* e = (DOUBLE*)malloc(sizeof(DOUBLE) * n);
    DOUBLE* o = (DOUBLE*)malloc(sizeof(DOUBLE) * n);
    DOUBLE* L = (DOUBLE*)malloc(sizeof(DOUBLE) * n * n);
    ensure_Ds();
    compute_D(Ds);
    if (my_dpotrf(Ds, n, n)) UNERR("Cholesky decomposition error");
    /* column j of L = L e_j, through the reference's own triangular product */
    for (j = 0; j < n; ++j)
    {
        memset(e, 0, sizeof(DOUBLE) * n);
        memset(o, 0, sizeof(DOUBLE) * n);
        e[j] = 1.0;
        my_dtrmv(Ds, e, o, n, n);
        for (i = 0; i < n; ++i) L[i + j * n] = o[i];
    }
    printf("L %d\n", n);
    for (i = 0; i < n; ++i)
        for (j = 0; j < n; ++j)
            printf("%.17g%c", (double)L[i + j * n], j + 1 == n ? '\n' : ' ');
    free(e); free(o); free(L);
}

int main(int argc, char* argv[])
{
    const char* mode;
    int sep, i, a1 = 0, a2 = 0;
    if (argc < 4) { fprintf(stderr, "usage: ref_harness MODE [ARGS] -- FILE.prm [--k=v]\n"); return 2; }
    mode = argv[1];
    for (sep = 2; sep < argc && strcmp(argv[sep], "--") != 0; ++sep) {}
    if (sep >= argc) { fprintf(stderr, "missing --\n"); return 2; }
    if (sep > 2) a1 = atoi(argv[2]);
    if (sep > 3) a2 = atoi(argv[3]);
#ifdef _OPENMP
    g_threads = omp_get_max_threads();
#endif
    /* parse_args expects argv[0] to be the program name */
    argv[sep] = argv[0];
    if (!parse_args(argc - sep, argv + sep)) return 2;
    read_prm_file();
    apply_options(0);
    init_rand(rand_seed);
    read_str_file();
    init_diff_tensor();
    init_forces();
    for (i = 0; i < DIMS1 * size; i += DIMS1) trans_to_box(coord + i); /* main_loop.c:61-64 */

    if (!strcmp(mode, "params"))
    {
        printf("size %d\n", size);
        printf("dt %.17g\nT %.17g\nvisc %.17g\nvfactor %.17g\n", (double)dt, (double)T, (double)visc, (double)vfactor);
        printf("preDii %.17g\npreDij %.17g\n", (double)preDii, (double)preDij);
        printf("cutoff_lj %.17g\ncutoff_c %.17g\nkappa_c %.17g\nepsilon_c %.17g\ngamma_c %.17g\nalpha_lj %.17g\n",
               (double)cutoff_lj, (double)cutoff_c, (double)kappa_c, (double)epsilon_c, (double)gamma_c, (double)alpha_lj);
        printf("bond_lj_scale %.17g\nbond_c_scale %.17g\n", (double)bond_lj_scale, (double)bond_c_scale);
        print_vec("coord", coord, DIMS1 * size);
        print_vec("Q", Q, size);
        print_vec("LJ", LJ, 2 * size);
    }
    else if (!strcmp(mode, "D"))
    {
        dump_full_D();
    }
    else if (!strcmp(mode, "forces"))
    {
        int rep = compute_forces(E, 0.0);
        printf("repeat %d\n", rep);
        print_vec("F", F[0], DIMS0 * size);
        print_vec("E", E, ENERGY_LAST);
    }
    else if (!strcmp(mode, "chol"))
    {
        dump_chol();
    }
    else if (!strcmp(mode, "gauss"))
    {
        int k;
        DOUBLE* g = (DOUBLE*)malloc(sizeof(DOUBLE) * (a1 > 0 ? a1 : 1));
        for (k = 0; k < (a2 > 0 ? a2 : 1); ++k)
        {
            gauss_rand(g, a1);
            print_vec("g", g, a1);
        }
        free(g);
    }
    else if (!strcmp(mode, "step1"))
    {
        if (compute_forces(NULL, 0.0)) UNERR("Wrong starting position");
        print_vec("F0", F[0], DIMS0 * size);
        save_curr_coord();
        save_curr_forces();
        curr_dt = dt;
        curr_iter++;
        compute_hydro_ermak(coord);
        print_vec("X", X, DIMS0 * size);
        print_vec("Rf", Rf[0], DIMS0 * size);
        print_vec("Rs", Rs[0], DIMS0 * size);
        print_vec("coord", coord, DIMS1 * size);
    }
    else if (!strcmp(mode, "traj"))
    {
        int k;
        if (compute_forces(NULL, 0.0)) UNERR("Wrong starting position");
        for (k = 1; k <= a1; ++k)
        {
            save_curr_coord();
            save_curr_forces();
            compute_hydro(NULL);
            print_vec("coord", coord, DIMS1 * size);
            printf("dt %.17g\n", (double)curr_dt);
        }
    }
    else
    {
        fprintf(stderr, "unknown mode %s\n", mode);
        return 2;
    }
    return 0;
}
