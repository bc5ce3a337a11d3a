/*
 * bdbox_hip.h -- C ABI of the MI355X (gfx950) backend for BD_BOX's Ermak-McCammon step
 * with hydrodynamic interactions.
 *
 * Plain C types only (pointers, sizes, doubles).  Host arrays use the reference's layouts
 * (src/data.h:252-253, src/data.c:85-109):
 *   coord  double[4N]  (x, y, z, a) per bead
 *   Q      double[N]
 *   LJ     double[2N]  (R, eps) per bead, R = half the LJ diameter of the .str file
 *   bonds  int[2B] bead indices, bond_parms double[3B] = (r0, r_max, H)
 *   D, L   column-major n x n (n = 3N), leading dimension ld >= n
 * All functions return 0 on success and a negative value on error; bdh_last_error() then
 * describes it.  Functions are synchronous: results are in host memory when they return.
 * One bdh_system may be driven by one host thread at a time (the reference's own
 * threading contract: one thread calls compute_hydro per step).
 *
 * The entry points replace the reference's backend interface (the TU set substituted at
 * link time in src/Makefile.am:40-70); each cites the function it stands in for.  The
 * adapter that maps the reference's global-state calls onto these entry points is
 * bd_box_amd/csrc/bdbox_dropin.c.
 */
#ifndef BDBOX_HIP_H
#define BDBOX_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct bdh_system bdh_system;

enum bdh_precision {
    BDH_FP64 = 0,   /* everything in FP64 (the comparator)                              */
    BDH_MIXED = 1,  /* FP32 tensor/forces/trailing updates, FP64 panels and positions   */
    BDH_MIXED_SPLIT = 2  /* as BDH_MIXED, trailing updates as split-bf16 products (three
                            bf16 MFMAs per FP32 product, FP32 accumulation)                 */
};

enum bdh_hydro { BDH_HYDRO_NONE = 0, BDH_HYDRO_CHOLESKY = 1, BDH_HYDRO_GEYER = 2 };

/* integration algorithm (src/diff_alg/diff_tensor.h:29-33); IG-T with hydro none or
 * cholesky */
enum bdh_algorithm { BDH_ERMAK_CONST = 0, BDH_ERMAK_VAR = 1, BDH_IGT_CONST = 3, BDH_IGT_VAR = 4 };

/* Control parameters after the reference's .prm parsing (src/input.c:353-503): DOUBLE
 * keywords already rounded through float (src/input.c:933). */
typedef struct bdh_config {
    double dt, T, visc, vfactor;             /* dt [ps], T [K], visc [P], vfactor       */
    double alpha_lj, cutoff_lj;              /* cutoff_lj <= 0: 1.122462 (R_i + R_j)    */
    int lj_6_term;
    int elec;
    double epsilon_c, kappa_c, gamma_c, cutoff_c;
    double bond_lj_scale, bond_c_scale;
    int check_overlap;
    int bc;                                  /* 0 none, 1 periodic cube                  */
    double box;                              /* cube edge when bc == 1                   */
    int hydro;                               /* enum bdh_hydro                           */
    int ewald_method;                        /* bc periodic: 0 Smith, 1 Beenakker        */
    int ewald_real;                          /* lattice sphere radius 0..5               */
    int precision;                           /* enum bdh_precision                       */
    int refactor_every;                      /* refactor D every k steps (reference: 1)  */
    uint32_t rand_seed;                      /* drand48 seed of the Gaussian stream      */
    int nb_list;                             /* 0 brute force; 1 bucket_space, 2 bucket_list:
                                                device cell list (src/potentials/bucket.h:40-46) */
    int algorithm;                           /* enum bdh_algorithm                       */
    int move_attempts;                       /* rejections allowed per step when
                                                check_overlap (diff_decl.c:71: 10000)    */
    double ewald_alpha;                      /* Smith splitting parameter in units of 1/L
                                                (diff_decl.c:79; Beenakker forces sqrt(pi)/L) */
} bdh_config;

/* Last error of the calling thread ("" if none). */
const char* bdh_last_error(void);

/* Select the HIP device for systems created afterwards by this thread.
 * Stands in for init_cuda (src/cuda.c:88-148). */
int bdh_init(int device);

/* Allocate device state and upload the bead table; computes the initial forces
 * (main_loop.c:91).  Stands in for init_diff_tensor (src/diff_alg/diff_decl.c:83-175)
 * and init_forces (src/potentials/calc_func.c:411-415). */
int bdh_create(const bdh_config* cfg, int N, const double* coord, const double* Q,
               const double* LJ, int nbonds, const int* bonds, const double* bond_parms,
               bdh_system** out);
/* Stands in for free_diff_tensor / free_forces (diff_decl.c:177-217, calc_func.c:417-421). */
void bdh_destroy(bdh_system* sys);

/* Advance `steps` Ermak-McCammon steps: Rf = (dt/kT) D F, Rs = sqrt(2 dt) L xi, r += Rf+Rs,
 * then forces at the new positions.  xi == NULL draws the Gaussians from the device
 * reproduction of the reference's drand48 stream (src/rand_move.c:132-168); otherwise
 * xi[steps * 3N] supplies them.  Stands in for compute_hydro with algorithm ermak_const
 * (src/diff_alg/diff_comp.c:152-179, 210-232; diff_equation.c:216-286; diff_rand.c:47-85). */
int bdh_step(bdh_system* sys, int steps, const double* xi);
/* With check_overlap, a step whose new positions overlap is undone (positions and forces
 * restored) and redrawn with the next Gaussians of the device stream (also when xi supplied
 * the first draw), ermak_var halving dt for each of the first 16 rejections; more than
 * move_attempts rejections of one step fail with "Wrong conformation, can't move".  The
 * factor of the step is reused across its attempts.  Follows comp_step_ermak
 * (src/diff_alg/diff_comp.c:152-179).  *dt receives the length of the last step taken (the
 * reference's curr_dt, which its caller adds to curr_time, diff_comp.c:212-216). */
int bdh_last_dt(bdh_system* sys, double* dt);

/* Assemble D at the current positions and copy it out (full symmetric matrix).
 * Stands in for compute_D / compute_single_D (src/diff_alg/diff_tensor.c:472-544, 561). */
int bdh_compute_D(bdh_system* sys, double* D, int ld);

/* Forces F[3N] and energies E[3] = (E_LJ, E_coulomb, E_bond) at the current positions;
 * *repeat receives the reference's "repeat step" flag (overlap/NaN with check_overlap).
 * Stands in for compute_forces (src/potentials/calc_func.c:106-186). */
int bdh_compute_forces(bdh_system* sys, double* F, double* E, int* repeat);

/* Pairs i < j with |r_j - r_i| <= cutoff (minimum image when periodic) at the current
 * positions, found through the device cell list; pairs[2k], pairs[2k+1] sorted ascending.
 * *count receives the number of pairs; at most cap are stored.  Stands in for the
 * neighbour queries of the reference's bucket lists (init_bucket_f / get_neighbours,
 * src/potentials/bucket.h:112-117; init_keys / can_interact, :119-121). */
int bdh_neighbour_pairs(bdh_system* sys, double cutoff, int* pairs, long long cap, long long* count);

/* Assemble and factor D = L L^T at the current positions (the refactor step), copy L out
 * (lower triangle; zeros above).  Stands in for my_dpotrf on D (src/myblas.c:45-92,
 * called from diff_rand.c:76-82). */
int bdh_factor(bdh_system* sys, double* L, int ld);

/* y = L x with the current factor.  Stands in for my_dtrmv (src/myblas.c:94-118). */
int bdh_noise_product(bdh_system* sys, const double* x, double* y);

/* D.F from blocks evaluated on the fly (never stored); Rf = (dt/kT) D F.
 * Stands in for compute_Rf (src/diff_alg/diff_equation.c:119-214). */
int bdh_compute_Rf(bdh_system* sys, double* Rf);

/* One 3x3 block of D at the current positions, tD[k + 3 l] = D(3 j + k, 3 i + l) (the
 * orientation of insert9, src/diff_alg/diff_tensor.c:425-435; the block is symmetric).
 * kind 0: the system's own tensor, compute_Dij (diff_tensor.c:437-470), i == j the diagonal
 * block; 1: free-space RPY, diff_tensor_calc_pos (:45-92, no minimum image, i != j);
 * 2 / 3: the Smith / Beenakker Ewald block and its diagonal (:102-263 / :265-422), for a
 * periodic system created with that method. */
int bdh_compute_Dij(bdh_system* sys, int i, int j, int kind, double* tD);

/* Rs[3N] of compute_r (src/diff_alg/diff_rand.c:47-236) at the current positions, from the
 * next 3N Gaussians of the device stream: sqrt(2 dt) L xi (L refreshed at the current
 * positions when refactor_every is 1, the reference's schedule), the Geyer-Winter TEA
 * displacement, or sqrt(2 dt diag_D) xi without hydrodynamics. */
int bdh_compute_r(bdh_system* sys, double* Rs);

/* One Ermak-McCammon move of length dt from the current positions and forces,
 * compute_hydro_ermak (src/diff_alg/diff_equation.c:216-264): coord_out[4N] = r + Rf + Rs
 * (wrapped into the box when periodic), Rs from the next Gaussians of the device stream.
 * The system's positions stay where they were (the caller evaluates forces at coord_out and
 * accepts or rejects, comp_step_ermak, diff_comp.c:152-179); bdh_get_last_displacement
 * returns the move's Rf and Rs. */
int bdh_ermak_move(bdh_system* sys, double dt, double* coord_out);

/* The two phases of an IG-T step of length dt for a system created with algorithm
 * igt_const / igt_var.  phase 0 stands in for compute_hydro_IGT (diff_equation.c:317-374):
 * the predictor from the current positions and forces, its forces; if they report an
 * overlap, *result = 1 and the positions are back at the start (the forces are those of the
 * predictor); otherwise the mean deterministic move and the factor of the mean tensor, then
 * phase 1.  phase 1 stands in for compute_hydro_IGT_second_phase (:288-315): the corrector
 * from the current positions (the step's start) with the next Gaussians and the mean
 * tensor's factor, its forces, *result = 2 on an overlap, else 0.  coord_out[4N] receives
 * the positions reached. */
int bdh_igt_phase(bdh_system* sys, int phase, double dt, double* coord_out, int* result);

/* beta and eps of the last Geyer-Winter TEA displacement (hydro geyer): the reference's
 * eps (the epsilon accumulator of diff_tensor.h:136, summed, normalised and clamped in compute_r,
 * src/diff_alg/diff_rand.c:197-206) and the beta it derives from it (:207-216). */
int bdh_tea_coefficients(bdh_system* sys, double* beta, double* eps);

/* Extension of the Beenakker-Ewald tensor (no counterpart in the reference, which forces
 * alpha = sqrt(pi)/L in init_tensor, src/diff_alg/diff_tensor.c:551-554, sums the real and
 * reciprocal lattices over one sphere, src/input.c:402-409, and centres the image sum on the
 * separation of the wrapped positions): for alpha_scale > 0, splitting parameter
 * alpha = alpha_scale sqrt(pi)/L, real-space images with h^2 + f^2 + g^2 <= R_real^2 around
 * the minimum-image separation and reciprocal vectors with h^2 + f^2 + g^2 <= R_recip^2
 * (radii 0..5), so the split can be chosen for a target truncation error.  alpha_scale 0
 * returns to the reference's tensor.  Periodic systems with ewald_method beenakker. */
int bdh_set_ewald_split(bdh_system* sys, double alpha_scale, int R_real, int R_recip);

/* Extension of the same split (no counterpart in the reference): real-space images farther
 * than r_cut (A) from the pair are dropped, the distance bound from which the radii of a
 * split for a target truncation error are derived; 0 keeps every image of the sphere.  After
 * bdh_set_ewald_split with alpha_scale > 0 (alpha_scale 0 clears it). */
int bdh_set_ewald_cutoff(bdh_system* sys, double r_cut);

/* A flow field: the velocity gradient tensor G[9] (row, column), the .prm keyword
 * vel_grad_tensor (src/diff_alg/diff_decl.c:72, parsed at src/input.c:1110-1125); NULL
 * removes it.  Every move then adds dt G r as compute_hydro_ermak does
 * (src/diff_alg/diff_equation.c:265-285), and the IG-T corrector adds half of the
 * predictor's FS = dt G r0 + dt G r* (:304-312).  bdh_step_vjp refuses systems with a flow
 * field. */
int bdh_set_flow(bdh_system* sys, const double* G);

/* Replace the forces F[3N] of the current positions, load_forces (diff_decl.c:234-237). */
int bdh_set_forces(bdh_system* sys, const double* F);

int bdh_get_coords(bdh_system* sys, double* coord);          /* 4N, for save_output       */
int bdh_set_coords(bdh_system* sys, const double* coord);    /* load_coord (diff_decl.c:229) */
int bdh_get_forces(bdh_system* sys, double* F);              /* F[0] of the last step     */
int bdh_get_last_displacement(bdh_system* sys, double* Rf, double* Rs);
int bdh_get_rng_state(bdh_system* sys, uint64_t* state);     /* get_state (rand_move.c:170) */
int bdh_set_rng_state(bdh_system* sys, uint64_t state);      /* set_state (rand_move.c:182) */

/* Host-matrix LA on the device, replacing myblas (src/myblas.c:45-118):
 * my_dpotrf(A, n, LD) factors in place; uplo 'U' returns U with A = U^T U in the upper
 * triangle (the LAPACK convention of the reference), 'L' returns L in the lower triangle.
 * Returns the LAPACK INFO (0, or the 1-based index of the failed pivot).
 * my_dtrmv: o = L v where A holds the factor as written by bdh_dpotrf with the same uplo. */
int bdh_dpotrf(char uplo, double* A, int n, int ld, int precision);
int bdh_dtrmv(char uplo, const double* A, const double* v, double* o, int n, int ld,
              int precision);

/* `count` Gaussians of the reference's stream continuing from *state (updated).
 * Stands in for gauss_rand (src/rand_move.c:132-168). */
int bdh_gauss(uint64_t* state, double* out, int count);
/* 48-bit drand48 state for srand48_r(seed) (src/rand_move.c:53-80). */
uint64_t bdh_seed_state(uint32_t seed);

/* bdh_gauss without any device allocation.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (tight serving loops, harnesses that forbid
 * hipMalloc).  The stream state and the output are staged through a fixed array that each
 * device holds until the process ends, so count <= BDH_GAUSS_NOALLOC_MAX (a larger count
 * is refused with an error).  Calls on one device share that array: calls that overlap on
 * one device race.  Results are identical to bdh_gauss for the same state and count. */
#define BDH_GAUSS_NOALLOC_MAX (1 << 20)
int bdh_gauss_noalloc(uint64_t* state, double* out, int count);

/* Test hook: one MFMA on a single wavefront with the operand layouts this backend assumes
 * (kind 0 f32 32x32x2, 1 f32 16x16x4, 2 f64 16x16x4, 3 bf16 32x32x16, 4 bf16 16x16x32).
 * A (MxK) and B (KxN) row-major host arrays (float, or double for kind 2); C = A B. */
int bdh_probe_mfma(int kind, const void* A, const void* B, void* C);

/* Measurement (no counterpart in the reference, which has no timers: diff_comp.c:58-74).
 * bdh_time_factor: mean ms of assembly + factorization over `reps` refactors.
 * bdh_time_steps : ms of `steps` device-stream steps, HIP events on the system's stream.
 * bdh_phase_timing(sys, 1) brackets every later launch with HIP events on the system's
 * stream; bdh_phase_times returns, per phase, summed ms and launch counts since the last
 * call, phases 0 noise, 1 assembly, 2 panel factor, 3 trailing update, 4 D.F, 5 L.xi,
 * 6 move, 7 forces (arrays of 8). */
int bdh_time_factor(bdh_system* sys, int reps, double* ms_per_factor);
int bdh_time_steps(bdh_system* sys, int steps, double* ms_total);
int bdh_phase_timing(bdh_system* sys, int enable);
int bdh_phase_times(bdh_system* sys, double* ms, long long* counts);
int bdh_sync(bdh_system* sys);

/* Reverse-mode derivative of one step (no counterpart in the reference, which has no
 * derivatives): for the FP64 Ermak-McCammon step with Cholesky noise refactored at coord,
 * r' = r + (dt/kT) D(r) F(r) + sqrt(2 dt) L(r) xi with the given xi[3N], rbar[3N] receives
 * the cotangent of r for the cotangent gbar[3N] of r' (forces, RPY assembly and the
 * Cholesky factor all differentiated).  Systems with hydro cholesky and bc none; for the
 * mixed precisions the factor differentiated is the FP32 factor the step drew its noise
 * with, promoted to FP64, and the rest of the reverse step runs in FP64.  The system is left
 * at coord with its forces and factor there. */
int bdh_step_vjp(bdh_system* sys, const double* coord, const double* xi, const double* gbar,
                 double* rbar);
/* Factor reuse for the reverse pass (no counterpart in the reference): bdh_factor_cache
 * allocates `slots` factors on the device (FP64, or FP32 for the mixed precisions); bdh_save_factor copies the factor of the
 * last step taken into a slot; bdh_step_vjp_cached is bdh_step_vjp with the factor at coord
 * taken from a slot instead of refactoring (the caller saved it when stepping from coord). */
int bdh_factor_cache(bdh_system* sys, int slots);
int bdh_save_factor(bdh_system* sys, int slot);
int bdh_step_vjp_cached(bdh_system* sys, int slot, const double* coord, const double* xi,
                        const double* gbar, double* rbar);
/* bdh_step_vjp (slot < 0) or bdh_step_vjp_cached (slot >= 0) that also returns the
 * cotangents of the force-field parameters for the same gbar (no counterpart in the
 * reference): eps_bar[N], R_bar[N] of the per-bead LJ epsilon and radius (LJ[2i+1], LJ[2i];
 * they enter at src/potentials/LJ.c:37-64 as sigma_ij = R_i + R_j, eps_ij = sqrt(eps_i
 * eps_j), set from the .str file at src/data.c:97-102), Q_bar[N] of the charges
 * (electro.c:36-56) and H_bar[nbonds] of each bond's force constant (bond.c:43-85).  Any
 * of the four may be NULL. */
int bdh_step_vjp_params(bdh_system* sys, int slot, const double* coord, const double* xi,
                        const double* gbar, double* rbar, double* eps_bar, double* R_bar,
                        double* Q_bar, double* H_bar);
/* Replace the per-bead charges Q[N], LJ (R, eps)[2N] and the per-bond parameters
 * bond_parms[3 * nbonds] (r0, r_max, H; the bond list given to bdh_create) of a system, any
 * of them NULL to keep; the forces at the current positions are recomputed.  The reference's
 * driver owns these host arrays (src/data.c:85-109) and its backend reads them at every
 * force evaluation, so a caller that changes them between steps sees the new values. */
int bdh_set_params(bdh_system* sys, const double* Q, const double* LJ, const double* bond_parms);

/* ---- Batches of independent chains.  No counterpart in the reference, which integrates one
 * system per process: nchains systems of uneven sizes that share one bdh_config advance
 * together, every phase of the step (assembly, each launch of the blocked factorization,
 * L.xi, D.F, move, forces, the Gaussian streams) running as one launch over all chains.
 * Chain c follows exactly the trajectory bdh_create + bdh_step give for its own beads with
 * rand_seed = seeds[c]: per chain the launch shapes and the arithmetic are the same.
 * Supported: hydro cholesky, bc none, ermak_const, check_overlap off, brute-force pairs.
 * Beads are concatenated chain after chain: coord[4 * sum N], Q, LJ[2 * sum N]; chain c has
 * nbonds[c] bonds, listed after those of chains 0..c-1 in bonds / bond_parms with
 * chain-local bead indices. */
typedef struct bdh_batch bdh_batch;
int bdh_batch_create(const bdh_config* cfg, int nchains, const int* N, const double* coord,
                     const double* Q, const double* LJ, const int* nbonds, const int* bonds,
                     const double* bond_parms, const uint32_t* seeds, bdh_batch** out);
void bdh_batch_destroy(bdh_batch* batch);
/* xi == NULL: each chain draws from its own stream; otherwise xi[steps][3 * sum N]. */
int bdh_batch_step(bdh_batch* batch, int steps, const double* xi);
int bdh_batch_get_coords(bdh_batch* batch, double* coord);   /* 4 * sum N */
int bdh_batch_get_forces(bdh_batch* batch, double* F);       /* 3 * sum N */
/* Measurement, as bdh_time_steps / bdh_phase_timing / bdh_phase_times. */
int bdh_batch_time_steps(bdh_batch* batch, int steps, double* ms_total);
int bdh_batch_phase_timing(bdh_batch* batch, int enable);
int bdh_batch_phase_times(bdh_batch* batch, double* ms, long long* counts);

/* Update schedule of the batch's factorizations (no counterpart in the reference, whose
 * CPU path calls LAPACK dpotrf).  Inside each 256-column outer panel, a chain whose padded
 * size n_pad is at most left_max updates left-looking (every panel of the outer panel so
 * far updates the next 64 columns); the others update right-looking (each panel updates all
 * later columns of the outer panel).  left_max < 0 (the default) gives every chain the
 * schedule a single system of its size uses: left-looking up to n_pad 4096, so chain c's
 * trajectory is then bitwise the single system's.  Schedules differ in rounding only. */
int bdh_batch_set_schedule(bdh_batch* batch, int left_max);
int bdh_batch_get_schedule(bdh_batch* batch, int* left_max);
/* Tile-schedule autotuner (no counterpart in the reference): times the refactor of every
 * chain (assembly + factorization, graph replays, fastest of `reps`) under each candidate
 * left_max -- 0, then up to BDH_TUNE_MAX - 1 of the batch's padded sizes, the largest
 * included -- and keeps the fastest.  cand / ms (BDH_TUNE_MAX entries each, may be NULL)
 * receive the candidates and their times, *ncand their count, *best the choice.  Call it
 * before the first step; the first step then refactors as usual. */
#define BDH_TUNE_MAX 8
int bdh_batch_autotune(bdh_batch* batch, int reps, int* cand, double* ms, int* ncand, int* best);

/* Tile shape of the batch's factorizations (no counterpart in the reference): the update
 * schedule as in bdh_batch_set_schedule, and the width of the outer panels (256, 512 or
 * 1024 columns) within which panels of 64 columns update, before the trailing matrix is
 * updated with the whole outer panel.  Widths and schedules differ in rounding only. */
int bdh_batch_set_tiles(bdh_batch* batch, int left_max, int outer);
int bdh_batch_get_tiles(bdh_batch* batch, int* left_max, int* outer);
/* bdh_batch_autotune over tile shapes: every schedule candidate of bdh_batch_autotune under
 * each outer panel width (256, then 512 and 1024 where the largest chain reaches them), at
 * most BDH_TUNE_TILES_MAX candidates.  cand_left / cand_outer / ms (may be NULL) receive
 * the candidates and their refactor times, *ncand their count, *best the index kept. */
#define BDH_TUNE_TILES_MAX 24
int bdh_batch_autotune_tiles(bdh_batch* batch, int reps, int* cand_left, int* cand_outer, double* ms,
                             int* ncand, int* best);

#ifdef __cplusplus
}
#endif
#endif
