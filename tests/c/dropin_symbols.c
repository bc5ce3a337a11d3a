/*
 * Link check of the drop-in backend: takes the address of every entry point and global the
 * reference's backend interface declares (src/diff_alg/diff_tensor.h:76-253,
 * src/potentials/calc_func.h:34-106, src/myblas.h:37-51, src/diff_alg/newton.h:33-37), so the
 * link fails if bd_box_amd/csrc/bdbox_dropin.c leaves any of them undefined.  Built by
 * `make -C oracle dropin_symbols` against the reference's headers; prints the count.
 */
#include <stdio.h>

#include "data.h"
#include "myblas.h"
#include "diff_alg/diff_tensor.h"
#include "diff_alg/newton.h"
#include "potentials/calc_func.h"

int main(void)
{
    const void* fns[] = {
        (const void*)&init_diff_tensor, (const void*)&init_tensor, (const void*)&free_diff_tensor,
        (const void*)&diff_tensor_calc_pos, (const void*)&diff_tensor_calc_ewald_smith,
        (const void*)&diff_tensor_calc_ewald_smith_diag, (const void*)&diff_tensor_calc_ewald_beenakker,
        (const void*)&diff_tensor_calc_ewald_beenakker_diag, (const void*)&compute_hydro,
        (const void*)&compute_hydro_ermak, (const void*)&compute_hydro_IGT_second_phase,
        (const void*)&compute_D, (const void*)&compute_Dij, (const void*)&compute_hydro_IGT,
        (const void*)&compute_r, (const void*)&compute_single_D, (const void*)&save_curr_coord,
        (const void*)&save_curr_forces, (const void*)&load_coord, (const void*)&load_forces,
        (const void*)&compute_forces, (const void*)&init_forces, (const void*)&free_forces,
        (const void*)&my_dpotrf, (const void*)&my_dtrmv, (const void*)&init_cholesky,
        (const void*)&free_cholesky, (const void*)&newton_moves, (const void*)&newton_init_vel,
        (const void*)&init_newton, (const void*)&free_newton,
    };
    const void* vars[] = {
        &visc, &vfactor, &hydro, &algorithm, &move_attempts, &vel_grad_tensor, &e_collision,
        &curr_dt, &ewald_method, &geyer_on_the_fly, &ewald, &preDii, &preDij, &X, &Rf, &Rs,
        &epsilon, &FS, &Rf0, &S, &save_iter, &curr_iter, &ewald_real, &ewald_recip, &ewald_alpha,
        &variable_visc, &variable_vfactor, &variable_hydro, &variable_algorithm,
        &variable_move_attempts, &variable_vel_grad_tensor, &variable_e_collision,
        &variable_curr_dt, &variable_ewald_method, &variable_geyer_on_the_fly,
        &variable_ewald_real, &variable_ewald_recip, &variable_ewald_alpha,
        &E, &ES, &F, &repeat, &bond_lj_scale, &bond_c_scale, &check_overlap,
        &variable_bond_lj_scale, &variable_bond_c_scale, &variable_check_overlap,
    };
    int k, n = 0;
    for (k = 0; k < (int)(sizeof(fns) / sizeof(fns[0])); ++k) n += fns[k] != NULL;
    for (k = 0; k < (int)(sizeof(vars) / sizeof(vars[0])); ++k) n += vars[k] != NULL;
    printf("%d symbols linked\n", n);
    return n == (int)(sizeof(fns) / sizeof(fns[0]) + sizeof(vars) / sizeof(vars[0])) ? 0 : 1;
}
