/*
 * Oracle (test infrastructure only): the DKD driver of the reference, instantiated for
 * REAL = double / float by nbody_d.c / nbody_f.c.  Restates NR/src/nbody/run_nbody.jl:104-172
 * (run_nbody with return_displacement = true) on top of the PMSolver / PMSolverLean forces.
 */
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

int SFX(orc_run_nbody)(const orc_cosmo *c, const orc_pm_config *cfg, int stepper, double a_ini,
                       double a_end, int n_steps, int time_var, const double *tv_array, int lean,
                       const REAL *psi_ini, const REAL *pi_ini, REAL *psi_out, REAL *p_out,
                       orc_step_cb cb, void *user) {
    int n = cfg->n_part;
    int64_t np3 = 3 * (int64_t)n * n * n;
    double *al = (double *)malloc(sizeof(double) * n_steps), *be = (double *)malloc(sizeof(double) * n_steps);
    double *d1 = (double *)malloc(sizeof(double) * n_steps), *d2 = (double *)malloc(sizeof(double) * n_steps);
    double *aa = (double *)malloc(sizeof(double) * (n_steps + 1));
    int err = orc_stepper_coefficients(c, stepper, a_ini, a_end, n_steps, time_var, tv_array, al, be, d1, d2, aa);
    REAL *Psi = psi_out, *Mom = (REAL *)malloc(sizeof(REAL) * np3), *acc = (REAL *)malloc(sizeof(REAL) * np3);
    if (err) goto done;
    memcpy(Psi, psi_ini, sizeof(REAL) * np3);
    /* run_nbody.jl:104: the symplectic integrator's momentum is Pi * F+(a) */
    if (stepper == ORC_STEPPER_SYMPLECTIC) {
        double F = orc_Fplus(c, a_ini);
        for (int64_t i = 0; i < np3; i++) Mom[i] = (REAL)((double)pi_ini[i] * F);
    } else {
        memcpy(Mom, pi_ini, sizeof(REAL) * np3);
    }
    if (cb) cb(0, aa[0], user);
    for (int k = 0; k < n_steps && !err; k++) {
        REAL t1 = (REAL)d1[k], ta = (REAL)al[k], tb = (REAL)be[k], t2 = (REAL)d2[k];
        for (int64_t i = 0; i < np3; i++) Psi[i] += t1 * Mom[i];
        err = lean ? SFX(orc_pm_acceleration_lean)(cfg, Psi, acc) : SFX(orc_pm_acceleration)(cfg, Psi, acc);
        for (int64_t i = 0; i < np3; i++) Mom[i] = ta * Mom[i] + tb * acc[i];
        for (int64_t i = 0; i < np3; i++) Psi[i] += t2 * Mom[i];
        if (cb) cb(k + 1, aa[k + 1], user);
    }
    /* canonical momentum P = Pi * F+(a_end) (run_nbody.jl:149, 170) */
    double F = orc_Fplus(c, a_end);
    for (int64_t i = 0; i < np3; i++) {
        double v = (double)Mom[i];
        if (stepper == ORC_STEPPER_SYMPLECTIC) v = v / F;
        p_out[i] = (REAL)(v * F);
    }
done:
    free(al); free(be); free(d1); free(d2); free(aa); free(Mom); free(acc);
    return err;
}
