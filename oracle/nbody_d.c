/* Oracle (test infrastructure only): FP64 instantiation of nbody_impl.h, plus bullfrog_ics */
#define REAL double
#define SFX(name) name##_d
#include "nbody_impl.h"

/* run_nbody.jl:217-225: one BullFrog step from a = 0 with the discreteness-suppressing defaults */
int orc_bullfrog_ics_d(const orc_cosmo *c, int n, double boxsize, int res_pm, double a_ini,
                       const double *psi1, double *psi_out, double *pi_out) {
    orc_pm_config cfg = {res_pm, n, boxsize, 4, 1, 1, 0, 0, 2, 0};
    int64_t np3 = 3 * (int64_t)n * n * n;
    double D0 = orc_Dplus(c, 0.0);
    double *p0 = (double *)malloc(sizeof(double) * np3);
    for (int64_t i = 0; i < np3; i++) p0[i] = D0 * psi1[i];
    int err = orc_run_nbody_d(c, &cfg, ORC_STEPPER_BULLFROG, 0.0, a_ini, 1, ORC_TV_D, NULL, 0, p0, psi1,
                              psi_out, pi_out, NULL, NULL);
    double F = orc_Fplus(c, a_ini);
    for (int64_t i = 0; i < np3; i++) pi_out[i] = pi_out[i] / F;
    free(p0);
    return err;
}
