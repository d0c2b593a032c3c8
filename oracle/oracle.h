/*
 * CPU oracle for the particle-mesh N-body step — TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference's algorithms for the PM path, used solely as the
 * checker of the HIP library (tests, smoke check) and as the timed CPU baseline.  The HIP
 * product never links, loads or calls anything here.  Every function names the reference
 * file:line it restates (paths relative to the reference root; NR = native/DiscoDJNative).
 *
 * Parity pins (what checks each part independently of this code):
 *   cosmology tables, lookups, stepper coefficients .... JAX golden vectors (/cosmo, /<cfg>/alpha..)
 *   PM force, every option ............................. golden /<cfg>/acc_ini (1e-12)
 *   DKD loop, bullfrog_ics, resampling ................. golden /<cfg>/psi_out, p_out (1e-11)
 *   2LPT compute_core + nbody_ics_lpt .................. golden /ics/psi_ini, pi_ini (1e-12)
 *   FFT ................................................ brute-force DFT, numpy.fft vectors, Parseval
 *   ranlxd1 / N-GenIC field ............................ runtests.jl:406-421 known answers
 *   Eisenstein-Hu T(k), sigma8 normalisation ........... closed forms (T->1 as k->0, sigma8 re-integral)
 *   power spectrum ..................................... white-noise amplitude, single-mode closed form
 *
 * Layouts follow the reference: particle arrays are (n,n,n,3) column-major, element (i,j,k,c)
 * at i + n*j + n^2*k + n^3*c; meshes (N,N,N) at ix + N*iy + N^2*iz; half spectra (N,N,N/2+1)
 * halved along z (the [3,1,2] rfft region of NR/src/lpt/nlpt_core.jl:35-37).
 */
#ifndef DJ_ORACLE_H
#define DJ_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- cosmology (FP64 host) */
enum { ORC_TAB_A = 0, ORC_TAB_CHI, ORC_TAB_D1, ORC_TAB_D2, ORC_TAB_F1, ORC_TAB_D3A, ORC_TAB_D3B,
       ORC_TAB_D3C, ORC_TAB_F2, ORC_TAB_F3A, ORC_TAB_F3B, ORC_TAB_SUPERCONFT, ORC_TAB_DPLUSDA,
       ORC_TAB_D2PLUSDA, ORC_NTAB };

typedef struct orc_cosmo orc_cosmo;

orc_cosmo *orc_cosmo_create(double Omega_c, double Omega_b, double h, double sigma8, double n_s,
                            double Omega_k, double w0, double wa, int n_pts, double a_ini);
void orc_cosmo_destroy(orc_cosmo *c);
int orc_cosmo_npts(const orc_cosmo *c);
void orc_cosmo_table(const orc_cosmo *c, int which, double *out);
double orc_cosmo_D1_unnormed_at_1(const orc_cosmo *c);
double orc_Omega_m(const orc_cosmo *c);
double orc_hubble_E(const orc_cosmo *c, double a);
double orc_jinterp(double x, const double *xp, const double *fp, int n);
double orc_Dplus(const orc_cosmo *c, double a);
double orc_D2plus(const orc_cosmo *c, double a);
double orc_Dplusda(const orc_cosmo *c, double a);
double orc_D2plusda(const orc_cosmo *c, double a);
double orc_superconft(const orc_cosmo *c, double a);
double orc_a_of_superconft(const orc_cosmo *c, double s);
double orc_a_of_Dplus(const orc_cosmo *c, double D);
double orc_Fplus(const orc_cosmo *c, double a);

enum { ORC_STEPPER_BULLFROG = 0, ORC_STEPPER_FASTPM = 1, ORC_STEPPER_SYMPLECTIC = 2 };
enum { ORC_TV_A = 0, ORC_TV_LOG_A = 1, ORC_TV_D = 2, ORC_TV_SUPERCONFT = 3, ORC_TV_ARRAY = 4 };

/* returns 0, or -1 on an invalid option; arrays have n_steps entries, all_a n_steps+1 */
int orc_stepper_coefficients(const orc_cosmo *c, int stepper, double a_ini, double a_end, int n_steps,
                             int time_var, const double *tv_array, double *alpha, double *beta,
                             double *dd1, double *dd2, double *all_a);

void orc_eisenstein_hu(const orc_cosmo *c, const double *k, int n, double Tcmb, double *T);
void orc_eh98_scales(const orc_cosmo *c, double Tcmb, double *out);
void orc_linear_power_spectrum(const orc_cosmo *c, double kmin, double kmax, int n_modes,
                               double *k_out, double *pk_out);

/* ---------------------------------------------------------------- FFT (own mixed radix) */
/* complex 1-D, any n >= 1; sign -1 forward, +1 backward (unnormalised); in/out interleaved re,im */
void orc_fft1_d(int n, const double *in, double *out, int sign);
void orc_fft1_f(int n, const float *in, float *out, int sign);
/* 3-D real <-> half-complex, [3,1,2] region: out is (N,N,N/2+1) interleaved complex */
void orc_rfft3_d(int N, const double *in, double *out);
void orc_rfft3_f(int N, const float *in, float *out);
/* unnormalised inverse (brfft); irfft = brfft / N^3 */
void orc_brfft3_d(int N, const double *in, double *out);
void orc_brfft3_f(int N, const float *in, float *out);

/* ---------------------------------------------------------------- PM force */
typedef struct {
    int32_t res_pm, n_part;
    double boxsize;
    int32_t worder, deconvolve, antialias, grad_order, lap_order, n_resample;
    int32_t resampling; /* 0 = fourier, 1 = linear */
} orc_pm_config;

/* pm_acceleration_reference (pm.jl:253-298): k-space gradient, one inverse FFT per component */
int orc_pm_acceleration_ref_d(const orc_pm_config *cfg, const double *psi, double *acc);
int orc_pm_acceleration_ref_f(const orc_pm_config *cfg, const float *psi, float *acc);
/* PMSolver form (pm.jl:385-521): combined K_phi, real-space FD stencil for orders 2/4/6 */
int orc_pm_acceleration_d(const orc_pm_config *cfg, const double *psi, double *acc);
int orc_pm_acceleration_f(const orc_pm_config *cfg, const float *psi, float *acc);
/* PMSolverLean form (pm.jl:569-641): aa = 0, n_resample = 1, FD gradient; acc assigned per axis */
int orc_pm_acceleration_lean_d(const orc_pm_config *cfg, const double *psi, double *acc);
int orc_pm_acceleration_lean_f(const orc_pm_config *cfg, const float *psi, float *acc);

/* unit-mass deposit (pm.jl:57-74) of np positions X (np,3) column-major, accumulated into mesh */
void orc_pm_scatter_d(double *mesh, const double *X, int64_t np, int res, double L, int W);
void orc_pm_scatter_f(float *mesh, const float *X, int64_t np, int res, float L, int W);
/* fixed-point deposit: every weight product w (computed exactly as the float scatter) is added
   as llrint(w * 2^frac_bits) to an int64 mesh — order independent, bit reproducible */
void orc_pm_scatter_fixed_f(int64_t *mesh, const float *X, int64_t np, int res, float L, int W,
                            int frac_bits);
void orc_pm_scatter_fixed_d(int64_t *mesh, const double *X, int64_t np, int res, double L, int W,
                            int frac_bits);
/* single-component gather (pm.jl:76-95) */
void orc_pm_gather_d(double *out, const double *mesh, const double *X, int64_t np, int res, double L, int W);
void orc_pm_gather_f(float *out, const float *mesh, const float *X, int64_t np, int res, float L, int W);
/* per-particle base cell and 1-D weights (pm.jl:49-68): base (np,3) int32, w (np,3,W) */
void orc_pm_cells_f(const float *X, int64_t np, int res, float L, int W, int32_t *base, float *w);
void orc_pm_cells_d(const double *X, int64_t np, int res, double L, int W, int32_t *base, double *w);

/* ---------------------------------------------------------------- N-body driver */
typedef void (*orc_step_cb)(int k, double a, void *user);
/* run_nbody (run_nbody.jl:124-172) with return_displacement = true: psi_out = Psi, p_out = P.
   lean != 0 selects the PMSolverLean force. cb (may be NULL) is called after every step k >= 0. */
int orc_run_nbody_d(const orc_cosmo *c, const orc_pm_config *cfg, int stepper, double a_ini,
                    double a_end, int n_steps, int time_var, const double *tv_array, int lean,
                    const double *psi_ini, const double *pi_ini, double *psi_out, double *p_out,
                    orc_step_cb cb, void *user);
int orc_run_nbody_f(const orc_cosmo *c, const orc_pm_config *cfg, int stepper, double a_ini,
                    double a_end, int n_steps, int time_var, const double *tv_array, int lean,
                    const float *psi_ini, const float *pi_ini, float *psi_out, float *p_out,
                    orc_step_cb cb, void *user);
/* bullfrog_ics (run_nbody.jl:217-225) with the reference's default settings */
int orc_bullfrog_ics_d(const orc_cosmo *c, int n, double boxsize, int res_pm, double a_ini,
                       const double *psi1, double *psi_out, double *pi_out);

/* ---------------------------------------------------------------- initial conditions */
/* nLPT compute_core at n_order = 2 (nlpt_core.jl:311-362 / nlpt_fast.jl:225-249) on ext = 3n/2:
   fphi (n,n,n/2+1) interleaved complex -> psi1, psi2 (n,n,n,3) */
int orc_compute_core2_d(int n, double boxsize, int ext, const double *fphi, double *psi1, double *psi2);
/* nbody_ics_lpt (run_nbody.jl:180-208), EdS growth, order 2 */
void orc_nbody_ics_lpt2_d(const orc_cosmo *c, int64_t np3, double a, const double *psi1,
                          const double *psi2, double *psi, double *pi);
/* ic_operator (grf.jl:118-128): scale (res,res,res/2+1) = -sqrt(P(k) (res/L)^3)/k^2, 0 at k = 0 */
void orc_ic_operator_d(int res, double boxsize, const double *k_tab, const double *pk_tab, int ntab,
                       double *scale);
/* white_noise_to_fphi (grf.jl:139-144): rfft(white) * scale */
void orc_white_noise_to_fphi_d(int res, const double *scale, const double *white, double *fphi);

/* ---------------------------------------------------------------- N-GenIC (GSL ranlxd1) */
void orc_ranlxd1_uniforms(uint32_t seed, int n, double *out);
/* ngenic_field_gsl (ngenic_gsl.jl:126-165): (res,res,res/2+1) interleaved complex */
void orc_ngenic_field_gsl(uint32_t seed, int res, double *out);
/* ngenic_wnoise_real (ngenic_gsl.jl:174-177) */
void orc_ngenic_wnoise_real(uint32_t seed, int res, double *out);

/* ---------------------------------------------------------------- power spectrum */
/* evaluate_power_spectrum (power_spectrum.jl:25-79); returns bins filled */
int orc_power_spectrum_d(int res, const double *field, double boxsize, int bins, int logarithmic,
                         double *k_out, double *pk_out, int64_t *nmodes_out);
/* linear bins of width k_f from k_f up to kmax (inclusive lower edges); deconvolve the CIC/TSC/PCS
   window of order W (0 = none) — the accuracy measurement of the PM path */
int orc_power_spectrum_lin_d(int res, const double *field, double boxsize, int W, int nbins,
                             double *k_out, double *pk_out, int64_t *nmodes_out);
/* evaluate_cross_power_spectrum (power_spectrum.jl:86-123) */
int orc_cross_power_spectrum_d(int res, const double *f1, const double *f2, double boxsize, int bins,
                               int logarithmic, double *k_out, double *pk_out);
/* pk_full of the study estimator (measure.jl:403-447) over given bin edges: P_delta of a density
   contrast on the deposit mesh (deconv_W: window order divided out, 0 none) and P_psi of a
   displacement (3, n^3) on the particle grid; NULL skips a part */
void orc_pk_binned_d(int res, const double *delta, int n, const double *psi, double boxsize, int nbins,
                     const double *edges, int deconv_W, double *pk_delta, double *pk_psi);

#ifdef __cplusplus
}
#endif
#endif
